// HDF5 tensor files for the C++ mirror: the reference's io/hdf5 module
// (cpp/corelib/lib/io/hdf5.cppm:14-19, hdf5.cpp:60-259) on the HDF5 C API.
// Separate from corelib.hpp so that only callers who include it need
// <hdf5.h> and -lhdf5.
//
//   import_tensor(path, dataset)          -> hdf5_tensor (dims + host data)
//   import_tensors(path, regex matchers)  -> {"/group/name": hdf5_tensor}, the
//                                            datasets whose full path matches
//                                            one regex (std::regex_match); all
//                                            when no matchers are given
//   export_tensor(tensor, path, dataset)   overwrites the file (File::Overwrite)
//
// Element types: float, double, and std::complex<float> / <double> stored as a
// compound of two float / double members (real part at offset 0; written with
// members "r" and "i", read with any two names).  Anything else throws
// "disallowed dtype", as the reference does.  The reference returns a
// hat::Tensor of runtime rank; here that is hdf5_tensor, whose as<TT, RANK>()
// gives a tensor<cpu_t, TT, RANK>.
#pragma once

#include <hdf5.h>

#include <optional>
#include <regex>
#include <string>
#include <unordered_map>
#include <variant>

#include "corelib.hpp"

namespace hasty {

struct hdf5_tensor {
    std::vector<int64_t> dims;
    std::variant<std::vector<f32_t>, std::vector<f64_t>, std::vector<c64_t>, std::vector<c128_t>> data;

    // the data as a tensor<cpu_t, TT, RANK>; throws when the element type or
    // the rank differs
    template <class TT, size_t RANK> tensor<cpu_t, TT, RANK> as() const {
        const auto* v = std::get_if<std::vector<TT>>(&data);
        if (!v) throw std::runtime_error("hdf5 tensor has another element type");
        if (dims.size() != RANK) throw std::runtime_error("Invalid number of dimensions");
        std::array<int64_t, RANK> shape{};
        for (size_t i = 0; i < RANK; ++i) shape[i] = dims[i];
        tensor<cpu_t, TT, RANK> t(shape);
        t.vec() = *v;
        return t;
    }
};

namespace detail_h5 {

// HDF5 handles closed on scope exit
struct handle {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    handle(hid_t i, herr_t (*c)(hid_t), const char* what) : id(i), close(c) {
        if (id < 0) throw std::runtime_error(std::string("HDF5: ") + what + " failed");
    }
    handle(const handle&) = delete;
    handle& operator=(const handle&) = delete;
    ~handle() {
        if (id >= 0) close(id);
    }
};

inline void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

template <class R> std::vector<std::complex<R>> read_complex(hid_t dset, hid_t ftype, size_t n) {
    // memory compound with the file's member names, real part first
    const size_t part = sizeof(R);
    std::string names[2];
    for (unsigned i = 0; i < 2; ++i) {
        char* nm = H5Tget_member_name(ftype, i);
        const size_t off = H5Tget_member_offset(ftype, i);
        if (!nm || (off != 0 && off != part)) {
            if (nm) H5free_memory(nm);
            throw std::runtime_error("disallowed dtype");
        }
        names[off == 0 ? 0 : 1] = nm;
        H5free_memory(nm);
    }
    handle mt(H5Tcreate(H5T_COMPOUND, 2 * part), H5Tclose, "H5Tcreate");
    const hid_t comp = sizeof(R) == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
    H5Tinsert(mt.id, names[0].c_str(), 0, comp);
    H5Tinsert(mt.id, names[1].c_str(), part, comp);
    std::vector<std::complex<R>> v(n);
    if (n && H5Dread(dset, mt.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
        throw std::runtime_error("HDF5: H5Dread failed");
    return v;
}

inline hdf5_tensor read(hid_t dset) {
    handle ftype(H5Dget_type(dset), H5Tclose, "H5Dget_type");
    handle space(H5Dget_space(dset), H5Sclose, "H5Dget_space");
    const int ndim = H5Sget_simple_extent_ndims(space.id);
    if (ndim < 0) throw std::runtime_error("HDF5: H5Sget_simple_extent_ndims failed");
    std::vector<hsize_t> d(size_t(std::max(ndim, 1)));
    if (ndim) H5Sget_simple_extent_dims(space.id, d.data(), nullptr);
    hdf5_tensor t;
    size_t n = 1;
    for (int i = 0; i < ndim; ++i) {
        t.dims.push_back(int64_t(d[size_t(i)]));
        n *= size_t(d[size_t(i)]);
    }
    const H5T_class_t cls = H5Tget_class(ftype.id);
    const size_t size = H5Tget_size(ftype.id);
    auto read_real = [&](auto tag, hid_t mem) {
        using R = decltype(tag);
        std::vector<R> v(n);
        if (n && H5Dread(dset, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
            throw std::runtime_error("HDF5: H5Dread failed");
        t.data = std::move(v);
    };
    if (cls == H5T_FLOAT && size == 4) {
        read_real(f32_t{}, H5T_NATIVE_FLOAT);
    } else if (cls == H5T_FLOAT && size == 8) {
        read_real(f64_t{}, H5T_NATIVE_DOUBLE);
    } else if (cls == H5T_COMPOUND && (size == 8 || size == 16) && H5Tget_nmembers(ftype.id) == 2) {
        for (unsigned i = 0; i < 2; ++i) {
            handle mt(H5Tget_member_type(ftype.id, i), H5Tclose, "H5Tget_member_type");
            if (H5Tget_class(mt.id) != H5T_FLOAT || H5Tget_size(mt.id) != size / 2)
                throw std::runtime_error("disallowed dtype");
        }
        if (size == 8) t.data = read_complex<float>(dset, ftype.id, n);
        else t.data = read_complex<double>(dset, ftype.id, n);
    } else {
        throw std::runtime_error("disallowed dtype");
    }
    return t;
}

inline void walk(hid_t gid, const std::string& prefix, const std::optional<std::vector<std::regex>>& matchers,
                 std::unordered_map<std::string, hdf5_tensor>& out) {
    H5G_info_t info;
    if (H5Gget_info(gid, &info) < 0) throw std::runtime_error("HDF5: H5Gget_info failed");
    for (hsize_t i = 0; i < info.nlinks; ++i) {
        const ssize_t len = H5Lget_name_by_idx(gid, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
        if (len < 0) throw std::runtime_error("HDF5: H5Lget_name_by_idx failed");
        std::string name(size_t(len) + 1, '\0');
        H5Lget_name_by_idx(gid, ".", H5_INDEX_NAME, H5_ITER_INC, i, name.data(), name.size(), H5P_DEFAULT);
        name.resize(size_t(len));
        handle obj(H5Oopen(gid, name.c_str(), H5P_DEFAULT), H5Oclose, "H5Oopen");
        const std::string full = prefix + "/" + name;
        const H5I_type_t kind = H5Iget_type(obj.id);
        if (kind == H5I_DATASET) {
            bool take = !matchers.has_value();
            if (matchers)
                for (const auto& m : *matchers)
                    if (std::regex_match(full, m)) { take = true; break; }
            if (take) out.emplace(full, read(obj.id));
        } else if (kind == H5I_GROUP) {
            walk(obj.id, full, matchers, out);
        } else {
            throw std::runtime_error("unsupported object type");
        }
    }
}

}  // namespace detail_h5

inline hdf5_tensor import_tensor(const std::string& filepath, const std::string& dataset) {
    detail_h5::quiet();
    detail_h5::handle f(H5Fopen(filepath.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "opening the file");
    detail_h5::handle d(H5Dopen2(f.id, dataset.c_str(), H5P_DEFAULT), H5Dclose, "opening the dataset");
    return detail_h5::read(d.id);
}

inline std::unordered_map<std::string, hdf5_tensor> import_tensors(
    const std::string& filepath, const std::optional<std::vector<std::regex>>& matchers) {
    detail_h5::quiet();
    detail_h5::handle f(H5Fopen(filepath.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "opening the file");
    std::unordered_map<std::string, hdf5_tensor> out;
    detail_h5::walk(f.id, "", matchers, out);
    return out;
}

template <class TT, size_t RANK>
void export_tensor(const tensor<cpu_t, TT, RANK>& t, const std::string& filepath, const std::string& dataset) {
    using namespace detail_h5;
    quiet();
    handle f(H5Fcreate(filepath.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "creating the file");
    std::array<hsize_t, RANK == 0 ? 1 : RANK> dims{};
    for (size_t i = 0; i < RANK; ++i) dims[i] = hsize_t(t.shape_arr()[i]);
    handle space(H5Screate_simple(int(RANK), dims.data(), nullptr), H5Sclose, "H5Screate_simple");
    handle lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose, "H5Pcreate");
    H5Pset_create_intermediate_group(lcpl.id, 1);
    auto write = [&](hid_t ftype, hid_t mtype) {
        handle d(H5Dcreate2(f.id, dataset.c_str(), ftype, space.id, lcpl.id, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
                 "creating the dataset");
        if (t.numel() && H5Dwrite(d.id, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, t.data_ptr()) < 0)
            throw std::runtime_error("HDF5: H5Dwrite failed");
    };
    if constexpr (std::is_same_v<TT, f32_t>) {
        write(H5T_IEEE_F32LE, H5T_NATIVE_FLOAT);
    } else if constexpr (std::is_same_v<TT, f64_t>) {
        write(H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE);
    } else {
        static_assert(std::is_same_v<TT, c64_t> || std::is_same_v<TT, c128_t>, "disallowed dtype");
        using R = real_t<TT>;
        const size_t part = sizeof(R);
        handle ft(H5Tcreate(H5T_COMPOUND, 2 * part), H5Tclose, "H5Tcreate");
        handle mt(H5Tcreate(H5T_COMPOUND, 2 * part), H5Tclose, "H5Tcreate");
        const hid_t fp = part == 4 ? H5T_IEEE_F32LE : H5T_IEEE_F64LE;
        const hid_t mp = part == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
        H5Tinsert(ft.id, "r", 0, fp);
        H5Tinsert(ft.id, "i", part, fp);
        H5Tinsert(mt.id, "r", 0, mp);
        H5Tinsert(mt.id, "i", part, mp);
        write(ft.id, mt.id);
    }
}

}  // namespace hasty
