// The generic_value wire format for the C++ mirror: tensors, lists, tuples and
// string-keyed dicts of them as bytes (the reference's
// cpp/corelib/lib/generic_value/generic_value.cppm:20-37, 261-437, 456-666).
// Host code only; the same bytes as hastycucompute_amd/generic_value.py.
//
// Layout (little-endian):
//   value   = i64 type tag (0 None, 1 tensor, 2 list, 3 dict, 4 tuple) + body
//   tensor  = 16-byte header {u8 device_type (0 cpu, 1 cuda), u8 ATen
//             ScalarType code, u8 ndim, i8 device_index (-1 on the host),
//             4 zero bytes of padding, i64 total_elements}, ndim i64 sizes,
//             then the contiguous element bytes
//   list / tuple = i64 count, elements
//   dict    = i64 count, then per entry i64 key length, key bytes, value
//
// The tag is 8 bytes as the reference's writer has it (:349); its reader takes
// one byte (:400), which would fall out of step with the writer, so the reader
// here takes 8.  Dicts are std::map (keys in sorted order, so equal values
// give equal bytes); the reference's unordered_map writes them in hash order.
#pragma once

#include <cstdint>
#include <cstring>
#include <istream>
#include <map>
#include <ostream>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "corelib.hpp"

namespace hasty {

enum class value_type : uint8_t { NONE = 0, TENSOR = 1, VECTOR = 2, DICT = 3, TUPLE = 4 };

namespace detail_gv {
// ATen ScalarType codes (c10/core/ScalarType.h)
template <class TT> constexpr uint8_t scalar_code() {
    if constexpr (std::is_same_v<TT, f32_t>) return 6;
    else if constexpr (std::is_same_v<TT, f64_t>) return 7;
    else if constexpr (std::is_same_v<TT, c64_t>) return 9;
    else if constexpr (std::is_same_v<TT, c128_t>) return 10;
    else if constexpr (std::is_same_v<TT, int16_t>) return 2;
    else if constexpr (std::is_same_v<TT, int32_t>) return 3;
    else if constexpr (std::is_same_v<TT, int64_t>) return 4;
    else if constexpr (std::is_same_v<TT, bool>) return 11;
    else if constexpr (std::is_same_v<TT, uint8_t>) return 0;
    else static_assert(sizeof(TT) == 0, "no ScalarType code for this element type");
}
inline size_t element_size(uint8_t code) {
    switch (code) {
        case 0: case 1: case 11: return 1;
        case 2: case 5: case 15: return 2;
        case 3: case 6: case 8: return 4;
        case 4: case 7: case 9: return 8;
        case 10: return 16;
        default: throw std::runtime_error("unknown scalar type code " + std::to_string(int(code)));
    }
}
// generic_value.cppm:575-605
inline std::string type_name(uint8_t code) {
    switch (code) {
        case 6: return "f32"; case 7: return "f64"; case 9: return "c64"; case 10: return "c128";
        case 2: return "i16"; case 3: return "i32"; case 4: return "i64"; case 11: return "b8"; case 0: return "u8";
        default: return "unknown";
    }
}
inline std::string long_name(uint8_t code) {
    switch (code) {
        case 6: return "Float32"; case 7: return "Float64"; case 9: return "Complex64"; case 10: return "Complex128";
        case 3: return "Int32"; case 4: return "Int64"; case 2: return "Int16"; case 0: return "UInt8";
        case 11: return "Bool";
        default: return "Unknown";
    }
}
inline void put_i64(std::vector<uint8_t>& out, int64_t v) {
    uint8_t b[8];
    std::memcpy(b, &v, 8);  // little-endian host
    out.insert(out.end(), b, b + 8);
}
struct reader {
    const uint8_t* p;
    size_t n, pos = 0;
    const uint8_t* take(size_t k, const char* what) {
        if (k > n - pos) throw std::runtime_error(std::string("Buffer too small to deserialize ") + what);
        const uint8_t* r = p + pos;
        pos += k;
        return r;
    }
    int64_t i64(const char* what) {
        int64_t v;
        std::memcpy(&v, take(8, what), 8);
        return v;
    }
};
}  // namespace detail_gv

// a tensor in the format: element type code, device and the host bytes
struct wire_tensor {
    uint8_t device_type = 0;  // 0 cpu, 1 cuda
    uint8_t scalar_type = 6;
    int8_t device_index = -1;
    std::vector<int64_t> shape;
    std::vector<uint8_t> bytes;

    int64_t numel() const {
        int64_t n = 1;
        for (auto s : shape) n *= s;
        return n;
    }
    template <class TT, size_t R> static wire_tensor from(const tensor<cpu_t, TT, R>& t) {
        wire_tensor w;
        w.scalar_type = detail_gv::scalar_code<TT>();
        w.shape.assign(t.shape_arr().begin(), t.shape_arr().end());
        w.bytes.resize(size_t(t.numel()) * sizeof(TT));
        if (!w.bytes.empty()) std::memcpy(w.bytes.data(), t.data_ptr(), w.bytes.size());
        return w;
    }
    // a device tensor, written from a host copy and marked with the current device
    template <class TT, size_t R> static wire_tensor from(const tensor<cuda_t, TT, R>& t) {
        tensor<cpu_t, TT, R> h(t.shape_arr());
        h.vec() = t.to_host();
        wire_tensor w = from(h);
        w.device_type = 1;
        w.device_index = int8_t(int(detail::current_device()));
        return w;
    }
    // the host data as tensor<cpu_t, TT, R>; throws on another type or rank
    template <class TT, size_t R> tensor<cpu_t, TT, R> as() const {
        if (scalar_type != detail_gv::scalar_code<TT>()) throw std::runtime_error("wire tensor has another element type");
        if (shape.size() != R) throw std::runtime_error("Invalid number of dimensions");
        std::array<int64_t, R> s{};
        for (size_t i = 0; i < R; ++i) s[i] = shape[i];
        tensor<cpu_t, TT, R> t(s);
        if (!bytes.empty()) std::memcpy(t.mutable_data_ptr(), bytes.data(), bytes.size());
        return t;
    }
};

class generic_value {
public:
    using list = std::vector<generic_value>;
    using dict = std::map<std::string, generic_value>;

    generic_value() = default;  // None
    generic_value(wire_tensor t) : _type(value_type::TENSOR), _data(std::move(t)) {}
    generic_value(list v) : _type(value_type::VECTOR), _data(std::move(v)) {}
    generic_value(dict d) : _type(value_type::DICT), _data(std::move(d)) {}
    static generic_value make_tuple(list v) {
        generic_value g(std::move(v));
        g._type = value_type::TUPLE;
        return g;
    }

    value_type type() const { return _type; }
    const wire_tensor& to_tensor() const { return get<wire_tensor>("tensor"); }
    const list& to_vector() const {
        if (_type != value_type::VECTOR) throw std::runtime_error("generic_value is not a list");
        return std::get<list>(_data);
    }
    const list& to_tuple() const {
        if (_type != value_type::TUPLE) throw std::runtime_error("generic_value is not a tuple");
        return std::get<list>(_data);
    }
    const dict& to_dict() const { return get<dict>("dict"); }

    std::vector<uint8_t> serialize() const {
        std::vector<uint8_t> out;
        write(out);
        return out;
    }
    void serialize(std::ostream& os) const {
        const auto b = serialize();
        os.write(reinterpret_cast<const char*>(b.data()), std::streamsize(b.size()));
    }
    static generic_value deserialize(const uint8_t* data, size_t n) {
        detail_gv::reader r{data, n};
        return read(r);
    }
    static generic_value deserialize(const std::vector<uint8_t>& b) { return deserialize(b.data(), b.size()); }
    static generic_value deserialize(std::istream& is) {
        std::vector<uint8_t> b((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
        return deserialize(b);
    }

    // generic_value.cppm:261-290
    std::string type_string() const {
        switch (_type) {
            case value_type::NONE: return "None";
            case value_type::TENSOR: {
                const auto& t = to_tensor();
                return "Tensor[" + std::string(t.device_type == 1 ? "cuda" : "cpu") + "," +
                       detail_gv::type_name(t.scalar_type) + "," + std::to_string(t.shape.size()) + "]";
            }
            case value_type::VECTOR: {
                const auto& v = to_vector();
                return v.empty() ? "List[]" : "List[" + v[0].type_string() + "]";
            }
            case value_type::TUPLE: {
                std::string s = "Tuple[";
                const auto& v = to_tuple();
                for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + v[i].type_string();
                return s + "]";
            }
            case value_type::DICT: {
                const auto& d = to_dict();
                return d.empty() ? "Dict[String,?]" : "Dict[String," + d.begin()->second.type_string() + "]";
            }
        }
        return "Unknown";
    }
    // generic_value.cppm:302-337
    std::string metadata_string() const {
        switch (_type) {
            case value_type::NONE: return "None";
            case value_type::TENSOR: {
                const auto& t = to_tensor();
                std::string s = "Tensor[" +
                                (t.device_type == 1 ? "cuda:" + std::to_string(int(t.device_index)) : std::string("cpu")) +
                                "," + detail_gv::long_name(t.scalar_type) + ",shape=(";
                for (size_t i = 0; i < t.shape.size(); ++i) s += (i ? "," : "") + std::to_string(t.shape[i]);
                return s + ")]";
            }
            case value_type::VECTOR: {
                const auto& v = to_vector();
                return "List[n=" + std::to_string(v.size()) + (v.empty() ? "" : ",elem=" + v[0].metadata_string()) + "]";
            }
            case value_type::TUPLE: {
                std::string s = "Tuple[";
                const auto& v = to_tuple();
                for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + v[i].metadata_string();
                return s + "]";
            }
            case value_type::DICT: {
                const auto& d = to_dict();
                return "Dict[n=" + std::to_string(d.size()) +
                       (d.empty() ? "" : ",value_type=" + d.begin()->second.metadata_string()) + "]";
            }
        }
        return "Unknown";
    }

private:
    template <class T> const T& get(const char* what) const {
        const T* p = std::get_if<T>(&_data);
        if (!p) throw std::runtime_error(std::string("generic_value is not a ") + what);
        return *p;
    }

    void write(std::vector<uint8_t>& out) const {
        detail_gv::put_i64(out, int64_t(_type));
        switch (_type) {
            case value_type::NONE: break;
            case value_type::TENSOR: {
                const auto& t = to_tensor();
                if (int64_t(t.bytes.size()) != t.numel() * int64_t(detail_gv::element_size(t.scalar_type)))
                    throw std::runtime_error("wire tensor bytes do not match its shape");
                uint8_t h[16] = {};
                h[0] = t.device_type;
                h[1] = t.scalar_type;
                h[2] = uint8_t(t.shape.size());
                h[3] = uint8_t(t.device_index);
                const int64_t total = t.numel();
                std::memcpy(h + 8, &total, 8);
                out.insert(out.end(), h, h + 16);
                for (auto s : t.shape) detail_gv::put_i64(out, s);
                out.insert(out.end(), t.bytes.begin(), t.bytes.end());
                break;
            }
            case value_type::VECTOR:
            case value_type::TUPLE: {
                const auto& v = std::get<list>(_data);
                detail_gv::put_i64(out, int64_t(v.size()));
                for (const auto& e : v) e.write(out);
                break;
            }
            case value_type::DICT: {
                const auto& d = to_dict();
                detail_gv::put_i64(out, int64_t(d.size()));
                for (const auto& [k, e] : d) {
                    detail_gv::put_i64(out, int64_t(k.size()));
                    out.insert(out.end(), k.begin(), k.end());
                    e.write(out);
                }
                break;
            }
        }
    }

    static generic_value read(detail_gv::reader& r) {
        const int64_t tag = r.i64("type tag");
        switch (tag) {
            case int64_t(value_type::NONE): return generic_value();
            case int64_t(value_type::TENSOR): {
                const uint8_t* h = r.take(16, "tensor header");
                wire_tensor t;
                t.device_type = h[0];
                t.scalar_type = h[1];
                t.device_index = int8_t(h[3]);
                int64_t total;
                std::memcpy(&total, h + 8, 8);
                t.shape.resize(h[2]);
                int64_t count = 1;   // checked product: a corrupt header must not overflow it
                for (auto& s : t.shape) {
                    s = r.i64("shape");
                    if (s < 0 || (s != 0 && count > INT64_MAX / s))
                        throw std::runtime_error("Tensor shape does not match total elements in header");
                    count *= s;
                }
                if (count != total) throw std::runtime_error("Tensor shape does not match total elements in header");
                const size_t es = detail_gv::element_size(t.scalar_type);
                if (size_t(total) > (r.n - r.pos) / es) throw std::runtime_error("Buffer too small to deserialize tensor data");
                const size_t nb = size_t(total) * es;
                const uint8_t* d = r.take(nb, "tensor data");
                t.bytes.assign(d, d + nb);
                return generic_value(std::move(t));
            }
            case int64_t(value_type::VECTOR):
            case int64_t(value_type::TUPLE): {
                const int64_t n = r.i64("length");
                if (n < 0) throw std::runtime_error("negative length");
                list v;
                for (int64_t i = 0; i < n; ++i) v.push_back(read(r));
                return tag == int64_t(value_type::TUPLE) ? make_tuple(std::move(v)) : generic_value(std::move(v));
            }
            case int64_t(value_type::DICT): {
                const int64_t n = r.i64("length");
                if (n < 0) throw std::runtime_error("negative length");
                dict d;
                for (int64_t i = 0; i < n; ++i) {
                    const int64_t kl = r.i64("key length");
                    if (kl < 0) throw std::runtime_error("negative length");
                    const uint8_t* k = r.take(size_t(kl), "key");
                    std::string key(reinterpret_cast<const char*>(k), size_t(kl));
                    d.emplace(std::move(key), read(r));
                }
                return generic_value(std::move(d));
            }
            default: throw std::runtime_error("unknown generic_value type tag " + std::to_string(tag));
        }
    }

    value_type _type = value_type::NONE;
    std::variant<std::monostate, wire_tensor, list, dict> _data;
};

}  // namespace hasty
