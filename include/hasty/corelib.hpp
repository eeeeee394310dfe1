// corelib.hpp -- C++17 host mirror of HastyCuCompute's corelib operator API
// for the non-Cartesian reconstruction path, over the C-ABI of hasty_hip.h.
//
// Names, template shapes, argument order and error behaviour follow the
// reference so that call sites read the same:
//   nufft_opts / nufft_plan<cuda_t,TT,DIM,NT>   cpp/corelib/lib/fft/nufft.cppm:42-310
//   nufft_backward_cpu_over_cuda               nufft.cppm:316-382
//   toeplitz_kernel / toeplitz_multiply        nufft.cppm:386-479
//   fft::ToeplitzMultType, fft::ToeplitzAccumulateType,
//   fft::transform_toeplitz_kernel, fft::toeplitz_multiplication
//                                              cpp/corelib/lib/fft/toeplitz.cppm:47-93
//   NORMAL_T_T1_OP<cuda_t,c64_t,DIM>          cpp/corelib/lib/mri/normal.cppm:43-93
//   NORMAL_T_T2_OP, NORMAL_IDT_T1_OP, NORMAL_IDT_T2_OP,
//   NORMAL_IDTW_T1_OP, NORMAL_IDTW_T2_OP       cpp/corelib/lib/mri/normal.cppm:113-735
//   sense_forward_offresonance                 cpp/corelib/lib/mri/forward.cppm:36-165
//   sense_adjoint_offresonance                 cpp/corelib/lib/mri/adjoint.cppm:33-171
//   conjugate_gradient(sptr<Op>, max_iter, tol) cpp/corelib/lib/op/cg.cppm:122-184
//   conjugate_gradient(sptr<Op>, sptr<PrecondOp>, inner, outer, tol)   cg.cppm:16-119
//   max_eig(A, x0, max_iter)                   cpp/corelib/lib/op/alg.cppm:13-30
//
// The reference's tensor<D,TT,RANK> wraps a torch tensor; here tensor<cuda_t,..>
// is a small device buffer (hipMalloc'd and reference-counted, or a view of
// caller memory) and tensor<cpu_t,..> a host vector, so the header needs only
// the HIP runtime API and libhasty_hip.so.  Build a program with
//   g++ -std=c++17 -D__HIP_PLATFORM_AMD__ -I<repo>/include -I/opt/rocm/include
//       prog.cpp -L<repo>/hastycucompute_amd -lhasty_hip -L/opt/rocm/lib -lamdhip64
// Every failure of the library is thrown as std::runtime_error carrying the
// library's message (hasty_last_error).
#ifndef HASTY_CORELIB_HPP
#define HASTY_CORELIB_HPP

#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../hasty_hip.h"
#include "threading.hpp"

namespace hasty {

// ---- device tags and element types (util_typing in the reference) --------
struct cuda_t {};
struct cpu_t {};
using f32_t = float;
using f64_t = double;
using c64_t = std::complex<float>;
using c128_t = std::complex<double>;
using i32 = int32_t;
template <class T> using sptr = std::shared_ptr<T>;
template <class T> using uptr = std::unique_ptr<T>;
template <class T> using optrefw = std::optional<std::reference_wrapper<T>>;
template <class T> using optcrefw = std::optional<std::reference_wrapper<const T>>;

template <class TT> struct complex_of;
template <> struct complex_of<f32_t> { using type = c64_t; };
template <> struct complex_of<f64_t> { using type = c128_t; };
template <class TT> using complex_t = typename complex_of<TT>::type;
template <class TT> struct real_of { using type = TT; };
template <class R> struct real_of<std::complex<R>> { using type = R; };
template <class TT> using real_t = typename real_of<TT>::type;

namespace detail {
inline void check(int status) {
    if (status != HASTY_OK) throw std::runtime_error(std::string("hasty: ") + hasty_last_error());
}
inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace detail

// ---- tensors ---------------------------------------------------------------
template <class D, class TT, size_t RANK> class tensor;

template <class TT, size_t RANK> class tensor<cuda_t, TT, RANK> {
public:
    tensor() = default;
    // view of caller-owned device memory
    tensor(TT* data, std::array<int64_t, RANK> shape) : _ptr(data), _shape(shape) {}

    static tensor empty(std::array<int64_t, RANK> shape) {
        tensor t;
        t._shape = shape;
        const size_t bytes = size_t(t.numel()) * sizeof(TT);
        void* p = nullptr;
        if (bytes) detail::check_hip(hipMalloc(&p, bytes), "hipMalloc");
        t._owner = std::shared_ptr<void>(p, [](void* q) { if (q) (void)hipFree(q); });
        t._ptr = static_cast<TT*>(p);
        return t;
    }
    static tensor zeros(std::array<int64_t, RANK> shape) {
        tensor t = empty(shape);
        t.fill_zero();
        return t;
    }
    static tensor from_host(const std::vector<TT>& h, std::array<int64_t, RANK> shape) {
        tensor t = empty(shape);
        if (int64_t(h.size()) != t.numel()) throw std::runtime_error("host data does not match shape");
        detail::check_hip(hipMemcpy(t._ptr, h.data(), h.size() * sizeof(TT), hipMemcpyHostToDevice),
                          "hipMemcpy H2D");
        return t;
    }
    std::vector<TT> to_host() const {
        std::vector<TT> h(static_cast<size_t>(numel()));
        detail::check_hip(hipMemcpy(h.data(), _ptr, h.size() * sizeof(TT), hipMemcpyDeviceToHost),
                          "hipMemcpy D2H");
        return h;
    }
    tensor clone() const {
        tensor t = empty(_shape);
        detail::check_hip(hipMemcpy(t._ptr, _ptr, size_t(numel()) * sizeof(TT), hipMemcpyDeviceToDevice),
                          "hipMemcpy D2D");
        return t;
    }
    void fill_zero() {
        if (numel()) detail::check_hip(hipMemset(_ptr, 0, size_t(numel()) * sizeof(TT)), "hipMemset");
    }

    template <size_t I> int64_t shape() const {
        static_assert(I < RANK, "shape index out of range");
        return _shape[I];
    }
    const std::array<int64_t, RANK>& shape_arr() const { return _shape; }
    int64_t numel() const {
        return std::accumulate(_shape.begin(), _shape.end(), int64_t(1), std::multiplies<int64_t>());
    }
    TT* mutable_data_ptr() { return _ptr; }
    const TT* data_ptr() const { return _ptr; }
    TT* unconsted_data_ptr() const { return _ptr; }

private:
    std::shared_ptr<void> _owner;
    TT* _ptr = nullptr;
    std::array<int64_t, RANK> _shape{};
};

template <class TT, size_t RANK> class tensor<cpu_t, TT, RANK> {
public:
    tensor() = default;
    explicit tensor(std::array<int64_t, RANK> shape)
        : _shape(shape), _data(static_cast<size_t>(std::accumulate(shape.begin(), shape.end(), int64_t(1),
                                                                    std::multiplies<int64_t>()))) {}
    template <size_t I> int64_t shape() const { return _shape[I]; }
    const std::array<int64_t, RANK>& shape_arr() const { return _shape; }
    int64_t numel() const { return int64_t(_data.size()); }
    TT* mutable_data_ptr() { return _data.data(); }
    const TT* data_ptr() const { return _data.data(); }
    std::vector<TT>& vec() { return _data; }
    const std::vector<TT>& vec() const { return _data; }

private:
    std::array<int64_t, RANK> _shape{};
    std::vector<TT> _data;
};

// ---- raw tensor files (util_io.cppm:47-84; the `.htc` cache blocks of
// mod_tensor_caching.cppm:84-96): the elements' bytes and nothing else, so
// the reader supplies the shape and element type.  A short file is an error
// here (the reference reads what is there).
template <class TT, size_t RANK>
void export_binary_tensor(const tensor<cpu_t, TT, RANK>& t, const std::string& path) {
    std::ofstream file(path, std::ios::binary);
    if (!file.is_open()) throw std::runtime_error("Could not open file for writing");
    file.write(reinterpret_cast<const char*>(t.data_ptr()), std::streamsize(t.numel() * int64_t(sizeof(TT))));
    if (!file) throw std::runtime_error("Exception opening/reading file");
}

template <class TT, size_t RANK>
tensor<cpu_t, TT, RANK> import_binary_tensor(const std::string& path, std::array<int64_t, RANK> sizes) {
    tensor<cpu_t, TT, RANK> t(sizes);
    std::ifstream file(path, std::ios::binary | std::ios::in);
    if (!file.is_open()) throw std::runtime_error("Could not open file for reading");
    const std::streamsize want = std::streamsize(t.numel() * int64_t(sizeof(TT)));
    file.read(reinterpret_cast<char*>(t.mutable_data_ptr()), want);
    if (file.gcount() != want) throw std::runtime_error("Exception opening/reading file");
    return t;
}

// ---- cache_tensor (mod_tensor_caching.cppm:14-335): one host tensor under a
// hash key, device copies made on first use and kept per device, and a
// spill to / reload from `cache_dir / "<hash>.htc"` (the raw tensor file
// above).  Copies of a cache_tensor share the same block.  get<cpu_t>()
// returns a copy of the host data (the reference returns a torch view);
// get_ptr<cpu_t>() shares it.
enum struct device_idx {
    CPU = -1, CUDA0 = 0, CUDA1, CUDA2, CUDA3, CUDA4, CUDA5, CUDA6, CUDA7, CUDA8, CUDA9, CUDA10, CUDA11,
    CUDA12, CUDA13, CUDA14, CUDA15, MAX_CUDA_DEVICES = 16
};

inline std::filesystem::path cache_dir = std::filesystem::temp_directory_path();

template <class TT, size_t RANK> class cache_tensor {
    struct block {
        std::mutex mutex;
        std::shared_ptr<tensor<cpu_t, TT, RANK>> cpu;
        std::array<int64_t, RANK> shape{};
        size_t hashidx = 0;
        std::array<std::shared_ptr<tensor<cuda_t, TT, RANK>>, size_t(device_idx::MAX_CUDA_DEVICES)> cuda;
    };
    std::shared_ptr<block> _block;

    void require(const char* what) const {
        if (!_block) throw std::runtime_error(std::string(what) + ": cache_tensor is not initialized");
    }
    std::filesystem::path file() const { return cache_dir / (std::to_string(_block->hashidx) + ".htc"); }
    std::shared_ptr<tensor<cpu_t, TT, RANK>>& cpu_ptr() {
        if (!_block->cpu)
            _block->cpu = std::make_shared<tensor<cpu_t, TT, RANK>>(
                import_binary_tensor<TT, RANK>(file().string(), _block->shape));
        return _block->cpu;
    }
    std::shared_ptr<tensor<cuda_t, TT, RANK>>& cuda_ptr(device_idx idx) {
        const int d = int(idx);
        if (d < 0 || d >= int(device_idx::MAX_CUDA_DEVICES)) throw std::runtime_error("invalid device index");
        auto& slot = _block->cuda[size_t(d)];
        if (!slot) {
            int prev = 0;
            detail::check_hip(hipGetDevice(&prev), "hipGetDevice");
            detail::check_hip(hipSetDevice(d), "hipSetDevice");
            try {
                slot = std::make_shared<tensor<cuda_t, TT, RANK>>(
                    tensor<cuda_t, TT, RANK>::from_host(cpu_ptr()->vec(), _block->shape));
            } catch (...) {
                (void)hipSetDevice(prev);
                throw;
            }
            detail::check_hip(hipSetDevice(prev), "hipSetDevice");
        }
        return slot;
    }

public:
    cache_tensor() = default;
    cache_tensor(tensor<cpu_t, TT, RANK>&& cputen, size_t hashidx) : _block(std::make_shared<block>()) {
        _block->shape = cputen.shape_arr();
        _block->cpu = std::make_shared<tensor<cpu_t, TT, RANK>>(std::move(cputen));
        _block->hashidx = hashidx;
    }
    template <size_t R> int64_t shape() const {
        static_assert(R < RANK, "shape index out of range");
        require("shape");
        return _block->shape[R];
    }
    std::array<int64_t, RANK> shape() const {
        require("shape");
        return _block->shape;
    }
    cache_tensor copy() const { return *this; }

    template <class D> std::shared_ptr<tensor<D, TT, RANK>> get_ptr(device_idx idx = device_idx::CPU) {
        require("get_ptr");
        std::unique_lock<std::mutex> lock(_block->mutex);
        if constexpr (std::is_same_v<D, cpu_t>) return cpu_ptr();
        else return cuda_ptr(idx);
    }
    template <class D> tensor<D, TT, RANK> get(device_idx idx = device_idx::CPU) { return *get_ptr<D>(idx); }

    void cache() {   // write the host data to cache_dir / "<hash>.htc"
        require("cache");
        std::unique_lock<std::mutex> lock(_block->mutex);
        if (!_block->cpu) throw std::runtime_error("cache_disk: no tensor to cache");
        export_binary_tensor(*_block->cpu, file().string());
    }
    void uncache() {  // reload the host data from the file
        require("uncache");
        std::unique_lock<std::mutex> lock(_block->mutex);
        _block->cpu = std::make_shared<tensor<cpu_t, TT, RANK>>(
            import_binary_tensor<TT, RANK>(file().string(), _block->shape));
    }
    void free_cpu() {
        require("free_cpu");
        std::unique_lock<std::mutex> lock(_block->mutex);
        _block->cpu = nullptr;
    }
    void free(device_idx idx) {
        require("free");
        std::unique_lock<std::mutex> lock(_block->mutex);
        const int d = int(idx);
        if (d < 0 || d >= int(device_idx::MAX_CUDA_DEVICES)) throw std::runtime_error("invalid device index");
        _block->cuda[size_t(d)] = nullptr;
    }
};

namespace detail {
inline device_idx current_device() {
    int d = 0;
    check_hip(hipGetDevice(&d), "hipGetDevice");
    return device_idx(d);
}
// a cache_tensor's copy on the current device (made once, then shared)
template <class TT, size_t RANK> tensor<cuda_t, TT, RANK> on_device(cache_tensor<TT, RANK>& c) {
    return c.template get<cuda_t>(current_device());
}
}  // namespace detail

template <class TT, size_t RANK>
tensor<cuda_t, TT, RANK> make_empty_tensor(std::array<int64_t, RANK> shape) {
    return tensor<cuda_t, TT, RANK>::empty(shape);
}

namespace detail {
template <class TT, size_t RANK> void require_c64(const tensor<cuda_t, TT, RANK>&) {
    static_assert(std::is_same_v<TT, c64_t>, "this operation is defined for c64_t tensors");
}
}  // namespace detail

// ---- enums (nufft.cppm:42-70) ------------------------------------------------
enum struct nufft_type { TYPE_1 = 1, TYPE_2 = 2, TYPE_3 = 3, FORWARD = TYPE_2, BACKWARD = TYPE_1 };
enum struct nufft_sign { POS = 1, NEG = -1, DEFAULT_TYPE_1 = POS, DEFAULT_TYPE_2 = NEG };
enum struct nufft_upsamp_cuda { UPSAMP_2_0, UPSAMP_1_25, DEFAULT };
enum struct nufft_method_cuda { GM_NO_SORT, GM_SORT, SM, DEFAULT };

// nufft.cppm:23-39 (same messages)
template <class TT, size_t DIM>
void verify_coords(const std::array<tensor<cuda_t, TT, 1>, DIM>& coords) {
    if constexpr (DIM == 2) {
        if (coords[0].template shape<0>() != coords[1].template shape<0>())
            throw std::runtime_error("x and y coords have different lengths");
    }
    if constexpr (DIM == 3) {
        if (coords[0].template shape<0>() != coords[1].template shape<0>() ||
            coords[1].template shape<0>() != coords[2].template shape<0>() ||
            coords[0].template shape<0>() != coords[2].template shape<0>())
            throw std::runtime_error("x,y  x,z  or  y,z  coords have different lengths");
    }
}

// ---- NUFFT options and plan (nufft.cppm:78-310) ------------------------------
template <class D, class TT, size_t DIM> struct nufft_opts;

template <class TT, size_t DIM> struct nufft_opts<cuda_t, TT, DIM> {
    static_assert(DIM == 2 || DIM == 3, "this library supports 2-D and 3-D transforms");
    std::array<int64_t, DIM> nmodes{};
    int32_t ntransf = 1;
    TT tol = TT(1e-6);
    nufft_sign sign = nufft_sign::DEFAULT_TYPE_1;
    nufft_upsamp_cuda upsamp = nufft_upsamp_cuda::DEFAULT;
    nufft_method_cuda method = nufft_method_cuda::DEFAULT;
    std::shared_ptr<hipStream_t> pstream;
    int32_t device_idx = 0;
    // tiling knobs of this library (0 = automatic)
    std::array<int32_t, 3> tile{0, 0, 0};
    int32_t coil_group = 0, max_batch = 0, max_subproblem = 0;

    int32_t get_sign() const { return static_cast<int>(sign); }
};

template <class D, class TT, size_t DIM, nufft_type NT> struct nufft_plan;

template <class TT, size_t DIM, nufft_type NT> struct nufft_plan<cuda_t, TT, DIM, NT> {
    static_assert(std::is_same_v<TT, f32_t> || std::is_same_v<TT, f64_t>, "TT must be f32_t or f64_t");

    static uptr<nufft_plan> make(const nufft_opts<cuda_t, TT, DIM>& opts) {
        if (NT == nufft_type::TYPE_3) throw std::runtime_error("nufft type 3 is not supported");
        uptr<nufft_plan> ret(new nufft_plan());
        ret->_opts = opts;
        hasty_nufft_opts o{};
        switch (opts.method) {
            case nufft_method_cuda::GM_NO_SORT: o.method = 0; break;
            case nufft_method_cuda::GM_SORT: o.method = 1; break;
            case nufft_method_cuda::SM: o.method = 2; break;
            case nufft_method_cuda::DEFAULT: o.method = 3; break;
            default: throw std::runtime_error("Invalid nufft_cuda_method");
        }
        switch (opts.upsamp) {
            case nufft_upsamp_cuda::UPSAMP_2_0: o.upsamp = 2.0; break;
            case nufft_upsamp_cuda::UPSAMP_1_25: o.upsamp = 1.25; break;
            case nufft_upsamp_cuda::DEFAULT: o.upsamp = 0.0; break;
            default: throw std::runtime_error("nufft_make_plan encountered invalid nufft_upsamp argument");
        }
        for (int a = 0; a < 3; ++a) o.tile[a] = opts.tile[size_t(a)];
        o.coil_group = opts.coil_group;
        o.max_batch = opts.max_batch;
        o.max_subproblem = opts.max_subproblem;
        void* stream = opts.pstream ? static_cast<void*>(*opts.pstream) : nullptr;
        detail::check_hip(hipSetDevice(opts.device_idx), "hipSetDevice");
        hasty_nufft_plan h = nullptr;
        detail::check(hasty_nufft_makeplan(static_cast<int>(NT), int(DIM), opts.nmodes.data(),
                                           opts.get_sign(), opts.ntransf, double(opts.tol),
                                           std::is_same_v<TT, f64_t> ? HASTY_F64 : HASTY_F32, &o,
                                           stream, &h));
        ret->_plan = h;
        return ret;
    }

    void setpts(const std::array<tensor<cuda_t, TT, 1>, DIM>& coords) {
        verify_coords(coords);
        const int64_t M = coords[0].template shape<0>();
        const void* c[3] = {nullptr, nullptr, nullptr};
        for (size_t a = 0; a < DIM; ++a) c[a] = coords[a].data_ptr();
        detail::check(hasty_nufft_setpts(_plan, M, c[0], c[1], c[2]));
        _M = M;
    }

    // type 1: [ntransf, M] -> [ntransf, N...]
    template <nufft_type U = NT, std::enable_if_t<U == nufft_type::TYPE_1, int> = 0>
    void execute(const tensor<cuda_t, complex_t<TT>, 2>& input,
                 tensor<cuda_t, complex_t<TT>, DIM + 1>& output) const {
        if (input.template shape<0>() != _opts.ntransf)
            throw std::runtime_error("input first dimension must match ntransf");
        check_modes(output.shape_arr());
        if (input.template shape<1>() != _M) throw std::runtime_error("input length must match setpts");
        detail::check(hasty_nufft_execute(_plan, input.unconsted_data_ptr(), output.mutable_data_ptr()));
    }

    // type 2: [ntransf, N...] -> [ntransf, M]
    template <nufft_type U = NT, std::enable_if_t<U == nufft_type::TYPE_2, int> = 0>
    void execute(const tensor<cuda_t, complex_t<TT>, DIM + 1>& input,
                 tensor<cuda_t, complex_t<TT>, 2>& output) const {
        check_modes(input.shape_arr());
        if (output.template shape<0>() != _opts.ntransf || output.template shape<1>() != _M)
            throw std::runtime_error("output must be [ntransf, M]");
        detail::check(hasty_nufft_execute(_plan, output.mutable_data_ptr(), input.unconsted_data_ptr()));
    }

    const std::array<int64_t, DIM>& nmodes() const { return _opts.nmodes; }
    int32_t ntransf() const { return _opts.ntransf; }
    int64_t npoints() const { return _M; }
    hasty_nufft_plan handle() const { return _plan; }
    TT tol() const { return _opts.tol; }
    nufft_opts<cuda_t, TT, DIM>& pa_opts() { return _opts; }
    const nufft_opts<cuda_t, TT, DIM>& pa_opts() const { return _opts; }

    void free() {
        if (_plan != nullptr) {
            const int st = hasty_nufft_destroy(_plan);
            _plan = nullptr;
            detail::check(st);
        }
    }
    ~nufft_plan() {
        if (_plan != nullptr) (void)hasty_nufft_destroy(_plan);
    }
    nufft_plan(const nufft_plan&) = delete;
    nufft_plan& operator=(const nufft_plan&) = delete;

private:
    nufft_plan() = default;
    template <size_t R> void check_modes(const std::array<int64_t, R>& s) const {
        if (s[0] != _opts.ntransf) throw std::runtime_error("mode array first dimension must match ntransf");
        for (size_t a = 0; a < DIM; ++a)
            if (s[a + 1] != _opts.nmodes[a]) throw std::runtime_error("mode array shape must match nmodes");
    }
    nufft_opts<cuda_t, TT, DIM> _opts{};
    hasty_nufft_plan _plan = nullptr;
    int64_t _M = 0;
};

// nufft.cppm:316-382: the plain per-coil path (one plan of ntransf 1, one
// host -> device -> host round trip per transform), kept for API parity.
template <class TT, class TTC, size_t DIM>
tensor<cpu_t, TT, DIM + 1> nufft_backward_cpu_over_cuda(const std::array<int64_t, DIM>& dimshape,
                                                        const tensor<cpu_t, TT, 2>& input,
                                                        const std::array<tensor<cuda_t, TTC, 1>, DIM>& coords) {
    static_assert(std::is_same_v<TT, complex_t<TTC>>, "input must be complex of the coordinate type");
    nufft_opts<cuda_t, TTC, DIM> options;
    options.nmodes = dimshape;
    options.ntransf = 1;
    options.tol = std::is_same_v<TTC, f32_t> ? TTC(1e-6) : TTC(1e-14);
    options.sign = nufft_sign::DEFAULT_TYPE_1;
    const int64_t ntransf = input.template shape<0>();
    const int64_t M = input.template shape<1>();
    std::array<int64_t, DIM + 1> full{};
    full[0] = ntransf;
    for (size_t a = 0; a < DIM; ++a) full[a + 1] = dimshape[a];
    tensor<cpu_t, TT, DIM + 1> output(full);
    auto plan = nufft_plan<cuda_t, TTC, DIM, nufft_type::BACKWARD>::make(options);
    plan->setpts(coords);
    std::array<int64_t, DIM + 1> one = full;
    one[0] = 1;
    auto din = tensor<cuda_t, TT, 2>::empty({1, M});
    auto dout = tensor<cuda_t, TT, DIM + 1>::empty(one);
    const int64_t per = dout.numel();
    for (int64_t i = 0; i < ntransf; ++i) {
        detail::check_hip(hipMemcpy(din.mutable_data_ptr(), input.data_ptr() + i * M, size_t(M) * sizeof(TT),
                                    hipMemcpyHostToDevice), "hipMemcpy H2D");
        plan->execute(din, dout);
        detail::check_hip(hipMemcpy(output.mutable_data_ptr() + i * per, dout.data_ptr(),
                                    size_t(per) * sizeof(TT), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    return output;
}

// ---- Toeplitz (nufft.cppm:386-479) ---------------------------------------------
// kernel [2N...] <- fftn(psf embedded at m + N - 1) / N^d; the PSF of weights
// nudata [1, M] is taken with an FP64 type-1 NUFFT (the reference uses the
// plan's own precision).
template <class TT, size_t DIM>
void toeplitz_kernel(const std::array<tensor<cuda_t, real_t<TT>, 1>, DIM>& coords,
                     tensor<cuda_t, TT, DIM>& kernel, tensor<cuda_t, TT, 2>& nudata) {
    detail::require_c64(kernel);
    verify_coords(coords);
    const int64_t M = coords[0].template shape<0>();
    if (nudata.template shape<1>() != M)
        throw std::runtime_error("shape<1>() of nudata didn't match length of coord vectors");
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) n[a] = kernel.shape_arr()[a] / 2;
    const void* c[3] = {nullptr, nullptr, nullptr};
    for (size_t a = 0; a < DIM; ++a) c[a] = coords[a].data_ptr();
    detail::check(hasty_toeplitz_kernel(int(DIM), n.data(), M, c[0], c[1], c[2], nudata.data_ptr(),
                                        kernel.mutable_data_ptr(), 0.0, nullptr));
}

// inout[b] <- crop_{N-1}(ifftn(fftn_{2N}(inout[b]) * kernel))
template <class TT, size_t DIM>
void toeplitz_multiply(tensor<cuda_t, TT, DIM + 1>& inout, const tensor<cuda_t, TT, DIM>& kernel) {
    detail::require_c64(kernel);
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) {
        n[a] = inout.shape_arr()[a + 1];
        if (kernel.shape_arr()[a] != 2 * n[a]) throw std::runtime_error("kernel must have shape 2N");
    }
    detail::check(hasty_toeplitz_multiply(int(DIM), n.data(), inout.template shape<0>(),
                                          inout.mutable_data_ptr(), kernel.data_ptr(), nullptr));
}

namespace fft {
// toeplitz.cppm:47-56
enum class ToeplitzMultType { NONE = 0, MULT = 1, MULT_CONJ = 2 };
enum class ToeplitzAccumulateType { NONE = 0, ACCUMULATE = 1 };

// toeplitz.cppm:510-609: the kernel stays in natural frequency order here,
// so the transform only validates the shape.
template <size_t DIM>
void transform_toeplitz_kernel(tensor<cuda_t, c64_t, DIM>& kernel, bool clear_vkfft_plan = false) {
    (void)clear_vkfft_plan;
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) n[a] = kernel.shape_arr()[a] / 2;
    detail::check(hasty_transform_toeplitz_kernel(int(DIM), n.data(), kernel.mutable_data_ptr(), nullptr));
}

// toeplitz.cppm:58-93 (3-D input [B, nx, ny] for DIM 2, [B, ...] for DIM 3)
template <size_t R>
void toeplitz_multiplication(const tensor<cuda_t, c64_t, R>& input, tensor<cuda_t, c64_t, R>& output,
                             const tensor<cuda_t, c64_t, R - 1>& kernel,
                             optrefw<tensor<cuda_t, c64_t, R - 1>> scratch,
                             optcrefw<tensor<cuda_t, c64_t, R - 1>> mult1,
                             optcrefw<tensor<cuda_t, c64_t, R - 1>> mult2,
                             ToeplitzMultType input_output_mult_type, ToeplitzMultType input_mult1_type,
                             ToeplitzMultType output_mult1_type, ToeplitzMultType input_mult2_type,
                             ToeplitzMultType output_mult2_type, ToeplitzAccumulateType accumulate_type) {
    constexpr size_t DIM = R - 1;
    static_assert(DIM == 2 || DIM == 3, "toeplitz_multiplication supports 2-D and 3-D");
    if (input.shape_arr() != output.shape_arr()) throw std::runtime_error("input and output shapes differ");
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) {
        n[a] = input.shape_arr()[a + 1];
        if (kernel.shape_arr()[a] != 2 * n[a]) throw std::runtime_error("kernel must have shape 2N");
        if (mult1 && mult1->get().shape_arr()[a] != n[a]) throw std::runtime_error("mult1 must have shape N");
        if (mult2 && mult2->get().shape_arr()[a] != n[a]) throw std::runtime_error("mult2 must have shape N");
    }
    const int32_t types[5] = {int32_t(input_output_mult_type), int32_t(input_mult1_type),
                              int32_t(output_mult1_type), int32_t(input_mult2_type),
                              int32_t(output_mult2_type)};
    void* scr = nullptr;
    int64_t scr_bytes = 0;
    if (scratch) {
        scr = scratch->get().mutable_data_ptr();
        scr_bytes = scratch->get().numel() * int64_t(sizeof(c64_t));
    }
    detail::check(hasty_toeplitz_multiplication(
        int(DIM), n.data(), input.template shape<0>(), input.data_ptr(), output.mutable_data_ptr(),
        kernel.data_ptr(), scr, scr_bytes, mult1 ? mult1->get().data_ptr() : nullptr,
        mult2 ? mult2->get().data_ptr() : nullptr, types, int(accumulate_type), nullptr));
}
}  // namespace fft

// ---- SENSE normal operator (normal.cppm:43-93) -----------------------------------
// The reference holds the kernel and maps as cache_tensors (and uploads the
// kernel on every apply); here both stay resident on the device.
template <class D, class TT, size_t DIM> class NORMAL_T_T1_OP;

template <size_t DIM> class NORMAL_T_T1_OP<cuda_t, c64_t, DIM> {
public:
    static constexpr const char* class_name = "NORMAL_T_T1_OP";
    using device_type_t = cuda_t;
    using input_tensor_type_t = c64_t;
    using output_tensor_type_t = c64_t;
    static constexpr size_t input_rank_t() { return DIM; }
    static constexpr size_t output_rank_t() { return DIM; }

    // the reference's constructor (normal.cppm:55-60): the kernel and maps as
    // cache_tensors, taken to the current device once
    NORMAL_T_T1_OP(cache_tensor<c64_t, DIM>&& kernel, cache_tensor<c64_t, DIM + 1>&& stacked_diags,
                   int coil_batch = 0)
        : NORMAL_T_T1_OP(kernel.template get<cuda_t>(detail::current_device()),
                         stacked_diags.template get<cuda_t>(detail::current_device()), coil_batch) {}

    NORMAL_T_T1_OP(tensor<cuda_t, c64_t, DIM>&& kernel, tensor<cuda_t, c64_t, DIM + 1>&& stacked_diags,
                   int coil_batch = 0)
        : _kernel(std::move(kernel)), _smaps(std::move(stacked_diags)) {
        for (size_t a = 0; a < DIM; ++a) {
            _n[a] = _smaps.shape_arr()[a + 1];
            if (_kernel.shape_arr()[a] != 2 * _n[a]) throw std::runtime_error("kernel must have shape 2N");
        }
        hasty_normal_op h = nullptr;
        detail::check(hasty_normal_create(int(DIM), _n.data(), int(_smaps.template shape<0>()),
                                          _smaps.data_ptr(), _kernel.data_ptr(), coil_batch, nullptr, &h));
        _op = std::shared_ptr<hasty_normal_op_s>(h, [](hasty_normal_op q) { (void)hasty_normal_destroy(q); });
    }

    tensor<cuda_t, c64_t, DIM> operator()(const tensor<cuda_t, c64_t, DIM>& x) const {
        auto out = tensor<cuda_t, c64_t, DIM>::empty(_n);
        apply(x, out);
        return out;
    }
    void apply(const tensor<cuda_t, c64_t, DIM>& x, tensor<cuda_t, c64_t, DIM>& out) const {
        if (x.shape_arr() != _n || out.shape_arr() != _n) throw std::runtime_error("x must have shape N");
        detail::check(hasty_normal_apply(_op.get(), x.data_ptr(), out.mutable_data_ptr()));
    }

private:
    tensor<cuda_t, c64_t, DIM> _kernel;
    tensor<cuda_t, c64_t, DIM + 1> _smaps;
    std::array<int64_t, DIM> _n{};
    std::shared_ptr<hasty_normal_op_s> _op;
};

// ---- multi-kernel / whitened normal operators (normal.cppm:113-735) -------------
// kind 1 (image -> image, coil sum fused):
//   out = sum_c conj(S'_c) . sum_l conj(R_l) . T_l(R_l . S_c . x)
// kind 2 (stacked images -> stacked images):
//   out_i = sum_j W_ij conj(X) . sum_l conj(R_l) . T_l(R_l . X . Psi_j)
// R (kerneldiags) may be empty (all ones); W is empty for the unweighted ops.
namespace detail {
template <size_t DIM> class normal_family {
protected:
    using K_T = tensor<cuda_t, c64_t, DIM + 1>;
    normal_family(K_T&& kernels, K_T&& kerneldiags)
        : _kernels(std::move(kernels)), _kerneldiags(std::move(kerneldiags)) {
        for (size_t a = 0; a < DIM; ++a) {
            if (_kernels.shape_arr()[a + 1] % 2) throw std::runtime_error("kernels must have extents 2N");
            _n[a] = _kernels.shape_arr()[a + 1] / 2;
        }
        if (_kerneldiags.data_ptr())
            for (size_t a = 0; a < DIM; ++a)
                if (_kerneldiags.shape_arr()[a + 1] != _n[a])
                    throw std::runtime_error("kerneldiags must have shape [L, N...]");
    }
    // the reference checks the counts when the operator is applied (normal.cppm:222-224)
    void check_counts() const {
        if (_kerneldiags.data_ptr() && _kerneldiags.template shape<0>() != _kernels.template shape<0>())
            throw std::runtime_error("Number of kernels and kernel diagonals must match.");
    }
    int nkernels() const { return int(_kernels.template shape<0>()); }
    void adopt(hasty_normal_op h) {
        _op = std::shared_ptr<hasty_normal_op_s>(h, [](hasty_normal_op q) { (void)hasty_normal_destroy(q); });
    }
    K_T _kernels, _kerneldiags;
    std::array<int64_t, DIM> _n{};
    std::shared_ptr<hasty_normal_op_s> _op;
};
}  // namespace detail

template <class D, class TT, size_t DIM> class NORMAL_IDT_T1_OP;
template <class D, class TT, size_t DIM> class NORMAL_IDTW_T1_OP;
template <class D, class TT, size_t DIM> class NORMAL_T_T2_OP;
template <class D, class TT, size_t DIM> class NORMAL_IDT_T2_OP;
template <class D, class TT, size_t DIM> class NORMAL_IDTW_T2_OP;

namespace detail {
template <size_t DIM> class normal_t1 : public normal_family<DIM> {
public:
    using device_type_t = cuda_t;
    using input_tensor_type_t = c64_t;
    using output_tensor_type_t = c64_t;
    static constexpr size_t input_rank_t() { return DIM; }
    static constexpr size_t output_rank_t() { return DIM; }

    tensor<cuda_t, c64_t, DIM> operator()(const tensor<cuda_t, c64_t, DIM>& x) const {
        auto out = tensor<cuda_t, c64_t, DIM>::empty(this->_n);
        apply(x, out);
        return out;
    }
    void apply(const tensor<cuda_t, c64_t, DIM>& x, tensor<cuda_t, c64_t, DIM>& out) const {
        this->check_counts();
        if (x.shape_arr() != this->_n || out.shape_arr() != this->_n) throw std::runtime_error("x must have shape N");
        detail::check(hasty_normal_apply(this->_op.get(), x.data_ptr(), out.mutable_data_ptr()));
    }

protected:
    normal_t1(tensor<cuda_t, c64_t, DIM + 1>&& kernels, tensor<cuda_t, c64_t, DIM + 1>&& kerneldiags,
              tensor<cuda_t, c64_t, DIM + 1>&& stacked_diags, tensor<cuda_t, c64_t, 2>&& weights, int coil_batch)
        : normal_family<DIM>(std::move(kernels), std::move(kerneldiags)), _smaps(std::move(stacked_diags)),
          _weights(std::move(weights)) {
        const int64_t C = _smaps.template shape<0>();
        for (size_t a = 0; a < DIM; ++a)
            if (_smaps.shape_arr()[a + 1] != this->_n[a]) throw std::runtime_error("stacked_diags must be [C, N...]");
        if (_weights.data_ptr() && (_weights.template shape<0>() != C || _weights.template shape<1>() != C))
            throw std::runtime_error("weights must be [C, C]");
        if (this->_kerneldiags.data_ptr() && this->_kerneldiags.template shape<0>() != this->nkernels()) return;
        hasty_normal_op h = nullptr;
        detail::check(hasty_normal_idt_create(int(DIM), this->_n.data(), int(C), _smaps.data_ptr(), this->nkernels(),
                                              this->_kernels.data_ptr(), this->_kerneldiags.data_ptr(),
                                              _weights.data_ptr(), coil_batch, nullptr, &h));
        this->adopt(h);
    }
    tensor<cuda_t, c64_t, DIM + 1> _smaps;
    tensor<cuda_t, c64_t, 2> _weights;
};

template <size_t DIM> class normal_t2 : public normal_family<DIM> {
public:
    using device_type_t = cuda_t;
    using input_tensor_type_t = c64_t;
    using output_tensor_type_t = c64_t;
    static constexpr size_t input_rank_t() { return DIM + 1; }
    static constexpr size_t output_rank_t() { return DIM + 1; }

    tensor<cuda_t, c64_t, DIM + 1> operator()(const tensor<cuda_t, c64_t, DIM + 1>& psi) {
        auto out = tensor<cuda_t, c64_t, DIM + 1>::empty(psi.shape_arr());
        apply(psi, out);
        return out;
    }
    void apply(const tensor<cuda_t, c64_t, DIM + 1>& psi, tensor<cuda_t, c64_t, DIM + 1>& out) {
        this->check_counts();
        for (size_t a = 0; a < DIM; ++a)
            if (psi.shape_arr()[a + 1] != this->_n[a]) throw std::runtime_error("stacked diagonals must be [B, N...]");
        if (out.shape_arr() != psi.shape_arr()) throw std::runtime_error("out must match the stacked diagonals");
        const int64_t B = psi.template shape<0>();
        if (_weights.data_ptr() && _weights.template shape<0>() != B)
            throw std::runtime_error("the number of stacked diagonals must match the weights");
        if (B != _B) {
            hasty_normal_op h = nullptr;
            detail::check(hasty_normal_t2_create(int(DIM), this->_n.data(), int(B), _x.data_ptr(), this->nkernels(),
                                                 this->_kernels.data_ptr(), this->_kerneldiags.data_ptr(),
                                                 _weights.data_ptr(), _stack_batch, nullptr, &h));
            this->adopt(h);
            _B = B;
        }
        detail::check(hasty_normal_apply(this->_op.get(), psi.data_ptr(), out.mutable_data_ptr()));
    }

protected:
    normal_t2(tensor<cuda_t, c64_t, DIM + 1>&& kernels, tensor<cuda_t, c64_t, DIM + 1>&& kerneldiags,
              tensor<cuda_t, c64_t, DIM>&& x, tensor<cuda_t, c64_t, 2>&& weights, int stack_batch)
        : normal_family<DIM>(std::move(kernels), std::move(kerneldiags)), _x(std::move(x)),
          _weights(std::move(weights)), _stack_batch(stack_batch) {
        if (_x.shape_arr() != this->_n) throw std::runtime_error("x must have shape N");
        if (_weights.data_ptr() && _weights.template shape<0>() != _weights.template shape<1>())
            throw std::runtime_error("weights must be square");
    }
    tensor<cuda_t, c64_t, DIM> _x;
    tensor<cuda_t, c64_t, 2> _weights;
    int _stack_batch = 0;
    int64_t _B = -1;
};
}  // namespace detail

// normal.cppm:185-272
template <size_t DIM> class NORMAL_IDT_T1_OP<cuda_t, c64_t, DIM> : public detail::normal_t1<DIM> {
public:
    static constexpr const char* class_name = "NORMAL_IDT_T1_OP";
    NORMAL_IDT_T1_OP(tensor<cuda_t, c64_t, DIM + 1>&& kernels, tensor<cuda_t, c64_t, DIM + 1>&& kerneldiags,
                     tensor<cuda_t, c64_t, DIM + 1>&& stacked_diags, bool store_module = false, int coil_batch = 0)
        : detail::normal_t1<DIM>(std::move(kernels), std::move(kerneldiags), std::move(stacked_diags),
                                 tensor<cuda_t, c64_t, 2>(), coil_batch) {
        (void)store_module;
    }
    // the reference's constructor (normal.cppm:206-210)
    NORMAL_IDT_T1_OP(cache_tensor<c64_t, DIM + 1>&& kernels, cache_tensor<c64_t, DIM + 1>&& kerneldiags,
                     cache_tensor<c64_t, DIM + 1>&& stacked_diags, bool store_module = false, int coil_batch = 0)
        : NORMAL_IDT_T1_OP(detail::on_device(kernels), detail::on_device(kerneldiags), detail::on_device(stacked_diags),
                           store_module, coil_batch) {}
};

// normal.cppm:405-564; the output-side maps S'_c = sum_i conj(W_ic) S_i are formed once
template <size_t DIM> class NORMAL_IDTW_T1_OP<cuda_t, c64_t, DIM> : public detail::normal_t1<DIM> {
public:
    static constexpr const char* class_name = "NORMAL_IDTW_T1_OP";
    NORMAL_IDTW_T1_OP(tensor<cuda_t, c64_t, DIM + 1>&& kernels, tensor<cuda_t, c64_t, DIM + 1>&& kerneldiags,
                      tensor<cuda_t, c64_t, DIM + 1>&& stacked_diags, tensor<cuda_t, c64_t, 2>&& weights,
                      bool store_module = false, int coil_batch = 0)
        : detail::normal_t1<DIM>(std::move(kernels), std::move(kerneldiags), std::move(stacked_diags),
                                 std::move(weights), coil_batch) {
        (void)store_module;
    }
    // the reference's constructor (normal.cppm:424-430)
    NORMAL_IDTW_T1_OP(cache_tensor<c64_t, DIM + 1>&& kernels, cache_tensor<c64_t, DIM + 1>&& kerneldiags,
                      cache_tensor<c64_t, DIM + 1>&& stacked_diags, cache_tensor<c64_t, 2>&& weights,
                      bool store_module = false, int coil_batch = 0)
        : NORMAL_IDTW_T1_OP(detail::on_device(kernels), detail::on_device(kerneldiags),
                            detail::on_device(stacked_diags), detail::on_device(weights), store_module, coil_batch) {}
};

// normal.cppm:113-160
template <size_t DIM> class NORMAL_T_T2_OP<cuda_t, c64_t, DIM> : public detail::normal_t2<DIM> {
public:
    static constexpr const char* class_name = "NORMAL_T_T2_OP";
    NORMAL_T_T2_OP(tensor<cuda_t, c64_t, DIM>&& kernel, tensor<cuda_t, c64_t, DIM>&& x, int stack_batch = 0)
        : detail::normal_t2<DIM>(as_stack(std::move(kernel)), tensor<cuda_t, c64_t, DIM + 1>(), std::move(x),
                                 tensor<cuda_t, c64_t, 2>(), stack_batch) {}
    // the reference's constructor (normal.cppm:124-126)
    NORMAL_T_T2_OP(cache_tensor<c64_t, DIM>&& kernel, cache_tensor<c64_t, DIM>&& x, int stack_batch = 0)
        : NORMAL_T_T2_OP(detail::on_device(kernel), detail::on_device(x), stack_batch) {}

private:
    static tensor<cuda_t, c64_t, DIM + 1> as_stack(tensor<cuda_t, c64_t, DIM>&& k) {
        std::array<int64_t, DIM + 1> s{};
        s[0] = 1;
        for (size_t a = 0; a < DIM; ++a) s[a + 1] = k.shape_arr()[a];
        auto t = tensor<cuda_t, c64_t, DIM + 1>::empty(s);
        detail::check_hip(hipMemcpy(t.mutable_data_ptr(), k.data_ptr(), size_t(k.numel()) * sizeof(c64_t),
                                    hipMemcpyDeviceToDevice), "hipMemcpy D2D");
        return t;
    }
};

// normal.cppm:292-375
template <size_t DIM> class NORMAL_IDT_T2_OP<cuda_t, c64_t, DIM> : public detail::normal_t2<DIM> {
public:
    static constexpr const char* class_name = "NORMAL_IDT_T2_OP";
    NORMAL_IDT_T2_OP(tensor<cuda_t, c64_t, DIM + 1>&& kernels, tensor<cuda_t, c64_t, DIM + 1>&& kerneldiags,
                     tensor<cuda_t, c64_t, DIM>&& x, int stack_batch = 0)
        : detail::normal_t2<DIM>(std::move(kernels), std::move(kerneldiags), std::move(x),
                                 tensor<cuda_t, c64_t, 2>(), stack_batch) {}
    // the reference's constructor (normal.cppm:311-314)
    NORMAL_IDT_T2_OP(cache_tensor<c64_t, DIM + 1>&& kernels, cache_tensor<c64_t, DIM + 1>&& kerneldiags,
                     cache_tensor<c64_t, DIM>&& x, int stack_batch = 0)
        : NORMAL_IDT_T2_OP(detail::on_device(kernels), detail::on_device(kerneldiags), detail::on_device(x),
                           stack_batch) {}
};

// normal.cppm:586-735
template <size_t DIM> class NORMAL_IDTW_T2_OP<cuda_t, c64_t, DIM> : public detail::normal_t2<DIM> {
public:
    static constexpr const char* class_name = "NORMAL_IDTW_T2_OP";
    NORMAL_IDTW_T2_OP(tensor<cuda_t, c64_t, DIM + 1>&& kernels, tensor<cuda_t, c64_t, DIM + 1>&& kerneldiags,
                      tensor<cuda_t, c64_t, DIM>&& x, tensor<cuda_t, c64_t, 2>&& weights, bool store_module = false,
                      int stack_batch = 0)
        : detail::normal_t2<DIM>(std::move(kernels), std::move(kerneldiags), std::move(x), std::move(weights),
                                 stack_batch) {
        (void)store_module;
    }
    // the reference's constructor (normal.cppm:602-607)
    NORMAL_IDTW_T2_OP(cache_tensor<c64_t, DIM + 1>&& kernels, cache_tensor<c64_t, DIM + 1>&& kerneldiags,
                      cache_tensor<c64_t, DIM>&& x, cache_tensor<c64_t, 2>&& weights, bool store_module = false,
                      int stack_batch = 0)
        : NORMAL_IDTW_T2_OP(detail::on_device(kernels), detail::on_device(kerneldiags), detail::on_device(x),
                            detail::on_device(weights), store_module, stack_batch) {}
};

// ---- time-segmented SENSE (forward.cppm:36-165, adjoint.cppm:33-171) -----------
// All coils of a chunk go through one plan of ntransf = coil_batch (0 = all);
// tolerance 1e-6 in single precision as in the reference, 1e-13 in double
// precision (the reference asks for 1e-14; FP64 plans here stop at 1e-13).  The
// reference takes host-cached tensors and a thread pool and returns host
// tensors; here everything stays on the device.
namespace detail {
template <class TTC, size_t DIM, nufft_type NT>
uptr<nufft_plan<cuda_t, TTC, DIM, NT>> sense_plan(const std::array<tensor<cuda_t, TTC, 1>, DIM>& coords,
                                                   const std::array<int64_t, DIM>& n, int64_t C, int coil_batch) {
    nufft_opts<cuda_t, TTC, DIM> o;
    o.nmodes = n;
    o.ntransf = int32_t(coil_batch > 0 ? std::min<int64_t>(coil_batch, C) : C);
    o.tol = std::is_same_v<TTC, f32_t> ? TTC(1e-6) : TTC(1e-13);
    o.sign = NT == nufft_type::TYPE_1 ? nufft_sign::DEFAULT_TYPE_1 : nufft_sign::DEFAULT_TYPE_2;
    o.device_idx = int32_t(current_device());
    auto plan = nufft_plan<cuda_t, TTC, DIM, NT>::make(o);
    plan->setpts(coords);
    return plan;
}
template <class TT, size_t DIM>
void check_offres(const tensor<cuda_t, TT, DIM + 1>& smaps, const tensor<cuda_t, TT, 2>& bil,
                  const tensor<cuda_t, TT, DIM + 1>& cjl, int64_t M) {
    if (bil.template shape<0>() != cjl.template shape<0>()) throw std::runtime_error("bil and cjl must have L rows");
    if (bil.template shape<1>() != M) throw std::runtime_error("bil must be [L, M]");
    for (size_t a = 0; a < DIM; ++a)
        if (cjl.shape_arr()[a + 1] != smaps.shape_arr()[a + 1]) throw std::runtime_error("cjl must be [L, N...]");
}
}  // namespace detail

template <class TT, class TTC, size_t DIM>
tensor<cuda_t, TT, 2> sense_forward_offresonance(const std::array<tensor<cuda_t, TTC, 1>, DIM>& coords,
                                                 const tensor<cuda_t, TT, DIM>& image,
                                                 const tensor<cuda_t, TT, DIM + 1>& smaps,
                                                 const tensor<cuda_t, TT, 2>& bil,
                                                 const tensor<cuda_t, TT, DIM + 1>& cjl, int coil_batch = 0) {
    static_assert(std::is_same_v<TT, complex_t<TTC>>, "data must be the plan's complex type");
    const int64_t C = smaps.template shape<0>(), M = coords[0].template shape<0>();
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) n[a] = smaps.shape_arr()[a + 1];
    if (image.shape_arr() != n) throw std::runtime_error("image must have the maps' spatial shape");
    detail::check_offres<TT, DIM>(smaps, bil, cjl, M);
    auto plan = detail::sense_plan<TTC, DIM, nufft_type::TYPE_2>(coords, n, C, coil_batch);
    auto out = tensor<cuda_t, TT, 2>::empty({C, M});
    detail::check(hasty_sense_forward(plan->handle(), int(C), int(bil.template shape<0>()), image.data_ptr(),
                                      smaps.data_ptr(), bil.data_ptr(), cjl.data_ptr(), out.mutable_data_ptr()));
    return out;
}

template <class TT, class TTC, size_t DIM>
tensor<cuda_t, TT, DIM + 1> sense_adjoint_offresonance(const std::array<tensor<cuda_t, TTC, 1>, DIM>& coords,
                                                       const tensor<cuda_t, TT, 2>& kdata,
                                                       const tensor<cuda_t, TT, DIM + 1>& smaps,
                                                       const tensor<cuda_t, TT, 2>& bil,
                                                       const tensor<cuda_t, TT, DIM + 1>& cjl, int coil_batch = 0) {
    static_assert(std::is_same_v<TT, complex_t<TTC>>, "data must be the plan's complex type");
    const int64_t C = smaps.template shape<0>(), M = coords[0].template shape<0>();
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) n[a] = smaps.shape_arr()[a + 1];
    if (kdata.template shape<0>() != C || kdata.template shape<1>() != M)
        throw std::runtime_error("kdata must be [C, M]");
    detail::check_offres<TT, DIM>(smaps, bil, cjl, M);
    auto plan = detail::sense_plan<TTC, DIM, nufft_type::TYPE_1>(coords, n, C, coil_batch);
    auto out = tensor<cuda_t, TT, DIM + 1>::empty(smaps.shape_arr());
    detail::check(hasty_sense_adjoint(plan->handle(), int(C), int(bil.template shape<0>()), kdata.data_ptr(),
                                      smaps.data_ptr(), bil.data_ptr(), cjl.data_ptr(), out.mutable_data_ptr(), 0));
    return out;
}

// The reference's form (forward.cppm:35-46, adjoint.cppm:33-43): cache_tensor
// arguments and a host result.  The coils run batched on the current device
// (there is no thread pool); free_bil / free_cjl drop the device copies of
// bil / cjl afterwards, and the plan lives for the call either way.
template <class D, class TT, class TTC, size_t DIM>
tensor<cpu_t, TT, 2> sense_forward_offresonance(std::array<cache_tensor<TTC, 1>, DIM>& coords,
                                                cache_tensor<TT, DIM>& image, cache_tensor<TT, DIM + 1>& smaps,
                                                cache_tensor<TT, 2>& bil, cache_tensor<TT, DIM + 1>& cjl,
                                                bool free_bil = true, bool free_cjl = true,
                                                bool free_nufft_plan = true) {
    static_assert(std::is_same_v<D, cuda_t>, "the time-segmented operators run on the GPU");
    (void)free_nufft_plan;
    const device_idx d = detail::current_device();
    std::array<tensor<cuda_t, TTC, 1>, DIM> dc;
    for (size_t a = 0; a < DIM; ++a) dc[a] = coords[a].template get<cuda_t>(d);
    auto out = sense_forward_offresonance<TT, TTC, DIM>(dc, image.template get<cuda_t>(d),
                                                        smaps.template get<cuda_t>(d), bil.template get<cuda_t>(d),
                                                        cjl.template get<cuda_t>(d));
    if (free_bil) bil.free(d);
    if (free_cjl) cjl.free(d);
    tensor<cpu_t, TT, 2> h(out.shape_arr());
    h.vec() = out.to_host();
    return h;
}

template <class D, class TT, class TTC, size_t DIM>
tensor<cpu_t, TT, DIM + 1> sense_adjoint_offresonance(std::array<cache_tensor<TTC, 1>, DIM>& coords,
                                                      cache_tensor<TT, 2>& kdata, cache_tensor<TT, DIM + 1>& smaps,
                                                      cache_tensor<TT, 2>& bil, cache_tensor<TT, DIM + 1>& cjl,
                                                      bool free_bil = true, bool free_cjl = true,
                                                      bool free_nufft_plan = true) {
    static_assert(std::is_same_v<D, cuda_t>, "the time-segmented operators run on the GPU");
    (void)free_nufft_plan;
    const device_idx d = detail::current_device();
    std::array<tensor<cuda_t, TTC, 1>, DIM> dc;
    for (size_t a = 0; a < DIM; ++a) dc[a] = coords[a].template get<cuda_t>(d);
    auto out = sense_adjoint_offresonance<TT, TTC, DIM>(dc, kdata.template get<cuda_t>(d),
                                                        smaps.template get<cuda_t>(d), bil.template get<cuda_t>(d),
                                                        cjl.template get<cuda_t>(d));
    if (free_bil) bil.free(d);
    if (free_cjl) cjl.free(d);
    tensor<cpu_t, TT, DIM + 1> h(out.shape_arr());
    h.vec() = out.to_host();
    return h;
}

// The reference's thread-pool form (forward.cppm:35-46, adjoint.cppm:33-43).
// The coils are split into one contiguous chunk per pool worker.  A worker
// runs on the device named by the device_idx entry "device_idx" of its
// storage and pushes its chunk through one plan of ntransf = chunk size,
// kept in its storage as "nufft_plan" until free_nufft_plan clears it (a kept
// plan is reused while the chunk size matches, whatever the coordinates, as in
// the reference).  The chunk's maps go to the device per call; the image, bil,
// cjl and coordinates come from the cache_tensors' device copies, and
// free_bil / free_cjl drop those of bil / cjl on every device used.  Chunks
// run concurrently on distinct devices; workers that share a device share its
// null stream.
namespace detail {
template <class NP> struct pooled_plan {
    uptr<NP> plan;
    int64_t ntransf = 0;
};
template <class TTC, size_t DIM, nufft_type NT>
nufft_plan<cuda_t, TTC, DIM, NT>& pool_plan(threading::storage& store, std::array<cache_tensor<TTC, 1>, DIM>& coords,
                                             const std::array<int64_t, DIM>& n, int64_t nt, device_idx d) {
    using E = pooled_plan<nufft_plan<cuda_t, TTC, DIM, NT>>;
    auto e = store.get_ptr<E>("nufft_plan");
    if (!e || e->ntransf != nt) {
        store.clear("nufft_plan");
        std::array<tensor<cuda_t, TTC, 1>, DIM> dc;
        for (size_t a = 0; a < DIM; ++a) dc[a] = coords[a].template get<cuda_t>(d);
        e = std::make_shared<E>();
        e->plan = sense_plan<TTC, DIM, NT>(dc, n, nt, 0);
        e->ntransf = nt;
        store.add<E>("nufft_plan", e);
    }
    return *e->plan;
}
// runs body(c0, c1, device) for contiguous coil chunks [c0, c1) on the pool;
// returns the devices used
template <class F>
std::vector<device_idx> run_coil_chunks(threading::storage_thread_pool_interface& pool, int64_t C, F body) {
    const int64_t K = std::max<int64_t>(1, std::min<int64_t>(C, int64_t(pool.nthreads())));
    const int64_t per = (C + K - 1) / std::max<int64_t>(K, 1);
    std::mutex m;
    std::vector<device_idx> used;
    std::vector<std::future<void>> futs;
    try {
        for (int64_t c0 = 0; c0 < C; c0 += per) {
            const int64_t c1 = std::min(C, c0 + per);
            futs.push_back(pool.enqueue(
                [&, c0, c1](threading::storage& store) {
                    const device_idx d = store.get_ref_throw<device_idx>("device_idx");
                    check_hip(hipSetDevice(int(d)), "hipSetDevice");
                    {
                        std::lock_guard<std::mutex> g(m);
                        if (std::find(used.begin(), used.end(), d) == used.end()) used.push_back(d);
                    }
                    body(c0, c1, d, store);
                },
                {}));
        }
    } catch (...) {  // the queued chunks use this frame: let them finish first
        for (auto& f : futs) f.wait();
        throw;
    }
    std::exception_ptr first;
    for (auto& f : futs) {  // wait for every chunk before reporting the first failure
        try {
            f.get();
        } catch (...) {
            if (!first) first = std::current_exception();
        }
    }
    if (first) std::rethrow_exception(first);
    return used;
}
template <class TT, size_t R> void free_on(cache_tensor<TT, R>& t, const std::vector<device_idx>& used) {
    for (auto d : used) t.free(d);
}
}  // namespace detail

template <class D, class TT, class TTC, size_t DIM>
tensor<cpu_t, TT, 2> sense_forward_offresonance(std::array<cache_tensor<TTC, 1>, DIM>& coords,
                                                cache_tensor<TT, DIM>& image, cache_tensor<TT, DIM + 1>& smaps,
                                                cache_tensor<TT, 2>& bil, cache_tensor<TT, DIM + 1>& cjl,
                                                threading::storage_thread_pool_interface& thread_pool,
                                                bool free_bil = true, bool free_cjl = true,
                                                bool free_nufft_plan = true) {
    static_assert(std::is_same_v<D, cuda_t>, "the time-segmented operators run on the GPU");
    static_assert(std::is_same_v<TT, complex_t<TTC>>, "data must be the plan's complex type");
    const auto sshape = smaps.shape();
    const int64_t C = sshape[0], M = coords[0].template shape<0>();
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) n[a] = sshape[a + 1];
    if (image.shape() != n) throw std::runtime_error("image must have the maps' spatial shape");
    const int64_t I = std::accumulate(n.begin(), n.end(), int64_t(1), std::multiplies<int64_t>());
    const int64_t L = bil.template shape<0>();
    if (bil.template shape<1>() != M || cjl.template shape<0>() != L)
        throw std::runtime_error("bil must be [L, M] and cjl [L, N...]");
    tensor<cpu_t, TT, 2> out({C, M});
    auto hmaps = smaps.template get_ptr<cpu_t>();
    auto used = detail::run_coil_chunks(thread_pool, C, [&](int64_t c0, int64_t c1, device_idx d,
                                                            threading::storage& store) {
        const int64_t nc = c1 - c0;
        auto& plan = detail::pool_plan<TTC, DIM, nufft_type::TYPE_2>(store, coords, n, nc, d);
        std::array<int64_t, DIM + 1> cs{};
        cs[0] = nc;
        for (size_t a = 0; a < DIM; ++a) cs[a + 1] = n[a];
        auto maps = tensor<cuda_t, TT, DIM + 1>::empty(cs);
        detail::check_hip(hipMemcpy(maps.mutable_data_ptr(), hmaps->data_ptr() + c0 * I, size_t(nc * I) * sizeof(TT),
                                    hipMemcpyHostToDevice), "hipMemcpy H2D");
        auto y = tensor<cuda_t, TT, 2>::empty({nc, M});
        auto img = image.template get<cuda_t>(d);
        auto b = bil.template get<cuda_t>(d);
        auto cj = cjl.template get<cuda_t>(d);
        detail::check(hasty_sense_forward(plan.handle(), int(nc), int(L), img.data_ptr(), maps.data_ptr(),
                                          b.data_ptr(), cj.data_ptr(), y.mutable_data_ptr()));
        detail::check_hip(hipMemcpy(out.mutable_data_ptr() + c0 * M, y.data_ptr(), size_t(nc * M) * sizeof(TT),
                                    hipMemcpyDeviceToHost), "hipMemcpy D2H");
    });
    if (free_bil) detail::free_on(bil, used);
    if (free_cjl) detail::free_on(cjl, used);
    if (free_nufft_plan) thread_pool.clear_from_storages({"nufft_plan"});
    return out;
}

template <class D, class TT, class TTC, size_t DIM>
tensor<cpu_t, TT, DIM + 1> sense_adjoint_offresonance(std::array<cache_tensor<TTC, 1>, DIM>& coords,
                                                      cache_tensor<TT, 2>& kdata, cache_tensor<TT, DIM + 1>& smaps,
                                                      cache_tensor<TT, 2>& bil, cache_tensor<TT, DIM + 1>& cjl,
                                                      threading::storage_thread_pool_interface& thread_pool,
                                                      bool free_bil = true, bool free_cjl = true,
                                                      bool free_nufft_plan = true) {
    static_assert(std::is_same_v<D, cuda_t>, "the time-segmented operators run on the GPU");
    static_assert(std::is_same_v<TT, complex_t<TTC>>, "data must be the plan's complex type");
    const auto sshape = smaps.shape();
    const int64_t C = sshape[0], M = coords[0].template shape<0>();
    std::array<int64_t, DIM> n{};
    for (size_t a = 0; a < DIM; ++a) n[a] = sshape[a + 1];
    if (kdata.template shape<0>() != C || kdata.template shape<1>() != M)
        throw std::runtime_error("kdata must be [C, M]");
    const int64_t I = std::accumulate(n.begin(), n.end(), int64_t(1), std::multiplies<int64_t>());
    const int64_t L = bil.template shape<0>();
    if (bil.template shape<1>() != M || cjl.template shape<0>() != L)
        throw std::runtime_error("bil must be [L, M] and cjl [L, N...]");
    tensor<cpu_t, TT, DIM + 1> out(sshape);
    auto hmaps = smaps.template get_ptr<cpu_t>();
    auto hk = kdata.template get_ptr<cpu_t>();
    auto used = detail::run_coil_chunks(thread_pool, C, [&](int64_t c0, int64_t c1, device_idx d,
                                                            threading::storage& store) {
        const int64_t nc = c1 - c0;
        auto& plan = detail::pool_plan<TTC, DIM, nufft_type::TYPE_1>(store, coords, n, nc, d);
        std::array<int64_t, DIM + 1> cs{};
        cs[0] = nc;
        for (size_t a = 0; a < DIM; ++a) cs[a + 1] = n[a];
        auto maps = tensor<cuda_t, TT, DIM + 1>::empty(cs);
        auto k = tensor<cuda_t, TT, 2>::empty({nc, M});
        detail::check_hip(hipMemcpy(maps.mutable_data_ptr(), hmaps->data_ptr() + c0 * I, size_t(nc * I) * sizeof(TT),
                                    hipMemcpyHostToDevice), "hipMemcpy H2D");
        detail::check_hip(hipMemcpy(k.mutable_data_ptr(), hk->data_ptr() + c0 * M, size_t(nc * M) * sizeof(TT),
                                    hipMemcpyHostToDevice), "hipMemcpy H2D");
        auto x = tensor<cuda_t, TT, DIM + 1>::empty(cs);
        auto b = bil.template get<cuda_t>(d);
        auto cj = cjl.template get<cuda_t>(d);
        detail::check(hasty_sense_adjoint(plan.handle(), int(nc), int(L), k.data_ptr(), maps.data_ptr(),
                                          b.data_ptr(), cj.data_ptr(), x.mutable_data_ptr(), 0));
        detail::check_hip(hipMemcpy(out.mutable_data_ptr() + c0 * I, x.data_ptr(), size_t(nc * I) * sizeof(TT),
                                    hipMemcpyDeviceToHost), "hipMemcpy D2H");
    });
    if (free_bil) detail::free_on(bil, used);
    if (free_cjl) detail::free_on(cjl, used);
    if (free_nufft_plan) thread_pool.clear_from_storages({"nufft_plan"});
    return out;
}

// ---- CG (cg.cppm:122-184) and power iteration (alg.cppm:13-30) -----------------
// Returns cg(x, b) -> x.  The step arithmetic runs in the library's FP64
// reduction kernels with every scalar on the device; the residual is read
// on the host once per iteration, as the reference does (cg.cppm:176).
template <class Op> auto conjugate_gradient(sptr<Op> A, i32 max_iter = 0, double tol = 1e-6) {
    constexpr size_t R = Op::input_rank_t();
    using T = tensor<cuda_t, c64_t, R>;
    return [A = std::move(A), max_iter, tol](T& x, const T& b) -> T& {
        const int64_t n = x.numel();
        auto state = tensor<cuda_t, double, 1>::zeros({HASTY_CG_STATE_DOUBLES});
        T r = b.clone();
        {
            T Ax = (*A)(x);
            detail::check(hasty_axpy(-1.0, 0.0, Ax.data_ptr(), r.mutable_data_ptr(), n, nullptr));
        }
        T p = T::empty(x.shape_arr());
        detail::check(hasty_cg_init(r.mutable_data_ptr(), p.mutable_data_ptr(), n, tol,
                                    state.mutable_data_ptr(), nullptr));
        double resid = 0.0;
        int done = 0;
        detail::check(hasty_cg_status(state.data_ptr(), &resid, &done, nullptr));
        T Ap = T::empty(x.shape_arr());
        for (i32 iter = 0; iter < max_iter && !done; ++iter) {
            A->apply(p, Ap);
            detail::check(hasty_cg_step(x.mutable_data_ptr(), r.mutable_data_ptr(), p.mutable_data_ptr(),
                                        Ap.data_ptr(), n, state.mutable_data_ptr(), nullptr));
            detail::check(hasty_cg_status(state.data_ptr(), &resid, &done, nullptr));
        }
        return x;
    };
}

// Preconditioned CG with restart callbacks (cg.cppm:16-119).  Returns
// pcg(x, b, should_restart_callback, after_restart_callback) -> x:
//   r = b - A x, z = P r, p = z, rz = Re<r, z>; each outer iteration runs up
//   to max_inner_iter steps while sqrt(rz) >= tol: alpha = rz / Re<p, Ap>,
//   x += alpha p, r -= alpha Ap, z = P r, rz' = Re<r, z>, p = z when
//   should_restart_callback(x) (which also ends the inner loop) else
//   z + (rz'/rz) p; then after_restart_callback(x).
// FP64 scalars on the device (hasty_pcg_*); the residual is read on the host
// once per step, as in the reference.  The reference's requires clause names
// output_tensor_rank_t, which its operators do not define (cg.cppm:17-19);
// the same constraints are checked here on the members that exist.
template <class Op, class PrecondOp>
auto conjugate_gradient(sptr<Op> A, sptr<PrecondOp> P, i32 max_inner_iter, i32 max_outer_iter,
                        double tol = 1e-6) {
    static_assert(std::is_same_v<typename Op::device_type_t, typename PrecondOp::device_type_t>,
                  "A and P must run on the same device type");
    static_assert(std::is_same_v<typename Op::output_tensor_type_t, typename PrecondOp::input_tensor_type_t>,
                  "P must take A's output element type");
    static_assert(Op::output_rank_t() == PrecondOp::input_rank_t(), "P must take A's output rank");
    constexpr size_t R = Op::input_rank_t();
    using T = tensor<cuda_t, c64_t, R>;
    return [A = std::move(A), P = std::move(P), max_inner_iter, max_outer_iter, tol](
               T& x, const T& b, std::optional<std::function<bool(T&)>> should_restart_callback = std::nullopt,
               std::optional<std::function<void(T&)>> after_restart_callback = std::nullopt) -> T& {
        const int64_t n = x.numel();
        auto state = tensor<cuda_t, double, 1>::zeros({HASTY_CG_STATE_DOUBLES});
        T r = b.clone();
        {
            T Ax = (*A)(x);
            detail::check(hasty_axpy(-1.0, 0.0, Ax.data_ptr(), r.mutable_data_ptr(), n, nullptr));
        }
        T z = T::empty(x.shape_arr());
        T p = T::empty(x.shape_arr());
        T Ap = T::empty(x.shape_arr());
        P->apply(r, z);
        detail::check(hasty_pcg_init(r.data_ptr(), z.data_ptr(), p.mutable_data_ptr(), n, tol,
                                     state.mutable_data_ptr(), nullptr));
        double resid = 0.0;
        int done = 0;
        detail::check(hasty_cg_status(state.data_ptr(), &resid, &done, nullptr));
        for (i32 outer = 0; outer < max_outer_iter; ++outer) {
            for (i32 inner = 0; inner < max_inner_iter && !done; ++inner) {
                A->apply(p, Ap);
                detail::check(hasty_pcg_update(x.mutable_data_ptr(), r.mutable_data_ptr(), p.data_ptr(),
                                               Ap.data_ptr(), n, state.mutable_data_ptr(), nullptr));
                P->apply(r, z);
                const bool restart = should_restart_callback.has_value() && (*should_restart_callback)(x);
                detail::check(hasty_pcg_direction(r.data_ptr(), z.data_ptr(), p.mutable_data_ptr(), n,
                                                  restart ? 1 : 0, state.mutable_data_ptr(), nullptr));
                detail::check(hasty_cg_status(state.data_ptr(), &resid, &done, nullptr));
                if (restart) break;
            }
            if (after_restart_callback.has_value()) (*after_restart_callback)(x);
        }
        return x;
    };
}

template <class Op, class T>
std::tuple<float, T> max_eig(const Op& A, T&& x0, i32 max_iter = 10) {
    auto nrm = tensor<cuda_t, double, 1>::zeros({2});
    T x = std::move(x0);
    for (i32 i = 0; i < max_iter; ++i) {
        T y = A(x);
        detail::check(hasty_norm(y.data_ptr(), y.numel(), nrm.mutable_data_ptr(), nullptr));
        detail::check(hasty_scale_by_device(y.mutable_data_ptr(), y.numel(), nrm.data_ptr(), 1, nullptr));
        x = std::move(y);
    }
    const double lam = nrm.to_host()[0];
    return {float(lam), std::move(x)};
}

}  // namespace hasty

#endif  // HASTY_CORELIB_HPP
