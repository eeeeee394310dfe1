// GPU test of the C++ corelib mirror (include/hasty/corelib.hpp): the calls
// read like the reference's own (cpp/corelib/interface/test_interface.cpp),
// and results are checked against a direct double-precision NDFT written
// here from the definition.  Prints one line per check; exit 0 = all passed.
#include <hasty/corelib.hpp>

#include <cstdio>
#include <random>

using namespace hasty;

namespace {

int failures = 0;

void report(const char* what, double err, double tol) {
    const bool ok = err <= tol;
    std::printf("[cpp] %-46s rel_l2 = %.3e (tol %.1e) %s\n", what, err, tol, ok ? "ok" : "FAIL");
    std::fflush(stdout);
    if (!ok) ++failures;
}

template <class A, class B> double rel_l2(const std::vector<A>& a, const std::vector<B>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        const std::complex<double> x(a[i].real(), a[i].imag()), y(b[i].real(), b[i].imag());
        num += std::norm(x - y);
        den += std::norm(y);
    }
    return std::sqrt(num / den);
}

struct Problem {
    std::array<int64_t, 3> n{12, 16, 10};
    int64_t M = 3000;
    int C = 3;
    std::array<std::vector<float>, 3> k;
    std::vector<c64_t> x, S;
};

Problem make_problem() {
    Problem p;
    std::mt19937_64 g(7);
    std::uniform_real_distribution<float> u(-3.14159f, 3.14159f);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (auto& ka : p.k) {
        ka.resize(size_t(p.M));
        for (auto& v : ka) v = u(g);
    }
    const int64_t I = p.n[0] * p.n[1] * p.n[2];
    p.x.resize(size_t(I));
    for (auto& v : p.x) v = c64_t(nd(g), nd(g));
    p.S.resize(size_t(p.C * I));
    for (auto& v : p.S) v = c64_t(nd(g), nd(g));
    return p;
}

// P = a I on the device, with the operator members the PCG overload checks
struct ScaleOp {
    using device_type_t = cuda_t;
    using input_tensor_type_t = c64_t;
    using output_tensor_type_t = c64_t;
    static constexpr size_t input_rank_t() { return 3; }
    static constexpr size_t output_rank_t() { return 3; }
    double a;
    void apply(const tensor<cuda_t, c64_t, 3>& x, tensor<cuda_t, c64_t, 3>& out) const {
        detail::check_hip(hipMemset(out.mutable_data_ptr(), 0, size_t(out.numel()) * sizeof(c64_t)), "hipMemset");
        detail::check(hasty_axpy(a, 0.0, x.data_ptr(), out.mutable_data_ptr(), x.numel(), nullptr));
    }
};

// c_j = sum_n f[n] exp(sign i k_j . n), centred modes, axis a <-> k[a]
std::vector<std::complex<double>> ndft2(const Problem& p, const std::vector<c64_t>& f, int sign) {
    std::vector<std::complex<double>> c(size_t(p.M));
    for (int64_t j = 0; j < p.M; ++j) {
        std::complex<double> acc = 0;
        for (int64_t a = 0; a < p.n[0]; ++a) {
            const double pa = double(p.k[0][size_t(j)]) * double(a - p.n[0] / 2);
            for (int64_t b = 0; b < p.n[1]; ++b) {
                const double pb = pa + double(p.k[1][size_t(j)]) * double(b - p.n[1] / 2);
                for (int64_t q = 0; q < p.n[2]; ++q) {
                    const double ph = pb + double(p.k[2][size_t(j)]) * double(q - p.n[2] / 2);
                    const c64_t v = f[size_t((a * p.n[1] + b) * p.n[2] + q)];
                    acc += std::complex<double>(v.real(), v.imag()) * std::polar(1.0, sign * ph);
                }
            }
        }
        c[size_t(j)] = acc;
    }
    return c;
}

std::vector<std::complex<double>> ndft1(const Problem& p, const std::vector<std::complex<double>>& c,
                                        int sign) {
    const int64_t I = p.n[0] * p.n[1] * p.n[2];
    std::vector<std::complex<double>> f(size_t(I), 0.0);
    for (int64_t a = 0; a < p.n[0]; ++a)
        for (int64_t b = 0; b < p.n[1]; ++b)
            for (int64_t q = 0; q < p.n[2]; ++q) {
                std::complex<double> acc = 0;
                for (int64_t j = 0; j < p.M; ++j) {
                    const double ph = double(p.k[0][size_t(j)]) * double(a - p.n[0] / 2) +
                                      double(p.k[1][size_t(j)]) * double(b - p.n[1] / 2) +
                                      double(p.k[2][size_t(j)]) * double(q - p.n[2] / 2);
                    acc += c[size_t(j)] * std::polar(1.0, sign * ph);
                }
                f[size_t((a * p.n[1] + b) * p.n[2] + q)] = acc;
            }
    return f;
}

std::array<tensor<cuda_t, f32_t, 1>, 3> device_coords(const Problem& p) {
    return {tensor<cuda_t, f32_t, 1>::from_host(p.k[0], {p.M}),
            tensor<cuda_t, f32_t, 1>::from_host(p.k[1], {p.M}),
            tensor<cuda_t, f32_t, 1>::from_host(p.k[2], {p.M})};
}

template <class F> void expect_throw(const char* what, const char* needle, F&& f) {
    try {
        f();
    } catch (const std::runtime_error& e) {
        const bool ok = std::string(e.what()).find(needle) != std::string::npos;
        std::printf("[cpp] %-46s threw \"%s\" %s\n", what, e.what(), ok ? "ok" : "FAIL");
        if (!ok) ++failures;
        return;
    }
    std::printf("[cpp] %-46s did not throw FAIL\n", what);
    ++failures;
}

}  // namespace

int main() {
    const Problem p = make_problem();
    const int64_t I = p.n[0] * p.n[1] * p.n[2];
    auto coords = device_coords(p);

    // type 2 over C transforms of S_c x
    std::vector<c64_t> Sx(size_t(p.C * I));
    for (int c = 0; c < p.C; ++c)
        for (int64_t i = 0; i < I; ++i) Sx[size_t(c * I + i)] = p.S[size_t(c * I + i)] * p.x[size_t(i)];
    nufft_opts<cuda_t, f32_t, 3> o2;
    o2.nmodes = p.n;
    o2.ntransf = p.C;
    o2.tol = 1e-5f;
    o2.sign = nufft_sign::DEFAULT_TYPE_2;
    auto plan2 = nufft_plan<cuda_t, f32_t, 3, nufft_type::FORWARD>::make(o2);
    plan2->setpts(coords);
    auto dimg = tensor<cuda_t, c64_t, 4>::from_host(Sx, {p.C, p.n[0], p.n[1], p.n[2]});
    auto dk = tensor<cuda_t, c64_t, 2>::empty({p.C, p.M});
    plan2->execute(dimg, dk);
    const auto hk = dk.to_host();
    std::vector<std::vector<std::complex<double>>> ref2(size_t(p.C));
    for (int c = 0; c < p.C; ++c) {
        std::vector<c64_t> f(Sx.begin() + c * I, Sx.begin() + (c + 1) * I);
        ref2[size_t(c)] = ndft2(p, f, -1);
        std::vector<c64_t> got(hk.begin() + c * p.M, hk.begin() + (c + 1) * p.M);
        report(("nufft_plan<TYPE_2> coil " + std::to_string(c) + " vs NDFT").c_str(),
               rel_l2(got, ref2[size_t(c)]), 2e-5);
    }

    // type 1 of the same data
    nufft_opts<cuda_t, f32_t, 3> o1 = o2;
    o1.sign = nufft_sign::DEFAULT_TYPE_1;
    auto plan1 = nufft_plan<cuda_t, f32_t, 3, nufft_type::BACKWARD>::make(o1);
    plan1->setpts(coords);
    auto dadj = tensor<cuda_t, c64_t, 4>::empty({p.C, p.n[0], p.n[1], p.n[2]});
    plan1->execute(dk, dadj);
    const auto hadj = dadj.to_host();
    {
        std::vector<std::complex<double>> c0(hk.begin(), hk.begin() + p.M);
        const auto ref1 = ndft1(p, c0, +1);
        std::vector<c64_t> got(hadj.begin(), hadj.begin() + I);
        report("nufft_plan<TYPE_1> coil 0 vs NDFT", rel_l2(got, ref1), 2e-5);
    }

    // the per-coil host path: each coil against the FP64 NDFT of its data
    std::vector<std::vector<std::complex<double>>> ref1(size_t(p.C));
    for (int c = 0; c < p.C; ++c)
        ref1[size_t(c)] = ndft1(p, std::vector<std::complex<double>>(hk.begin() + c * p.M, hk.begin() + (c + 1) * p.M),
                                +1);
    {
        tensor<cpu_t, c64_t, 2> hin({p.C, p.M});
        hin.vec() = hk;
        auto hout = nufft_backward_cpu_over_cuda<c64_t, f32_t, 3>(p.n, hin, coords);
        const auto hv = hout.vec();
        for (int c = 0; c < p.C; ++c)
            report(("nufft_backward_cpu_over_cuda coil " + std::to_string(c) + " vs NDFT").c_str(),
                   rel_l2(std::vector<c64_t>(hv.begin() + c * I, hv.begin() + (c + 1) * I), ref1[size_t(c)]), 2e-5);
    }

    // Toeplitz kernel + SENSE normal operator vs NDFT: sum_c conj(S_c) (1/N^3) A^H A (S_c x)
    auto kernel = tensor<cuda_t, c64_t, 3>::empty({2 * p.n[0], 2 * p.n[1], 2 * p.n[2]});
    auto ones = tensor<cuda_t, c64_t, 2>::from_host(std::vector<c64_t>(size_t(p.M), c64_t(1.f, 0.f)), {1, p.M});
    toeplitz_kernel<c64_t, 3>(coords, kernel, ones);
    std::vector<std::complex<double>> ref_n(size_t(I), 0.0);
    for (int c = 0; c < p.C; ++c) {
        const auto back = ndft1(p, ref2[size_t(c)], +1);
        for (int64_t i = 0; i < I; ++i) {
            const c64_t s = p.S[size_t(c * I + i)];
            ref_n[size_t(i)] += std::conj(std::complex<double>(s.real(), s.imag())) * back[size_t(i)] / double(I);
        }
    }
    auto smaps = tensor<cuda_t, c64_t, 4>::from_host(p.S, {p.C, p.n[0], p.n[1], p.n[2]});
    auto op = std::make_shared<NORMAL_T_T1_OP<cuda_t, c64_t, 3>>(kernel.clone(), std::move(smaps));
    auto dx = tensor<cuda_t, c64_t, 3>::from_host(p.x, p.n);
    const auto hn = (*op)(dx).to_host();
    report("NORMAL_T_T1_OP vs NDFT normal operator", rel_l2(hn, ref_n), 2e-5);

    // the reference's constructor: kernel and maps as cache_tensors (host data
    // taken to the device once); the same operator, bit for bit
    {
        tensor<cpu_t, c64_t, 3> hk(kernel.shape_arr());
        hk.vec() = kernel.to_host();
        tensor<cpu_t, c64_t, 4> hs({p.C, p.n[0], p.n[1], p.n[2]});
        hs.vec() = p.S;
        cache_tensor<c64_t, 3> ck(std::move(hk), 101);
        cache_tensor<c64_t, 4> cs(std::move(hs), 102);
        NORMAL_T_T1_OP<cuda_t, c64_t, 3> cop(ck.copy(), cs.copy());
        const auto hc = cop(dx).to_host();
        bool same = hc.size() == hn.size();
        for (size_t i = 0; same && i < hc.size(); ++i) same = hc[i] == hn[i];
        report("NORMAL_T_T1_OP from cache_tensors vs from device tensors (0 = bitwise)", same ? 0.0 : 1.0, 0.5);
        // the cached device copy is shared, not made again
        const bool shared = ck.get_ptr<cuda_t>(device_idx::CUDA0).get() == ck.get_ptr<cuda_t>(device_idx::CUDA0).get();
        report("cache_tensor keeps one device copy (0 = yes)", shared ? 0.0 : 1.0, 0.5);
    }

    // toeplitz_multiplication with the reference's NORMAL_T_T1_OP modes, then sum
    {
        auto in = tensor<cuda_t, c64_t, 4>::from_host(p.S, {p.C, p.n[0], p.n[1], p.n[2]});
        auto out = tensor<cuda_t, c64_t, 4>::empty(in.shape_arr());
        fft::toeplitz_multiplication<4>(in, out, kernel, std::nullopt, std::cref(dx), std::nullopt,
                                        fft::ToeplitzMultType::MULT_CONJ, fft::ToeplitzMultType::MULT,
                                        fft::ToeplitzMultType::NONE, fft::ToeplitzMultType::NONE,
                                        fft::ToeplitzMultType::NONE, fft::ToeplitzAccumulateType::NONE);
        const auto h = out.to_host();
        std::vector<c64_t> sum(size_t(I), c64_t(0.f, 0.f));
        for (int c = 0; c < p.C; ++c)
            for (int64_t i = 0; i < I; ++i) sum[size_t(i)] += h[size_t(c * I + i)];
        report("toeplitz_multiplication + sum vs NDFT", rel_l2(sum, ref_n), 2e-5);
    }

    // toeplitz_multiply on S_0 x vs (1/N^3) A^H A (S_0 x)
    {
        std::vector<c64_t> s0(Sx.begin(), Sx.begin() + I);
        auto t = tensor<cuda_t, c64_t, 4>::from_host(s0, {1, p.n[0], p.n[1], p.n[2]});
        toeplitz_multiply<c64_t, 3>(t, kernel);
        auto back = ndft1(p, ref2[0], +1);
        for (auto& v : back) v /= double(I);
        report("toeplitz_multiply vs NDFT", rel_l2(t.to_host(), back), 2e-5);
    }

    // multi-kernel operators: kernels {T, T}, diagonals {1, R} with a spatial phase R
    std::vector<c64_t> R(static_cast<size_t>(I));
    for (int64_t i = 0; i < I; ++i) R[size_t(i)] = std::polar(1.f, 0.37f * float(i % 17) - 0.11f * float(i / 160));
    auto idt_ref = ref_n;
    for (int c = 0; c < p.C; ++c) {
        std::vector<c64_t> f(static_cast<size_t>(I));
        for (int64_t i = 0; i < I; ++i) f[size_t(i)] = R[size_t(i)] * Sx[size_t(c * I + i)];
        const auto back = ndft1(p, ndft2(p, f, -1), +1);
        for (int64_t i = 0; i < I; ++i) {
            const std::complex<double> s(p.S[size_t(c * I + i)].real(), p.S[size_t(c * I + i)].imag());
            const std::complex<double> r(R[size_t(i)].real(), R[size_t(i)].imag());
            idt_ref[size_t(i)] += std::conj(s) * std::conj(r) * back[size_t(i)] / double(I);
        }
    }
    const int64_t nK = kernel.numel();
    std::vector<c64_t> hkern = kernel.to_host(), kk(size_t(2 * nK)), dd(size_t(2 * I));
    std::copy(hkern.begin(), hkern.end(), kk.begin());
    std::copy(hkern.begin(), hkern.end(), kk.begin() + nK);
    std::fill(dd.begin(), dd.begin() + I, c64_t(1.f, 0.f));
    std::copy(R.begin(), R.end(), dd.begin() + I);
    const std::array<int64_t, 4> kshape{2, 2 * p.n[0], 2 * p.n[1], 2 * p.n[2]}, dshape{2, p.n[0], p.n[1], p.n[2]};
    const std::array<int64_t, 4> sshape{p.C, p.n[0], p.n[1], p.n[2]};
    {
        NORMAL_IDT_T1_OP<cuda_t, c64_t, 3> idt(tensor<cuda_t, c64_t, 4>::from_host(kk, kshape),
                                               tensor<cuda_t, c64_t, 4>::from_host(dd, dshape),
                                               tensor<cuda_t, c64_t, 4>::from_host(p.S, sshape));
        report("NORMAL_IDT_T1_OP (L=2) vs NDFT", rel_l2(idt(dx).to_host(), idt_ref), 2e-5);
        std::vector<c64_t> W(size_t(p.C * p.C), c64_t(0.f, 0.f));
        for (int c = 0; c < p.C; ++c) W[size_t(c * p.C + c)] = c64_t(2.f, 0.f);
        NORMAL_IDTW_T1_OP<cuda_t, c64_t, 3> idtw(tensor<cuda_t, c64_t, 4>::from_host(kk, kshape),
                                                 tensor<cuda_t, c64_t, 4>::from_host(dd, dshape),
                                                 tensor<cuda_t, c64_t, 4>::from_host(p.S, sshape),
                                                 tensor<cuda_t, c64_t, 2>::from_host(W, {p.C, p.C}));
        auto twice = idt_ref;
        for (auto& v : twice) v *= 2.0;
        report("NORMAL_IDTW_T1_OP (W = 2 I) vs NDFT", rel_l2(idtw(dx).to_host(), twice), 2e-5);
        // <Psi, T2(Psi)> with Psi = S equals <x, sum_c conj(S_c) T(S_c x)>
        NORMAL_T_T2_OP<cuda_t, c64_t, 3> t2(kernel.clone(), dx.clone());
        auto psi = tensor<cuda_t, c64_t, 4>::from_host(p.S, sshape);
        const auto ht2 = t2(psi).to_host();
        std::complex<double> lhs = 0, rhs = 0;
        for (int64_t i = 0; i < p.C * I; ++i)
            lhs += std::conj(std::complex<double>(p.S[size_t(i)])) * std::complex<double>(ht2[size_t(i)]);
        for (int64_t i = 0; i < I; ++i) rhs += std::conj(std::complex<double>(p.x[size_t(i)])) * ref_n[size_t(i)];
        report("NORMAL_T_T2_OP: <S, T2(S)> vs <x, N x>", std::abs(lhs - rhs) / std::abs(rhs), 2e-5);

        // the reference's cache_tensor constructors build the same operators
        auto cache = [](const std::vector<c64_t>& v, auto shape, size_t h) {
            tensor<cpu_t, c64_t, std::tuple_size_v<decltype(shape)>> t(shape);
            t.vec() = v;
            return cache_tensor<c64_t, std::tuple_size_v<decltype(shape)>>(std::move(t), h);
        };
        const std::array<int64_t, 3> kshape3{2 * p.n[0], 2 * p.n[1], 2 * p.n[2]};
        const std::array<int64_t, 2> wshape{p.C, p.C};
        NORMAL_IDT_T1_OP<cuda_t, c64_t, 3> idt_c(cache(kk, kshape, 301), cache(dd, dshape, 302),
                                                 cache(p.S, sshape, 303));
        NORMAL_IDTW_T1_OP<cuda_t, c64_t, 3> idtw_c(cache(kk, kshape, 301), cache(dd, dshape, 302),
                                                   cache(p.S, sshape, 303), cache(W, wshape, 304));
        NORMAL_T_T2_OP<cuda_t, c64_t, 3> t2_c(cache(hkern, kshape3, 305), cache(p.x, p.n, 306));
        NORMAL_IDT_T2_OP<cuda_t, c64_t, 3> idt2(tensor<cuda_t, c64_t, 4>::from_host(kk, kshape),
                                                tensor<cuda_t, c64_t, 4>::from_host(dd, dshape), dx.clone());
        NORMAL_IDT_T2_OP<cuda_t, c64_t, 3> idt2_c(cache(kk, kshape, 301), cache(dd, dshape, 302),
                                                  cache(p.x, p.n, 306));
        NORMAL_IDTW_T2_OP<cuda_t, c64_t, 3> idtw2(tensor<cuda_t, c64_t, 4>::from_host(kk, kshape),
                                                  tensor<cuda_t, c64_t, 4>::from_host(dd, dshape), dx.clone(),
                                                  tensor<cuda_t, c64_t, 2>::from_host(W, {p.C, p.C}));
        NORMAL_IDTW_T2_OP<cuda_t, c64_t, 3> idtw2_c(cache(kk, kshape, 301), cache(dd, dshape, 302),
                                                    cache(p.x, p.n, 306), cache(W, wshape, 304));
        const bool same = idt_c(dx).to_host() == idt(dx).to_host() && idtw_c(dx).to_host() == idtw(dx).to_host() &&
                          t2_c(psi).to_host() == ht2 && idt2_c(psi).to_host() == idt2(psi).to_host() &&
                          idtw2_c(psi).to_host() == idtw2(psi).to_host();
        report("normal operators from cache_tensors vs device tensors (0 = bitwise)", same ? 0.0 : 1.0, 0.5);
        expect_throw("kernel / diagonal count mismatch", "Number of kernels and kernel diagonals must match.", [&] {
            NORMAL_IDT_T1_OP<cuda_t, c64_t, 3> bad(tensor<cuda_t, c64_t, 4>::from_host(kk, kshape),
                                                   tensor<cuda_t, c64_t, 4>::from_host(R, {1, p.n[0], p.n[1], p.n[2]}),
                                                   tensor<cuda_t, c64_t, 4>::from_host(p.S, sshape));
            (void)bad(dx);
        });
    }

    // time-segmented SENSE with one segment of unit factors is plain SENSE
    {
        auto bil = tensor<cuda_t, c64_t, 2>::from_host(std::vector<c64_t>(size_t(p.M), c64_t(1.f, 0.f)), {1, p.M});
        auto cjl = tensor<cuda_t, c64_t, 4>::from_host(std::vector<c64_t>(size_t(I), c64_t(1.f, 0.f)),
                                                       {1, p.n[0], p.n[1], p.n[2]});
        auto maps = tensor<cuda_t, c64_t, 4>::from_host(p.S, sshape);
        const auto y = sense_forward_offresonance<c64_t, f32_t, 3>(coords, dx, maps, bil, cjl, 2).to_host();
        std::vector<c64_t> y1(y.begin() + p.M, y.begin() + 2 * p.M);
        report("sense_forward_offresonance coil 1 vs NDFT", rel_l2(y1, ref2[1]), 2e-5);
        const auto xa = sense_adjoint_offresonance<c64_t, f32_t, 3>(coords, dk, maps, bil, cjl).to_host();
        std::vector<c64_t> k0(hk.begin(), hk.begin() + p.M);
        std::vector<std::complex<double>> k0d(k0.begin(), k0.end());
        auto back = ndft1(p, k0d, +1);
        for (int64_t i = 0; i < I; ++i)
            back[size_t(i)] *= std::conj(std::complex<double>(p.S[size_t(i)]));
        std::vector<c64_t> x0(xa.begin(), xa.begin() + I);
        report("sense_adjoint_offresonance coil 0 vs NDFT", rel_l2(x0, back), 2e-5);

        // the reference's form: cache_tensor arguments, host results, the same numbers
        auto host = [](const auto& dt, size_t h) {
            constexpr size_t R = std::tuple_size_v<std::decay_t<decltype(dt.shape_arr())>>;
            using E = std::decay_t<decltype(*dt.data_ptr())>;
            tensor<cpu_t, E, R> t(dt.shape_arr());
            t.vec() = dt.to_host();
            return cache_tensor<E, R>(std::move(t), h);
        };
        std::array<cache_tensor<f32_t, 1>, 3> ccoords{host(coords[0], 201), host(coords[1], 202), host(coords[2], 203)};
        auto cimg = host(dx, 204);
        auto cmaps = host(maps, 205);
        auto cbil = host(bil, 206);
        auto ccjl = host(cjl, 207);
        auto ckd = host(dk, 208);
        const auto yc = sense_forward_offresonance<cuda_t, c64_t, f32_t, 3>(ccoords, cimg, cmaps, cbil, ccjl);
        const auto xc = sense_adjoint_offresonance<cuda_t, c64_t, f32_t, 3>(ccoords, ckd, cmaps, cbil, ccjl);
        report("sense_*_offresonance from cache_tensors vs device tensors (0 = bitwise)",
               (yc.vec() == y && xc.vec() == xa) ? 0.0 : 1.0, 0.5);

        // the thread-pool form: two workers on this device split the 3 coils
        // into chunks of 2 and 1; a kept plan stays in the storages until cleared
        std::vector<threading::storage> st(2);
        for (auto& s : st) s.add<device_idx>("device_idx", std::make_shared<device_idx>(device_idx::CUDA0));
        auto pool = std::make_shared<threading::storage_thread_pool>(std::move(st));
        threading::storage_thread_pool_interface iface(pool, {"device_idx"});
        const auto yp = sense_forward_offresonance<cuda_t, c64_t, f32_t, 3>(ccoords, cimg, cmaps, cbil, ccjl, iface,
                                                                            false, false, false);
        bool kept = true;
        for (const auto& s : pool->storages()) kept &= s.exist("nufft_plan");
        const auto yp2 = sense_forward_offresonance<cuda_t, c64_t, f32_t, 3>(ccoords, cimg, cmaps, cbil, ccjl, iface);
        bool cleared = true;
        for (const auto& s : pool->storages()) cleared &= !s.exist("nufft_plan");
        const auto xp = sense_adjoint_offresonance<cuda_t, c64_t, f32_t, 3>(ccoords, ckd, cmaps, cbil, ccjl, iface);
        // every coil of the pool forms against the FP64 NDFT: y_c = A (S_c x) and
        // x_c = conj(S_c) A^H y_c (one segment of unit factors)
        const auto ypv = yp.vec(), xpv = xp.vec();
        for (int c = 0; c < p.C; ++c) {
            report(("sense_forward_offresonance 2-worker pool coil " + std::to_string(c) + " vs NDFT").c_str(),
                   rel_l2(std::vector<c64_t>(ypv.begin() + c * p.M, ypv.begin() + (c + 1) * p.M), ref2[size_t(c)]),
                   2e-5);
            std::vector<std::complex<double>> xref = ref1[size_t(c)];
            for (int64_t i = 0; i < I; ++i) xref[size_t(i)] *= std::conj(std::complex<double>(p.S[size_t(c * I + i)]));
            report(("sense_adjoint_offresonance 2-worker pool coil " + std::to_string(c) + " vs NDFT").c_str(),
                   rel_l2(std::vector<c64_t>(xpv.begin() + c * I, xpv.begin() + (c + 1) * I), xref), 2e-5);
        }
        report("sense_forward_offresonance with the kept plan (0 = bitwise)", yp2.vec() == yp.vec() ? 0.0 : 1.0, 0.5);
        report("pool keeps the plan, then free_nufft_plan clears it (0 = yes)", kept && cleared ? 0.0 : 1.0, 0.5);
        expect_throw("pool worker without a device_idx", "Storage does not contain key: device_idx", [&] {
            std::vector<threading::storage> bare(1);
            threading::storage_thread_pool_interface none(
                std::make_shared<threading::storage_thread_pool>(std::move(bare)));
            (void)sense_forward_offresonance<cuda_t, c64_t, f32_t, 3>(ccoords, cimg, cmaps, cbil, ccjl, none);
        });
    }

    // CG on the normal equations: 30 iterations reduce the residual by 1e3
    {
        auto b = (*op)(dx);
        auto x = tensor<cuda_t, c64_t, 3>::zeros(p.n);
        auto cg = conjugate_gradient(op, 30, 0.0);
        cg(x, b);
        const auto Ax = (*op)(x).to_host();
        const auto hb = b.to_host();
        report("conjugate_gradient: ||A x - b|| / ||b||", rel_l2(Ax, hb), 1e-3);
        // preconditioned CG (cg.cppm:16-119) with P = 0.5 I takes the plain CG
        // steps (z = r / 2 rescales alpha by 2 and p by 1/2); a restart on the
        // 4th step and the after-restart callback once per outer iteration
        auto xp = tensor<cuda_t, c64_t, 3>::zeros(p.n);
        auto pcg = conjugate_gradient(op, std::make_shared<ScaleOp>(ScaleOp{0.5}), 10, 3, 0.0);
        int steps = 0, outers = 0;
        auto xc = tensor<cuda_t, c64_t, 3>::zeros(p.n);
        conjugate_gradient(op, 10, 0.0)(xc, b);
        auto x10 = tensor<cuda_t, c64_t, 3>::zeros(p.n);
        conjugate_gradient(op, std::make_shared<ScaleOp>(ScaleOp{0.5}), 10, 1, 0.0)(x10, b);
        report("PCG (P = I/2), 10 steps vs plain CG", rel_l2(x10.to_host(), xc.to_host()), 1e-5);
        pcg(xp, b, std::function<bool(tensor<cuda_t, c64_t, 3>&)>([&](tensor<cuda_t, c64_t, 3>&) { return ++steps == 4; }),
            std::function<void(tensor<cuda_t, c64_t, 3>&)>([&](tensor<cuda_t, c64_t, 3>&) { ++outers; }));
        const auto Axp = (*op)(xp).to_host();
        report("PCG with a restart: ||A x - b|| / ||b||", rel_l2(Axp, hb), 1e-3);
        report("PCG callbacks: 24 steps, 3 outer iterations (0 = yes)", steps == 24 && outers == 3 ? 0.0 : 1.0, 0.5);
        auto [lam, v] = max_eig(*op, dx.clone(), 20);
        std::printf("[cpp] max_eig estimate %.6f\n", double(lam));
        if (!(lam > 0.f) || !std::isfinite(lam)) ++failures;
    }

    // error behaviour mirrors the reference's messages
    expect_throw("verify_coords length mismatch", "coords have different lengths", [&] {
        std::array<tensor<cuda_t, f32_t, 1>, 3> bad = {coords[0], coords[1],
                                                       tensor<cuda_t, f32_t, 1>::empty({p.M - 1})};
        plan2->setpts(bad);
    });
    expect_throw("execute with wrong ntransf", "input first dimension must match ntransf", [&] {
        auto small = tensor<cuda_t, c64_t, 2>::empty({1, p.M});
        plan1->execute(small, dadj);
    });
    expect_throw("library error surfaces as runtime_error", "hasty", [&] {
        nufft_opts<cuda_t, f32_t, 3> bad = o2;
        bad.tol = 0.f;
        (void)nufft_plan<cuda_t, f32_t, 3, nufft_type::FORWARD>::make(bad);
    });

    plan1->free();
    plan2->free();
    std::printf("[cpp] %s (%d failure%s)\n", failures ? "FAILED" : "ALL OK", failures, failures == 1 ? "" : "s");
    return failures ? 1 : 0;
}
