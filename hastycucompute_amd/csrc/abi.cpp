// C-ABI entry points for the NUFFT plan plus library-wide utilities.
// Each function converts hasty::Error into a status code (guarded()).
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "kb.hpp"
#include "nufft.hpp"

#include <map>
#include <vector>

namespace hasty {
namespace {
thread_local std::string g_last_error;
struct ProfEntry { std::string name; hipEvent_t start, stop; };
thread_local bool g_prof_on = false;
thread_local std::vector<ProfEntry> g_prof;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
bool profiling_enabled() { return g_prof_on; }
void profile_record(const char* name, hipEvent_t start, hipEvent_t stop) {
    g_prof.push_back({name, start, stop});
}
}  // namespace hasty

using namespace hasty;

extern "C" {

const char* hasty_last_error(void) { return g_last_error.c_str(); }

int hasty_version(void) { return 100; }

int hasty_profile_enable(int on) {
    g_prof_on = on != 0;
    return HASTY_OK;
}

int hasty_profile_read(char* names, int64_t names_len, double* ms, int32_t* counts,
                       int32_t max_entries, int32_t* n_entries) {
    return guarded([&] {
        std::map<std::string, std::pair<double, int>> agg;
        std::vector<std::string> order;
        for (auto& e : g_prof) {
            HASTY_CHECK_HIP(hipEventSynchronize(e.stop));
            float t = 0.f;
            HASTY_CHECK_HIP(hipEventElapsedTime(&t, e.start, e.stop));
            if (!agg.count(e.name)) order.push_back(e.name);
            agg[e.name].first += t;
            agg[e.name].second += 1;
            (void)hipEventDestroy(e.start);
            (void)hipEventDestroy(e.stop);
        }
        g_prof.clear();
        std::string joined;
        int n = 0;
        for (auto& nm : order) {
            if (n >= max_entries) break;
            if (ms) ms[n] = agg[nm].first;
            if (counts) counts[n] = agg[nm].second;
            joined += nm;
            joined += '\n';
            ++n;
        }
        if (n_entries) *n_entries = n;
        if (names && names_len > 0) {
            const size_t k = std::min<size_t>(joined.size(), size_t(names_len - 1));
            std::memcpy(names, joined.data(), k);
            names[k] = 0;
        }
    });
}

int hasty_device_synchronize(void) {
    return guarded([] { HASTY_CHECK_HIP(hipDeviceSynchronize()); });
}

int hasty_nufft_makeplan(int type, int dim, const int64_t* nmodes, int sign, int ntransf,
                         double tol, int precision, const hasty_nufft_opts* opts, void* stream,
                         hasty_nufft_plan* plan_out) {
    return guarded([&] {
        HASTY_REQUIRE(plan_out != nullptr, "plan output pointer is NULL");
        HASTY_REQUIRE(precision == HASTY_F32 || precision == HASTY_F64, "invalid precision");
        *plan_out = reinterpret_cast<hasty_nufft_plan>(
            nufft_makeplan(type, dim, nmodes, sign, ntransf, tol, precision == HASTY_F64, opts,
                           static_cast<hipStream_t>(stream)));
    });
}

int hasty_nufft_setpts(hasty_nufft_plan plan, int64_t M, const void* c0, const void* c1,
                       const void* c2) {
    return guarded([&] {
        HASTY_REQUIRE(plan != nullptr, "plan is NULL");
        nufft_setpts(reinterpret_cast<NufftPlan*>(plan), M, c0, c1, c2);
    });
}

int hasty_nufft_execute(hasty_nufft_plan plan, void* c, void* f) {
    return guarded([&] {
        HASTY_REQUIRE(plan != nullptr, "plan is NULL");
        nufft_execute(reinterpret_cast<NufftPlan*>(plan), c, f);
    });
}

int hasty_nufft_set_stream(hasty_nufft_plan plan, void* stream) {
    return guarded([&] {
        HASTY_REQUIRE(plan != nullptr, "plan is NULL");
        reinterpret_cast<NufftPlan*>(plan)->stream = static_cast<hipStream_t>(stream);
    });
}

int hasty_nufft_destroy(hasty_nufft_plan plan) {
    return guarded([&] { delete reinterpret_cast<NufftPlan*>(plan); });
}

int hasty_nufft_coord_grad(hasty_nufft_plan plan, const void* h, const void* w, void* out) {
    return guarded([&] {
        HASTY_REQUIRE(plan != nullptr, "plan is NULL");
        nufft_coord_grad(reinterpret_cast<NufftPlan*>(plan), h, w, out);
    });
}

int hasty_sense_forward(hasty_nufft_plan plan, int ncoils, int L, const void* image,
                        const void* smaps, const void* bil, const void* cjl, void* kdata) {
    return guarded([&] {
        HASTY_REQUIRE(plan != nullptr, "plan is NULL");
        sense_offres_forward(reinterpret_cast<NufftPlan*>(plan), ncoils, L, image, smaps, bil, cjl,
                             kdata);
    });
}

int hasty_sense_adjoint(hasty_nufft_plan plan, int ncoils, int L, const void* kdata,
                        const void* smaps, const void* bil, const void* cjl, void* images,
                        int combine) {
    return guarded([&] {
        HASTY_REQUIRE(plan != nullptr, "plan is NULL");
        sense_offres_adjoint(reinterpret_cast<NufftPlan*>(plan), ncoils, L, kdata, smaps, bil, cjl,
                             images, combine ? 1 : 0);
    });
}

int hasty_nufft_default_opts(hasty_nufft_opts* opts) {
    return guarded([&] {
        HASTY_REQUIRE(opts != nullptr, "opts is NULL");
        *opts = hasty_nufft_opts{};
        opts->upsamp = 2.0;
        opts->method = HASTY_METHOD_DEFAULT;
    });
}

int hasty_nufft_kernel_width(double tol, double upsamp, int dim, int precision, int* width) {
    return guarded([&] {
        HASTY_REQUIRE(width != nullptr, "width is NULL");
        HASTY_REQUIRE(tol > 0 && tol < 1, "tol must be in (0, 1)");
        HASTY_REQUIRE(dim == 2 || dim == 3, "nufft dim must be 2 or 3");
        HASTY_REQUIRE(precision == HASTY_F32 || precision == HASTY_F64, "precision must be HASTY_F32 or HASTY_F64");
        *width = kb_width_for(tol, upsamp > 0 ? upsamp : 2.0, dim, precision == HASTY_F64);
    });
}

int hasty_nufft_info(hasty_nufft_plan plan, int* width, double* beta, int64_t* nf3, int32_t* tile3,
                     int64_t* device_bytes) {
    return guarded([&] {
        HASTY_REQUIRE(plan != nullptr, "plan is NULL");
        const NufftPlan* p = reinterpret_cast<const NufftPlan*>(plan);
        if (width) *width = p->w;
        if (beta) *beta = p->beta;
        if (nf3) for (int s = 0; s < 3; ++s) nf3[s] = p->nf[s];
        if (tile3) for (int s = 0; s < 3; ++s) tile3[s] = p->T[s];
        if (device_bytes) *device_bytes = p->device_bytes;
    });
}

}  // extern "C"
