// recover.cpp -- SOL 101 result recovery on the device-resident static state:
// reactions r = K u - F and the CQUAD4 / CTRIA3 / CBAR forces, stresses and
// strains (kernels in elements.hip).  Replaces the result build of
// Solver.solve_case (JFEM/src/solver/solve_case.jl:126-186, 866-868) and
// recover_shell_stresses! / recover_bar_stresses!
// (JFEM/src/solver/stress_recovery.jl:163-513).
#include "internal.h"

#include <algorithm>
#include <cstring>

using namespace s105;

namespace {

constexpr int Q4_COLS = 44, T3_COLS = 20, BAR_COLS = 18;

uint64_t edge_key(int32_t a, int32_t b) {
    const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return ((uint64_t)lo << 32) | hi;
}

// Macro-moment blend weight of each CTRIA3: clamp((edges shared with another
// shell - 1) / 2, 0, 1) over the CQUAD4 and CTRIA3 edges of the model.
// stress_recovery.jl:20-51.
void t3_blend_weights(const sol105_model *M, std::vector<double> &w) {
    std::vector<uint64_t> keys;
    keys.reserve(4 * (size_t)M->n_q4 + 3 * (size_t)M->n_t3);
    for (int e = 0; e < M->n_q4; e++)
        for (int k = 0; k < 4; k++) keys.push_back(edge_key(M->q4[4 * e + k], M->q4[4 * e + (k + 1) % 4]));
    for (int e = 0; e < M->n_t3; e++)
        for (int k = 0; k < 3; k++) keys.push_back(edge_key(M->t3[3 * e + k], M->t3[3 * e + (k + 1) % 3]));
    std::sort(keys.begin(), keys.end());
    w.assign(M->n_t3, 0.0);
    for (int e = 0; e < M->n_t3; e++) {
        int shared = 0;
        for (int k = 0; k < 3; k++) {
            const uint64_t key = edge_key(M->t3[3 * e + k], M->t3[3 * e + (k + 1) % 3]);
            const auto r = std::equal_range(keys.begin(), keys.end(), key);
            shared += (r.second - r.first) > 1 ? 1 : 0;
        }
        w[e] = std::min(1.0, std::max(0.0, (shared - 1) / 2.0));
    }
}

// events of the optional per-kernel timing, destroyed on every return path
struct Events {
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    ~Events() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace

extern "C" int sol105_static_recover(sol105_ctx *ctx, sol105_model *M, const double *bar_points, double *residual,
                                     double *q4_out, double *t3_out, double *bar_out, double *kernel_ms) {
    if (!ctx || !M) return set_err(ctx, SOL105_E_INVALID, "null argument");
    (void)hipSetDevice(ctx->device);
    if (!M->u_valid) return set_err(ctx, SOL105_E_STATE, "no static displacement: run sol105_static_solve first");
    if (residual && !M->K_valid)
        return set_err(ctx, SOL105_E_STATE, "no assembled stiffness for the reactions: run sol105_static_solve first");
    hipStream_t s = ctx->stream;
    if ((int)M->t3w.size() != M->n_t3 || (M->n_t3 > 0 && !M->d_t3w.p)) {
        t3_blend_weights(M, M->t3w);
        HIPCHK(M->d_t3w.upload(M->t3w.data(), M->t3w.size(), s));
    }
    const size_t nb = (size_t)M->n_bar;
    if (nb) {
        HIPCHK(M->d_barpts.alloc(8 * nb));
        if (bar_points)
            HIPCHK(hipMemcpyAsync(M->d_barpts.p, bar_points, 8 * nb * sizeof(double), hipMemcpyHostToDevice, s));
        else
            HIPCHK(hipMemsetAsync(M->d_barpts.p, 0, 8 * nb * sizeof(double), s));
    }
    const size_t ndof = (size_t)M->ndof(), nq = (size_t)Q4_COLS * M->n_q4, nt = (size_t)T3_COLS * M->n_t3,
                 nbr = (size_t)BAR_COLS * nb;
    HIPCHK(M->d_rec.alloc(ndof + nq + nt + nbr));
    double *d_res = M->d_rec.p, *d_q4 = d_res + ndof, *d_t3 = d_q4 + nq, *d_bar = d_t3 + nt;
    ElemArgs a = elem_args(M);
    Events evs;
    hipEvent_t *ev = evs.ev;
    if (kernel_ms)
        for (int i = 0; i < 5; i++) HIPCHK(hipEventCreate(&ev[i]));
    auto mark = [&](int i) -> hipError_t { return kernel_ms ? hipEventRecord(ev[i], s) : hipSuccess; };
    HIPCHK(mark(0));
    if (residual)
        launch_bsr_residual(M->n_nodes, M->d_rp.p, M->d_col.p, M->d_K.p, M->d_u.p, M->d_force.p, d_res, s);
    HIPCHK(mark(1));
    if (q4_out) launch_q4_recover(a, M->d_u.p, d_q4, s);
    HIPCHK(mark(2));
    if (t3_out) launch_t3_recover(a, M->d_u.p, M->d_t3w.p, d_t3, s);
    HIPCHK(mark(3));
    if (bar_out) launch_bar_recover(a, M->d_u.p, M->d_barpts.p, d_bar, s);
    HIPCHK(mark(4));
    HIPCHK(hipGetLastError());
    if (residual) HIPCHK(hipMemcpyAsync(residual, d_res, ndof * sizeof(double), hipMemcpyDeviceToHost, s));
    if (q4_out && nq) HIPCHK(hipMemcpyAsync(q4_out, d_q4, nq * sizeof(double), hipMemcpyDeviceToHost, s));
    if (t3_out && nt) HIPCHK(hipMemcpyAsync(t3_out, d_t3, nt * sizeof(double), hipMemcpyDeviceToHost, s));
    if (bar_out && nbr) HIPCHK(hipMemcpyAsync(bar_out, d_bar, nbr * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (kernel_ms)
        for (int i = 0; i < 4; i++) kernel_ms[i] = ev_ms(ev[i], ev[i + 1]);
    return SOL105_OK;
}
