// tune.cpp -- launch shapes of the triangular-solve products per tree height
// and the autotuner that picks them by measurement.
//
// A height's product launch is shaped by (kmax, ks, pch, kg): fronts whose
// inner dimension is at most kmax are one k-slice, wider ones are cut in
// ks-wide slices (reduced by a finish launch), pch four-k steps of factor
// loads go out per chunk, and kg groups of four wavefronts split each tile's
// steps.  The best shape depends on the height's front sizes and counts (the
// leaf height streams most of the factor and wants occupancy; the few large
// fronts near the root want more workgroups or fewer launches), so the
// autotuner times each height of each direction on the model's own factor for
// a set of candidates, confirms the fastest against the starting shape in an
// alternating re-timing, and keeps the new shapes only if the whole solve
// gets faster.
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace s105;

namespace {

int plan_rebuild(sol105_ctx *ctx, sol105_model *M) {
    std::vector<int> ns(M->sym.fronts.size()), m(M->sym.fronts.size());
    for (size_t f = 0; f < ns.size(); f++) {
        ns[f] = M->sym.fronts[f].ns;
        m[f] = M->sym.fronts[f].m;
    }
    solve_plan_build(M->plan, ns, m, M->flist_host, M->splan, M->prod_cfg);
    return upload_solve_tiles(ctx, M);
}

// median device time of `reps` repetitions of height h, direction bwd
double time_height(sol105_ctx *ctx, sol105_model *M, int bwd, int h, const double *xin, double *xout, int nrhs,
                   int reps, hipEvent_t e0, hipEvent_t e1) {
    hipStream_t s = ctx->stream;
    const double *A = bwd ? M->d_PT.p : M->d_P.p;
    std::vector<float> t;
    for (int r = 0; r < reps + 1; r++) {
        (void)hipEventRecord(e0, s);
        if (solve_height(M->plan, M->dev, M->dsolve, M->splan, bwd, h, A, M->d_W.p, M->d_part.p, xin, xout, nrhs,
                         s) != hipSuccess)
            return -1.0;
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) t.push_back(ms);  // the first repetition warms caches
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

}  // namespace

extern "C" {

int sol105_solve_cfg_get(sol105_ctx *ctx, const sol105_model *M, int32_t *cfg, int32_t *n_heights) {
    if (!ctx || !M || !n_heights) return set_err(ctx, SOL105_E_INVALID, "null argument");
    if (!M->have_sym) return set_err(ctx, SOL105_E_STATE, "no symbolic analysis yet: solve once first");
    const int nh = M->plan.height + 1;
    if (cfg) {
        if (*n_heights < nh) return set_err(ctx, SOL105_E_INVALID, "cfg holds %d heights, the plan has %d", *n_heights, nh);
        for (int h = 0; h < nh; h++) {
            const SolveHeight &SH = M->splan.heights[h];
            const ProdCfg c[2] = {SH.fwd_cfg, SH.bwd_cfg};
            for (int d = 0; d < 2; d++) {
                int32_t *o = cfg + 8 * h + 4 * d;
                o[0] = c[d].kmax;
                o[1] = c[d].ks;
                o[2] = c[d].pch;
                o[3] = c[d].kg;
            }
        }
    }
    *n_heights = nh;
    return SOL105_OK;
}

int sol105_solve_heights(sol105_ctx *ctx, const sol105_model *M, int64_t *shape, int32_t *n_heights) {
    if (!ctx || !M || !n_heights) return set_err(ctx, SOL105_E_INVALID, "null argument");
    if (!M->have_sym) return set_err(ctx, SOL105_E_STATE, "no symbolic analysis yet: solve once first");
    const int nh = M->plan.height + 1;
    if (shape) {
        if (*n_heights < nh) return set_err(ctx, SOL105_E_INVALID, "shape holds %d heights, the plan has %d", *n_heights, nh);
        for (int h = 0; h < nh; h++) {
            const HeightInfo &H = M->plan.heights[h];
            const SolveHeight &SH = M->splan.heights[h];
            shape[4 * h] = H.n;
            shape[4 * h + 1] = H.max_m;
            shape[4 * h + 2] = H.max_ns;
            shape[4 * h + 3] = SH.sum_nz;
        }
    }
    *n_heights = nh;
    return SOL105_OK;
}

int sol105_solve_cfg_set(sol105_ctx *ctx, sol105_model *M, const int32_t *cfg, int32_t n_heights) {
    if (!ctx || !M || (n_heights > 0 && !cfg) || n_heights < 0) return set_err(ctx, SOL105_E_INVALID, "null argument");
    std::vector<std::pair<ProdCfg, ProdCfg>> c(n_heights);
    for (int h = 0; h < n_heights; h++) {
        ProdCfg *d[2] = {&c[h].first, &c[h].second};
        for (int k = 0; k < 2; k++) {
            const int32_t *v = cfg + 8 * h + 4 * k;
            if (v[0] > 0) d[k]->kmax = v[0];
            if (v[1] > 0) d[k]->ks = v[1];
            if (v[2] != 0 && v[2] != 8 && v[2] != 16 && v[2] != 32)
                return set_err(ctx, SOL105_E_INVALID, "pch must be 0, 8, 16 or 32 (height %d)", h);
            if (v[3] != 0 && v[3] != 1 && v[3] != 2 && v[3] != 4)
                return set_err(ctx, SOL105_E_INVALID, "kg must be 0, 1, 2 or 4 (height %d)", h);
            if (v[3] > 1 && v[2] == 32)
                return set_err(ctx, SOL105_E_INVALID, "pch 32 runs with one k-group only (height %d)", h);
            d[k]->pch = v[2];
            d[k]->kg = v[3] > 0 ? v[3] : 1;
        }
    }
    (void)hipSetDevice(ctx->device);
    M->prod_cfg = c;
    if (M->have_sym) return plan_rebuild(ctx, M);
    return SOL105_OK;
}

int sol105_solve_autotune(sol105_ctx *ctx, sol105_model *M, int32_t reps, double *report) {
    if (!ctx || !M) return set_err(ctx, SOL105_E_INVALID, "null argument");
    if (!M->have_sym || !M->factor_valid)
        return set_err(ctx, SOL105_E_STATE, "no factorisation: run sol105_static_solve or a solve first");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    reps = std::max(3, std::min(reps, 50));
    const int n = M->sym.n_free, nrhs = MAXR, nh = M->plan.height + 1;
    // a block of right-hand sides and one full solve, so that every height's
    // inputs are in place
    DevBuf<double> X, Y;
    HIPCHK(X.alloc((size_t)n * nrhs));
    HIPCHK(Y.alloc((size_t)n * nrhs));
    {
        std::vector<double> h((size_t)n * nrhs);
        uint64_t z = 0x9e3779b97f4a7c15ull;
        for (auto &v : h) {
            z ^= z << 13; z ^= z >> 7; z ^= z << 17;
            v = (double)(z >> 11) / 9007199254740992.0 - 0.5;
        }
        HIPCHK(X.upload(h.data(), h.size(), s));
        HIPCHK(hipStreamSynchronize(s));
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    auto full = [&]() -> hipError_t {
        hipError_t e = solve_fwd(M->plan, M->dev, M->dsolve, M->splan, M->d_P.p, M->d_W.p, M->d_part.p, X.p, Y.p,
                                 nrhs, s);
        if (e != hipSuccess) return e;
        return solve_bwd(M->plan, M->dev, M->dsolve, M->splan, M->d_PT.p, M->d_W.p, M->d_part.p, Y.p, X.p, nrhs, s);
    };
    // whole block solves (forward + backward) before and after, for the report
    auto time_full = [&]() -> double {
        std::vector<float> t;
        for (int r = 0; r < reps + 1; r++) {
            (void)hipEventRecord(e0, s);
            if (full() != hipSuccess) return -1.0;
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r > 0) t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    };
    // the backward solve overwrites X: restore it before each full solve
    std::vector<double> x0((size_t)n * nrhs);
    HIPCHK(hipMemcpy(x0.data(), X.p, x0.size() * sizeof(double), hipMemcpyDeviceToHost));
    auto reset_x = [&]() -> hipError_t {
        return hipMemcpyAsync(X.p, x0.data(), x0.size() * sizeof(double), hipMemcpyHostToDevice, s);
    };
    HIPCHK(reset_x());
    const double before = time_full();
    if (before < 0) return set_err(ctx, SOL105_E_DEVICE, "autotune: a timed solve failed");
    if ((int)M->prod_cfg.size() < nh) M->prod_cfg.resize(nh);
    const std::vector<std::pair<ProdCfg, ProdCfg>> initial = M->prod_cfg;
    // sets one height's shape; the tiles depend on (kmax, ks) only, so a
    // change of pch or kg goes straight into the plan's copy of the shape
    auto set_shape = [&](int bwd, int h, const ProdCfg &c) -> int {
        ProdCfg &pc = bwd ? M->prod_cfg[h].second : M->prod_cfg[h].first;
        const bool retile = pc.kmax != c.kmax || pc.ks != c.ks;
        pc = c;
        if (retile) return plan_rebuild(ctx, M);
        SolveHeight &SH = M->splan.heights[h];
        ProdCfg &lc = bwd ? SH.bwd_cfg : SH.fwd_cfg;
        lc.pch = c.pch;
        lc.kg = c.pch == 32 ? 1 : c.kg;
        return SOL105_OK;
    };
    auto same = [](const ProdCfg &a, const ProdCfg &b) {
        return a.kmax == b.kmax && a.ks == b.ks && a.pch == b.pch && a.kg == b.kg;
    };
    const int kmaxs[3] = {128, 256, SOLVE_KMAX}, kss[3] = {128, 256, SOLVE_KMAX}, pchs[3] = {8, 16, 32},
              kgs[3] = {1, 2, 4};
    int timed = 0;
    for (int bwd = 0; bwd < 2; bwd++)
        for (int hh = 0; hh < nh; hh++) {
            // forward heights bottom-up, backward heights top-down: the
            // inputs of the timed height are those of a real solve
            const int h = bwd ? nh - 1 - hh : hh;
            if (M->plan.heights[h].n == 0) continue;
            HIPCHK(reset_x());
            HIPCHK(full() == hipSuccess ? hipSuccess : hipErrorUnknown);
            // forward timing reads X and writes Y; backward reads Y, writes X
            const double *xin = bwd ? Y.p : X.p;
            double *xout = bwd ? X.p : Y.p;
            const ProdCfg start = bwd ? M->prod_cfg[h].second : M->prod_cfg[h].first;
            ProdCfg best = start;
            int rc = plan_rebuild(ctx, M);
            if (rc) return rc;
            double best_t = time_height(ctx, M, bwd, h, xin, xout, nrhs, reps, e0, e1);
            timed++;
            const SolveHeight &SH0 = M->splan.heights[h];
            const int inner_max = bwd ? SH0.max_m : SH0.max_ns;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    if (kss[b] > kmaxs[a]) continue;
                    // slices wider than every front of the height: ks is moot
                    if (inner_max <= kmaxs[a] && b > 0) continue;
                    for (int c = 0; c < 3; c++)
                        for (int g = 0; g < 3; g++) {
                            if (pchs[c] == 32 && kgs[g] > 1) continue;
                            const ProdCfg cand{kmaxs[a], kss[b], pchs[c], kgs[g]};
                            if ((rc = set_shape(bwd, h, cand))) return rc;
                            const double t = time_height(ctx, M, bwd, h, xin, xout, nrhs, reps, e0, e1);
                            timed++;
                            if (t > 0 && t < 0.97 * best_t) {
                                best_t = t;
                                best = cand;
                            }
                        }
                }
            if (!same(best, start)) {
                // the fastest of many noisy timings is biased low: time the
                // winner against the starting shape again, alternating, and
                // keep it only if it still wins by 3 %
                std::vector<double> ta, tb;
                for (int r = 0; r < 3; r++) {
                    if ((rc = set_shape(bwd, h, start))) return rc;
                    ta.push_back(time_height(ctx, M, bwd, h, xin, xout, nrhs, reps, e0, e1));
                    if ((rc = set_shape(bwd, h, best))) return rc;
                    tb.push_back(time_height(ctx, M, bwd, h, xin, xout, nrhs, reps, e0, e1));
                    timed += 2;
                }
                std::sort(ta.begin(), ta.end());
                std::sort(tb.begin(), tb.end());
                if (!(tb[1] > 0 && tb[1] < 0.97 * ta[1])) best = start;
            }
            if ((rc = set_shape(bwd, h, best))) return rc;
        }
    int rc = plan_rebuild(ctx, M);
    if (rc) return rc;
    HIPCHK(reset_x());
    double after = time_full();
    if (after < 0) return set_err(ctx, SOL105_E_DEVICE, "autotune: a timed solve failed");
    if (after >= before) {
        // the heights' gains did not add up over the whole solve: keep the
        // shapes the model had
        M->prod_cfg = initial;
        if ((rc = plan_rebuild(ctx, M))) return rc;
        after = before;
    }
    if (report) {
        report[0] = before;
        report[1] = after;
        report[2] = timed;
        report[3] = nh;
    }
    return SOL105_OK;
}

}  // extern "C"
