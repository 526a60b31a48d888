// solve.h -- GEMM-form triangular solves with the multifrontal factor.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "factor.h"
#include "plan.h"

// One product tile and the front data its workgroup needs, in one 64-byte
// record (the workgroup reads it with one scalar load): front, row tile,
// k-slice, pivots ns, rows m, first pivot, k-slice width, k-slices of the
// front, whether the front has children, and the front's vector slot, factor
// offset (P or PT) and row-list offset.
struct alignas(16) TileDesc {
    int32_t f, rt, ks, ns, m, piv0, kw, nks, parent, pad;
    int64_t vec_off, poff, row_off;
};
static_assert(sizeof(TileDesc) == 64, "TileDesc is one 64-byte record");

struct DevSolve {
    const int64_t *poff;
    const TileDesc *fwd_desc, *bwd_desc;  // per tile, in the order of fwd_tiles / bwd_tiles
    const int32_t *fwd_tiles, *bwd_tiles;
    const int32_t *fwd_tbase_rel, *bwd_tbase_rel;
    const int32_t *fwd_nks, *bwd_nks;
    const int32_t *fwd_kw, *bwd_kw;
    const int32_t *pull_ptr, *pull_src;  // pull map: per vector slot, the child slots it adds (absolute)
    const int32_t *lblk;  // (front, first column) of every 32-wide L11 column block
    int n_lblk;
    // the pull map as (first, second) child slot per vector slot, -1 for
    // none; nullptr when some row has more than two entries
    const int32_t *pull2;
};

// Optional HIP-event timing of the solve product launches (diagnostics):
// every product launch is bracketed by two events on its stream; collect()
// waits for them and accumulates the elapsed time and the launches'
// algorithmic bytes (factor entries + vector rows in and out).
struct ProdTimer {
    std::vector<hipEvent_t> pool;
    std::vector<double> pending_bytes;
    std::vector<int> pending_tag;
    size_t used = 0;
    // per tag: 0 = leaf height, 1 = the heights above it
    double total_ms[2] = {0.0, 0.0}, total_bytes[2] = {0.0, 0.0};
    long launches[2] = {0, 0};
    void begin(hipStream_t s, double bytes, int tag);
    void end(hipStream_t s);
    void collect();
    void reset();
    ~ProdTimer();
};

void solve_prepare(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S, const double *F,
                   double *Pbuf, double *PTbuf, hipStream_t s);
// Forward (L y = b) and backward (L^T x = y) solves of nrhs <= MAXR
// right-hand sides, column-major with ld n_free, from xin into xout.  The two
// arrays must be distinct (hipErrorInvalidValue otherwise): every product
// launch reads xin while it writes xout.
hipError_t solve_fwd(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S,
                     const double *Pbuf, double *W, double *part, const double *xin, double *xout, int nrhs,
                     hipStream_t s, ProdTimer *pt = nullptr);
hipError_t solve_bwd(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S,
                     const double *PTbuf, double *W, double *part, const double *xin, double *xout, int nrhs,
                     hipStream_t s, ProdTimer *pt = nullptr);
// One tree height of one direction (the product launch and, for split
// fronts, the finish launch); A is P (forward) or PT (backward).  With the
// lower heights (forward) or the higher ones (backward) done, repeating a
// height's launches rewrites the same values.
hipError_t solve_height(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S, int bwd,
                        int h, const double *A, double *W, double *part, const double *xin, double *xout, int nrhs,
                        hipStream_t s, ProdTimer *pt = nullptr);
