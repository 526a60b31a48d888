// solve.h -- GEMM-form triangular solves with the multifrontal factor.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "factor.h"
#include "plan.h"

struct DevSolve {
    const int64_t *poff;
    const int32_t *fwd_tiles, *bwd_tiles;
    const int32_t *fwd_tbase_rel, *bwd_tbase_rel;
    const int32_t *fwd_nks, *bwd_nks;
    const int32_t *pull_ptr, *pull_src;
    const int32_t *lblk;  // (front, first column) of every 32-wide L11 column block
    int n_lblk;
};

void solve_prepare(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S, const double *F,
                   double *Pbuf, double *PTbuf, hipStream_t s);
void solve_fwd(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S, const double *Pbuf,
               double *W, double *part, double *X, int nrhs, hipStream_t s);
void solve_bwd(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S, const double *PTbuf,
               double *W, double *part, double *X, int nrhs, hipStream_t s);
