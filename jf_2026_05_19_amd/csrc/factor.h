// factor.h -- device-side description of the multifrontal factor.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "plan.h"

// Per-front arrays on the device (indexed by front id).
struct DevFronts {
    const int32_t *ns, *nu, *m, *piv0;
    const int64_t *off, *row_off, *vec_off, *rel_off;
    const int32_t *rows, *rel;
    const int32_t *child_ptr, *child_idx;
    const int32_t *flist;  // front ids grouped by height
    const int32_t *k_src;
    const int64_t *k_dst;
};

void dev_fill(double *p, int64_t n, double v, hipStream_t s);
void factor_numeric(const FactorPlan &P, const DevFronts &D, const double *kvals, double *F, double *piv, int *err,
                    hipStream_t s);
void launch_gather_elim(int n_free, int ndof, int nrhs, const int32_t *dof_of_elim, const double *xf, double *xe,
                        hipStream_t s);
void launch_scatter_elim(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof, const double *xe, double *xf,
                         hipStream_t s);
// y = scale * A x, A 6x6 BSR; x and y interleaved: x[dof * nrhs + r] (nrhs <= 32)
void launch_bsr_spmv(int n_nodes, const int32_t *rp, const int32_t *col, const double *val, const double *x, double *y,
                     int nrhs, double scale, hipStream_t s);
// elimination order (column-major, ld n_free) <-> interleaved DOF order
void launch_scatter_elim_il(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof, const double *xe, double *xf,
                            hipStream_t s);
void launch_gather_elim_il(int n_free, int nrhs, const int32_t *dof_of_elim, const double *xf, double *xe,
                           hipStream_t s);
void launch_kdiag_elim(int n_free, const int32_t *dof_of_elim, const int32_t *diag_blk, const double *kv, double *out,
                       hipStream_t s);
void launch_pivot_ratio(int n, const double *piv, const double *kdiag_elim, int *count, uint8_t *flag, hipStream_t s);
// kv (BSR values) += shift on every diagonal entry
void launch_add_diag(int n_nodes, const int32_t *diag_blk, double shift, double *kv, hipStream_t s);
// *out = max |x_i| (device scalar)
void launch_maxabs(int64_t n, const double *x, double *out, hipStream_t s);
