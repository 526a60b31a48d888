// common.h -- shared definitions of the MI355X SOL 105 library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#define S105_WAVE 64

// Element-block layout: every element writes its global stiffness as dense
// 6x6 node blocks (row-major), CQUAD4 first (16 blocks per element, index
// e*16 + a*4 + b for corner nodes a, b), then CTRIA3 (9 blocks per element,
// index q4_blocks + e*9 + a*3 + b), then CBAR (4 blocks per bar, index
// q4_blocks + t3_blocks + b*4 + a*2 + c).  The block gather sums these into
// the 6x6 block-sparse-row matrix in a fixed element order.
struct ElemArgs {
    int n_q4;
    int n_t3;
    int n_bar;
    const double *xyz;        // n_nodes x 3
    const int32_t *q4;        // n_q4 x 4
    const double *q4prop;     // n_q4 x 6
    const int32_t *t3;        // n_t3 x 3
    const double *t3prop;     // n_t3 x 6
    const int32_t *bar;       // n_bar x 2
    const double *bar_vref;   // n_bar x 3
    const double *barprop;    // n_bar x 8
    double k6rot;
};

// launchers (elements.hip)
void launch_q4_ke(const ElemArgs &a, double *eb, hipStream_t s);
void launch_q4_kg(const ElemArgs &a, const double *u, double *eb, hipStream_t s);
void launch_t3_ke(const ElemArgs &a, double *eb, hipStream_t s);
void launch_t3_kg(const ElemArgs &a, const double *u, double *eb, hipStream_t s);
void launch_bar_ke(const ElemArgs &a, double *eb, hipStream_t s);
void launch_bar_kg(const ElemArgs &a, const double *u, double *eb, hipStream_t s);
void launch_block_gather(int64_t n_blocks, const int32_t *cptr, const int32_t *cidx, const double *eb,
                         double *vals, hipStream_t s);
// constraint elimination: K'[tgt] += coef K[src] (CSR by target), then the
// dependent entries zpos are cleared; dependent DOFs of ncols columns of x
// (leading dimension ld) recovered from the independent ones
void launch_mpc_apply(int64_t nt, const int64_t *tgt, const int32_t *tptr, const int64_t *src, const double *coef,
                      int64_t nz, const int64_t *zpos, double *K, hipStream_t s);
void launch_mpc_recover(int ndep, const int32_t *dep, const int32_t *ptr, const int32_t *ind, const double *coef,
                        double *x, int64_t ld, int ncols, hipStream_t s);
// a vector with entries on the dependent DOFs folded onto the independent
// ones (v[ind] += coef v[dep] on free receivers), then zeroed off the free set
void launch_mpc_fold(int nrecv, const int32_t *recv, const int32_t *ptr, const int32_t *dep, const double *coef,
                     const uint8_t *free_mask, int ndof, double *v, hipStream_t s);

// SOL 101 result recovery (elements.hip); outputs component-major
// (column c of element e at out[c * n + e]): CQUAD4 44, CTRIA3 20, CBAR 18
// columns.  r = K u - F over the 6x6 BSR matrix.
void launch_bsr_residual(int n_nodes, const int32_t *rp, const int32_t *col, const double *K, const double *u,
                         const double *F, double *r, hipStream_t s);
void launch_q4_recover(const ElemArgs &a, const double *u, double *out, hipStream_t s);
void launch_t3_recover(const ElemArgs &a, const double *u, const double *t3w /* n_t3 */, double *out,
                       hipStream_t s);
void launch_bar_recover(const ElemArgs &a, const double *u, const double *pts /* n_bar x 8 */, double *out,
                        hipStream_t s);

// dlambda/dt kernels (elements.hip)
// terms[3e..3e+2] = phi' dKe/dt phi, phi' KGe phi / t, psi' dKe/dt u (global frame vectors)
void launch_q4_sens_terms(const ElemArgs &a, const double *u, const double *phi, const double *psi,
                          double *terms /* n_q4 x 3 */, hipStream_t s);
// dldt[e] = -(terms[3e] + lam terms[3e+1] - terms[3e+2]) / phiKGphi
void launch_sens_combine(int n_q4, const double *terms, double lam, double phiKGphi, double *dldt, hipStream_t s);
void launch_q4_adjoint_rhs(const ElemArgs &a, const double *u, const double *phi, double lam,
                           double *eg /* n_q4 x 24 */, hipStream_t s);
// the reference's own analytic path (legacy dKe/dt, unit-basis adjoint load;
// six condensed membrane modes when the model has no bars, else four)
void launch_q4_sens_terms_ref(const ElemArgs &a, const double *u, const double *phi, const double *psi,
                              bool six_modes, double *terms, hipStream_t s);
void launch_q4_adjoint_rhs_ref(const ElemArgs &a, const double *u, const double *phi, double lam, bool six_modes,
                               double *eg, hipStream_t s);
void launch_bar_adjoint_rhs(const ElemArgs &a, const double *phi, double lam, double *eg /* n_bar x 12 */,
                            hipStream_t s);
// CTRIA3 rows n_q4 + e of terms: [0], [2] (launch_t3_sens_terms) and [1]
// with the element adjoint loads eg (n_t3 x 18) (launch_t3_kg_sens)
void launch_t3_sens_terms(const ElemArgs &a, const double *u, const double *phi, const double *psi, double *terms,
                          hipStream_t s);
void launch_t3_kg_sens(const ElemArgs &a, const double *u, const double *phi, double lam, double *terms, double *eg,
                       hipStream_t s);
