// factor.hip -- multifrontal sparse Cholesky (FP64) and its triangular solves.
//
// Fronts are dense column-major m x m blocks (m = pivots + update rows).  All
// fronts of one tree height are independent and are processed by the same
// launches: K entries are scattered into the fronts once, children's update
// matrices are extend-added into their parent in a fixed child order (no
// atomics, bit-reproducible), then each front is partially factored with a
// right-looking blocked algorithm (panel width 32):
//   potrf of the 32x32 diagonal block (one workgroup per front),
//   trsm of the panel rows below it, and
//   a 64x64-tiled syrk update of the trailing lower triangle.
// The first ns columns of every front then hold L11 / L21.
//
// The reference factors the same K_ff with CHOLMOD
// (JFEM/src/solver/boundary_conditions.jl:740) and reuses the factor for the
// shift-invert operator (JFEM/src/solver/solve_case.jl:2213-2233).
#include "factor.h"

#include <algorithm>

#include <cmath>

namespace {

constexpr int NB = 32;   // panel width
constexpr int TS = 64;   // syrk tile

__global__ void fill_kernel(double *p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_scatter_kernel(int64_t n, const int32_t *src, const int64_t *dst, const double *kv, double *F) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) F[dst[i]] = kv[src[i]];
}

__global__ void extend_add_kernel(DevFronts D, const int32_t *flist, int slot, double *F) {
    int f = flist[blockIdx.y];
    int cb = D.child_ptr[f], ce = D.child_ptr[f + 1];
    if (slot >= ce - cb) return;
    int c = D.child_idx[cb + slot];
    int nu = D.nu[c], nsc = D.ns[c], mc = D.m[c], mp = D.m[f];
    int64_t total = (int64_t)nu * nu;
    const int32_t *rel = D.rel + D.rel_off[c];
    const double *Fc = F + D.off[c];
    double *Fp = F + D.off[f];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(t % nu), j = (int)(t / nu);
        if (i < j) continue;
        Fp[rel[i] + (int64_t)rel[j] * mp] += Fc[(nsc + i) + (int64_t)(nsc + j) * mc];
    }
}

__global__ void potrf_diag_kernel(DevFronts D, const int32_t *flist, int k0, double *F, double *piv, int *err) {
    int f = flist[blockIdx.x];
    int ns = D.ns[f];
    if (k0 >= ns) return;
    int nb = min(NB, ns - k0), m = D.m[f];
    double *A = F + D.off[f];
    __shared__ double L[NB][NB + 1];
    int tid = threadIdx.x;
    for (int t = tid; t < nb * nb; t += blockDim.x) {
        int i = t % nb, j = t / nb;
        L[i][j] = (i >= j) ? A[(k0 + i) + (int64_t)(k0 + j) * m] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nb; j++) {
        if (tid == 0) {
            double d = L[j][j];
            piv[D.piv0[f] + k0 + j] = d;
            if (!(d > 0.0)) { atomicOr(err, 1); d = 1.0; }
            L[j][j] = sqrt(d);
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < nb; i += blockDim.x) L[i][j] /= L[j][j];
        __syncthreads();
        int w = nb - j - 1;
        for (int t = tid; t < w * w; t += blockDim.x) {
            int ii = j + 1 + t % w, ll = j + 1 + t / w;
            if (ii >= ll) L[ii][ll] -= L[ii][j] * L[ll][j];
        }
        __syncthreads();
    }
    for (int t = tid; t < nb * nb; t += blockDim.x) {
        int i = t % nb, j = t / nb;
        if (i >= j) A[(k0 + i) + (int64_t)(k0 + j) * m] = L[i][j];
    }
}

__global__ void trsm_panel_kernel(DevFronts D, const int32_t *flist, int k0, double *F) {
    int f = flist[blockIdx.y];
    int ns = D.ns[f];
    if (k0 >= ns) return;
    int nb = min(NB, ns - k0), m = D.m[f];
    double *A = F + D.off[f];
    __shared__ double L[NB][NB + 1];
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
        int i = t % NB, j = t / NB;
        L[i][j] = (i < nb && j < nb && i >= j) ? A[(k0 + i) + (int64_t)(k0 + j) * m] : 0.0;
    }
    __syncthreads();
    int i = k0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double x[NB];
#pragma unroll
    for (int l = 0; l < NB; l++) x[l] = (l < nb) ? A[i + (int64_t)(k0 + l) * m] : 0.0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j < nb) {
            double s = x[j];
#pragma unroll
            for (int l = 0; l < j; l++) s -= x[l] * L[j][l];
            x[j] = s / L[j][j];
        }
    }
#pragma unroll
    for (int l = 0; l < NB; l++)
        if (l < nb) A[i + (int64_t)(k0 + l) * m] = x[l];
}

// Trailing update A[i][j] -= sum_l A[i][k0+l] A[j][k0+l] for i >= j >= k1 on
// 64x64 tiles of the lower triangle.  256 threads, each owning a 4 x 4
// micro-tile (rows tx + 16a, columns ty + 16b) fed from LDS copies of the
// two 64 x 32 panel slices.
__global__ void __launch_bounds__(256) syrk_update_kernel(DevFronts D, const int32_t *flist, int k0, double *F) {
    int f = flist[blockIdx.y];
    int ns = D.ns[f];
    if (k0 >= ns) return;
    int nb = min(NB, ns - k0), m = D.m[f];
    int k1 = k0 + nb;
    int n2 = m - k1;
    if (n2 <= 0) return;
    int T = (n2 + TS - 1) / TS;
    int t = blockIdx.x;
    if (t >= T * (T + 1) / 2) return;
    int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) / 2.0);
    while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
    while (ti * (ti + 1) / 2 > t) ti--;
    int tj = t - ti * (ti + 1) / 2;
    int r0 = k1 + ti * TS, c0 = k1 + tj * TS;
    double *A = F + D.off[f];
    __shared__ double Lr[NB][TS + 1];
    __shared__ double Lc[NB][TS + 1];
    for (int q = threadIdx.x; q < NB * TS; q += blockDim.x) {
        int r = q % TS, l = q / TS;
        bool ok = l < nb;
        Lr[l][r] = (ok && r0 + r < m) ? A[(r0 + r) + (int64_t)(k0 + l) * m] : 0.0;
        Lc[l][r] = (ok && c0 + r < m) ? A[(c0 + r) + (int64_t)(k0 + l) * m] : 0.0;
    }
    __syncthreads();
    int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
    for (int l = 0; l < NB; l++) {
        double ra[4], cb[4];
#pragma unroll
        for (int a = 0; a < 4; a++) ra[a] = Lr[l][tx + 16 * a];
#pragma unroll
        for (int b = 0; b < 4; b++) cb[b] = Lc[l][ty + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) acc[a][b] += ra[a] * cb[b];
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
        int j = c0 + ty + 16 * b;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            int i = r0 + tx + 16 * a;
            if (i < m && j < m && i >= j) A[i + (int64_t)j * m] -= acc[a][b];
        }
    }
}

// pivot ratio check of the post-factorisation refix (boundary_conditions.jl:742-782):
// flag elimination indices whose d_i / K_ii < 1e-7.
__global__ void pivot_ratio_kernel(int n, const double *piv, const double *kdiag_elim, int *count, uint8_t *flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double kd = fabs(kdiag_elim[i]);
    double r = kd > 1e-30 ? piv[i] / kd : 1.0;
    uint8_t bad = r < 1e-7;
    flag[i] = bad;
    if (bad) atomicAdd(count, 1);
}

__global__ void gather_elim_kernel(int n_free, int ndof, int nrhs, const int32_t *dof_of_elim, const double *xf,
                                   double *xe) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_free * nrhs) return;
    int e = (int)(t % n_free), r = (int)(t / n_free);
    xe[t] = xf[dof_of_elim[e] + (int64_t)r * ndof];
}

__global__ void scatter_elim_kernel(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof, const double *xe,
                                    double *xf) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ndof * nrhs) return;
    int d = (int)(t % ndof), r = (int)(t / ndof);
    int e = elim_of_dof[d];
    xf[t] = e >= 0 ? xe[e + (int64_t)r * n_free] : 0.0;
}

// Interleaved DOF-order vectors (x[d * nrhs + r]) for the BSR product: the
// nrhs values of one DOF are adjacent, so a block's 6 x nrhs inputs are
// 6 * nrhs contiguous doubles.
__global__ void scatter_elim_il_kernel(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof, const double *xe,
                                       double *xf) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ndof * nrhs) return;
    int d = (int)(t / nrhs), r = (int)(t % nrhs);
    int e = elim_of_dof[d];
    xf[t] = e >= 0 ? xe[e + (int64_t)r * n_free] : 0.0;
}

__global__ void gather_elim_il_kernel(int n_free, int nrhs, const int32_t *dof_of_elim, const double *xf, double *xe) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_free * nrhs) return;
    int e = (int)(t % n_free), r = (int)(t / n_free);
    xe[t] = xf[(int64_t)dof_of_elim[e] * nrhs + r];
}

// y = s * A x for a 6x6 BSR matrix, x and y interleaved (x[d * nrhs + r]).
// 6 * nrhs lanes per block row, lane (ci, r): the block-row values of row ci
// are read by the nrhs lanes of that row together, the inputs x[6J + cj][r]
// by consecutive lanes.
__global__ void __launch_bounds__(192) bsr_spmv_kernel(int n_nodes, const int32_t *rp, const int32_t *col,
                                                       const double *val, const double *x, double *y, int nrhs,
                                                       double s) {
    const int tpr = 6 * nrhs, rpw = max(1, 192 / tpr);
    const int lr = threadIdx.x / tpr, l = threadIdx.x % tpr;
    const int I = blockIdx.x * rpw + lr;
    if (lr >= rpw || I >= n_nodes) return;
    const int ci = l / nrhs, r = l % nrhs;
    double acc = 0.0;
    for (int k = rp[I]; k < rp[I + 1]; k++) {
        const double *v = val + (int64_t)k * 36 + ci * 6;
        const double *xb = x + (int64_t)col[k] * 6 * nrhs + r;
        double a[6], xv[6];
#pragma unroll
        for (int cj = 0; cj < 6; cj++) {
            a[cj] = v[cj];
            xv[cj] = xb[cj * nrhs];
        }
#pragma unroll
        for (int cj = 0; cj < 6; cj++) acc += a[cj] * xv[cj];
    }
    y[((int64_t)I * 6 + ci) * nrhs + r] = s * acc;
}

__global__ void kdiag_elim_kernel(int n_free, const int32_t *dof_of_elim, const int32_t *diag_blk, const double *kv,
                                  double *out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_free) return;
    int d = dof_of_elim[e];
    int c = d % 6;
    out[e] = kv[(int64_t)diag_blk[d / 6] * 36 + c * 6 + c];
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void dev_fill(double *p, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    fill_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
}

void factor_numeric(const FactorPlan &P, const DevFronts &D, const double *kvals, double *F, double *piv, int *err,
                    hipStream_t s) {
    dev_fill(F, P.front_storage, 0.0, s);
    if (P.n_kmap > 0) k_scatter_kernel<<<nblk(P.n_kmap, 256), 256, 0, s>>>(P.n_kmap, D.k_src, D.k_dst, kvals, F);
    for (int h = 0; h <= P.height; h++) {
        const HeightInfo &H = P.heights[h];
        const int32_t *fl = D.flist + H.list_off;
        if (H.max_child_nu2 > 0) {
            unsigned gx = nblk(H.max_child_nu2, 256);
            if (gx > 1024) gx = 1024;
            for (int slot = 0; slot < H.max_children; slot++)
                extend_add_kernel<<<dim3(gx, H.n), 256, 0, s>>>(D, fl, slot, F);
        }
        for (int k0 = 0; k0 < H.max_ns; k0 += NB) {
            potrf_diag_kernel<<<H.n, 256, 0, s>>>(D, fl, k0, F, piv, err);
            // rows below the panel: m - k0 - nb for each front (nb <= NB), so
            // max_m - k0 bounds every front of this height; kernels clip to m.
            int rows = H.max_m - k0;
            if (rows > 0) {
                trsm_panel_kernel<<<dim3(nblk(rows, 256), H.n), 256, 0, s>>>(D, fl, k0, F);
                int T = (rows + TS - 1) / TS;
                syrk_update_kernel<<<dim3((unsigned)(T * (T + 1) / 2), H.n), 256, 0, s>>>(D, fl, k0, F);
            }
        }
    }
}

void launch_gather_elim(int n_free, int ndof, int nrhs, const int32_t *dof_of_elim, const double *xf, double *xe,
                        hipStream_t s) {
    int64_t n = (int64_t)n_free * nrhs;
    if (n > 0) gather_elim_kernel<<<nblk(n, 256), 256, 0, s>>>(n_free, ndof, nrhs, dof_of_elim, xf, xe);
}

void launch_scatter_elim(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof, const double *xe, double *xf,
                         hipStream_t s) {
    int64_t n = (int64_t)ndof * nrhs;
    if (n > 0) scatter_elim_kernel<<<nblk(n, 256), 256, 0, s>>>(n_free, ndof, nrhs, elim_of_dof, xe, xf);
}

void launch_bsr_spmv(int n_nodes, const int32_t *rp, const int32_t *col, const double *val, const double *x, double *y,
                     int nrhs, double scale, hipStream_t s) {
    const int rpw = std::max(1, 192 / (6 * nrhs));
    bsr_spmv_kernel<<<nblk(n_nodes, rpw), 192, 0, s>>>(n_nodes, rp, col, val, x, y, nrhs, scale);
}

void launch_scatter_elim_il(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof, const double *xe, double *xf,
                            hipStream_t s) {
    int64_t n = (int64_t)ndof * nrhs;
    if (n > 0) scatter_elim_il_kernel<<<nblk(n, 256), 256, 0, s>>>(n_free, ndof, nrhs, elim_of_dof, xe, xf);
}

void launch_gather_elim_il(int n_free, int nrhs, const int32_t *dof_of_elim, const double *xf, double *xe,
                           hipStream_t s) {
    int64_t n = (int64_t)n_free * nrhs;
    if (n > 0) gather_elim_il_kernel<<<nblk(n, 256), 256, 0, s>>>(n_free, nrhs, dof_of_elim, xf, xe);
}

void launch_kdiag_elim(int n_free, const int32_t *dof_of_elim, const int32_t *diag_blk, const double *kv, double *out,
                       hipStream_t s) {
    if (n_free > 0) kdiag_elim_kernel<<<nblk(n_free, 256), 256, 0, s>>>(n_free, dof_of_elim, diag_blk, kv, out);
}

void launch_pivot_ratio(int n, const double *piv, const double *kdiag_elim, int *count, uint8_t *flag, hipStream_t s) {
    if (n > 0) pivot_ratio_kernel<<<nblk(n, 256), 256, 0, s>>>(n, piv, kdiag_elim, count, flag);
}
