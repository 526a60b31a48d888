// factor.hip -- multifrontal sparse Cholesky (FP64) and its triangular solves.
//
// Fronts are dense column-major m x m blocks (m = pivots + update rows).  All
// fronts of one tree height are independent and are processed by the same
// launches: K entries are scattered into the fronts once, children's update
// matrices are extend-added into their parent in a fixed child order (no
// atomics, bit-reproducible), then each front is partially factored with a
// right-looking blocked algorithm (panel width 32):
//   potrf of the 32x32 diagonal block (one wavefront per front; after the
//   first panel, by the update that completed the block),
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

typedef double dbl4 __attribute__((ext_vector_type(4)));

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

__device__ __forceinline__ double readlane_d(double x, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Cholesky of an nb x nb diagonal block (nb <= 32) by one wavefront: lane l
// holds row l of the block in registers (a[j], j <= l); the diagonal and
// column j are broadcast with v_readlane (scalar registers), so the 32
// columns need no barrier and no LDS round trip.  Pivots go to pv, a
// non-positive one sets err |= 1.
__device__ __forceinline__ void potrf_regs(double (&a)[NB], int nb, double *pv, int *err) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j < nb) {
            const double d = readlane_d(a[j], j);
            if (l == 0) {
                pv[j] = d;
                if (!(d > 0.0)) atomicOr(err, 1);
            }
            const double sd = sqrt((d > 0.0) ? d : 1.0), rs = 1.0 / sd;
            const double lij = (l == j) ? sd : ((l > j) ? a[j] * rs : 0.0);
            a[j] = lij;
#pragma unroll
            for (int k = j + 1; k < NB; k++) {
                const double lkj = readlane_d(lij, k);
                if (k <= l) a[k] -= lij * lkj;
            }
        }
    }
}

// The first panel's diagonal block of every front of a height (the later
// panels' blocks are factored by the trailing update that completes them).
__global__ void __launch_bounds__(64) potrf_diag_kernel(DevFronts D, const int32_t *flist, int k0, double *F,
                                                        double *piv, int *err) {
    const int f = flist[blockIdx.x];
    const int ns = D.ns[f];
    if (k0 >= ns) return;
    const int nb = min(NB, ns - k0), m = D.m[f];
    double *A = F + D.off[f];
    const int l = threadIdx.x, lr = min(l, nb - 1);
    double a[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) a[j] = A[(k0 + lr) + (int64_t)(k0 + min(j, nb - 1)) * m];
#pragma unroll
    for (int j = 0; j < NB; j++) a[j] = (l < nb && j <= l && j < nb) ? a[j] : 0.0;
    potrf_regs(a, nb, piv + D.piv0[f] + k0, err);
    if (l < nb)
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (j <= l) A[(k0 + l) + (int64_t)(k0 + j) * m] = a[j];
}

// rows below the diagonal block: x L^T = a, one row per thread, the factored
// block and its reciprocal diagonal in LDS
__global__ void trsm_panel_kernel(DevFronts D, const int32_t *flist, int k0, double *F) {
    int f = flist[blockIdx.y];
    int ns = D.ns[f];
    if (k0 >= ns) return;
    int nb = min(NB, ns - k0), m = D.m[f];
    double *A = F + D.off[f];
    __shared__ double L[NB][NB + 1];
    __shared__ double rdiag[NB];
    const int tid = threadIdx.x;
    const int i = k0 + nb + blockIdx.x * blockDim.x + tid;
    // this thread's row is loaded with the block (clamped), before the barrier
    double x[NB];
#pragma unroll
    for (int l = 0; l < NB; l++) x[l] = A[min(i, m - 1) + (int64_t)(k0 + min(l, nb - 1)) * m];
    {
        constexpr int U = NB * NB / 256;
        double tmp[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = tid + 256 * u, r = t % NB, c = t / NB;
            tmp[u] = A[(k0 + min(r, nb - 1)) + (int64_t)(k0 + min(c, nb - 1)) * m];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = tid + 256 * u, r = t % NB, c = t / NB;
            L[r][c] = (r < nb && c < nb && r >= c) ? tmp[u] : 0.0;
            if (r == c) rdiag[r] = (r < nb) ? 1.0 / tmp[u] : 0.0;
        }
    }
    __syncthreads();
    if (i >= m) return;
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j < nb) {
            double sacc = x[j];
#pragma unroll
            for (int l = 0; l < j; l++) sacc -= x[l] * L[j][l];
            x[j] = sacc * rdiag[j];
        }
    }
#pragma unroll
    for (int l = 0; l < NB; l++)
        if (l < nb) A[i + (int64_t)(k0 + l) * m] = x[l];
}

// Trailing update A[i][j] -= sum_l A[i][k0+l] A[j][k0+l] for i >= j >= k1 on
// 64x64 tiles of the lower triangle, on the matrix cores.  The two 64 x 32
// panel slices are staged in LDS; wavefront w takes the 32 x 32 quadrant
// (rows 32 (w & 1), columns 32 (w >> 1)) as 2 x 2 blocks of
// v_mfma_f64_16x16x4f64 over the panel width, computing the transposed block
// D^T[j][i] = sum_l Lc[l][j] Lr[l][i] so that each lane's outputs
// (D^T[l/16 + 4 v][l%16]) run down consecutive rows of the column-major front.
// The old tile is read before the product (indices clamped into the front, no
// conditional read) and parked in LDS, so the read-modify-write costs one
// memory latency per tile however few tiles a CU holds.
// Tile 0 (rows and columns [k1, k1 + 64)) holds the next panel's diagonal
// block once updated: its quadrant (0, 0) is wavefront 0's, which keeps the
// new values in LDS and factors the block there (pivots, error flag), so the
// next panel needs no separate diagonal launch.  No other workgroup of the
// launch reads that block.
__global__ void __launch_bounds__(256) syrk_update_kernel(DevFronts D, const int32_t *flist, int k0, double *F,
                                                          double *piv, int *err) {
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
    __shared__ double lds[2 * NB * (TS + 1)];  // the panels, then the old tile (TS x (TS + 1))
    double (*Lr)[TS + 1] = reinterpret_cast<double (*)[TS + 1]>(lds);
    double (*Lc)[TS + 1] = reinterpret_cast<double (*)[TS + 1]>(lds + NB * (TS + 1));
    double (*Ot)[TS + 1] = reinterpret_cast<double (*)[TS + 1]>(lds);
    const int tid = threadIdx.x;
    double oldr[TS / 4];
    {
        const int64_t i = min(r0 + (tid & 63), m - 1);
#pragma unroll
        for (int u = 0; u < TS / 4; u++) oldr[u] = A[i + (int64_t)min(c0 + (tid >> 6) + 4 * u, m - 1) * m];
    }
    // panel slices: all loads issued together (clamped indices, the masks
    // applied to the loaded values) before the LDS writes
    constexpr int PL = NB * TS / 256;
    double lr[PL], lc[PL];
#pragma unroll
    for (int u = 0; u < PL; u++) {
        const int q = tid + 256 * u, r = q % TS, l = q / TS;
        const int64_t cl = (int64_t)(k0 + min(l, nb - 1)) * m;
        lr[u] = A[min(r0 + r, m - 1) + cl];
        lc[u] = A[min(c0 + r, m - 1) + cl];
    }
#pragma unroll
    for (int u = 0; u < PL; u++) {
        const int q = tid + 256 * u, r = q % TS, l = q / TS;
        Lr[l][r] = (l < nb && r0 + r < m) ? lr[u] : 0.0;
        Lc[l][r] = (l < nb && c0 + r < m) ? lc[u] : 0.0;
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6, kq = lane >> 4, col = lane & 15;
    const int qi = 32 * (wv & 1), qj = 32 * (wv >> 1);
    dbl4 acc[2][2];
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
        for (int ib = 0; ib < 2; ib++) acc[jb][ib] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < NB / 4; kk++) {
        const int l = 4 * kk + kq;
        const double a0 = Lc[l][qj + col], a1 = Lc[l][qj + 16 + col];
        const double b0 = Lr[l][qi + col], b1 = Lr[l][qi + 16 + col];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();  // the panels are no longer read
#pragma unroll
    for (int u = 0; u < TS / 4; u++) Ot[tid & 63][(tid >> 6) + 4 * u] = oldr[u];
    __syncthreads();
    const bool next = t == 0 && k1 < ns;  // workgroup-uniform
    const int nb2 = min(NB, ns - k1);
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
        for (int ib = 0; ib < 2; ib++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int il = qi + 16 * ib + col, jl = qj + 16 * jb + kq + 4 * v;
                const int i = r0 + il, j = c0 + jl;
                const double val = Ot[il][jl] - acc[jb][ib][v];
                const bool diag = next && il < nb2 && jl < nb2;  // written factored below
                if (i < m && j < m && i >= j && !diag) A[i + (int64_t)j * m] = val;
                if (next && wv == 0) Ot[il][jl] = val;  // quadrant (0, 0): wavefront 0's own entries
            }
    if (next && wv == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int l = lane, lr = min(l, nb2 - 1);
        double a[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) a[j] = (l < nb2 && j <= l && j < nb2) ? Ot[lr][min(j, nb2 - 1)] : 0.0;
        potrf_regs(a, nb2, piv + D.piv0[f] + k1, err);
        if (l < nb2)
#pragma unroll
            for (int j = 0; j < NB; j++)
                if (j <= l) A[(k1 + l) + (int64_t)(k1 + j) * m] = a[j];
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
// Both permutations go through a 64-DOF tile in LDS so that each side is
// read or written along its contiguous direction: the elimination-order side
// (column-major, runs of a node's DOFs) with the DOF index fastest, the
// interleaved side (one row of nrhs values per DOF) with the column fastest.
__global__ void __launch_bounds__(256) scatter_elim_il_kernel(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof,
                                                              const double *xe, double *xf) {
    __shared__ double tile[64][17];
    const int d0 = blockIdx.x * 64, t = threadIdx.x, dl = t & 63, rg = t >> 6;
    const int d = min(d0 + dl, ndof - 1);
    const int e = elim_of_dof[d];
    double v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int r = rg + 4 * j;
        v[j] = (e >= 0 && r < nrhs) ? xe[max(e, 0) + (int64_t)min(r, nrhs - 1) * n_free] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) tile[dl][rg + 4 * j] = v[j];
    __syncthreads();
    const int nd = min(64, ndof - d0);
    for (int idx = t; idx < nd * nrhs; idx += 256) xf[(int64_t)d0 * nrhs + idx] = tile[idx / nrhs][idx % nrhs];
}

__global__ void __launch_bounds__(256) gather_elim_il_kernel(int n_free, int nrhs, const int32_t *dof_of_elim,
                                                             const double *xf, double *xe) {
    __shared__ double tile[64][17];
    const int e0 = blockIdx.x * 64, t = threadIdx.x;
    const int ne = min(64, n_free - e0);
    for (int idx = t; idx < ne * nrhs; idx += 256) {
        const int el = idx / nrhs, r = idx % nrhs;
        tile[el][r] = xf[(int64_t)dof_of_elim[e0 + el] * nrhs + r];
    }
    __syncthreads();
    const int el = t & 63, rg = t >> 6;
    if (el < ne)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = rg + 4 * j;
            if (r < nrhs) xe[(e0 + el) + (int64_t)r * n_free] = tile[el][r];
        }
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

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// y = s * A x for 16 right-hand sides: one wavefront per block row, blocks
// taken two at a time.  The wave stages the two blocks' 72 values in LDS
// (column-major inside a block) with two coalesced loads; lane (r, q) then
// takes the (block, column) pairs 3q..3q+2 of the 12, reads x[6J + cj][r]
// (each input of the two blocks is read by exactly one lane) and adds the
// column times that input into six row sums.  The four partial sums of each
// output are added through LDS at the end of the row, and the row's 96
// outputs are written as one contiguous run.  The next pair's values and
// inputs are loaded before the current pair is multiplied.
__global__ void __launch_bounds__(256) bsr_spmv16_kernel(int n_nodes, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ x, double *__restrict__ y,
                                                         double s) {
    __shared__ double sA[4][72];
    __shared__ double sR[4][384];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
    const int I = blockIdx.x * 4 + w;
    if (I >= n_nodes) return;  // the whole wave leaves; no block barrier follows
    const int r = l & 15, q = l >> 4, kk = q >> 1, cb = 3 * (q & 1);
    const int k0 = rp[I], k1 = rp[I + 1];
    double *A = sA[w];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // values of the pair at k (zero past the row end) and this lane's inputs
    auto load = [&](int k, double &a0, double &a1, double (&xv)[3]) {
        const int nv = min(2, k1 - k) * 36;
        const double *v = val + (int64_t)k * 36;
        a0 = v[min(l, nv - 1)];
        a1 = v[min(l + 64, nv - 1)];
        a0 = l < nv ? a0 : 0.0;
        a1 = l + 64 < nv ? a1 : 0.0;
        const int J = col[min(k + kk, k1 - 1)];
        const double *xb = x + (int64_t)J * 96 + cb * 16 + r;
#pragma unroll
        for (int j = 0; j < 3; j++) xv[j] = xb[j * 16];
#pragma unroll
        for (int j = 0; j < 3; j++) xv[j] = k + kk < k1 ? xv[j] : 0.0;
    };
    // value t = kk*36 + ci*6 + cj of a pair goes to A[kk*36 + cj*6 + ci]
    const int t0 = l, t1 = l + 64;
    const int d0 = (t0 / 36) * 36 + (t0 % 6) * 6 + (t0 % 36) / 6;
    const int d1 = (t1 / 36) * 36 + (t1 % 6) * 6 + (t1 % 36) / 6;
    double a0 = 0.0, a1 = 0.0, xv[3] = {0.0, 0.0, 0.0};
    if (k0 < k1) load(k0, a0, a1, xv);
    for (int k = k0; k < k1; k += 2) {
        A[d0] = a0;
        if (t1 < 72) A[d1] = a1;
        wave_sync();
        double cx[3] = {xv[0], xv[1], xv[2]};
        if (k + 2 < k1) load(k + 2, a0, a1, xv);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double *a = A + kk * 36 + (cb + j) * 6;
#pragma unroll
            for (int ci = 0; ci < 6; ci++) acc[ci] += a[ci] * cx[j];
        }
        wave_sync();
    }
    double *R = sR[w];
#pragma unroll
    for (int ci = 0; ci < 6; ci++) R[q * 96 + ci * 16 + r] = acc[ci];
    wave_sync();
    y[(int64_t)I * 96 + l] = s * ((R[l] + R[96 + l]) + (R[192 + l] + R[288 + l]));
    if (l < 32) y[(int64_t)I * 96 + 64 + l] = s * ((R[64 + l] + R[160 + l]) + (R[256 + l] + R[352 + l]));
}

__global__ void kdiag_elim_kernel(int n_free, const int32_t *dof_of_elim, const int32_t *diag_blk, const double *kv,
                                  double *out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_free) return;
    int d = dof_of_elim[e];
    int c = d % 6;
    out[e] = kv[(int64_t)diag_blk[d / 6] * 36 + c * 6 + c];
}

// K + shift I on the 6x6 diagonal blocks (the shifted probes of the
// factorisation AUTOSPC, boundary_conditions.jl:795-801)
__global__ void add_diag_kernel(int n_nodes, const int32_t *diag_blk, double shift, double *kv) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * n_nodes) return;
    const int v = t / 6, c = t % 6;
    kv[(int64_t)diag_blk[v] * 36 + c * 7] += shift;
}

// max |x| into *out (initialised to +0 by the caller); the bit patterns of
// non-negative doubles order like the values, so an integer atomic max works
__global__ void maxabs_kernel(int64_t n, const double *x, unsigned long long *out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_add_diag(int n_nodes, const int32_t *diag_blk, double shift, double *kv, hipStream_t s) {
    if (n_nodes > 0) add_diag_kernel<<<nblk(6 * (int64_t)n_nodes, 256), 256, 0, s>>>(n_nodes, diag_blk, shift, kv);
}

void launch_maxabs(int64_t n, const double *x, double *out, hipStream_t s) {
    dev_fill(out, 1, 0.0, s);
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 2048);
    maxabs_kernel<<<(unsigned)blocks, 256, 0, s>>>(n, x, reinterpret_cast<unsigned long long *>(out));
}

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
            // later panels' diagonal blocks: factored by the previous update
            if (k0 == 0) potrf_diag_kernel<<<H.n, 64, 0, s>>>(D, fl, k0, F, piv, err);
            // rows below the panel: m - k0 - nb for each front (nb <= NB), so
            // max_m - k0 bounds every front of this height; kernels clip to m.
            int rows = H.max_m - k0;
            if (rows > 0) {
                trsm_panel_kernel<<<dim3(nblk(rows, 256), H.n), 256, 0, s>>>(D, fl, k0, F);
                int T = (rows + TS - 1) / TS;
                syrk_update_kernel<<<dim3((unsigned)(T * (T + 1) / 2), H.n), 256, 0, s>>>(D, fl, k0, F, piv, err);
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
    if (nrhs == 16) {
        bsr_spmv16_kernel<<<nblk(n_nodes, 4), 256, 0, s>>>(n_nodes, rp, col, val, x, y, scale);
        return;
    }
    const int rpw = std::max(1, 192 / (6 * nrhs));
    bsr_spmv_kernel<<<nblk(n_nodes, rpw), 192, 0, s>>>(n_nodes, rp, col, val, x, y, nrhs, scale);
}

void launch_scatter_elim_il(int n_free, int ndof, int nrhs, const int32_t *elim_of_dof, const double *xe, double *xf,
                            hipStream_t s) {
    if (ndof > 0) scatter_elim_il_kernel<<<nblk(ndof, 64), 256, 0, s>>>(n_free, ndof, nrhs, elim_of_dof, xe, xf);
}

void launch_gather_elim_il(int n_free, int nrhs, const int32_t *dof_of_elim, const double *xf, double *xe,
                           hipStream_t s) {
    if (n_free > 0) gather_elim_il_kernel<<<nblk(n_free, 64), 256, 0, s>>>(n_free, nrhs, dof_of_elim, xf, xe);
}

void launch_kdiag_elim(int n_free, const int32_t *dof_of_elim, const int32_t *diag_blk, const double *kv, double *out,
                       hipStream_t s) {
    if (n_free > 0) kdiag_elim_kernel<<<nblk(n_free, 256), 256, 0, s>>>(n_free, dof_of_elim, diag_blk, kv, out);
}

void launch_pivot_ratio(int n, const double *piv, const double *kdiag_elim, int *count, uint8_t *flag, hipStream_t s) {
    if (n > 0) pivot_ratio_kernel<<<nblk(n, 256), 256, 0, s>>>(n, piv, kdiag_elim, count, flag);
}
