// solve.hip -- triangular solves with the multifrontal factor as GEMMs.
//
// After the numeric factorization every front f holds [L11; L21] (m x ns).
// Once per factorization we form
//     P_f = [ L11^-1 ; L21 L11^-1 ]      (m x ns, column-major)
// and its transpose PT_f (ns x m, column-major).  Then, per tree height,
//   forward:  v = [x_piv ; 0] + sum_children w_c ;  z = P v_top ;
//             x_piv = z_top ;  w_f = v_bot - z_bot     (w_f goes to the parent)
//   backward: u = [x_piv ; -x(update rows)] ;  x_piv = PT u
// so each direction is one tall-skinny product per height, split into
// 64-row tiles and 128-wide k-slices over many workgroups (memory bound: the
// factor is streamed once per solve for up to MAXR right-hand sides), with a
// fixed-order reduction of the k-slices (bit-reproducible).
#include "factor.h"
#include "solve.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr int TB = 32;        // inversion block
constexpr int RT = SOLVE_RT;  // rows per product tile
constexpr int KS = SOLVE_KS;  // k-slice of the product

typedef double dbl4 __attribute__((ext_vector_type(4)));

// --- inverse of L11 ----------------------------------------------------------
// Every front is final once the factorization has finished, so all fronts'
// inverses are built in two launches: the 32 x 32 diagonal blocks first, then
// each 32-wide column block of L11^-1 by block forward substitution
//     X[i,j] = -inv(L[i,i]) * sum_{k=j}^{i-1} L[i,k] X[k,j]
// (one workgroup per column block, longest chains dispatched first).

// inverse of one 32 x 32 lower-triangular diagonal block, one wavefront:
// lane j keeps column j of the inverse in registers, rows of L broadcast from LDS
__global__ void __launch_bounds__(64) diag_inv_kernel(DevFronts D, DevSolve DS, const double *F, double *P) {
    const int f = DS.lblk[2 * blockIdx.x], b0 = DS.lblk[2 * blockIdx.x + 1];
    const int ns = D.ns[f], m = D.m[f];
    const int nb = min(TB, ns - b0);
    const double *A = F + D.off[f];
    double *Pf = P + DS.poff[f];
    __shared__ double L[TB][TB + 1];
    {
        // all loads issued together (clamped indices), masks applied after
        constexpr int U = TB * TB / 64;
        double tmp[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = threadIdx.x + 64 * u, i = t % TB, j = t / TB;
            tmp[u] = A[(b0 + min(i, nb - 1)) + (int64_t)(b0 + min(j, nb - 1)) * m];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = threadIdx.x + 64 * u, i = t % TB, j = t / TB;
            L[i][j] = (i < nb && j < nb && i >= j) ? tmp[u] : ((i == j) ? 1.0 : 0.0);
        }
    }
    __syncthreads();
    const int j = threadIdx.x;
    double x[TB];
#pragma unroll
    for (int i = 0; i < TB; i++) {
        double sacc = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int l = 0; l < i; l++) sacc -= L[i][l] * x[l];
        x[i] = (i >= j) ? sacc / L[i][i] : 0.0;
    }
    if (j < nb)
#pragma unroll
        for (int i = 0; i < TB; i++)
            if (i < nb) Pf[(b0 + i) + (int64_t)(b0 + j) * m] = x[i];
}

// column block [j0, j0 + 32) of L11^-1 below its diagonal block; rows above it are zero
__global__ void __launch_bounds__(256) linv_col_kernel(DevFronts D, DevSolve DS, const double *F, double *P) {
    const int f = DS.lblk[2 * blockIdx.x], j0 = DS.lblk[2 * blockIdx.x + 1];
    const int ns = D.ns[f], m = D.m[f];
    const int nbj = min(TB, ns - j0);
    const double *A = F + D.off[f];
    double *Pf = P + DS.poff[f];
    __shared__ double Lt[TB][TB + 1], Xt[TB][TB + 1];
    const int tid = threadIdx.x, ti = tid % TB, tq = tid / TB;  // output row, column quad
    for (int t = tid; t < j0 * nbj; t += blockDim.x) {
        int i = t % j0, jj = t / j0;
        Pf[i + (int64_t)(j0 + jj) * m] = 0.0;
    }
    for (int i0 = j0 + TB; i0 < ns; i0 += TB) {
        const int nbi = min(TB, ns - i0);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = j0; k0 < i0; k0 += TB) {
            const int nbk = min(TB, ns - k0);
            {
                // all loads issued together (clamped indices), masks applied after
                constexpr int U = TB * TB / 256;
                double tl[U], tx[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int t = tid + 256 * u, i = t % TB, jj = t / TB;
                    tl[u] = A[(i0 + min(i, nbi - 1)) + (int64_t)(k0 + min(jj, nbk - 1)) * m];
                    tx[u] = Pf[(k0 + min(i, nbk - 1)) + (int64_t)(j0 + min(jj, nbj - 1)) * m];
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int t = tid + 256 * u, i = t % TB, jj = t / TB;
                    Lt[i][jj] = (i < nbi && jj < nbk) ? tl[u] : 0.0;
                    Xt[i][jj] = (i < nbk && jj < nbj) ? tx[u] : 0.0;
                }
            }
            __syncthreads();
#pragma unroll 8
            for (int l = 0; l < TB; l++) {
                const double a = Lt[ti][l];
#pragma unroll
                for (int u = 0; u < 4; u++) acc[u] += a * Xt[l][4 * tq + u];
            }
            __syncthreads();
        }
        // X[i0, j0] = -inv(L[i0,i0]) acc, inv(L[i0,i0]) from the diagonal-block pass
#pragma unroll
        for (int u = 0; u < 4; u++) Xt[ti][4 * tq + u] = acc[u];
        {
            constexpr int U = TB * TB / 256;
            double tl[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = tid + 256 * u, i = t % TB, jj = t / TB;
                tl[u] = Pf[(i0 + min(i, nbi - 1)) + (int64_t)(i0 + min(jj, nbi - 1)) * m];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = tid + 256 * u, i = t % TB, jj = t / TB;
                Lt[i][jj] = (i < nbi && jj < nbi && jj <= i) ? tl[u] : 0.0;
            }
        }
        __syncthreads();
        if (ti < nbi)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = 4 * tq + u;
                if (c < nbj) {
                    double sacc = 0.0;
                    for (int l = 0; l <= ti; l++) sacc += Lt[ti][l] * Xt[l][c];
                    Pf[(i0 + ti) + (int64_t)(j0 + c) * m] = -sacc;
                }
            }
        // the block just written is read back (rows k0 of later steps) only by
        // this workgroup: the barrier's workgroup-scope release/acquire orders
        // it; a device-scope fence here would write back and invalidate L2
        __syncthreads();
    }
}

// Workgroup barrier that orders LDS only (this wavefront's LDS operations
// complete, then s_barrier): global loads in flight stay in flight, where
// __syncthreads() would wait for them.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// P[bottom] = L21 * L11^-1 : 64 x 64 output tiles, k from the tile column
// start (L11^-1 is lower triangular), on the matrix cores.  Wavefront w takes
// the 32 x 32 quadrant (rows 32 (w & 1), columns 32 (w >> 1)) as 2 x 2 blocks
// of v_mfma_f64_16x16x4f64, computing the transposed block
// D^T[j][i] = sum_k Bs[k][j] As[i][k] so that each lane's outputs
// (D^T[l/16 + 4 v][l%16]) run down consecutive rows of the column-major P.
__global__ void __launch_bounds__(256) m21_kernel(DevFronts D, const int32_t *flist, const double *F, double *P,
                                                  const int64_t *poff) {
    int f = flist[blockIdx.z];
    int ns = D.ns[f], m = D.m[f], nu = m - ns;
    int r0 = ns + blockIdx.x * 64, c0 = blockIdx.y * 64;
    if (nu <= 0 || r0 >= m || c0 >= ns) return;
    const double *A = F + D.off[f];
    double *Pf = P + poff[f];
    __shared__ double As[64][33], Bs[32][65];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, kq = lane >> 4, col = lane & 15;
    const int qi = 32 * (wv & 1), qj = 32 * (wv >> 1);
    dbl4 acc[2][2];
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
        for (int ib = 0; ib < 2; ib++) acc[jb][ib] = dbl4{0.0, 0.0, 0.0, 0.0};
    // the next k-step's tiles are loaded (all loads together, clamped
    // indices, masks applied after) while the current one is multiplied
    constexpr int U = 64 * 32 / 256;
    double ta[U], tb[U];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = tid + 256 * u;
            const int r = t % 64, k = t / 64, kk = t % 32, c = t / 32;
            ta[u] = A[min(r0 + r, m - 1) + (int64_t)min(k0 + k, ns - 1) * m];
            tb[u] = Pf[min(k0 + kk, ns - 1) + (int64_t)min(c0 + c, ns - 1) * m];
        }
    };
    load(c0);
    for (int k0 = c0; k0 < ns; k0 += 32) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = tid + 256 * u;
            const int r = t % 64, k = t / 64, kk = t % 32, c = t / 32;
            As[r][k] = (r0 + r < m && k0 + k < ns) ? ta[u] : 0.0;
            Bs[kk][c] = (k0 + kk < ns && c0 + c < ns) ? tb[u] : 0.0;
        }
        lds_barrier();
        if (k0 + 32 < ns) load(k0 + 32);
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            const int k = 4 * kk + kq;
            const double b0 = Bs[k][qj + col], b1 = Bs[k][qj + 16 + col];
            const double a0 = As[qi + col][k], a1 = As[qi + 16 + col][k];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
        }
        lds_barrier();  // the tiles are read; the next loads stay in flight
    }
#pragma unroll
    for (int jb = 0; jb < 2; jb++)
#pragma unroll
        for (int ib = 0; ib < 2; ib++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = r0 + qi + 16 * ib + col, j = c0 + qj + 16 * jb + kq + 4 * v;
                if (i < m && j < ns) Pf[i + (int64_t)j * m] = acc[jb][ib][v];
            }
}

// PT[j + i*ns] = P[i + j*m]  (64 x 64 tiles through LDS)
__global__ void transpose_kernel(DevFronts D, const int32_t *flist, const double *P, double *PT, const int64_t *poff) {
    int f = flist[blockIdx.z];
    int ns = D.ns[f], m = D.m[f];
    int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    if (i0 >= m || j0 >= ns) return;
    const double *Pf = P + poff[f];
    double *Tf = PT + poff[f];
    __shared__ double T[64][65];
    {
        // all loads issued together (clamped indices), masks applied after
        constexpr int U = 64 * 64 / 256;
        double tv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = threadIdx.x + 256 * u, i = t % 64, j = t / 64;
            tv[u] = Pf[min(i0 + i, m - 1) + (int64_t)min(j0 + j, ns - 1) * m];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = threadIdx.x + 256 * u, i = t % 64, j = t / 64;
            T[i][j] = (i0 + i < m && j0 + j < ns) ? tv[u] : 0.0;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 64 * 64; t += blockDim.x) {
        int j = t % 64, i = t / 64;
        if (i0 + i < m && j0 + j < ns) Tf[(j0 + j) + (int64_t)(i0 + i) * ns] = T[i][j];
    }
}

// --- solves ----------------------------------------------------------------

// Working vectors: front f owns m_f rows at W + vec_off[f] * ldv, row-major
// with row stride ldv = MAXR for a block of right-hand sides (one 128-byte
// line per row) and 1 for a single one.  Only the update rows [ns, m) are
// written: w_f, which the parent's forward product adds in.
//
// The vector operand of a product is assembled by each workgroup for its own
// k-slice, in LDS, from
//   forward:  v[k] = x_in[piv0 + k] + the child update rows that the pull map
//             lists for row k (fixed child order);
//   backward: v[k] = x_in[piv0 + k] for k < ns and -x_out[rows[k]] for the
//             update rows (ancestor unknowns, solved by earlier launches).
// x_in and x_out are distinct column-major arrays (ld n_free): while some
// workgroups of a launch read x_in, others write their pivots' x_out.

// v plus the child update rows pulled into vector slot `slot`, column c, in
// the pull map's order.  With the pair table both rows are loaded together
// (slot 0 stands in for a missing one and is masked).
__device__ __forceinline__ double pulled(const DevSolve &DS, const double *W, int ldv, int64_t slot, int c,
                                         double v) {
    if (DS.pull2) {
        const int2 e = reinterpret_cast<const int2 *>(DS.pull2)[slot];
        const double wa = W[(int64_t)max(e.x, 0) * ldv + c], wb = W[(int64_t)max(e.y, 0) * ldv + c];
        v = e.x >= 0 ? v + wa : v;
        return e.y >= 0 ? v + wb : v;
    }
    for (int q = DS.pull_ptr[slot]; q < DS.pull_ptr[slot + 1]; q++) v += W[(int64_t)DS.pull_src[q] * ldv + c];
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8), each with its own L2.  This bijection hands every XCD a
// contiguous run of logical tiles, so the row tiles of one front (which all
// read that front's V slice) share an L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n / 8, r = n % 8, x = b % 8, i = b / 8;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

__global__ void __launch_bounds__(256) prod1_kernel(DevFronts D, DevSolve DS, const int32_t *tiles, const double *A,
                                                    double *W, double *part, int bwd, const int32_t *kwid,
                                                    const int32_t *nks, const double *xin, double *xout, int n_free) {
    constexpr int NR = 1, nrhs = 1;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    int f = tiles[3 * tile], rt = tiles[3 * tile + 1], ks = tiles[3 * tile + 2];
    int ns = D.ns[f], m = D.m[f], piv0 = D.piv0[f];
    int nrows = bwd ? ns : m;
    int inner = bwd ? m : ns;
    int lda = bwd ? ns : m;
    const double *Af = A + DS.poff[f];
    const int64_t slot0 = D.vec_off[f];
    double *V = W + slot0 * nrhs;
    const int kw = kwid[f];
    const int KW = (kw + 3) / 4;  // k per wavefront
    int k0 = ks * kw, k1 = min(inner, k0 + kw);
    // the vector slice and, after the products, the cross-wavefront partial
    // sums share one LDS buffer
    constexpr int VS_N = SOLVE_KMAX * NR, RED_N = 3 * RT * (NR + 1);
    __shared__ double lds[VS_N > RED_N ? VS_N : RED_N];
    double (*Vs)[NR] = reinterpret_cast<double (*)[NR]>(lds);
    double (*red)[RT][NR + 1] = reinterpret_cast<double (*)[RT][NR + 1]>(lds);
    const int32_t *rows = D.rows + D.row_off[f];
    for (int k = threadIdx.x; k < k1 - k0; k += blockDim.x) {
        const int kk = k0 + k;
        double v;
        if (bwd) v = (kk < ns) ? xin[piv0 + kk] : -xout[rows[kk]];
        else v = pulled(DS, W, 1, slot0 + kk, 0, xin[piv0 + kk]);
        Vs[k][0] = v;
    }
    __syncthreads();
    int row = threadIdx.x % RT, g = threadIdx.x / RT;
    int i = rt * RT + row;
    int kb = k0 + g * KW, ke = min(k1, kb + KW);
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = 0.0;
    if (i < nrows && kb < ke) {
        const double *a = Af + i + (int64_t)kb * lda;
        int k = kb;
        for (; k + 16 <= ke; k += 16) {
            double av[16];
#pragma unroll
            for (int u = 0; u < 16; u++) av[u] = a[(int64_t)(k - kb + u) * lda];
#pragma unroll
            for (int u = 0; u < 16; u++)
#pragma unroll
                for (int r = 0; r < NR; r++) acc[r] += av[u] * Vs[k - k0 + u][r];
        }
        for (; k < ke; k++) {
            double av = a[(int64_t)(k - kb) * lda];
#pragma unroll
            for (int r = 0; r < NR; r++) acc[r] += av * Vs[k - k0][r];
        }
    }
    __syncthreads();  // every wavefront is done with Vs
    if (g > 0) {
#pragma unroll
        for (int r = 0; r < NR; r++) red[g - 1][row][r] = acc[r];
    }
    __syncthreads();
    if (g == 0 && i < nrows) {
        if (nks[f] == 1) {
            // single k-slice: the sum is final.  Nothing else in this launch
            // reads this front's pivots of x_out or its update rows.
            for (int r = 0; r < nrhs; r++) {
                double s = acc[r] + red[0][row][r] + red[1][row][r] + red[2][row][r];
                if (bwd || i < ns) xout[(piv0 + i) + (int64_t)r * n_free] = s;
                else V[(int64_t)i * nrhs + r] = pulled(DS, W, 1, slot0 + i, r, 0.0) - s;
            }
        } else {
            for (int r = 0; r < nrhs; r++) {
                double s = acc[r] + red[0][row][r] + red[1][row][r] + red[2][row][r];
                part[((int64_t)tile * RT + row) * MAXR + r] = s;
            }
        }
    }
}

// The same product for a block of right-hand sides (row stride MAXR) on the
// matrix cores, v_mfma_f64_16x16x4f64 (D[16x16] += A[16x4] B[4x16]; lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and holds D[l/16 + 4v][l%16],
// v = 0..3).  Wavefront w owns rows [16 w, 16 w + 16) of the 64-row tile
// over the whole k-slice, so no partial sums cross wavefronts: the factor is
// read in the column-major front layout (128 contiguous bytes per block
// column), the B values (slice row k + l/16, column l%16) come from the
// slice in LDS, and a chunk of PCH four-k steps (PCH factor loads per lane)
// is issued before its products; the first chunk goes out before the vector
// slice is assembled, so the slice's loads overlap it.  Results go out from
// the accumulator layout: update rows and partial sums are row-major (one
// 128-byte row per 16 lanes); rows of the column-major x_out pass through a
// 16 x 16 transpose in LDS so that each column's 16 rows leave together.
// The factor loads are not masked: rows past the front re-read the front's
// last row (cache hits; they reach only discarded outputs), steps past the
// slice re-read its last step (B is zero there), and the last step's up to 3
// k past the slice read factor data of later fronts or the zero tail of the
// buffer (p_pad), finite either way and met by a zero B value.
// A front cut in several k-slices writes partial sums, which the finish
// kernels add in slice order.  (Reducing them in the last slice to finish,
// behind a device-scope release and acquire per wavefront, measured slower:
// each such release writes back the XCD's L2.)
// PCH: four-k steps per chunk of factor loads (8: more wavefronts per SIMD
// for short slices on large grids; 16 or 32: fewer dependent rounds)
// KG: k-groups of four wavefronts per workgroup (256 KG threads).  The
// wavefronts of k-group g take the g-th contiguous part of the slice's
// four-k steps for their 16 rows; k-groups 1..KG-1 leave their partial sums
// in LDS and k-group 0 adds them in group order (fixed order, bit-
// reproducible for a given KG) before the write-back.  Long slices on
// launches with few tiles (near the root) finish in 1 / KG of the time.
template <int PCH, bool FPULL, int KG>
__global__ void __launch_bounds__(256 * KG) prod_mfma_kernel(DevFronts D, DevSolve DS, const TileDesc *desc,
                                                        const double *A, double *W, double *part, int nrhs, int bwd,
                                                        const double *xin, double *xout, int n_free) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const TileDesc T = desc[tile];  // workgroup-uniform: one scalar load
    const int rt = T.rt, ks = T.ks, ns = T.ns, m = T.m, piv0 = T.piv0;
    const int nrows = bwd ? ns : m, inner = bwd ? m : ns, lda = bwd ? ns : m;
    const bool whole = T.nks == 1;
    const int64_t slot0 = T.vec_off;
    double *V = W + slot0 * MAXR;
    const int kw = T.kw;
    const int k0 = ks * kw, kl = min(inner, k0 + kw) - k0;
    const int nst = (kl + 3) >> 2;  // four-k steps of the slice
    const int wall = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wv = wall & 3, kg = wall >> 2;     // row block, k-group
    const int t = threadIdx.x, lane = t & 63, kq = lane >> 4, col = lane & 15;
    const int r0 = rt * RT, rb = r0 + 16 * wv;  // this wavefront's first row
    const bool live = rb < nrows;                // wavefront-uniform
    const double *Aw = A + T.poff + (int64_t)k0 * lda;
    const unsigned aoff = (unsigned)(min(rb + (lane & 15), nrows - 1) + kq * lda);
    // L11^-1 is lower triangular: forward pivot rows i use columns k <= i,
    // backward rows use k >= i below ns.  Four-k steps that are zero for all
    // 16 rows of the wavefront are skipped (they would add exact zeros):
    // steps [s_lo, s_hi) remain.
    int s_lo = 0, s_hi = nst;
    if (!bwd && rb + 16 <= ns) s_hi = rb + 15 < k0 ? 0 : min(nst, ((rb + 15 - k0) >> 2) + 1);
    if (bwd && rb > k0) s_lo = min(nst, (rb - k0) >> 2);
    if (KG > 1) {
        // this k-group's contiguous part of the steps
        const int per = (max(s_hi - s_lo, 0) + KG - 1) / KG;
        const int g_lo = min(s_hi, s_lo + kg * per);
        s_hi = min(s_hi, g_lo + per);
        s_lo = g_lo;
    }
    // Loads go out in dependency order so that every wait covers only what it
    // needs (vector loads complete in order): first the index loads (pull
    // pairs of the forward slice rows and of this wavefront's update rows,
    // ancestor row numbers of the backward slice), then the first chunk of
    // factor values and the right-hand-side values, then the loads that
    // depend on the indices (child update rows, ancestor values).
    const int2 *pairs = reinterpret_cast<const int2 *>(DS.pull2);
    const bool parent = T.parent != 0;
    // FPULL: a forward launch above the leaves (every front has children)
    const bool fpull = FPULL && parent && pairs;  // forward slice rows add child rows (pair table)
    const int kpad = 4 * nst;                     // slice rows, <= 4 * ceil(SOLVE_KMAX / 4) < 512
    const int32_t *rows = D.rows + T.row_off;
    int2 ep[2] = {{-1, -1}, {-1, -1}};
    int arow[2] = {0, 0};
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int k = t + 256 * KG * j, kk = k0 + min(k, kl - 1);
        if (fpull && k < kl) ep[j] = pairs[slot0 + k0 + k];
        if (bwd && k < kpad && kk >= ns) arow[j] = rows[kk];
    }
    // forward, single k-slice: the child rows pulled into this wavefront's
    // update rows (written back below)
    const bool uprows = FPULL && whole && parent && live && kg == 0 && rb + 16 > ns;
    int2 eu[4] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};
    if (uprows && pairs)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = rb + kq + 4 * v;
            if (i >= ns && i < nrows && col < nrhs) eu[v] = pairs[slot0 + i];
        }
    // factor values in two register sets: the next chunk's loads are in
    // flight while the current chunk is multiplied
    double av[PCH], an[PCH];
    auto load_a = [&](double (&dst)[PCH], int s) {
#pragma unroll
        for (int u = 0; u < PCH; u++) dst[u] = Aw[(int64_t)(4 * min(s + u, s_hi - 1)) * lda + aoff];
    };
    if (live && s_lo < s_hi) load_a(av, s_lo);
    // child update rows (128 contiguous bytes each, in the pull map's order)
    double upd[4] = {0.0, 0.0, 0.0, 0.0};
    if (uprows) {
        if (pairs) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double wa = W[(int64_t)max(eu[v].x, 0) * MAXR + col], wb = W[(int64_t)max(eu[v].y, 0) * MAXR + col];
                upd[v] = eu[v].x >= 0 ? upd[v] + wa : upd[v];
                upd[v] = eu[v].y >= 0 ? upd[v] + wb : upd[v];
            }
        } else {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = rb + kq + 4 * v;
                if (i >= ns && i < nrows && col < nrhs) upd[v] = pulled(DS, W, MAXR, slot0 + i, col, 0.0);
            }
        }
    }
    // LDS (dynamic, sized for the launch's widest slice): the vector slice
    // (4 nst rows, padded columns), then one 16 x 17 transpose tile per wavefront
    extern __shared__ double lds[];
    double (*Vs)[MAXR + 1] = reinterpret_cast<double (*)[MAXR + 1]>(lds);
    // the slice rows of this thread (t, and t + 256 when the slice is wider):
    // right-hand side or ancestor values, k fastest (x columns are
    // contiguous), plus the pulled child rows; rows past the slice and
    // columns past nrhs are zero
#pragma unroll 1
    for (int j = 0; j < 2; j++) {
        const int k = t + 256 * KG * j;
        if (k >= kpad) break;
        const int kk = k0 + min(k, kl - 1);
        const bool anc = bwd && kk >= ns;
        const double *src = anc ? xout + (j ? arow[1] : arow[0]) : xin + piv0 + kk;
        const double sg = anc ? -1.0 : 1.0;
        double xv[MAXR];
#pragma unroll
        for (int c = 0; c < MAXR; c++) xv[c] = src[(int64_t)min(c, nrhs - 1) * n_free];
        if (fpull) {
            const int2 e = j ? ep[1] : ep[0];
            const double *wa = W + (int64_t)max(e.x, 0) * MAXR, *wb = W + (int64_t)max(e.y, 0) * MAXR;
#pragma unroll
            for (int h = 0; h < MAXR; h += 8) {
                double xa[8], xb[8];
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    xa[c] = wa[h + c];
                    xb[c] = wb[h + c];
                }
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    double v = xv[h + c];
                    v = e.x >= 0 ? v + xa[c] : v;
                    v = e.y >= 0 ? v + xb[c] : v;
                    Vs[k][h + c] = (k < kl && h + c < nrhs) ? v : 0.0;
                }
            }
        } else if (FPULL && parent) {
            // a row with more than two child contributions somewhere: the
            // general pull map
#pragma unroll
            for (int c = 0; c < MAXR; c++) {
                double v = xv[c];
                if (k < kl && c < nrhs) v = pulled(DS, W, MAXR, slot0 + k0 + k, c, v);
                Vs[k][c] = (k < kl && c < nrhs) ? v : 0.0;
            }
        } else {
#pragma unroll
            for (int c = 0; c < MAXR; c++) Vs[k][c] = (k < kl && c < nrhs) ? sg * xv[c] : 0.0;
        }
    }
    __syncthreads();
    if (KG == 1 && !live) return;  // a whole wavefront past the front; no barrier follows
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    auto mult = [&](const double (&a)[PCH], int s) {
        double bv[PCH];
#pragma unroll
        for (int u = 0; u < PCH; u++) {
            const int ss = min(s + u, s_hi - 1);
            bv[u] = (s + u < s_hi) ? Vs[4 * ss + kq][col] : 0.0;  // wavefront-uniform condition
        }
#pragma unroll
        for (int u = 0; u < PCH; u++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bv[u], acc, 0, 0, 0);
    };
    // chunks alternate between the two register sets, each chunk's loads
    // issued one chunk ahead
    if (live)
        for (int s = s_lo; s < s_hi; s += 2 * PCH) {
            if (s + PCH < s_hi) load_a(an, s + PCH);
            mult(av, s);
            if (s + 2 * PCH < s_hi) load_a(av, s + 2 * PCH);
            if (s + PCH < s_hi) mult(an, s + PCH);
        }
    if (KG > 1) {
        // partial sums of k-groups 1..KG-1 (in the slice's LDS, which every
        // wavefront has finished reading), added by k-group 0 in group order
        double (*red)[64][MAXR] = reinterpret_cast<double (*)[64][MAXR]>(lds);
        __syncthreads();
        if (kg > 0 && live)
#pragma unroll
            for (int v = 0; v < 4; v++) red[kg - 1][16 * wv + kq + 4 * v][col] = acc[v];
        __syncthreads();
        if (kg > 0 || !live) return;  // no barrier follows
#pragma unroll
        for (int g = 0; g < KG - 1; g++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[v] += red[g][16 * wv + kq + 4 * v][col];
    }
    // lane holds D[kq + 4 v][col]: row rb + kq + 4 v, right-hand side col
    if (!whole) {
#pragma unroll
        for (int v = 0; v < 4; v++)
            part[((int64_t)tile * RT + 16 * wv + kq + 4 * v) * MAXR + col] = acc[v];
        return;
    }
    if (!bwd && rb + 16 > ns)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = rb + kq + 4 * v;
            if (i >= ns && i < nrows && col < nrhs) V[(int64_t)i * MAXR + col] = upd[v] - acc[v];
        }
    if (rb < (bwd ? nrows : min(nrows, ns))) {
        // pivot rows to x_out through the wavefront's transpose tile
        // past the slice and the k-groups' partial sums (both may still be read)
        const size_t ts_off = max((size_t)4 * nst * (MAXR + 1), (size_t)(KG - 1) * 64 * MAXR);
        double (*Ts)[MAXR + 1] = reinterpret_cast<double (*)[MAXR + 1]>(lds + ts_off) + 16 * wv;
#pragma unroll
        for (int v = 0; v < 4; v++) Ts[kq + 4 * v][col] = acc[v];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int r = lane & 15, i = rb + r;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c = (lane >> 4) + 4 * j;
            if (i < nrows && c < nrhs && (bwd || i < ns)) xout[(piv0 + i) + (int64_t)c * n_free] = Ts[r][c];
        }
    }
}

// fixed-order reduction of the k-slices and write-back
__global__ void fwd_finish_kernel(DevFronts D, DevSolve DS, const int32_t *flist, const int32_t *tbase,
                                  const int32_t *nks, const double *part, double *W, double *xout, int n_free,
                                  int nrhs, int ldv) {
    int f = flist[blockIdx.y];
    int ns = D.ns[f], m = D.m[f], piv0 = D.piv0[f];
    const int64_t slot0 = D.vec_off[f];
    double *V = W + slot0 * ldv;
    int nk = nks[f];
    if (nk == 1) return;  // written by the product kernel
    int64_t base = tbase[f];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m * nrhs; t += gridDim.x * blockDim.x) {
        int i = t / nrhs, r = t % nrhs;
        int rt = i / RT, rr = i % RT;
        double s = 0.0;
        for (int k = 0; k < nk; k++) s += part[((base + (int64_t)rt * nk + k) * RT + rr) * MAXR + r];
        if (i < ns) xout[(piv0 + i) + (int64_t)r * n_free] = s;
        else V[(int64_t)i * ldv + r] = pulled(DS, W, ldv, slot0 + i, r, 0.0) - s;
    }
}

__global__ void bwd_finish_kernel(DevFronts D, const int32_t *flist, const int32_t *tbase, const int32_t *nks,
                                  const double *part, double *xout, int n_free, int nrhs) {
    int f = flist[blockIdx.y];
    int ns = D.ns[f], piv0 = D.piv0[f];
    int nk = nks[f];
    if (nk == 1) return;  // written by the product kernel
    int64_t base = tbase[f];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ns * nrhs; t += gridDim.x * blockDim.x) {
        int i = t % ns, r = t / ns;
        int rt = i / RT, rr = i % RT;
        double s = 0.0;
        for (int k = 0; k < nk; k++) s += part[((base + (int64_t)rt * nk + k) * RT + rr) * MAXR + r];
        xout[(piv0 + i) + (int64_t)r * n_free] = s;
    }
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

void solve_prepare(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S, const double *F,
                   double *Pbuf, double *PTbuf, hipStream_t s) {
    if (DS.n_lblk > 0) {
        diag_inv_kernel<<<DS.n_lblk, 64, 0, s>>>(D, DS, F, Pbuf);
        linv_col_kernel<<<DS.n_lblk, 256, 0, s>>>(D, DS, F, Pbuf);
    }
    for (int h = 0; h <= P.height; h++) {
        const HeightInfo &H = P.heights[h];
        const int32_t *fl = D.flist + H.list_off;
        if (H.n == 0) continue;
        // update rows of each front: m - ns <= max_m (kernel clips)
        m21_kernel<<<dim3(cdiv(H.max_m, 64), cdiv(H.max_ns, 64), H.n), 256, 0, s>>>(D, fl, F, Pbuf, DS.poff);
        transpose_kernel<<<dim3(cdiv(H.max_m, 64), cdiv(H.max_ns, 64), H.n), 256, 0, s>>>(D, fl, Pbuf, PTbuf, DS.poff);
    }
}

void ProdTimer::begin(hipStream_t s, double bytes, int tag) {
    while (pool.size() < used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        pool.push_back(e);
    }
    (void)hipEventRecord(pool[used], s);
    pending_bytes.push_back(bytes);
    pending_tag.push_back(tag);
}

void ProdTimer::end(hipStream_t s) {
    if (pool.size() < used + 2) return;
    (void)hipEventRecord(pool[used + 1], s);
    used += 2;
}

void ProdTimer::collect() {
    for (size_t i = 0; i + 1 < used; i += 2) {
        float ms = 0.f;
        const int tag = pending_tag[i / 2];
        (void)hipEventSynchronize(pool[i + 1]);
        if (hipEventElapsedTime(&ms, pool[i], pool[i + 1]) == hipSuccess) total_ms[tag] += ms;
        total_bytes[tag] += pending_bytes[i / 2];
        launches[tag]++;
    }
    used = 0;
    pending_bytes.clear();
    pending_tag.clear();
}

void ProdTimer::reset() {
    collect();
    total_ms[0] = total_ms[1] = total_bytes[0] = total_bytes[1] = 0.0;
    launches[0] = launches[1] = 0;
}

ProdTimer::~ProdTimer() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

// LDS bytes of a block-RHS product launch whose widest k-slice is kmax: the
// slice rows (rounded up to four-k steps), then four 16 x 17 transpose tiles
static size_t prod_lds(int kmax, int kg) {
    const size_t slice = (size_t)4 * ((kmax + 3) / 4) * (MAXR + 1), red = (size_t)(kg - 1) * 64 * MAXR;
    return sizeof(double) * (std::max(slice, red) + (size_t)4 * 16 * (MAXR + 1));
}

// one block-RHS product launch over n_tiles tiles whose widest k-slice is
// kmax, pch four-k steps per chunk of factor loads (0: by the launch's size)
template <bool FPULL, int KG>
static void launch_prod_k(int n_tiles, int kmax, int pch, hipStream_t s, const DevFronts &D, const DevSolve &DS,
                          const TileDesc *td, const double *A, double *W, double *part, int nrhs, int bwd,
                          const double *xin, double *xout, int n_free) {
    const size_t lds = prod_lds(kmax, KG);
    if (pch == 8)
        prod_mfma_kernel<8, FPULL, KG><<<n_tiles, 256 * KG, lds, s>>>(D, DS, td, A, W, part, nrhs, bwd, xin, xout,
                                                                      n_free);
    else if (pch == 32 && KG == 1)
        prod_mfma_kernel<32, FPULL, 1><<<n_tiles, 256, lds, s>>>(D, DS, td, A, W, part, nrhs, bwd, xin, xout, n_free);
    else
        prod_mfma_kernel<16, FPULL, KG><<<n_tiles, 256 * KG, lds, s>>>(D, DS, td, A, W, part, nrhs, bwd, xin, xout,
                                                                       n_free);
}

template <bool FPULL>
static void launch_prod_t(int n_tiles, int kmax, int pch, int kg, hipStream_t s, const DevFronts &D,
                          const DevSolve &DS, const TileDesc *td, const double *A, double *W, double *part, int nrhs,
                          int bwd, const double *xin, double *xout, int n_free) {
    if (kg == 4)
        launch_prod_k<FPULL, 4>(n_tiles, kmax, pch, s, D, DS, td, A, W, part, nrhs, bwd, xin, xout, n_free);
    else if (kg == 2)
        launch_prod_k<FPULL, 2>(n_tiles, kmax, pch, s, D, DS, td, A, W, part, nrhs, bwd, xin, xout, n_free);
    else
        launch_prod_k<FPULL, 1>(n_tiles, kmax, pch, s, D, DS, td, A, W, part, nrhs, bwd, xin, xout, n_free);
}

// pull: a forward launch above the leaves (its fronts add child update rows);
// kg: k-groups per workgroup (1, 2, 4)
static void launch_prod(int n_tiles, int kmax, int pch, int kg, bool pull, hipStream_t s, const DevFronts &D,
                        const DevSolve &DS, const TileDesc *td, const double *A, double *W, double *part, int nrhs,
                        int bwd, const double *xin, double *xout, int n_free) {
    if (pch == 0) pch = (kmax <= 192 && n_tiles >= 380) ? 8 : 16;
    if (pull && !bwd)
        launch_prod_t<true>(n_tiles, kmax, pch, kg, s, D, DS, td, A, W, part, nrhs, bwd, xin, xout, n_free);
    else
        launch_prod_t<false>(n_tiles, kmax, pch, kg, s, D, DS, td, A, W, part, nrhs, bwd, xin, xout, n_free);
}

// algorithmic bytes of one product launch: the factor's nonzero entries
// (the lower triangle of L11^-1 and L21 L11^-1), vector rows in and out
static double prod_bytes(const SolveHeight &SH, int nrhs) {
    return 8.0 * ((double)SH.sum_nz + (double)nrhs * (double)(SH.sum_m + SH.sum_ns));
}

hipError_t solve_height(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S, int bwd,
                        int h, const double *A, double *W, double *part, const double *xin, double *xout, int nrhs,
                        hipStream_t s, ProdTimer *pt) {
    if (xin == xout || nrhs < 1 || nrhs > MAXR || h < 0 || h > P.height) return hipErrorInvalidValue;
    const HeightInfo &H = P.heights[h];
    const SolveHeight &SH = S.heights[h];
    if (H.n == 0) return hipSuccess;
    const int32_t *fl = D.flist + H.list_off;
    const bool timed = pt && nrhs > 1;
    if (timed) pt->begin(s, prod_bytes(SH, nrhs), h == 0 ? 0 : 1);
    if (!bwd) {
        const int32_t *tl = DS.fwd_tiles + 3 * SH.fwd_off;
        if (nrhs == 1)
            prod1_kernel<<<SH.n_fwd, 256, 0, s>>>(D, DS, tl, A, W, part, 0, DS.fwd_kw, DS.fwd_nks, xin, xout,
                                                  P.n_free);
        else
            launch_prod(SH.n_fwd, SH.fwd_kmax, SH.fwd_cfg.pch, SH.fwd_cfg.kg, h > 0, s, D, DS,
                        DS.fwd_desc + SH.fwd_off, A, W, part, nrhs, 0, xin, xout, P.n_free);
        if (timed) pt->end(s);
        // single-slice fronts were finished by the product itself
        if (SH.fwd_split)
            fwd_finish_kernel<<<dim3(cdiv((int64_t)H.max_m * nrhs, 256), H.n), 256, 0, s>>>(
                D, DS, fl, DS.fwd_tbase_rel, DS.fwd_nks, part, W, xout, P.n_free, nrhs, nrhs > 1 ? MAXR : 1);
    } else {
        const int32_t *tl = DS.bwd_tiles + 3 * SH.bwd_off;
        if (nrhs == 1)
            prod1_kernel<<<SH.n_bwd, 256, 0, s>>>(D, DS, tl, A, W, part, 1, DS.bwd_kw, DS.bwd_nks, xin, xout,
                                                  P.n_free);
        else
            launch_prod(SH.n_bwd, SH.bwd_kmax, SH.bwd_cfg.pch, SH.bwd_cfg.kg, false, s, D, DS,
                        DS.bwd_desc + SH.bwd_off, A, W, part, nrhs, 1, xin, xout, P.n_free);
        if (timed) pt->end(s);
        if (SH.bwd_split)
            bwd_finish_kernel<<<dim3(cdiv((int64_t)H.max_ns * nrhs, 256), H.n), 256, 0, s>>>(
                D, fl, DS.bwd_tbase_rel, DS.bwd_nks, part, xout, P.n_free, nrhs);
    }
    return hipSuccess;
}

hipError_t solve_fwd(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S,
                     const double *Pbuf, double *W, double *part, const double *xin, double *xout, int nrhs,
                     hipStream_t s, ProdTimer *pt) {
    if (xin == xout || nrhs < 1 || nrhs > MAXR) return hipErrorInvalidValue;
    for (int h = 0; h <= P.height; h++) {
        const hipError_t e = solve_height(P, D, DS, S, 0, h, Pbuf, W, part, xin, xout, nrhs, s, pt);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t solve_bwd(const FactorPlan &P, const DevFronts &D, const DevSolve &DS, const SolvePlan &S,
                     const double *PTbuf, double *W, double *part, const double *xin, double *xout, int nrhs,
                     hipStream_t s, ProdTimer *pt) {
    if (xin == xout || nrhs < 1 || nrhs > MAXR) return hipErrorInvalidValue;
    for (int h = P.height; h >= 0; h--) {
        const hipError_t e = solve_height(P, D, DS, S, 1, h, PTbuf, W, part, xin, xout, nrhs, s, pt);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
