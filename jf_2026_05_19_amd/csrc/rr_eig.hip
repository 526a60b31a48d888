// rr_eig.hip -- the Rayleigh-Ritz step of the block Lanczos eigensolver:
// the k eigenpairs of largest |theta| of the symmetric projected matrix
// (m x m, m <= 1024), in one 1024-thread workgroup.
//   1. Householder tridiagonalisation (Golub & Van Loan, Matrix Computations,
//      4th ed., Alg. 5.1.1 `house` and Alg. 8.3.1).  The matrix stays in
//      global memory (L2); every wavefront handles CPI columns per round
//      with all of its loads issued before any use, so one L2 round trip
//      covers CPI columns of the symmetric matrix-vector product and of the
//      rank-2 update;
//   2. eigenvalues of the tridiagonal matrix at the 2k extreme indices by
//      Sturm-sequence multisection (Wilkinson, The Algebraic Eigenvalue
//      Problem, ch. 5): G lanes per index evaluate G interior points per
//      round, a ballot picks the subinterval (log2(G+1) bits per round);
//   3. eigenvectors by inverse iteration with a partially pivoted
//      tridiagonal LU (Golub & Van Loan sec. 4.3), one lane per cluster of
//      close eigenvalues with its LU factors and vector in LDS, modified
//      Gram-Schmidt inside a cluster (the scheme of Wilkinson and of
//      LAPACK's stein);
//   4. back-transformation of the k vectors (held in LDS) with the stored
//      reflectors, the next reflector prefetched while the current one is
//      applied.
// Only k vectors are formed, so the cost is the O(m^3) reduction of step 1
// plus O(k m^2); the large cluster of near-zero Ritz values that a Lanczos
// basis carries does not slow any step down.
#include "dense.h"

#include <cfloat>
#include <cmath>

namespace {

constexpr int NT = 1024;
constexpr int NW = NT / 64;
constexpr double EPS = DBL_EPSILON;
constexpr int SCR = 12288;  // LDS scratch (doubles) for steps 3 and 4

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ double block_sum(double v, double *wred) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) wred[wid] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < NW; i++) r += wred[i];
    __syncthreads();
    return r;
}

// number of eigenvalues of the tridiagonal (d, e2 = e^2) below x
__device__ int sturm_count(int m, const double *d, const double *e2, double x, double pivmin) {
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int c = (q < 0.0);
    for (int i = 1; i < m; i++) {
        q = d[i] - x - e2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        c += (q < 0.0);
    }
    return c;
}

__device__ __forceinline__ double hash_unit(uint64_t k) {
    uint64_t z = k + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// RPL: rows per lane of a column (m <= 64 RPL + 1); CPI: columns per wavefront round
// Batched launches: problem list[blockIdx.x] at the given strides (list ==
// nullptr: the single problem at the base pointers).
struct SymeigBatch {
    const int *list;
    int64_t sA, sT, sZ, sW;
};

template <int RPL, int CPI>
__global__ void __launch_bounds__(NT) symeig_top_kernel(int m, double *A, int k, double *theta, double *Zs,
                                                        double *work, int *info, SymeigBatch bt) {
    if (bt.list) {
        const int64_t p = bt.list[blockIdx.x];
        A += p * bt.sA;
        theta += p * bt.sT;
        Zs += p * bt.sZ;
        work += p * bt.sW;
        info += p;
    }
    __shared__ double vsh[NT];
    __shared__ double psh[NT];
    __shared__ double dd[NT], ee[NT], e2[NT];
    __shared__ double scr[SCR];
    __shared__ double ev[2 * SYMEIG_MAX_K];  // eigenvalues at the extreme indices
    __shared__ int sel[SYMEIG_MAX_K];        // chosen index, descending |theta|
    __shared__ int asc[SYMEIG_MAX_K];        // chosen indices, ascending
    __shared__ int ocol[SYMEIG_MAX_K];       // output column of asc[t]
    __shared__ int cl_start[SYMEIG_MAX_K + 1];
    __shared__ double wred[NW], wlo[NW], whi[NW], wem[NW];
    __shared__ double bnd[4];
    __shared__ double sx1;
    __shared__ int sint[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double *tau = work;

    // ---- 1. tridiagonalisation: column j of A below the diagonal becomes the
    //         reflector v (v[0] = 1), tau[j] its beta, ee[j] the subdiagonal ----
    for (int j = 0; j + 2 < m; j++) {
        const int n2 = m - j - 1;
        double *x = A + (j + 1) + (int64_t)j * m;
        const double xi = (tid < n2) ? x[tid] : 0.0;
        if (tid == 0) sx1 = xi;
        const double sigma = block_sum((tid >= 1 && tid < n2) ? xi * xi : 0.0, wred);
        const double x1 = sx1;
        double beta = 0.0, mu = x1, v1 = 1.0;
        if (sigma > 0.0) {
            mu = sqrt(x1 * x1 + sigma);
            v1 = (x1 <= 0.0) ? x1 - mu : -sigma / (x1 + mu);
            beta = 2.0 * v1 * v1 / (sigma + v1 * v1);
        }
        if (tid < n2) {
            const double v = (tid == 0) ? 1.0 : ((sigma > 0.0) ? xi / v1 : 0.0);
            vsh[tid] = v;
            x[tid] = v;
        }
        if (tid == 0) {
            dd[j] = A[j + (int64_t)j * m];
            ee[j] = mu;
            tau[j] = beta;
        }
        __syncthreads();
        if (beta != 0.0) {
            // p = beta * A22 v (A22 symmetric: column dot products)
            for (int cb = wid; cb < n2; cb += NW * CPI) {
                double a[CPI][RPL];
#pragma unroll
                for (int q = 0; q < CPI; q++) {
                    const int c = min(cb + NW * q, n2 - 1);
                    const double *col = A + (j + 1) + (int64_t)(j + 1 + c) * m;
#pragma unroll
                    for (int r = 0; r < RPL; r++) {
                        const int i = lane + 64 * r;
                        a[q][r] = (i < n2) ? col[i] : 0.0;
                    }
                }
                double s[CPI];
#pragma unroll
                for (int q = 0; q < CPI; q++) s[q] = 0.0;
#pragma unroll
                for (int r = 0; r < RPL; r++) {
                    const int i = lane + 64 * r;
                    const double vi = (i < n2) ? vsh[i] : 0.0;
#pragma unroll
                    for (int q = 0; q < CPI; q++) s[q] += a[q][r] * vi;
                }
#pragma unroll
                for (int q = 0; q < CPI; q++) {
                    const double t = wave_sum(s[q]);
                    const int c = cb + NW * q;
                    if (lane == 0 && c < n2) psh[c] = beta * t;
                }
            }
            __syncthreads();
            double pv = 0.0;
            for (int i = tid; i < n2; i += NT) pv += psh[i] * vsh[i];
            const double K = 0.5 * beta * block_sum(pv, wred);
            for (int i = tid; i < n2; i += NT) psh[i] -= K * vsh[i];
            __syncthreads();
            // A22 -= v w^T + w v^T
            for (int cb = wid; cb < n2; cb += NW * CPI) {
                double a[CPI][RPL];
#pragma unroll
                for (int q = 0; q < CPI; q++) {
                    const int c = cb + NW * q;
                    const double *col = A + (j + 1) + (int64_t)(j + 1 + min(c, n2 - 1)) * m;
#pragma unroll
                    for (int r = 0; r < RPL; r++) {
                        const int i = lane + 64 * r;
                        a[q][r] = (i < n2 && c < n2) ? col[i] : 0.0;
                    }
                }
#pragma unroll
                for (int q = 0; q < CPI; q++) {
                    const int c = cb + NW * q;
                    if (c < n2) {
                        double *col = A + (j + 1) + (int64_t)(j + 1 + c) * m;
                        const double vc = vsh[c], pc = psh[c];
#pragma unroll
                        for (int r = 0; r < RPL; r++) {
                            const int i = lane + 64 * r;
                            if (i < n2) col[i] = a[q][r] - (vsh[i] * pc + psh[i] * vc);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (m >= 2) {
            dd[m - 2] = A[(m - 2) + (int64_t)(m - 2) * m];
            ee[m - 2] = A[(m - 1) + (int64_t)(m - 2) * m];
            tau[m - 2] = 0.0;
        }
        dd[m - 1] = A[(m - 1) + (int64_t)(m - 1) * m];
        ee[m - 1] = 0.0;
        sint[0] = 0;
    }
    __syncthreads();

    // ---- 2. Gershgorin interval, then multisection at the extreme indices ----
    for (int i = tid; i < m; i += NT) e2[i] = ee[i] * ee[i];
    {
        double lo = INFINITY, hi = -INFINITY, em = 0.0;
        for (int i = tid; i < m; i += NT) {
            double r = fabs(ee[i]) + (i > 0 ? fabs(ee[i - 1]) : 0.0);
            lo = fmin(lo, dd[i] - r);
            hi = fmax(hi, dd[i] + r);
            em = fmax(em, ee[i] * ee[i]);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo = fmin(lo, __shfl_xor(lo, o, 64));
            hi = fmax(hi, __shfl_xor(hi, o, 64));
            em = fmax(em, __shfl_xor(em, o, 64));
        }
        if (lane == 0) { wlo[wid] = lo; whi[wid] = hi; wem[wid] = em; }
        __syncthreads();
        if (tid == 0) {
            lo = wlo[0]; hi = whi[0]; em = wem[0];
            for (int i = 1; i < NW; i++) {
                lo = fmin(lo, wlo[i]);
                hi = fmax(hi, whi[i]);
                em = fmax(em, wem[i]);
            }
            double tnorm = fmax(fabs(lo), fabs(hi));
            double pivmin = DBL_MIN * fmax(1.0, em);
            double pad = 2.0 * EPS * tnorm * m + 4.0 * pivmin;
            bnd[0] = lo - pad;
            bnd[1] = hi + pad;
            bnd[2] = tnorm;
            bnd[3] = pivmin;
        }
        __syncthreads();
    }
    const double tnorm = bnd[2], pivmin = bnd[3];
    const int both = (2 * k >= m);          // every index is an extreme one
    const int nidx = both ? m : 2 * k;
    {
        int G = 32;
        while (G > 1 && nidx * G > NT) G >>= 1;
        const int grp = tid / G, gl = tid % G;
        const bool active = grp < nidx;
        const int idx = both ? grp : (grp < k ? grp : m - 2 * k + grp);
        const int shift = (lane / G) * G;
        const uint64_t gmask = (1ull << G) - 1;
        double lo = bnd[0], hi = bnd[1];
        const double abstol = EPS * tnorm;
        for (int it = 0; it < 200; it++) {
            const double tol = fmax(abstol, 2.0 * EPS * fmax(fabs(lo), fabs(hi)));
            if (!active || hi - lo <= tol) break;
            const double w = hi - lo;
            const double xg = lo + w * (double)(gl + 1) / (double)(G + 1);
            const bool le = sturm_count(m, dd, e2, xg, pivmin) <= idx;
            const uint64_t bal = (uint64_t)__ballot(le);
            const int s = __popcll((bal >> shift) & gmask);
            const double nlo = (s > 0) ? lo + w * (double)s / (double)(G + 1) : lo;
            const double nhi = (s < G) ? lo + w * (double)(s + 1) / (double)(G + 1) : hi;
            lo = nlo;
            hi = fmax(nhi, nlo);
        }
        if (active && gl == 0) ev[grp] = 0.5 * (lo + hi);
    }
    __syncthreads();

    // ---- selection of the k largest |theta| and clusters (one lane) ----
    // eigenvalue of ascending index i, for i in [0, k) or [m - k, m)
    auto ev_at = [&](int i) { return both ? ev[i] : (i < k ? ev[i] : ev[i - (m - 2 * k)]); };
    if (tid == 0) {
        int a = 0, b = m - 1;
        for (int t = 0; t < k; t++) {
            if (fabs(ev_at(b)) >= fabs(ev_at(a))) sel[t] = b--;
            else sel[t] = a++;
        }
        // ascending list: [0, a) then (b, m)
        int na = 0;
        for (int i = 0; i < a; i++) asc[na++] = i;
        for (int i = b + 1; i < m; i++) asc[na++] = i;
        for (int t = 0; t < k; t++)
            for (int u = 0; u < k; u++)
                if (sel[u] == asc[t]) ocol[t] = u;
        // clusters: consecutive indices with a gap below 1e-3 * ||T||
        const double ortol = 1e-3 * tnorm;
        int nc = 0;
        for (int t = 0; t < k; t++) {
            bool join = t > 0 && asc[t] == asc[t - 1] + 1 && (ev_at(asc[t]) - ev_at(asc[t - 1])) <= ortol;
            if (!join) cl_start[nc++] = t;
        }
        cl_start[nc] = k;
        sint[1] = nc;
        for (int t = 0; t < k; t++) theta[t] = ev_at(sel[t]);
    }
    __syncthreads();

    // ---- 3. inverse iteration, one lane per cluster, LU and vector in LDS ----
    const int nc = sint[1];
    const int CB = SCR / (6 * m);  // clusters per batch (>= 2 for m <= 1024)
    for (int c0 = 0; c0 < nc; c0 += CB) {
        const int cid = c0 + tid;
        if (tid < CB && cid < nc) {
            double *D = scr + (size_t)tid * 6 * m;
            double *DL = D + m, *DU = DL + m, *DU2 = DU + m, *PV = DU2 + m, *z = PV + m;
            const double tiny = EPS * tnorm + pivmin;
            double xprev = 0.0;
            for (int t = cl_start[cid]; t < cl_start[cid + 1]; t++) {
                double x = ev_at(asc[t]);
                if (t > cl_start[cid]) {
                    double pertol = 10.0 * EPS * fmax(fabs(x), tnorm * 1e-3);
                    if (x <= xprev + pertol) x = xprev + pertol;
                }
                xprev = x;
                // LU with partial pivoting of T - x I (rows i, i+1 swapped when PV[i] = 1);
                // dcur/ucur carry D[i]/DU[i] of the row being eliminated
                double dcur = dd[0] - x, ucur = ee[0];
                for (int i = 0; i + 1 < m; i++) {
                    const double l = ee[i];
                    const double dn = dd[i + 1] - x, un = (i + 2 < m) ? ee[i + 1] : 0.0;
                    if (fabs(dcur) >= fabs(l)) {
                        const double f = (dcur != 0.0) ? l / dcur : 0.0;
                        D[i] = dcur; DL[i] = f; DU[i] = ucur; DU2[i] = 0.0; PV[i] = 0.0;
                        dcur = dn - f * ucur;
                        ucur = un;
                    } else {
                        const double f = dcur / l;
                        D[i] = l; DL[i] = f; DU[i] = dn; DU2[i] = un; PV[i] = 1.0;
                        dcur = ucur - f * dn;
                        ucur = -f * un;
                    }
                }
                D[m - 1] = dcur;
                DU[m - 1] = 0.0;
                for (int i = 0; i < m; i++)
                    if (fabs(D[i]) < tiny) D[i] = (D[i] < 0.0) ? -tiny : tiny;
                for (int i = 0; i < m; i++) z[i] = hash_unit(((uint64_t)(asc[t] + 1) << 32) ^ (uint64_t)i);
                // the eigenvalue is accurate to rounding, so two solves give the
                // vector to working precision; a clustered one gets a third with
                // re-orthogonalisation.  Solves are linear: normalise once at the end.
                const int nit = (cl_start[cid + 1] - cl_start[cid] > 1) ? 3 : 2;
                for (int it = 0; it < nit; it++) {
                    // L solve, carrying the current row's value
                    double cur = z[0];
                    for (int i = 0; i + 1 < m; i++) {
                        const double nxt = z[i + 1];
                        if (PV[i] == 0.0) {
                            z[i] = cur;
                            cur = nxt - DL[i] * cur;
                        } else {
                            z[i] = nxt;
                            cur = cur - DL[i] * nxt;
                        }
                    }
                    z[m - 1] = cur;
                    // U solve, carrying z[i+1], z[i+2]
                    double z1 = z[m - 1] / D[m - 1], z2 = 0.0;
                    z[m - 1] = z1;
                    for (int i = m - 2; i >= 0; i--) {
                        const double zi = (z[i] - DU[i] * z1 - DU2[i] * z2) / D[i];
                        z[i] = zi;
                        z2 = z1;
                        z1 = zi;
                    }
                    // orthogonalise against the cluster's earlier (unit) vectors
                    for (int u = cl_start[cid]; u < t; u++) {
                        const double *zu = Zs + (int64_t)ocol[u] * m;
                        double sdot = 0.0;
                        for (int i = 0; i < m; i++) sdot += zu[i] * z[i];
                        for (int i = 0; i < m; i++) z[i] -= sdot * zu[i];
                    }
                }
                double nrm = 0.0;
                for (int i = 0; i < m; i++) nrm += z[i] * z[i];
                nrm = sqrt(nrm);
                if (!(nrm > 0.0) || !isfinite(nrm)) {
                    atomicOr(&sint[0], 2);
                    nrm = 1.0;
                }
                const double inv = 1.0 / nrm;
                for (int i = 0; i < m; i++) z[i] *= inv;
                double *zo = Zs + (int64_t)ocol[t] * m;
                for (int i = 0; i < m; i++) zo[i] = z[i];
            }
        }
        __syncthreads();
    }

    // ---- 4. back-transformation Zs <- H_0 ... H_{m-3} Zs, vectors in LDS ----
    const int cbat = SCR / m;  // vectors per batch
    for (int c0 = 0; c0 < k; c0 += cbat) {
        const int nv = min(cbat, k - c0);
        for (int t = tid; t < nv * m; t += NT) scr[t] = Zs[(int64_t)c0 * m + t];
        // reflector j lives in A[j+1:, j] (length m-j-1); double buffer vsh/psh
        double *bcur = vsh, *bnext = psh;
        if (m >= 3 && tid < 2) bcur[tid] = A[(m - 2) + tid + (int64_t)(m - 3) * m];
        __syncthreads();
        for (int j = m - 3; j >= 0; j--) {
            // prefetch reflector j-1 (length m-j) into a register
            double pre = 0.0;
            const int n2n = m - j;
            if (j > 0 && tid < n2n) pre = A[j + tid + (int64_t)(j - 1) * m];
            const double beta = tau[j];
            const int n2 = m - j - 1;
            const double *v = bcur;
            if (beta != 0.0)
                for (int c = wid; c < nv; c += NW) {
                    double *z = scr + (size_t)c * m + (j + 1);
                    double sacc = 0.0;
                    for (int i = lane; i < n2; i += 64) sacc += v[i] * z[i];
                    sacc = beta * wave_sum(sacc);
                    for (int i = lane; i < n2; i += 64) z[i] -= sacc * v[i];
                }
            if (j > 0 && tid < n2n) bnext[tid] = pre;
            double *tsw = bcur;
            bcur = bnext;
            bnext = tsw;
            __syncthreads();
        }
        for (int t = tid; t < nv * m; t += NT) Zs[(int64_t)c0 * m + t] = scr[t];
        __syncthreads();
    }
    if (tid == 0) *info = sint[0];
}

}  // namespace

int64_t symeig_work_size(int m, int k) {
    (void)k;
    return (int64_t)m;
}

void sym_eig_top(int m, double *A, int k, double *theta, double *Zs, double *work, int *info, hipStream_t s) {
    const SymeigBatch one{nullptr, 0, 0, 0, 0};
    if (m <= 385)
        symeig_top_kernel<6, 4><<<1, NT, 0, s>>>(m, A, k, theta, Zs, work, info, one);
    else
        symeig_top_kernel<16, 1><<<1, NT, 0, s>>>(m, A, k, theta, Zs, work, info, one);
}

void sym_eig_top_batched(int m, double *A, int64_t sA, int k, double *theta, int64_t sT, double *Zs, int64_t sZ,
                         double *work, int64_t sW, int *info, const int *list, int nlist, hipStream_t s) {
    if (nlist <= 0) return;
    const SymeigBatch bt{list, sA, sT, sZ, sW};
    if (m <= 385)
        symeig_top_kernel<6, 4><<<nlist, NT, 0, s>>>(m, A, k, theta, Zs, work, info, bt);
    else
        symeig_top_kernel<16, 1><<<nlist, NT, 0, s>>>(m, A, k, theta, Zs, work, info, bt);
}
