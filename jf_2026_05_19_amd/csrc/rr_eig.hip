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

#include <algorithm>
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

// 1/q from the hardware reciprocal estimate and two Newton-Raphson steps
// r <- r + r (1 - q r) (each doubles the correct bits; the result is within
// an ulp or so, not correctly rounded): a shorter dependent chain than the
// correctly rounded division.  |q| >= pivmin keeps 1/q finite.
__device__ __forceinline__ double recip_nr(double q) {
    double r = __builtin_amdgcn_rcp(q);
    r = fma(fma(-q, r, 1.0), r, r);
    r = fma(fma(-q, r, 1.0), r, r);
    return r;
}

// number of eigenvalues of the tridiagonal (d, e2 = e^2) below x; each step
// multiplies by the reciprocal above, so the count is that of a matrix
// perturbed at the rounding level, as with a division
__device__ int sturm_count(int m, const double *d, const double *e2, double x, double pivmin) {
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int c = (q < 0.0);
    // unrolled so the LDS reads of d and e2 run ahead of the recurrence
#pragma unroll 8
    for (int i = 1; i < m; i++) {
        q = d[i] - x - e2[i - 1] * recip_nr(q);
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
    int alg;  // 1: the k largest theta instead of the k largest |theta|
};

// LDS arrays shared by the tridiagonal stages (steps 2 and 3)
struct TriLds {
    double *dd, *ee, *e2, *scr, *ev, *wlo, *whi, *wem, *bnd;
    int *sel, *asc, *ocol, *cl_start, *sint;
    int scr_n;  // doubles of scr
    int alg;    // select the k largest theta (descending) instead of |theta|
};

// Steps 2 and 3 on the tridiagonal (dd, ee; ee[m-1] = 0) held in LDS: the
// eigenvalues at the 2k extreme indices by multisection, the k of largest
// |theta| (or, with L.alg, the k largest) into theta (descending |theta|,
// or descending) and their eigenvectors by inverse iteration into Zs
// (m x k, ld m).  All NT threads of the workgroup call it;
// sint[0] collects failures (bit 2: a vector did not come out finite).
__device__ void tridiag_top_pairs(int m, int k, const TriLds &L, double *theta, double *Zs) {
    double *dd = L.dd, *ee = L.ee, *e2 = L.e2, *scr = L.scr, *ev = L.ev, *wlo = L.wlo, *whi = L.whi, *wem = L.wem,
           *bnd = L.bnd;
    int *sel = L.sel, *asc = L.asc, *ocol = L.ocol, *cl_start = L.cl_start, *sint = L.sint;
    const int scr_n = L.scr_n;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
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
            if (L.alg || fabs(ev_at(b)) >= fabs(ev_at(a))) sel[t] = b--;
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
    // clusters per batch (>= 2 for m <= 1024): a chain keeps four arrays of
    // m doubles and m pivot bytes; its stride is odd, so the chains of a
    // batch start on different LDS banks.  The second superdiagonal of U is
    // not stored: it is ee[i+1] where rows i, i+1 were swapped, else 0.
    const int CS = (4 * m + (m + 7) / 8) | 1;
    const int CB = scr_n / CS;
    // slot s of a batch runs on lane s of wavefront 0: the LU and the solves
    // are branch-free, so one instruction advances every chain of the batch
    // (one lane per wavefront would pay a whole wavefront's issue per chain)
    const int slot = tid;
    for (int c0 = 0; c0 < nc; c0 += CB) {
        const int cid = c0 + slot;
        if (slot < CB && cid < nc) {
            double *D = scr + (size_t)slot * CS;
            double *DL = D + m, *DU = DL + m, *z = DU + m;
            unsigned char *PV = reinterpret_cast<unsigned char *>(z + m);
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
                // the O(m) loops below are unrolled so that their LDS reads run
                // ahead of the recurrence
                // (branch-free: both pivot choices are selects, so the loop
                // body is one block and the reads run ahead of the recurrence)
#pragma unroll 4
                for (int i = 0; i + 1 < m; i++) {
                    const double l = ee[i];
                    const double dn = dd[i + 1] - x, un = (i + 2 < m) ? ee[i + 1] : 0.0;
                    const bool keep = fabs(dcur) >= fabs(l);  // no row swap
                    const double num = keep ? l : dcur, den = keep ? dcur : l;
                    const double f = (den != 0.0) ? num / den : 0.0;
                    D[i] = den;
                    DL[i] = f;
                    DU[i] = keep ? ucur : dn;
                    PV[i] = keep ? 0 : 1;
                    const double nd = keep ? dn - f * ucur : ucur - f * dn;
                    ucur = keep ? un : -f * un;
                    dcur = nd;
                }
                D[m - 1] = dcur;
                DU[m - 1] = 0.0;
                // D keeps the reciprocal pivots from here on: the U solves multiply
#pragma unroll 4
                for (int i = 0; i < m; i++) {
                    const double di = D[i];
                    D[i] = 1.0 / ((fabs(di) < tiny) ? ((di < 0.0) ? -tiny : tiny) : di);
                }
                for (int i = 0; i < m; i++) z[i] = hash_unit(((uint64_t)(asc[t] + 1) << 32) ^ (uint64_t)i);
                // the eigenvalue is accurate to rounding, so two solves give the
                // vector to working precision; a clustered one gets a third with
                // re-orthogonalisation.  Solves are linear: normalise once at the end.
                const int nit = (cl_start[cid + 1] - cl_start[cid] > 1) ? 3 : 2;
                for (int it = 0; it < nit; it++) {
                    // L solve, carrying the current row's value
                    double cur = z[0];
#pragma unroll 4
                    for (int i = 0; i + 1 < m; i++) {
                        const double nxt = z[i + 1], li = DL[i];
                        const bool keep = PV[i] == 0;
                        z[i] = keep ? cur : nxt;
                        cur = keep ? nxt - li * cur : cur - li * nxt;
                    }
                    z[m - 1] = cur;
                    // U solve, carrying z[i+1], z[i+2]
                    double z1 = z[m - 1] * D[m - 1], z2 = 0.0;
                    z[m - 1] = z1;
#pragma unroll 4
                    for (int i = m - 2; i >= 0; i--) {
                        const double du2 = (PV[i] && i + 2 < m) ? ee[i + 1] : 0.0;
                        const double zi = (z[i] - DU[i] * z1 - du2 * z2) * D[i];
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

}

template <int RPL, int CPI>
__global__ void __launch_bounds__(NT) symeig_top_kernel(int m, double *A, int k, double *theta, double *Zs,
                                                        double *work, int *info, SymeigBatch bt, int after_band) {
    if (bt.list) {
        const int64_t p = bt.list[blockIdx.x];
        A += p * bt.sA;
        theta += p * bt.sT;
        Zs += p * bt.sZ;
        work += p * bt.sW;
        info += p;
    }
    // behind the band kernel: run only where it declined (info < 0)
    if (after_band && *info >= 0) return;
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

    {
        const TriLds L{dd, ee, e2, scr, ev, wlo, whi, wem, bnd, sel, asc, ocol, cl_start, sint, SCR, bt.alg};
        tridiag_top_pairs(m, k, L, theta, Zs);
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


// ---------------------------------------------------------------------------
// Banded projected matrices.  Without restarts the block Lanczos matrix is
// block tridiagonal: half-bandwidth bw = the block size.  Its top pairs come
// from a band-to-tridiagonal reduction by bulge chasing (Schwarz's method in
// the Householder form of Bischof, Lang and Sun's successive band reduction,
// the task structure of LAPACK's dsb2st_kernels): sweep i annihilates column
// i below the subdiagonal with one reflector acting on rows [i+1, i+bw],
// then chases the bulge down in blocks of bw rows; each step applies the
// previous reflector from the right to the block below, annihilates that
// block's first column with a new reflector and applies it from both sides
// to the next diagonal block.  Sweeps run pipelined, sweep i two steps behind
// sweep i-1: step s of sweep i touches rows up to i + (s+1) bw, step s+2 of
// sweep i-1 rows from i + (s+2) bw on (their row ranges meet only at index
// i + (s+1) bw, a column of the one and the last diagonal row of the other,
// no common entry), and every pair of tasks that does share entries keeps
// its sequential order; one wavefront per active sweep and one workgroup
// barrier per step, the band in LDS.  Steps 2 and 3 are those
// of the dense path; the vectors are then mapped back through the stored
// reflectors, sweep by sweep in reverse (one sweep's reflectors act on
// disjoint rows and are applied together).  O(m^2 bw) work instead of the
// dense path's O(m^3).

constexpr int LDB = 2 * SYMEIG_BAND_MAX_W + 1;                 // band column: diagonal + bw + bulge
constexpr int BAND_LDS = SYMEIG_BAND_MAX_M * LDB > SCR ? SYMEIG_BAND_MAX_M * LDB : SCR;
constexpr int SCR_GB = 14336;  // LDS scratch of steps 3 and 4 when the band lives in the workspace
// LDS scratch of steps 3 and 4 with the band in LDS: 12 inverse-iteration
// chains of m = 304 (4 m + m/8 doubles each) in one batch
constexpr int SCR_BAND = BAND_LDS > 15104 ? BAND_LDS : 15104;
constexpr int NSLOT = 64;                                      // previous-reflector slots (>= active sweeps)

// Workgroup barrier that orders LDS only: this wavefront's LDS operations
// complete, then s_barrier.  Global loads and stores in flight stay in flight
// (__syncthreads() starts with a release fence that waits for them); used
// where the barrier guards LDS data alone.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ __forceinline__ int band_nsteps(int m, int bw, int i) { return (m - 1 - i + bw - 1) / bw; }
// sum_{j=1}^{n} ceil(j / bw)
__device__ __forceinline__ int64_t band_F(int n, int bw) {
    const int64_t q = n / bw, r = n % bw;
    return bw * q * (q + 1) / 2 + r * (q + 1);
}
// first task index of sweep i (tasks of sweeps 0..i-1)
__device__ __forceinline__ int64_t band_toff(int m, int bw, int i) { return band_F(m - 1, bw) - band_F(m - 1 - i, bw); }

// v from the lane that the DPP control ctrl selects within the row of 16
template <int ctrl>
__device__ __forceinline__ double dpp_d(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, ctrl, 0xF, 0xF, false);
    const uint32_t hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), ctrl, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// sum over the 16 lanes of a row group (lanes sharing l / 16) by a DPP
// butterfly: quad_perm(1,0,3,2), quad_perm(2,3,0,1), row_half_mirror,
// row_mirror; every step adds a pair symmetrically, so all 16 lanes hold the
// same bits
__device__ __forceinline__ double sum16(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return v;
}

__device__ __forceinline__ double readlane_d(double x, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// sum over the 4 lanes sharing l % 16
__device__ __forceinline__ double sum4(double v) {
    v += __shfl_xor(v, 16, 64);
    return v + __shfl_xor(v, 32, 64);
}

// Householder reflector H = I - tau v v^T with v[0] = 1 and H x = mu e_0 for
// the vector x (lane l holds x[l % 16], length L); Golub & Van Loan Alg. 5.1.1.
__device__ __forceinline__ void house16(double x, int p, int L, double &v, double &tau, double &mu) {
    const double alpha = readlane_d(x, 0);
    const double sigma = sum16((p >= 1 && p < L) ? x * x : 0.0);
    if (sigma > 0.0) {
        mu = sqrt(alpha * alpha + sigma);
        const double v1 = (alpha <= 0.0) ? alpha - mu : -sigma / (alpha + mu);
        tau = 2.0 * v1 * v1 / (sigma + v1 * v1);
        v = (p == 0) ? 1.0 : ((p < L) ? x / v1 : 0.0);
    } else {
        tau = 0.0;
        mu = alpha;
        v = (p == 0) ? 1.0 : 0.0;
    }
}

// One task of sweep i, step s, by one wavefront.  AB: lower band, A[r][c]
// (r >= c) at AB[c * LDB + r - c].  slot: the sweep's previous reflector
// (v[0..15], tau at [16]); rec: this task's record for the back-transformation.
__device__ void band_task(int m, int bw, int i, int s, double *AB, double *slot, double *rec) {
    const int lane = threadIdx.x & 63, p = lane & 15, qg = lane >> 4;
    auto at = [&](int r, int c) -> double & { return AB[c * LDB + (r - c)]; };
    double v, tau, mu;
    int r0, L;
    if (s == 0) {
        r0 = i + 1;
        L = min(bw, m - r0);
        const double x = (p < L) ? at(r0 + p, i) : 0.0;
        house16(x, p, L, v, tau, mu);
        if (qg == 0 && p < L) at(r0 + p, i) = (p == 0) ? mu : 0.0;
    } else {
        const int cs = i + 1 + (s - 1) * bw, Lp = min(bw, m - cs);
        r0 = i + 1 + s * bw;
        L = min(bw, m - r0);
        const double tp = slot[16];
        double c[4], vq[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = qg + 4 * j;
            vq[j] = (q < Lp) ? slot[q] : 0.0;
            c[j] = (p < L && q < Lp) ? at(r0 + p, cs + q) : 0.0;
        }
        // the block below the previous reflector's rows: C <- C H_prev
        const double y = sum4(c[0] * vq[0] + c[1] * vq[1] + c[2] * vq[2] + c[3] * vq[3]);
#pragma unroll
        for (int j = 0; j < 4; j++) c[j] -= tp * y * vq[j];
        // annihilate C's first column below its first row (lanes with qg == 0 hold it in c[0])
        const double x = __shfl(c[0], p, 64);
        house16(x, p, L, v, tau, mu);
        if (qg == 0) c[0] = (p == 0) ? mu : 0.0;
        // and apply the new reflector from the left to C's other columns
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double z = sum16(v * c[j]);
            if (qg + 4 * j >= 1) c[j] -= tau * v * z;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = qg + 4 * j;
            if (p < L && q < Lp) at(r0 + p, cs + q) = c[j];
        }
    }
    // two-sided on the diagonal block S = A[r0:r0+L, r0:r0+L]:
    // p = tau S v, K = tau/2 p^T v, w = p - K v, S <- S - v w^T - w v^T
    double sv[4], vq[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = qg + 4 * j;
        vq[j] = __shfl(v, q, 64);
        sv[j] = (p < L && q < L) ? at(r0 + max(p, q), r0 + min(p, q)) : 0.0;
    }
    const double pv = tau * sum4(sv[0] * vq[0] + sv[1] * vq[1] + sv[2] * vq[2] + sv[3] * vq[3]);
    const double K = 0.5 * tau * sum16((p < L) ? pv * v : 0.0);
    const double w = pv - K * v;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = qg + 4 * j;
        const double wq = __shfl(w, q, 64);
        if (p < L && q < L && p >= q) at(r0 + p, r0 + q) = sv[j] - (v * wq + w * vq[j]);
    }
    if (qg == 0) {
        slot[p] = v;
        if (p < bw) rec[p] = v;
    }
    if (lane == 0) {
        slot[16] = tau;
        rec[bw] = tau;
    }
}

// GB: the band lives in the workspace (L2-resident; m up to SYMEIG_MAX_M)
// instead of LDS (m up to SYMEIG_BAND_MAX_M), and the step barriers then
// order global memory too.
template <bool GB>
__global__ void __launch_bounds__(NT) symeig_band_kernel(int m, int bw, const double *A, int k, double *theta,
                                                         double *Zs, double *work, int *info, SymeigBatch bt) {
    if (bt.list) {
        const int64_t p = bt.list[blockIdx.x];
        A += p * bt.sA;
        theta += p * bt.sT;
        Zs += p * bt.sZ;
        work += p * bt.sW;
        info += p;
    }
    constexpr int MAXM = GB ? SYMEIG_MAX_M : SYMEIG_BAND_MAX_M;
    constexpr int SCRN = GB ? SCR_GB : SCR_BAND;
    __shared__ double lscr[SCRN];  // the band (LDS form); afterwards scratch of steps 3 and 4
    __shared__ double dd[MAXM], ee[MAXM], e2[MAXM];
    __shared__ double slots[NSLOT][17];
    __shared__ int sw_nst[MAXM], sw_off[MAXM];  // per sweep: steps, first record
    __shared__ double ev[2 * SYMEIG_MAX_K];
    __shared__ int sel[SYMEIG_MAX_K], asc[SYMEIG_MAX_K], ocol[SYMEIG_MAX_K], cl_start[SYMEIG_MAX_K + 1];
    __shared__ double wlo[NW], whi[NW], wem[NW];
    __shared__ double bnd[4];
    __shared__ int sint[4];
    double *AB = GB ? work : lscr;                          // the band
    double *recs0 = GB ? work + (int64_t)m * LDB : work;    // reflector records
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;

    // ---- 1. band into place (bulge room zeroed); decline a matrix that is not banded ----
    for (int t = tid; t < m * LDB; t += NT) {
        const int c = t / LDB, d = t % LDB;
        AB[t] = (d <= bw && c + d < m) ? A[(c + d) + (int64_t)c * m] : 0.0;
    }
    {
        double offm = 0.0, allm = 0.0;
        for (int c = wid; c < m; c += NW)
            for (int r = c + lane; r < m; r += 64) {
                const double a = fabs(A[r + (int64_t)c * m]);
                allm = fmax(allm, a);
                if (r - c > bw) offm = fmax(offm, a);
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            offm = fmax(offm, __shfl_xor(offm, o, 64));
            allm = fmax(allm, __shfl_xor(allm, o, 64));
        }
        if (lane == 0) { wlo[wid] = offm; whi[wid] = allm; }
        __syncthreads();
        if (tid == 0) {
            double o = 0.0, a = 0.0;
            for (int i = 0; i < NW; i++) { o = fmax(o, wlo[i]); a = fmax(a, whi[i]); }
            sint[0] = 0;
            sint[2] = !(o <= 1e-12 * a);  // off-band entries beyond rounding (or NaN): the dense path
        }
        __syncthreads();
        if (sint[2]) {
            if (tid == 0) *info = -1;
            return;
        }
    }

    // ---- bulge chasing: time step ts runs step ts - 2i of every active sweep i ----
    const int nsw = m - 2;  // sweeps 0 .. m-3
    for (int i = tid; i < nsw; i += NT) {
        sw_nst[i] = band_nsteps(m, bw, i);
        sw_off[i] = (int)band_toff(m, bw, i);
    }
    __syncthreads();
    if (nsw > 0) {
        const int tend = 2 * (nsw - 1) + sw_nst[nsw - 1];
        // this wavefront's sweeps are i == wid (mod NW); ihi = min(nsw - 1, ts / 2)
        // is tracked incrementally, and older sweeps finish first; at most
        // ceil(steps of sweep 0 / 2) + 1 sweeps are active, <= NSLOT (band_path)
        int ihi = 0, ph = 0;
        for (int ts = 0; ts < tend; ts++) {
            for (int i = ihi - ((ihi - wid) % NW + NW) % NW; i >= 0; i -= NW) {
                const int st = ts - 2 * i;
                if (st >= sw_nst[i]) break;
                band_task(m, bw, i, st, AB, slots[i % NSLOT], recs0 + (int64_t)(sw_off[i] + st) * (bw + 1));
            }
            // an LDS band and the slots need LDS order only (the record stores,
            // read after the full barriers below, need not complete here); a
            // band in the workspace needs the full barrier
            if (GB) __syncthreads();
            else lds_barrier();
            if (++ph == 2) {
                ph = 0;
                ihi = min(nsw - 1, ihi + 1);
            }
        }
    }
    for (int j = tid; j < m; j += NT) {
        dd[j] = AB[j * LDB];
        ee[j] = (j + 1 < m) ? AB[j * LDB + 1] : 0.0;
    }
    __syncthreads();

    // ---- 2 and 3 on the tridiagonal ----
    {
        const TriLds L{dd, ee, e2, lscr, ev, wlo, whi, wem, bnd, sel, asc, ocol, cl_start, sint, SCRN, bt.alg};
        tridiag_top_pairs(m, k, L, theta, Zs);
    }
    __syncthreads();

    // ---- 4. back-transformation Zs <- Q Zs, Q the product of the sweeps' reflectors ----
    // task (st, c) applies step st of the current sweep to vector c.  When a
    // sweep's tasks fit the workgroup (always for bw = 16), thread tid keeps
    // task (tid / nv, tid % nv) and loads the records of the next sweep
    // (i - 1) while this one is applied; otherwise the tasks are strided.
    // vectors in LDS with column stride m + 1 (odd), so that the tasks of a
    // sweep, which differ in the vector, hit different banks
    const int ldz = m + 1, cbat = SCRN / ldz;
    auto apply = [&](int i, int st, int c, const double *vr, double tau, double *Z) {
        const int r0 = i + 1 + st * bw, L = min(bw, m - r0);
        double *z = Z + (size_t)c * ldz + r0;
        double dot = 0.0;
#pragma unroll
        for (int q = 0; q < SYMEIG_BAND_MAX_W; q++) dot += ((q < L) ? vr[q] : 0.0) * z[min(q, L - 1)];
        dot *= tau;
        // z is read again from LDS rather than kept in registers
#pragma unroll
        for (int q = 0; q < SYMEIG_BAND_MAX_W; q++)
            if (q < L) z[q] -= dot * vr[q];
    };
    for (int c0 = 0; c0 < k; c0 += cbat) {
        const int nv = min(cbat, k - c0);
        double *Z = lscr;  // nv columns of length m, stride ldz
        for (int t = tid; t < nv * m; t += NT) Z[(t / m) * ldz + t % m] = Zs[(int64_t)c0 * m + t];
        __syncthreads();
        if (nsw > 0 && sw_nst[0] * nv <= NT) {
            const int st = tid / nv, c = tid % nv;
            const bool act = tid < sw_nst[0] * nv;  // threads without a task load nothing
            double v[SYMEIG_BAND_MAX_W], tau = 0.0;
#pragma unroll
            for (int q = 0; q < SYMEIG_BAND_MAX_W; q++) v[q] = 0.0;
            // fixed-width record loads, clamped into sweep i's records
            auto load_rec = [&](int i) {
                const double *rv = recs0 + ((int64_t)sw_off[i] + min(st, sw_nst[i] - 1)) * (bw + 1);
#pragma unroll
                for (int q = 0; q < SYMEIG_BAND_MAX_W; q++) v[q] = rv[min(q, bw)];
                tau = rv[bw];
            };
            if (act) load_rec(nsw - 1);
            for (int i = nsw - 1; i >= 0; i--) {
                double vc[SYMEIG_BAND_MAX_W];
#pragma unroll
                for (int q = 0; q < SYMEIG_BAND_MAX_W; q++) vc[q] = v[q];
                const double tc = tau;
                if (act && i > 0) load_rec(i - 1);
                if (st < sw_nst[i]) apply(i, st, c, vc, tc, Z);
                lds_barrier();  // Z is in LDS; the next records stay in flight
            }
        } else {
            for (int i = nsw - 1; i >= 0; i--) {
                const int nst = sw_nst[i];
                const double *recs = recs0 + (int64_t)sw_off[i] * (bw + 1);
                for (int t = tid; t < nst * nv; t += NT) {
                    const int st = t / nv, c = t % nv;
                    const double *rv = recs + st * (bw + 1);
                    double v[SYMEIG_BAND_MAX_W];
#pragma unroll
                    for (int q = 0; q < SYMEIG_BAND_MAX_W; q++) v[q] = rv[min(q, bw)];
                    apply(i, st, c, v, rv[bw], Z);
                }
                __syncthreads();
            }
        }
        for (int t = tid; t < nv * m; t += NT) Zs[(int64_t)c0 * m + t] = Z[(t / m) * ldz + t % m];
        __syncthreads();
    }
    if (tid == 0) *info = sint[0];
}

}  // namespace

// band path (4 <= bw <= 16) at size m: one record of bw + 1 doubles per task,
// at most m^2 (bw + 1) / (2 bw) + m (bw + 1) <= 5 m^2 / 8 + 17 m; above
// SYMEIG_BAND_MAX_M the band itself (m columns of LDB) comes first
static int64_t band_work(int m) {
    return (m > SYMEIG_BAND_MAX_M ? (int64_t)m * LDB : 0) + (int64_t)m + (5 * (int64_t)m * m) / 8 +
           17 * (int64_t)m + 64;
}

// a workspace sized for m serves every call at sizes up to m: the dense path
// needs m doubles, the band path band_work(m') for m' <= m (band_work grows with m)
int64_t symeig_work_size(int m, int k) {
    (void)k;
    return std::max<int64_t>(m, band_work(std::min(m, SYMEIG_MAX_M)));
}

// the sweeps active in one time step (at most ceil(steps of sweep 0 / 2) + 1)
// must fit the previous-reflector slots
static bool band_path(int m, int bw, int64_t work_n) {
    if (bw < 4 || bw > SYMEIG_BAND_MAX_W || m < 3 || m > SYMEIG_MAX_M || work_n < band_work(m)) return false;
    const int nst0 = (m - 1 + bw - 1) / bw;
    return (nst0 + 1) / 2 + 1 <= NSLOT;
}

void sym_eig_top(int m, double *A, int k, double *theta, double *Zs, double *work, int64_t work_n, int *info,
                 hipStream_t s, int bw, int algebraic) {
    const SymeigBatch one{nullptr, 0, 0, 0, 0, algebraic ? 1 : 0};
    const int band = band_path(m, bw, work_n);
    if (band && m <= SYMEIG_BAND_MAX_M) symeig_band_kernel<false><<<1, NT, 0, s>>>(m, bw, A, k, theta, Zs, work, info, one);
    else if (band) symeig_band_kernel<true><<<1, NT, 0, s>>>(m, bw, A, k, theta, Zs, work, info, one);
    if (m <= 385)
        symeig_top_kernel<6, 4><<<1, NT, 0, s>>>(m, A, k, theta, Zs, work, info, one, band);
    else
        symeig_top_kernel<16, 1><<<1, NT, 0, s>>>(m, A, k, theta, Zs, work, info, one, band);
}

void sym_eig_top_batched(int m, double *A, int64_t sA, int k, double *theta, int64_t sT, double *Zs, int64_t sZ,
                         double *work, int64_t sW, int *info, const int *list, int nlist, hipStream_t s, int bw) {
    if (nlist <= 0) return;
    const SymeigBatch bt{list, sA, sT, sZ, sW, 0};
    const int band = band_path(m, bw, sW);
    if (band && m <= SYMEIG_BAND_MAX_M)
        symeig_band_kernel<false><<<nlist, NT, 0, s>>>(m, bw, A, k, theta, Zs, work, info, bt);
    else if (band)
        symeig_band_kernel<true><<<nlist, NT, 0, s>>>(m, bw, A, k, theta, Zs, work, info, bt);
    if (m <= 385)
        symeig_top_kernel<6, 4><<<nlist, NT, 0, s>>>(m, A, k, theta, Zs, work, info, bt, band);
    else
        symeig_top_kernel<16, 1><<<nlist, NT, 0, s>>>(m, A, k, theta, Zs, work, info, bt, band);
}
