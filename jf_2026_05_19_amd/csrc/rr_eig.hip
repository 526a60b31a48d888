// rr_eig.hip -- the Rayleigh-Ritz step of the block Lanczos eigensolver:
// the k eigenpairs of largest |theta| of the symmetric projected matrix
// (m x m, m <= 1024), in one 1024-thread workgroup.
//   1. Householder tridiagonalisation (Golub & Van Loan, Matrix Computations,
//      4th ed., Alg. 5.1.1 `house` and Alg. 8.3.1): one wavefront per column
//      for the symmetric matrix-vector product, coalesced rank-2 update;
//   2. eigenvalues of the tridiagonal matrix at the 2k extreme indices by
//      Sturm-sequence bisection (Wilkinson, The Algebraic Eigenvalue Problem,
//      ch. 5), one lane per index; the k of largest |theta| are kept;
//   3. eigenvectors by inverse iteration with a partially pivoted
//      tridiagonal LU (Golub & Van Loan Alg. 4.3.x / 8.2), one lane per
//      cluster of close eigenvalues, modified Gram-Schmidt inside a cluster
//      (the scheme of Wilkinson and of LAPACK's stein);
//   4. back-transformation of the k vectors with the stored reflectors.
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

__global__ void __launch_bounds__(NT) symeig_top_kernel(int m, double *A, int k, double *theta, double *Zs,
                                                        double *work, int *info) {
    __shared__ double vsh[NT];
    __shared__ double psh[NT];
    __shared__ double dd[NT], ee[NT], e2[NT];
    __shared__ double ev[2 * SYMEIG_MAX_K];  // eigenvalues at the extreme indices
    __shared__ int sel[SYMEIG_MAX_K];        // chosen index, descending |theta|
    __shared__ int asc[SYMEIG_MAX_K];        // chosen indices, ascending
    __shared__ int ocol[SYMEIG_MAX_K];       // output column of asc[t]
    __shared__ int cl_start[SYMEIG_MAX_K + 1];
    __shared__ double wred[NW], wlo[NW], whi[NW], wem[NW];
    __shared__ double bnd[4];
    __shared__ int sint[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double *tau = work;

    // ---- 1. tridiagonalisation: column j of A below the diagonal becomes the
    //         reflector v (v[0] = 1), tau[j] its beta, ee[j] the subdiagonal ----
    for (int j = 0; j + 2 < m; j++) {
        const int n2 = m - j - 1;
        double *x = A + (j + 1) + (int64_t)j * m;
        double part = 0.0;
        for (int i = 1 + tid; i < n2; i += NT) part += x[i] * x[i];
        const double sigma = block_sum(part, wred);
        const double x1 = x[0];
        double beta = 0.0, mu = x1;
        if (sigma > 0.0) {
            mu = sqrt(x1 * x1 + sigma);
            double v1 = (x1 <= 0.0) ? x1 - mu : -sigma / (x1 + mu);
            beta = 2.0 * v1 * v1 / (sigma + v1 * v1);
            for (int i = tid; i < n2; i += NT) vsh[i] = (i == 0) ? 1.0 : x[i] / v1;
        } else {
            for (int i = tid; i < n2; i += NT) vsh[i] = (i == 0) ? 1.0 : 0.0;
        }
        __syncthreads();
        if (tid == 0) {
            dd[j] = A[j + (int64_t)j * m];
            ee[j] = mu;
            tau[j] = beta;
        }
        for (int i = tid; i < n2; i += NT) x[i] = vsh[i];
        if (beta != 0.0) {
            // p = beta * A22 v, one wavefront per column of the symmetric A22
            for (int i = wid; i < n2; i += NW) {
                const double *col = A + (j + 1) + (int64_t)(j + 1 + i) * m;
                double sacc = 0.0;
                for (int r = lane; r < n2; r += 64) sacc += col[r] * vsh[r];
                sacc = wave_sum(sacc);
                if (lane == 0) psh[i] = beta * sacc;
            }
            __syncthreads();
            double pv = 0.0;
            for (int i = tid; i < n2; i += NT) pv += psh[i] * vsh[i];
            const double K = 0.5 * beta * block_sum(pv, wred);
            for (int i = tid; i < n2; i += NT) psh[i] -= K * vsh[i];
            __syncthreads();
            // A22 -= v w^T + w v^T
            for (int64_t t = tid; t < (int64_t)n2 * n2; t += NT) {
                int i = (int)(t % n2), c = (int)(t / n2);
                A[(j + 1 + i) + (int64_t)(j + 1 + c) * m] -= vsh[i] * psh[c] + psh[i] * vsh[c];
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

    // ---- 2. Gershgorin interval, then bisection at the extreme indices ----
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
    if (tid < nidx) {
        int idx = both ? tid : (tid < k ? tid : m - 2 * k + tid);
        double lo = bnd[0], hi = bnd[1];
        const double abstol = EPS * tnorm;
        for (int it = 0; it < 200; it++) {
            double tol = fmax(abstol, 2.0 * EPS * fmax(fabs(lo), fabs(hi)));
            if (hi - lo <= tol) break;
            double mid = 0.5 * (lo + hi);
            if (sturm_count(m, dd, e2, mid, pivmin) <= idx) lo = mid;
            else hi = mid;
        }
        ev[tid] = 0.5 * (lo + hi);
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

    // ---- 3. inverse iteration, one lane per cluster ----
    const int nc = sint[1];
    if (tid < nc) {
        double *D = work + m + (int64_t)tid * 5 * m;
        double *DL = D + m, *DU = DL + m, *DU2 = DU + m, *PV = DU2 + m;
        const double tiny = EPS * tnorm + pivmin;
        double xprev = 0.0;
        for (int t = cl_start[tid]; t < cl_start[tid + 1]; t++) {
            double x = ev_at(asc[t]);
            if (t > cl_start[tid]) {
                double pertol = 10.0 * EPS * fmax(fabs(x), tnorm * 1e-3);
                if (x <= xprev + pertol) x = xprev + pertol;
            }
            xprev = x;
            // LU with partial pivoting of T - x I (rows i, i+1 swapped when PV[i] = 1)
            for (int i = 0; i < m; i++) {
                D[i] = dd[i] - x;
                DL[i] = ee[i];
                DU[i] = ee[i];
                DU2[i] = 0.0;
                PV[i] = 0.0;
            }
            for (int i = 0; i + 1 < m; i++) {
                if (fabs(D[i]) >= fabs(DL[i])) {
                    double f = (D[i] != 0.0) ? DL[i] / D[i] : 0.0;
                    DL[i] = f;
                    D[i + 1] -= f * DU[i];
                } else {
                    double f = D[i] / DL[i];
                    D[i] = DL[i];
                    DL[i] = f;
                    double tmp = DU[i];
                    DU[i] = D[i + 1];
                    D[i + 1] = tmp - f * D[i + 1];
                    if (i + 2 < m) {
                        DU2[i] = DU[i + 1];
                        DU[i + 1] = -f * DU[i + 1];
                    }
                    PV[i] = 1.0;
                }
            }
            for (int i = 0; i < m; i++)
                if (fabs(D[i]) < tiny) D[i] = (D[i] < 0.0) ? -tiny : tiny;
            double *z = Zs + (int64_t)ocol[t] * m;
            for (int i = 0; i < m; i++) z[i] = hash_unit(((uint64_t)(asc[t] + 1) << 32) ^ (uint64_t)i);
            for (int it = 0; it < 4; it++) {
                // L solve
                for (int i = 0; i + 1 < m; i++) {
                    if (PV[i] == 0.0) {
                        z[i + 1] -= DL[i] * z[i];
                    } else {
                        double tmp = z[i];
                        z[i] = z[i + 1];
                        z[i + 1] = tmp - DL[i] * z[i];
                    }
                }
                // U solve
                z[m - 1] /= D[m - 1];
                if (m >= 2) z[m - 2] = (z[m - 2] - DU[m - 2] * z[m - 1]) / D[m - 2];
                for (int i = m - 3; i >= 0; i--) z[i] = (z[i] - DU[i] * z[i + 1] - DU2[i] * z[i + 2]) / D[i];
                // orthogonalise against the cluster's earlier vectors
                for (int u = cl_start[tid]; u < t; u++) {
                    const double *zu = Zs + (int64_t)ocol[u] * m;
                    double s = 0.0;
                    for (int i = 0; i < m; i++) s += zu[i] * z[i];
                    for (int i = 0; i < m; i++) z[i] -= s * zu[i];
                }
                double nrm = 0.0;
                for (int i = 0; i < m; i++) nrm += z[i] * z[i];
                nrm = sqrt(nrm);
                if (!(nrm > 0.0) || !isfinite(nrm)) {
                    atomicOr(&sint[0], 2);
                    break;
                }
                double inv = 1.0 / nrm;
                for (int i = 0; i < m; i++) z[i] *= inv;
            }
        }
    }
    __syncthreads();

    // ---- 4. back-transformation Zs <- H_0 ... H_{m-3} Zs, one wavefront per vector ----
    for (int c = wid; c < k; c += NW) {
        double *z = Zs + (int64_t)c * m;
        for (int j = m - 3; j >= 0; j--) {
            const double beta = tau[j];
            if (beta == 0.0) continue;
            const double *v = A + (j + 1) + (int64_t)j * m;
            const int n2 = m - j - 1;
            double sacc = 0.0;
            for (int i = lane; i < n2; i += 64) sacc += v[i] * z[j + 1 + i];
            sacc = beta * wave_sum(sacc);
            for (int i = lane; i < n2; i += 64) z[j + 1 + i] -= sacc * v[i];
            // the next reflector's dot product reads entries other lanes of
            // this wavefront just wrote: keep the compiler from reordering
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (tid == 0) *info = sint[0];
}

}  // namespace

int64_t symeig_work_size(int m, int k) { return (int64_t)m + (int64_t)k * 5 * m; }

void sym_eig_top(int m, double *A, int k, double *theta, double *Zs, double *work, int *info, hipStream_t s) {
    symeig_top_kernel<<<1, NT, 0, s>>>(m, A, k, theta, Zs, work, info);
}
