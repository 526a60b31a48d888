// elements.hip -- element stiffness / geometric stiffness kernels for gfx950.
//
// CQUAD4 (MacNeal, default flat-isotropic branch) is computed by one 64-lane
// wavefront per element: lane l owns the 3x3 sub-block (l/8, l%8) of the
// 24x24 element matrix, so the per-element set-up (Gauss-point Jacobians,
// incompatible-mode couplings, MacNeal shear compliances) is staged once in
// LDS and the 576 outputs are produced without redundancy.  Results go to an
// element-block buffer through an LDS transpose so every wavefront stores
// 4.6 KB contiguously.  Four elements share a 256-thread workgroup.
//
// Mathematics restated from the reference:
//   frame .......... JFEM/src/solver/helpers.jl:1134-1193 (:diag mode)
//   Ke ............. JFEM/src/FEMKernels.jl:3185-4034 (macneal, twist row,
//                    Hughes-Brezzi drilling k6rot/1e5, Wilson membrane and
//                    bending incompatible modes condensed)
//   MacNeal shear .. JFEM/src/FEMKernels.jl:4715-4978 (eps 0.04, zb 0.65)
//   T'KT ........... JFEM/src/solver/assembly.jl:3808-3837
//   KG ............. JFEM/src/FEMKernels.jl:6341-6539, 6708-6785, 6891-6972,
//                    7554-7985; JFEM/src/solver/assembly.jl:5175-6007
//   CTRIA3 ......... JFEM/src/FEMKernels.jl:7129-7209, 7315-7432, 7989-8103;
//                    JFEM/src/solver/assembly.jl:3862-3955, 6020-6160
//   CBAR ........... JFEM/src/FEMKernels.jl:140-229, 7441-7471;
//                    JFEM/src/solver/assembly.jl:3960-4020, 6272-6345
#include "common.h"
#include <cmath>

namespace {

constexpr double PT = 0.57735026918962576451;  // 1/sqrt(3)
constexpr int EPB = 4;                         // elements (wavefronts) per workgroup

// sum over the 64 lanes of a wavefront (butterfly, same order on every lane)
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void normalize3(double *a) {
    double n = sqrt(dot3(a, a));
    a[0] /= n; a[1] /= n; a[2] /= n;
}

// CQUAD4 element frame (rows of R = v1, v2, v3) and local corner coordinates.
__device__ void q4_frame(const double *p /*12*/, double *R /*9*/, double *lc /*8*/) {
    double gxi[3], geta[3], n[3], d13[3], d24[3];
    for (int i = 0; i < 3; i++) {
        gxi[i] = 0.25 * (-p[i] + p[3 + i] + p[6 + i] - p[9 + i]);
        geta[i] = 0.25 * (-p[i] - p[3 + i] + p[6 + i] + p[9 + i]);
        d13[i] = p[6 + i] - p[i];
        d24[i] = p[9 + i] - p[3 + i];
    }
    cross3(gxi, geta, n);
    if (dot3(n, n) <= 1e-24) cross3(d13, d24, n);
    normalize3(n);
    normalize3(d13);
    normalize3(d24);
    double x[3] = {d13[0] - d24[0], d13[1] - d24[1], d13[2] - d24[2]};
    if (dot3(x, x) < 1e-20) { x[0] = d13[0] + d24[0]; x[1] = d13[1] + d24[1]; x[2] = d13[2] + d24[2]; }
    double xd = dot3(x, n);
    double xp[3] = {x[0] - xd * n[0], x[1] - xd * n[1], x[2] - xd * n[2]};
    if (dot3(xp, xp) <= 1e-24) {
        double gd = dot3(geta, n);
        for (int i = 0; i < 3; i++) xp[i] = geta[i] - gd * n[i];
        if (dot3(xp, xp) <= 1e-24) {
            double e[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
            double ed = dot3(e, n);
            for (int i = 0; i < 3; i++) xp[i] = e[i] - ed * n[i];
        }
    }
    normalize3(xp);
    double v2[3];
    cross3(n, xp, v2);
    for (int i = 0; i < 3; i++) { R[i] = xp[i]; R[3 + i] = v2[i]; R[6 + i] = n[i]; }
    double c[3];
    for (int i = 0; i < 3; i++) c[i] = (p[i] + p[3 + i] + p[6 + i] + p[9 + i]) / 4.0;
    for (int k = 0; k < 4; k++) {
        double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
        lc[2 * k] = dot3(d, xp);
        lc[2 * k + 1] = dot3(d, v2);
    }
}

__device__ __forceinline__ void shape_derivs(double r, double s, double *dNr, double *dNs) {
    dNr[0] = -0.25 * (1 - s); dNr[1] = 0.25 * (1 - s); dNr[2] = 0.25 * (1 + s); dNr[3] = -0.25 * (1 + s);
    dNs[0] = -0.25 * (1 - r); dNs[1] = -0.25 * (1 + r); dNs[2] = 0.25 * (1 + r); dNs[3] = 0.25 * (1 - r);
}
__device__ __forceinline__ void jacobian(const double *lc, const double *dNr, const double *dNs, double &J11,
                                         double &J12, double &J21, double &J22) {
    J11 = dNr[0] * lc[0] + dNr[1] * lc[2] + dNr[2] * lc[4] + dNr[3] * lc[6];
    J12 = dNr[0] * lc[1] + dNr[1] * lc[3] + dNr[2] * lc[5] + dNr[3] * lc[7];
    J21 = dNs[0] * lc[0] + dNs[1] * lc[2] + dNs[2] * lc[4] + dNs[3] * lc[6];
    J22 = dNs[0] * lc[1] + dNs[1] * lc[3] + dNs[2] * lc[5] + dNs[3] * lc[7];
}

// 4x4 inverse by Gauss-Jordan with partial pivoting (single lane).
__device__ void inv4(const double *a, double *ai) {
    // Gauss-Jordan with partial pivoting; every index is a compile-time
    // constant after unrolling (the row swap is a select over the candidate
    // rows), so the working matrix stays in registers
    double m[4][8];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) { m[i][j] = a[i * 4 + j]; m[i][4 + j] = (i == j) ? 1.0 : 0.0; }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int p = c;
        double best = fabs(m[c][c]);
#pragma unroll
        for (int r = c + 1; r < 4; r++)
            if (fabs(m[r][c]) > best) { best = fabs(m[r][c]); p = r; }
#pragma unroll
        for (int r = c + 1; r < 4; r++) {
            const bool sw = (r == p);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double t = m[c][j];
                m[c][j] = sw ? m[r][j] : t;
                m[r][j] = sw ? t : m[r][j];
            }
        }
        double d = m[c][c];
#pragma unroll
        for (int j = 0; j < 8; j++) m[c][j] /= d;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            double f = m[r][c];
#pragma unroll
            for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) ai[i * 4 + j] = m[i][4 + j];
}

// Per-element LDS staging of the CQUAD4 kernels.
struct Q4Stage {
    double lc[8];
    double R[9];
    double gdx[4][4], gdy[4][4], gN[4][4];  // per Gauss point, per corner
    double gadj[4];                          // |det J| (clamped at 1e-12)
    double gdp[4][4];                        // dphi1/dx, dphi1/dy, dphi2/dx, dphi2/dy
    double dNdx_c[4], dNdy_c[4];
    double Kab[24][4];                       // membrane incompatible coupling
    double Kabb[24][4];                      // bending incompatible coupling
    double Kbb[16], Kbbb[16];
    double Kbbi[16], Kbbbi[16];
    double D[4][12];                         // MacNeal shear strain-displacement
    double Jp[4];
    double Zi[16];                           // (Zs + Zb)^-1
    double Zs[16], Zb[16];
    double ZiD[4][12];
    double out[576];                         // 24x24 staging (element-block order)
};

// membrane B column (3-vector) of element DOF a at Gauss point g
__device__ __forceinline__ void bm_col(const Q4Stage &S, int g, int a, double *b) {
    int k = a / 6, c = a % 6;
    double dx = S.gdx[g][k], dy = S.gdy[g][k];
    b[0] = (c == 0) ? dx : 0.0;
    b[1] = (c == 1) ? dy : 0.0;
    b[2] = (c == 0) ? dy : ((c == 1) ? dx : 0.0);
}
// bending B column with the MacNeal twist row (S: Q4Stage or a virtual quad
// of the CTRIA3 macro element, which carry the same geometry members)
template <class ST>
__device__ __forceinline__ void bb_col(const ST &S, int g, int a, double *b) {
    int k = a / 6, c = a % 6;
    double dx = S.gdx[g][k], dy = S.gdy[g][k];
    if (c == 4) { b[0] = dx; b[1] = 0.0; b[2] = 2.0 * dy - S.dNdy_c[k]; }
    else if (c == 3) { b[0] = 0.0; b[1] = -dy; b[2] = 2.0 * (-dx) - (-S.dNdx_c[k]); }
    else { b[0] = 0.0; b[1] = 0.0; b[2] = 0.0; }
}
__device__ __forceinline__ double bd_val(const Q4Stage &S, int g, int a) {
    int k = a / 6, c = a % 6;
    if (c == 0) return 0.5 * S.gdy[g][k];
    if (c == 1) return -0.5 * S.gdx[g][k];
    if (c == 5) return S.gN[g][k];
    return 0.0;
}
// Wilson membrane mode column m
__device__ __forceinline__ void bi_col(const Q4Stage &S, int g, int m, double *b) {
    double p1x = S.gdp[g][0], p1y = S.gdp[g][1], p2x = S.gdp[g][2], p2y = S.gdp[g][3];
    switch (m) {
        case 0: b[0] = p1x; b[1] = 0.0; b[2] = p1y; break;
        case 1: b[0] = 0.0; b[1] = p1y; b[2] = p1x; break;
        case 2: b[0] = p2x; b[1] = 0.0; b[2] = p2y; break;
        default: b[0] = 0.0; b[1] = p2y; b[2] = p2x; break;
    }
}
// bending incompatible mode column m
__device__ __forceinline__ void bib_col(const Q4Stage &S, int g, int m, double *b) {
    double p1x = S.gdp[g][0], p1y = S.gdp[g][1], p2x = S.gdp[g][2], p2y = S.gdp[g][3];
    switch (m) {
        case 0: b[0] = 0.0; b[1] = -p1y; b[2] = -p1x; break;
        case 1: b[0] = 0.0; b[1] = -p2y; b[2] = -p2x; break;
        case 2: b[0] = p1x; b[1] = 0.0; b[2] = p1y; break;
        default: b[0] = p2x; b[1] = 0.0; b[2] = p2y; break;
    }
}
__device__ __forceinline__ void mat3v(const double *C, const double *b, double *o) {
    o[0] = C[0] * b[0] + C[1] * b[1] + C[2] * b[2];
    o[1] = C[3] * b[0] + C[4] * b[1] + C[5] * b[2];
    o[2] = C[6] * b[0] + C[7] * b[1] + C[8] * b[2];
}

struct ShellProps {
    double Cm[9], Cb[9], gs;  // Cs = gs * I (isotropic)
    double h, E, alpha_drill;
    bool bend;                // 12I/T^3 > 1e-12
};

__device__ void shell_props(const double *prop, double k6rot, ShellProps &P) {
    double E = prop[0], nu = prop[1], h = prop[2], br = prop[3], tst = prop[4], Gs = prop[5];
    double cm = E * h / (1 - nu * nu);
    double cb = br * (E * h * h * h) / (12 * (1 - nu * nu));
    const double D[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
    for (int i = 0; i < 9; i++) { P.Cm[i] = cm * D[i]; P.Cb[i] = cb * D[i]; }
    P.bend = br > 1e-12;
    P.gs = P.bend ? tst * h * Gs : 0.0;
    P.h = h;
    P.E = E;
    double ek6 = P.bend ? k6rot : 0.0;
    double eds = P.bend ? 1.0 : 0.0;
    double Gd = P.Cm[8] / h;
    if (Gd < 1e-6) Gd = E / (2 * 3.0);
    P.alpha_drill = eds * (ek6 / 1e5) * Gd * h;
}

// Stage Gauss-point data (lanes 0-3), centre derivatives (lane 4) and MacNeal
// sample points (lanes 8-11).
template <class ST>
__device__ void q4_stage_geometry(ST &S, int lane) {
    if (lane < 4) {
        // Gauss points (-,-), (+,-), (+,+), (-,+)
        const double r = (lane == 1 || lane == 2) ? PT : -PT, s = (lane >= 2) ? PT : -PT;
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jacobian(S.lc, dNr, dNs, J11, J12, J21, J22);
        double detJ = J11 * J22 - J12 * J21;
        double adj = fabs(detJ);
        if (adj < 1e-12) adj = 1e-12;
        double id = 1.0 / detJ;
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            S.gdx[lane][k] = iJ11 * dNr[k] + iJ12 * dNs[k];
            S.gdy[lane][k] = iJ21 * dNr[k] + iJ22 * dNs[k];
            S.gN[lane][k] = 0.25 * (1 + ((k == 1 || k == 2) ? r : -r)) * (1 + (k >= 2 ? s : -s));
        }
        S.gadj[lane] = adj;
        S.gdp[lane][0] = iJ11 * (-2.0 * r);
        S.gdp[lane][1] = iJ21 * (-2.0 * r);
        S.gdp[lane][2] = iJ12 * (-2.0 * s);
        S.gdp[lane][3] = iJ22 * (-2.0 * s);
    } else if (lane == 4) {
        const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
        double J11, J12, J21, J22;
        jacobian(S.lc, dNr_c, dNs_c, J11, J12, J21, J22);
        double id = 1.0 / (J11 * J22 - J12 * J21);
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
        for (int k = 0; k < 4; k++) {
            S.dNdx_c[k] = iJ11 * dNr_c[k] + iJ12 * dNs_c[k];
            S.dNdy_c[k] = iJ21 * dNr_c[k] + iJ22 * dNs_c[k];
        }
    } else if (lane >= 8 && lane < 12) {
        int p = lane - 8;
        // sample points (0,-), (0,+), (-,0), (+,0)
        const double xi = p < 2 ? 0.0 : (p == 2 ? -PT : PT), eta = p == 0 ? -PT : (p == 1 ? PT : 0.0);
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(xi, eta, dNr, dNs);
        jacobian(S.lc, dNr, dNs, J11, J12, J21, J22);
        double detJ = J11 * J22 - J12 * J21;
        if (fabs(detJ) < 1e-14) detJ = detJ < 0.0 ? -1e-14 : 1e-14;
        S.Jp[p] = fabs(detJ);
        double id = 1.0 / detJ;
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
        double N[4] = {0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta), 0.25 * (1 + xi) * (1 + eta),
                       0.25 * (1 - xi) * (1 + eta)};
        for (int k = 0; k < 4; k++) {
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k];
            double dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            if (p < 2) { S.D[p][3 * k] = dx; S.D[p][3 * k + 1] = 0.0; S.D[p][3 * k + 2] = N[k]; }
            else { S.D[p][3 * k] = dy; S.D[p][3 * k + 1] = -N[k]; S.D[p][3 * k + 2] = 0.0; }
        }
    }
}

// MacNeal compliance Zs, Zb and Zi = (Zs + Zb)^-1 (single lane).
template <class ST>
__device__ void q4_macneal_z(ST &S, const ShellProps &P) {
    const double *lc = S.lc;
    double Dx = 0.5 * (lc[2] + lc[4] - lc[0] - lc[6]);
    double Dy = 0.5 * (lc[5] + lc[7] - lc[1] - lc[3]);
    double Dx2 = Dx * Dx, Dy2 = Dy * Dy;
    const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11, J12, J21, J22;
    jacobian(lc, dNr_c, dNs_c, J11, J12, J21, J22);
    double A_elem = 4.0 * fabs(J11 * J22 - J12 * J21);
    const double eps = 0.04, zb = 0.65;
    double a_par = eps / (eps + (1.0 - eps) * Dx2 / fmax(Dy2, 1e-30));
    double b_par = eps / (eps + (1.0 - eps) * Dy2 / fmax(Dx2, 1e-30));
    double flex_x = 1.0 / fmax(fabs(P.Cb[0]), 1e-30);
    double flex_y = 1.0 / fmax(fabs(P.Cb[4]), 1e-30);
    double inv12A = 1.0 / (12.0 * A_elem);
    double zbx_u = zb * inv12A * Dx2 * flex_x;
    double zbx_d = zb * inv12A * a_par * Dx2 * flex_x;
    double zby_u = zb * inv12A * Dy2 * flex_y;
    double zby_d = zb * inv12A * b_par * Dy2 * flex_y;
    for (int i = 0; i < 16; i++) S.Zb[i] = 0.0;
    S.Zb[0] = zbx_u + zbx_d; S.Zb[1] = zbx_u - zbx_d; S.Zb[4] = S.Zb[1]; S.Zb[5] = S.Zb[0];
    S.Zb[10] = zby_u + zby_d; S.Zb[11] = zby_u - zby_d; S.Zb[14] = S.Zb[11]; S.Zb[15] = S.Zb[10];
    // [V G V] is diagonal for an isotropic Cs (G_xy = 0)
    double VGV[16];
    for (int i = 0; i < 16; i++) VGV[i] = 0.0;
    for (int i = 0; i < 4; i++) {
        double Jf = sqrt(2.0 * S.Jp[i]) * sqrt(2.0 * S.Jp[i]);
        double v = Jf * P.gs;
        VGV[i * 5] = v < 1e-30 ? 1e-30 : v;
    }
    double Zs[16];
    inv4(VGV, Zs);
    double Z[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double zs = 0.5 * (Zs[i * 4 + j] + Zs[j * 4 + i]);
            S.Zs[i * 4 + j] = zs;
        }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double a = S.Zs[i * 4 + j] + S.Zb[i * 4 + j];
            double b = S.Zs[j * 4 + i] + S.Zb[j * 4 + i];
            Z[i * 4 + j] = 0.5 * (a + b);
        }
    inv4(Z, S.Zi);
}

// Rotate a local 3x3 block to the global frame: out = R^T blk R.
__device__ __forceinline__ void rot_block(const double *R, const double *blk, double *out) {
    double tmp[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tmp[i * 3 + j] = blk[i * 3] * R[j] + blk[i * 3 + 1] * R[3 + j] + blk[i * 3 + 2] * R[6 + j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i * 3 + j] = R[i] * tmp[j] + R[3 + i] * tmp[3 + j] + R[6 + i] * tmp[6 + j];
}

// position of entry (a, b) of a 24x24 element matrix in the 16 x 36 block layout
__device__ __forceinline__ int eb_pos(int a, int b) {
    return ((a / 6) * 4 + (b / 6)) * 36 + (a % 6) * 6 + (b % 6);
}

// Stores S.out (576 values in element-block order) coalesced.
__device__ __forceinline__ void store_576(const Q4Stage &S, double *dst, int lane) {
    for (int t = lane; t < 576; t += S105_WAVE) dst[t] = S.out[t];
}

// Sum over the element's 2x2 Gauss points of the Ke entry (a, b) parts that
// do not involve condensation or shear: membrane + bending + drilling.
// Components are returned separately (needed by the thickness derivative).
__device__ void q4_gauss_entry(const Q4Stage &S, const ShellProps &P, int a, int b, double &km, double &kb,
                               double &kd) {
    km = kb = kd = 0.0;
    for (int g = 0; g < 4; g++) {
        double adj = S.gadj[g];
        double ba[3], bb[3], t[3];
        bm_col(S, g, a, ba);
        bm_col(S, g, b, bb);
        mat3v(P.Cm, bb, t);
        double vm = ba[0] * t[0] + ba[1] * t[1] + ba[2] * t[2];
        bb_col(S, g, a, ba);
        bb_col(S, g, b, bb);
        mat3v(P.Cb, bb, t);
        double vb = ba[0] * t[0] + ba[1] * t[1] + ba[2] * t[2];
        km += adj * vm;
        kb += adj * vb;
        kd += adj * P.alpha_drill * bd_val(S, g, a) * bd_val(S, g, b);
    }
}

// Condensation corrections (membrane, bending) of entry (a, b).
__device__ __forceinline__ void q4_condense_entry(const Q4Stage &S, int a, int b, bool bend_c, double &cm,
                                                  double &cb) {
    cm = 0.0;
    cb = 0.0;
    for (int l = 0; l < 4; l++) {
        double tm = 0.0, tb = 0.0;
        for (int q = 0; q < 4; q++) {
            tm += S.Kbbi[l * 4 + q] * S.Kab[b][q];
            tb += S.Kbbbi[l * 4 + q] * S.Kabb[b][q];
        }
        cm += S.Kab[a][l] * tm;
        cb += S.Kabb[a][l] * tb;
    }
    if (!bend_c) cb = 0.0;
}

// plate DOF index (w, thx, thy per node) of element DOF a, or -1
__device__ __forceinline__ int plate_idx(int a) {
    int c = a % 6;
    return (c >= 2 && c <= 4) ? 3 * (a / 6) + (c - 2) : -1;
}

// Common per-element set-up of the Ke / sensitivity kernels (all lanes of the
// element's wavefront participate; contains the workgroup barriers).
__device__ bool q4_setup(Q4Stage &S, ShellProps &P, const ElemArgs &A, int e, int lane, bool &bend_c,
                         bool &has_shear) {
    bool valid = e < A.n_q4;
    double p[12];
    if (valid) {
        for (int k = 0; k < 4; k++) {
            int nd = A.q4[4 * e + k];
            for (int i = 0; i < 3; i++) p[3 * k + i] = A.xyz[3 * nd + i];
        }
        shell_props(A.q4prop + 6 * e, A.k6rot, P);
        if (lane == 0) q4_frame(p, S.R, S.lc);
    }
    __syncthreads();
    if (valid) q4_stage_geometry(S, lane);
    __syncthreads();
    if (valid) {
        // couplings: 96 membrane + 96 bending entries, then 16 + 16 self terms
        for (int t = lane; t < 224; t += S105_WAVE) {
            if (t < 192) {
                int bend = t >= 96;
                int idx = bend ? t - 96 : t;
                int a = idx / 4, m = idx % 4;
                double s = 0.0;
                for (int g = 0; g < 4; g++) {
                    double ba[3], bi[3], cbi[3];
                    if (bend) { bb_col(S, g, a, ba); bib_col(S, g, m, bi); mat3v(P.Cb, bi, cbi); }
                    else { bm_col(S, g, a, ba); bi_col(S, g, m, bi); mat3v(P.Cm, bi, cbi); }
                    s += S.gadj[g] * (ba[0] * cbi[0] + ba[1] * cbi[1] + ba[2] * cbi[2]);
                }
                if (bend) S.Kabb[a][m] = s; else S.Kab[a][m] = s;
            } else {
                int idx = t - 192;
                int bend = idx >= 16;
                int l = (idx % 16) / 4, m = idx % 4;
                double s = 0.0;
                for (int g = 0; g < 4; g++) {
                    double bl[3], bm[3], cbm[3];
                    if (bend) { bib_col(S, g, l, bl); bib_col(S, g, m, bm); mat3v(P.Cb, bm, cbm); }
                    else { bi_col(S, g, l, bl); bi_col(S, g, m, bm); mat3v(P.Cm, bm, cbm); }
                    s += S.gadj[g] * (bl[0] * cbm[0] + bl[1] * cbm[1] + bl[2] * cbm[2]);
                }
                if (bend) S.Kbbb[l * 4 + m] = s; else S.Kbb[l * 4 + m] = s;
            }
        }
    }
    __syncthreads();
    has_shear = valid && P.h >= 1e-30 && P.gs >= 1e-30;
    if (valid) {
        if (lane == 0) inv4(S.Kbb, S.Kbbi);
        if (lane == 1) inv4(S.Kbbb, S.Kbbbi);
        if (lane == 2 && has_shear) q4_macneal_z(S, P);
    }
    double kmax = 0.0;
    if (valid)
        for (int i = 0; i < 16; i++) kmax = fmax(kmax, fabs(S.Kbbb[i]));
    bend_c = kmax > 0.0;
    __syncthreads();
    if (valid && has_shear)
        for (int t = lane; t < 48; t += S105_WAVE) {
            int i = t / 12, c = t % 12;
            double s = 0.0;
            for (int l = 0; l < 4; l++) s += S.Zi[i * 4 + l] * S.D[l][c];
            S.ZiD[i][c] = s;
        }
    __syncthreads();
    return valid;
}

// Element stiffness of lane's 3x3 sub-block in the local frame.
__device__ void q4_local_block(const Q4Stage &S, const ShellProps &P, int lane, bool bend_c, bool has_shear,
                               double *blk) {
    int bi = lane / 8, bj = lane % 8;
    // kept rolled: unrolled, the nine entries' Gauss-point sums need more
    // registers than a lane has (spills, one wavefront per SIMD)
#pragma unroll 1
    for (int i = 0; i < 3; i++)
#pragma unroll 1
        for (int j = 0; j < 3; j++) {
            int a = 3 * bi + i, b = 3 * bj + j;
            double km, kb, kd, cm, cb;
            q4_gauss_entry(S, P, a, b, km, kb, kd);
            double v = km + kb + kd;
            if (has_shear) {
                int pa = plate_idx(a), pb = plate_idx(b);
                if (pa >= 0 && pb >= 0) {
                    double s1 = 0.0, s2 = 0.0;
                    for (int l = 0; l < 4; l++) {
                        s1 += S.D[l][pa] * S.ZiD[l][pb];
                        s2 += S.D[l][pb] * S.ZiD[l][pa];
                    }
                    v += 0.5 * (s1 + s2);
                }
            }
            q4_condense_entry(S, a, b, bend_c, cm, cb);
            v -= cb;
            v -= cm;
            blk[i * 3 + j] = v;
        }
}

__global__ void __launch_bounds__(256) q4_ke_kernel(ElemArgs A, double *eb) {
    __shared__ Q4Stage stage[EPB];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    Q4Stage &S = stage[w];
    ShellProps P;
    bool bend_c, has_shear;
    bool valid = q4_setup(S, P, A, e, lane, bend_c, has_shear);
    if (valid) {
        double blk[9], g[9];
        q4_local_block(S, P, lane, bend_c, has_shear, blk);
        rot_block(S.R, blk, g);
        int bi = lane / 8, bj = lane % 8;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) S.out[eb_pos(3 * bi + i, 3 * bj + j)] = g[i * 3 + j];
    }
    __syncthreads();
    if (valid) store_576(S, eb + (size_t)e * 576, lane);
}

// ---------------------------------------------------------------------------
// CQUAD4 geometric stiffness
// ---------------------------------------------------------------------------

// NM condensed membrane modes: the four Wilson modes (1 - r^2), (1 - s^2) on
// ux and uy of the forward KG, or six with the r s mode added (the
// "enhanced" set of the reference's adjoint path, FEMKernels.jl:1861-1896)
template <int NM>
struct Q4KGStageT {
    double lc[8];
    double R[9];
    double ul[24];
    double gdx[4][4], gdy[4][4], gadj[4], gdp[4][NM == 6 ? 6 : 4];  // + d(rs)/dx, d(rs)/dy with six modes
    double Kab[24][NM];
    double Kbb[NM * NM], Kbbi[NM * NM];
    double Am[NM][24];  // condensation map A = -Kbb^-1 Kab^T
    double alpha[NM];
    double sig[4][3];
    double pr[4][6];    // principal stresses per GP: l1 c1 s1 l2 c2 s2
    double va[8][24];   // (s gx - c gy) per (gp, direction)
    double vb[8][24];   // (c gzx + s gzy)
    double wt[8];
    double out[NM == 4 ? 576 : 1];  // the six-mode stage (adjoint kernels) writes no element block
};
using Q4KGStage = Q4KGStageT<4>;

template <int NM>
__device__ void q4kg_gp_stage(Q4KGStageT<NM> &S, int lane) {
    const double r = (lane == 1 || lane == 2) ? PT : -PT, s = (lane >= 2) ? PT : -PT;
    double dNr[4], dNs[4], J11, J12, J21, J22;
    shape_derivs(r, s, dNr, dNs);
    jacobian(S.lc, dNr, dNs, J11, J12, J21, J22);
    double detJ = J11 * J22 - J12 * J21;
    double adj = fabs(detJ);
    if (adj < 1e-12) adj = 1e-12;
    double id = 1.0 / detJ;
    double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
    for (int k = 0; k < 4; k++) {
        S.gdx[lane][k] = iJ11 * dNr[k] + iJ12 * dNs[k];
        S.gdy[lane][k] = iJ21 * dNr[k] + iJ22 * dNs[k];
    }
    S.gadj[lane] = adj;
    S.gdp[lane][0] = iJ11 * (-2.0 * r);
    S.gdp[lane][1] = iJ21 * (-2.0 * r);
    S.gdp[lane][2] = iJ12 * (-2.0 * s);
    S.gdp[lane][3] = iJ22 * (-2.0 * s);
    if constexpr (NM == 6) {
        S.gdp[lane][4] = iJ11 * s + iJ12 * r;
        S.gdp[lane][5] = iJ21 * s + iJ22 * r;
    }
}

// principal stress decomposition, radius <= 1e-30 returns the axes
__device__ void principal2d(double sxx, double syy, double sxy, double *o) {
    double mean = 0.5 * (sxx + syy), hd = 0.5 * (sxx - syy);
    double rad = sqrt(hd * hd + sxy * sxy);
    if (rad <= 1e-30) { o[0] = mean; o[1] = 1.0; o[2] = 0.0; o[3] = mean; o[4] = 0.0; o[5] = 1.0; return; }
    double th = 0.5 * atan2(2.0 * sxy, sxx - syy);
    double c = cos(th), s = sin(th);
    o[0] = mean + rad; o[1] = c; o[2] = s; o[3] = mean - rad; o[4] = -s; o[5] = c;
}

// membrane strain column (3-vector) of condensed mode m at GP g
template <int NM>
__device__ __forceinline__ void q4kg_mode_col(const Q4KGStageT<NM> &S, int g, int m, double *b) {
    const int j = 2 * (m / 2);  // first gdp entry of the mode's shape function
    const double px = S.gdp[g][j], py = S.gdp[g][j + 1];
    if (m % 2 == 0) { b[0] = px; b[1] = 0.0; b[2] = py; }
    else { b[0] = 0.0; b[1] = py; b[2] = px; }
}

// gradient vectors of (ux, uy, uz) along x/y at GP g for element DOF a,
// including the condensed modes.
template <int NM>
__device__ __forceinline__ void q4kg_grads(const Q4KGStageT<NM> &S, int g, int a, double *gr /*6*/) {
    int k = a / 6, c = a % 6;
    double dx = S.gdx[g][k], dy = S.gdy[g][k];
    double uxx = (c == 0) ? dx : 0.0, uxy = (c == 0) ? dy : 0.0;
    double uyx = (c == 1) ? dx : 0.0, uyy = (c == 1) ? dy : 0.0;
    double uzx = (c == 2) ? dx : 0.0, uzy = (c == 2) ? dy : 0.0;
    double p1x = S.gdp[g][0], p1y = S.gdp[g][1], p2x = S.gdp[g][2], p2y = S.gdp[g][3];
    double a1 = S.Am[0][a], a2 = S.Am[1][a], a3 = S.Am[2][a], a4 = S.Am[3][a];
    if constexpr (NM == 6) {
        double qx = S.gdp[g][4], qy = S.gdp[g][5], a5 = S.Am[4][a], a6 = S.Am[5][a];
        uxx += p1x * a1 + p2x * a3 + qx * a5;
        uxy += p1y * a1 + p2y * a3 + qy * a5;
        uyx += p1x * a2 + p2x * a4 + qx * a6;
        uyy += p1y * a2 + p2y * a4 + qy * a6;
    } else {
        uxx += p1x * a1 + p2x * a3;
        uxy += p1y * a1 + p2y * a3;
        uyx += p1x * a2 + p2x * a4;
        uyy += p1y * a2 + p2y * a4;
    }
    gr[0] = uxx; gr[1] = uxy; gr[2] = uyx; gr[3] = uyy; gr[4] = uzx; gr[5] = uzy;
}

// n x n inverse by Gauss-Jordan with partial pivoting (single lane; the
// six-mode Kbb)
template <int N>
__device__ void inv_gj(const double *a, double *ai) {
    double m[N][2 * N];
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) { m[i][j] = a[i * N + j]; m[i][N + j] = (i == j) ? 1.0 : 0.0; }
    for (int c = 0; c < N; c++) {
        int p = c;
        double best = fabs(m[c][c]);
        for (int r = c + 1; r < N; r++)
            if (fabs(m[r][c]) > best) { best = fabs(m[r][c]); p = r; }
        if (p != c)
            for (int j = 0; j < 2 * N; j++) { double t = m[c][j]; m[c][j] = m[p][j]; m[p][j] = t; }
        double d = m[c][c];
        for (int j = 0; j < 2 * N; j++) m[c][j] /= d;
        for (int r = 0; r < N; r++) {
            if (r == c) continue;
            double f = m[r][c];
            for (int j = 0; j < 2 * N; j++) m[r][j] -= f * m[c][j];
        }
    }
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) ai[i * N + j] = m[i][N + j];
}

// Shared set-up of the KG and sensitivity kernels: frame, local u, GP data,
// condensation map, membrane stresses and principal directions.
template <int NM>
__device__ bool q4kg_setup(Q4KGStageT<NM> &S, const ElemArgs &A, const double *u, int e, int lane, double &h) {
    bool valid = e < A.n_q4;
    double Cm[9];
    h = 1.0;
    if (valid) {
        const double *prop = A.q4prop + 6 * e;
        double E = prop[0], nu = prop[1];
        h = prop[2];
        double cm = E / (1 - nu * nu);
        const double D[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
        for (int i = 0; i < 9; i++) Cm[i] = (cm * D[i]) * h;
        if (lane == 0) {
            double p[12];
            for (int k = 0; k < 4; k++) {
                int nd = A.q4[4 * e + k];
                for (int i = 0; i < 3; i++) p[3 * k + i] = A.xyz[3 * nd + i];
            }
            q4_frame(p, S.R, S.lc);
        }
    }
    __syncthreads();
    if (valid) {
        if (lane < 4) q4kg_gp_stage(S, lane);
        if (lane >= 8 && lane < 16) {
            int b = lane - 8;  // 3-vector block b of the element DOFs
            int nd = A.q4[4 * e + b / 2];
            const double *ug = u + 6 * (size_t)nd + 3 * (b % 2);
            for (int i = 0; i < 3; i++) S.ul[3 * b + i] = S.R[i * 3] * ug[0] + S.R[i * 3 + 1] * ug[1] + S.R[i * 3 + 2] * ug[2];
        }
    }
    __syncthreads();
    if (valid) {
        for (int t = lane; t < 24 * NM + NM * NM; t += S105_WAVE) {
            if (t < 24 * NM) {
                int a = t / NM, m = t % NM;
                int k = a / 6, c = a % 6;
                double s = 0.0;
                for (int g = 0; g < 4; g++) {
                    double dx = S.gdx[g][k], dy = S.gdy[g][k];
                    double ba[3] = {(c == 0) ? dx : 0.0, (c == 1) ? dy : 0.0, (c == 0) ? dy : ((c == 1) ? dx : 0.0)};
                    double bi[3];
                    q4kg_mode_col(S, g, m, bi);
                    double cbi[3];
                    mat3v(Cm, bi, cbi);
                    s += (ba[0] * cbi[0] + ba[1] * cbi[1] + ba[2] * cbi[2]) * S.gadj[g];
                }
                S.Kab[a][m] = s;
            } else {
                int idx = t - 24 * NM;
                int l = idx / NM, m = idx % NM;
                double s = 0.0;
                for (int g = 0; g < 4; g++) {
                    double bl[3], bm[3], cbm[3];
                    q4kg_mode_col(S, g, l, bl);
                    q4kg_mode_col(S, g, m, bm);
                    mat3v(Cm, bm, cbm);
                    s += (bl[0] * cbm[0] + bl[1] * cbm[1] + bl[2] * cbm[2]) * S.gadj[g];
                }
                S.Kbb[l * NM + m] = s;
            }
        }
    }
    __syncthreads();
    if (valid && lane == 0) {
        if constexpr (NM == 4) inv4(S.Kbb, S.Kbbi);
        else inv_gj<NM>(S.Kbb, S.Kbbi);
    }
    __syncthreads();
    if (valid)
        for (int t = lane; t < NM * 24; t += S105_WAVE) {
            int l = t / 24, a = t % 24;
            double s = 0.0;
            for (int q = 0; q < NM; q++) s += S.Kbbi[l * NM + q] * S.Kab[a][q];
            S.Am[l][a] = -s;
        }
    __syncthreads();
    if (valid && lane < NM) {
        double s = 0.0;
        for (int a = 0; a < 24; a++) s += S.Am[lane][a] * S.ul[a];
        S.alpha[lane] = s;
    }
    __syncthreads();
    if (valid && lane < 4) {
        int g = lane;
        double eps[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 4; k++) {
            double dx = S.gdx[g][k], dy = S.gdy[g][k];
            eps[0] += dx * S.ul[6 * k];
            eps[1] += dy * S.ul[6 * k + 1];
            eps[2] += dy * S.ul[6 * k] + dx * S.ul[6 * k + 1];
        }
        double p1x = S.gdp[g][0], p1y = S.gdp[g][1], p2x = S.gdp[g][2], p2y = S.gdp[g][3];
        eps[0] += p1x * S.alpha[0] + p2x * S.alpha[2];
        eps[1] += p1y * S.alpha[1] + p2y * S.alpha[3];
        eps[2] += p1y * S.alpha[0] + p1x * S.alpha[1] + p2y * S.alpha[2] + p2x * S.alpha[3];
        if constexpr (NM == 6) {
            double qx = S.gdp[g][4], qy = S.gdp[g][5];
            eps[0] += qx * S.alpha[4];
            eps[1] += qy * S.alpha[5];
            eps[2] += qy * S.alpha[4] + qx * S.alpha[5];
        }
        double N[3];
        mat3v(Cm, eps, N);
        for (int i = 0; i < 3; i++) S.sig[g][i] = N[i] / h;
        principal2d(S.sig[g][0], S.sig[g][1], S.sig[g][2], S.pr[g]);
    }
    __syncthreads();
    return valid;
}

// va/vb vectors and weights for the 8 (gp, direction) rank-2 updates.
template <int NM>
__device__ void q4kg_vectors(Q4KGStageT<NM> &S, int lane, double h) {
    for (int t = lane; t < 8 * 24; t += S105_WAVE) {
        int q = t / 24, a = t % 24;
        int g = q / 2, k = q % 2;
        double c = S.pr[g][3 * k + 1], s = S.pr[g][3 * k + 2];
        double gr[6];
        q4kg_grads(S, g, a, gr);
        double gx = c * gr[0] + s * gr[1], gy = c * gr[2] + s * gr[3], gz = c * gr[4] + s * gr[5];
        S.va[q][a] = s * gx - c * gy;
        S.vb[q][a] = gz;
    }
    if (lane < 8) {
        int g = lane / 2, k = lane % 2;
        double lam = S.pr[g][3 * k];
        S.wt[lane] = (fabs(lam) <= 1e-30) ? 0.0 : h * S.gadj[g] * lam;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) q4_kg_kernel(ElemArgs A, const double *u, double *eb) {
    __shared__ Q4KGStage stage[EPB];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    Q4KGStage &S = stage[w];
    double h;
    bool valid = q4kg_setup(S, A, u, e, lane, h);
    if (valid) q4kg_vectors(S, lane, h);
    else __syncthreads();
    if (valid) {
        int bi = lane / 8, bj = lane % 8;
        double blk[9], g[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int a = 3 * bi + i, b = 3 * bj + j;
                double v = 0.0;
                for (int q = 0; q < 8; q++) v += S.wt[q] * (S.va[q][a] * S.va[q][b] + S.vb[q][a] * S.vb[q][b]);
                blk[i * 3 + j] = v;
            }
        rot_block(S.R, blk, g);
        // drilling rows/columns (global component 6 of each node) are zeroed
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                bool drill = ((bi % 2) == 1 && i == 2) || ((bj % 2) == 1 && j == 2);
                S.out[eb_pos(3 * bi + i, 3 * bj + j)] = drill ? 0.0 : g[i * 3 + j];
            }
    }
    __syncthreads();
    if (valid)
        for (int t = lane; t < 576; t += S105_WAVE) eb[(size_t)e * 576 + t] = S.out[t];
}

// ---------------------------------------------------------------------------
// CBAR (one thread per bar; 12x12 local matrices are analytic and sparse)
// ---------------------------------------------------------------------------

struct BarLocal {
    double R[9];
    double L, X, T, ay, by, cy, dy, az, bz, cz, dz;
};

__device__ void bar_setup(const ElemArgs &A, int b, BarLocal &B) {
    const double *p1 = A.xyz + 3 * A.bar[2 * b], *p2 = A.xyz + 3 * A.bar[2 * b + 1];
    double vx[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    B.L = sqrt(dot3(vx, vx));
    normalize3(vx);
    double vr[3] = {A.bar_vref[3 * b], A.bar_vref[3 * b + 1], A.bar_vref[3 * b + 2]};
    if (sqrt(dot3(vr, vr)) < 1e-6) {
        vr[0] = 0.0; vr[1] = 0.0; vr[2] = 1.0;
        if (fabs(dot3(vx, vr)) > 0.9) { vr[1] = 1.0; vr[2] = 0.0; }
    }
    double vz[3], vy[3];
    cross3(vx, vr, vz);
    normalize3(vz);
    cross3(vz, vx, vy);
    for (int i = 0; i < 3; i++) { B.R[i] = vx[i]; B.R[3 + i] = vy[i]; B.R[6 + i] = vz[i]; }
    const double *bp = A.barprop + 8 * b;
    double Ar = bp[0], I1 = bp[1], I2 = bp[2], J = bp[3], K1 = bp[4], K2 = bp[5], E = bp[6], G = bp[7];
    double Iy = I2, Iz = I1;
    if (Iy == 0.0) Iy = Iz;
    if (Iz == 0.0) Iz = Iy;
    double L = B.L;
    bool ok = L >= 1e-9;
    B.X = ok ? E * Ar / L : 0.0;
    B.T = ok ? G * J / L : 0.0;
    double Py = (K2 > 0.0) ? 12 * E * Iy / (G * (K2 * Ar) * L * L) : 0.0;
    double Pz = (K1 > 0.0) ? 12 * E * Iz / (G * (K1 * Ar) * L * L) : 0.0;
    B.ay = ok ? 12 * E * Iy / (L * L * L * (1 + Py)) : 0.0;
    B.by = ok ? 6 * E * Iy / (L * L * (1 + Py)) : 0.0;
    B.cy = ok ? (4 + Py) * E * Iy / (L * (1 + Py)) : 0.0;
    B.dy = ok ? (2 - Py) * E * Iy / (L * (1 + Py)) : 0.0;
    B.az = ok ? 12 * E * Iz / (L * L * L * (1 + Pz)) : 0.0;
    B.bz = ok ? 6 * E * Iz / (L * L * (1 + Pz)) : 0.0;
    B.cz = ok ? (4 + Pz) * E * Iz / (L * (1 + Pz)) : 0.0;
    B.dz = ok ? (2 - Pz) * E * Iz / (L * (1 + Pz)) : 0.0;
}

// local Timoshenko stiffness entry (0-based i, j)
__device__ double bar_k(const BarLocal &B, int i, int j) {
    int I = i + 1, J = j + 1;
    auto pair = [&](int a, int b) { return (I == a && J == b) || (I == b && J == a); };
    if (pair(1, 1) || pair(7, 7)) return B.X;
    if (pair(1, 7)) return -B.X;
    if (pair(4, 4) || pair(10, 10)) return B.T;
    if (pair(4, 10)) return -B.T;
    if (pair(3, 3) || pair(9, 9)) return B.ay;
    if (pair(3, 9)) return -B.ay;
    if (pair(3, 5) || pair(3, 11)) return -B.by;
    if (pair(9, 5) || pair(9, 11)) return B.by;
    if (pair(5, 5) || pair(11, 11)) return B.cy;
    if (pair(5, 11)) return B.dy;
    if (pair(2, 2) || pair(8, 8)) return B.az;
    if (pair(2, 8)) return -B.az;
    if (pair(2, 6) || pair(2, 12)) return B.bz;
    if (pair(8, 6) || pair(8, 12)) return -B.bz;
    if (pair(6, 6) || pair(12, 12)) return B.cz;
    if (pair(6, 12)) return B.dz;
    return 0.0;
}

// local beam-column KG entry for axial force P
__device__ double bar_kg(double L, double P, int i, int j) {
    if (L < 1e-9 || fabs(P) < 1e-30) return 0.0;
    double c1 = 6.0 * P / (5.0 * L), c2 = P / 10.0, c3 = 2.0 * P * L / 15.0, c4 = -P * L / 30.0;
    int I = i + 1, J = j + 1;
    auto pair = [&](int a, int b) { return (I == a && J == b) || (I == b && J == a); };
    if (pair(2, 2) || pair(8, 8) || pair(3, 3) || pair(9, 9)) return c1;
    if (pair(2, 8) || pair(3, 9)) return -c1;
    if (pair(2, 6) || pair(2, 12)) return c2;
    if (pair(6, 6) || pair(12, 12) || pair(5, 5) || pair(11, 11)) return c3;
    if (pair(6, 12) || pair(5, 11)) return c4;
    if (pair(8, 6) || pair(8, 12)) return -c2;
    if (pair(3, 5) || pair(3, 11)) return -c2;
    if (pair(9, 5) || pair(9, 11)) return c2;
    return 0.0;
}

__global__ void bar_ke_kernel(ElemArgs A, double *eb) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.n_bar) return;
    BarLocal B;
    bar_setup(A, b, B);
    double *dst = eb + (size_t)b * 144;
    for (int bi = 0; bi < 4; bi++)
        for (int bj = 0; bj < 4; bj++) {
            double blk[9], g[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) blk[i * 3 + j] = bar_k(B, 3 * bi + i, 3 * bj + j);
            rot_block(B.R, blk, g);
            int na = bi / 2, nb = bj / 2;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    dst[(na * 2 + nb) * 36 + ((bi % 2) * 3 + i) * 6 + (bj % 2) * 3 + j] = g[i * 3 + j];
        }
}

__device__ double bar_axial(const ElemArgs &A, const BarLocal &B, int b, const double *u) {
    double ul[12];
    for (int blk = 0; blk < 4; blk++) {
        int nd = A.bar[2 * b + blk / 2];
        const double *ug = u + 6 * (size_t)nd + 3 * (blk % 2);
        for (int i = 0; i < 3; i++)
            ul[3 * blk + i] = B.R[i * 3] * ug[0] + B.R[i * 3 + 1] * ug[1] + B.R[i * 3 + 2] * ug[2];
    }
    double P = 0.0;
    for (int j = 0; j < 12; j++) P += bar_k(B, 6, j) * ul[j];
    return P;
}

__global__ void bar_kg_kernel(ElemArgs A, const double *u, double *eb) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.n_bar) return;
    BarLocal B;
    bar_setup(A, b, B);
    double P = bar_axial(A, B, b, u);
    double *dst = eb + (size_t)b * 144;
    for (int bi = 0; bi < 4; bi++)
        for (int bj = 0; bj < 4; bj++) {
            double blk[9], g[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) blk[i * 3 + j] = bar_kg(B.L, P, 3 * bi + i, 3 * bj + j);
            rot_block(B.R, blk, g);
            int na = bi / 2, nb = bj / 2;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    dst[(na * 2 + nb) * 36 + ((bi % 2) * 3 + i) * 6 + (bj % 2) * 3 + j] = g[i * 3 + j];
        }
}

// ---------------------------------------------------------------------------
// CTRIA3, one 64-lane wavefront per element, four elements per workgroup.
//   Ke = CST membrane Bm' Cm Bm A on (u, v)
//      + macro plate block on (w, thx, thy): the plate rows/columns of three
//        virtual MacNeal CQUAD4s on the corners, mid-sides and centroid
//        (kernel default flags: bending incompatible modes kept), assembled
//        21 x 21 and the 12 interior DOFs condensed by partially pivoted LU
//      + drilling alpha (Bd' Bd) A, alpha = (k6rot/1e5) G_ref h, at the centroid.
//   FEMKernels.jl:7129-7209, 7315-7385; assembly.jl:3862-3955.
//   KG = h A (sxx b_i b_j + syy c_i c_j + sxy (b_i c_j + c_i b_j)) on u, v, w
//   from the constant-strain stress.  FEMKernels.jl:7391-7432, 7989-8103;
//   assembly.jl:6020-6160.
// ---------------------------------------------------------------------------

// One virtual quad of the macro plate block; member names follow Q4Stage so
// the CQUAD4 staging helpers apply unchanged.
struct T3Quad {
    double lc[8];
    double gdx[4][4], gdy[4][4], gN[4][4];
    double gadj[4];
    double gdp[4][4];
    double dNdx_c[4], dNdy_c[4];
    double D[4][12];
    double Jp[4];
    double Zi[16], Zs[16], Zb[16];
    double ZiD[4][12];
};

struct T3Stage {
    double R[9], lc[6];
    double bv[3], cv[3], area;
    T3Quad q[3];
    double K21[21 * 21];   // assembled plate matrix of the 7 macro points
    double M[12 * 21];     // [Kii | Kie] -> LU factors | Kii^-1 Kie
    double Kc[81];         // condensed plate block (corner w, thx, thy)
    double out[324];       // 9 global node blocks of 36
};

// Virtual quads (corner, mid-side, centroid, mid-side) over the points
// 0-2 corners, 3-5 mid-sides 12/23/31, 6 centroid.
__constant__ int T3_QUAD[3][4] = {{0, 3, 6, 5}, {1, 4, 6, 3}, {2, 5, 6, 4}};

// position of macro point n in virtual quad q, or -1
__device__ __forceinline__ int t3_pos(int q, int n) {
    for (int k = 0; k < 4; k++)
        if (T3_QUAD[q][k] == n) return k;
    return -1;
}

// CTRIA3 frame (v1 = e12, v3 = v1 x (p3 - p1), v2 = v3 x v1) and corner
// coordinates about the centroid.  helpers.jl:1195-1204, assembly.jl:3927-3930.
__device__ void t3_frame(const double *p /*9*/, double *R /*9*/, double *lc /*6*/) {
    double v1[3], v2[3], v3[3], d13[3];
    for (int i = 0; i < 3; i++) { v1[i] = p[3 + i] - p[i]; d13[i] = p[6 + i] - p[i]; }
    normalize3(v1);
    cross3(v1, d13, v3);
    normalize3(v3);
    cross3(v3, v1, v2);
    for (int i = 0; i < 3; i++) { R[i] = v1[i]; R[3 + i] = v2[i]; R[6 + i] = v3[i]; }
    double c[3];
    for (int i = 0; i < 3; i++) c[i] = (p[i] + p[3 + i] + p[6 + i]) / 3.0;
    for (int k = 0; k < 3; k++) {
        double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
        lc[2 * k] = dot3(d, v1);
        lc[2 * k + 1] = dot3(d, v2);
    }
}

// area and the constant shape-function derivatives b = dN/dx, c = dN/dy
__device__ double t3_derivs(const double *lc, double *bv, double *cv) {
    double x0 = lc[0], y0 = lc[1], x1 = lc[2], y1 = lc[3], x2 = lc[4], y2 = lc[5];
    double A = 0.5 * fabs(x0 * (y1 - y2) + x1 * (y2 - y0) + x2 * (y0 - y1));
    bv[0] = (y1 - y2) / (2 * A); bv[1] = (y2 - y0) / (2 * A); bv[2] = (y0 - y1) / (2 * A);
    cv[0] = (x2 - x1) / (2 * A); cv[1] = (x0 - x2) / (2 * A); cv[2] = (x1 - x0) / (2 * A);
    return A;
}

// Plate-block entry (a, b), plate indices 0..11 (w, thx, thy per corner),
// of a virtual quad: 2x2 Gauss bending with the MacNeal twist row plus the
// MacNeal shear D' Z^-1 D (symmetrised as the kernel does).
__device__ double t3_quad_plate_entry(const T3Quad &Q, const ShellProps &P, int a, int b, bool has_shear) {
    int a24 = 6 * (a / 3) + 2 + a % 3, b24 = 6 * (b / 3) + 2 + b % 3;
    double kb = 0.0;
    for (int g = 0; g < 4; g++) {
        double ba[3], bb[3], t[3];
        bb_col(Q, g, a24, ba);
        bb_col(Q, g, b24, bb);
        mat3v(P.Cb, bb, t);
        kb += Q.gadj[g] * (ba[0] * t[0] + ba[1] * t[1] + ba[2] * t[2]);
    }
    if (has_shear) {
        double s1 = 0.0, s2 = 0.0;
        for (int l = 0; l < 4; l++) {
            s1 += Q.D[l][a] * Q.ZiD[l][b];
            s2 += Q.D[l][b] * Q.ZiD[l][a];
        }
        kb += 0.5 * (s1 + s2);
    }
    return kb;
}

// Entry (a, b) of the local 18 x 18 CTRIA3 matrix: (membrane + drilling) *
// md_scale + the plate block Kp (condensed macro block or its derivative).
__device__ double t3_local_entry(const T3Stage &S, const ShellProps &P, double alpha, bool nonzero, bool plate,
                                 const double *Kp, double md_scale, int a, int b) {
    if (!nonzero) return 0.0;
    int I = a / 6, J = b / 6, ca = a % 6, cb = b % 6;
    const double ar = S.area;
    double v = 0.0;
    if (ca < 2 && cb < 2) {
        double ba[3] = {ca == 0 ? S.bv[I] : 0.0, ca == 1 ? S.cv[I] : 0.0, ca == 0 ? S.cv[I] : S.bv[I]};
        double bb[3] = {cb == 0 ? S.bv[J] : 0.0, cb == 1 ? S.cv[J] : 0.0, cb == 0 ? S.cv[J] : S.bv[J]};
        double t[3];
        mat3v(P.Cm, bb, t);
        v = (ba[0] * t[0] + ba[1] * t[1] + ba[2] * t[2]) * ar;
    }
    double da = ca == 0 ? 0.5 * S.cv[I] : (ca == 1 ? -0.5 * S.bv[I] : (ca == 5 ? 1.0 / 3.0 : 0.0));
    double db = cb == 0 ? 0.5 * S.cv[J] : (cb == 1 ? -0.5 * S.bv[J] : (cb == 5 ? 1.0 / 3.0 : 0.0));
    v += alpha * (da * db) * ar;
    v *= md_scale;
    if (plate && ca >= 2 && ca <= 4 && cb >= 2 && cb <= 4) v += Kp[(3 * I + ca - 2) * 9 + 3 * J + cb - 2];
    return v;
}

// Shared set-up of the CTRIA3 kernels (all lanes of the element's wavefront
// take part; contains the workgroup barriers): frame, area and shape
// derivatives, the three virtual quads, K21, the LU elimination of Kii with
// X = Kii^-1 Kie left in M[:, 12:21], and the condensed plate block Kc.
struct T3Flags {
    bool valid, nonzero, plate, has_shear;
};

// With `recovery` the plate block takes the result-recovery constitutive set
// (Cb = D h^3/12, Cs = 5/6 G h, G = E/2(1+nu), FEMKernels.jl:7271-7279)
// instead of the element's stiffness set.
__device__ T3Flags t3_setup(T3Stage &S, ShellProps &P, double &Gref, const ElemArgs &A, int e, int lane,
                           bool recovery = false) {
    bool valid = e < A.n_t3;
    Gref = 0.0;
    if (valid) {
        double p[9];
        for (int k = 0; k < 3; k++) {
            int nd = A.t3[3 * e + k];
            for (int i = 0; i < 3; i++) p[3 * k + i] = A.xyz[3 * nd + i];
        }
        if (recovery) {
            const double *pr = A.t3prop + 6 * e;
            const double rp[6] = {pr[0], pr[1], pr[2], 1.0, 5.0 / 6.0, pr[0] / (2.0 * (1.0 + pr[1]))};
            shell_props(rp, 100.0, P);
            Gref = rp[5];
        } else {
            shell_props(A.t3prop + 6 * e, A.k6rot, P);
            Gref = A.t3prop[6 * e + 5];
        }
        if (lane == 0) {
            t3_frame(p, S.R, S.lc);
            S.area = t3_derivs(S.lc, S.bv, S.cv);
            double pts[7][2];
            for (int k = 0; k < 3; k++) { pts[k][0] = S.lc[2 * k]; pts[k][1] = S.lc[2 * k + 1]; }
            for (int d = 0; d < 2; d++) {
                pts[3][d] = (S.lc[d] + S.lc[2 + d]) / 2;
                pts[4][d] = (S.lc[2 + d] + S.lc[4 + d]) / 2;
                pts[5][d] = (S.lc[4 + d] + S.lc[d]) / 2;
                pts[6][d] = (S.lc[d] + S.lc[2 + d] + S.lc[4 + d]) / 3;
            }
            for (int q = 0; q < 3; q++)
                for (int k = 0; k < 4; k++) {
                    S.q[q].lc[2 * k] = pts[T3_QUAD[q][k]][0];
                    S.q[q].lc[2 * k + 1] = pts[T3_QUAD[q][k]][1];
                }
        }
    }
    __syncthreads();
    bool nonzero = valid && S.area >= 1e-12;
    bool plate = nonzero && P.bend;
    bool has_shear = plate && P.h >= 1e-30 && P.gs >= 1e-30;
    if (plate && lane < 48) q4_stage_geometry(S.q[lane / 16], lane % 16);
    __syncthreads();
    if (has_shear && lane < 3) q4_macneal_z(S.q[lane], P);
    __syncthreads();
    if (has_shear)
        for (int t = lane; t < 3 * 48; t += S105_WAVE) {
            T3Quad &Q = S.q[t / 48];
            int r = t % 48, i = r / 12, c = r % 12;
            double s = 0.0;
            for (int l = 0; l < 4; l++) s += Q.Zi[i * 4 + l] * Q.D[l][c];
            Q.ZiD[i][c] = s;
        }
    __syncthreads();
    // K21 in quad order (the reference adds the three quads in sequence)
    if (plate)
        for (int t = lane; t < 441; t += S105_WAVE) {
            int i = t / 21, j = t % 21, ni = i / 3, nj = j / 3;
            double v = 0.0;
            for (int q = 0; q < 3; q++) {
                int pa = t3_pos(q, ni), pb = t3_pos(q, nj);
                if (pa >= 0 && pb >= 0) v += t3_quad_plate_entry(S.q[q], P, 3 * pa + i % 3, 3 * pb + j % 3, has_shear);
            }
            S.K21[t] = v;
        }
    __syncthreads();
    if (plate)
        for (int t = lane; t < 252; t += S105_WAVE) {
            int i = t / 21, j = t % 21;
            S.M[t] = j < 12 ? S.K21[(9 + i) * 21 + 9 + j] : S.K21[(9 + i) * 21 + (j - 12)];
        }
    __syncthreads();
    // Gaussian elimination with partial pivoting (first largest |pivot|)
    for (int k = 0; k < 12; k++) {
        int piv = k;
        if (plate) {
            double best = fabs(S.M[k * 21 + k]);
            for (int i = k + 1; i < 12; i++) {
                double v = fabs(S.M[i * 21 + k]);
                if (v > best) { best = v; piv = i; }
            }
        }
        __syncthreads();
        if (plate && piv != k && lane < 21) {
            double tmp = S.M[k * 21 + lane];
            S.M[k * 21 + lane] = S.M[piv * 21 + lane];
            S.M[piv * 21 + lane] = tmp;
        }
        __syncthreads();
        if (plate) {
            int nc = 20 - k, n = (11 - k) * nc;
            for (int t = lane; t < n; t += S105_WAVE) {
                int i = k + 1 + t / nc, j = k + 1 + t % nc;
                double l = S.M[i * 21 + k] / S.M[k * 21 + k];
                S.M[i * 21 + j] -= l * S.M[k * 21 + j];
            }
        }
        __syncthreads();
    }
    // back substitution, one right-hand side (corner DOF) per lane
    if (plate && lane < 9) {
        int c = 12 + lane;
        for (int k = 11; k >= 0; k--) {
            double s = S.M[k * 21 + c];
            for (int l = k + 1; l < 12; l++) s -= S.M[k * 21 + l] * S.M[l * 21 + c];
            S.M[k * 21 + c] = s / S.M[k * 21 + k];
        }
    }
    __syncthreads();
    if (plate)
        for (int t = lane; t < 81; t += S105_WAVE) {
            int i = t / 9, j = t % 9;
            double s = 0.0;
            for (int l = 0; l < 12; l++) s += S.K21[i * 21 + 9 + l] * S.M[l * 21 + 12 + j];
            S.Kc[t] = S.K21[i * 21 + j] - s;
        }
    __syncthreads();
    return T3Flags{valid, nonzero, plate, has_shear};
}

__global__ void __launch_bounds__(256) t3_ke_kernel(ElemArgs A, double *eb) {
    __shared__ T3Stage stage[EPB];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    T3Stage &S = stage[w];
    ShellProps P;
    double Gref;
    T3Flags F = t3_setup(S, P, Gref, A, e, lane);
    const bool valid = F.valid, nonzero = F.nonzero, plate = F.plate;
    // local 18 x 18 by 3 x 3 sub-blocks (lane < 36), rotated to the global frame
    if (valid && lane < 36) {
        int blk = lane / 4, sub = lane % 4;
        int I = blk / 3, J = blk % 3, si = sub / 2, sj = sub % 2;
        double alpha = ((P.bend ? A.k6rot : 0.0) / 1e5) * Gref * P.h;
        double loc[9], g[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                loc[i * 3 + j] = t3_local_entry(S, P, alpha, nonzero, plate, S.Kc, 1.0, 6 * I + 3 * si + i,
                                                6 * J + 3 * sj + j);
        rot_block(S.R, loc, g);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) S.out[(I * 3 + J) * 36 + (3 * si + i) * 6 + 3 * sj + j] = g[i * 3 + j];
    }
    __syncthreads();
    if (valid)
        for (int t = lane; t < 324; t += S105_WAVE) eb[(size_t)e * 324 + t] = S.out[t];
}

// Thickness derivative of the plate-block entry (a, b) of a virtual quad:
// bending scales with t^3, the MacNeal shear block is D' Z^-1 D with
// Z = Zs + Zb, Zs ~ 1/t and Zb ~ 1/t^3.
__device__ double t3_quad_plate_dentry(const T3Quad &Q, const ShellProps &P, int a, int b, bool has_shear) {
    const double t = P.h;
    int a24 = 6 * (a / 3) + 2 + a % 3, b24 = 6 * (b / 3) + 2 + b % 3;
    double kb = 0.0;
    for (int g = 0; g < 4; g++) {
        double ba[3], bb[3], w[3];
        bb_col(Q, g, a24, ba);
        bb_col(Q, g, b24, bb);
        mat3v(P.Cb, bb, w);
        kb += Q.gadj[g] * (ba[0] * w[0] + ba[1] * w[1] + ba[2] * w[2]);
    }
    double v = 3.0 * kb / t;
    if (has_shear) {
        double s = 0.0;
        for (int l = 0; l < 4; l++)
            for (int q = 0; q < 4; q++)
                s += Q.ZiD[l][a] * (Q.Zs[l * 4 + q] / t + 3.0 * Q.Zb[l * 4 + q] / t) * Q.ZiD[q][b];
        v += s;
    }
    return v;
}

// terms[3 (n_q4 + e)] = phi' dKe/dt phi, terms[3 (n_q4 + e) + 2] = psi' dKe/dt u
// of CTRIA3 e (global-frame vectors).  Membrane and drilling scale with t;
// the condensed plate block is differentiated as a Schur complement:
// dKc = dKee - dKei X - X' dKie + X' dKii X, X = Kii^-1 Kie.
__global__ void __launch_bounds__(256) t3_dke_kernel(ElemArgs A, const double *u, const double *phi,
                                                     const double *psi, double *terms) {
    __shared__ T3Stage stage[EPB];
    __shared__ double dK[EPB][441];
    __shared__ double Yd[EPB][108];
    __shared__ double dKc[EPB][81];
    __shared__ double vl[EPB][3][18];  // phi, psi, u in the element frame
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    T3Stage &S = stage[w];
    ShellProps P;
    double Gref;
    T3Flags F = t3_setup(S, P, Gref, A, e, lane);
    if (F.plate)
        for (int t = lane; t < 441; t += S105_WAVE) {
            int i = t / 21, j = t % 21, ni = i / 3, nj = j / 3;
            double v = 0.0;
            for (int q = 0; q < 3; q++) {
                int pa = t3_pos(q, ni), pb = t3_pos(q, nj);
                if (pa >= 0 && pb >= 0)
                    v += t3_quad_plate_dentry(S.q[q], P, 3 * pa + i % 3, 3 * pb + j % 3, F.has_shear);
            }
            dK[w][t] = v;
        }
    if (F.valid && lane < 54) {
        int vec = lane / 18, a = lane % 18, node = a / 6, blk = (a % 6) / 3, i = a % 3;
        const double *x = vec == 0 ? phi : (vec == 1 ? psi : u);
        const double *xg = x + 6 * (size_t)A.t3[3 * e + node] + 3 * blk;
        vl[w][vec][a] = S.R[i * 3] * xg[0] + S.R[i * 3 + 1] * xg[1] + S.R[i * 3 + 2] * xg[2];
    }
    __syncthreads();
    // X = M[:, 12:21]; Yd = dKii X
    if (F.plate)
        for (int t = lane; t < 108; t += S105_WAVE) {
            int i = t / 9, j = t % 9;
            double sum = 0.0;
            for (int l = 0; l < 12; l++) sum += dK[w][(9 + i) * 21 + 9 + l] * S.M[l * 21 + 12 + j];
            Yd[w][t] = sum;
        }
    __syncthreads();
    if (F.plate)
        for (int t = lane; t < 81; t += S105_WAVE) {
            int i = t / 9, j = t % 9;
            double v = dK[w][i * 21 + j];
            for (int l = 0; l < 12; l++) {
                double xli = S.M[l * 21 + 12 + i], xlj = S.M[l * 21 + 12 + j];
                v += -dK[w][i * 21 + 9 + l] * xlj - xli * dK[w][(9 + l) * 21 + j] + xli * Yd[w][l * 9 + j];
            }
            dKc[w][t] = v;
        }
    __syncthreads();
    double f1 = 0.0, f2 = 0.0;
    if (F.valid) {
        double alpha = ((P.bend ? A.k6rot : 0.0) / 1e5) * Gref * P.h;
        for (int t = lane; t < 324; t += S105_WAVE) {
            int a = t / 18, b = t % 18;
            double v = t3_local_entry(S, P, alpha, F.nonzero, F.plate, dKc[w], 1.0 / P.h, a, b);
            f1 += vl[w][0][a] * v * vl[w][0][b];
            f2 += vl[w][1][a] * v * vl[w][2][b];
        }
    }
    f1 = wave_sum(f1);
    f2 = wave_sum(f2);
    if (F.valid && lane == 0) {
        terms[3 * ((size_t)A.n_q4 + e)] = f1;
        terms[3 * ((size_t)A.n_q4 + e) + 2] = f2;
    }
}

// One thread per CTRIA3: terms[3 (n_q4 + e) + 1] = phi' KGe phi / t and the
// element's adjoint load eg[18 e ..] = lam d(phi' KGe(u) phi)/du (global
// frame, rotations zero).  With Phx = sum_i b_i phi_i, Phy = sum_i c_i phi_i
// (translations), phi' KGe phi = t A (sxx |Phx|^2 + syy |Phy|^2 + 2 sxy
// Phx.Phy) and the stress is linear in u: sigma = D Bm u.
__global__ void t3_kg_sens_kernel(ElemArgs A, const double *u, const double *phi, double lam, double *terms,
                                  double *eg) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.n_t3) return;
    double p[9], R[9], lc[6], bv[3], cv[3];
    for (int k = 0; k < 3; k++) {
        int nd = A.t3[3 * e + k];
        for (int i = 0; i < 3; i++) p[3 * k + i] = A.xyz[3 * nd + i];
    }
    t3_frame(p, R, lc);
    const double *pr = A.t3prop + 6 * e;
    double E = pr[0], nu = pr[1], h = pr[2];
    double ar = t3_derivs(lc, bv, cv);
    double *g = eg + 18 * (size_t)e;
    if (!(ar >= 1e-12 && h >= 1e-30)) {
        terms[3 * ((size_t)A.n_q4 + e) + 1] = 0.0;
        for (int i = 0; i < 18; i++) g[i] = 0.0;
        return;
    }
    double exx = 0.0, eyy = 0.0, gxy = 0.0, Px[3] = {0, 0, 0}, Py[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
        const size_t base = 6 * (size_t)A.t3[3 * e + k];
        const double *ug = u + base, *fg = phi + base;
        double ul = R[0] * ug[0] + R[1] * ug[1] + R[2] * ug[2];
        double vl = R[3] * ug[0] + R[4] * ug[1] + R[5] * ug[2];
        exx += bv[k] * ul;
        eyy += cv[k] * vl;
        gxy += cv[k] * ul + bv[k] * vl;
        for (int d = 0; d < 3; d++) { Px[d] += bv[k] * fg[d]; Py[d] += cv[k] * fg[d]; }
    }
    double c = E / (1 - nu * nu);
    double D[9] = {c, c * nu, 0.0, c * nu, c, 0.0, 0.0, 0.0, c * ((1 - nu) / 2)};
    double sig[3];
    for (int i = 0; i < 3; i++) sig[i] = ((D[3 * i] * exx + D[3 * i + 1] * eyy + D[3 * i + 2] * gxy) * h) / h;
    double wv[3] = {Px[0] * Px[0] + Px[1] * Px[1] + Px[2] * Px[2], Py[0] * Py[0] + Py[1] * Py[1] + Py[2] * Py[2],
                    2.0 * (Px[0] * Py[0] + Px[1] * Py[1] + Px[2] * Py[2])};
    terms[3 * ((size_t)A.n_q4 + e) + 1] = ar * (sig[0] * wv[0] + sig[1] * wv[1] + sig[2] * wv[2]);
    double Dw[3];
    for (int i = 0; i < 3; i++) Dw[i] = D[3 * i] * wv[0] + D[3 * i + 1] * wv[1] + D[3 * i + 2] * wv[2];
    const double f = lam * h * ar;
    for (int k = 0; k < 3; k++) {
        double gu = f * (bv[k] * Dw[0] + cv[k] * Dw[2]);
        double gv = f * (cv[k] * Dw[1] + bv[k] * Dw[2]);
        for (int d = 0; d < 3; d++) {
            g[6 * k + d] = gu * R[d] + gv * R[3 + d];
            g[6 * k + 3 + d] = 0.0;
        }
    }
}

// One thread per (element, node-block pair): CST stress and the 3 x 3
// translational block v I rotated to the global frame; rotations carry none.
__global__ void t3_kg_kernel(ElemArgs A, const double *u, double *eb) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)A.n_t3 * 9) return;
    int e = (int)(gid / 9), blk = (int)(gid % 9), I = blk / 3, J = blk % 3;
    double p[9], R[9], lc[6], bv[3], cv[3];
    for (int k = 0; k < 3; k++) {
        int nd = A.t3[3 * e + k];
        for (int i = 0; i < 3; i++) p[3 * k + i] = A.xyz[3 * nd + i];
    }
    t3_frame(p, R, lc);
    const double *pr = A.t3prop + 6 * e;
    double E = pr[0], nu = pr[1], h = pr[2];
    double ar = t3_derivs(lc, bv, cv);
    double v = 0.0;
    if (ar >= 1e-12 && h >= 1e-30) {
        double exx = 0.0, eyy = 0.0, gxy = 0.0;
        for (int k = 0; k < 3; k++) {
            const double *ug = u + 6 * (size_t)A.t3[3 * e + k];
            double ul = R[0] * ug[0] + R[1] * ug[1] + R[2] * ug[2];
            double vl = R[3] * ug[0] + R[4] * ug[1] + R[5] * ug[2];
            exx += bv[k] * ul;
            eyy += cv[k] * vl;
            gxy += cv[k] * ul + bv[k] * vl;
        }
        double c = E / (1 - nu * nu);
        double D[9] = {c, c * nu, 0.0, c * nu, c, 0.0, 0.0, 0.0, c * ((1 - nu) / 2)};
        double sig[3];
        for (int i = 0; i < 3; i++) sig[i] = ((D[3 * i] * exx + D[3 * i + 1] * eyy + D[3 * i + 2] * gxy) * h) / h;
        v = h * ar * (sig[0] * bv[I] * bv[J] + sig[1] * cv[I] * cv[J] + sig[2] * (bv[I] * cv[J] + cv[I] * bv[J]));
    }
    double loc[9] = {v, 0.0, 0.0, 0.0, v, 0.0, 0.0, 0.0, v}, g[9];
    rot_block(R, loc, g);
    double *dst = eb + (size_t)gid * 36;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) dst[i * 6 + j] = (i < 3 && j < 3) ? g[i * 3 + j] : 0.0;
}

// ---------------------------------------------------------------------------
// Deterministic block gather: K[blk][t] = sum of contributing element blocks
// in element order (no atomics: the summation order is fixed run to run).
// ---------------------------------------------------------------------------
__global__ void block_gather_kernel(int64_t n_blocks, const int32_t *cptr, const int32_t *cidx,
                                    const double *eb, double *vals) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_blocks * 36) return;
    int64_t blk = gid / 36;
    int t = (int)(gid % 36);
    double s = 0.0;
    for (int k = cptr[blk]; k < cptr[blk + 1]; k++) s += eb[(size_t)cidx[k] * 36 + t];
    vals[gid] = s;
}

// ---------------------------------------------------------------------------
// Thickness sensitivity of one buckling mode (per CQUAD4):
//   dlam/dt_e = -(phi' dKe phi + lam phi' dKGe|u phi - psi' dKe u) / (phi' KG phi)
// with dKe/dt = (Km + Kd)/t + 3 Kb/t + D' Z^-1 (Zs/t + 3 Zb/t) Z^-1 D
// (membrane and drilling terms scale with t, bending with t^3, the MacNeal
// shear compliances with 1/t and 1/t^3) and dKGe/dt|u = KGe/t.
// Restates the quantities of JFEM/src/solver/adjoint_buckling.jl:2406-2637
// for one design variable per element, differentiating the forward kernel.
// ---------------------------------------------------------------------------

// element DOF a of element e in the global DOF vector
__device__ __forceinline__ double q4_gvec(const ElemArgs &A, const double *x, int e, int a) {
    return x[6 * (size_t)A.q4[4 * e + a / 6] + a % 6];
}

__global__ void __launch_bounds__(256) q4_dke_kernel(ElemArgs A, const double *u, const double *phi,
                                                     const double *psi, double *terms /* n_q4 x 3 */) {
    __shared__ Q4Stage stage[EPB];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    Q4Stage &S = stage[w];
    ShellProps P;
    bool bend_c, has_shear;
    bool valid = q4_setup(S, P, A, e, lane, bend_c, has_shear);
    double f1 = 0.0, f2 = 0.0;
    if (valid) {
        double t = P.h;
        int bi = lane / 8, bj = lane % 8;
        double blk[9], g[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int a = 3 * bi + i, b = 3 * bj + j;
                double km, kb, kd, cm, cb;
                q4_gauss_entry(S, P, a, b, km, kb, kd);
                q4_condense_entry(S, a, b, bend_c, cm, cb);
                double v = (km - cm + kd) / t + 3.0 * (kb - cb) / t;
                if (has_shear) {
                    int pa = plate_idx(a), pb = plate_idx(b);
                    if (pa >= 0 && pb >= 0) {
                        // D' Zi (Zs/t + 3 Zb/t) Zi D, symmetrised like the forward block
                        double za[4], zb[4];
                        for (int l = 0; l < 4; l++) { za[l] = S.ZiD[l][pa]; zb[l] = S.ZiD[l][pb]; }
                        double s = 0.0;
                        for (int l = 0; l < 4; l++)
                            for (int q = 0; q < 4; q++)
                                s += za[l] * (S.Zs[l * 4 + q] / t + 3.0 * S.Zb[l * 4 + q] / t) * zb[q];
                        v += s;
                    }
                }
                blk[i * 3 + j] = v;
            }
        rot_block(S.R, blk, g);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double pa = q4_gvec(A, phi, e, 3 * bi + i), pb = q4_gvec(A, phi, e, 3 * bj + j);
                double sa = q4_gvec(A, psi, e, 3 * bi + i), ub = q4_gvec(A, u, e, 3 * bj + j);
                f1 += pa * g[i * 3 + j] * pb;
                f2 += sa * g[i * 3 + j] * ub;
            }
    }
    f1 = wave_sum(f1);
    f2 = wave_sum(f2);
    if (valid && lane == 0) { terms[3 * e] = f1; terms[3 * e + 2] = f2; }
}

// phi' KGe phi / t per element (the dKG/dt|u term), KGe with NM condensed
// membrane modes (four: the forward KG)
template <int NM>
__global__ void __launch_bounds__(256) q4_kgt_kernel(ElemArgs A, const double *u, const double *phi,
                                                     double *terms /* n_q4 x 3 */) {
    __shared__ Q4KGStageT<NM> stage[EPB];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    Q4KGStageT<NM> &S = stage[w];
    double h;
    bool valid = q4kg_setup(S, A, u, e, lane, h);
    if (valid) q4kg_vectors(S, lane, h);
    else __syncthreads();
    double f = 0.0;
    if (valid) {
        int bi = lane / 8, bj = lane % 8;
        double blk[9], g[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int a = 3 * bi + i, b = 3 * bj + j;
                double v = 0.0;
                for (int q = 0; q < 8; q++) v += S.wt[q] * (S.va[q][a] * S.va[q][b] + S.vb[q][a] * S.vb[q][b]);
                blk[i * 3 + j] = v;
            }
        rot_block(S.R, blk, g);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                bool drill = ((bi % 2) == 1 && i == 2) || ((bj % 2) == 1 && j == 2);
                if (drill) continue;
                f += q4_gvec(A, phi, e, 3 * bi + i) * g[i * 3 + j] * q4_gvec(A, phi, e, 3 * bj + j);
            }
    }
    f = wave_sum(f);
    if (valid && lane == 0) terms[3 * e + 1] = f / h;
}

// Forward-mode dual number with the three membrane-stress directions.
struct Dual3 {
    double v, d[3];
};
__device__ __forceinline__ Dual3 dconst(double v) { return Dual3{v, {0.0, 0.0, 0.0}}; }
__device__ __forceinline__ Dual3 operator+(const Dual3 &a, const Dual3 &b) {
    return Dual3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ __forceinline__ Dual3 operator-(const Dual3 &a, const Dual3 &b) {
    return Dual3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ __forceinline__ Dual3 operator*(const Dual3 &a, const Dual3 &b) {
    return Dual3{a.v * b.v, {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ __forceinline__ Dual3 operator*(double s, const Dual3 &a) {
    return Dual3{s * a.v, {s * a.d[0], s * a.d[1], s * a.d[2]}};
}
__device__ __forceinline__ Dual3 dsqrt(const Dual3 &a) {
    double r = sqrt(a.v);
    double f = r > 0.0 ? 0.5 / r : 0.0;
    return Dual3{r, {f * a.d[0], f * a.d[1], f * a.d[2]}};
}
__device__ __forceinline__ Dual3 datan2(const Dual3 &y, const Dual3 &x) {
    double den = x.v * x.v + y.v * y.v;
    double fx = den > 0.0 ? -y.v / den : 0.0, fy = den > 0.0 ? x.v / den : 0.0;
    return Dual3{atan2(y.v, x.v), {fx * x.d[0] + fy * y.d[0], fx * x.d[1] + fy * y.d[1], fx * x.d[2] + fy * y.d[2]}};
}
__device__ __forceinline__ Dual3 dcos(const Dual3 &a) {
    double s = -sin(a.v);
    return Dual3{cos(a.v), {s * a.d[0], s * a.d[1], s * a.d[2]}};
}
__device__ __forceinline__ Dual3 dsin(const Dual3 &a) {
    double c = cos(a.v);
    return Dual3{sin(a.v), {c * a.d[0], c * a.d[1], c * a.d[2]}};
}

// Adjoint right-hand side per element: d(phi' KGe(u) phi)/du_e (24, global
// frame), differentiated through the principal-stress decomposition.
__global__ void __launch_bounds__(256) q4_adjrhs_kernel(ElemArgs A, const double *u, const double *phi, double lam,
                                                        double *eg /* n_q4 x 24 */) {
    __shared__ Q4KGStage stage[EPB];
    __shared__ double dphi_s[EPB][4][3];  // dPhi_gp/dsigma
    __shared__ double dsdu[EPB][4][3][24];
    __shared__ double phil[EPB][24];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    Q4KGStage &S = stage[w];
    double h;
    bool valid = q4kg_setup(S, A, u, e, lane, h);
    // phi in the local frame with the global drilling component removed
    if (valid && lane < 8) {
        int b = lane;
        int nd = A.q4[4 * e + b / 2];
        double pg[3];
        for (int i = 0; i < 3; i++) pg[i] = phi[6 * (size_t)nd + 3 * (b % 2) + i];
        if (b % 2 == 1) pg[2] = 0.0;
        for (int i = 0; i < 3; i++) phil[w][3 * b + i] = S.R[i * 3] * pg[0] + S.R[i * 3 + 1] * pg[1] + S.R[i * 3 + 2] * pg[2];
    }
    __syncthreads();
    if (valid && lane < 4) {
        int g = lane;
        // gradients of phi at this GP (incl. condensed Wilson modes)
        double X1 = 0, X2 = 0, Y1 = 0, Y2 = 0, Z1 = 0, Z2 = 0;
        for (int a = 0; a < 24; a++) {
            double gr[6];
            q4kg_grads(S, g, a, gr);
            double pa = phil[w][a];
            X1 += gr[0] * pa; X2 += gr[1] * pa; Y1 += gr[2] * pa; Y2 += gr[3] * pa; Z1 += gr[4] * pa; Z2 += gr[5] * pa;
        }
        Dual3 sxx = Dual3{S.sig[g][0], {1.0, 0.0, 0.0}};
        Dual3 syy = Dual3{S.sig[g][1], {0.0, 1.0, 0.0}};
        Dual3 sxy = Dual3{S.sig[g][2], {0.0, 0.0, 1.0}};
        Dual3 mean = 0.5 * (sxx + syy), hd = 0.5 * (sxx - syy);
        Dual3 rad = dsqrt(hd * hd + sxy * sxy);
        Dual3 l[2], c[2], s[2];
        if (rad.v <= 1e-30) {
            l[0] = mean; l[1] = mean;
            c[0] = dconst(1.0); s[0] = dconst(0.0);
            c[1] = dconst(0.0); s[1] = dconst(1.0);
        } else {
            Dual3 th = 0.5 * datan2(2.0 * sxy, sxx - syy);
            Dual3 cc = dcos(th), ss = dsin(th);
            l[0] = mean + rad; c[0] = cc; s[0] = ss;
            l[1] = mean - rad; c[1] = dconst(0.0) - ss; s[1] = cc;
        }
        // The forward operator drops a principal stress with |l| <= 1e-30;
        // its value is then zero either way, but its derivative dl/dsigma * q
        // is not, so every direction is kept here (uniaxial states have a
        // minor principal stress at rounding level, often exactly 0).
        Dual3 Phi = dconst(0.0);
        for (int k = 0; k < 2; k++) {
            Dual3 gx = c[k] * dconst(X1) + s[k] * dconst(X2);
            Dual3 gy = c[k] * dconst(Y1) + s[k] * dconst(Y2);
            Dual3 gz = c[k] * dconst(Z1) + s[k] * dconst(Z2);
            Dual3 a = s[k] * gx - c[k] * gy;
            Phi = Phi + l[k] * (a * a + gz * gz);
        }
        double sc = h * S.gadj[g];
        for (int i = 0; i < 3; i++) dphi_s[w][g][i] = sc * Phi.d[i];
    }
    // dsigma_gp/du_local (3 x 24 per GP): Cm (Bm + Bi A) / h
    if (valid) {
        const double *prop = A.q4prop + 6 * e;
        double E = prop[0], nu = prop[1];
        double cm = E / (1 - nu * nu);
        const double D[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
        double Cm[9];
        for (int i = 0; i < 9; i++) Cm[i] = (cm * D[i]) * h;
        for (int t = lane; t < 96; t += S105_WAVE) {
            int g = t / 24, a = t % 24;
            int k = a / 6, cpt = a % 6;
            double dx = S.gdx[g][k], dy = S.gdy[g][k];
            double eps[3] = {(cpt == 0) ? dx : 0.0, (cpt == 1) ? dy : 0.0, (cpt == 0) ? dy : ((cpt == 1) ? dx : 0.0)};
            double p1x = S.gdp[g][0], p1y = S.gdp[g][1], p2x = S.gdp[g][2], p2y = S.gdp[g][3];
            double a1 = S.Am[0][a], a2 = S.Am[1][a], a3 = S.Am[2][a], a4 = S.Am[3][a];
            eps[0] += p1x * a1 + p2x * a3;
            eps[1] += p1y * a2 + p2y * a4;
            eps[2] += p1y * a1 + p1x * a2 + p2y * a3 + p2x * a4;
            double N[3];
            mat3v(Cm, eps, N);
            for (int i = 0; i < 3; i++) dsdu[w][g][i][a] = N[i] / h;
        }
    }
    __syncthreads();
    // d/du_local = sum_gp dPhi/dsigma . dsigma/du ; rotate to global: g_glob = R^T g_loc
    if (valid && lane < 8) {
        int b = lane;
        double gl[3];
        for (int i = 0; i < 3; i++) {
            int a = 3 * b + i;
            double s = 0.0;
            for (int g = 0; g < 4; g++)
                for (int q = 0; q < 3; q++) s += dphi_s[w][g][q] * dsdu[w][g][q][a];
            gl[i] = s;
        }
        for (int j = 0; j < 3; j++) {
            double v = S.R[j] * gl[0] + S.R[3 + j] * gl[1] + S.R[6 + j] * gl[2];
            eg[(size_t)e * 24 + 3 * b + j] = lam * v;
        }
    }
}

// ---------------------------------------------------------------------------
// The reference's own analytic thickness sensitivity (its default adjoint
// path for flat isotropic CQUAD4, adjoint_buckling.jl:2406-2637), selected by
// SOL105_OPT_SENS_MODE:
//  * dKe/dt of the legacy element stiffness_quad4_generic (dKdx.jl:219-247,
//    FEMKernels.jl:1473-1658), not of the forward MacNeal element: Wilson
//    membrane modes, standard bending B, MITC4 shear with
//    phi^2 = min(1, 10 t^2 / (4 |det J_c|)), drilling, with the generic
//    defaults (bending ratio 1, TS/T 5/6, K6ROT 100).  Membrane, drilling and
//    the condensation scale with t, bending with t^3, shear with phi^2(t) t,
//    so dKe/dt = (Km + Kd - Kc)/t + 3 Kb/t + d(phi^2 t)/dt Ks/t exactly;
//  * the adjoint load f_j = phi' KG(sigma(e_j)) phi over the 24 unit nodal
//    displacements (_flat_iso_quad4_rhs_local, adjoint_buckling.jl:1024-1042)
//    and phi' KG phi / t, with four Wilson modes when the model has bars and
//    six (plus the r s mode) otherwise (adjoint_buckling.jl:930-1005).
// ---------------------------------------------------------------------------

struct Q4LegStage {
    double lc[8];
    double R[9];
    double gdx[4][4], gdy[4][4], gN[4][4], gadj[4], gdp[4][4];
    double giJ[4][4];   // inverse Jacobian per GP: iJ11 iJ12 iJ21 iJ22
    double btp[4][12];  // MITC4 tying rows on (w, thx, thy) of the corners
    double Kab[24][4], Kbb[16], Kbbi[16];
    double dphi2t;      // d(phi^2 t)/dt
};

__global__ void __launch_bounds__(256) q4_dke_ref_kernel(ElemArgs A, const double *u, const double *phi,
                                                         const double *psi, double *terms /* n_q4 x 3 */) {
    __shared__ Q4LegStage stage[EPB];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    Q4LegStage &S = stage[w];
    const bool valid = e < A.n_q4;
    double E = 1.0, nu = 0.0, t = 1.0;
    if (valid) {
        const double *prop = A.q4prop + 6 * e;
        E = prop[0];
        nu = prop[1];
        t = prop[2];
        if (lane == 0) {
            double p[12];
            for (int k = 0; k < 4; k++) {
                int nd = A.q4[4 * e + k];
                for (int i = 0; i < 3; i++) p[3 * k + i] = A.xyz[3 * nd + i];
            }
            q4_frame(p, S.R, S.lc);
        }
    }
    __syncthreads();
    const double G = E / (2.0 * (1.0 + nu));
    const double cmf = E * t / (1 - nu * nu), cbf = E * t * t * t / (12 * (1 - nu * nu));
    const double D[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
    double Cm[9], Cb[9];
    for (int i = 0; i < 9; i++) { Cm[i] = cmf * D[i]; Cb[i] = cbf * D[i]; }
    const double tsg = (5.0 / 6.0) * G * t, alpha = (100.0 / 1e5) * G * t;
    if (valid) {
        if (lane < 4) {
            const double r = (lane == 1 || lane == 2) ? PT : -PT, s = (lane >= 2) ? PT : -PT;
            double dNr[4], dNs[4], J11, J12, J21, J22;
            shape_derivs(r, s, dNr, dNs);
            jacobian(S.lc, dNr, dNs, J11, J12, J21, J22);
            double detJ = J11 * J22 - J12 * J21;
            double adj = fabs(detJ);
            if (adj < 1e-12) adj = 1e-12;
            double id = 1.0 / detJ;
            double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
            for (int k = 0; k < 4; k++) {
                S.gdx[lane][k] = iJ11 * dNr[k] + iJ12 * dNs[k];
                S.gdy[lane][k] = iJ21 * dNr[k] + iJ22 * dNs[k];
                S.gN[lane][k] = 0.25 * (1 + ((k == 1 || k == 2) ? r : -r)) * (1 + (k >= 2 ? s : -s));
            }
            S.gadj[lane] = adj;
            S.gdp[lane][0] = iJ11 * (-2.0 * r);
            S.gdp[lane][1] = iJ21 * (-2.0 * r);
            S.gdp[lane][2] = iJ12 * (-2.0 * s);
            S.gdp[lane][3] = iJ22 * (-2.0 * s);
            S.giJ[lane][0] = iJ11; S.giJ[lane][1] = iJ12; S.giJ[lane][2] = iJ21; S.giJ[lane][3] = iJ22;
        } else if (lane < 8) {
            // tying points (0,-1), (0,1) (r-shear) and (-1,0), (1,0) (s-shear)
            const int tp = lane - 4;
            const double r = tp < 2 ? 0.0 : (tp == 2 ? -1.0 : 1.0), s = tp == 0 ? -1.0 : (tp == 1 ? 1.0 : 0.0);
            double dNr[4], dNs[4], J11, J12, J21, J22;
            shape_derivs(r, s, dNr, dNs);
            jacobian(S.lc, dNr, dNs, J11, J12, J21, J22);
            const double N[4] = {0.25 * (1 - r) * (1 - s), 0.25 * (1 + r) * (1 - s), 0.25 * (1 + r) * (1 + s),
                                 0.25 * (1 - r) * (1 + s)};
            for (int k = 0; k < 4; k++) {
                if (tp < 2) { S.btp[tp][3 * k] = dNr[k]; S.btp[tp][3 * k + 1] = -J12 * N[k]; S.btp[tp][3 * k + 2] = J11 * N[k]; }
                else { S.btp[tp][3 * k] = dNs[k]; S.btp[tp][3 * k + 1] = -J22 * N[k]; S.btp[tp][3 * k + 2] = J21 * N[k]; }
            }
        } else if (lane == 8) {
            const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
            double J11, J12, J21, J22;
            jacobian(S.lc, dNr_c, dNs_c, J11, J12, J21, J22);
            double trial = 10.0 * t * t / fmax(4.0 * fabs(J11 * J22 - J12 * J21), 1e-30);
            S.dphi2t = trial < 1.0 ? 3.0 * trial : 1.0;
        }
    }
    __syncthreads();
    if (valid)
        for (int q = lane; q < 112; q += S105_WAVE) {
            if (q < 96) {
                int a = q / 4, m = q % 4, k = a / 6, c = a % 6;
                double sacc = 0.0;
                for (int g = 0; g < 4; g++) {
                    double dx = S.gdx[g][k], dy = S.gdy[g][k];
                    double ba[3] = {(c == 0) ? dx : 0.0, (c == 1) ? dy : 0.0, (c == 0) ? dy : ((c == 1) ? dx : 0.0)};
                    const int j = 2 * (m / 2);
                    const double px = S.gdp[g][j], py = S.gdp[g][j + 1];
                    double bi[3] = {m % 2 == 0 ? px : 0.0, m % 2 == 0 ? 0.0 : py, m % 2 == 0 ? py : px}, cbi[3];
                    mat3v(Cm, bi, cbi);
                    sacc += (ba[0] * cbi[0] + ba[1] * cbi[1] + ba[2] * cbi[2]) * S.gadj[g];
                }
                S.Kab[a][m] = sacc;
            } else {
                int l = (q - 96) / 4, m = (q - 96) % 4;
                double sacc = 0.0;
                for (int g = 0; g < 4; g++) {
                    const int jl = 2 * (l / 2), jm = 2 * (m / 2);
                    double bl[3] = {l % 2 == 0 ? S.gdp[g][jl] : 0.0, l % 2 == 0 ? 0.0 : S.gdp[g][jl + 1],
                                    l % 2 == 0 ? S.gdp[g][jl + 1] : S.gdp[g][jl]};
                    double bm[3] = {m % 2 == 0 ? S.gdp[g][jm] : 0.0, m % 2 == 0 ? 0.0 : S.gdp[g][jm + 1],
                                    m % 2 == 0 ? S.gdp[g][jm + 1] : S.gdp[g][jm]}, cbm[3];
                    mat3v(Cm, bm, cbm);
                    sacc += (bl[0] * cbm[0] + bl[1] * cbm[1] + bl[2] * cbm[2]) * S.gadj[g];
                }
                S.Kbb[l * 4 + m] = sacc;
            }
        }
    __syncthreads();
    if (valid && lane == 0) inv4(S.Kbb, S.Kbbi);
    __syncthreads();
    double f1 = 0.0, f2 = 0.0;
    if (valid) {
        const int bi = lane / 8, bj = lane % 8;
        double blk[9], gbl[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const int a = 3 * bi + i, b = 3 * bj + j;
                const int ka = a / 6, ca = a % 6, kb = b / 6, cb = b % 6;
                double km = 0.0, kbend = 0.0, kd = 0.0, ks = 0.0;
                for (int g = 0; g < 4; g++) {
                    const double adj = S.gadj[g];
                    const double dxa = S.gdx[g][ka], dya = S.gdy[g][ka], dxb = S.gdx[g][kb], dyb = S.gdy[g][kb];
                    // membrane
                    double ma[3] = {ca == 0 ? dxa : 0.0, ca == 1 ? dya : 0.0, ca == 0 ? dya : (ca == 1 ? dxa : 0.0)};
                    double mb[3] = {cb == 0 ? dxb : 0.0, cb == 1 ? dyb : 0.0, cb == 0 ? dyb : (cb == 1 ? dxb : 0.0)};
                    double cmb[3];
                    mat3v(Cm, mb, cmb);
                    km += adj * (ma[0] * cmb[0] + ma[1] * cmb[1] + ma[2] * cmb[2]);
                    // standard bending: kx = d thy/dx, ky = -d thx/dy, kxy = d thy/dy - d thx/dx
                    double ba_[3] = {ca == 4 ? dxa : 0.0, ca == 3 ? -dya : 0.0, ca == 4 ? dya : (ca == 3 ? -dxa : 0.0)};
                    double bb_[3] = {cb == 4 ? dxb : 0.0, cb == 3 ? -dyb : 0.0, cb == 4 ? dyb : (cb == 3 ? -dxb : 0.0)};
                    double cbb[3];
                    mat3v(Cb, bb_, cbb);
                    kbend += adj * (ba_[0] * cbb[0] + ba_[1] * cbb[1] + ba_[2] * cbb[2]);
                    // drilling
                    double da = ca == 0 ? 0.5 * dya : (ca == 1 ? -0.5 * dxa : (ca == 5 ? S.gN[g][ka] : 0.0));
                    double db = cb == 0 ? 0.5 * dyb : (cb == 1 ? -0.5 * dxb : (cb == 5 ? S.gN[g][kb] : 0.0));
                    kd += adj * alpha * da * db;
                    // MITC4 shear: covariant strains interpolated from the tying points
                    if (ca >= 2 && ca <= 4 && cb >= 2 && cb <= 4) {
                        const double r = (g == 1 || g == 2) ? PT : -PT, s = (g >= 2) ? PT : -PT;
                        const int ia = 3 * ka + (ca - 2), ib = 3 * kb + (cb - 2);
                        const double gra = 0.5 * (1 - s) * S.btp[0][ia] + 0.5 * (1 + s) * S.btp[1][ia];
                        const double gsa = 0.5 * (1 - r) * S.btp[2][ia] + 0.5 * (1 + r) * S.btp[3][ia];
                        const double grb = 0.5 * (1 - s) * S.btp[0][ib] + 0.5 * (1 + s) * S.btp[1][ib];
                        const double gsb = 0.5 * (1 - r) * S.btp[2][ib] + 0.5 * (1 + r) * S.btp[3][ib];
                        const double i11 = S.giJ[g][0], i12 = S.giJ[g][1], i21 = S.giJ[g][2], i22 = S.giJ[g][3];
                        const double c11 = tsg * (i11 * i11 + i21 * i21), c12 = tsg * (i11 * i12 + i21 * i22),
                                     c22 = tsg * (i12 * i12 + i22 * i22);
                        ks += adj * (gra * (c11 * grb + c12 * gsb) + gsa * (c12 * grb + c22 * gsb));
                    }
                }
                double kc = 0.0;
                for (int l = 0; l < 4; l++)
                    for (int m = 0; m < 4; m++) kc += S.Kab[a][l] * S.Kbbi[l * 4 + m] * S.Kab[b][m];
                blk[i * 3 + j] = (km + kd - kc) / t + 3.0 * kbend / t + S.dphi2t * ks / t;
            }
        rot_block(S.R, blk, gbl);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double pa = q4_gvec(A, phi, e, 3 * bi + i), pb = q4_gvec(A, phi, e, 3 * bj + j);
                double sa = q4_gvec(A, psi, e, 3 * bi + i), ub = q4_gvec(A, u, e, 3 * bj + j);
                f1 += pa * gbl[i * 3 + j] * pb;
                f2 += sa * gbl[i * 3 + j] * ub;
            }
    }
    f1 = wave_sum(f1);
    f2 = wave_sum(f2);
    if (valid && lane == 0) { terms[3 * e] = f1; terms[3 * e + 2] = f2; }
}

// Unit-basis adjoint load per element: f_a = phi_l' KG(sigma(e_a)) phi_l for
// the local unit displacements e_a (zero for w and the rotations), rotated to
// the global frame and scaled by lambda.
template <int NM>
__global__ void __launch_bounds__(256) q4_adjrhs_ref_kernel(ElemArgs A, const double *u, const double *phi,
                                                            double lam, double *eg /* n_q4 x 24 */) {
    __shared__ Q4KGStageT<NM> stage[EPB];
    __shared__ double grd[EPB][4][6];        // gradients of phi per GP
    __shared__ double dsdu[EPB][4][3][24];   // sigma_gp(e_a)
    __shared__ double phil[EPB][24];
    __shared__ double fc[EPB][4][24];        // per-GP parts of f_a
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    Q4KGStageT<NM> &S = stage[w];
    double h;
    bool valid = q4kg_setup(S, A, u, e, lane, h);
    if (valid && lane < 8) {
        int b = lane;
        int nd = A.q4[4 * e + b / 2];
        double pg[3];
        for (int i = 0; i < 3; i++) pg[i] = phi[6 * (size_t)nd + 3 * (b % 2) + i];
        for (int i = 0; i < 3; i++) phil[w][3 * b + i] = S.R[i * 3] * pg[0] + S.R[i * 3 + 1] * pg[1] + S.R[i * 3 + 2] * pg[2];
    }
    __syncthreads();
    if (valid && lane < 4) {
        int g = lane;
        double X[6] = {0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 24; a++) {
            double gr[6];
            q4kg_grads(S, g, a, gr);
            for (int i = 0; i < 6; i++) X[i] += gr[i] * phil[w][a];
        }
        for (int i = 0; i < 6; i++) grd[w][g][i] = X[i];
    }
    if (valid) {
        const double *prop = A.q4prop + 6 * e;
        double E = prop[0], nu = prop[1];
        double cm = E / (1 - nu * nu);
        const double D[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
        double Cm[9];
        for (int i = 0; i < 9; i++) Cm[i] = (cm * D[i]) * h;
        for (int q = lane; q < 96; q += S105_WAVE) {
            int g = q / 24, a = q % 24;
            int k = a / 6, cpt = a % 6;
            double dx = S.gdx[g][k], dy = S.gdy[g][k];
            double eps[3] = {(cpt == 0) ? dx : 0.0, (cpt == 1) ? dy : 0.0, (cpt == 0) ? dy : ((cpt == 1) ? dx : 0.0)};
            for (int m = 0; m < NM; m++) {
                double bm[3];
                q4kg_mode_col(S, g, m, bm);
                const double am = S.Am[m][a];
                eps[0] += bm[0] * am;
                eps[1] += bm[1] * am;
                eps[2] += bm[2] * am;
            }
            double N[3];
            mat3v(Cm, eps, N);
            for (int i = 0; i < 3; i++) dsdu[w][g][i][a] = N[i] / h;
        }
    }
    __syncthreads();
    if (valid)
        for (int q = lane; q < 96; q += S105_WAVE) {
            int g = q / 24, a = q % 24;
            double pr[6];
            principal2d(dsdu[w][g][0][a], dsdu[w][g][1][a], dsdu[w][g][2][a], pr);
            const double *X = grd[w][g];
            double val = 0.0;
            for (int k = 0; k < 2; k++) {
                const double l = pr[3 * k], c = pr[3 * k + 1], sn = pr[3 * k + 2];
                if (fabs(l) <= 1e-30) continue;
                const double gx = c * X[0] + sn * X[1], gy = c * X[2] + sn * X[3], gz = c * X[4] + sn * X[5];
                const double av = sn * gx - c * gy;
                val += l * (av * av + gz * gz);
            }
            fc[w][g][a] = h * S.gadj[g] * val;
        }
    __syncthreads();
    if (valid && lane < 8) {
        int b = lane;
        double gl[3];
        for (int i = 0; i < 3; i++) {
            int a = 3 * b + i;
            gl[i] = fc[w][0][a] + fc[w][1][a] + fc[w][2][a] + fc[w][3][a];
        }
        for (int j = 0; j < 3; j++) {
            double v = S.R[j] * gl[0] + S.R[3 + j] * gl[1] + S.R[6 + j] * gl[2];
            eg[(size_t)e * 24 + 3 * b + j] = lam * v;
        }
    }
}

// bar adjoint RHS: d(phi' KGb(P(u)) phi)/du = (phi_l' KG_l(P=1) phi_l) dP/du
__global__ void bar_adjrhs_kernel(ElemArgs A, const double *phi, double lam, double *eg /* n_bar x 12 */) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.n_bar) return;
    BarLocal B;
    bar_setup(A, b, B);
    double pl[12];
    for (int blk = 0; blk < 4; blk++) {
        int nd = A.bar[2 * b + blk / 2];
        const double *pg = phi + 6 * (size_t)nd + 3 * (blk % 2);
        for (int i = 0; i < 3; i++) pl[3 * blk + i] = B.R[i * 3] * pg[0] + B.R[i * 3 + 1] * pg[1] + B.R[i * 3 + 2] * pg[2];
    }
    double q = 0.0;
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) q += pl[i] * bar_kg(B.L, 1.0, i, j) * pl[j];
    // dP/du_global = row 7 of k_local times T
    for (int blk = 0; blk < 4; blk++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int i = 0; i < 3; i++) s += bar_k(B, 6, 3 * blk + i) * B.R[i * 3 + j];
            eg[(size_t)b * 12 + 3 * blk + j] = lam * q * s;
        }
}

// terms[3e..3e+2] = phi' dKe/dt phi, phi' KGe phi / t, psi' dKe/dt u
__global__ void sens_combine_kernel(int n, const double *terms, double lam, double phiKGphi, double *dldt) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double num = terms[3 * e] + lam * terms[3 * e + 1] - terms[3 * e + 2];
    dldt[e] = -num / phiKGphi;
}

// ---------------------------------------------------------------------------
// SOL 101 result recovery (stress_recovery.jl:163-513).  Outputs are
// component-major: column c of element e at out[c * n + e], so the threads of
// a wavefront store each column contiguously.
//   CQUAD4 (44): N(3) M(3) Q(2) sigma_z1(3) sigma_z2(3) eps_z1(3) eps_z2(3)
//                corner N (4 x 3) corner M (4 x 3)
//   CTRIA3 (20): the first 20 columns of the CQUAD4 layout
//   CBAR   (18): axial shear_1 shear_2 torque moment_a1 moment_a2 moment_b1
//                moment_b2, axial stress, end A stresses at C D E F, end B
//                stresses at C D E F, 1 (0 for a bar shorter than 1e-9)
// ---------------------------------------------------------------------------

// MITC4 centre transverse-shear resultant of a quad with corner coordinates
// lc and plate DOFs (w, thx, thy) of corner k at wp[3k..3k+2]: covariant
// shears tied at the edge mid-points (0, -+1) and (-+1, 0), mapped through
// the centre Jacobian, phi2 = min(1, 10 h^2 / (4 |det Jc|)), ts/t = 5/6.
// FEMKernels.jl:5927-6001.
__device__ void mitc4_center_shear(const double *lc, const double *wp, double G, double h, double *Q) {
    const double tx[4] = {0.0, 0.0, -1.0, 1.0}, ty[4] = {-1.0, 1.0, 0.0, 0.0};
    double g[2] = {0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const double xi = tx[t], eta = ty[t];
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(xi, eta, dNr, dNs);
        jacobian(lc, dNr, dNs, J11, J12, J21, J22);
        const double Nk[4] = {0.25 * (1.0 - xi) * (1.0 - eta), 0.25 * (1.0 + xi) * (1.0 - eta),
                              0.25 * (1.0 + xi) * (1.0 + eta), 0.25 * (1.0 - xi) * (1.0 + eta)};
        const bool along_xi = t < 2;
        const double cx = along_xi ? -J12 : -J22, cy = along_xi ? J11 : J21;
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            v += (along_xi ? dNr[k] : dNs[k]) * wp[3 * k] + cx * Nk[k] * wp[3 * k + 1] + cy * Nk[k] * wp[3 * k + 2];
        g[along_xi ? 0 : 1] += 0.5 * v;
    }
    double dNr[4], dNs[4], J11, J12, J21, J22;
    shape_derivs(0.0, 0.0, dNr, dNs);
    jacobian(lc, dNr, dNs, J11, J12, J21, J22);
    const double det = J11 * J22 - J12 * J21;
    const double adet = fmax(fabs(det), 1e-30);
    const double id = 1.0 / det;
    const double i11 = J22 * id, i12 = -J12 * id, i21 = -J21 * id, i22 = J11 * id;
    const double phi2 = fmin(1.0, 10.0 * h * h / fmax(4.0 * adet, 1e-30));
    const double k = (5.0 / 6.0) * G * h;
    const double t11 = k * i11, t12 = k * i12, t21 = k * i21, t22 = k * i22;
    const double q0 = phi2 * ((i11 * t11 + i21 * t21) * g[0] + (i11 * t12 + i21 * t22) * g[1]);
    const double q1 = phi2 * ((i12 * t11 + i22 * t21) * g[0] + (i12 * t12 + i22 * t22) * g[1]);
    Q[0] = J11 * q0 + J21 * q1;
    Q[1] = J12 * q0 + J22 * q1;
}

// One thread per CQUAD4: frame and local displacements, the Wilson membrane
// modes condensed at the |det J|-weighted 2x2 Gauss points, Gauss-averaged
// membrane strain and curvature (N = Cm eps, M = -br D kappa h^3/12), MITC4
// shear, fibre stresses/strains at -+h/2, corner N and M extrapolated from
// the Gauss points, the equilibrium shear (dMx/dx, dMy/dy) replacing each
// shear component it undercuts, and the twisting moment held at its centroid
// value over the corners.  FEMKernels.jl:6003-6338; stress_recovery.jl:89-133,
// 242-311.
__global__ void __launch_bounds__(128) q4_recover_kernel(ElemArgs A, const double *u, double *out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.n_q4) return;
    const int n = A.n_q4;
    double p[12];
    int nd[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        nd[k] = A.q4[4 * e + k];
#pragma unroll
        for (int i = 0; i < 3; i++) p[3 * k + i] = A.xyz[3 * nd[k] + i];
    }
    double R[9], lc[8];
    q4_frame(p, R, lc);
    double ul[24];
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const double *ug = u + 6 * (size_t)nd[b / 2] + 3 * (b % 2);
        const double x = ug[0], y = ug[1], z = ug[2];
#pragma unroll
        for (int i = 0; i < 3; i++) ul[3 * b + i] = R[3 * i] * x + R[3 * i + 1] * y + R[3 * i + 2] * z;
    }
    const double *prop = A.q4prop + 6 * e;
    const double E = prop[0], nu = prop[1], h = prop[2], br = prop[3];
    const double dm = E / (1.0 - nu * nu);
    const double D[9] = {dm, dm * nu, 0.0, dm * nu, dm, 0.0, 0.0, 0.0, dm * ((1.0 - nu) / 2.0)};
    double Cm[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Cm[i] = D[i] * h;
    // Gauss points (-,-), (+,-), (+,+), (-,+): compatible strain, curvature,
    // Wilson-mode derivatives and the condensation system
    double adj[4], ec[4][3], kp[4][3], dp[4][4];
    double Kbb[16], rhs[4];
#pragma unroll
    for (int i = 0; i < 16; i++) Kbb[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) rhs[i] = 0.0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const double r = (g == 1 || g == 2) ? PT : -PT, s = (g >= 2) ? PT : -PT;
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jacobian(lc, dNr, dNs, J11, J12, J21, J22);
        const double det = J11 * J22 - J12 * J21;
        adj[g] = fmax(fabs(det), 1e-12);
        const double id = 1.0 / det;
        const double i11 = J22 * id, i12 = -J12 * id, i21 = -J21 * id, i22 = J11 * id;
        double e0 = 0.0, e1 = 0.0, e2 = 0.0, k0 = 0.0, k1 = 0.0, k2 = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double dx = i11 * dNr[k] + i12 * dNs[k], dy = i21 * dNr[k] + i22 * dNs[k];
            const double ux = ul[6 * k], uy = ul[6 * k + 1], tx = ul[6 * k + 3], ty = ul[6 * k + 4];
            e0 += dx * ux;
            e1 += dy * uy;
            e2 += dy * ux + dx * uy;
            k0 += dx * ty;
            k1 += -dy * tx;
            k2 += dy * ty - dx * tx;
        }
        ec[g][0] = e0; ec[g][1] = e1; ec[g][2] = e2;
        kp[g][0] = k0; kp[g][1] = k1; kp[g][2] = k2;
        dp[g][0] = i11 * (-2.0 * r);  // d(1 - xi^2)/dx
        dp[g][1] = i21 * (-2.0 * r);  // d(1 - xi^2)/dy
        dp[g][2] = i12 * (-2.0 * s);  // d(1 - eta^2)/dx
        dp[g][3] = i22 * (-2.0 * s);  // d(1 - eta^2)/dy
        const double cols[4][3] = {{dp[g][0], 0.0, dp[g][1]}, {0.0, dp[g][1], dp[g][0]},
                                   {dp[g][2], 0.0, dp[g][3]}, {0.0, dp[g][3], dp[g][2]}};
        double t[3];
        mat3v(Cm, ec[g], t);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            rhs[m] += adj[g] * (cols[m][0] * t[0] + cols[m][1] * t[1] + cols[m][2] * t[2]);
            double cm[3];
            mat3v(Cm, cols[m], cm);
#pragma unroll
            for (int l = 0; l < 4; l++)
                Kbb[l * 4 + m] += adj[g] * (cols[l][0] * cm[0] + cols[l][1] * cm[1] + cols[l][2] * cm[2]);
        }
    }
    double Kbi[16], al[4];
    inv4(Kbb, Kbi);
#pragma unroll
    for (int l = 0; l < 4; l++)
        al[l] = -(Kbi[4 * l] * rhs[0] + Kbi[4 * l + 1] * rhs[1] + Kbi[4 * l + 2] * rhs[2] + Kbi[4 * l + 3] * rhs[3]);
    double Cb[9];
    const double cbs = h * h * h / 12.0;
#pragma unroll
    for (int i = 0; i < 9; i++) Cb[i] = (br * D[i]) * cbs;
    double eps[3] = {0.0, 0.0, 0.0}, kap[3] = {0.0, 0.0, 0.0}, area = 0.0;
    double Ng[4][3], Mg[4][3];
#pragma unroll
    for (int g = 0; g < 4; g++) {
        double eg[3] = {ec[g][0] + dp[g][0] * al[0] + dp[g][2] * al[2],
                        ec[g][1] + dp[g][1] * al[1] + dp[g][3] * al[3],
                        ec[g][2] + dp[g][1] * al[0] + dp[g][0] * al[1] + dp[g][3] * al[2] + dp[g][2] * al[3]};
#pragma unroll
        for (int i = 0; i < 3; i++) {
            eps[i] += eg[i] * adj[g];
            kap[i] += kp[g][i] * adj[g];
        }
        area += adj[g];
        double mk[3];
        mat3v(Cm, eg, Ng[g]);
        mat3v(Cb, kp[g], mk);
#pragma unroll
        for (int i = 0; i < 3; i++) Mg[g][i] = -mk[i];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) { eps[i] /= area; kap[i] /= area; }
    double N[3], Dk[3], M[3];
    mat3v(Cm, eps, N);
    mat3v(D, kap, Dk);
#pragma unroll
    for (int i = 0; i < 3; i++) M[i] = (-br * Dk[i]) * cbs;
    double Q[2] = {0.0, 0.0};
    if (br > 1e-12) {
        double wp[12];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 3; i++) wp[3 * k + i] = ul[6 * k + 2 + i];
        mitc4_center_shear(lc, wp, E / (2.0 * (1.0 + nu)), h, Q);
    }
    // corner values: 2x2 Gauss extrapolation
    double Nc[4][3], Mc[4][3];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const double xi = (c == 1 || c == 2) ? 1.0 : -1.0, eta = (c >= 2) ? 1.0 : -1.0;
        const double lx1 = (PT - xi) / (2.0 * PT), lx2 = (xi + PT) / (2.0 * PT);
        const double ly1 = (PT - eta) / (2.0 * PT), ly2 = (eta + PT) / (2.0 * PT);
        const double w[4] = {lx1 * ly1, lx2 * ly1, lx2 * ly2, lx1 * ly2};
#pragma unroll
        for (int i = 0; i < 3; i++) {
            Nc[c][i] = w[0] * Ng[0][i] + w[1] * Ng[1][i] + w[2] * Ng[2][i] + w[3] * Ng[3][i];
            Mc[c][i] = w[0] * Mg[0][i] + w[1] * Mg[1][i] + w[2] * Mg[2][i] + w[3] * Mg[3][i];
        }
    }
    // equilibrium shear from the corner moments
    if (fabs(br) > 1e-12) {
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(0.0, 0.0, dNr, dNs);
        jacobian(lc, dNr, dNs, J11, J12, J21, J22);
        const double det = J11 * J22 - J12 * J21;
        if (fabs(det) >= 1e-12) {
            double rx = 0.0, sx = 0.0, ry = 0.0, sy = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                rx += dNr[k] * Mc[k][0]; sx += dNs[k] * Mc[k][0];
                ry += dNr[k] * Mc[k][1]; sy += dNs[k] * Mc[k][1];
            }
            const double eq[2] = {(J22 / det) * rx + (-J12 / det) * sx, (-J21 / det) * ry + (J11 / det) * sy};
#pragma unroll
            for (int i = 0; i < 2; i++)
                if (isfinite(eq[i]) && fabs(eq[i]) < fabs(Q[i])) Q[i] = eq[i];
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) Mc[c][2] = M[2];
    double row[20];
    const double z1 = -h / 2.0, z2 = h / 2.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        row[i] = N[i];
        row[3 + i] = M[i];
        row[14 + i] = eps[i] + z1 * kap[i];
        row[17 + i] = eps[i] + z2 * kap[i];
    }
    row[6] = Q[0];
    row[7] = Q[1];
    mat3v(D, row + 14, row + 8);
    mat3v(D, row + 17, row + 11);
#pragma unroll
    for (int c = 0; c < 20; c++) out[(size_t)c * n + e] = row[c];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            out[(size_t)(20 + 3 * c + i) * n + e] = Nc[c][i];
            out[(size_t)(32 + 3 * c + i) * n + e] = Mc[c][i];
        }
}

// One 64-lane wavefront per CTRIA3 (the macro plate set-up of the stiffness
// kernel with the recovery constitutive set): constant-strain membrane and
// curvature, the macro shear (interior DOFs -Kii^-1 Kie u, the three
// virtual quads' MITC4 shears averaged by area) and, with weight w > 0, the
// macro moment blend M = (1 - w) M + w M_macro with the fibre strains
// rebuilt from the blended moment.  FEMKernels.jl:7213-7311, 7391-7432;
// stress_recovery.jl:52-87, 356-366.
__global__ void __launch_bounds__(256) t3_recover_kernel(ElemArgs A, const double *u, const double *wt,
                                                         double *out) {
    __shared__ T3Stage stage[EPB];
    int w = threadIdx.x / S105_WAVE, lane = threadIdx.x % S105_WAVE;
    int e = blockIdx.x * EPB + w;
    T3Stage &S = stage[w];
    ShellProps P;
    double Gref;
    T3Flags F = t3_setup(S, P, Gref, A, e, lane, true);
    // S.out is free here: local DOFs at [0, 18), the 7-point plate field at
    // [24, 45), per virtual quad (Qx, Qy, Mx, My, Mxy, area) at [48, 66)
    double *ul = S.out, *ua = S.out + 24, *qr = S.out + 48;
    const double br = F.valid ? A.t3prop[6 * e + 3] : 0.0;
    const double E = F.valid ? A.t3prop[6 * e] : 0.0, nu = F.valid ? A.t3prop[6 * e + 1] : 0.0;
    const double h = F.valid ? A.t3prop[6 * e + 2] : 1.0;
    if (F.valid && lane < 6) {
        const double *ug = u + 6 * (size_t)A.t3[3 * e + lane / 2] + 3 * (lane % 2);
        for (int i = 0; i < 3; i++) ul[3 * lane + i] = S.R[i * 3] * ug[0] + S.R[i * 3 + 1] * ug[1] + S.R[i * 3 + 2] * ug[2];
    }
    __syncthreads();
    if (F.plate && lane < 21) {
        if (lane < 9) {
            ua[lane] = ul[6 * (lane / 3) + 2 + lane % 3];
        } else {
            const int i = lane - 9;
            double s = 0.0;
            for (int j = 0; j < 9; j++) s += S.M[i * 21 + 12 + j] * ul[6 * (j / 3) + 2 + j % 3];
            ua[lane] = -s;
        }
    }
    __syncthreads();
    if (F.plate && lane < 3) {
        const T3Quad &Q = S.q[lane];
        double wp[12];
        for (int a = 0; a < 4; a++)
            for (int i = 0; i < 3; i++) wp[3 * a + i] = ua[3 * T3_QUAD[lane][a] + i];
        double kap[3] = {0.0, 0.0, 0.0}, area = 0.0;
        for (int g = 0; g < 4; g++) {
            for (int k = 0; k < 4; k++) {
                const double dx = Q.gdx[g][k], dy = Q.gdy[g][k], tx = wp[3 * k + 1], ty = wp[3 * k + 2];
                kap[0] += Q.gadj[g] * (dx * ty);
                kap[1] += Q.gadj[g] * (-dy * tx);
                kap[2] += Q.gadj[g] * (dy * ty - dx * tx);
            }
            area += Q.gadj[g];
        }
        const double dm = E / (1.0 - nu * nu);
        const double k0 = kap[0] / area, k1 = kap[1] / area, k2 = kap[2] / area;
        const double cbs = h * h * h / 12.0;
        double Qs[2];
        mitc4_center_shear(Q.lc, wp, E / (2.0 * (1.0 + nu)), h, Qs);
        const double *qc = Q.lc;
        const double a1 = (qc[2] - qc[0]) * (qc[7] - qc[1]) - (qc[6] - qc[0]) * (qc[3] - qc[1]);
        const double a2 = (qc[4] - qc[2]) * (qc[7] - qc[3]) - (qc[6] - qc[2]) * (qc[5] - qc[3]);
        qr[6 * lane] = Qs[0];
        qr[6 * lane + 1] = Qs[1];
        qr[6 * lane + 2] = (-br * (dm * k0 + dm * nu * k1)) * cbs;
        qr[6 * lane + 3] = (-br * (dm * nu * k0 + dm * k1)) * cbs;
        qr[6 * lane + 4] = (-br * (dm * ((1.0 - nu) / 2.0) * k2)) * cbs;
        qr[6 * lane + 5] = (fabs(a1) + fabs(a2)) / 2.0;
    }
    __syncthreads();
    if (F.valid && lane == 0) {
        const int n = A.n_t3;
        double row[20];
        for (int c = 0; c < 20; c++) row[c] = 0.0;
        if (F.nonzero) {
            const double dm = E / (1.0 - nu * nu);
            const double D[9] = {dm, dm * nu, 0.0, dm * nu, dm, 0.0, 0.0, 0.0, dm * ((1.0 - nu) / 2.0)};
            double eps[3] = {0.0, 0.0, 0.0}, kap[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < 3; i++) {
                const double b = S.bv[i], c = S.cv[i];
                const double ux = ul[6 * i], uy = ul[6 * i + 1], tx = ul[6 * i + 3], ty = ul[6 * i + 4];
                eps[0] += b * ux; eps[1] += c * uy; eps[2] += c * ux + b * uy;
                kap[0] += b * ty; kap[1] += -c * tx; kap[2] += c * ty - b * tx;
            }
            const double cbs = h * h * h / 12.0;
            double De[3], Dk[3], M[3];
            mat3v(D, eps, De);
            mat3v(D, kap, Dk);
            for (int i = 0; i < 3; i++) {
                row[i] = De[i] * h;
                M[i] = (-br * Dk[i]) * cbs;
            }
            double Qm[2] = {0.0, 0.0}, Mm[3] = {0.0, 0.0, 0.0};
            if (br > 1e-12) {
                const double As = qr[5] + qr[11] + qr[17];
                if (As > 1e-12) {
                    for (int i = 0; i < 2; i++) Qm[i] = (qr[5] * qr[i] + qr[11] * qr[6 + i] + qr[17] * qr[12 + i]) / As;
                    for (int i = 0; i < 3; i++)
                        Mm[i] = (qr[5] * qr[2 + i] + qr[11] * qr[8 + i] + qr[17] * qr[14 + i]) / As;
                }
            }
            double e1[3], e2[3];
            for (int i = 0; i < 3; i++) {
                e1[i] = eps[i] + (-h / 2.0) * kap[i];
                e2[i] = eps[i] + (h / 2.0) * kap[i];
            }
            const double wgt = wt[e];
            if (wgt > 0.0 && fabs(br) > 1e-12 && isfinite(Mm[0]) && isfinite(Mm[1]) && isfinite(Mm[2])) {
                for (int i = 0; i < 3; i++) M[i] = (1.0 - wgt) * M[i] + wgt * Mm[i];
                // kappa = -(12 / (br h^3)) D^-1 M, D^-1 of the plane-stress matrix
                const double f = -(12.0 / (br * h * h * h)), dinv = 1.0 / dm;
                const double den = 1.0 - nu * nu;
                const double kb[3] = {f * dinv * (M[0] - nu * M[1]) / den, f * dinv * (M[1] - nu * M[0]) / den,
                                      f * dinv * M[2] / ((1.0 - nu) / 2.0)};
                for (int i = 0; i < 3; i++) {
                    const double em = (e1[i] + e2[i]) / 2.0;
                    e1[i] = em + (-h / 2.0) * kb[i];
                    e2[i] = em + (h / 2.0) * kb[i];
                }
            }
            for (int i = 0; i < 3; i++) {
                row[3 + i] = M[i];
                row[14 + i] = e1[i];
                row[17 + i] = e2[i];
            }
            row[6] = Qm[0];
            row[7] = Qm[1];
            mat3v(D, e1, row + 8);
            mat3v(D, e2, row + 11);
        }
        for (int c = 0; c < 20; c++) out[(size_t)c * n + e] = row[c];
    }
}

// One thread per CBAR: end forces k u in the bar frame and the stresses at the
// recovery points C, D, E, F (y, z pairs in pts, n_bar x 8) with the BAR
// section's sign -1 and I12 = 0.  stress_recovery.jl:432-513,
// FEMKernels.jl:215-229.
__global__ void __launch_bounds__(128) bar_recover_kernel(ElemArgs A, const double *u, const double *pts,
                                                          double *out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= A.n_bar) return;
    const int n = A.n_bar;
    BarLocal B;
    bar_setup(A, b, B);
    double ul[12];
    for (int blk = 0; blk < 4; blk++) {
        const double *ug = u + 6 * (size_t)A.bar[2 * b + blk / 2] + 3 * (blk % 2);
        for (int i = 0; i < 3; i++) ul[3 * blk + i] = B.R[i * 3] * ug[0] + B.R[i * 3 + 1] * ug[1] + B.R[i * 3 + 2] * ug[2];
    }
    // f = k u with the nonzero pattern of the local Timoshenko matrix (bar_k)
    const double *q = ul;
    double f[12];
    f[0] = B.X * (q[0] - q[6]);
    f[6] = B.X * (q[6] - q[0]);
    f[3] = B.T * (q[3] - q[9]);
    f[9] = B.T * (q[9] - q[3]);
    f[2] = B.ay * (q[2] - q[8]) - B.by * (q[4] + q[10]);
    f[8] = B.ay * (q[8] - q[2]) + B.by * (q[4] + q[10]);
    f[4] = B.by * (q[8] - q[2]) + B.cy * q[4] + B.dy * q[10];
    f[10] = B.by * (q[8] - q[2]) + B.dy * q[4] + B.cy * q[10];
    f[1] = B.az * (q[1] - q[7]) + B.bz * (q[5] + q[11]);
    f[7] = B.az * (q[7] - q[1]) - B.bz * (q[5] + q[11]);
    f[5] = B.bz * (q[1] - q[7]) + B.cz * q[5] + B.dz * q[11];
    f[11] = B.bz * (q[1] - q[7]) + B.dz * q[5] + B.cz * q[11];
    const double *bp = A.barprop + 8 * b;
    const double Ar = bp[0];
    double Iy = bp[2], Iz = bp[1];
    if (Iy == 0.0) Iy = Iz;
    if (Iz == 0.0) Iz = Iy;
    double row[18];
    row[0] = f[6]; row[1] = -f[1]; row[2] = -f[2]; row[3] = -f[3];
    row[4] = -f[5]; row[5] = f[4]; row[6] = f[11]; row[7] = -f[10];
    row[8] = fabs(Ar) > 1e-30 ? f[6] / Ar : 0.0;
    const double den = Iz * Iy;
    for (int j = 0; j < 4; j++) {
        const double y = pts[8 * b + 2 * j], z = pts[8 * b + 2 * j + 1];
        for (int end = 0; end < 2; end++) {
            const double m1 = row[4 + 2 * end], m2 = row[5 + 2 * end];
            const double s = fabs(den) > 1e-30 ? (m1 * Iy * y + m2 * Iz * z) / den
                                              : m1 * y / fmax(Iz, 1e-30) + m2 * z / fmax(Iy, 1e-30);
            row[9 + 4 * end + j] = -s;
        }
    }
    const bool ok = B.L >= 1e-9;
    for (int c = 0; c < 17; c++) out[(size_t)c * n + b] = ok ? row[c] : 0.0;
    out[(size_t)17 * n + b] = ok ? 1.0 : 0.0;
}

// r = K u - F over the 6x6 block rows: 6 lanes per block row (one per row of
// the block), 32 block rows per 192-thread workgroup.
__global__ void __launch_bounds__(192) bsr_residual_kernel(int n_nodes, const int32_t *rp, const int32_t *col,
                                                           const double *val, const double *x, const double *F,
                                                           double *r) {
    const int I = blockIdx.x * 32 + threadIdx.x / 6, ci = threadIdx.x % 6;
    if (I >= n_nodes) return;
    double acc = 0.0;
    for (int k = rp[I]; k < rp[I + 1]; k++) {
        const double *v = val + (int64_t)k * 36 + ci * 6;
        const double *xb = x + (int64_t)col[k] * 6;
#pragma unroll
        for (int cj = 0; cj < 6; cj++) acc += v[cj] * xb[cj];
    }
    r[(int64_t)I * 6 + ci] = acc - F[(int64_t)I * 6 + ci];
}

}  // namespace

void launch_q4_ke(const ElemArgs &a, double *eb, hipStream_t s) {
    if (a.n_q4 == 0) return;
    q4_ke_kernel<<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, eb);
}
void launch_q4_kg(const ElemArgs &a, const double *u, double *eb, hipStream_t s) {
    if (a.n_q4 == 0) return;
    q4_kg_kernel<<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, eb);
}
void launch_t3_ke(const ElemArgs &a, double *eb, hipStream_t s) {
    if (a.n_t3 == 0) return;
    t3_ke_kernel<<<(a.n_t3 + EPB - 1) / EPB, 256, 0, s>>>(a, eb + (size_t)a.n_q4 * 576);
}
void launch_t3_kg(const ElemArgs &a, const double *u, double *eb, hipStream_t s) {
    if (a.n_t3 == 0) return;
    int64_t n = (int64_t)a.n_t3 * 9;
    t3_kg_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, u, eb + (size_t)a.n_q4 * 576);
}
void launch_t3_sens_terms(const ElemArgs &a, const double *u, const double *phi, const double *psi, double *terms,
                          hipStream_t s) {
    if (a.n_t3 == 0) return;
    t3_dke_kernel<<<(a.n_t3 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, psi, terms);
}
void launch_t3_kg_sens(const ElemArgs &a, const double *u, const double *phi, double lam, double *terms, double *eg,
                       hipStream_t s) {
    if (a.n_t3 == 0) return;
    t3_kg_sens_kernel<<<(a.n_t3 + 127) / 128, 128, 0, s>>>(a, u, phi, lam, terms, eg);
}
void launch_bar_ke(const ElemArgs &a, double *eb, hipStream_t s) {
    if (a.n_bar == 0) return;
    bar_ke_kernel<<<(a.n_bar + 127) / 128, 128, 0, s>>>(a, eb + (size_t)a.n_q4 * 576 + (size_t)a.n_t3 * 324);
}
void launch_bar_kg(const ElemArgs &a, const double *u, double *eb, hipStream_t s) {
    if (a.n_bar == 0) return;
    bar_kg_kernel<<<(a.n_bar + 127) / 128, 128, 0, s>>>(a, u, eb + (size_t)a.n_q4 * 576 + (size_t)a.n_t3 * 324);
}
// ---------------------------------------------------------------------------
// Constraint elimination (RBE2 / RBE3): entries of dependent rows and columns
// move to the independent DOFs in a fixed order per target (targets are
// entries of independent rows and columns, never sources, so the update is
// in place without races), then the dependent rows and columns are cleared;
// dependent DOFs of a solution are recovered from the independent ones.
// ---------------------------------------------------------------------------
__global__ void mpc_apply_kernel(int64_t nt, const int64_t *tgt, const int32_t *tptr, const int64_t *src,
                                 const double *coef, double *K) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double acc = K[tgt[t]];
    for (int32_t p = tptr[t]; p < tptr[t + 1]; p++) acc += coef[p] * K[src[p]];
    K[tgt[t]] = acc;
}

__global__ void mpc_zero_kernel(int64_t nz, const int64_t *pos, double *K) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nz) K[pos[t]] = 0.0;
}

// x[dep[i] + c ld] = sum_p coef[p] x[ind[p] + c ld] for columns c < ncols
__global__ void mpc_recover_kernel(int ndep, const int32_t *dep, const int32_t *ptr, const int32_t *ind,
                                   const double *coef, double *x, int64_t ld, int ncols) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ndep * ncols) return;
    const int i = (int)(t % ndep), c = (int)(t / ndep);
    double acc = 0.0;
    for (int32_t p = ptr[i]; p < ptr[i + 1]; p++) acc += coef[p] * x[ind[p] + c * ld];
    x[dep[i] + c * ld] = acc;
}

// v[recv[k]] += sum_p coef[p] v[dep[p]] on free receivers (fixed ones get
// zero); dependent DOFs are never receivers, so reads and writes do not meet
__global__ void mpc_fold_kernel(int nrecv, const int32_t *recv, const int32_t *ptr, const int32_t *dep,
                                const double *coef, const uint8_t *free_mask, double *v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrecv) return;
    const int a = recv[k];
    double acc = v[a];
    for (int32_t p = ptr[k]; p < ptr[k + 1]; p++) acc += coef[p] * v[dep[p]];
    v[a] = free_mask[a] ? acc : 0.0;
}

__global__ void vec_free_mask_kernel(int n, const uint8_t *free_mask, double *v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && !free_mask[t]) v[t] = 0.0;
}

void launch_mpc_fold(int nrecv, const int32_t *recv, const int32_t *ptr, const int32_t *dep, const double *coef,
                     const uint8_t *free_mask, int ndof, double *v, hipStream_t s) {
    if (nrecv > 0) mpc_fold_kernel<<<(unsigned)((nrecv + 255) / 256), 256, 0, s>>>(nrecv, recv, ptr, dep, coef,
                                                                                    free_mask, v);
    vec_free_mask_kernel<<<(unsigned)((ndof + 255) / 256), 256, 0, s>>>(ndof, free_mask, v);
}

void launch_mpc_apply(int64_t nt, const int64_t *tgt, const int32_t *tptr, const int64_t *src, const double *coef,
                      int64_t nz, const int64_t *zpos, double *K, hipStream_t s) {
    if (nt > 0) mpc_apply_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(nt, tgt, tptr, src, coef, K);
    if (nz > 0) mpc_zero_kernel<<<(unsigned)((nz + 255) / 256), 256, 0, s>>>(nz, zpos, K);
}

void launch_mpc_recover(int ndep, const int32_t *dep, const int32_t *ptr, const int32_t *ind, const double *coef,
                        double *x, int64_t ld, int ncols, hipStream_t s) {
    const int64_t n = (int64_t)ndep * ncols;
    if (n > 0) mpc_recover_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ndep, dep, ptr, ind, coef, x, ld, ncols);
}

void launch_block_gather(int64_t n_blocks, const int32_t *cptr, const int32_t *cidx, const double *eb,
                         double *vals, hipStream_t s) {
    int64_t n = n_blocks * 36;
    if (n == 0) return;
    block_gather_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n_blocks, cptr, cidx, eb, vals);
}
void launch_q4_sens_terms(const ElemArgs &a, const double *u, const double *phi, const double *psi,
                          double *terms /* n_q4 x 3 */, hipStream_t s) {
    if (a.n_q4 == 0) return;
    q4_dke_kernel<<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, psi, terms);
    q4_kgt_kernel<4><<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, terms);
}
void launch_q4_sens_terms_ref(const ElemArgs &a, const double *u, const double *phi, const double *psi,
                              bool six_modes, double *terms /* n_q4 x 3 */, hipStream_t s) {
    if (a.n_q4 == 0) return;
    q4_dke_ref_kernel<<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, psi, terms);
    if (six_modes) q4_kgt_kernel<6><<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, terms);
    else q4_kgt_kernel<4><<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, terms);
}
void launch_q4_adjoint_rhs_ref(const ElemArgs &a, const double *u, const double *phi, double lam, bool six_modes,
                               double *eg, hipStream_t s) {
    if (a.n_q4 == 0) return;
    if (six_modes) q4_adjrhs_ref_kernel<6><<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, lam, eg);
    else q4_adjrhs_ref_kernel<4><<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, lam, eg);
}
void launch_sens_combine(int n_q4, const double *terms, double lam, double phiKGphi, double *dldt, hipStream_t s) {
    if (n_q4 == 0) return;
    sens_combine_kernel<<<(n_q4 + 255) / 256, 256, 0, s>>>(n_q4, terms, lam, phiKGphi, dldt);
}
void launch_q4_adjoint_rhs(const ElemArgs &a, const double *u, const double *phi, double lam, double *eg,
                           hipStream_t s) {
    if (a.n_q4 == 0) return;
    q4_adjrhs_kernel<<<(a.n_q4 + EPB - 1) / EPB, 256, 0, s>>>(a, u, phi, lam, eg);
}
void launch_bar_adjoint_rhs(const ElemArgs &a, const double *phi, double lam, double *eg, hipStream_t s) {
    if (a.n_bar == 0) return;
    bar_adjrhs_kernel<<<(a.n_bar + 127) / 128, 128, 0, s>>>(a, phi, lam, eg);
}
void launch_bsr_residual(int n_nodes, const int32_t *rp, const int32_t *col, const double *K, const double *u,
                         const double *F, double *r, hipStream_t s) {
    if (n_nodes > 0) bsr_residual_kernel<<<(n_nodes + 31) / 32, 192, 0, s>>>(n_nodes, rp, col, K, u, F, r);
}
void launch_q4_recover(const ElemArgs &a, const double *u, double *out, hipStream_t s) {
    if (a.n_q4 > 0) q4_recover_kernel<<<(a.n_q4 + 127) / 128, 128, 0, s>>>(a, u, out);
}
void launch_t3_recover(const ElemArgs &a, const double *u, const double *t3w, double *out, hipStream_t s) {
    if (a.n_t3 > 0) t3_recover_kernel<<<(a.n_t3 + EPB - 1) / EPB, 256, 0, s>>>(a, u, t3w, out);
}
void launch_bar_recover(const ElemArgs &a, const double *u, const double *pts, double *out, hipStream_t s) {
    if (a.n_bar > 0) bar_recover_kernel<<<(a.n_bar + 127) / 128, 128, 0, s>>>(a, u, pts, out);
}
