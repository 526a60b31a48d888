/*
 * sol105_oracle.c -- CPU restatement of OpenJFEM's default SOL 105 element and
 * assembly arithmetic, used ONLY as a test oracle (tests/, bench.py's
 * cpu_baseline leg and __graft_entry__.smoke()).  Nothing in the product path
 * links or calls this file.
 *
 * Scope restated here (default flags, flat isotropic PSHELL + PBARL BAR):
 *   - CQUAD4 element frame (:diag mode)       JFEM/src/solver/helpers.jl:1134-1193
 *   - MacNeal CQUAD4 stiffness                 JFEM/src/FEMKernels.jl:3185-4034
 *       (2x2 Gauss membrane/bending, MacNeal twist row, Hughes-Brezzi drilling,
 *        Wilson membrane + bending incompatible modes condensed)
 *   - MacNeal residual-bending-flexibility shear  JFEM/src/FEMKernels.jl:4715-4978
 *   - shell T'KT transform and COO order       JFEM/src/solver/assembly.jl:3808-3851
 *   - CTRIA3 frame                             JFEM/src/solver/helpers.jl:1195-1204
 *   - CTRIA3 stiffness: CST membrane, 3-quad macro plate block with the
 *     12 interior DOFs condensed, centroid drilling
 *                                              JFEM/src/FEMKernels.jl:7129-7209, 7315-7385,
 *                                              JFEM/src/solver/assembly.jl:3862-3955
 *   - CTRIA3 geometric stiffness (CST stress; principal-transverse falls
 *     through to the all-translations branch)  JFEM/src/FEMKernels.jl:7391-7432, 7989-8103,
 *                                              JFEM/src/solver/assembly.jl:6020-6160
 *   - CBAR Timoshenko 12x12 + frame            JFEM/src/FEMKernels.jl:140-213,
 *                                              JFEM/src/solver/assembly.jl:3960-4020
 *   - Q4 membrane force field (Wilson modes)   JFEM/src/FEMKernels.jl:6341-6539
 *   - Q4 condensation map A = -Kbb^-1 Kab^T    JFEM/src/FEMKernels.jl:6708-6785
 *   - Q4 principal-transverse geometric stiffness
 *                                              JFEM/src/FEMKernels.jl:6891-6972, 7554-7985
 *   - Q4 KG assembly (u_e = T u, sigma = N_gp/h, T'KGT, drilling zeroed)
 *                                              JFEM/src/solver/assembly.jl:5175-5185,
 *                                              5383-5394, 5958-6007
 *   - CBAR axial force + beam-column KG        JFEM/src/FEMKernels.jl:215-229, 7441-7471,
 *                                              JFEM/src/solver/assembly.jl:6272-6345
 *
 * Every routine is plain double arithmetic in the order the reference states
 * the mathematics; it is deliberately unoptimised.  Matrices are row-major.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define PT (0.57735026918962576451) /* 1/sqrt(3) */

/* ------------------------------------------------------------------------ */
/* small dense helpers                                                      */
/* ------------------------------------------------------------------------ */

static void v3_cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double v3_dot(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void v3_normalize(double a[3]) {
    double n = sqrt(v3_dot(a, a));
    a[0] /= n; a[1] /= n; a[2] /= n;
}

/* Gauss-Jordan inverse with partial pivoting of an n x n matrix (n <= 16). */
static int inv_n(int n, const double *a, double *ainv) {
    double m[16][32];
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) { m[i][j] = a[i * n + j]; m[i][n + j] = (i == j) ? 1.0 : 0.0; }
    }
    for (int c = 0; c < n; c++) {
        int p = c; double best = fabs(m[c][c]);
        for (int r = c + 1; r < n; r++) if (fabs(m[r][c]) > best) { best = fabs(m[r][c]); p = r; }
        if (best == 0.0) return -1;
        if (p != c) for (int j = 0; j < 2 * n; j++) { double t = m[c][j]; m[c][j] = m[p][j]; m[p][j] = t; }
        double d = m[c][c];
        for (int j = 0; j < 2 * n; j++) m[c][j] /= d;
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            double f = m[r][c];
            if (f != 0.0) for (int j = 0; j < 2 * n; j++) m[r][j] -= f * m[c][j];
        }
    }
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) ainv[i * n + j] = m[i][n + j];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* element frame                                                            */
/* ------------------------------------------------------------------------ */

/* CQUAD4 frame, :diag mode.  helpers.jl:1134-1193.
 * v3 = normalize(g_xi x g_eta) (fallback d13 x d24); x_raw = d13/|d13| - d24/|d24|
 * (fallback sum), projected on the plane, v1 = normalize, v2 = v3 x v1. */
void or_q4_frame(const double p[12], double v1[3], double v2[3], double v3[3]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6, *p4 = p + 9;
    double gxi[3], geta[3], n[3];
    for (int i = 0; i < 3; i++) {
        gxi[i] = 0.25 * (-p1[i] + p2[i] + p3[i] - p4[i]);
        geta[i] = 0.25 * (-p1[i] - p2[i] + p3[i] + p4[i]);
    }
    v3_cross(gxi, geta, n);
    if (v3_dot(n, n) <= 1e-24) {
        double d13[3], d24[3];
        for (int i = 0; i < 3; i++) { d13[i] = p3[i] - p1[i]; d24[i] = p4[i] - p2[i]; }
        v3_cross(d13, d24, n);
    }
    v3_normalize(n);
    for (int i = 0; i < 3; i++) v3[i] = n[i];
    double d13[3], d24[3], x[3];
    for (int i = 0; i < 3; i++) { d13[i] = p3[i] - p1[i]; d24[i] = p4[i] - p2[i]; }
    v3_normalize(d13); v3_normalize(d24);
    for (int i = 0; i < 3; i++) x[i] = d13[i] - d24[i];
    if (v3_dot(x, x) < 1e-20) for (int i = 0; i < 3; i++) x[i] = d13[i] + d24[i];
    double xd = v3_dot(x, v3);
    double xp[3];
    for (int i = 0; i < 3; i++) xp[i] = x[i] - xd * v3[i];
    if (v3_dot(xp, xp) <= 1e-24) {
        double gd = v3_dot(geta, v3);
        for (int i = 0; i < 3; i++) xp[i] = geta[i] - gd * v3[i];
        if (v3_dot(xp, xp) <= 1e-24) {
            double e[3]; for (int i = 0; i < 3; i++) e[i] = p2[i] - p1[i];
            double ed = v3_dot(e, v3);
            for (int i = 0; i < 3; i++) xp[i] = e[i] - ed * v3[i];
        }
    }
    for (int i = 0; i < 3; i++) v1[i] = xp[i];
    v3_normalize(v1);
    v3_cross(v3, v1, v2);
}

/* local 2-D corner coordinates about the centroid.  assembly.jl:2693-2698 */
void or_q4_local_coords(const double p[12], const double v1[3], const double v2[3], double lc[8]) {
    double c[3];
    for (int i = 0; i < 3; i++) c[i] = (p[i] + p[3 + i] + p[6 + i] + p[9 + i]) / 4.0;
    for (int k = 0; k < 4; k++) {
        double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
        lc[2 * k] = v3_dot(d, v1);
        lc[2 * k + 1] = v3_dot(d, v2);
    }
}

/* ------------------------------------------------------------------------ */
/* CQUAD4 MacNeal stiffness                                                 */
/* ------------------------------------------------------------------------ */

static void shape_derivs(double r, double s, double dNr[4], double dNs[4]) {
    dNr[0] = -0.25 * (1 - s); dNr[1] = 0.25 * (1 - s); dNr[2] = 0.25 * (1 + s); dNr[3] = -0.25 * (1 + s);
    dNs[0] = -0.25 * (1 - r); dNs[1] = -0.25 * (1 + r); dNs[2] = 0.25 * (1 + r); dNs[3] = 0.25 * (1 - r);
}
static void shape_vals(double r, double s, double N[4]) {
    N[0] = 0.25 * (1 - r) * (1 - s); N[1] = 0.25 * (1 + r) * (1 - s);
    N[2] = 0.25 * (1 + r) * (1 + s); N[3] = 0.25 * (1 - r) * (1 + s);
}
static void jac(const double lc[8], const double dNr[4], const double dNs[4],
                double *J11, double *J12, double *J21, double *J22) {
    *J11 = dNr[0] * lc[0] + dNr[1] * lc[2] + dNr[2] * lc[4] + dNr[3] * lc[6];
    *J12 = dNr[0] * lc[1] + dNr[1] * lc[3] + dNr[2] * lc[5] + dNr[3] * lc[7];
    *J21 = dNs[0] * lc[0] + dNs[1] * lc[2] + dNs[2] * lc[4] + dNs[3] * lc[6];
    *J22 = dNs[0] * lc[1] + dNs[1] * lc[3] + dNs[2] * lc[5] + dNs[3] * lc[7];
}

/* MacNeal (1978) residual-bending-flexibility transverse shear, default
 * constants eps = 0.04, zb_scale = 0.65.  FEMKernels.jl:4715-4978.  Adds
 * K_plate = D^T (Zs + Zb)^-1 D on the plate DOFs (w, thx, thy) of Ke. */
void or_macneal_shear(const double lc[8], const double Cb[9], const double Cs[4], double h,
                      double *Ke /*24x24*/) {
    double csmax = fmax(fmax(fabs(Cs[0]), fabs(Cs[1])), fmax(fabs(Cs[2]), fabs(Cs[3])));
    if (h < 1e-30 || csmax < 1e-30) return;
    const double sp[4][3] = {{0.0, -PT, 1}, {0.0, PT, 1}, {-PT, 0.0, 2}, {PT, 0.0, 2}};
    double D[4][12];
    double Jp[4];
    memset(D, 0, sizeof(D));
    for (int p = 0; p < 4; p++) {
        double xi = sp[p][0], eta = sp[p][1];
        int comp = (int)sp[p][2];
        double dNr[4], dNs[4], N[4], J11, J12, J21, J22;
        shape_derivs(xi, eta, dNr, dNs);
        shape_vals(xi, eta, N);
        jac(lc, dNr, dNs, &J11, &J12, &J21, &J22);
        double detJ = J11 * J22 - J12 * J21;
        if (fabs(detJ) < 1e-14) detJ = detJ < 0.0 ? -1e-14 : 1e-14;
        Jp[p] = fabs(detJ);
        double id = 1.0 / detJ;
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
        for (int k = 0; k < 4; k++) {
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k];
            double dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            if (comp == 1) { D[p][3 * k] = dx; D[p][3 * k + 2] = N[k]; }      /* gxz = w,x + thy */
            else           { D[p][3 * k] = dy; D[p][3 * k + 1] = -N[k]; }     /* gyz = w,y - thx */
        }
    }
    double Dx = 0.5 * (lc[2] + lc[4] - lc[0] - lc[6]);
    double Dy = 0.5 * (lc[5] + lc[7] - lc[1] - lc[3]);
    double Dx2 = Dx * Dx, Dy2 = Dy * Dy;
    double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11c, J12c, J21c, J22c;
    jac(lc, dNr_c, dNs_c, &J11c, &J12c, &J21c, &J22c);
    double A_elem = 4.0 * fabs(J11c * J22c - J12c * J21c);
    double eps = 0.04;
    double a_par = eps / (eps + (1.0 - eps) * Dx2 / fmax(Dy2, 1e-30));
    double b_par = eps / (eps + (1.0 - eps) * Dy2 / fmax(Dx2, 1e-30));
    double flex_x = 1.0 / fmax(fabs(Cb[0]), 1e-30);
    double flex_y = 1.0 / fmax(fabs(Cb[4]), 1e-30);
    double inv12A = 1.0 / (12.0 * A_elem);
    double zb = 0.65;
    double zbx_u = zb * inv12A * Dx2 * flex_x;
    double zbx_d = zb * inv12A * a_par * Dx2 * flex_x;
    double zby_u = zb * inv12A * Dy2 * flex_y;
    double zby_d = zb * inv12A * b_par * Dy2 * flex_y;
    double Zb[16] = {0};
    Zb[0] = zbx_u + zbx_d; Zb[1] = zbx_u - zbx_d; Zb[4] = Zb[1]; Zb[5] = Zb[0];
    Zb[10] = zby_u + zby_d; Zb[11] = zby_u - zby_d; Zb[14] = Zb[11]; Zb[15] = Zb[10];
    double Gxx = Cs[0], Gyy = Cs[3], Gxy = Cs[1];
    const int comps[4] = {1, 1, 2, 2};
    double VGV[16];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) {
        double Jf = sqrt(2.0 * Jp[i]) * sqrt(2.0 * Jp[j]);
        double v;
        if (comps[i] == comps[j]) v = (i == j) ? Jf * (comps[i] == 1 ? Gxx : Gyy) : 0.0;
        else v = 0.5 * Jf * Gxy;
        VGV[i * 4 + j] = v;
    }
    for (int i = 0; i < 4; i++) if (VGV[i * 5] < 1e-30) VGV[i * 5] = 1e-30;
    double VS[16], Zs[16];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) VS[i * 4 + j] = 0.5 * (VGV[i * 4 + j] + VGV[j * 4 + i]);
    inv_n(4, VS, Zs);
    double Z[16], Zt[16], Zi[16];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) Zt[i * 4 + j] = 0.5 * (Zs[i * 4 + j] + Zs[j * 4 + i]) + Zb[i * 4 + j];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) Z[i * 4 + j] = 0.5 * (Zt[i * 4 + j] + Zt[j * 4 + i]);
    inv_n(4, Z, Zi);
    double ZD[4][12];
    for (int i = 0; i < 4; i++) for (int c = 0; c < 12; c++) {
        double s = 0; for (int l = 0; l < 4; l++) s += Zi[i * 4 + l] * D[l][c];
        ZD[i][c] = s;
    }
    double Kp[12][12];
    for (int a = 0; a < 12; a++) for (int b = 0; b < 12; b++) {
        double s = 0; for (int l = 0; l < 4; l++) s += D[l][a] * ZD[l][b];
        Kp[a][b] = s;
    }
    const int pd[12] = {2, 3, 4, 8, 9, 10, 14, 15, 16, 20, 21, 22};
    for (int a = 0; a < 12; a++) for (int b = 0; b < 12; b++)
        Ke[pd[a] * 24 + pd[b]] += 0.5 * (Kp[a][b] + Kp[b][a]);
}

/* MacNeal CQUAD4 element stiffness in element-local coordinates, default
 * flags: 2x2 Gauss, twist correction, drilling, MacNeal shear, incompatible
 * modes condensed.  FEMKernels.jl:3185-4034 (default branch). */
void or_q4_ke_local(const double lc[8], const double Cm[9], const double Cb[9], const double Cs[4],
                    double h, double Eref, double k6rot, double drill_scale,
                    int membrane_incomp, int bending_incomp, double *Ke /*24x24*/) {
    double Kab[24][4], Kbb[4][4], Kabb[24][4], Kbbb[4][4];
    memset(Kab, 0, sizeof(Kab)); memset(Kbb, 0, sizeof(Kbb));
    memset(Kabb, 0, sizeof(Kabb)); memset(Kbbb, 0, sizeof(Kbbb));
    memset(Ke, 0, sizeof(double) * 576);
    double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11c, J12c, J21c, J22c;
    jac(lc, dNr_c, dNs_c, &J11c, &J12c, &J21c, &J22c);
    double detJc = J11c * J22c - J12c * J21c;
    double idc = 1.0 / detJc;
    double iJ11c = J22c * idc, iJ12c = -J12c * idc, iJ21c = -J21c * idc, iJ22c = J11c * idc;
    double dNdx_c[4], dNdy_c[4];
    for (int k = 0; k < 4; k++) {
        dNdx_c[k] = iJ11c * dNr_c[k] + iJ12c * dNs_c[k];
        dNdy_c[k] = iJ21c * dNr_c[k] + iJ22c * dNs_c[k];
    }
    double G_drill = Cm[8] / h;
    if (G_drill < 1e-6) G_drill = Eref / (2 * 3.0);
    double alpha_drill = drill_scale * (k6rot / 1e5) * G_drill * h;
    const double gps[4][2] = {{-PT, -PT}, {PT, -PT}, {PT, PT}, {-PT, PT}};
    for (int g = 0; g < 4; g++) {
        double r = gps[g][0], s = gps[g][1];
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jac(lc, dNr, dNs, &J11, &J12, &J21, &J22);
        double detJ = J11 * J22 - J12 * J21;
        double adj = fabs(detJ);
        if (adj < 1e-12) adj = 1e-12;
        double id = 1.0 / detJ;
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
        double Bm[3][24], Bb[3][24], Bd[24], Bi[3][4], Bib[3][4];
        memset(Bm, 0, sizeof(Bm)); memset(Bb, 0, sizeof(Bb)); memset(Bd, 0, sizeof(Bd));
        memset(Bi, 0, sizeof(Bi)); memset(Bib, 0, sizeof(Bib));
        for (int k = 0; k < 4; k++) {
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k];
            double dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            int idx = 6 * k;
            double Nk = 0.25 * (1 + ((k == 1 || k == 2) ? r : -r)) * (1 + (k >= 2 ? s : -s));
            Bm[0][idx] = dx; Bm[1][idx + 1] = dy; Bm[2][idx] = dy; Bm[2][idx + 1] = dx;
            Bb[0][idx + 4] = dx; Bb[1][idx + 3] = -dy;
            /* MacNeal twist row: chi_xy(gp) -> 2 chi_xy(gp) - chi_xy(0) */
            Bb[2][idx + 4] = 2.0 * dy - dNdy_c[k];
            Bb[2][idx + 3] = 2.0 * (-dx) - (-dNdx_c[k]);
            Bd[idx] = 0.5 * dy; Bd[idx + 1] = -0.5 * dx; Bd[idx + 5] = Nk;
        }
        double dp1x = iJ11 * (-2.0 * r), dp1y = iJ21 * (-2.0 * r);
        double dp2x = iJ12 * (-2.0 * s), dp2y = iJ22 * (-2.0 * s);
        Bi[0][0] = dp1x; Bi[2][0] = dp1y; Bi[1][1] = dp1y; Bi[2][1] = dp1x;
        Bi[0][2] = dp2x; Bi[2][2] = dp2y; Bi[1][3] = dp2y; Bi[2][3] = dp2x;
        Bib[1][0] = -dp1y; Bib[1][1] = -dp2y; Bib[0][2] = dp1x; Bib[0][3] = dp2x;
        Bib[2][0] = -dp1x; Bib[2][1] = -dp2x; Bib[2][2] = dp1y; Bib[2][3] = dp2y;
        /* CmBm, CbBb */
        double CmBm[3][24], CbBb[3][24], CmBi[3][4], CbBib[3][4];
        for (int i = 0; i < 3; i++) for (int j = 0; j < 24; j++) {
            CmBm[i][j] = Cm[i * 3] * Bm[0][j] + Cm[i * 3 + 1] * Bm[1][j] + Cm[i * 3 + 2] * Bm[2][j];
            CbBb[i][j] = Cb[i * 3] * Bb[0][j] + Cb[i * 3 + 1] * Bb[1][j] + Cb[i * 3 + 2] * Bb[2][j];
        }
        for (int i = 0; i < 3; i++) for (int j = 0; j < 4; j++) {
            CmBi[i][j] = Cm[i * 3] * Bi[0][j] + Cm[i * 3 + 1] * Bi[1][j] + Cm[i * 3 + 2] * Bi[2][j];
            CbBib[i][j] = Cb[i * 3] * Bib[0][j] + Cb[i * 3 + 1] * Bib[1][j] + Cb[i * 3 + 2] * Bib[2][j];
        }
        for (int a = 0; a < 24; a++) for (int b = 0; b < 24; b++) {
            double v = 0;
            for (int i = 0; i < 3; i++) v += Bm[i][a] * CmBm[i][b] + Bb[i][a] * CbBb[i][b];
            Ke[a * 24 + b] += adj * v + adj * alpha_drill * Bd[a] * Bd[b];
        }
        if (membrane_incomp) {
            for (int a = 0; a < 24; a++) for (int m = 0; m < 4; m++) {
                double v = 0; for (int i = 0; i < 3; i++) v += Bm[i][a] * CmBi[i][m];
                Kab[a][m] += adj * v;
            }
            for (int l = 0; l < 4; l++) for (int m = 0; m < 4; m++) {
                double v = 0; for (int i = 0; i < 3; i++) v += Bi[i][l] * CmBi[i][m];
                Kbb[l][m] += adj * v;
            }
        }
        for (int a = 0; a < 24; a++) for (int m = 0; m < 4; m++) {
            double v = 0; for (int i = 0; i < 3; i++) v += Bb[i][a] * CbBib[i][m];
            Kabb[a][m] += adj * v;
        }
        for (int l = 0; l < 4; l++) for (int m = 0; m < 4; m++) {
            double v = 0; for (int i = 0; i < 3; i++) v += Bib[i][l] * CbBib[i][m];
            Kbbb[l][m] += adj * v;
        }
    }
    or_macneal_shear(lc, Cb, Cs, h, Ke);
    double kbbmax = 0; for (int l = 0; l < 16; l++) kbbmax = fmax(kbbmax, fabs(((double *)Kbbb)[l]));
    int do_bend = bending_incomp && kbbmax > 0.0;
    if (do_bend) {
        double inv[16];
        inv_n(4, (double *)Kbbb, inv);
        for (int i = 0; i < 24; i++) for (int j = 0; j < 24; j++) {
            double sb = 0;
            for (int l = 0; l < 4; l++) {
                double t = 0; for (int q = 0; q < 4; q++) t += inv[l * 4 + q] * Kabb[j][q];
                sb += Kabb[i][l] * t;
            }
            Ke[i * 24 + j] -= sb;
        }
    }
    if (membrane_incomp) {
        double inv[16];
        inv_n(4, (double *)Kbb, inv);
        for (int i = 0; i < 24; i++) for (int j = 0; j < 24; j++) {
            double sm = 0;
            for (int l = 0; l < 4; l++) {
                double t = 0; for (int q = 0; q < 4; q++) t += inv[l * 4 + q] * Kab[j][q];
                sm += Kab[i][l] * t;
            }
            Ke[i * 24 + j] -= sm;
        }
    }
}

/* isotropic PSHELL constitutive matrices, assembly.jl:2460-2476, 1808-1857 */
static void shell_constitutive(double E, double nu, double Gs, double h, double br, double tst,
                               double Cm[9], double Cb[9], double Cs[4]) {
    double cm = E * h / (1 - nu * nu);
    double cb = br * (E * h * h * h) / (12 * (1 - nu * nu));
    const double D[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
    for (int i = 0; i < 9; i++) { Cm[i] = cm * D[i]; Cb[i] = cb * D[i]; }
    Cs[0] = tst * h * Gs; Cs[1] = 0; Cs[2] = 0; Cs[3] = tst * h * Gs;
    if (br <= 1e-12) { Cs[0] = Cs[1] = Cs[2] = Cs[3] = 0.0; }
}

/* T'AT for a block-diagonal T with nb 3x3 copies of R (rows = local axes). */
static void rotate_blocks(int n, const double R[9], const double *A, double *out) {
    /* out = T^T A T, T = blockdiag(R,...,R) */
    for (int bi = 0; bi < n / 3; bi++) for (int bj = 0; bj < n / 3; bj++) {
        double blk[9];
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) blk[i * 3 + j] = A[(bi * 3 + i) * n + bj * 3 + j];
        double tmp[9];
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
            double s = 0; for (int l = 0; l < 3; l++) s += blk[i * 3 + l] * R[l * 3 + j];
            tmp[i * 3 + j] = s;
        }
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
            double s = 0; for (int l = 0; l < 3; l++) s += R[l * 3 + i] * tmp[l * 3 + j];
            out[(bi * 3 + i) * n + bj * 3 + j] = s;
        }
    }
}

/* Element frame + local stiffness + T'KT for one flat isotropic CQUAD4.
 * prop = [E, nu, h, bend_ratio, ts_t, G_shear].  assembly.jl:2657-3851. */
void or_q4_ke_global(const double p[12], const double prop[6], double k6rot, double *Keg /*24x24*/,
                     double R_out[9], double lc_out[8]) {
    double v1[3], v2[3], v3[3], lc[8];
    or_q4_frame(p, v1, v2, v3);
    or_q4_local_coords(p, v1, v2, lc);
    double E = prop[0], nu = prop[1], h = prop[2], br = prop[3], tst = prop[4], Gs = prop[5];
    double Cm[9], Cb[9], Cs[4];
    shell_constitutive(E, nu, Gs, h, br, tst, Cm, Cb, Cs);
    double ek6 = br <= 1e-12 ? 0.0 : k6rot;
    double eds = br <= 1e-12 ? 0.0 : 1.0;
    double Ke[576];
    or_q4_ke_local(lc, Cm, Cb, Cs, h, E, ek6, eds, 1, 1, Ke);
    double R[9] = {v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
    rotate_blocks(24, R, Ke, Keg);
    if (R_out) memcpy(R_out, R, sizeof(R));
    if (lc_out) memcpy(lc_out, lc, sizeof(lc));
}

/* ------------------------------------------------------------------------ */
/* CTRIA3                                                                   */
/* ------------------------------------------------------------------------ */

/* CTRIA3 frame: v1 = e12, v3 = normalize(v1 x (p3 - p1)), v2 = v3 x v1
 * (helpers.jl:1195-1204); local corner coordinates about the centroid
 * (assembly.jl:3927-3930). */
void or_t3_frame(const double p[9], double R[9], double lc[6]) {
    double v1[3], v3[3], v2[3], d13[3];
    for (int i = 0; i < 3; i++) { v1[i] = p[3 + i] - p[i]; d13[i] = p[6 + i] - p[i]; }
    v3_normalize(v1);
    v3_cross(v1, d13, v3);
    v3_normalize(v3);
    v3_cross(v3, v1, v2);
    for (int i = 0; i < 3; i++) { R[i] = v1[i]; R[3 + i] = v2[i]; R[6 + i] = v3[i]; }
    double c[3];
    for (int i = 0; i < 3; i++) c[i] = (p[i] + p[3 + i] + p[6 + i]) / 3.0;
    for (int k = 0; k < 3; k++) {
        double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
        lc[2 * k] = v3_dot(d, v1);
        lc[2 * k + 1] = v3_dot(d, v2);
    }
}

/* Gaussian elimination with partial pivoting (Golub & Van Loan 3.4): solves
 * A X = B in place, A n x n, B n x nrhs, row-major.  Returns 0 on success. */
static int lu_solve_n(int n, double *A, int nrhs, double *B) {
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++) if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
        if (A[p * n + k] == 0.0) return -1;
        if (p != k) {
            for (int j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
            for (int j = 0; j < nrhs; j++) { double t = B[k * nrhs + j]; B[k * nrhs + j] = B[p * nrhs + j]; B[p * nrhs + j] = t; }
        }
        for (int i = k + 1; i < n; i++) {
            double l = A[i * n + k] / A[k * n + k];
            A[i * n + k] = l;
            for (int j = k + 1; j < n; j++) A[i * n + j] -= l * A[k * n + j];
            for (int j = 0; j < nrhs; j++) B[i * nrhs + j] -= l * B[k * nrhs + j];
        }
    }
    for (int k = n - 1; k >= 0; k--)
        for (int j = 0; j < nrhs; j++) {
            double s = B[k * nrhs + j];
            for (int l = k + 1; l < n; l++) s -= A[k * n + l] * B[l * nrhs + j];
            B[k * nrhs + j] = s / A[k * n + k];
        }
    return 0;
}

/* The condensed plate block of the CTRIA3 macro element.  FEMKernels.jl:
 * 7129-7209.  Seven points (corners, mid-sides 12/23/31, centroid) carry
 * (w, thx, thy); three virtual CQUAD4s (1,4,7,6), (2,5,7,4), (3,6,7,5)
 * contribute the plate rows/columns of the MacNeal stiffness with the
 * kernel's default flags (bending incompatible modes NOT condensed); the 12
 * interior DOFs are condensed: Kcond = Kee - Kei Kii^-1 Kie.  Also returns
 * the 21x21 assembled matrix when K21 is non-NULL. */
void or_t3_plate_macro(const double lc[6], const double Cm[9], const double Cb[9], const double Cs[4],
                       double h, double Gref, double k6rot, double *Kcond /*9x9*/, double *K21 /*21x21*/) {
    static const int quads[3][4] = {{0, 3, 6, 5}, {1, 4, 6, 3}, {2, 5, 6, 4}};
    static const int pd[12] = {2, 3, 4, 8, 9, 10, 14, 15, 16, 20, 21, 22};
    double pts[7][2];
    for (int k = 0; k < 3; k++) { pts[k][0] = lc[2 * k]; pts[k][1] = lc[2 * k + 1]; }
    for (int d = 0; d < 2; d++) {
        pts[3][d] = (lc[d] + lc[2 + d]) / 2;
        pts[4][d] = (lc[2 + d] + lc[4 + d]) / 2;
        pts[5][d] = (lc[4 + d] + lc[d]) / 2;
        pts[6][d] = (lc[d] + lc[2 + d] + lc[4 + d]) / 3;
    }
    double K[21 * 21];
    memset(K, 0, sizeof(K));
    for (int q = 0; q < 3; q++) {
        double qc[8];
        for (int k = 0; k < 4; k++) { qc[2 * k] = pts[quads[q][k]][0]; qc[2 * k + 1] = pts[quads[q][k]][1]; }
        double Kq[576];
        or_q4_ke_local(qc, Cm, Cb, Cs, h, Gref, k6rot, 1.0, 1, 0, Kq);
        for (int a = 0; a < 12; a++) for (int b = 0; b < 12; b++) {
            int ga = 3 * quads[q][a / 3] + a % 3, gb = 3 * quads[q][b / 3] + b % 3;
            K[ga * 21 + gb] += Kq[pd[a] * 24 + pd[b]];
        }
    }
    if (K21) memcpy(K21, K, sizeof(K));
    double Kii[144], X[12 * 9];
    for (int i = 0; i < 12; i++) {
        for (int j = 0; j < 12; j++) Kii[i * 12 + j] = K[(9 + i) * 21 + 9 + j];
        for (int j = 0; j < 9; j++) X[i * 9 + j] = K[(9 + i) * 21 + j];
    }
    lu_solve_n(12, Kii, 9, X);
    for (int i = 0; i < 9; i++) for (int j = 0; j < 9; j++) {
        double s = 0; for (int l = 0; l < 12; l++) s += K[i * 21 + 9 + l] * X[l * 9 + j];
        Kcond[i * 9 + j] = K[i * 21 + j] - s;
    }
}

/* CTRIA3 stiffness in element-local coordinates.  FEMKernels.jl:7315-7385:
 * CST membrane Bm' Dm Bm A on (u, v), the macro plate block on (w, thx, thy),
 * Hughes-Brezzi drilling alpha = (k6rot/1e5) G_ref h at the centroid. */
void or_t3_ke_local(const double lc[6], const double Cm[9], const double Cb[9], const double Cs[4], double h,
                    double Gref, double br, double k6rot, double *Ke /*18x18*/) {
    memset(Ke, 0, sizeof(double) * 324);
    const double *x0 = lc;
    double x[3] = {x0[0], x0[2], x0[4]}, y[3] = {x0[1], x0[3], x0[5]};
    double A2 = x[0] * (y[1] - y[2]) + x[1] * (y[2] - y[0]) + x[2] * (y[0] - y[1]);
    double A = 0.5 * fabs(A2);
    if (A < 1e-12) return;
    double bv[3] = {(y[1] - y[2]) / (2 * A), (y[2] - y[0]) / (2 * A), (y[0] - y[1]) / (2 * A)};
    double cv[3] = {(x[2] - x[1]) / (2 * A), (x[0] - x[2]) / (2 * A), (x[1] - x[0]) / (2 * A)};
    double Bm[3][6];
    memset(Bm, 0, sizeof(Bm));
    for (int i = 0; i < 3; i++) {
        Bm[0][2 * i] = bv[i]; Bm[1][2 * i + 1] = cv[i]; Bm[2][2 * i] = cv[i]; Bm[2][2 * i + 1] = bv[i];
    }
    static const int mi[6] = {0, 1, 6, 7, 12, 13};
    for (int a = 0; a < 6; a++) for (int b = 0; b < 6; b++) {
        double s = 0;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) s += Bm[i][a] * Cm[i * 3 + j] * Bm[j][b];
        Ke[mi[a] * 18 + mi[b]] = s * A;
    }
    if (br > 1e-12) {
        double Kc[81];
        or_t3_plate_macro(lc, Cm, Cb, Cs, h, Gref, k6rot, Kc, 0);
        static const int bi[9] = {2, 3, 4, 8, 9, 10, 14, 15, 16};
        for (int a = 0; a < 9; a++) for (int b = 0; b < 9; b++) Ke[bi[a] * 18 + bi[b]] += Kc[a * 9 + b];
    }
    double alpha = (k6rot / 1e5) * Gref * h;
    double Bd[18];
    memset(Bd, 0, sizeof(Bd));
    for (int i = 0; i < 3; i++) { Bd[6 * i] = 0.5 * cv[i]; Bd[6 * i + 1] = -0.5 * bv[i]; Bd[6 * i + 5] = 1.0 / 3.0; }
    for (int a = 0; a < 18; a++) for (int b = 0; b < 18; b++) Ke[a * 18 + b] += alpha * (Bd[a] * Bd[b]) * A;
}

/* Frame + local stiffness + T'KT for one flat isotropic CTRIA3.
 * prop = [E, nu, h, bend_ratio, ts_t, G_shear]; G_ref = G_shear
 * (assembly.jl:2454-2460, 2515); k6rot -> 0 for a membrane-only shell
 * (assembly.jl:3930-3933).  assembly.jl:3862-3955. */
void or_t3_ke_global(const double p[9], const double prop[6], double k6rot, double *Keg /*18x18*/,
                     double R_out[9], double lc_out[6]) {
    double R[9], lc[6];
    or_t3_frame(p, R, lc);
    double E = prop[0], nu = prop[1], h = prop[2], br = prop[3], tst = prop[4], Gs = prop[5];
    double Cm[9], Cb[9], Cs[4];
    shell_constitutive(E, nu, Gs, h, br, tst, Cm, Cb, Cs);
    double ek6 = br <= 1e-12 ? 0.0 : k6rot;
    double Ke[324];
    or_t3_ke_local(lc, Cm, Cb, Cs, h, Gs, br, ek6, Ke);
    rotate_blocks(18, R, Ke, Keg);
    if (R_out) memcpy(R_out, R, sizeof(R));
    if (lc_out) memcpy(lc_out, lc, sizeof(lc));
}

/* CTRIA3 geometric stiffness.  Membrane stress of the constant-strain
 * triangle sigma = N/h with N = (D eps) h, D = E/(1-nu^2)[1 nu 0; nu 1 0;
 * 0 0 (1-nu)/2] (FEMKernels.jl:7391-7432); KG = h A (sxx b_i b_j + syy c_i c_j
 * + sxy (b_i c_j + c_i b_j)) on u, v and w of each node pair -- the default
 * principal-transverse mode falls through to this branch (FEMKernels.jl:
 * 7989-8103); T'KG T with the drilling rows zeroed (assembly.jl:6020-6160).
 * Also returns the membrane stress when sig_out is non-NULL. */
void or_t3_kg_global(const double p[9], const double prop[6], const double *ug /*18*/, double *KGg /*18x18*/,
                     double sig_out[3]) {
    double R[9], lc[6];
    or_t3_frame(p, R, lc);
    double ul[18];
    for (int b = 0; b < 6; b++) for (int i = 0; i < 3; i++) {
        double s = 0; for (int j = 0; j < 3; j++) s += R[i * 3 + j] * ug[b * 3 + j];
        ul[b * 3 + i] = s;
    }
    double E = prop[0], nu = prop[1], h = prop[2];
    double x[3] = {lc[0], lc[2], lc[4]}, y[3] = {lc[1], lc[3], lc[5]};
    double A = 0.5 * fabs(x[0] * (y[1] - y[2]) + x[1] * (y[2] - y[0]) + x[2] * (y[0] - y[1]));
    double KG[324];
    memset(KG, 0, sizeof(KG));
    double sig[3] = {0, 0, 0};
    if (A >= 1e-12 && h >= 1e-30) {
        double bv[3] = {(y[1] - y[2]) / (2 * A), (y[2] - y[0]) / (2 * A), (y[0] - y[1]) / (2 * A)};
        double cv[3] = {(x[2] - x[1]) / (2 * A), (x[0] - x[2]) / (2 * A), (x[1] - x[0]) / (2 * A)};
        double exx = 0, eyy = 0, gxy = 0;
        for (int i = 0; i < 3; i++) {
            double ui = ul[6 * i], vi = ul[6 * i + 1];
            exx += bv[i] * ui;
            eyy += cv[i] * vi;
            gxy += cv[i] * ui + bv[i] * vi;
        }
        double c = E / (1 - nu * nu);
        double D[9] = {c, c * nu, 0.0, c * nu, c, 0.0, 0.0, 0.0, c * ((1 - nu) / 2)};
        double N[3];
        for (int i = 0; i < 3; i++) N[i] = (D[3 * i] * exx + D[3 * i + 1] * eyy + D[3 * i + 2] * gxy) * h;
        for (int i = 0; i < 3; i++) sig[i] = N[i] / h;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
            double v = h * A * (sig[0] * bv[i] * bv[j] + sig[1] * cv[i] * cv[j] + sig[2] * (bv[i] * cv[j] + cv[i] * bv[j]));
            for (int d = 0; d < 3; d++) KG[(6 * i + d) * 18 + 6 * j + d] += v;
        }
    }
    rotate_blocks(18, R, KG, KGg);
    for (int k = 0; k < 3; k++) {
        int d = 6 * k + 5;
        for (int r = 0; r < 18; r++) { KGg[r * 18 + d] = 0.0; KGg[d * 18 + r] = 0.0; }
    }
    if (sig_out) memcpy(sig_out, sig, sizeof(sig));
}

/* ------------------------------------------------------------------------ */
/* CBAR                                                                     */
/* ------------------------------------------------------------------------ */

/* Timoshenko 3-D frame stiffness.  FEMKernels.jl:140-213 (I12 = 0). */
void or_frame3d_ke(double L, double A, double Iy, double Iz, double J, double E, double G,
                   double Asy, double Asz, double *k /*12x12*/) {
    memset(k, 0, sizeof(double) * 144);
    if (L < 1e-9) return;
#define K(i, j) k[((i) - 1) * 12 + (j) - 1]
    double X = E * A / L;
    K(1, 1) = X; K(1, 7) = -X; K(7, 1) = -X; K(7, 7) = X;
    double T = G * J / L;
    K(4, 4) = T; K(4, 10) = -T; K(10, 4) = -T; K(10, 10) = T;
    double Py = isinf(Asz) ? 0.0 : 12 * E * Iy / (G * Asz * L * L);
    double Pz = isinf(Asy) ? 0.0 : 12 * E * Iz / (G * Asy * L * L);
    double ay = 12 * E * Iy / (L * L * L * (1 + Py)), by = 6 * E * Iy / (L * L * (1 + Py));
    double cy = (4 + Py) * E * Iy / (L * (1 + Py)), dy = (2 - Py) * E * Iy / (L * (1 + Py));
    K(3, 3) = ay; K(3, 9) = -ay; K(9, 3) = -ay; K(9, 9) = ay;
    K(3, 5) = -by; K(3, 11) = -by; K(5, 3) = -by; K(11, 3) = -by;
    K(9, 5) = by; K(9, 11) = by; K(5, 9) = by; K(11, 9) = by;
    K(5, 5) = cy; K(5, 11) = dy; K(11, 5) = dy; K(11, 11) = cy;
    double az = 12 * E * Iz / (L * L * L * (1 + Pz)), bz = 6 * E * Iz / (L * L * (1 + Pz));
    double cz = (4 + Pz) * E * Iz / (L * (1 + Pz)), dz = (2 - Pz) * E * Iz / (L * (1 + Pz));
    K(2, 2) = az; K(2, 8) = -az; K(8, 2) = -az; K(8, 8) = az;
    K(2, 6) = bz; K(2, 12) = bz; K(6, 2) = bz; K(12, 2) = bz;
    K(8, 6) = -bz; K(8, 12) = -bz; K(6, 8) = -bz; K(12, 8) = -bz;
    K(6, 6) = cz; K(6, 12) = dz; K(12, 6) = dz; K(12, 12) = cz;
#undef K
}

/* beam-column geometric stiffness.  FEMKernels.jl:7441-7471 */
void or_frame3d_kg(double L, double P, double *kg /*12x12*/) {
    memset(kg, 0, sizeof(double) * 144);
    if (L < 1e-9 || fabs(P) < 1e-30) return;
#define K(i, j) kg[((i) - 1) * 12 + (j) - 1]
    double c1 = 6.0 * P / (5.0 * L), c2 = P / 10.0, c3 = 2.0 * P * L / 15.0, c4 = -P * L / 30.0;
    K(2, 2) = c1; K(8, 8) = c1; K(2, 8) = -c1; K(8, 2) = -c1;
    K(2, 6) = c2; K(6, 2) = c2; K(2, 12) = c2; K(12, 2) = c2;
    K(6, 6) = c3; K(12, 12) = c3; K(6, 12) = c4; K(12, 6) = c4;
    K(8, 6) = -c2; K(6, 8) = -c2; K(8, 12) = -c2; K(12, 8) = -c2;
    K(3, 3) = c1; K(9, 9) = c1; K(3, 9) = -c1; K(9, 3) = -c1;
    K(3, 5) = -c2; K(5, 3) = -c2; K(3, 11) = -c2; K(11, 3) = -c2;
    K(5, 5) = c3; K(11, 11) = c3; K(5, 11) = c4; K(11, 5) = c4;
    K(9, 5) = c2; K(5, 9) = c2; K(9, 11) = c2; K(11, 9) = c2;
#undef K
}

/* CBAR frame (no offsets): vx along the bar, vz = vx x v_ref, vy = vz x vx.
 * assembly.jl:3978-3988 */
void or_bar_frame(const double p1[3], const double p2[3], const double vref_in[3], double R[9], double *L) {
    double vx[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    *L = sqrt(v3_dot(vx, vx));
    v3_normalize(vx);
    double vr[3] = {vref_in[0], vref_in[1], vref_in[2]};
    if (sqrt(v3_dot(vr, vr)) < 1e-6) {
        vr[0] = 0; vr[1] = 0; vr[2] = 1;
        if (fabs(v3_dot(vx, vr)) > 0.9) { vr[1] = 1; vr[2] = 0; }
    }
    double vz[3], vy[3];
    v3_cross(vx, vr, vz); v3_normalize(vz);
    v3_cross(vz, vx, vy);
    for (int i = 0; i < 3; i++) { R[i] = vx[i]; R[3 + i] = vy[i]; R[6 + i] = vz[i]; }
}

/* bar prop = [A, I1, I2, J, K1, K2, E, G] */
void or_bar_ke_global(const double p1[3], const double p2[3], const double vref[3], const double bp[8],
                      double *Keg /*12x12*/) {
    double R[9], L;
    or_bar_frame(p1, p2, vref, R, &L);
    double A = bp[0], I1 = bp[1], I2 = bp[2], J = bp[3], K1 = bp[4], K2 = bp[5], E = bp[6], G = bp[7];
    double Iy = I2, Iz = I1;
    if (Iy == 0.0) Iy = Iz;
    if (Iz == 0.0) Iz = Iy;
    double Asy = K1 > 0.0 ? K1 * A : INFINITY, Asz = K2 > 0.0 ? K2 * A : INFINITY;
    double k[144];
    or_frame3d_ke(L, A, Iy, Iz, J, E, G, Asy, Asz, k);
    rotate_blocks(12, R, k, Keg);
}

/* ------------------------------------------------------------------------ */
/* CQUAD4 geometric stiffness (principal-transverse, Wilson-mode consistent) */
/* ------------------------------------------------------------------------ */

/* K_ab, K_bb of the Wilson membrane modes with |det J| >= 1e-12 weights. */
static void q4_membrane_kab_kbb(const double lc[8], const double Cm[9], double Kab[24][4], double Kbb[4][4]) {
    memset(Kab, 0, sizeof(double) * 96); memset(Kbb, 0, sizeof(double) * 16);
    const double gps[4][2] = {{-PT, -PT}, {PT, -PT}, {PT, PT}, {-PT, PT}};
    for (int g = 0; g < 4; g++) {
        double r = gps[g][0], s = gps[g][1];
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jac(lc, dNr, dNs, &J11, &J12, &J21, &J22);
        double detJ = J11 * J22 - J12 * J21;
        double adj = fabs(detJ); if (adj < 1e-12) adj = 1e-12;
        double id = 1.0 / detJ;
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
        double Bm[3][24], Bi[3][4];
        memset(Bm, 0, sizeof(Bm)); memset(Bi, 0, sizeof(Bi));
        for (int k = 0; k < 4; k++) {
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k], dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            Bm[0][6 * k] = dx; Bm[1][6 * k + 1] = dy; Bm[2][6 * k] = dy; Bm[2][6 * k + 1] = dx;
        }
        double dp1x = iJ11 * (-2.0 * r), dp1y = iJ21 * (-2.0 * r);
        double dp2x = iJ12 * (-2.0 * s), dp2y = iJ22 * (-2.0 * s);
        Bi[0][0] = dp1x; Bi[2][0] = dp1y; Bi[1][1] = dp1y; Bi[2][1] = dp1x;
        Bi[0][2] = dp2x; Bi[2][2] = dp2y; Bi[1][3] = dp2y; Bi[2][3] = dp2x;
        double CmBi[3][4];
        for (int i = 0; i < 3; i++) for (int j = 0; j < 4; j++)
            CmBi[i][j] = Cm[i * 3] * Bi[0][j] + Cm[i * 3 + 1] * Bi[1][j] + Cm[i * 3 + 2] * Bi[2][j];
        for (int a = 0; a < 24; a++) for (int m = 0; m < 4; m++) {
            double v = 0; for (int i = 0; i < 3; i++) v += Bm[i][a] * CmBi[i][m];
            Kab[a][m] += v * adj;
        }
        for (int l = 0; l < 4; l++) for (int m = 0; m < 4; m++) {
            double v = 0; for (int i = 0; i < 3; i++) v += Bi[i][l] * CmBi[i][m];
            Kbb[l][m] += v * adj;
        }
    }
}

/* A = -Kbb^-1 Kab^T (4x24).  FEMKernels.jl:6708-6785 */
void or_q4_condensation_map(const double lc[8], const double Cm[9], double *Amap /*4x24*/) {
    double Kab[24][4], Kbb[4][4], inv[16];
    q4_membrane_kab_kbb(lc, Cm, Kab, Kbb);
    inv_n(4, (double *)Kbb, inv);
    for (int l = 0; l < 4; l++) for (int a = 0; a < 24; a++) {
        double s = 0; for (int q = 0; q < 4; q++) s += inv[l * 4 + q] * Kab[a][q];
        Amap[l * 24 + a] = -s;
    }
}

/* membrane resultants at the 4 Gauss points with condensed Wilson modes.
 * FEMKernels.jl:6341-6539 (use_incompatible_modes = true). */
void or_q4_membrane_force_field(const double lc[8], const double *u /*24*/, const double Cm[9],
                                double Ngp[12] /*4x3*/) {
    double Amap[96];
    or_q4_condensation_map(lc, Cm, Amap);
    double alpha[4];
    for (int l = 0; l < 4; l++) { double s = 0; for (int a = 0; a < 24; a++) s += Amap[l * 24 + a] * u[a]; alpha[l] = s; }
    const double gps[4][2] = {{-PT, -PT}, {PT, -PT}, {PT, PT}, {-PT, PT}};
    for (int g = 0; g < 4; g++) {
        double r = gps[g][0], s = gps[g][1];
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jac(lc, dNr, dNs, &J11, &J12, &J21, &J22);
        double detJ = J11 * J22 - J12 * J21;
        double id = 1.0 / detJ;
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
        double eps[3] = {0, 0, 0};
        for (int k = 0; k < 4; k++) {
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k], dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            eps[0] += dx * u[6 * k];
            eps[1] += dy * u[6 * k + 1];
            eps[2] += dy * u[6 * k] + dx * u[6 * k + 1];
        }
        double dp1x = iJ11 * (-2.0 * r), dp1y = iJ21 * (-2.0 * r);
        double dp2x = iJ12 * (-2.0 * s), dp2y = iJ22 * (-2.0 * s);
        eps[0] += dp1x * alpha[0] + dp2x * alpha[2];
        eps[1] += dp1y * alpha[1] + dp2y * alpha[3];
        eps[2] += dp1y * alpha[0] + dp1x * alpha[1] + dp2y * alpha[2] + dp2x * alpha[3];
        for (int i = 0; i < 3; i++) Ngp[g * 3 + i] = Cm[i * 3] * eps[0] + Cm[i * 3 + 1] * eps[1] + Cm[i * 3 + 2] * eps[2];
    }
}

/* principal stresses: (l1, c1, s1, l2, c2, s2).  FEMKernels.jl:6891-6902 */
static void principal2d(double sxx, double syy, double sxy, double out[6]) {
    double mean = 0.5 * (sxx + syy), hd = 0.5 * (sxx - syy);
    double rad = sqrt(hd * hd + sxy * sxy);
    if (rad <= 1e-30) { out[0] = mean; out[1] = 1; out[2] = 0; out[3] = mean; out[4] = 0; out[5] = 1; return; }
    double th = 0.5 * atan2(2.0 * sxy, sxx - syy);
    double c = cos(th), s = sin(th);
    out[0] = mean + rad; out[1] = c; out[2] = s; out[3] = mean - rad; out[4] = -s; out[5] = c;
}

/* principal-transverse geometric stiffness with the Wilson-mode displacement
 * gradients (membrane_A path).  FEMKernels.jl:7554-7985 with
 * trans_mode = :principal_transverse, factors 1.0. */
void or_q4_kg_local(const double lc[8], const double sig[12] /*4x3*/, double h, const double *Amap /*4x24*/,
                    double *KG /*24x24*/) {
    memset(KG, 0, sizeof(double) * 576);
    if (h < 1e-30) return;
    const double gps[4][2] = {{-PT, -PT}, {PT, -PT}, {PT, PT}, {-PT, PT}};
    for (int g = 0; g < 4; g++) {
        double sxx = sig[g * 3], syy = sig[g * 3 + 1], sxy = sig[g * 3 + 2];
        double r = gps[g][0], s = gps[g][1];
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jac(lc, dNr, dNs, &J11, &J12, &J21, &J22);
        double detJ = J11 * J22 - J12 * J21;
        double adj = fabs(detJ); if (adj < 1e-12) adj = 1e-12;
        double id = 1.0 / detJ;
        double iJ11 = J22 * id, iJ12 = -J12 * id, iJ21 = -J21 * id, iJ22 = J11 * id;
        double uxx[24] = {0}, uxy[24] = {0}, uyx[24] = {0}, uyy[24] = {0}, uzx[24] = {0}, uzy[24] = {0};
        for (int i = 0; i < 4; i++) {
            double dx = iJ11 * dNr[i] + iJ12 * dNs[i], dy = iJ21 * dNr[i] + iJ22 * dNs[i];
            uxx[6 * i] = dx; uxy[6 * i] = dy;
            uyx[6 * i + 1] = dx; uyy[6 * i + 1] = dy;
            uzx[6 * i + 2] = dx; uzy[6 * i + 2] = dy;
        }
        if (Amap) {
            double dp1x = iJ11 * (-2.0 * r), dp1y = iJ21 * (-2.0 * r);
            double dp2x = iJ12 * (-2.0 * s), dp2y = iJ22 * (-2.0 * s);
            for (int a = 0; a < 24; a++) {
                double a1 = Amap[a], a2 = Amap[24 + a], a3 = Amap[48 + a], a4 = Amap[72 + a];
                uxx[a] += dp1x * a1 + dp2x * a3;
                uxy[a] += dp1y * a1 + dp2y * a3;
                uyx[a] += dp1x * a2 + dp2x * a4;
                uyy[a] += dp1y * a2 + dp2y * a4;
            }
        }
        double scale = h * adj;
        double pr[6];
        principal2d(sxx, syy, sxy, pr);
        for (int k = 0; k < 2; k++) {
            double lam = pr[3 * k], c = pr[3 * k + 1], sn = pr[3 * k + 2];
            if (fabs(lam) <= 1e-30) continue;
            double p11 = sn * sn, p22 = c * c, p12 = -c * sn, f = scale * lam;
            for (int i = 0; i < 24; i++) {
                double gxi = c * uxx[i] + sn * uxy[i], gyi = c * uyx[i] + sn * uyy[i], gzi = c * uzx[i] + sn * uzy[i];
                for (int j = 0; j < 24; j++) {
                    double gxj = c * uxx[j] + sn * uxy[j], gyj = c * uyx[j] + sn * uyy[j], gzj = c * uzx[j] + sn * uzy[j];
                    KG[i * 24 + j] += f * (p11 * gxi * gxj + p22 * gyi * gyj + p12 * (gxi * gyj + gyi * gxj) + gzi * gzj);
                }
            }
        }
    }
}

/* Full Q4 KG: u_e = T u_glob, sigma = N_gp/h, KG_loc, T'KG T, drilling rows and
 * columns zeroed.  prop = [E, nu, h, ...]. assembly.jl:5175-6007 */
void or_q4_kg_global(const double p[12], const double prop[6], const double *ug /*24 global*/,
                     double *KGg /*24x24*/) {
    double v1[3], v2[3], v3[3], lc[8];
    or_q4_frame(p, v1, v2, v3);
    or_q4_local_coords(p, v1, v2, lc);
    double R[9] = {v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
    double ul[24];
    for (int b = 0; b < 8; b++) for (int i = 0; i < 3; i++) {
        double s = 0; for (int j = 0; j < 3; j++) s += R[i * 3 + j] * ug[b * 3 + j];
        ul[b * 3 + i] = s;
    }
    double E = prop[0], nu = prop[1], h = prop[2];
    double cm = E / (1 - nu * nu);
    const double D[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
    double Cm[9];
    for (int i = 0; i < 9; i++) Cm[i] = (cm * D[i]) * h;
    double Ngp[12];
    or_q4_membrane_force_field(lc, ul, Cm, Ngp);
    double sig[12];
    for (int i = 0; i < 12; i++) sig[i] = Ngp[i] / h;
    double Amap[96];
    or_q4_condensation_map(lc, Cm, Amap);
    double KG[576];
    or_q4_kg_local(lc, sig, h, Amap, KG);
    rotate_blocks(24, R, KG, KGg);
    for (int k = 0; k < 4; k++) {
        int d = 6 * k + 5;
        for (int r = 0; r < 24; r++) { KGg[r * 24 + d] = 0.0; KGg[d * 24 + r] = 0.0; }
    }
}

/* CBAR KG: P = (k u_local)[7], beam-column KG, T'KG T.  assembly.jl:6272-6345 */
void or_bar_kg_global(const double p1[3], const double p2[3], const double vref[3], const double bp[8],
                      const double *ug /*12 global*/, double *KGg /*12x12*/, double *P_out) {
    double R[9], L;
    or_bar_frame(p1, p2, vref, R, &L);
    double ul[12];
    for (int b = 0; b < 4; b++) for (int i = 0; i < 3; i++) {
        double s = 0; for (int j = 0; j < 3; j++) s += R[i * 3 + j] * ug[b * 3 + j];
        ul[b * 3 + i] = s;
    }
    double A = bp[0], I1 = bp[1], I2 = bp[2], J = bp[3], K1 = bp[4], K2 = bp[5], E = bp[6], G = bp[7];
    double Iy = I2, Iz = I1;
    if (Iy == 0.0) Iy = Iz;
    if (Iz == 0.0) Iz = Iy;
    double Asy = K1 > 0.0 ? K1 * A : INFINITY, Asz = K2 > 0.0 ? K2 * A : INFINITY;
    double k[144];
    or_frame3d_ke(L, A, Iy, Iz, J, E, G, Asy, Asz, k);
    double P = 0; for (int j = 0; j < 12; j++) P += k[6 * 12 + j] * ul[j];
    double kg[144];
    or_frame3d_kg(L, P, kg);
    rotate_blocks(12, R, kg, KGg);
    if (P_out) *P_out = P;
}

/* ------------------------------------------------------------------------ */
/* model-level COO assembly (element order: CQUAD4, CTRIA3, then CBAR)      */
/* ------------------------------------------------------------------------ */

/* COO triplets exactly as the reference emits them: 576 per Q4 column-major
 * over (rr, cc), then 324 per CTRIA3, then 144 per bar.  assembly.jl:3839-3851,
 * 3949-3953, 4016-4020.  Returns the number of triplets written.  Node
 * indices are 0-based ranks. */
int64_t or_assemble_K(int n_q4, const int32_t *q4, const double *xyz, const double *q4prop, double k6rot,
                      int n_t3, const int32_t *t3, const double *t3prop,
                      int n_bar, const int32_t *bar, const double *bar_vref, const double *barprop,
                      int64_t *I, int64_t *J, double *V) {
    int64_t cnt = 0;
    for (int e = 0; e < n_q4; e++) {
        double p[12];
        for (int k = 0; k < 4; k++) for (int i = 0; i < 3; i++) p[3 * k + i] = xyz[3 * q4[4 * e + k] + i];
        double Keg[576];
        or_q4_ke_global(p, q4prop + 6 * e, k6rot, Keg, 0, 0);
        for (int cc = 0; cc < 24; cc++) for (int rr = 0; rr < 24; rr++) {
            I[cnt] = 6 * (int64_t)q4[4 * e + rr / 6] + rr % 6;
            J[cnt] = 6 * (int64_t)q4[4 * e + cc / 6] + cc % 6;
            V[cnt] = Keg[rr * 24 + cc];
            cnt++;
        }
    }
    for (int e = 0; e < n_t3; e++) {
        double p[9];
        for (int k = 0; k < 3; k++) for (int i = 0; i < 3; i++) p[3 * k + i] = xyz[3 * t3[3 * e + k] + i];
        double Keg[324];
        or_t3_ke_global(p, t3prop + 6 * e, k6rot, Keg, 0, 0);
        for (int cc = 0; cc < 18; cc++) for (int rr = 0; rr < 18; rr++) {
            I[cnt] = 6 * (int64_t)t3[3 * e + rr / 6] + rr % 6;
            J[cnt] = 6 * (int64_t)t3[3 * e + cc / 6] + cc % 6;
            V[cnt] = Keg[rr * 18 + cc];
            cnt++;
        }
    }
    for (int b = 0; b < n_bar; b++) {
        const double *p1 = xyz + 3 * bar[2 * b], *p2 = xyz + 3 * bar[2 * b + 1];
        double Keg[144];
        or_bar_ke_global(p1, p2, bar_vref + 3 * b, barprop + 8 * b, Keg);
        for (int cc = 0; cc < 12; cc++) for (int rr = 0; rr < 12; rr++) {
            I[cnt] = 6 * (int64_t)bar[2 * b + rr / 6] + rr % 6;
            J[cnt] = 6 * (int64_t)bar[2 * b + cc / 6] + cc % 6;
            V[cnt] = Keg[rr * 12 + cc];
            cnt++;
        }
    }
    return cnt;
}

int64_t or_assemble_KG(int n_q4, const int32_t *q4, const double *xyz, const double *q4prop,
                       int n_t3, const int32_t *t3, const double *t3prop, int n_bar, const int32_t *bar, const double *bar_vref, const double *barprop,
                       const double *u, int64_t *I, int64_t *J, double *V) {
    int64_t cnt = 0;
    for (int e = 0; e < n_q4; e++) {
        double p[12], ug[24];
        for (int k = 0; k < 4; k++) {
            int nd = q4[4 * e + k];
            for (int i = 0; i < 3; i++) p[3 * k + i] = xyz[3 * nd + i];
            for (int i = 0; i < 6; i++) ug[6 * k + i] = u[6 * (int64_t)nd + i];
        }
        double KGg[576];
        or_q4_kg_global(p, q4prop + 6 * e, ug, KGg);
        for (int cc = 0; cc < 24; cc++) for (int rr = 0; rr < 24; rr++) {
            I[cnt] = 6 * (int64_t)q4[4 * e + rr / 6] + rr % 6;
            J[cnt] = 6 * (int64_t)q4[4 * e + cc / 6] + cc % 6;
            V[cnt] = KGg[rr * 24 + cc];
            cnt++;
        }
    }
    for (int e = 0; e < n_t3; e++) {
        double p[9], ug[18];
        for (int k = 0; k < 3; k++) {
            int nd = t3[3 * e + k];
            for (int i = 0; i < 3; i++) p[3 * k + i] = xyz[3 * nd + i];
            for (int i = 0; i < 6; i++) ug[6 * k + i] = u[6 * (int64_t)nd + i];
        }
        double KGg[324];
        or_t3_kg_global(p, t3prop + 6 * e, ug, KGg, 0);
        for (int cc = 0; cc < 18; cc++) for (int rr = 0; rr < 18; rr++) {
            I[cnt] = 6 * (int64_t)t3[3 * e + rr / 6] + rr % 6;
            J[cnt] = 6 * (int64_t)t3[3 * e + cc / 6] + cc % 6;
            V[cnt] = KGg[rr * 18 + cc];
            cnt++;
        }
    }
    for (int b = 0; b < n_bar; b++) {
        int na = bar[2 * b], nb = bar[2 * b + 1];
        double ug[12];
        for (int i = 0; i < 6; i++) { ug[i] = u[6 * (int64_t)na + i]; ug[6 + i] = u[6 * (int64_t)nb + i]; }
        double KGg[144];
        or_bar_kg_global(xyz + 3 * na, xyz + 3 * nb, bar_vref + 3 * b, barprop + 8 * b, ug, KGg, 0);
        for (int cc = 0; cc < 12; cc++) for (int rr = 0; rr < 12; rr++) {
            I[cnt] = 6 * (int64_t)bar[2 * b + rr / 6] + rr % 6;
            J[cnt] = 6 * (int64_t)bar[2 * b + cc / 6] + cc % 6;
            V[cnt] = KGg[rr * 12 + cc];
            cnt++;
        }
    }
    return cnt;
}
