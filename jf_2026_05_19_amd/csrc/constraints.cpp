// constraints.cpp -- RBE1 / RBE2 / RBE3 / RSPLINE / MPC constraint map, pattern pairs and the
// deterministic redistribution plan (host code).
//
// Restates JFEM/src/solver/constraints.jl:37-364 for RBE2 (and RBAR as
// RBE2) and RBE3 with basic-frame nodes (node_R = I):
//  * RBE2: slave component c of a master-slave pair depends on the master's
//    six DOFs through the rigid row [e_c, (O (p_s - p_m))_c] for c <= 3
//    (u_s = u_m + theta_m x (p_s - p_m)) and [0, e_{c-3}] for rotations;
//  * RBE3: with G_i (n_comp x 6) the rigid rows of grid i's components about
//    the reference point, A6 = sum_i w_i G_i' G_i, the reference DOF r takes
//    coefficients (pinv(A6) w_i G_i')[r, :] (pinv with relative cut 1e-10);
//  * coefficients with |c| <= 1e-15 are dropped; RBE2 entries replace RBE3
//    entries of the same DOF (the reference's merge! order).
// Element stiffness entries in a dependent row or column then move to the
// independent DOFs (constraints.jl:318-360): K'[a][j] += c_ia K[i][j],
// K'[i][b] += K[i][j] c_jb, K'[a][b] += c_ia c_jb K[i][j]; KG takes the same
// redistribution (assembly.jl:6617-6620).
#include "constraints.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <set>

namespace {

std::vector<int> digits(int32_t code) {
    std::vector<int> d;
    char buf[16];
    std::snprintf(buf, sizeof buf, "%d", code < 0 ? -code : code);
    for (const char *p = buf; *p; p++)
        if (*p >= '1' && *p <= '6') d.push_back(*p - '0');
    return d;
}

// rigid row of component c (1..6) of a point at offset (dx, dy, dz) from the
// reference point, on the reference's six DOFs
void rigid_row(double dx, double dy, double dz, int c, double out[6]) {
    for (int k = 0; k < 6; k++) out[k] = 0.0;
    if (c <= 3) {
        const double O[3][3] = {{0.0, dz, -dy}, {-dz, 0.0, dx}, {dy, -dx, 0.0}};
        out[c - 1] = 1.0;
        for (int k = 0; k < 3; k++) out[3 + k] = O[c - 1][k];
    } else {
        out[c - 1] = 1.0;
    }
}

// pseudo-inverse of a symmetric positive semi-definite 6x6 matrix (cyclic
// Jacobi; eigenvalues at or below rcut times the largest are dropped)
void pinv6(const double A[36], double P[36], double rcut = 1e-10) {
    double a[6][6], v[6][6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) { a[i][j] = 0.5 * (A[i * 6 + j] + A[j * 6 + i]); v[i][j] = i == j ? 1.0 : 0.0; }
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, dia = 0.0;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) (i == j ? dia : off) += a[i][j] * a[i][j];
        if (off <= 1e-34 * dia || off == 0.0) break;
        for (int p = 0; p < 5; p++)
            for (int q = p + 1; q < 6; q++) {
                if (a[p][q] == 0.0) continue;
                double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 6; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 6; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 6; k++) {
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double lmax = 0.0;
    for (int i = 0; i < 6; i++) lmax = std::max(lmax, std::fabs(a[i][i]));
    const double cut = rcut * lmax;
    for (int i = 0; i < 36; i++) P[i] = 0.0;
    for (int k = 0; k < 6; k++) {
        double l = a[k][k];
        if (!(std::fabs(l) > cut) || l == 0.0) continue;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) P[i * 6 + j] += v[i][k] * v[j][k] / l;
    }
}

}  // namespace

int build_constraint_map(const sol105_model_desc *d, ConstraintMap &C, std::string &err) {
    C = ConstraintMap();
    const int n = d->n_nodes;
    std::map<int32_t, std::vector<std::pair<int32_t, double>>> rbe3m, rbe2m;
    auto P = [&](int i, int k) { return d->xyz[3 * (size_t)i + k]; };
    char buf[160];
    if (d->n_rbe3 > 0 && (!d->rbe3 || !d->rbe3_grp || !d->rbe3_wt || !d->rbe3_node)) {
        err = "missing RBE3 arrays";
        return SOL105_E_INVALID;
    }
    if (d->n_rbe2 > 0 && !d->rbe2) {
        err = "missing RBE2 arrays";
        return SOL105_E_INVALID;
    }
    for (int e = 0; e < d->n_rbe3; e++) {
        const int32_t *r = d->rbe3 + 4 * (size_t)e;
        const int ref = r[0], g0 = r[2], ng = r[3];
        if (ref < 0 || ref >= n || g0 < 0 || ng < 0 || g0 + ng > d->n_rbe3_grp) {
            std::snprintf(buf, sizeof buf, "RBE3 %d: reference rank or groups out of range", e);
            err = buf;
            return SOL105_E_INVALID;
        }
        const std::vector<int> refc = [&] { auto v = digits(r[1]); std::sort(v.begin(), v.end()); return v; }();
        if (refc.empty()) continue;
        double A6[36] = {0};
        struct GridRows { int node; std::vector<int> comps; std::vector<double> G; double wt; };
        std::vector<GridRows> rows;
        for (int g = g0; g < g0 + ng; g++) {
            const int32_t *gr = d->rbe3_grp + 3 * (size_t)g;
            const double wt = d->rbe3_wt[g];
            std::vector<int> comps = digits(gr[0]);
            std::sort(comps.begin(), comps.end());
            if (comps.empty()) continue;
            if (gr[1] < 0 || gr[2] < 0 || gr[1] + gr[2] > d->n_rbe3_node) {
                std::snprintf(buf, sizeof buf, "RBE3 %d: group node list out of range", e);
                err = buf;
                return SOL105_E_INVALID;
            }
            for (int k = gr[1]; k < gr[1] + gr[2]; k++) {
                const int di = d->rbe3_node[k];
                if (di < 0 || di >= n) {
                    std::snprintf(buf, sizeof buf, "RBE3 %d: grid rank %d out of range", e, di);
                    err = buf;
                    return SOL105_E_INVALID;
                }
                GridRows R{di, comps, std::vector<double>(comps.size() * 6), wt};
                for (size_t jj = 0; jj < comps.size(); jj++)
                    rigid_row(P(di, 0) - P(ref, 0), P(di, 1) - P(ref, 1), P(di, 2) - P(ref, 2), comps[jj],
                              &R.G[jj * 6]);
                for (int a = 0; a < 6; a++)
                    for (int b = 0; b < 6; b++) {
                        double s = 0.0;
                        for (size_t jj = 0; jj < comps.size(); jj++) s += R.G[jj * 6 + a] * R.G[jj * 6 + b];
                        A6[a * 6 + b] += wt * s;
                    }
                rows.push_back(std::move(R));
            }
        }
        if (rows.empty()) continue;
        double Ai[36];
        pinv6(A6, Ai);
        for (const GridRows &R : rows)
            for (int rd : refc) {
                const int32_t ref_dof = 6 * ref + rd - 1;
                auto it = rbe3m.find(ref_dof);
                if (it == rbe3m.end()) it = rbe3m.emplace(ref_dof, std::vector<std::pair<int32_t, double>>()).first;
                else if (&R == &rows.front()) {
                    std::snprintf(buf, sizeof buf, "RBE3 %d: reference DOF %d is already dependent", e, ref_dof);
                    err = buf;
                    return SOL105_E_UNSUPPORTED;
                }
                for (size_t jj = 0; jj < R.comps.size(); jj++) {
                    double c = 0.0;
                    for (int k = 0; k < 6; k++) c += Ai[(rd - 1) * 6 + k] * (R.wt * R.G[jj * 6 + k]);
                    if (std::fabs(c) > 1e-15) it->second.emplace_back(6 * R.node + R.comps[jj] - 1, c);
                }
            }
    }
    for (int e = 0; e < d->n_rbe2; e++) {
        const int32_t *r = d->rbe2 + 3 * (size_t)e;
        const int m = r[0], sl = r[2];
        if (m < 0 || m >= n || sl < 0 || sl >= n) {
            std::snprintf(buf, sizeof buf, "RBE2 pair %d: node rank out of range", e);
            err = buf;
            return SOL105_E_INVALID;
        }
        for (int c : digits(r[1])) {
            double row[6];
            rigid_row(P(sl, 0) - P(m, 0), P(sl, 1) - P(m, 1), P(sl, 2) - P(m, 2), c, row);
            std::vector<std::pair<int32_t, double>> pairs;
            for (int k = 0; k < 6; k++)
                if (std::fabs(row[k]) > 1e-15) pairs.emplace_back(6 * m + k, row[k]);
            if (pairs.empty()) continue;
            const int32_t sd = 6 * sl + c - 1;
            if (rbe2m.count(sd)) {
                std::snprintf(buf, sizeof buf, "RBE2: DOF %d is the slave of two constraints", sd);
                err = buf;
                return SOL105_E_UNSUPPORTED;
            }
            rbe2m.emplace(sd, std::move(pairs));
        }
    }
    for (auto &kv : rbe2m) rbe3m[kv.first] = kv.second;
    // RBE1 (constraints.jl:76-143): the rigid motion q about the centroid of
    // the independent GRIDs with A q = u_indep (A's rows the rigid rows of
    // the independent DOFs), q = pinv(A) u_indep; a dependent DOF with rigid
    // row b takes pinv(A)' b.  pinv(A) = pinv(A'A) A', singular values below
    // 1e-10 of the largest dropped (eigenvalues of A'A below 1e-20).
    if (d->n_rbe1 > 0 && (!d->rbe1 || !d->rbe1_indep || !d->rbe1_dep)) {
        err = "missing RBE1 arrays";
        return SOL105_E_INVALID;
    }
    std::map<int32_t, std::vector<std::pair<int32_t, double>>> rbe1m;
    for (int e = 0; e < d->n_rbe1; e++) {
        const int32_t *r = d->rbe1 + 4 * (size_t)e;
        const int i0 = r[0], ni = r[1], d0 = r[2], ndp = r[3];
        bool ok = i0 >= 0 && ni >= 1 && ni <= 6 && i0 + ni <= d->n_rbe1_indep && d0 >= 0 && ndp >= 0 &&
                  d0 + ndp <= d->n_rbe1_dep;
        for (int k = 0; ok && k < ni + ndp; k++) {
            const int32_t *gc = k < ni ? d->rbe1_indep + 2 * (size_t)(i0 + k) : d->rbe1_dep + 2 * (size_t)(d0 + k - ni);
            ok = gc[0] >= 0 && gc[0] < n && gc[1] >= 1 && gc[1] <= 6;
        }
        if (!ok) {
            std::snprintf(buf, sizeof buf, "RBE1 %d: ranks, components or counts out of range", e);
            err = buf;
            return SOL105_E_INVALID;
        }
        const int32_t *ind = d->rbe1_indep + 2 * (size_t)i0;
        std::vector<int> grids;
        for (int k = 0; k < ni; k++)
            if (std::find(grids.begin(), grids.end(), ind[2 * k]) == grids.end()) grids.push_back(ind[2 * k]);
        double xr[3] = {0.0, 0.0, 0.0};
        for (int g : grids)
            for (int a = 0; a < 3; a++) xr[a] += P(g, a);
        for (int a = 0; a < 3; a++) xr[a] /= (double)grids.size();
        std::vector<double> A((size_t)ni * 6);
        for (int k = 0; k < ni; k++)
            rigid_row(P(ind[2 * k], 0) - xr[0], P(ind[2 * k], 1) - xr[1], P(ind[2 * k], 2) - xr[2], ind[2 * k + 1],
                      &A[(size_t)k * 6]);
        double AtA[36] = {0}, Pi[36];
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++)
                for (int k = 0; k < ni; k++) AtA[a * 6 + b] += A[(size_t)k * 6 + a] * A[(size_t)k * 6 + b];
        pinv6(AtA, Pi, 1e-20);
        for (int k = 0; k < ndp; k++) {
            const int g = d->rbe1_dep[2 * (size_t)(d0 + k)], c = d->rbe1_dep[2 * (size_t)(d0 + k) + 1];
            double b[6], w[6];
            rigid_row(P(g, 0) - xr[0], P(g, 1) - xr[1], P(g, 2) - xr[2], c, b);
            for (int a = 0; a < 6; a++) {
                w[a] = 0.0;
                for (int q = 0; q < 6; q++) w[a] += Pi[a * 6 + q] * b[q];
            }
            std::vector<std::pair<int32_t, double>> pairs;
            for (int j = 0; j < ni; j++) {
                double cj = 0.0;
                for (int a = 0; a < 6; a++) cj += A[(size_t)j * 6 + a] * w[a];
                if (std::fabs(cj) > 1e-15) pairs.emplace_back(6 * ind[2 * j] + ind[2 * j + 1] - 1, cj);
            }
            if (pairs.empty()) continue;
            const int32_t dd = 6 * g + c - 1;
            if (rbe1m.count(dd)) {
                std::snprintf(buf, sizeof buf, "RBE1: DOF %d is dependent in two elements", dd);
                err = buf;
                return SOL105_E_UNSUPPORTED;
            }
            rbe1m.emplace(dd, std::move(pairs));
        }
    }
    for (auto &kv : rbe1m) rbe3m[kv.first] = kv.second;
    // RSPLINE (constraints.jl:146-213): a dependent point is projected on the
    // polyline of the independent GRIDs (the first nearest segment, t
    // clamped to [0, 1]) and follows that segment's ends with 1 - t and t
    if (d->n_rspline > 0 && (!d->rspline || !d->rspline_indep || !d->rspline_dep)) {
        err = "missing RSPLINE arrays";
        return SOL105_E_INVALID;
    }
    std::map<int32_t, std::vector<std::pair<int32_t, double>>> rsplm;
    for (int e = 0; e < d->n_rspline; e++) {
        const int32_t *r = d->rspline + 4 * (size_t)e;
        const int i0 = r[0], ni = r[1], d0 = r[2], ndp = r[3];
        bool ok = i0 >= 0 && ni >= 0 && i0 + ni <= d->n_rspline_indep && d0 >= 0 && ndp >= 0 &&
                  d0 + ndp <= d->n_rspline_dep;
        for (int k = 0; ok && k < ni; k++) ok = d->rspline_indep[i0 + k] >= 0 && d->rspline_indep[i0 + k] < n;
        for (int k = 0; ok && k < ndp; k++) {
            const int g = d->rspline_dep[2 * (size_t)(d0 + k)], c = d->rspline_dep[2 * (size_t)(d0 + k) + 1];
            ok = g >= 0 && g < n && c >= 1 && c <= 6;
        }
        if (!ok) {
            std::snprintf(buf, sizeof buf, "RSPLINE %d: ranks or components out of range", e);
            err = buf;
            return SOL105_E_INVALID;
        }
        if (ni < 2) continue;
        const int32_t *ind = d->rspline_indep + i0;
        double total = 0.0;
        for (int s = 1; s < ni; s++) {
            const double dx = P(ind[s], 0) - P(ind[s - 1], 0), dy = P(ind[s], 1) - P(ind[s - 1], 1),
                         dz = P(ind[s], 2) - P(ind[s - 1], 2);
            total += std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        if (total < 1e-30) continue;
        for (int k = 0; k < ndp; k++) {
            const int g = d->rspline_dep[2 * (size_t)(d0 + k)], c = d->rspline_dep[2 * (size_t)(d0 + k) + 1];
            const double px = P(g, 0), py = P(g, 1), pz = P(g, 2);
            int best = 0;
            double bt = 0.0, bd = INFINITY;
            for (int s = 0; s + 1 < ni; s++) {
                const double ax = P(ind[s], 0), ay = P(ind[s], 1), az = P(ind[s], 2);
                const double dx = P(ind[s + 1], 0) - ax, dy = P(ind[s + 1], 1) - ay, dz = P(ind[s + 1], 2) - az;
                const double l2 = dx * dx + dy * dy + dz * dz;
                if (l2 < 1e-30) continue;
                const double t = std::min(std::max(((px - ax) * dx + (py - ay) * dy + (pz - az) * dz) / l2, 0.0), 1.0);
                const double cx = ax + t * dx, cy = ay + t * dy, cz = az + t * dz;
                const double dist = std::sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy) + (pz - cz) * (pz - cz));
                if (dist < bd) {
                    best = s;
                    bt = t;
                    bd = dist;
                }
            }
            std::vector<std::pair<int32_t, double>> pairs;
            if (std::fabs(1.0 - bt) > 1e-15) pairs.emplace_back(6 * ind[best] + c - 1, 1.0 - bt);
            if (std::fabs(bt) > 1e-15) pairs.emplace_back(6 * ind[best + 1] + c - 1, bt);
            if (pairs.empty()) continue;
            const int32_t dd = 6 * g + c - 1;
            if (rsplm.count(dd)) {
                std::snprintf(buf, sizeof buf, "RSPLINE: DOF %d is dependent in two elements", dd);
                err = buf;
                return SOL105_E_UNSUPPORTED;
            }
            rsplm.emplace(dd, std::move(pairs));
        }
    }
    for (auto &kv : rsplm) rbe3m[kv.first] = kv.second;
    // explicit MPC equations (constraints.jl:290-316): u_dep = sum -A_i/A_1
    // u_i; they replace RBE2 / RBE3 entries, a later equation an earlier one
    if (d->n_mpc > 0 && (!d->mpc_eq || !d->mpc_term || !d->mpc_coef)) {
        err = "missing MPC arrays";
        return SOL105_E_INVALID;
    }
    for (int e = 0; e < d->n_mpc; e++) {
        const int t0 = d->mpc_eq[2 * (size_t)e], nt = d->mpc_eq[2 * (size_t)e + 1];
        bool ok = t0 >= 0 && nt >= 2 && t0 + nt <= d->n_mpc_term;
        for (int t = t0; ok && t < t0 + nt; t++) {
            const int g = d->mpc_term[2 * (size_t)t], c = d->mpc_term[2 * (size_t)t + 1];
            ok = g >= 0 && g < n && c >= 1 && c <= 6;
        }
        if (!ok || !(std::fabs(d->mpc_coef[t0]) >= 1e-30)) {
            std::snprintf(buf, sizeof buf, "MPC %d: terms out of range or a zero dependent coefficient", e);
            err = buf;
            return SOL105_E_INVALID;
        }
        const double a0 = d->mpc_coef[t0];
        std::vector<std::pair<int32_t, double>> pairs;
        for (int t = t0 + 1; t < t0 + nt; t++)
            pairs.emplace_back(6 * d->mpc_term[2 * (size_t)t] + d->mpc_term[2 * (size_t)t + 1] - 1,
                               -d->mpc_coef[t] / a0);
        rbe3m[6 * d->mpc_term[2 * (size_t)t0] + d->mpc_term[2 * (size_t)t0 + 1] - 1] = std::move(pairs);
    }
    C.ptr.push_back(0);
    for (auto &kv : rbe3m) {
        C.dep.push_back(kv.first);
        for (auto &pc : kv.second) {
            C.ind.push_back(pc.first);
            C.coef.push_back(pc.second);
        }
        C.ptr.push_back((int32_t)C.ind.size());
    }
    for (int32_t v : C.ind)
        if (rbe3m.count(v)) {
            std::snprintf(buf, sizeof buf, "constraint chain: independent DOF %d is itself dependent", v);
            err = buf;
            return SOL105_E_UNSUPPORTED;
        }
    return SOL105_OK;
}

void constraint_pattern_pairs(const ConstraintMap &C, const BlockPattern &elem, int n_nodes,
                              std::vector<std::pair<int32_t, int32_t>> &pairs) {
    pairs.clear();
    if (C.empty()) return;
    std::vector<std::vector<int32_t>> masters(n_nodes);
    for (size_t i = 0; i < C.dep.size(); i++)
        for (int p = C.ptr[i]; p < C.ptr[i + 1]; p++) masters[C.dep[i] / 6].push_back(C.ind[p] / 6);
    for (auto &m : masters) {
        std::sort(m.begin(), m.end());
        m.erase(std::unique(m.begin(), m.end()), m.end());
    }
    for (int a = 0; a < n_nodes; a++)
        for (int32_t k = elem.row_ptr[a]; k < elem.row_ptr[a + 1]; k++) {
            const int b = elem.col[k];
            if (elem.cptr[k + 1] == elem.cptr[k]) continue;  // no element contribution
            if (masters[a].empty() && masters[b].empty()) continue;
            std::vector<int32_t> A = masters[a], B = masters[b];
            A.push_back(a);
            B.push_back(b);
            for (int32_t x : A)
                for (int32_t y : B) pairs.emplace_back(x, y);
        }
}

void build_constraint_plan(const ConstraintMap &C, const BlockPattern &elem, const BlockPattern &P, int n_nodes,
                           ConstraintPlan &out) {
    out = ConstraintPlan();
    if (C.empty()) return;
    std::vector<int32_t> slot(6 * (size_t)n_nodes, -1);
    for (size_t i = 0; i < C.dep.size(); i++) slot[C.dep[i]] = (int32_t)i;
    auto pos = [&](int i, int j) {
        int64_t b = P.find(i / 6, j / 6);
        return b * 36 + (int64_t)(i % 6) * 6 + (j % 6);
    };
    struct Item { int64_t tgt, src; double coef; };
    std::vector<Item> items;
    std::vector<uint8_t> node_dep(n_nodes, 0);
    for (int32_t dd : C.dep) node_dep[dd / 6] = 1;
    for (int a = 0; a < n_nodes; a++)
        for (int32_t k = elem.row_ptr[a]; k < elem.row_ptr[a + 1]; k++) {
            const int b = elem.col[k];
            if (elem.cptr[k + 1] == elem.cptr[k] || (!node_dep[a] && !node_dep[b])) continue;
            for (int r = 0; r < 6; r++)
                for (int c = 0; c < 6; c++) {
                    const int i = 6 * a + r, j = 6 * b + c;
                    const int si = slot[i], sj = slot[j];
                    if (si < 0 && sj < 0) continue;
                    const int64_t src = pos(i, j);
                    if (si >= 0 && sj < 0) {
                        for (int p = C.ptr[si]; p < C.ptr[si + 1]; p++) items.push_back({pos(C.ind[p], j), src, C.coef[p]});
                    } else if (si < 0) {
                        for (int p = C.ptr[sj]; p < C.ptr[sj + 1]; p++) items.push_back({pos(i, C.ind[p]), src, C.coef[p]});
                    } else {
                        for (int p = C.ptr[si]; p < C.ptr[si + 1]; p++)
                            for (int q = C.ptr[sj]; q < C.ptr[sj + 1]; q++)
                                items.push_back({pos(C.ind[p], C.ind[q]), src, C.coef[p] * C.coef[q]});
                    }
                }
        }
    std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) { return x.tgt < y.tgt; });
    out.tptr.push_back(0);
    for (size_t t = 0; t < items.size(); t++) {
        if (t == 0 || items[t].tgt != items[t - 1].tgt) {
            if (t > 0) out.tptr.push_back((int32_t)t);
            out.tgt.push_back(items[t].tgt);
        }
        out.src.push_back(items[t].src);
        out.coef.push_back(items[t].coef);
    }
    if (!items.empty()) out.tptr.push_back((int32_t)items.size());
    for (int a = 0; a < n_nodes; a++)
        for (int32_t k = P.row_ptr[a]; k < P.row_ptr[a + 1]; k++) {
            const int b = P.col[k];
            if (!node_dep[a] && !node_dep[b]) continue;
            for (int r = 0; r < 6; r++)
                for (int c = 0; c < 6; c++)
                    if (slot[6 * a + r] >= 0 || slot[6 * b + c] >= 0) out.zero.push_back((int64_t)k * 36 + r * 6 + c);
        }
}

void build_constraint_fold(const ConstraintMap &C, ConstraintFold &out) {
    std::map<int32_t, std::vector<std::pair<int32_t, double>>> by_ind;
    for (size_t i = 0; i < C.dep.size(); i++)
        for (int32_t p = C.ptr[i]; p < C.ptr[i + 1]; p++) by_ind[C.ind[p]].push_back({C.dep[i], C.coef[p]});
    out = ConstraintFold();
    out.ptr.push_back(0);
    for (const auto &kv : by_ind) {
        out.recv.push_back(kv.first);
        for (const auto &t : kv.second) {
            out.dep.push_back(t.first);
            out.coef.push_back(t.second);
        }
        out.ptr.push_back((int32_t)out.dep.size());
    }
}

void redistribute_force(const ConstraintMap &C, std::vector<double> &F) {
    for (size_t i = 0; i < C.dep.size(); i++) {
        const double f = F[C.dep[i]];
        if (!(std::fabs(f) > 1e-30)) continue;
        for (int p = C.ptr[i]; p < C.ptr[i + 1]; p++) F[C.ind[p]] += f * C.coef[p];
        F[C.dep[i]] = 0.0;
    }
}
