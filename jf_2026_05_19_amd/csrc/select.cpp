// select.cpp -- output semantics of the buckling eigensolve (select.h).
#include "select.h"

#include <algorithm>
#include <cmath>

int ModeSelect::requested() const {
    int r = has_range() ? (int)std::ceil(3.0 * nd) : (dense() ? 3 * nd : nd);
    return std::min(r, std::max(n_free - 2, 1));
}

std::vector<int> select_candidates(const ModeSelect &s, const double *theta, int k, std::vector<double> &lam) {
    const bool dense = s.dense();
    std::vector<int> cand;
    lam.assign(k, 0.0);
    for (int i = 0; i < k; i++) {
        const double th = theta[i];
        lam[i] = 1.0 / th;
        const bool ok = dense ? (std::isfinite(th) && std::fabs(th) > 1e-12)
                              : (std::fabs(th) >= 1e-14 && std::fabs(lam[i]) > 1e-6);
        if (ok) cand.push_back(i);
    }
    std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return std::fabs(lam[a]) < std::fabs(lam[b]); });
    // the dense path keeps its whole spectrum for the range filter: every
    // pair computed here belongs to it
    const int keep = (dense && s.has_range()) ? (int)cand.size() : s.requested();
    if ((int)cand.size() > keep) cand.resize(keep);
    const double tol = 1e-10, lo = std::max(s.v1, tol);
    std::vector<int> pos;
    for (int i : cand)
        if (lam[i] > tol && (!s.has_range() || (lam[i] >= lo && lam[i] <= s.v2))) pos.push_back(i);
    std::stable_sort(pos.begin(), pos.end(), [&](int a, int b) { return lam[a] < lam[b]; });
    return pos;
}

int localization_scan_limit(const ModeSelect &s, int n_candidates) {
    if (s.n_shell_cards < 100) return 0;
    return std::min(n_candidates, std::max(s.nd + 16, s.nd));
}

int cluster_skip(const ModeSelect &s, const std::vector<int> &surv, const std::vector<double> &lam) {
    const int full_n = (int)surv.size(), dense_n = 3;
    if (!s.has_range() || s.v2 < 1.0 || s.n_shell_cards < 100 || full_n < dense_n + 2) return 0;
    for (int i : surv)
        if (!(lam[i] > 0.0)) return 0;
    int n_skip = 0;
    for (int i = 0; i < full_n - dense_n; i++) {
        const double jump = lam[surv[i + 1]] / lam[surv[i]];
        if (jump < (i == 0 ? 3.0 : 1.25) || jump > 8.0) continue;
        if (i >= 1 && lam[surv[i]] / lam[surv[0]] > 5.0) continue;
        double cmax = lam[surv[i + 1]], cmin = cmax;
        for (int j = i + 1; j <= i + dense_n; j++) {
            cmax = std::max(cmax, lam[surv[j]]);
            cmin = std::min(cmin, lam[surv[j]]);
        }
        if ((cmax - cmin) / cmin <= 0.30) n_skip = i + 1;
    }
    return n_skip;
}

std::vector<int> select_output(const ModeSelect &s, const std::vector<int> &surv, const std::vector<double> &lam) {
    const int cap = s.has_range() ? s.nd : s.requested();
    const int n_skip = cluster_skip(s, surv, lam);
    std::vector<int> out(surv.begin() + n_skip, surv.end());
    if ((int)out.size() > cap) out.resize(cap);
    return out;
}

bool selection_decided(const ModeSelect &s, const std::vector<int> &surv, const std::vector<double> &lam,
                       double lam_cover) {
    if (!s.has_range() || lam_cover > s.v2) return true;
    const int K = (int)surv.size();
    if (s.v2 >= 1.0 && s.n_shell_cards >= 100) {
        // the filter looks at jumps i with lam_i <= 5 lam_0 and their windows
        // i + 1 .. i + 3: all of them must lie among the known survivors
        int first_far = -1;
        for (int i = 1; i < K && first_far < 0; i++)
            if (lam[surv[i]] / lam[surv[0]] > 5.0) first_far = i;
        if (first_far < 0 || K < std::max(5, first_far + 3)) return false;
    }
    return K >= cluster_skip(s, surv, lam) + s.nd;
}
