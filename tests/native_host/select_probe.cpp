// Host-only test probe: exposes the buckling output selection of the library
// (jf_2026_05_19_amd/csrc/select.cpp) to the CPU test-suite, so the device
// path's candidate, range, cluster and decidedness rules can be checked
// against the oracle's restatement on hand-built spectra without a GPU.
#include <cstddef>
#include <vector>

#include "../../jf_2026_05_19_amd/csrc/select.h"

extern "C" {

// theta: k Ritz values of C = L^-1 (-KG) L^-T (any order).  share[i] > 0.10
// drops pair i when it lies within the localization scan.  out_idx receives the selected pair indices;
// returns their count.  info: [in-range candidates, scan limit, cluster skip,
// decided].
int probe_select(int nd, int n_free, double v1, double v2, int n_shell_cards, int all_pairs, const double *theta,
                 int k, const double *share, double lam_cover, int *out_idx, int *info) {
    ModeSelect s;
    s.nd = nd;
    s.n_free = n_free;
    s.v1 = v1;
    s.v2 = v2;
    s.n_shell_cards = n_shell_cards;
    s.all_pairs = all_pairs != 0;
    std::vector<double> lam;
    std::vector<int> pos = select_candidates(s, theta, k, lam);
    const int limit = localization_scan_limit(s, (int)pos.size());
    std::vector<int> surv;
    for (int t = 0; t < (int)pos.size(); t++)
        if (t >= limit || !(share[pos[t]] > 0.10)) surv.push_back(pos[t]);
    std::vector<int> out = select_output(s, surv, lam);
    info[0] = (int)pos.size();
    info[1] = limit;
    info[2] = cluster_skip(s, surv, lam);
    info[3] = selection_decided(s, surv, lam, lam_cover) ? 1 : 0;
    for (std::size_t t = 0; t < out.size(); t++) out_idx[t] = out[t];
    return (int)out.size();
}

int probe_requested(int nd, int n_free, double v1, double v2) {
    ModeSelect s;
    s.nd = nd;
    s.n_free = n_free;
    s.v1 = v1;
    s.v2 = v2;
    return s.requested();
}

}  // extern "C"
