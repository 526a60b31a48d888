// select.h -- which Ritz pairs of the buckling eigensolve become results:
// the output semantics of JFEM/src/solver/solve_case.jl (candidate floors
// :2244-2252 and :2174-2179, positive filter :2414, EIGRL range :2422-2436,
// mode count :1938-1945 and :2966-2987, cluster filter :2995-3139).  Host
// code only; the localization filter between the two steps needs the mode
// shapes and runs on the device.
#pragma once
#include <vector>

struct ModeSelect {
    int nd = 3;               // EIGRL ND
    int n_free = 0;
    double v1 = 0.0, v2 = 0.0;  // EIGRL V1, V2
    int n_shell_cards = 0;    // CSHELL cards of the model (the filters' >= 100 guard)
    bool all_pairs = false;   // the eigensolve returned every pair (dense path of a small model)

    bool dense() const { return n_free <= 4000; }
    bool has_range() const { return (v1 != 0.0 || v2 != 0.0) && v2 > v1; }
    // pairs the eigensolve is asked for
    int requested() const;
};

// Candidates in output order: lambda = 1/theta with the path's floors,
// the requested count by |lambda| (all of them for a dense path with a
// range), positive, inside [max(V1, 1e-10), V2] with a range, ascending.
// lam receives 1/theta for every pair.
std::vector<int> select_candidates(const ModeSelect &s, const double *theta, int k, std::vector<double> &lam);

// How many leading candidates the localization filter looks at.
int localization_scan_limit(const ModeSelect &s, int n_candidates);

// Leading survivors the cluster filter skips (a range with V2 >= 1, >= 100
// shell cards; else 0): the last spectral jump (>= 3.0 after a single mode,
// >= 1.25 after more, <= 8.0) that starts three modes within 30 %, the modes
// before it within a factor 5.
int cluster_skip(const ModeSelect &s, const std::vector<int> &surv, const std::vector<double> &lam);

// Final pairs from the candidates that survived the localization filter:
// the mode count (ND with a range, else the requested count) after the
// cluster filter.
std::vector<int> select_output(const ModeSelect &s, const std::vector<int> &surv, const std::vector<double> &lam);

// Whether the output from these survivors is the one the complete spectrum
// gives, when every eigenpair with |lambda| < lam_cover was computed (the
// others can only add positive eigenvalues >= lam_cover, after every known
// one in ascending order).  Always true without a range.
bool selection_decided(const ModeSelect &s, const std::vector<int> &surv, const std::vector<double> &lam,
                       double lam_cover);
