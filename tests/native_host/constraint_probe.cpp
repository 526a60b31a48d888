// Host-only test probe: exposes the constraint elimination of the library
// (jf_2026_05_19_amd/csrc/constraints.cpp) to the CPU test-suite: the
// dependent-DOF map of a model description, the redistributed load, and K'
// of a dense element-assembled K through the block pattern and the
// redistribution plan (the device kernels' arithmetic emulated on the host).
#include <cstring>
#include <string>
#include <vector>

#include "../../jf_2026_05_19_amd/csrc/constraints.h"

extern "C" {

// map sizes: out[0] dependent DOFs, out[1] pairs; returns the status code
// (err receives the message)
int probe_map_sizes(const sol105_model_desc *d, int64_t *out, char *err, int err_len) {
    ConstraintMap C;
    std::string e;
    const int rc = build_constraint_map(d, C, e);
    if (err && err_len > 0) {
        std::strncpy(err, e.c_str(), (size_t)err_len - 1);
        err[err_len - 1] = 0;
    }
    out[0] = (int64_t)C.dep.size();
    out[1] = (int64_t)C.ind.size();
    return rc;
}

int probe_map(const sol105_model_desc *d, int32_t *dep, int32_t *ptr, int32_t *ind, double *coef, double *force_out) {
    ConstraintMap C;
    std::string e;
    const int rc = build_constraint_map(d, C, e);
    if (rc) return rc;
    std::copy(C.dep.begin(), C.dep.end(), dep);
    std::copy(C.ptr.begin(), C.ptr.end(), ptr);
    std::copy(C.ind.begin(), C.ind.end(), ind);
    std::copy(C.coef.begin(), C.coef.end(), coef);
    std::vector<double> F(d->force, d->force + 6 * (size_t)d->n_nodes);
    redistribute_force(C, F);
    std::copy(F.begin(), F.end(), force_out);
    return 0;
}

// K' (dense ndof x ndof, row-major) of a dense element-assembled K: entries
// gathered into the constraint-extended block pattern, the plan applied as
// the device does (targets in order, then the dependent entries cleared)
int probe_redistribute(const sol105_model_desc *d, const double *K, double *Kout) {
    ConstraintMap C;
    std::string e;
    int rc = build_constraint_map(d, C, e);
    if (rc) return rc;
    const int n = d->n_nodes;
    const int64_t nd = 6 * (int64_t)n;
    BlockPattern elem, P;
    build_block_pattern(n, d->n_q4, d->q4_nodes, d->n_t3, d->t3_nodes, d->n_bar, d->bar_nodes, elem);
    std::vector<std::pair<int32_t, int32_t>> extra;
    constraint_pattern_pairs(C, elem, n, extra);
    build_block_pattern(n, d->n_q4, d->q4_nodes, d->n_t3, d->t3_nodes, d->n_bar, d->bar_nodes, P, &extra);
    ConstraintPlan plan;
    build_constraint_plan(C, elem, P, n, plan);
    std::vector<double> vals((size_t)P.n_blocks() * 36, 0.0);
    for (int a = 0; a < n; a++)
        for (int32_t k = elem.row_ptr[a]; k < elem.row_ptr[a + 1]; k++) {
            const int b = elem.col[k];
            const int64_t blk = P.find(a, b);
            for (int r = 0; r < 6; r++)
                for (int c = 0; c < 6; c++) vals[blk * 36 + r * 6 + c] = K[(6 * a + r) * nd + 6 * b + c];
        }
    for (size_t t = 0; t < plan.tgt.size(); t++) {
        double acc = vals[plan.tgt[t]];
        for (int32_t p = plan.tptr[t]; p < plan.tptr[t + 1]; p++) acc += plan.coef[p] * vals[plan.src[p]];
        vals[plan.tgt[t]] = acc;
    }
    for (int64_t z : plan.zero) vals[z] = 0.0;
    std::memset(Kout, 0, sizeof(double) * nd * nd);
    for (int a = 0; a < n; a++)
        for (int32_t k = P.row_ptr[a]; k < P.row_ptr[a + 1]; k++) {
            const int b = P.col[k];
            for (int r = 0; r < 6; r++)
                for (int c = 0; c < 6; c++) Kout[(6 * a + r) * nd + 6 * b + c] = vals[(int64_t)k * 36 + r * 6 + c];
        }
    return 0;
}

}  // extern "C"
