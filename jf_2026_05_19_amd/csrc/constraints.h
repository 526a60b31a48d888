// constraints.h -- rigid (RBE1, RBE2), interpolation (RBE3, RSPLINE) and
// explicit MPC constraints as dependent-DOF elimination: the constraint
// map, the extra block-pattern pairs it needs and the entry-level
// redistribution plan of K.  Host code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sol105.h"
#include "symbolic.h"

// u[dep[i]] = sum_p coef[p] u[ind[p]], p in [ptr[i], ptr[i+1]); dep ascending
struct ConstraintMap {
    std::vector<int32_t> dep, ptr, ind;
    std::vector<double> coef;
    bool empty() const { return dep.empty(); }
};

// Builds the map of a model description.  Returns SOL105_OK, or an error
// code with the reason in err: SOL105_E_INVALID for out-of-range ranks or
// groups, SOL105_E_UNSUPPORTED for a dependent DOF claimed twice by one
// RBE1 / RBE2 / RBE3 / RSPLINE family or an independent DOF that is itself
// dependent (the reference resolves neither deterministically).  Families
// merge in the order RBE3, RBE2, RBE1, RSPLINE, MPC, each replacing the
// entries before it; among MPC equations the later card wins.
int build_constraint_map(const sol105_model_desc *d, ConstraintMap &C, std::string &err);

// Node pairs (i, j) to add to the element block pattern so that the
// redistributed entries have a block: for every element pair (a, b) the
// products of {a} + masters(a) and {b} + masters(b).
void constraint_pattern_pairs(const ConstraintMap &C, const BlockPattern &elem, int n_nodes,
                              std::vector<std::pair<int32_t, int32_t>> &pairs);

// Entry-level plan on the extended pattern P (entry position = block * 36 +
// 6 * (row DOF % 6) + col DOF % 6): K'[tgt[t]] = K[tgt[t]] + sum_p coef[p]
// K[src[p]] over p in [tptr[t], tptr[t+1]) for entries of the element
// pattern with a dependent row or column, then the dependent rows and
// columns (zero) are cleared.
struct ConstraintPlan {
    std::vector<int64_t> tgt, src, zero;
    std::vector<int32_t> tptr;
    std::vector<double> coef;
};
void build_constraint_plan(const ConstraintMap &C, const BlockPattern &elem, const BlockPattern &P, int n_nodes,
                           ConstraintPlan &out);

// The map transposed for folding a vector onto the independent DOFs:
// v[recv[k]] += sum_p coef[p] v[dep[p]], p in [ptr[k], ptr[k+1]); recv
// ascending, the terms of one receiver in ascending dependent order.
struct ConstraintFold {
    std::vector<int32_t> recv, ptr, dep;
    std::vector<double> coef;
};
void build_constraint_fold(const ConstraintMap &C, ConstraintFold &out);

// F'[ind] += coef F[dep], F'[dep] = 0 (the load redistribution of
// _assemble_applied_force, JFEM/src/solver/solve_case.jl:40-55)
void redistribute_force(const ConstraintMap &C, std::vector<double> &F);
