// symbolic.h -- host-side analysis: block pattern, element->block gather map,
// nested-dissection ordering and the supernodal (multifrontal) structure.
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

struct BlockPattern {
    int n_nodes = 0;
    std::vector<int32_t> row_ptr;  // n_nodes + 1
    std::vector<int32_t> col;      // n_blocks (sorted per row)
    std::vector<int32_t> cptr;     // n_blocks + 1: element-block contributions
    std::vector<int32_t> cidx;     // element-block index, in element order
    int64_t n_blocks() const { return (int64_t)col.size(); }
    int64_t find(int i, int j) const;  // block index of (i, j) or -1
};

// Build the pattern and the deterministic gather map: CQUAD4 blocks first
// (element order, a-major), then CTRIA3 blocks, then CBAR blocks.
// extra: further node pairs to include (constraint couplings), no element
// contributions.
void build_block_pattern(int n_nodes, int n_q4, const int32_t *q4, int n_t3, const int32_t *t3, int n_bar,
                         const int32_t *bar, BlockPattern &P,
                         const std::vector<std::pair<int32_t, int32_t>> *extra = nullptr);

struct FrontInfo {
    int ns = 0;          // pivot rows (contiguous elimination indices piv0..piv0+ns-1)
    int nu = 0;          // update rows
    int m = 0;           // ns + nu
    int piv0 = 0;
    int64_t off = 0;     // column-major m x m storage offset (doubles)
    int64_t row_off = 0; // offset of the m row indices
    int64_t vec_off = 0; // offset of the m-vector workspace (per right-hand side)
    int parent = -1;
    int height = 0;
    int64_t rel_off = 0; // offset of nu positions in the parent's rows
};

struct Symbolic {
    int n_free = 0;
    std::vector<int32_t> elim_of_dof;   // ndof: elimination index or -1 (fixed)
    std::vector<int32_t> dof_of_elim;   // n_free
    std::vector<FrontInfo> fronts;      // postorder
    std::vector<int32_t> rows;          // concatenated front rows (elimination indices)
    std::vector<int32_t> rel;           // concatenated child->parent positions
    std::vector<std::vector<int32_t>> children;
    int height = 0;
    std::vector<std::vector<int32_t>> by_height;  // front ids per height
    int64_t front_storage = 0;          // doubles
    int64_t vec_storage = 0;            // doubles per right-hand side
    int64_t nnz_L = 0;
    double flops = 0.0;
    // K -> front scatter (lower triangle of the permuted free matrix)
    std::vector<int32_t> k_src;         // index into BSR values
    std::vector<int64_t> k_dst;         // index into front storage
    int max_ns = 0, max_m = 0, max_nu = 0;
};

// Nested dissection on the node graph (geometric bisection with a one-layer
// vertex separator), supernodes = separators / leaf subdomains, postorder.
void symbolic_analyze(const BlockPattern &P, const double *xyz, const uint8_t *free_mask, int leaf_nodes,
                      Symbolic &S);
