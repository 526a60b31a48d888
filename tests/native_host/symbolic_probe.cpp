// Host-only test probe: exposes the symbolic analysis of the library
// (jf_2026_05_19_amd/csrc/symbolic.cpp) to the CPU test-suite so the
// multifrontal maps can be checked against a dense emulation without a GPU.
#include <cstring>

#include "../../jf_2026_05_19_amd/csrc/symbolic.h"

struct Probe {
    BlockPattern P;
    Symbolic S;
};

extern "C" {

void *probe_analyze(int n_nodes, int n_q4, const int32_t *q4, int n_bar, const int32_t *bar, const double *xyz,
                    const uint8_t *free_mask, int leaf) {
    Probe *p = new Probe();
    build_block_pattern(n_nodes, n_q4, q4, n_bar, bar, p->P);
    symbolic_analyze(p->P, xyz, free_mask, leaf, p->S);
    return p;
}

void probe_free(void *h) { delete (Probe *)h; }

// sizes: [n_blocks, n_contrib, n_free, n_fronts, n_rows, n_rel, n_kmap, front_storage, height]
void probe_sizes(void *h, int64_t *out) {
    Probe *p = (Probe *)h;
    out[0] = p->P.n_blocks();
    out[1] = (int64_t)p->P.cidx.size();
    out[2] = p->S.n_free;
    out[3] = (int64_t)p->S.fronts.size();
    out[4] = (int64_t)p->S.rows.size();
    out[5] = (int64_t)p->S.rel.size();
    out[6] = (int64_t)p->S.k_src.size();
    out[7] = p->S.front_storage;
    out[8] = p->S.height;
}

void probe_pattern(void *h, int32_t *row_ptr, int32_t *col, int32_t *cptr, int32_t *cidx) {
    Probe *p = (Probe *)h;
    std::memcpy(row_ptr, p->P.row_ptr.data(), p->P.row_ptr.size() * 4);
    std::memcpy(col, p->P.col.data(), p->P.col.size() * 4);
    std::memcpy(cptr, p->P.cptr.data(), p->P.cptr.size() * 4);
    std::memcpy(cidx, p->P.cidx.data(), p->P.cidx.size() * 4);
}

// per front: ns, nu, m, piv0, parent, height (int64 x 6), off, row_off, rel_off (int64 x 3)
void probe_fronts(void *h, int64_t *info, int32_t *rows, int32_t *rel, int32_t *elim_of_dof, int32_t *k_src,
                  int64_t *k_dst) {
    Probe *p = (Probe *)h;
    const Symbolic &S = p->S;
    for (size_t f = 0; f < S.fronts.size(); f++) {
        const FrontInfo &F = S.fronts[f];
        int64_t *o = info + 9 * f;
        o[0] = F.ns; o[1] = F.nu; o[2] = F.m; o[3] = F.piv0; o[4] = F.parent; o[5] = F.height;
        o[6] = F.off; o[7] = F.row_off; o[8] = F.rel_off;
    }
    std::memcpy(rows, S.rows.data(), S.rows.size() * 4);
    std::memcpy(rel, S.rel.data(), S.rel.size() * 4);
    std::memcpy(elim_of_dof, S.elim_of_dof.data(), S.elim_of_dof.size() * 4);
    std::memcpy(k_src, S.k_src.data(), S.k_src.size() * 4);
    std::memcpy(k_dst, S.k_dst.data(), S.k_dst.size() * 8);
}
}
