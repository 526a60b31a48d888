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

void *probe_analyze(int n_nodes, int n_q4, const int32_t *q4, int n_t3, const int32_t *t3, int n_bar,
                    const int32_t *bar, const double *xyz, const uint8_t *free_mask, int leaf) {
    Probe *p = new Probe();
    build_block_pattern(n_nodes, n_q4, q4, n_t3, t3, n_bar, bar, p->P);
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

#include "../../jf_2026_05_19_amd/csrc/plan.h"

struct PlanProbe {
    FactorPlan P;
    SolvePlan SP;
    std::vector<int32_t> flist;
};

extern "C" {

void *probe_plans(void *h) {
    Probe *p = (Probe *)h;
    PlanProbe *q = new PlanProbe();
    build_plans(p->S, q->P, q->SP, q->flist);
    return q;
}

void probe_plans_free(void *q) { delete (PlanProbe *)q; }

// [n_fwd_tiles, n_bwd_tiles, height, max_height_tiles, p_storage]
void probe_plan_sizes(void *qh, int64_t *out) {
    PlanProbe *q = (PlanProbe *)qh;
    out[0] = (int64_t)q->SP.fwd_tiles.size() / 3;
    out[1] = (int64_t)q->SP.bwd_tiles.size() / 3;
    out[2] = q->P.height;
    out[3] = q->SP.max_height_tiles;
    out[4] = q->SP.p_storage;
}

// per height (8 values): list_off, n, fwd_off, bwd_off, n_fwd, n_bwd, max_m, max_ns
void probe_plan(void *qh, int32_t *ft, int32_t *bt, int32_t *ftb, int32_t *btb, int32_t *fnk, int32_t *bnk,
                int32_t *flist, int64_t *hinfo, int64_t *poff) {
    PlanProbe *q = (PlanProbe *)qh;
    const SolvePlan &S = q->SP;
    std::memcpy(ft, S.fwd_tiles.data(), S.fwd_tiles.size() * 4);
    std::memcpy(bt, S.bwd_tiles.data(), S.bwd_tiles.size() * 4);
    std::memcpy(ftb, S.fwd_tbase.data(), S.fwd_tbase.size() * 4);
    std::memcpy(btb, S.bwd_tbase.data(), S.bwd_tbase.size() * 4);
    std::memcpy(fnk, S.fwd_nks.data(), S.fwd_nks.size() * 4);
    std::memcpy(bnk, S.bwd_nks.data(), S.bwd_nks.size() * 4);
    std::memcpy(flist, q->flist.data(), q->flist.size() * 4);
    for (int h = 0; h <= q->P.height; h++) {
        const HeightInfo &H = q->P.heights[h];
        const SolveHeight &SH = S.heights[h];
        int64_t *o = hinfo + 8 * h;
        o[0] = H.list_off; o[1] = H.n; o[2] = SH.fwd_off; o[3] = SH.bwd_off;
        o[4] = SH.n_fwd; o[5] = SH.n_bwd; o[6] = H.max_m; o[7] = H.max_ns;
    }
    std::memcpy(poff, S.poff.data(), S.poff.size() * 8);
}
}

extern "C" {

// pull map of the forward assembly: sizes [n_slots + 1, n_entries]
void probe_pull_sizes(void *qh, int64_t *out) {
    PlanProbe *q = (PlanProbe *)qh;
    out[0] = (int64_t)q->SP.pull_ptr.size();
    out[1] = (int64_t)q->SP.pull_src.size() / 2;
}

void probe_pull(void *qh, int32_t *ptr, int32_t *src) {
    PlanProbe *q = (PlanProbe *)qh;
    std::memcpy(ptr, q->SP.pull_ptr.data(), q->SP.pull_ptr.size() * 4);
    std::memcpy(src, q->SP.pull_src.data(), q->SP.pull_src.size() * 4);
}
}
