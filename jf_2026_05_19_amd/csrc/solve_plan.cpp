// solve_plan.cpp -- tile plan of the GEMM-form triangular solves (host).
//
// Per front f (pivots ns, rows m): the forward product P_f (m x ns) is cut
// in 64-row tiles x 128-column k-slices, the backward product PT_f (ns x m)
// in 64-row tiles x 128-column k-slices.  Tiles of one tree height form one
// launch; the per-front tile base is relative to the height's first tile so
// the partial-sum buffer only has to hold the largest height.
#include "plan.h"
#include "symbolic.h"

#include <algorithm>

void solve_plan_build(const FactorPlan &P, const std::vector<int> &ns, const std::vector<int> &m,
                      const std::vector<int32_t> &flist, SolvePlan &S,
                      const std::vector<std::pair<ProdCfg, ProdCfg>> &cfg) {
    const int RT = SOLVE_RT, KS = SOLVE_KS;
    int nf = (int)ns.size();
    S.poff.assign(nf, 0);
    int64_t off = 0;
    int max_m = 0;
    for (int f = 0; f < nf; f++) {
        S.poff[f] = off;
        off += (int64_t)m[f] * ns[f];
        max_m = std::max(max_m, m[f]);
    }
    S.p_storage = off;
    S.p_pad = 64 * ((int64_t)max_m + 1);
    S.fwd_tiles.clear();
    S.bwd_tiles.clear();
    S.fwd_tbase.assign(nf, 0);
    S.bwd_tbase.assign(nf, 0);
    S.fwd_nks.assign(nf, 0);
    S.bwd_nks.assign(nf, 0);
    S.fwd_kw.assign(nf, KS);
    S.bwd_kw.assign(nf, KS);
    S.heights.assign(P.height + 1, SolveHeight());
    S.max_height_tiles = 0;
    for (int h = 0; h <= P.height; h++) {
        const HeightInfo &H = P.heights[h];
        SolveHeight &SH = S.heights[h];
        if (h < (int)cfg.size()) {
            SH.fwd_cfg = cfg[h].first;
            SH.bwd_cfg = cfg[h].second;
        }
        // the product kernel stages at most SOLVE_KMAX slice rows (its LDS
        // and its two slice rows per thread are sized for that)
        for (ProdCfg *c : {&SH.fwd_cfg, &SH.bwd_cfg}) {
            c->kmax = std::max(1, std::min(c->kmax, SOLVE_KMAX));
            c->ks = std::max(4, std::min(c->ks, SOLVE_KMAX));
            // one, two or four k-groups; the 32-step chunk has the registers
            // of one group only
            c->kg = c->kg >= 4 ? 4 : c->kg >= 2 ? 2 : 1;
            if (c->pch == 32) c->kg = 1;
        }
        const ProdCfg &CF = SH.fwd_cfg, &CB = SH.bwd_cfg;
        SH.fwd_off = (int64_t)S.fwd_tiles.size() / 3;
        SH.bwd_off = (int64_t)S.bwd_tiles.size() / 3;
        for (int q = 0; q < H.n; q++) {
            int f = flist[H.list_off + q];
            SH.sum_mns += (int64_t)m[f] * ns[f];
            SH.sum_nz += (int64_t)ns[f] * (ns[f] + 1) / 2 + (int64_t)(m[f] - ns[f]) * ns[f];
            SH.sum_m += m[f];
            SH.sum_ns += ns[f];
            // one k-slice (and no partial sums) when the inner dimension fits
            const int kwf = ns[f] <= CF.kmax ? std::max(ns[f], 1) : CF.ks;
            const int kwb = m[f] <= CB.kmax ? std::max(m[f], 1) : CB.ks;
            S.fwd_kw[f] = kwf;
            S.bwd_kw[f] = kwb;
            SH.fwd_kmax = std::max(SH.fwd_kmax, std::min(kwf, std::max(ns[f], 1)));
            SH.bwd_kmax = std::max(SH.bwd_kmax, std::min(kwb, std::max(m[f], 1)));
            int nrt = (m[f] + RT - 1) / RT, nk = (ns[f] + kwf - 1) / kwf;
            S.fwd_tbase[f] = (int32_t)(S.fwd_tiles.size() / 3 - SH.fwd_off);
            S.fwd_nks[f] = nk;
            SH.fwd_split = SH.fwd_split || nk > 1;
            for (int rt = 0; rt < nrt; rt++)
                for (int k = 0; k < nk; k++) {
                    S.fwd_tiles.push_back(f);
                    S.fwd_tiles.push_back(rt);
                    S.fwd_tiles.push_back(k);
                }
            int nrtb = (ns[f] + RT - 1) / RT, nkb = (m[f] + kwb - 1) / kwb;
            S.bwd_tbase[f] = (int32_t)(S.bwd_tiles.size() / 3 - SH.bwd_off);
            S.bwd_nks[f] = nkb;
            SH.bwd_split = SH.bwd_split || nkb > 1;
            for (int rt = 0; rt < nrtb; rt++)
                for (int k = 0; k < nkb; k++) {
                    S.bwd_tiles.push_back(f);
                    S.bwd_tiles.push_back(rt);
                    S.bwd_tiles.push_back(k);
                }
        }
        SH.n_fwd = (int)(S.fwd_tiles.size() / 3 - SH.fwd_off);
        SH.n_bwd = (int)(S.bwd_tiles.size() / 3 - SH.bwd_off);
        SH.max_m = H.max_m;
        SH.max_ns = H.max_ns;
        S.max_height_tiles = std::max<int64_t>(S.max_height_tiles, std::max(SH.n_fwd, SH.n_bwd));
    }
    // column blocks of the L11 inverses, longest dependency chain first
    std::vector<std::pair<int, std::pair<int32_t, int32_t>>> blocks;
    for (int f = 0; f < nf; f++) {
        int nb = (ns[f] + 31) / 32;
        for (int j = 0; j < nb; j++) blocks.push_back({nb - j, {f, 32 * j}});
    }
    std::stable_sort(blocks.begin(), blocks.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    S.lblk.clear();
    for (auto &b : blocks) {
        S.lblk.push_back(b.second.first);
        S.lblk.push_back(b.second.second);
    }
}

// Factorization plan (per-height launch bounds) and solve plan from the
// symbolic analysis.  flist receives the front ids grouped by height.
void build_plans(const Symbolic &S, FactorPlan &P, SolvePlan &SP, std::vector<int32_t> &flist) {
    int nf = (int)S.fronts.size();
    P.n_free = S.n_free;
    P.height = S.height;
    P.front_storage = S.front_storage;
    P.vec_storage = S.vec_storage;
    P.n_kmap = (int64_t)S.k_src.size();
    P.heights.assign(S.height + 1, HeightInfo());
    flist.clear();
    for (int h = 0; h <= S.height; h++) {
        HeightInfo &H = P.heights[h];
        H.n = (int)S.by_height[h].size();
        H.list_off = (int64_t)flist.size();
        for (int f : S.by_height[h]) {
            flist.push_back(f);
            H.max_ns = std::max(H.max_ns, S.fronts[f].ns);
            H.max_m = std::max(H.max_m, S.fronts[f].m);
            H.max_children = std::max(H.max_children, (int)S.children[f].size());
            for (int c : S.children[f])
                H.max_child_nu2 = std::max<int64_t>(H.max_child_nu2, (int64_t)S.fronts[c].nu * S.fronts[c].nu);
        }
    }
    std::vector<int> ns(nf), m(nf);
    for (int f = 0; f < nf; f++) { ns[f] = S.fronts[f].ns; m[f] = S.fronts[f].m; }
    solve_plan_build(P, ns, m, flist, SP);
    // pull map: count, prefix sum, fill (children in their stored order)
    SP.pull_ptr.assign(S.vec_storage + 1, 0);
    for (int f = 0; f < nf; f++)
        for (int c : S.children[f]) {
            const FrontInfo &C = S.fronts[c];
            for (int j = 0; j < C.nu; j++) SP.pull_ptr[S.fronts[f].vec_off + S.rel[C.rel_off + j] + 1]++;
        }
    for (int64_t i = 0; i < S.vec_storage; i++) SP.pull_ptr[i + 1] += SP.pull_ptr[i];
    SP.pull_src.assign(2 * (size_t)SP.pull_ptr[S.vec_storage], 0);
    std::vector<int32_t> cur(SP.pull_ptr.begin(), SP.pull_ptr.end() - 1);
    for (int f = 0; f < nf; f++)
        for (int c : S.children[f]) {
            const FrontInfo &C = S.fronts[c];
            for (int j = 0; j < C.nu; j++) {
                int32_t q = cur[S.fronts[f].vec_off + S.rel[C.rel_off + j]]++;
                SP.pull_src[2 * (size_t)q] = c;
                SP.pull_src[2 * (size_t)q + 1] = C.ns + j;
            }
        }
}
