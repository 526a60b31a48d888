// symbolic.cpp -- block pattern, deterministic gather map and the
// nested-dissection supernodal analysis of the free-DOF stiffness matrix.
//
// The reference sums COO triplets with Julia's sparse() in element order
// (JFEM/src/solver/assembly.jl:4371) and hands K_ff to CHOLMOD with its own
// fill-reducing ordering (JFEM/src/solver/boundary_conditions.jl:740).  Here
// the summation order is fixed by a precomputed element->block map, and the
// ordering is a geometric nested dissection over mesh nodes whose separators
// become the supernodes (fronts) of a multifrontal Cholesky.
#include "symbolic.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

int64_t BlockPattern::find(int i, int j) const {
    auto b = col.begin() + row_ptr[i], e = col.begin() + row_ptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    if (it == e || *it != j) return -1;
    return (int64_t)(it - col.begin());
}

void build_block_pattern(int n_nodes, int n_q4, const int32_t *q4, int n_t3, const int32_t *t3, int n_bar,
                         const int32_t *bar, BlockPattern &P,
                         const std::vector<std::pair<int32_t, int32_t>> *extra) {
    P.n_nodes = n_nodes;
    std::vector<std::vector<int32_t>> nb(n_nodes);
    for (int e = 0; e < n_q4; e++)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) nb[q4[4 * e + a]].push_back(q4[4 * e + b]);
    for (int e = 0; e < n_t3; e++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) nb[t3[3 * e + a]].push_back(t3[3 * e + b]);
    for (int e = 0; e < n_bar; e++)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) nb[bar[2 * e + a]].push_back(bar[2 * e + b]);
    if (extra)
        for (const auto &pr : *extra) nb[pr.first].push_back(pr.second);
    P.row_ptr.assign(n_nodes + 1, 0);
    for (int i = 0; i < n_nodes; i++) {
        auto &v = nb[i];
        v.push_back(i);  // every node keeps its diagonal block
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        P.row_ptr[i + 1] = P.row_ptr[i] + (int32_t)v.size();
    }
    P.col.resize(P.row_ptr[n_nodes]);
    for (int i = 0; i < n_nodes; i++) std::copy(nb[i].begin(), nb[i].end(), P.col.begin() + P.row_ptr[i]);
    int64_t nbk = P.n_blocks();
    std::vector<int32_t> cnt(nbk + 1, 0);
    std::vector<int64_t> q4blk((size_t)n_q4 * 16), t3blk((size_t)n_t3 * 9), barblk((size_t)n_bar * 4);
    for (int e = 0; e < n_q4; e++)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                int64_t k = P.find(q4[4 * e + a], q4[4 * e + b]);
                q4blk[(size_t)e * 16 + a * 4 + b] = k;
                cnt[k + 1]++;
            }
    for (int e = 0; e < n_t3; e++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                int64_t k = P.find(t3[3 * e + a], t3[3 * e + b]);
                t3blk[(size_t)e * 9 + a * 3 + b] = k;
                cnt[k + 1]++;
            }
    for (int e = 0; e < n_bar; e++)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                int64_t k = P.find(bar[2 * e + a], bar[2 * e + b]);
                barblk[(size_t)e * 4 + a * 2 + b] = k;
                cnt[k + 1]++;
            }
    for (int64_t k = 0; k < nbk; k++) cnt[k + 1] += cnt[k];
    P.cptr = cnt;
    P.cidx.resize(cnt[nbk]);
    std::vector<int32_t> cur(cnt.begin(), cnt.end() - 1);
    for (int64_t t = 0; t < (int64_t)q4blk.size(); t++) P.cidx[cur[q4blk[t]]++] = (int32_t)t;
    for (int64_t t = 0; t < (int64_t)t3blk.size(); t++)
        P.cidx[cur[t3blk[t]]++] = (int32_t)((int64_t)n_q4 * 16 + t);
    for (int64_t t = 0; t < (int64_t)barblk.size(); t++)
        P.cidx[cur[barblk[t]]++] = (int32_t)((int64_t)n_q4 * 16 + (int64_t)n_t3 * 9 + t);
}

namespace {

struct NDTree {
    std::vector<std::vector<int32_t>> nodes;  // supernode node lists
    std::vector<int32_t> left, right;
};

struct NDContext {
    const BlockPattern *P;
    const double *xyz;
    const std::vector<uint8_t> *active;
    std::vector<int32_t> part;  // scratch marks
    int stamp = 0;
    int leaf;
    NDTree T;
};

int nd_leaf(NDContext &C, std::vector<int32_t> &nodes) {
    C.T.nodes.push_back(std::move(nodes));
    C.T.left.push_back(-1);
    C.T.right.push_back(-1);
    return (int)C.T.nodes.size() - 1;
}

int nd_recurse(NDContext &C, std::vector<int32_t> nodes) {
    if ((int)nodes.size() <= C.leaf) return nd_leaf(C, nodes);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v : nodes)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], C.xyz[3 * v + d]);
            hi[d] = std::max(hi[d], C.xyz[3 * v + d]);
        }
    int ax = 0;
    for (int d = 1; d < 3; d++)
        if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    if (!(hi[ax] - lo[ax] > 0.0)) return nd_leaf(C, nodes);
    std::sort(nodes.begin(), nodes.end(), [&](int a, int b) {
        double xa = C.xyz[3 * a + ax], xb = C.xyz[3 * b + ax];
        return xa < xb || (xa == xb && a < b);
    });
    size_t half = nodes.size() / 2;
    // move the cut to a coordinate boundary when one is close (keeps grid
    // lines whole, which gives one-line separators on structured meshes)
    {
        double xc = C.xyz[3 * nodes[half] + ax];
        size_t lo_i = half, hi_i = half;
        while (lo_i > 0 && C.xyz[3 * nodes[lo_i - 1] + ax] == xc) lo_i--;
        while (hi_i < nodes.size() && C.xyz[3 * nodes[hi_i] + ax] == xc) hi_i++;
        size_t best = (half - lo_i <= hi_i - half) ? lo_i : hi_i;
        if (best > 0 && best < nodes.size()) half = best;
    }
    int sa = ++C.stamp, sb = ++C.stamp;
    for (size_t i = 0; i < nodes.size(); i++) C.part[nodes[i]] = i < half ? sa : sb;
    std::vector<int32_t> sepA, sepB;
    const BlockPattern &P = *C.P;
    for (size_t i = 0; i < nodes.size(); i++) {
        int v = nodes[i];
        int other = i < half ? sb : sa;
        for (int k = P.row_ptr[v]; k < P.row_ptr[v + 1]; k++)
            if (C.part[P.col[k]] == other) {
                (i < half ? sepA : sepB).push_back(v);
                break;
            }
    }
    bool useB = sepB.size() <= sepA.size();
    std::vector<int32_t> &sep = useB ? sepB : sepA;
    if (sep.empty()) sep.push_back(nodes.back());
    int ss = ++C.stamp;
    for (int v : sep) C.part[v] = ss;
    std::vector<int32_t> A, B;
    for (size_t i = 0; i < nodes.size(); i++) {
        int v = nodes[i];
        if (C.part[v] == ss) continue;
        (i < half ? A : B).push_back(v);
    }
    if (A.empty() || B.empty()) return nd_leaf(C, nodes);
    std::vector<int32_t> sepc = sep;
    int l = nd_recurse(C, std::move(A));
    int r = nd_recurse(C, std::move(B));
    C.T.nodes.push_back(std::move(sepc));
    C.T.left.push_back(l);
    C.T.right.push_back(r);
    return (int)C.T.nodes.size() - 1;
}

}  // namespace

void symbolic_analyze(const BlockPattern &P, const double *xyz, const uint8_t *free_mask, int leaf_nodes,
                      Symbolic &S) {
    int n = P.n_nodes;
    std::vector<uint8_t> active(n, 0);
    std::vector<int32_t> act_nodes;
    for (int v = 0; v < n; v++) {
        int c = 0;
        for (int k = 0; k < 6; k++) c += free_mask[6 * v + k] ? 1 : 0;
        active[v] = c > 0;
        if (c > 0) act_nodes.push_back(v);
    }
    // A subgraph restricted to active nodes: inactive nodes never get a part
    // mark, so they are ignored by the separator search.
    NDContext C;
    C.P = &P;
    C.xyz = xyz;
    C.active = &active;
    C.part.assign(n, 0);
    C.leaf = std::max(leaf_nodes, 1);
    // split into connected components first (each its own tree)
    std::vector<int32_t> comp(n, -1);
    std::vector<int> roots;
    {
        int nc = 0;
        std::vector<int32_t> stack;
        for (int v : act_nodes) {
            if (comp[v] >= 0) continue;
            std::vector<int32_t> members;
            comp[v] = nc;
            stack.push_back(v);
            while (!stack.empty()) {
                int x = stack.back();
                stack.pop_back();
                members.push_back(x);
                for (int k = P.row_ptr[x]; k < P.row_ptr[x + 1]; k++) {
                    int y = P.col[k];
                    if (active[y] && comp[y] < 0) { comp[y] = nc; stack.push_back(y); }
                }
            }
            roots.push_back(nd_recurse(C, std::move(members)));
            nc++;
        }
    }
    NDTree &T = C.T;
    int nf = (int)T.nodes.size();
    std::vector<int32_t> parent(nf, -1);
    for (int f = 0; f < nf; f++) {
        if (T.left[f] >= 0) parent[T.left[f]] = f;
        if (T.right[f] >= 0) parent[T.right[f]] = f;
    }
    // node elimination numbers in postorder (tree ids are already postorder)
    std::vector<int32_t> node_elim(n, -1);
    std::vector<int32_t> first(nf), last(nf);
    int cnt = 0;
    for (int f = 0; f < nf; f++) {
        first[f] = cnt;
        for (int v : T.nodes[f]) node_elim[v] = cnt++;
        last[f] = cnt - 1;
    }
    // node-level update structures
    std::vector<std::vector<int32_t>> U(nf);
    std::vector<int32_t> mark(n, -1);
    for (int f = 0; f < nf; f++) {
        std::vector<int32_t> cand;
        auto add = [&](int w) {
            if (node_elim[w] > last[f] && mark[w] != f) { mark[w] = f; cand.push_back(w); }
        };
        for (int v : T.nodes[f])
            for (int k = P.row_ptr[v]; k < P.row_ptr[v + 1]; k++)
                if (active[P.col[k]]) add(P.col[k]);
        for (int ch : {T.left[f], T.right[f]})
            if (ch >= 0)
                for (int w : U[ch]) add(w);
        std::sort(cand.begin(), cand.end(), [&](int a, int b) { return node_elim[a] < node_elim[b]; });
        U[f] = std::move(cand);
    }
    // DOF level
    std::vector<int32_t> order(cnt);
    for (int v = 0; v < n; v++)
        if (node_elim[v] >= 0) order[node_elim[v]] = v;
    S.elim_of_dof.assign((size_t)6 * n, -1);
    S.dof_of_elim.clear();
    std::vector<int32_t> node_first_dof(n, -1), node_nfree(n, 0);
    for (int v : order) {
        node_first_dof[v] = (int32_t)S.dof_of_elim.size();
        for (int k = 0; k < 6; k++)
            if (free_mask[6 * v + k]) {
                S.elim_of_dof[6 * v + k] = (int32_t)S.dof_of_elim.size();
                S.dof_of_elim.push_back(6 * v + k);
                node_nfree[v]++;
            }
    }
    S.n_free = (int)S.dof_of_elim.size();
    S.fronts.assign(nf, FrontInfo());
    S.rows.clear();
    S.children.assign(nf, {});
    int64_t off = 0, voff = 0;
    S.nnz_L = 0;
    S.flops = 0.0;
    S.max_ns = S.max_m = S.max_nu = 0;
    for (int f = 0; f < nf; f++) {
        FrontInfo &F = S.fronts[f];
        F.parent = parent[f];
        F.piv0 = node_first_dof[T.nodes[f][0]];
        F.ns = 0;
        for (int v : T.nodes[f]) F.ns += node_nfree[v];
        F.row_off = (int64_t)S.rows.size();
        for (int i = 0; i < F.ns; i++) S.rows.push_back(F.piv0 + i);
        for (int w : U[f])
            for (int k = 0; k < node_nfree[w]; k++) S.rows.push_back(node_first_dof[w] + k);
        F.m = (int)((int64_t)S.rows.size() - F.row_off);
        F.nu = F.m - F.ns;
        F.off = off;
        off += (int64_t)F.m * F.m;
        F.vec_off = voff;
        voff += F.m;
        int h = 0;
        for (int ch : {T.left[f], T.right[f]})
            if (ch >= 0) { h = std::max(h, S.fronts[ch].height + 1); S.children[f].push_back(ch); }
        F.height = h;
        double ns = F.ns, nu = F.nu;
        S.nnz_L += (int64_t)(ns * (ns + 1) / 2 + ns * nu);
        S.flops += ns * ns * ns / 3.0 + ns * ns * nu + ns * nu * nu;
        S.max_ns = std::max(S.max_ns, F.ns);
        S.max_m = std::max(S.max_m, F.m);
        S.max_nu = std::max(S.max_nu, F.nu);
    }
    S.front_storage = off;
    S.vec_storage = voff;
    // child -> parent relative positions
    S.rel.clear();
    for (int f = 0; f < nf; f++) {
        FrontInfo &F = S.fronts[f];
        F.rel_off = (int64_t)S.rel.size();
        if (F.parent < 0) {
            if (F.nu != 0) throw std::runtime_error("root front with update rows");
            continue;
        }
        const FrontInfo &Pf = S.fronts[F.parent];
        const int32_t *pr = S.rows.data() + Pf.row_off;
        for (int i = 0; i < F.nu; i++) {
            int32_t r = S.rows[F.row_off + F.ns + i];
            const int32_t *it = std::lower_bound(pr, pr + Pf.m, r);
            if (it == pr + Pf.m || *it != r) throw std::runtime_error("update row missing in parent front");
            S.rel.push_back((int32_t)(it - pr));
        }
    }
    S.height = 0;
    for (auto &F : S.fronts) S.height = std::max(S.height, F.height);
    S.by_height.assign(S.height + 1, {});
    for (int f = 0; f < nf; f++) S.by_height[S.fronts[f].height].push_back(f);
    // K -> front scatter map (lower triangle in elimination order)
    std::vector<int32_t> front_of_elim(S.n_free);
    for (int f = 0; f < nf; f++)
        for (int i = 0; i < S.fronts[f].ns; i++) front_of_elim[S.fronts[f].piv0 + i] = f;
    S.k_src.clear();
    S.k_dst.clear();
    for (int I = 0; I < n; I++) {
        if (!active[I]) continue;
        for (int k = P.row_ptr[I]; k < P.row_ptr[I + 1]; k++) {
            int J = P.col[k];
            if (!active[J]) continue;
            for (int ci = 0; ci < 6; ci++) {
                int pi = S.elim_of_dof[6 * I + ci];
                if (pi < 0) continue;
                for (int cj = 0; cj < 6; cj++) {
                    int pj = S.elim_of_dof[6 * J + cj];
                    if (pj < 0 || pi < pj) continue;
                    int f = front_of_elim[pj];
                    const FrontInfo &F = S.fronts[f];
                    const int32_t *fr = S.rows.data() + F.row_off;
                    const int32_t *it = std::lower_bound(fr, fr + F.m, pi);
                    if (it == fr + F.m || *it != pi) throw std::runtime_error("K entry outside front structure");
                    int64_t lrow = it - fr, lcol = pj - F.piv0;
                    S.k_src.push_back((int32_t)(k * 36 + ci * 6 + cj));
                    S.k_dst.push_back(F.off + lrow + lcol * F.m);
                }
            }
        }
    }
}
