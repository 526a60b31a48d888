// plan.h -- host-only plans of the multifrontal factorization and solves
// (no HIP types, so host tools and tests can build them with a plain C++
// compiler).
#pragma once
#include <cstdint>
#include <vector>

#define MAXR 16          // right-hand sides per triangular-solve call
constexpr int SOLVE_RT = 64;   // rows per product tile of the solves
constexpr int SOLVE_KS = 128;  // k-slice of the solve products (fronts split over k)
constexpr int SOLVE_KMAX = 384;  // fronts with inner dimension up to this are not split

struct HeightInfo {
    int n = 0;              // fronts at this height
    int64_t list_off = 0;   // offset in flist
    int max_ns = 0, max_m = 0;
    int max_children = 0;
    int64_t max_child_nu2 = 0;
};

struct FactorPlan {
    int n_free = 0;
    int height = 0;
    int64_t front_storage = 0;
    int64_t vec_storage = 0;
    int64_t n_kmap = 0;
    std::vector<HeightInfo> heights;
};

// Shape of one height's product launch: fronts whose inner dimension is at
// most kmax form one k-slice, wider ones are cut in ks-wide slices (reduced
// in slice order by a finish launch); pch four-k steps per chunk of factor
// loads (0: chosen from the launch's size); kg k-groups of four wavefronts
// per workgroup (1, 2 or 4).
struct ProdCfg {
    int kmax = SOLVE_KMAX, ks = SOLVE_KS, pch = 0, kg = 1;
};

struct SolveHeight {
    int64_t fwd_off = 0, bwd_off = 0;  // first tile of this height
    int n_fwd = 0, n_bwd = 0;          // tiles of this height
    int max_m = 0, max_ns = 0;
    int64_t sum_mns = 0, sum_m = 0, sum_ns = 0;  // factor entries and vector rows of this height
    int64_t sum_nz = 0;                          // nonzero factor entries: ns (ns + 1) / 2 + (m - ns) ns
    bool fwd_split = false, bwd_split = false;   // some front's product is split over k (needs a finish pass)
    int fwd_kmax = 0, bwd_kmax = 0;              // widest k-slice of this height's products
    ProdCfg fwd_cfg, bwd_cfg;
};

struct SolvePlan {
    std::vector<int64_t> poff;                  // per-front offset of P / PT (m * ns doubles)
    int64_t p_storage = 0;
    // zero-filled tail of P / PT: the block-RHS product loads whole 64 x 64
    // chunks of a front without masks, reaching at most 64 (max m + 1)
    // entries past the front's end
    int64_t p_pad = 0;
    std::vector<int32_t> fwd_tiles, bwd_tiles;  // (front, row tile, k slice) triples
    std::vector<int32_t> fwd_tbase, bwd_tbase;  // per front: first tile, relative to its height
    std::vector<int32_t> fwd_nks, bwd_nks;      // per front: k slices
    std::vector<int32_t> fwd_kw, bwd_kw;        // per front: k-slice width
    std::vector<SolveHeight> heights;
    int64_t max_height_tiles = 0;
    // forward-assembly pull map: row i of front f (vector slot vec_off[f] + i)
    // adds the child update rows pull_src[pull_ptr[slot] .. pull_ptr[slot+1])
    // as (child front, row in the child) pairs, in fixed child order
    std::vector<int32_t> pull_ptr, pull_src;
    // 32-wide column blocks of every L11, (front, first column) pairs, longest
    // block-column chain first (the explicit inverse is built from these)
    std::vector<int32_t> lblk;
};

inline int64_t solve_part_size(const SolvePlan &S) { return S.max_height_tiles * SOLVE_RT * MAXR; }
// working-vector storage: rows of MAXR values plus the 64 rows the block-RHS
// product may read past the last front (masked, never used)
inline int64_t solve_w_size(int64_t vec_storage) { return (vec_storage + 64) * MAXR; }

// cfg: per height (fwd, bwd) launch shapes; empty = defaults everywhere
void solve_plan_build(const FactorPlan &P, const std::vector<int> &ns, const std::vector<int> &m,
                      const std::vector<int32_t> &flist, SolvePlan &S,
                      const std::vector<std::pair<ProdCfg, ProdCfg>> &cfg = {});

struct Symbolic;
void build_plans(const Symbolic &S, FactorPlan &P, SolvePlan &SP, std::vector<int32_t> &flist);
