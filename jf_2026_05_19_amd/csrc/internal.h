// internal.h -- shared internals of the C ABI implementation: the context
// and model structures, device buffers and the host-side helpers that both
// the single-model entry points (api.cpp) and the batched entry points
// (batch.cpp) use.  Not part of the public interface.
#pragma once
#include "../../include/sol105.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "constraints.h"
#include "dense.h"
#include "factor.h"
#include "solve.h"
#include "symbolic.h"

struct sol105_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    double timings[8] = {0};
    double stats[8] = {0};
    hipEvent_t ev[9] = {nullptr};
    bool profiling = false;
    int lanczos_max_vectors = 1024;  // SOL105_OPT_LANCZOS_MAX_VECTORS
    int lanczos_max_restarts = 40;   // SOL105_OPT_LANCZOS_MAX_RESTARTS
    bool warm_start = false;         // SOL105_OPT_WARM_START
    bool graphs = true;              // SOL105_OPT_STEP_GRAPHS
    int sens_mode = 0;               // SOL105_OPT_SENS_MODE
    ProdTimer ptimer;  // HIP-event timing of the block triangular-solve products
    ProdTimer *timer() { return profiling ? &ptimer : nullptr; }
};

namespace s105 {

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void free_() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free_(); }
    hipError_t alloc(size_t count) {
        if (count <= n && p) return hipSuccess;
        free_();
        n = count;
        return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
    }
    hipError_t upload(const T *h, size_t count, hipStream_t s) {
        hipError_t e = alloc(count);
        if (e != hipSuccess) return e;
        if (count == 0) return hipSuccess;
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

}  // namespace s105

// HIP graphs of the block Lanczos expansion step, one per block index.  A
// step's launches (both triangular solves over every tree height, the KG
// product, the two orthogonalisation passes and Cholesky-QR) depend only on
// the block index and on buffer addresses, sizes and launch shapes that stay
// fixed for a model between analyses; `key` hashes all of them.  The first
// solve under a key launches directly and arms the cache, the next one
// captures each step as it reaches it, and later steps and solves replay.
struct StepGraphs {
    std::vector<hipGraphExec_t> exec;  // per block index (nullptr: not captured)
    uint64_t key = 0;
    bool armed = false;
    void clear() {
        for (auto &g : exec)
            if (g) (void)hipGraphExecDestroy(g);
        exec.clear();
        key = 0;
        armed = false;
    }
    ~StepGraphs() { clear(); }
};

// Eigensolver workspace kept with a model, so repeated solves (design
// loops, served requests) re-use the device buffers instead of allocating
// and freeing them per solve.
struct LanczosWork {
    s105::DevBuf<double> V, Wb, xf, yf, T, A, Zs, thd, ework, part, small, H2, Ytmp, Y, Ys, Xf, scr, sh;
    s105::DevBuf<int> ierr, sw;
    StepGraphs graphs;
};

// FNV-1a over the values a captured launch sequence depends on.
struct KeyHash {
    uint64_t h = 1469598103934665603ull;
    void bytes(const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 1099511628211ull;
    }
    template <typename T>
    void add(const T &v) {
        bytes(&v, sizeof(T));
    }
};

struct sol105_model {
    // host copies
    int n_nodes = 0, n_q4 = 0, n_t3 = 0, n_bar = 0;
    std::vector<double> xyz, q4prop, t3prop, bar_vref, barprop, force;
    std::vector<int32_t> q4, t3, bar;
    std::vector<uint8_t> spc_mask, has_shell, has_bar;
    double k6rot = 100.0;
    int autospc = 1, nd = 3;
    double v1 = 0.0, v2 = 0.0;  // EIGRL V1, V2
    int n_shell_cards = 0;      // shell cards of the deck (the mode filters' mesh-size guard)
    double sel_info[4] = {0, 0, 0, 1};  // last buckling: pairs, in-range candidates, cluster skip, decided
    BlockPattern pat;
    std::vector<int32_t> diag_blk;
    // RBE2 / RBE3 constraints: the map (dependent DOF -> independent DOFs)
    // and its device copies, the redistribution plan of K on the pattern
    ConstraintMap mpc;
    int64_t mpc_nt = 0, mpc_nz = 0;
    s105::DevBuf<int32_t> d_mpc_dep, d_mpc_ptr, d_mpc_ind, d_mpc_tptr;
    s105::DevBuf<double> d_mpc_coef, d_mpc_tcoef;
    s105::DevBuf<int64_t> d_mpc_tgt, d_mpc_src, d_mpc_zero;
    // the map transposed (receiving independent DOFs), for the adjoint load
    int32_t mpc_nrecv = 0;
    s105::DevBuf<int32_t> d_mpc_frecv, d_mpc_fptr, d_mpc_fdep;
    s105::DevBuf<double> d_mpc_fcoef;
    bool has_mpc() const { return !mpc.empty(); }
    // device: model + assembly
    s105::DevBuf<double> d_xyz, d_q4prop, d_t3prop, d_bar_vref, d_barprop, d_force;
    s105::DevBuf<int32_t> d_q4, d_t3, d_bar, d_rp, d_col, d_cptr, d_cidx, d_diag;
    s105::DevBuf<double> d_eb, d_K, d_KG, d_u, d_kdiag;
    s105::DevBuf<double> d_Kshift;  // K + shift I of the factorisation AUTOSPC probes (allocated on first use)
    // factor
    bool have_sym = false;
    std::vector<uint8_t> free_mask;
    Symbolic sym;
    FactorPlan plan;
    s105::DevBuf<int32_t> f_ns, f_nu, f_m, f_piv0, f_rows, f_rel, f_cptr, f_cidx, f_flist, f_ksrc, d_elim_of_dof,
        d_dof_of_elim;
    s105::DevBuf<int64_t> f_off, f_rowoff, f_vecoff, f_reloff, f_kdst;
    s105::DevBuf<uint8_t> d_free, d_flag;
    s105::DevBuf<double> d_front, d_piv, d_W, d_kdiag_elim, d_tsol;  // d_tsol: the intermediate y of tri_solve
    s105::DevBuf<int> d_err;
    DevFronts dev;
    SolvePlan splan;
    s105::DevBuf<int64_t> s_poff;
    s105::DevBuf<int32_t> s_ftiles, s_btiles, s_ftbase, s_btbase, s_fnks, s_bnks, s_fkw, s_bkw, s_pullptr, s_pullsrc, s_pull2, s_lblk;
    bool pull2_ok = false;
    s105::DevBuf<TileDesc> s_fdesc, s_bdesc;
    // product launch shapes per height (forward, backward; empty: defaults)
    // and the front ids grouped by height of the current analysis
    std::vector<std::pair<ProdCfg, ProdCfg>> prod_cfg;
    std::vector<int32_t> flist_host;
    s105::DevBuf<double> d_P, d_PT, d_part;
    DevSolve dsolve;
    // state
    bool K_valid = false, factor_valid = false, u_valid = false, KG_valid = false;
    // eigen results
    std::vector<double> lambda;
    s105::DevBuf<double> d_modes;  // ndof x n_modes (normalised, global frame)
    int n_modes = 0;
    double lanczos_m = 0, lanczos_steps = 0;
    int warm_m = 0, warm_b = 0;  // Lanczos size at the last convergence (first check there next time)
    // symbolic analyses so far (the free set and elimination order change
    // with each); the Ritz vectors kept in lw.Y (y_cols columns, elimination
    // order) belong to analysis y_sym
    uint64_t sym_gen = 0, y_sym = 0;
    int y_cols = 0;
    bool warm_first_ok = true;  // the last warm-started solve converged at its first check
    // a batch model is the block-diagonal union of independent members; their
    // node ranges (n_members + 1 offsets; empty for a single model) scope the
    // per-model AUTOSPC maxima
    std::vector<int32_t> mem_node_off;
    LanczosWork lw;
    // result recovery: CTRIA3 macro-blend weights (host copy sized n_t3 once
    // computed), bar recovery points and the device outputs
    std::vector<double> t3w;
    s105::DevBuf<double> d_t3w, d_barpts, d_rec;
    int ndof() const { return 6 * n_nodes; }
};

namespace s105 {

int set_err(sol105_ctx *ctx, int code, const char *fmt, ...);

#define HIPCHK(call)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess) return s105::set_err(ctx, SOL105_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

ElemArgs elem_args(sol105_model *M);
// AUTOSPC on the diagonal, per member of a batch model
void autospc_mask(const sol105_model *M, const std::vector<double> &kd, std::vector<uint8_t> &free_out);
int upload_symbolic(sol105_ctx *ctx, sol105_model *M);
int upload_solve_tiles(sol105_ctx *ctx, sol105_model *M);
// x <- K_ff^-1 x (elimination order, nrhs <= MAXR columns)
hipError_t tri_solve(sol105_model *M, double *X, int nrhs, hipStream_t s);
double ev_ms(hipEvent_t a, hipEvent_t b);
// host copies, validation, block pattern and device upload of a model
int model_init(sol105_ctx *ctx, const sol105_model_desc *d, sol105_model *M);
// AUTOSPC + symbolic (when the free set changed) + numeric factorisation with
// the pivot-ratio refix and, when shifted_fallback is set, the factorisation
// AUTOSPC of a failed Cholesky (shifted probes)
int factorize(sol105_ctx *ctx, sol105_model *M, bool shifted_fallback);
// the launch shapes and device addresses that the triangular solves and the
// KG product of a model depend on (key of its Lanczos step graphs)
void hash_solve_state(KeyHash &k, const sol105_model *M);
// device address of the no-allocation scratch array on `device` (current),
// looked up once per device (noalloc.hip)
hipError_t noalloc_setup(int device);

}  // namespace s105
