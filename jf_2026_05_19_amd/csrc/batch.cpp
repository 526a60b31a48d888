// batch.cpp -- batched SOL 105 solves (include/sol105.h, sol105_batch_*).
//
// The reference runs a batch as a loop of independent solves
// (JFEM/tools/manifest_batch_core.jl:409-478).  Here the members of a batch
// are stacked into one block-diagonal model.  Element kernels, assembly, the
// static solve and KG then run once for all members, and the multifrontal
// factorisation and the triangular solves run once per elimination-tree
// height: the symbolic analysis keeps every connected component in its own
// tree, so a member's free DOFs form one contiguous range of the elimination
// order (a segment).  The block Lanczos iteration runs on all segments in
// lockstep through the segmented tall-skinny kernels (inner products, block
// orthonormalisation and Rayleigh-Ritz per member), each member stopping at
// its own convergence.  Per member, the eigen semantics are those of
// sol105_buckling (mode count, filters, normalisation).  A member whose block
// turns rank-deficient, or whose iteration or batch factorisation fails, is
// re-solved alone by the single-model path, which has the rank-revealing
// rebuild.
#include "internal.h"
#include "select.h"

#include <cmath>
#include <cstring>
#include <memory>

using namespace s105;

struct sol105_batch {
    sol105_model M;  // block-diagonal union of the members
    int n_mem = 0;
    std::vector<int32_t> node_off, q4_off, t3_off, bar_off;  // n_mem + 1
    std::vector<int32_t> nd, n_shell_cards;
    std::vector<double> v1, v2;  // EIGRL range per member
    // results of the last solve
    std::vector<std::vector<double>> lambda;
    std::vector<int32_t> status, n_out;
    std::vector<int64_t> mode_off;  // member's modes at d_modes + mode_off[i], 6 n_i x n_out[i]
    DevBuf<double> d_modes;
    // segment table (device)
    DevBuf<int32_t> d_seg_off, d_ch_r0, d_ch_seg, d_ch_ptr, d_chs_r0, d_chs_seg, d_chs_ptr;
    DevBuf<int32_t> d_q4_off, d_t3_off;
    DevBuf<uint8_t> d_active, d_conv, d_idle;
    // stacked-iteration workspace, kept between solves of the batch (the
    // step graphs replay on its addresses)
    DevBuf<double> V, Wb, xf, yf, Tm, A, Zs, thd, ework, part, G1, G2, G0, Rt, H2, Y, zl, rl;
    DevBuf<int> ierr, info, dlist;
    StepGraphs graphs;
    double last_ms = 0.0;
    int last_m = 0, last_steps = 0, n_fallback = 0;
};

namespace {

// A member's model description sliced out of the union's host arrays.
struct MemberDesc {
    std::vector<double> xyz, q4prop, t3prop, vref, barprop, force;
    std::vector<int32_t> q4, t3, bar;
    std::vector<uint8_t> spc, hs, hb;
    sol105_model_desc d{};
};

void member_desc(const sol105_batch *B, int i, MemberDesc &D) {
    const sol105_model &M = B->M;
    int n0 = B->node_off[i], n1 = B->node_off[i + 1];
    int q0 = B->q4_off[i], q1 = B->q4_off[i + 1], t0 = B->t3_off[i], t1 = B->t3_off[i + 1];
    int b0 = B->bar_off[i], b1 = B->bar_off[i + 1];
    D.xyz.assign(M.xyz.begin() + 3 * n0, M.xyz.begin() + 3 * n1);
    D.q4.resize(4 * (size_t)(q1 - q0));
    for (size_t k = 0; k < D.q4.size(); k++) D.q4[k] = M.q4[4 * (size_t)q0 + k] - n0;
    D.q4prop.assign(M.q4prop.begin() + 6 * (size_t)q0, M.q4prop.begin() + 6 * (size_t)q1);
    D.t3.resize(3 * (size_t)(t1 - t0));
    for (size_t k = 0; k < D.t3.size(); k++) D.t3[k] = M.t3[3 * (size_t)t0 + k] - n0;
    D.t3prop.assign(M.t3prop.begin() + 6 * (size_t)t0, M.t3prop.begin() + 6 * (size_t)t1);
    D.bar.resize(2 * (size_t)(b1 - b0));
    for (size_t k = 0; k < D.bar.size(); k++) D.bar[k] = M.bar[2 * (size_t)b0 + k] - n0;
    D.vref.assign(M.bar_vref.begin() + 3 * (size_t)b0, M.bar_vref.begin() + 3 * (size_t)b1);
    D.barprop.assign(M.barprop.begin() + 8 * (size_t)b0, M.barprop.begin() + 8 * (size_t)b1);
    D.spc.assign(M.spc_mask.begin() + 6 * (size_t)n0, M.spc_mask.begin() + 6 * (size_t)n1);
    D.force.assign(M.force.begin() + 6 * (size_t)n0, M.force.begin() + 6 * (size_t)n1);
    D.hs.assign(M.has_shell.begin() + n0, M.has_shell.begin() + n1);
    D.hb.assign(M.has_bar.begin() + n0, M.has_bar.begin() + n1);
    sol105_model_desc &d = D.d;
    d.n_nodes = n1 - n0;
    d.xyz = D.xyz.data();
    d.n_q4 = q1 - q0;
    d.q4_nodes = D.q4.data();
    d.q4_prop = D.q4prop.data();
    d.n_t3 = t1 - t0;
    d.t3_nodes = D.t3.data();
    d.t3_prop = D.t3prop.data();
    d.n_bar = b1 - b0;
    d.bar_nodes = D.bar.data();
    d.bar_vref = D.vref.data();
    d.bar_prop = D.barprop.data();
    d.spc_mask = D.spc.data();
    d.force = D.force.data();
    d.node_has_shell = D.hs.data();
    d.node_has_bar = D.hb.data();
    d.k6rot = M.k6rot;
    d.autospc = M.autospc;
    d.nd = B->nd[i];
    d.eigrl_v1 = B->v1[i];
    d.eigrl_v2 = B->v2[i];
    d.n_shell_cards = B->n_shell_cards[i];
}

ModeSelect member_select(const sol105_batch *B, int i, int n_free) {
    ModeSelect sel;
    sel.nd = B->nd[i];
    sel.n_free = n_free;
    sel.v1 = B->v1[i];
    sel.v2 = B->v2[i];
    sel.n_shell_cards = B->n_shell_cards[i];
    return sel;
}

// Per-member eigen semantics of sol105_buckling (solve_case.jl:1938-1945,
// 2131-2196, 2244-2252, 2414, 2568).
struct MemberEig {
    int n = 0;             // free DOFs (segment length)
    ModeSelect sel;
    int req = 0, kneed = 0, mmax = 0;
    bool active = false, fallback = false;
    int code = SOL105_OK;  // error of this member (when not re-solved)
    std::vector<double> theta;
};

// Segmented block Lanczos + post-processing of all members.
int batch_buckling(sol105_ctx *ctx, sol105_batch *B, int max_modes) {
    sol105_model *M = &B->M;
    hipStream_t s = ctx->stream;
    const int nm = B->n_mem, ntot = M->sym.n_free, ndof = M->ndof();
    const double tol = 1e-10, dep_tol2 = 1e-14;
    std::vector<MemberEig> E(nm);
    // segments: each member's free DOFs are one contiguous range of the
    // elimination order (components are eliminated one after another)
    std::vector<int32_t> seg_off(nm + 1, 0);
    {
        int cur = 0;
        for (int i = 0; i < nm; i++) {
            int lo = INT32_MAX, hi = -1, cnt = 0;
            for (int d = 6 * B->node_off[i]; d < 6 * B->node_off[i + 1]; d++) {
                int e = M->sym.elim_of_dof[d];
                if (e < 0) continue;
                lo = std::min(lo, e);
                hi = std::max(hi, e);
                cnt++;
            }
            if (cnt > 0 && (lo != cur || hi - lo + 1 != cnt))
                return set_err(ctx, SOL105_E_INVALID, "batch member %d is not contiguous in the elimination order", i);
            seg_off[i] = cur;
            cur += cnt;
            E[i].n = cnt;
        }
        seg_off[nm] = cur;
        if (cur != ntot) return set_err(ctx, SOL105_E_INVALID, "batch segments do not cover the free set");
    }
    int b = 8;
    for (int i = 0; i < nm; i++) {
        MemberEig &e = E[i];
        e.sel = member_select(B, i, e.n);
        e.req = e.sel.requested();
        if (e.n == 0) { e.code = SOL105_E_SINGULAR; continue; }
        // a range (whole-spectrum semantics) or more pairs than the stacked
        // iteration carries: the member is solved alone
        if (e.sel.has_range() || e.req > MAXR) { e.fallback = true; continue; }
        if (e.req > 8) b = 16;
    }
    int mmax = 0, kmax = 0;
    for (int i = 0; i < nm; i++) {
        MemberEig &e = E[i];
        if (e.code != SOL105_OK || e.fallback) continue;
        e.mmax = std::min(384, (e.n / b) * b);
        if (e.mmax < b) { e.fallback = true; continue; }  // too small for the stacked block size
        e.kneed = std::min(e.req + 2, e.mmax);
        e.active = true;
        mmax = std::max(mmax, e.mmax);
        kmax = std::max(kmax, e.kneed);
    }
    // chunk tables
    std::vector<int32_t> ch_r0, ch_seg, ch_ptr(nm + 1, 0), chs_r0, chs_seg, chs_ptr(nm + 1, 0);
    for (int i = 0; i < nm; i++) {
        for (int r = seg_off[i]; r < seg_off[i + 1]; r += SEG_CH) { ch_r0.push_back(r); ch_seg.push_back(i); }
        ch_ptr[i + 1] = (int32_t)ch_r0.size();
        for (int r = seg_off[i]; r < seg_off[i + 1]; r += SEG_CHS) { chs_r0.push_back(r); chs_seg.push_back(i); }
        chs_ptr[i + 1] = (int32_t)chs_r0.size();
    }
    std::vector<uint8_t> act(nm), conv(nm, 0), idle(nm);
    for (int i = 0; i < nm; i++) {
        act[i] = E[i].active;
        idle[i] = !E[i].active;
    }
    HIPCHK(B->d_seg_off.upload(seg_off.data(), seg_off.size(), s));
    HIPCHK(B->d_ch_r0.upload(ch_r0.data(), ch_r0.size(), s));
    HIPCHK(B->d_ch_seg.upload(ch_seg.data(), ch_seg.size(), s));
    HIPCHK(B->d_ch_ptr.upload(ch_ptr.data(), ch_ptr.size(), s));
    HIPCHK(B->d_chs_r0.upload(chs_r0.data(), chs_r0.size(), s));
    HIPCHK(B->d_chs_seg.upload(chs_seg.data(), chs_seg.size(), s));
    HIPCHK(B->d_chs_ptr.upload(chs_ptr.data(), chs_ptr.size(), s));
    HIPCHK(B->d_active.upload(act.data(), act.size(), s));
    HIPCHK(B->d_conv.upload(conv.data(), conv.size(), s));
    HIPCHK(B->d_idle.upload(idle.data(), idle.size(), s));
    SegTable T;
    T.n_seg = nm;
    T.n_ch = (int)ch_r0.size();
    T.n_chs = (int)chs_r0.size();
    T.seg_off = B->d_seg_off.p;
    T.ch_r0 = B->d_ch_r0.p; T.ch_seg = B->d_ch_seg.p; T.ch_ptr = B->d_ch_ptr.p;
    T.chs_r0 = B->d_chs_r0.p; T.chs_seg = B->d_chs_seg.p; T.chs_ptr = B->d_chs_ptr.p;
    T.active = B->d_active.p;
    SegTable Tc = T;  // members converged at the current check
    Tc.active = B->d_conv.p;
    SegTable Ti = T;  // members outside the iteration
    Ti.active = B->d_idle.p;
    const int n = ntot;
    const int ldT = mmax + b;
    DevBuf<double> &V = B->V, &Wb = B->Wb, &xf = B->xf, &yf = B->yf, &Tm = B->Tm, &A = B->A, &Zs = B->Zs,
                   &thd = B->thd, &ework = B->ework, &part = B->part, &G1 = B->G1, &G2 = B->G2, &G0 = B->G0,
                   &Rt = B->Rt, &H2 = B->H2, &Y = B->Y, &zl = B->zl, &rl = B->rl;
    DevBuf<int> &ierr = B->ierr, &info = B->info, &dlist = B->dlist;
    int nact = 0;
    for (int i = 0; i < nm; i++) nact += E[i].active;
    const int kdim = std::max(kmax, 1);
    HIPCHK(Y.alloc((size_t)std::max(n, 1) * MAXR));
    dev_fill(Y.p, (int64_t)std::max(n, 1) * MAXR, 0.0, s);
    if (nact > 0) {
        HIPCHK(V.alloc((size_t)n * (mmax + b)));
        HIPCHK(Wb.alloc((size_t)n * b));
        HIPCHK(xf.alloc((size_t)ndof * b));
        HIPCHK(yf.alloc((size_t)ndof * b));
        HIPCHK(Tm.alloc((size_t)nm * ldT * ldT));
        HIPCHK(A.alloc((size_t)nm * mmax * mmax));
        HIPCHK(Zs.alloc((size_t)nm * mmax * kdim));
        HIPCHK(thd.alloc((size_t)nm * kdim));
        HIPCHK(ework.alloc((size_t)nm * symeig_work_size(mmax, kdim)));
        HIPCHK(part.alloc(std::max<int64_t>(
            std::max(seg_gemm_tn_part_size(T.n_ch, T.n_chs, mmax + b, b), seg_gemm_nn_tn_part_size(T.n_ch, mmax + b, b)),
            1)));
        HIPCHK(G1.alloc((size_t)nm * 256));
        HIPCHK(G2.alloc((size_t)nm * 256));
        HIPCHK(G0.alloc((size_t)nm * 256));
        HIPCHK(Rt.alloc((size_t)nm * 256));
        HIPCHK(H2.alloc((size_t)nm * (mmax + b) * b));
        HIPCHK(zl.alloc((size_t)nm * b * kdim));
        HIPCHK(rl.alloc((size_t)nm * b * b));
        HIPCHK(ierr.alloc(2 * (size_t)nm));
        HIPCHK(info.alloc(nm));
        HIPCHK(dlist.alloc(nm));
        if (kmax > SYMEIG_MAX_K || mmax > SYMEIG_MAX_M)
            return set_err(ctx, SOL105_E_UNSUPPORTED, "Rayleigh-Ritz size out of range (k=%d, m=%d)", kmax, mmax);
        dev_fill(Tm.p, (int64_t)nm * ldT * ldT, 0.0, s);
        HIPCHK(hipMemsetAsync(ierr.p, 0, sizeof(int) * 2 * nm, s));
    }
    const int64_t sT = (int64_t)ldT * ldT, s16 = 256, sH2 = (int64_t)(mmax + b) * b;
    auto blk = [&](int jj) { return V.p + (size_t)jj * b * n; };
    // CholQR2 per segment; g1_ready: G1 = Q^T Q was formed by the kernel
    // that wrote Q.  The first triangular solve forms the second pass's Gram
    // matrix, the second one R = R2 R1.
    auto cholqr2 = [&](double *Q, double *Rout, int64_t sR, int ldr, const double *g0, bool g1_ready) {
        if (!g1_ready) seg_gemm_tn(T, b, b, Q, n, Q, n, G1.p, s16, 16, 1.0, 0.0, part.p, s);
        seg_chol_upper(T, b, G1.p, s16, 16, g0, s16, 16, g0 ? dep_tol2 : 0.0, ierr.p, s);
        seg_trsm_right_upper_gram(T, b, Q, n, G1.p, s16, 16, part.p, G2.p, s16, 16, s);
        seg_chol_upper(T, b, G2.p, s16, 16, nullptr, 0, 0, 0.0, ierr.p, s);
        if (Rout) seg_trsm_right_upper_mul(T, b, Q, n, G2.p, s16, 16, G1.p, s16, 16, Rout, sR, ldr, s);
        else seg_trsm_right_upper(T, b, Q, n, G2.p, s16, 16, s);
    };
    auto step = [&](int jj) {
        HIPCHK(solve_bwd(M->plan, M->dev, M->dsolve, M->splan, M->d_PT.p, M->d_W.p, M->d_part.p, blk(jj), Wb.p, b, s));
        launch_scatter_elim_il(n, ndof, b, M->d_elim_of_dof.p, Wb.p, xf.p, s);
        launch_bsr_spmv(M->n_nodes, M->d_rp.p, M->d_col.p, M->d_KG.p, xf.p, yf.p, b, -1.0, s);
        launch_gather_elim_il(n, b, M->d_dof_of_elim.p, yf.p, Wb.p, s);
        // into column block jj+1 of V, which the products of this step do not read
        double *Wn = blk(jj + 1);
        HIPCHK(solve_fwd(M->plan, M->dev, M->dsolve, M->splan, M->d_P.p, M->d_W.p, M->d_part.p, Wb.p, Wn, b, s));
        int kcols = (jj + 1) * b;
        double *Tcol = Tm.p + (size_t)jj * b * ldT;
        // first pass: coefficients into T, update; G0 = Gram matrix after it
        seg_gemm_tn(T, kcols, b, V.p, n, Wn, n, Tcol, sT, ldT, 1.0, 0.0, part.p, s);
        // ... and the second pass's coefficients (also added into T), in the
        // same pass over V as the update
        seg_gemm_nn_gram_tn(T, kcols, b, V.p, n, Tcol, sT, ldT, Wn, n, part.p, G0.p, s16, 16, H2.p, sH2, kcols, Tcol, sT,
                            ldT, s);
        // second pass update; G1 = Gram matrix of the result (the first
        // Cholesky-QR pass of the new block)
        seg_gemm_nn_gram(T, kcols, b, V.p, n, H2.p, sH2, kcols, Wn, n, -1.0, 1.0, part.p, G1.p, s16, 16, s);
        cholqr2(blk(jj + 1), Tm.p + (size_t)(jj + 1) * b + (size_t)jj * b * ldT, sT, ldT, G0.p, true);
        return SOL105_OK;
    };
    // step graphs (as for a single model): the key covers the solve state,
    // the workspace and segment-table addresses and the stacked sizes; the
    // members' activity flags are read by the kernels at run time
    StepGraphs &G = B->graphs;
    const bool use_graphs = ctx->graphs && !ctx->profiling && nact > 0;
    if (use_graphs) {
        KeyHash k;
        hash_solve_state(k, M);
        for (const void *p : {(const void *)V.p, (const void *)Wb.p, (const void *)xf.p, (const void *)yf.p,
                              (const void *)Tm.p, (const void *)part.p, (const void *)G0.p, (const void *)G1.p,
                              (const void *)G2.p, (const void *)H2.p, (const void *)ierr.p, (const void *)T.seg_off,
                              (const void *)T.ch_r0, (const void *)T.ch_seg, (const void *)T.ch_ptr,
                              (const void *)T.chs_r0, (const void *)T.chs_seg, (const void *)T.chs_ptr,
                              (const void *)T.active})
            k.add(p);
        for (int v : {T.n_seg, T.n_ch, T.n_chs, n, ndof, b, ldT, mmax}) k.add(v);
        if (G.key != k.h) {
            G.clear();
            G.key = k.h;
        } else {
            G.armed = true;
        }
    }
    auto run_step = [&](int jj) {
        if (!use_graphs || !G.armed) return step(jj);
        if ((size_t)jj >= G.exec.size()) G.exec.resize(jj + 1, nullptr);
        if (!G.exec[jj]) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = step(jj);
            hipError_t e = hipStreamEndCapture(s, &g);
            hipGraphExec_t ex = nullptr;
            if (rc == SOL105_OK && e == hipSuccess) e = hipGraphInstantiateWithFlags(&ex, g, 0);
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
            HIPCHK(e);
            G.exec[jj] = ex;
        }
        HIPCHK(hipGraphLaunch(G.exec[jj], s));
        return (int)SOL105_OK;
    };
    int j = 0, m = 0;
    if (nact > 0) {
        seg_start_block(T, b, V.p, n, 0, s);
        seg_fill(Ti, b, V.p, n, 0.0, s);
        cholqr2(V.p, nullptr, 0, 0, nullptr, false);
    }
    int next_check = std::min(mmax, std::max(2 * b, ((kmax * 3 + b - 1) / b) * b));
    std::vector<double> hth, hzl, hrl;
    std::vector<int> herr, hinfo, list;
    while (nact > 0) {
        int rc = run_step(j);
        if (rc) return rc;
        j++;
        m = j * b;
        bool due = m >= next_check;
        for (int i = 0; i < nm; i++)
            if (E[i].active && m >= E[i].mmax) due = true;
        if (!due) continue;
        list.clear();
        for (int i = 0; i < nm; i++)
            if (E[i].active) list.push_back(i);
        const int kchk = std::min(kmax, m);
        HIPCHK(hipMemcpyAsync(dlist.p, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, s));
        seg_symmetrize_copy(T, m, ldT, Tm.p, sT, A.p, (int64_t)mmax * mmax, s);
        sym_eig_top_batched(m, A.p, (int64_t)mmax * mmax, kchk, thd.p, kdim, Zs.p, (int64_t)mmax * kdim, ework.p,
                            symeig_work_size(mmax, kdim), info.p, dlist.p, (int)list.size(), s, b);
        // last block row of the Ritz vectors and the trailing coupling block R_j
        seg_copy(T, b, kchk, Zs.p + (m - b), (int64_t)mmax * kdim, m, zl.p, (int64_t)b * kdim, b, s);
        seg_copy(T, b, b, Tm.p + (size_t)m + (size_t)(m - b) * ldT, sT, ldT, rl.p, (int64_t)b * b, b, s);
        HIPCHK(hipGetLastError());
        hth.resize((size_t)nm * kdim);
        hzl.resize((size_t)nm * b * kdim);
        hrl.resize((size_t)nm * b * b);
        herr.resize(2 * (size_t)nm);
        hinfo.resize(nm);
        HIPCHK(hipMemcpyAsync(hth.data(), thd.p, sizeof(double) * hth.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(hzl.data(), zl.p, sizeof(double) * hzl.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(hrl.data(), rl.p, sizeof(double) * hrl.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(herr.data(), ierr.p, sizeof(int) * herr.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(hinfo.data(), info.p, sizeof(int) * hinfo.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::fill(conv.begin(), conv.end(), 0);
        bool any_conv = false, changed = false;
        for (int i : list) {
            MemberEig &e = E[i];
            if ((herr[2 * i] & 6) || hinfo[i] != 0) {
                // rank-deficient block or breakdown: re-solved alone afterwards
                e.active = false;
                e.fallback = true;
                changed = true;
                continue;
            }
            const double *th = &hth[(size_t)i * kdim];
            const double *Z = &hzl[(size_t)i * b * kdim];
            const double *R = &hrl[(size_t)i * b * b];
            auto resid = [&](int t) {
                double r2 = 0.0;
                for (int p = 0; p < b; p++) {
                    double v = 0.0;
                    for (int q = 0; q < b; q++) v += R[(size_t)q * b + p] * Z[(size_t)t * b + q];
                    r2 += v * v;
                }
                return std::sqrt(r2);
            };
            const int kc = std::min(e.kneed, m);
            bool ok = true;
            for (int t = 0; t < kc; t++) {
                double r = resid(t);
                if (!std::isfinite(th[t]) || !std::isfinite(r)) { ok = false; break; }
                if (std::fabs(th[t]) < 1e-300) continue;
                if (r > tol * std::fabs(th[t])) { ok = false; break; }
            }
            if (ok) {
                int k = std::min(e.req, kc);
                e.theta.assign(th, th + k);
                conv[i] = 1;
                any_conv = true;
                e.active = false;
                changed = true;
            } else if (m >= e.mmax) {
                e.active = false;
                e.fallback = true;
                changed = true;
            }
        }
        if (any_conv) {
            // Ritz vectors of the members that converged now: Y_s = V_s Z_s
            HIPCHK(hipMemcpyAsync(B->d_conv.p, conv.data(), nm, hipMemcpyHostToDevice, s));
            seg_gemm_nn(Tc, m, std::min(kchk, MAXR), V.p, n, Zs.p, (int64_t)mmax * kdim, m, Y.p, n, 1.0, 0.0, s);
        }
        if (changed) {
            // finished members leave the iteration: their rows of the newest
            // block are zeroed so the stacked operator keeps them at zero
            for (int i = 0; i < nm; i++) {
                act[i] = E[i].active;
                idle[i] = !E[i].active;
            }
            HIPCHK(hipMemcpyAsync(B->d_active.p, act.data(), nm, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(B->d_idle.p, idle.data(), nm, hipMemcpyHostToDevice, s));
            seg_fill(Ti, b, blk(j), n, 0.0, s);
        }
        nact = 0;
        for (int i = 0; i < nm; i++) nact += E[i].active;
        next_check = m + 2 * b;
        if (m >= mmax) break;
    }
    for (int i = 0; i < nm; i++)
        if (E[i].active) { E[i].active = false; E[i].fallback = true; }
    B->last_m = m;
    B->last_steps = j;
    // x = L^-T y for every member at once, expanded to all DOFs
    int kcols = 0;
    for (int i = 0; i < nm; i++) kcols = std::max(kcols, (int)E[i].theta.size());
    DevBuf<double> Xf;
    HIPCHK(Xf.alloc((size_t)ndof * std::max(kcols, 1)));
    if (kcols > 0) {
        DevBuf<double> Ys;
        HIPCHK(Ys.alloc((size_t)n * kcols));
        HIPCHK(solve_bwd(M->plan, M->dev, M->dsolve, M->splan, M->d_PT.p, M->d_W.p, M->d_part.p, Y.p, Ys.p, kcols, s));
        launch_scatter_elim(n, ndof, kcols, M->d_elim_of_dof.p, Ys.p, Xf.p, s);
    }
    // eigenvalue semantics per member
    std::vector<std::vector<int>> pos(nm);
    std::vector<std::vector<double>> lam(nm);
    for (int i = 0; i < nm; i++) {
        MemberEig &e = E[i];
        pos[i] = select_candidates(e.sel, e.theta.data(), (int)e.theta.size(), lam[i]);
    }
    // localization filter (>= 100 shells per member, share 0.10, scan ND + 16)
    std::vector<int32_t> pairs;
    std::vector<int> pair_mem, pair_t;
    for (int i = 0; i < nm; i++) {
        const int limit = localization_scan_limit(E[i].sel, (int)pos[i].size());
        if (limit == 0) continue;
        for (int t = 0; t < limit; t++) {
            pairs.push_back(i);
            pairs.push_back(pos[i][t]);
            pair_mem.push_back(i);
            pair_t.push_back(t);
        }
    }
    std::vector<std::vector<int>> surv(nm);
    for (int i = 0; i < nm; i++) surv[i] = pos[i];
    if (!pairs.empty()) {
        int np = (int)pairs.size() / 2;
        DevBuf<int32_t> dp;
        DevBuf<double> sh;
        HIPCHK(dp.upload(pairs.data(), pairs.size(), s));
        HIPCHK(sh.alloc(2 * (size_t)np));
        HIPCHK(B->d_q4_off.upload(B->q4_off.data(), B->q4_off.size(), s));
        HIPCHK(B->d_t3_off.upload(B->t3_off.data(), B->t3_off.size(), s));
        seg_energy_share(np, dp.p, B->d_q4_off.p, M->d_q4.p, B->d_t3_off.p, M->d_t3.p, ndof, Xf.p, sh.p, s);
        std::vector<double> hsh(2 * (size_t)np);
        HIPCHK(hipMemcpyAsync(hsh.data(), sh.p, hsh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<std::vector<uint8_t>> drop(nm);
        for (int i = 0; i < nm; i++) drop[i].assign(pos[i].size(), 0);
        for (int p = 0; p < np; p++) {
            double mx = hsh[2 * p], tot = hsh[2 * p + 1];
            if (tot > 0 && mx / tot > 0.10) drop[pair_mem[p]][pair_t[p]] = 1;
        }
        for (int i = 0; i < nm; i++) {
            surv[i].clear();
            for (size_t t = 0; t < pos[i].size(); t++)
                if (!drop[i][t]) surv[i].push_back(pos[i][t]);
        }
    }
    // outputs: normalised modes of every member, contiguous per member
    B->lambda.assign(nm, {});
    B->n_out.assign(nm, 0);
    B->status.assign(nm, SOL105_OK);
    B->mode_off.assign(nm + 1, 0);
    for (int i = 0; i < nm; i++) {
        int no = 0;
        if (!E[i].fallback && E[i].code == SOL105_OK) {
            surv[i] = select_output(E[i].sel, surv[i], lam[i]);
            no = std::min((int)surv[i].size(), max_modes);
        }
        B->n_out[i] = no;
        B->mode_off[i + 1] = B->mode_off[i] + (int64_t)6 * (B->node_off[i + 1] - B->node_off[i]) *
                                                  std::max(no, E[i].fallback ? std::min(max_modes, SYMEIG_MAX_K) : 0);
    }
    HIPCHK(B->d_modes.alloc(std::max<int64_t>(B->mode_off[nm], 1)));
    std::vector<int32_t> npairs;
    std::vector<int64_t> dsto, dofo(nm + 1);
    for (int i = 0; i <= nm; i++) dofo[i] = 6 * (int64_t)B->node_off[i];
    for (int i = 0; i < nm; i++) {
        B->lambda[i].resize(B->n_out[i]);
        for (int t = 0; t < B->n_out[i]; t++) {
            int c = surv[i][t];
            B->lambda[i][t] = lam[i][c];
            npairs.push_back(i);
            npairs.push_back(c);
            dsto.push_back(B->mode_off[i] + (int64_t)t * (dofo[i + 1] - dofo[i]));
        }
        if (E[i].code != SOL105_OK) B->status[i] = E[i].code;
    }
    if (!npairs.empty()) {
        DevBuf<int32_t> dp;
        DevBuf<int64_t> ddof, ddst;
        HIPCHK(dp.upload(npairs.data(), npairs.size(), s));
        HIPCHK(ddof.upload(dofo.data(), dofo.size(), s));
        HIPCHK(ddst.upload(dsto.data(), dsto.size(), s));
        seg_normalize((int)npairs.size() / 2, dp.p, ddof.p, ddst.p, ndof, Xf.p, B->d_modes.p, s);
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipGetLastError());
    // members the stacked iteration could not finish: solved alone
    B->n_fallback = 0;
    for (int i = 0; i < nm; i++) {
        if (!E[i].fallback) continue;
        B->n_fallback++;
        MemberDesc D;
        member_desc(B, i, D);
        sol105_model Mi;
        int rc = model_init(ctx, &D.d, &Mi);
        int32_t no = 0;
        std::vector<double> l(std::min(max_modes, SYMEIG_MAX_K));
        if (rc == SOL105_OK) rc = sol105_solve(ctx, &Mi, (int32_t)l.size(), l.data(), nullptr, &no);
        B->status[i] = rc;
        if (rc != SOL105_OK) continue;
        B->n_out[i] = no;
        B->lambda[i].assign(l.begin(), l.begin() + no);
        HIPCHK(hipMemcpyAsync(B->d_modes.p + B->mode_off[i], Mi.d_modes.p, sizeof(double) * Mi.ndof() * no,
                              hipMemcpyDeviceToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return SOL105_OK;
}

}  // namespace

extern "C" {

int sol105_batch_create(sol105_ctx *ctx, const sol105_model_desc *mem, int32_t n_members, sol105_batch **out) {
    if (!ctx || !mem || !out || n_members <= 0) return set_err(ctx, SOL105_E_INVALID, "null argument or empty batch");
    *out = nullptr;
    auto B = std::make_unique<sol105_batch>();
    B->n_mem = n_members;
    B->node_off.assign(n_members + 1, 0);
    B->q4_off.assign(n_members + 1, 0);
    B->t3_off.assign(n_members + 1, 0);
    B->bar_off.assign(n_members + 1, 0);
    B->nd.resize(n_members);
    B->v1.resize(n_members);
    B->v2.resize(n_members);
    B->n_shell_cards.resize(n_members);
    for (int i = 0; i < n_members; i++) {
        const sol105_model_desc &d = mem[i];
        if (d.n_nodes <= 0 || d.n_q4 < 0 || d.n_t3 < 0 || d.n_bar < 0 || d.nd <= 0)
            return set_err(ctx, SOL105_E_INVALID, "batch member %d: invalid sizes", i);
        if (d.k6rot != mem[0].k6rot || (d.autospc != 0) != (mem[0].autospc != 0))
            return set_err(ctx, SOL105_E_UNSUPPORTED, "batch members must share PARAM K6ROT and AUTOSPC");
        if (d.n_rbe2 > 0 || d.n_rbe3 > 0)
            return set_err(ctx, SOL105_E_UNSUPPORTED, "batch member %d: RBE2/RBE3 constraints are solved alone", i);
        B->node_off[i + 1] = B->node_off[i] + d.n_nodes;
        B->q4_off[i + 1] = B->q4_off[i] + d.n_q4;
        B->t3_off[i + 1] = B->t3_off[i] + d.n_t3;
        B->bar_off[i + 1] = B->bar_off[i] + d.n_bar;
        if (!std::isfinite(d.eigrl_v1) || !std::isfinite(d.eigrl_v2) || d.n_shell_cards < 0)
            return set_err(ctx, SOL105_E_INVALID, "batch member %d: invalid EIGRL range or shell card count", i);
        B->nd[i] = d.nd;
        B->v1[i] = d.eigrl_v1;
        B->v2[i] = d.eigrl_v2;
        B->n_shell_cards[i] = d.n_shell_cards > 0 ? d.n_shell_cards : d.n_q4 + d.n_t3;
    }
    const int N = B->node_off[n_members];
    std::vector<double> xyz(3 * (size_t)N), q4p(6 * (size_t)B->q4_off[n_members]), t3p(6 * (size_t)B->t3_off[n_members]);
    std::vector<double> vref(3 * (size_t)B->bar_off[n_members]), barp(8 * (size_t)B->bar_off[n_members]);
    std::vector<double> force(6 * (size_t)N);
    std::vector<int32_t> q4(4 * (size_t)B->q4_off[n_members]), t3(3 * (size_t)B->t3_off[n_members]);
    std::vector<int32_t> bar(2 * (size_t)B->bar_off[n_members]);
    std::vector<uint8_t> spc(6 * (size_t)N), hs(N), hb(N);
    int nd_max = 1;
    for (int i = 0; i < n_members; i++) {
        const sol105_model_desc &d = mem[i];
        const int n0 = B->node_off[i];
        if (!d.xyz || !d.spc_mask || !d.force || !d.node_has_shell || !d.node_has_bar ||
            (d.n_q4 > 0 && (!d.q4_nodes || !d.q4_prop)) || (d.n_t3 > 0 && (!d.t3_nodes || !d.t3_prop)) ||
            (d.n_bar > 0 && (!d.bar_nodes || !d.bar_prop || !d.bar_vref)))
            return set_err(ctx, SOL105_E_INVALID, "batch member %d: missing arrays", i);
        auto shift = [&](const int32_t *src, int count, int32_t *dst) {
            for (int k = 0; k < count; k++) {
                if (src[k] < 0 || src[k] >= d.n_nodes) return false;
                dst[k] = src[k] + n0;
            }
            return true;
        };
        std::copy(d.xyz, d.xyz + 3 * (size_t)d.n_nodes, xyz.begin() + 3 * (size_t)n0);
        std::copy(d.force, d.force + 6 * (size_t)d.n_nodes, force.begin() + 6 * (size_t)n0);
        std::copy(d.spc_mask, d.spc_mask + 6 * (size_t)d.n_nodes, spc.begin() + 6 * (size_t)n0);
        std::copy(d.node_has_shell, d.node_has_shell + d.n_nodes, hs.begin() + n0);
        std::copy(d.node_has_bar, d.node_has_bar + d.n_nodes, hb.begin() + n0);
        if (!shift(d.q4_nodes, 4 * d.n_q4, q4.data() + 4 * (size_t)B->q4_off[i]) ||
            !shift(d.t3_nodes, 3 * d.n_t3, t3.data() + 3 * (size_t)B->t3_off[i]) ||
            !shift(d.bar_nodes, 2 * d.n_bar, bar.data() + 2 * (size_t)B->bar_off[i]))
            return set_err(ctx, SOL105_E_INVALID, "batch member %d: node rank out of range", i);
        if (d.n_q4) std::copy(d.q4_prop, d.q4_prop + 6 * (size_t)d.n_q4, q4p.begin() + 6 * (size_t)B->q4_off[i]);
        if (d.n_t3) std::copy(d.t3_prop, d.t3_prop + 6 * (size_t)d.n_t3, t3p.begin() + 6 * (size_t)B->t3_off[i]);
        if (d.n_bar) {
            std::copy(d.bar_vref, d.bar_vref + 3 * (size_t)d.n_bar, vref.begin() + 3 * (size_t)B->bar_off[i]);
            std::copy(d.bar_prop, d.bar_prop + 8 * (size_t)d.n_bar, barp.begin() + 8 * (size_t)B->bar_off[i]);
        }
        nd_max = std::max(nd_max, d.nd);
    }
    sol105_model_desc u{};
    u.n_nodes = N;
    u.xyz = xyz.data();
    u.n_q4 = B->q4_off[n_members];
    u.q4_nodes = q4.data();
    u.q4_prop = q4p.data();
    u.n_t3 = B->t3_off[n_members];
    u.t3_nodes = t3.data();
    u.t3_prop = t3p.data();
    u.n_bar = B->bar_off[n_members];
    u.bar_nodes = bar.data();
    u.bar_vref = vref.data();
    u.bar_prop = barp.data();
    u.spc_mask = spc.data();
    u.force = force.data();
    u.node_has_shell = hs.data();
    u.node_has_bar = hb.data();
    u.k6rot = mem[0].k6rot;
    u.autospc = mem[0].autospc;
    u.nd = nd_max;
    int rc = model_init(ctx, &u, &B->M);
    if (rc) return rc;
    B->M.mem_node_off = B->node_off;
    *out = B.release();
    return SOL105_OK;
}

int sol105_batch_free(sol105_ctx *ctx, sol105_batch *batch) {
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    delete batch;
    return SOL105_OK;
}

int sol105_batch_solve(sol105_ctx *ctx, sol105_batch *B, int32_t max_modes, double *lambda_out, int32_t *n_out,
                       int32_t *status) {
    if (!ctx || !B || max_modes <= 0) return set_err(ctx, SOL105_E_INVALID, "null argument");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    HIPCHK(hipEventRecord(ctx->ev[7], s));  // events 0-6 belong to sol105_solve (member fallbacks)
    sol105_model *M = &B->M;
    int rc = sol105_assemble_K(ctx, M, nullptr);
    if (rc == SOL105_OK) rc = factorize(ctx, M, false);  // a failed stack: members re-solved alone
    if (rc == SOL105_OK) rc = sol105_static_solve(ctx, M, nullptr, nullptr);
    if (rc == SOL105_OK) rc = sol105_assemble_KG(ctx, M, nullptr, nullptr);
    if (rc == SOL105_OK) {
        rc = batch_buckling(ctx, B, max_modes);
        if (rc) return rc;
    } else if (rc == SOL105_E_SINGULAR || rc == SOL105_E_NOCONV) {
        // a member made the stacked system singular: every member alone
        const int nm = B->n_mem;
        B->lambda.assign(nm, {});
        B->n_out.assign(nm, 0);
        B->status.assign(nm, SOL105_OK);
        B->mode_off.assign(nm + 1, 0);
        const int mm = std::min(max_modes, SYMEIG_MAX_K);
        for (int i = 0; i < nm; i++) B->mode_off[i + 1] = B->mode_off[i] + 6 * (int64_t)(B->node_off[i + 1] - B->node_off[i]) * mm;
        HIPCHK(B->d_modes.alloc(std::max<int64_t>(B->mode_off[nm], 1)));
        B->n_fallback = nm;
        for (int i = 0; i < nm; i++) {
            MemberDesc D;
            member_desc(B, i, D);
            sol105_model Mi;
            int r = model_init(ctx, &D.d, &Mi);
            int32_t no = 0;
            std::vector<double> l(mm);
            if (r == SOL105_OK) r = sol105_solve(ctx, &Mi, mm, l.data(), nullptr, &no);
            B->status[i] = r;
            if (r != SOL105_OK) continue;
            B->n_out[i] = no;
            B->lambda[i].assign(l.begin(), l.begin() + no);
            HIPCHK(hipMemcpyAsync(B->d_modes.p + B->mode_off[i], Mi.d_modes.p, sizeof(double) * Mi.ndof() * no,
                                  hipMemcpyDeviceToDevice, s));
            HIPCHK(hipStreamSynchronize(s));
        }
    } else {
        return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev[8], s));
    HIPCHK(hipEventSynchronize(ctx->ev[8]));
    B->last_ms = ev_ms(ctx->ev[7], ctx->ev[8]);
    for (int i = 0; i < B->n_mem; i++) {
        if (n_out) n_out[i] = B->n_out[i];
        if (status) status[i] = B->status[i];
        if (lambda_out)
            for (int t = 0; t < max_modes; t++)
                lambda_out[(size_t)i * max_modes + t] = t < B->n_out[i] ? B->lambda[i][t] : 0.0;
    }
    return SOL105_OK;
}

int sol105_batch_modes(sol105_ctx *ctx, const sol105_batch *B, int32_t member, double *modes_out) {
    if (!ctx || !B || !modes_out) return set_err(ctx, SOL105_E_INVALID, "null argument");
    if (member < 0 || member >= B->n_mem || (int)B->n_out.size() != B->n_mem)
        return set_err(ctx, SOL105_E_STATE, "member %d has no results", member);
    (void)hipSetDevice(ctx->device);
    size_t cnt = (size_t)6 * (B->node_off[member + 1] - B->node_off[member]) * B->n_out[member];
    if (cnt == 0) return SOL105_OK;
    HIPCHK(hipMemcpy(modes_out, B->d_modes.p + B->mode_off[member], cnt * sizeof(double), hipMemcpyDeviceToHost));
    return SOL105_OK;
}

sol105_model *sol105_batch_model(sol105_batch *B) { return B ? &B->M : nullptr; }

int sol105_batch_autotune(sol105_ctx *ctx, sol105_batch *B, int32_t reps, double *report) {
    if (!ctx || !B) return set_err(ctx, SOL105_E_INVALID, "null argument");
    return sol105_solve_autotune(ctx, &B->M, reps, report);
}

int sol105_batch_stats(const sol105_ctx *ctx, const sol105_batch *B, double *out) {
    (void)ctx;
    if (!B || !out) return SOL105_E_INVALID;
    out[0] = B->last_ms;
    out[1] = B->last_m;
    out[2] = B->last_steps;
    out[3] = B->n_fallback;
    out[4] = B->M.sym.n_free;
    out[5] = (double)B->M.sym.front_storage * 8.0;
    out[6] = B->M.sym.flops;
    out[7] = B->M.sym.height;
    return SOL105_OK;
}

}  // extern "C"
