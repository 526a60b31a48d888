// api.cpp -- the C ABI (include/sol105.h): context, model upload, and the
// orchestration of one SOL 105 solve on a single HIP stream.
//
// Phase map to the reference:
//   sol105_assemble_K    Solver.assemble_stiffness     JFEM/src/solver/assembly.jl:1953-4375
//   sol105_static_solve  Solver.apply_bc_and_solve     JFEM/src/solver/boundary_conditions.jl:530-1012
//   sol105_assemble_KG   Solver.assemble_geometric_stiffness  JFEM/src/solver/assembly.jl:4382-6635
//   sol105_buckling      Solver.solve_buckling         JFEM/src/solver/solve_case.jl:1863-3205
//   sol105_sensitivity_thickness  solve_adjoint_buckling  JFEM/src/solver/adjoint_buckling.jl:2406-2725
#include "internal.h"
#include "select.h"

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>

using namespace s105;

namespace s105 {

int set_err(sol105_ctx *ctx, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

ElemArgs elem_args(sol105_model *M) {
    ElemArgs a;
    a.n_q4 = M->n_q4;
    a.n_t3 = M->n_t3;
    a.n_bar = M->n_bar;
    a.xyz = M->d_xyz.p;
    a.q4 = M->d_q4.p;
    a.q4prop = M->d_q4prop.p;
    a.t3 = M->d_t3.p;
    a.t3prop = M->d_t3prop.p;
    a.bar = M->d_bar.p;
    a.bar_vref = M->d_bar_vref.p;
    a.barprop = M->d_barprop.p;
    a.k6rot = M->k6rot;
    return a;
}

// AUTOSPC on the diagonal (JFEM/src/solver/boundary_conditions.jl:296-383);
// the translational / rotational maxima are those of each member's own K.
void autospc_mask(const sol105_model *M, const std::vector<double> &kd, std::vector<uint8_t> &free_out) {
    int ndof = M->ndof();
    std::vector<uint8_t> fixed(M->spc_mask.begin(), M->spc_mask.end());
    std::vector<int32_t> off = M->mem_node_off;
    if (off.empty()) off = {0, M->n_nodes};
    for (size_t mb = 0; M->autospc && mb + 1 < off.size(); mb++) {
        int i0 = 6 * off[mb], i1 = 6 * off[mb + 1];
        double max_t = 1.0, max_r = 1.0;
        for (int i = i0; i < i1; i++) {
            if (i % 6 < 3) max_t = std::max(max_t, std::fabs(kd[i]));
            else max_r = std::max(max_r, std::fabs(kd[i]));
        }
        for (int i = i0; i < i1; i++) {
            if (fixed[i]) continue;
            int c = i % 6, v = i / 6;
            double th;
            if (c < 3) th = 1e-8 * std::max(max_t, 1.0);
            else {
                double mult = 1.0;
                bool hs = M->has_shell[v], hb = M->has_bar[v];
                if (hs && hb) mult = 0.1;
                else if (hs) mult = 4.0;
                else if (hb) mult = 0.25;
                th = 1e-8 * std::max(max_r, 1.0) * mult;
            }
            if (std::fabs(kd[i]) < th || kd[i] < 0) fixed[i] = 1;
        }
    }
    free_out.resize(ndof);
    for (int i = 0; i < ndof; i++) free_out[i] = fixed[i] ? 0 : 1;
}

void hash_solve_state(KeyHash &k, const sol105_model *M) {
    static_assert(sizeof(DevFronts) == 15 * sizeof(void *), "DevFronts holds device pointers only");
    k.add(M->dev);
    const DevSolve &DS = M->dsolve;
    for (const void *p : {(const void *)DS.poff, (const void *)DS.fwd_desc, (const void *)DS.bwd_desc,
                          (const void *)DS.fwd_tiles, (const void *)DS.bwd_tiles, (const void *)DS.fwd_tbase_rel,
                          (const void *)DS.bwd_tbase_rel, (const void *)DS.fwd_nks, (const void *)DS.bwd_nks,
                          (const void *)DS.fwd_kw, (const void *)DS.bwd_kw, (const void *)DS.pull_ptr,
                          (const void *)DS.pull_src, (const void *)DS.lblk, (const void *)DS.pull2})
        k.add(p);
    k.add(DS.n_lblk);
    const FactorPlan &P = M->plan;
    k.add(P.n_free);
    k.add(P.height);
    for (const HeightInfo &H : P.heights) {
        k.add(H.n);
        k.add(H.list_off);
        k.add(H.max_ns);
        k.add(H.max_m);
    }
    for (const SolveHeight &S : M->splan.heights) {
        k.add(S.fwd_off);
        k.add(S.bwd_off);
        k.add(S.n_fwd);
        k.add(S.n_bwd);
        k.add(S.max_m);
        k.add(S.max_ns);
        k.add(S.fwd_split);
        k.add(S.bwd_split);
        k.add(S.fwd_kmax);
        k.add(S.bwd_kmax);
        for (const ProdCfg *c : {&S.fwd_cfg, &S.bwd_cfg}) {
            k.add(c->kmax);
            k.add(c->ks);
            k.add(c->pch);
            k.add(c->kg);
        }
    }
    for (const void *p : {(const void *)M->d_P.p, (const void *)M->d_PT.p, (const void *)M->d_W.p,
                          (const void *)M->d_part.p, (const void *)M->d_KG.p, (const void *)M->d_rp.p,
                          (const void *)M->d_col.p, (const void *)M->d_elim_of_dof.p, (const void *)M->d_dof_of_elim.p})
        k.add(p);
    k.add(M->n_nodes);
    k.add(M->sym_gen);
}

// Device copies of the tile-dependent parts of the solve plan (tiles, their
// descriptors, per-front slice data, the partial-sum buffer): at the symbolic
// analysis and whenever the product launch shapes change.
int upload_solve_tiles(sol105_ctx *ctx, sol105_model *M) {
    hipStream_t s = ctx->stream;
    const SolvePlan &SP = M->splan;
    const Symbolic &S = M->sym;
    HIPCHK(M->s_ftiles.upload(SP.fwd_tiles.data(), SP.fwd_tiles.size(), s));
    HIPCHK(M->s_btiles.upload(SP.bwd_tiles.data(), SP.bwd_tiles.size(), s));
    HIPCHK(M->s_ftbase.upload(SP.fwd_tbase.data(), SP.fwd_tbase.size(), s));
    HIPCHK(M->s_btbase.upload(SP.bwd_tbase.data(), SP.bwd_tbase.size(), s));
    HIPCHK(M->s_fnks.upload(SP.fwd_nks.data(), SP.fwd_nks.size(), s));
    HIPCHK(M->s_bnks.upload(SP.bwd_nks.data(), SP.bwd_nks.size(), s));
    HIPCHK(M->s_fkw.upload(SP.fwd_kw.data(), SP.fwd_kw.size(), s));
    HIPCHK(M->s_bkw.upload(SP.bwd_kw.data(), SP.bwd_kw.size(), s));
    // tile descriptors (the factor offset is that of P and of PT alike)
    auto build = [&](const std::vector<int32_t> &tl, const std::vector<int32_t> &kw, const std::vector<int32_t> &nks,
                     std::vector<TileDesc> &out) {
        out.resize(tl.size() / 3);
        for (size_t t = 0; t < out.size(); t++) {
            const int f = tl[3 * t];
            const FrontInfo &F = S.fronts[f];
            TileDesc &d = out[t];
            d.f = f; d.rt = tl[3 * t + 1]; d.ks = tl[3 * t + 2];
            d.ns = F.ns; d.m = F.m; d.piv0 = F.piv0; d.kw = kw[f]; d.nks = nks[f];
            d.parent = S.children[f].empty() ? 0 : 1;
            d.pad = 0;
            d.vec_off = F.vec_off; d.poff = SP.poff[f]; d.row_off = F.row_off;
        }
    };
    std::vector<TileDesc> fd, bd;
    build(SP.fwd_tiles, SP.fwd_kw, SP.fwd_nks, fd);
    build(SP.bwd_tiles, SP.bwd_kw, SP.bwd_nks, bd);
    HIPCHK(M->s_fdesc.upload(fd.data(), fd.size(), s));
    HIPCHK(M->s_bdesc.upload(bd.data(), bd.size(), s));
    HIPCHK(M->d_part.alloc(std::max<int64_t>(solve_part_size(SP), 1)));
    HIPCHK(hipStreamSynchronize(s));  // fd and bd are host temporaries
    DevSolve &DS = M->dsolve;
    DS.fwd_tiles = M->s_ftiles.p;
    DS.bwd_tiles = M->s_btiles.p;
    DS.fwd_desc = M->s_fdesc.p;
    DS.bwd_desc = M->s_bdesc.p;
    DS.fwd_tbase_rel = M->s_ftbase.p;
    DS.bwd_tbase_rel = M->s_btbase.p;
    DS.fwd_nks = M->s_fnks.p;
    DS.bwd_nks = M->s_bnks.p;
    DS.fwd_kw = M->s_fkw.p;
    DS.bwd_kw = M->s_bkw.p;
    return SOL105_OK;
}

int upload_symbolic(sol105_ctx *ctx, sol105_model *M) {
    hipStream_t s = ctx->stream;
    Symbolic &S = M->sym;
    int nf = (int)S.fronts.size();
    std::vector<int32_t> ns(nf), nu(nf), m(nf), piv0(nf), cptr(nf + 1, 0), cidx, flist;
    std::vector<int64_t> off(nf), rowoff(nf), vecoff(nf), reloff(nf);
    for (int f = 0; f < nf; f++) {
        const FrontInfo &F = S.fronts[f];
        ns[f] = F.ns; nu[f] = F.nu; m[f] = F.m; piv0[f] = F.piv0;
        off[f] = F.off; rowoff[f] = F.row_off; vecoff[f] = F.vec_off; reloff[f] = F.rel_off;
        cptr[f + 1] = cptr[f] + (int32_t)S.children[f].size();
        for (int c : S.children[f]) cidx.push_back(c);
    }
    FactorPlan &P = M->plan;
    build_plans(S, P, M->splan, flist);
    if (!M->prod_cfg.empty()) solve_plan_build(P, ns, m, flist, M->splan, M->prod_cfg);
    M->flist_host = flist;
    HIPCHK(M->f_ns.upload(ns.data(), nf, s));
    HIPCHK(M->f_nu.upload(nu.data(), nf, s));
    HIPCHK(M->f_m.upload(m.data(), nf, s));
    HIPCHK(M->f_piv0.upload(piv0.data(), nf, s));
    HIPCHK(M->f_off.upload(off.data(), nf, s));
    HIPCHK(M->f_rowoff.upload(rowoff.data(), nf, s));
    HIPCHK(M->f_vecoff.upload(vecoff.data(), nf, s));
    HIPCHK(M->f_reloff.upload(reloff.data(), nf, s));
    HIPCHK(M->f_rows.upload(S.rows.data(), S.rows.size(), s));
    HIPCHK(M->f_rel.upload(S.rel.data(), S.rel.size(), s));
    HIPCHK(M->f_cptr.upload(cptr.data(), cptr.size(), s));
    HIPCHK(M->f_cidx.upload(cidx.data(), cidx.size(), s));
    HIPCHK(M->f_flist.upload(flist.data(), flist.size(), s));
    HIPCHK(M->f_ksrc.upload(S.k_src.data(), S.k_src.size(), s));
    HIPCHK(M->f_kdst.upload(S.k_dst.data(), S.k_dst.size(), s));
    HIPCHK(M->d_elim_of_dof.upload(S.elim_of_dof.data(), S.elim_of_dof.size(), s));
    HIPCHK(M->d_dof_of_elim.upload(S.dof_of_elim.data(), S.dof_of_elim.size(), s));
    HIPCHK(M->d_free.upload(M->free_mask.data(), M->free_mask.size(), s));
    HIPCHK(M->d_front.alloc(std::max<int64_t>(S.front_storage, 1)));
    HIPCHK(M->d_piv.alloc(std::max(S.n_free, 1)));
    HIPCHK(M->d_kdiag_elim.alloc(std::max(S.n_free, 1)));
    HIPCHK(M->d_flag.alloc(std::max(S.n_free, 1)));
    HIPCHK(M->d_W.alloc(solve_w_size(S.vec_storage)));
    HIPCHK(M->d_err.alloc(4));
    DevFronts &D = M->dev;
    D.ns = M->f_ns.p; D.nu = M->f_nu.p; D.m = M->f_m.p; D.piv0 = M->f_piv0.p;
    D.off = M->f_off.p; D.row_off = M->f_rowoff.p; D.vec_off = M->f_vecoff.p; D.rel_off = M->f_reloff.p;
    D.rows = M->f_rows.p; D.rel = M->f_rel.p; D.child_ptr = M->f_cptr.p; D.child_idx = M->f_cidx.p;
    D.flist = M->f_flist.p; D.k_src = M->f_ksrc.p; D.k_dst = M->f_kdst.p;
    // GEMM-form solve plan
    {
        SolvePlan &SP = M->splan;
        HIPCHK(M->s_poff.upload(SP.poff.data(), SP.poff.size(), s));
        HIPCHK(M->s_pullptr.upload(SP.pull_ptr.data(), SP.pull_ptr.size(), s));
        {
            // the device pull map holds absolute vector slots (child's vec_off
            // + row): one dependent load less per contribution
            std::vector<int32_t> slot(SP.pull_src.size() / 2);
            for (size_t q = 0; q < slot.size(); q++)
                slot[q] = (int32_t)(S.fronts[SP.pull_src[2 * q]].vec_off + SP.pull_src[2 * q + 1]);
            HIPCHK(M->s_pullsrc.upload(slot.data(), slot.size(), s));
            // the same map as (first, second) child slot per row, -1 for none,
            // when no row has more than two entries (a binary tree): one
            // dependent load per row instead of two
            std::vector<int32_t> pair(2 * (size_t)S.vec_storage, -1);
            bool pairs_ok = true;
            for (int64_t v = 0; v < S.vec_storage && pairs_ok; v++) {
                const int q0 = SP.pull_ptr[v], cnt = SP.pull_ptr[v + 1] - q0;
                if (cnt > 2) pairs_ok = false;
                for (int j = 0; j < cnt && j < 2; j++) pair[2 * v + j] = slot[q0 + j];
            }
            if (pairs_ok) HIPCHK(M->s_pull2.upload(pair.data(), pair.size(), s));
            M->pull2_ok = pairs_ok;
            HIPCHK(hipStreamSynchronize(s));  // slot and pair are host temporaries
        }
        HIPCHK(M->s_lblk.upload(SP.lblk.data(), SP.lblk.size(), s));
        HIPCHK(M->d_P.alloc(SP.p_storage + SP.p_pad));
        HIPCHK(M->d_PT.alloc(SP.p_storage + SP.p_pad));
        HIPCHK(hipMemsetAsync(M->d_P.p + SP.p_storage, 0, sizeof(double) * SP.p_pad, s));
        HIPCHK(hipMemsetAsync(M->d_PT.p + SP.p_storage, 0, sizeof(double) * SP.p_pad, s));
        DevSolve &DS = M->dsolve;
        DS.poff = M->s_poff.p;
        int rc = upload_solve_tiles(ctx, M);
        if (rc) return rc;
        DS.pull_ptr = M->s_pullptr.p;
        DS.pull_src = M->s_pullsrc.p;
        DS.lblk = M->s_lblk.p;
        DS.n_lblk = (int)(SP.lblk.size() / 2);
        DS.pull2 = M->pull2_ok ? M->s_pull2.p : nullptr;
    }
    M->have_sym = true;
    return SOL105_OK;
}

// x <- K_ff^-1 x (elimination order, nrhs <= MAXR columns)
hipError_t tri_solve(sol105_model *M, double *X, int nrhs, hipStream_t s) {
    hipError_t e = M->d_tsol.alloc((size_t)std::max(M->sym.n_free, 1) * nrhs);
    if (e != hipSuccess) return e;
    e = solve_fwd(M->plan, M->dev, M->dsolve, M->splan, M->d_P.p, M->d_W.p, M->d_part.p, X, M->d_tsol.p, nrhs, s);
    if (e != hipSuccess) return e;
    return solve_bwd(M->plan, M->dev, M->dsolve, M->splan, M->d_PT.p, M->d_W.p, M->d_part.p, M->d_tsol.p, X, nrhs, s);
}

double ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

int model_init(sol105_ctx *ctx, const sol105_model_desc *d, sol105_model *M) {
    if (d->n_nodes <= 0 || !d->xyz) return set_err(ctx, SOL105_E_INVALID, "model has no nodes");
    if (d->n_q4 < 0 || d->n_t3 < 0 || d->n_bar < 0) return set_err(ctx, SOL105_E_INVALID, "negative element count");
    if ((d->n_q4 > 0 && (!d->q4_nodes || !d->q4_prop)) || (d->n_t3 > 0 && (!d->t3_nodes || !d->t3_prop)) ||
        (d->n_bar > 0 && (!d->bar_nodes || !d->bar_prop || !d->bar_vref)))
        return set_err(ctx, SOL105_E_INVALID, "missing element arrays");
    if (!d->spc_mask || !d->force || !d->node_has_shell || !d->node_has_bar)
        return set_err(ctx, SOL105_E_INVALID, "missing SPC / load / topology arrays");
    if (d->nd <= 0) return set_err(ctx, SOL105_E_INVALID, "EIGRL ND must be positive");
    if (!std::isfinite(d->eigrl_v1) || !std::isfinite(d->eigrl_v2) || d->n_shell_cards < 0)
        return set_err(ctx, SOL105_E_INVALID, "EIGRL V1/V2 must be finite and the shell card count non-negative");
    (void)hipSetDevice(ctx->device);
    int n = d->n_nodes;
    M->n_nodes = n;
    M->n_q4 = d->n_q4;
    M->n_t3 = d->n_t3;
    M->n_bar = d->n_bar;
    M->xyz.assign(d->xyz, d->xyz + 3 * (size_t)n);
    M->q4.assign(d->q4_nodes, d->q4_nodes + 4 * (size_t)d->n_q4);
    M->q4prop.assign(d->q4_prop, d->q4_prop + 6 * (size_t)d->n_q4);
    M->t3.assign(d->t3_nodes, d->t3_nodes + 3 * (size_t)d->n_t3);
    M->t3prop.assign(d->t3_prop, d->t3_prop + 6 * (size_t)d->n_t3);
    M->bar.assign(d->bar_nodes, d->bar_nodes + 2 * (size_t)d->n_bar);
    M->bar_vref.assign(d->bar_vref, d->bar_vref + 3 * (size_t)d->n_bar);
    M->barprop.assign(d->bar_prop, d->bar_prop + 8 * (size_t)d->n_bar);
    M->spc_mask.assign(d->spc_mask, d->spc_mask + 6 * (size_t)n);
    M->force.assign(d->force, d->force + 6 * (size_t)n);
    M->has_shell.assign(d->node_has_shell, d->node_has_shell + n);
    M->has_bar.assign(d->node_has_bar, d->node_has_bar + n);
    M->k6rot = d->k6rot;
    M->autospc = d->autospc;
    M->nd = d->nd;
    M->v1 = d->eigrl_v1;
    M->v2 = d->eigrl_v2;
    M->n_shell_cards = d->n_shell_cards > 0 ? d->n_shell_cards : d->n_q4 + d->n_t3;
    for (int32_t v : M->q4)
        if (v < 0 || v >= n) return set_err(ctx, SOL105_E_INVALID, "CQUAD4 node rank %d out of range", v);
    for (int32_t v : M->t3)
        if (v < 0 || v >= n) return set_err(ctx, SOL105_E_INVALID, "CTRIA3 node rank %d out of range", v);
    for (int32_t v : M->bar)
        if (v < 0 || v >= n) return set_err(ctx, SOL105_E_INVALID, "CBAR node rank %d out of range", v);
    for (int e = 0; e < M->n_q4; e++) {
        const double *p = &M->q4prop[6 * (size_t)e];
        if (!(p[2] > 0.0) || !(p[0] > 0.0)) return set_err(ctx, SOL105_E_INVALID, "CQUAD4 %d: T and E must be positive", e);
    }
    for (int e = 0; e < M->n_t3; e++) {
        const double *p = &M->t3prop[6 * (size_t)e];
        if (!(p[2] > 0.0) || !(p[0] > 0.0)) return set_err(ctx, SOL105_E_INVALID, "CTRIA3 %d: T and E must be positive", e);
    }
    {
        std::string cerr;
        int crc = build_constraint_map(d, M->mpc, cerr);
        if (crc) return set_err(ctx, crc, "%s", cerr.c_str());
    }
    ConstraintPlan cplan;
    if (M->has_mpc()) {
        // the element pattern, then the pattern with the constraint couplings
        BlockPattern elem;
        build_block_pattern(n, M->n_q4, M->q4.data(), M->n_t3, M->t3.data(), M->n_bar, M->bar.data(), elem);
        std::vector<std::pair<int32_t, int32_t>> extra;
        constraint_pattern_pairs(M->mpc, elem, n, extra);
        build_block_pattern(n, M->n_q4, M->q4.data(), M->n_t3, M->t3.data(), M->n_bar, M->bar.data(), M->pat,
                            &extra);
        build_constraint_plan(M->mpc, elem, M->pat, n, cplan);
        // dependent DOFs are fixed; their loads move to the independent DOFs
        for (int32_t dd : M->mpc.dep) M->spc_mask[dd] = 1;
        redistribute_force(M->mpc, M->force);
    } else {
        build_block_pattern(n, M->n_q4, M->q4.data(), M->n_t3, M->t3.data(), M->n_bar, M->bar.data(), M->pat);
    }
    M->diag_blk.resize(n);
    for (int v = 0; v < n; v++) M->diag_blk[v] = (int32_t)M->pat.find(v, v);
    hipStream_t s = ctx->stream;
    if (M->has_mpc()) {
        const ConstraintMap &C = M->mpc;
        HIPCHK(M->d_mpc_dep.upload(C.dep.data(), C.dep.size(), s));
        HIPCHK(M->d_mpc_ptr.upload(C.ptr.data(), C.ptr.size(), s));
        HIPCHK(M->d_mpc_ind.upload(C.ind.data(), std::max<size_t>(C.ind.size(), 1), s));
        HIPCHK(M->d_mpc_coef.upload(C.coef.data(), std::max<size_t>(C.coef.size(), 1), s));
        M->mpc_nt = (int64_t)cplan.tgt.size();
        M->mpc_nz = (int64_t)cplan.zero.size();
        HIPCHK(M->d_mpc_tgt.upload(cplan.tgt.data(), cplan.tgt.size(), s));
        HIPCHK(M->d_mpc_tptr.upload(cplan.tptr.data(), cplan.tptr.size(), s));
        HIPCHK(M->d_mpc_src.upload(cplan.src.data(), cplan.src.size(), s));
        HIPCHK(M->d_mpc_tcoef.upload(cplan.coef.data(), cplan.coef.size(), s));
        HIPCHK(M->d_mpc_zero.upload(cplan.zero.data(), cplan.zero.size(), s));
        ConstraintFold fold;
        build_constraint_fold(C, fold);
        M->mpc_nrecv = (int32_t)fold.recv.size();
        HIPCHK(M->d_mpc_frecv.upload(fold.recv.data(), fold.recv.size(), s));
        HIPCHK(M->d_mpc_fptr.upload(fold.ptr.data(), fold.ptr.size(), s));
        HIPCHK(M->d_mpc_fdep.upload(fold.dep.data(), fold.dep.size(), s));
        HIPCHK(M->d_mpc_fcoef.upload(fold.coef.data(), fold.coef.size(), s));
        HIPCHK(hipStreamSynchronize(s));  // cplan and fold are host temporaries
    }
    HIPCHK(M->d_xyz.upload(M->xyz.data(), M->xyz.size(), s));
    HIPCHK(M->d_q4.upload(M->q4.data(), M->q4.size(), s));
    HIPCHK(M->d_q4prop.upload(M->q4prop.data(), M->q4prop.size(), s));
    HIPCHK(M->d_t3.upload(M->t3.data(), M->t3.size(), s));
    HIPCHK(M->d_t3prop.upload(M->t3prop.data(), M->t3prop.size(), s));
    HIPCHK(M->d_bar.upload(M->bar.data(), M->bar.size(), s));
    HIPCHK(M->d_bar_vref.upload(M->bar_vref.data(), M->bar_vref.size(), s));
    HIPCHK(M->d_barprop.upload(M->barprop.data(), M->barprop.size(), s));
    HIPCHK(M->d_force.upload(M->force.data(), M->force.size(), s));
    HIPCHK(M->d_rp.upload(M->pat.row_ptr.data(), M->pat.row_ptr.size(), s));
    HIPCHK(M->d_col.upload(M->pat.col.data(), M->pat.col.size(), s));
    HIPCHK(M->d_cptr.upload(M->pat.cptr.data(), M->pat.cptr.size(), s));
    HIPCHK(M->d_cidx.upload(M->pat.cidx.data(), M->pat.cidx.size(), s));
    HIPCHK(M->d_diag.upload(M->diag_blk.data(), M->diag_blk.size(), s));
    size_t neb = ((size_t)M->n_q4 * 16 + (size_t)M->n_t3 * 9 + (size_t)M->n_bar * 4) * 36;
    HIPCHK(M->d_eb.alloc(neb));
    HIPCHK(M->d_K.alloc(M->pat.n_blocks() * 36));
    HIPCHK(M->d_KG.alloc(M->pat.n_blocks() * 36));
    HIPCHK(M->d_u.alloc(6 * (size_t)n));
    HIPCHK(M->d_kdiag.alloc(6 * (size_t)n));
    HIPCHK(hipStreamSynchronize(s));
    return SOL105_OK;
}

}  // namespace s105

extern "C" {

int sol105_ctx_create(int device, sol105_ctx **out) {
    if (!out) return SOL105_E_INVALID;
    *out = nullptr;
    auto ctx = std::make_unique<sol105_ctx>();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SOL105_E_DEVICE;
    if (device < 0 || device >= ndev) return SOL105_E_INVALID;
    ctx->device = device;
    // SOL105_STEP_GRAPHS=0 in the environment: new contexts launch the
    // Lanczos steps directly (for kernel-trace profilers that do not follow
    // stream capture)
    if (const char *g = std::getenv("SOL105_STEP_GRAPHS")) ctx->graphs = std::strcmp(g, "0") != 0;
    if (hipSetDevice(device) != hipSuccess) return SOL105_E_DEVICE;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return SOL105_E_DEVICE;
    // device address of the no-allocation scratch array, once per device; a
    // failure is kept and reported by sol105_assemble_K_noalloc
    (void)noalloc_setup(device);
    for (auto &e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) return SOL105_E_DEVICE;
    *out = ctx.release();
    return SOL105_OK;
}

int sol105_ctx_destroy(sol105_ctx *ctx) {
    if (!ctx) return SOL105_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SOL105_OK;
}

const char *sol105_last_error(const sol105_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sol105_ctx_set_option(sol105_ctx *ctx, int32_t option, double value) {
    if (!ctx) return SOL105_E_INVALID;
    if (!(value == std::floor(value)))
        return set_err(ctx, SOL105_E_INVALID, "option %d: value %g is not an integer", option, value);
    switch (option) {
        case SOL105_OPT_LANCZOS_MAX_VECTORS:
            if (value < 32 || value > 1024)
                return set_err(ctx, SOL105_E_INVALID, "Lanczos basis size %g out of range 32..1024", value);
            ctx->lanczos_max_vectors = (int)value;
            return SOL105_OK;
        case SOL105_OPT_WARM_START:
            if (value != 0 && value != 1) return set_err(ctx, SOL105_E_INVALID, "warm start is 0 or 1");
            ctx->warm_start = value != 0;
            return SOL105_OK;
        case SOL105_OPT_STEP_GRAPHS:
            if (value != 0 && value != 1) return set_err(ctx, SOL105_E_INVALID, "step graphs is 0 or 1");
            ctx->graphs = value != 0;
            return SOL105_OK;
        case SOL105_OPT_SENS_MODE:
            if (value != 0 && value != 1) return set_err(ctx, SOL105_E_INVALID, "sensitivity mode is 0 or 1");
            ctx->sens_mode = (int)value;
            return SOL105_OK;
        case SOL105_OPT_LANCZOS_MAX_RESTARTS:
            if (value < 0 || value > 1000)
                return set_err(ctx, SOL105_E_INVALID, "Lanczos restarts %g out of range 0..1000", value);
            ctx->lanczos_max_restarts = (int)value;
            return SOL105_OK;
        default:
            return set_err(ctx, SOL105_E_INVALID, "unknown option %d", option);
    }
}

int sol105_model_upload(sol105_ctx *ctx, const sol105_model_desc *d, sol105_model **out) {
    if (!ctx || !d || !out) return set_err(ctx, SOL105_E_INVALID, "null argument");
    *out = nullptr;
    auto M = std::make_unique<sol105_model>();
    int rc = model_init(ctx, d, M.get());
    if (rc) return rc;
    *out = M.release();
    return SOL105_OK;
}

int sol105_model_free(sol105_ctx *ctx, sol105_model *model) {
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    delete model;
    return SOL105_OK;
}

int sol105_model_set_q4_thickness(sol105_ctx *ctx, sol105_model *M, const double *t) {
    if (!ctx || !M || (!t && M->n_q4 > 0)) return set_err(ctx, SOL105_E_INVALID, "null argument");
    for (int e = 0; e < M->n_q4; e++) {
        if (!(t[e] > 0.0)) return set_err(ctx, SOL105_E_INVALID, "thickness of CQUAD4 %d must be positive", e);
        M->q4prop[6 * (size_t)e + 2] = t[e];
    }
    (void)hipSetDevice(ctx->device);
    HIPCHK(M->d_q4prop.upload(M->q4prop.data(), M->q4prop.size(), ctx->stream));
    M->K_valid = M->factor_valid = M->u_valid = M->KG_valid = false;
    M->n_modes = 0;
    return SOL105_OK;
}

int sol105_model_set_shell_thickness(sol105_ctx *ctx, sol105_model *M, const double *t) {
    if (!ctx || !M || (!t && M->n_q4 + M->n_t3 > 0)) return set_err(ctx, SOL105_E_INVALID, "null argument");
    for (int e = 0; e < M->n_q4 + M->n_t3; e++)
        if (!(t[e] > 0.0)) return set_err(ctx, SOL105_E_INVALID, "thickness of shell %d must be positive", e);
    for (int e = 0; e < M->n_q4; e++) M->q4prop[6 * (size_t)e + 2] = t[e];
    for (int e = 0; e < M->n_t3; e++) M->t3prop[6 * (size_t)e + 2] = t[M->n_q4 + e];
    (void)hipSetDevice(ctx->device);
    HIPCHK(M->d_q4prop.upload(M->q4prop.data(), M->q4prop.size(), ctx->stream));
    HIPCHK(M->d_t3prop.upload(M->t3prop.data(), M->t3prop.size(), ctx->stream));
    M->K_valid = M->factor_valid = M->u_valid = M->KG_valid = false;
    M->n_modes = 0;
    return SOL105_OK;
}

int sol105_pattern(sol105_ctx *ctx, const sol105_model *M, int64_t *n_blocks, int32_t *rp, int32_t *col) {
    if (!M) return set_err(ctx, SOL105_E_INVALID, "null model");
    if (n_blocks) *n_blocks = M->pat.n_blocks();
    if (rp) std::memcpy(rp, M->pat.row_ptr.data(), M->pat.row_ptr.size() * sizeof(int32_t));
    if (col) std::memcpy(col, M->pat.col.data(), M->pat.col.size() * sizeof(int32_t));
    return SOL105_OK;
}

int sol105_assemble_K(sol105_ctx *ctx, sol105_model *M, double *vals) {
    if (!ctx || !M) return set_err(ctx, SOL105_E_INVALID, "null argument");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    ElemArgs a = elem_args(M);
    launch_q4_ke(a, M->d_eb.p, s);
    launch_t3_ke(a, M->d_eb.p, s);
    launch_bar_ke(a, M->d_eb.p, s);
    launch_block_gather(M->pat.n_blocks(), M->d_cptr.p, M->d_cidx.p, M->d_eb.p, M->d_K.p, s);
    if (M->has_mpc())
        launch_mpc_apply(M->mpc_nt, M->d_mpc_tgt.p, M->d_mpc_tptr.p, M->d_mpc_src.p, M->d_mpc_tcoef.p, M->mpc_nz,
                         M->d_mpc_zero.p, M->d_K.p, s);
    HIPCHK(hipGetLastError());
    M->K_valid = true;
    M->factor_valid = M->u_valid = M->KG_valid = false;
    if (vals) {
        HIPCHK(hipMemcpyAsync(vals, M->d_K.p, M->pat.n_blocks() * 36 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return SOL105_OK;
}

}  // extern "C"

namespace s105 {

namespace {

// One numeric factorisation of K_ff + shift I on the free set fm (a new
// symbolic analysis when fm changed).  herr[0] != 0: a pivot was not
// positive (the reference's cholesky throws); with shift == 0, herr[1]
// counts the pivots with d_i / K_ii < 1e-7 (flags per elimination index) and
// the explicit-inverse solve operators are formed.
int factor_free_set(sol105_ctx *ctx, sol105_model *M, const std::vector<uint8_t> &fm, double shift, int herr[4]) {
    hipStream_t s = ctx->stream;
    if (!M->have_sym || fm != M->free_mask) {
        M->free_mask = fm;
        M->factor_valid = false;
        try {
            symbolic_analyze(M->pat, M->xyz.data(), M->free_mask.data(), 32, M->sym);
        } catch (const std::exception &ex) {
            return set_err(ctx, SOL105_E_INVALID, "symbolic analysis failed: %s", ex.what());
        }
        if (M->sym.n_free == 0) return set_err(ctx, SOL105_E_SINGULAR, "no free DOFs");
        M->sym_gen++;
        int rc = upload_symbolic(ctx, M);
        if (rc) return rc;
    }
    HIPCHK(hipMemsetAsync(M->d_err.p, 0, sizeof(int) * 4, s));
    const double *kv = M->d_K.p;
    if (shift != 0.0) {
        const size_t nv = M->pat.n_blocks() * 36;
        HIPCHK(M->d_Kshift.alloc(nv));
        HIPCHK(hipMemcpyAsync(M->d_Kshift.p, M->d_K.p, nv * sizeof(double), hipMemcpyDeviceToDevice, s));
        launch_add_diag(M->n_nodes, M->d_diag.p, shift, M->d_Kshift.p, s);
        kv = M->d_Kshift.p;
    }
    factor_numeric(M->plan, M->dev, kv, M->d_front.p, M->d_piv.p, M->d_err.p, s);
    if (shift == 0.0) {
        // post-factorisation singular-pivot check: d_i / K_ii < 1e-7
        launch_kdiag_elim(M->sym.n_free, M->d_dof_of_elim.p, M->d_diag.p, M->d_K.p, M->d_kdiag_elim.p, s);
        launch_pivot_ratio(M->sym.n_free, M->d_piv.p, M->d_kdiag_elim.p, M->d_err.p + 1, M->d_flag.p, s);
        solve_prepare(M->plan, M->dev, M->dsolve, M->splan, M->d_front.p, M->d_P.p, M->d_PT.p, s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(herr, M->d_err.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SOL105_OK;
}

// Largest |entry| of the element stiffness matrices before assembly
// (max_elem_stiff, JFEM/src/solver/assembly.jl:4352-4356): the element
// kernels run again into the model's staging buffer.
int max_elem_stiff(sol105_ctx *ctx, sol105_model *M, double &out) {
    hipStream_t s = ctx->stream;
    ElemArgs a = elem_args(M);
    launch_q4_ke(a, M->d_eb.p, s);
    launch_t3_ke(a, M->d_eb.p, s);
    launch_bar_ke(a, M->d_eb.p, s);
    const int64_t neb = ((int64_t)M->n_q4 * 16 + (int64_t)M->n_t3 * 9 + (int64_t)M->n_bar * 4) * 36;
    HIPCHK(M->d_Kshift.alloc(std::max<size_t>(M->pat.n_blocks() * 36, 1)));
    launch_maxabs(neb, M->d_eb.p, M->d_Kshift.p, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&out, M->d_Kshift.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SOL105_OK;
}

// DOFs (by DOF index) of the last unshifted factorisation whose pivot is
// zero to working precision, |d_i / K_ii| < 1e-7 (non-positive pivots of a
// failed factorisation included only when they are that small).
int near_zero_pivots(sol105_ctx *ctx, sol105_model *M, std::vector<int32_t> &dofs) {
    const int nf = M->sym.n_free;
    std::vector<double> piv(nf), kd(nf);
    HIPCHK(hipMemcpy(piv.data(), M->d_piv.p, nf * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(kd.data(), M->d_kdiag_elim.p, nf * sizeof(double), hipMemcpyDeviceToHost));
    dofs.clear();
    for (int i = 0; i < nf; i++) {
        const double k = std::fabs(kd[i]);
        if (k > 1e-30 && std::fabs(piv[i]) < 1e-7 * k) dofs.push_back(M->sym.dof_of_elim[i]);
    }
    return SOL105_OK;
}

// Factorisation AUTOSPC after a failed Cholesky (boundary_conditions.jl:
// 787-881): probe K_ff + shift I with shift = max(max_elem_stiff, 1) 10^e,
// e = -12, -10, -8, -6, until a probe factors; fix the DOFs whose pivot
// L_ii is below min(3 sqrt(shift), 1e-4 median(L_ii)) unless they are more
// than half of the free set; then factor K_ff once more.  Where that still
// fails the reference turns to an LU factorisation of K_ff (and, for the
// buckling solve, to a dense generalized eigensolver); here the DOFs whose
// pivots came out zero to working precision in the failed factorisation
// (`zero_piv`, the refix rule on |d_i / K_ii|) are fixed instead and K_ff is
// factored once more.  On exact mechanisms (zero-energy patterns with no
// load and no geometric stiffness) that gives the reference's translations
// and buckling loads; an indefinite K_ff (large negative pivots) stays
// SOL105_E_SINGULAR.
int factor_autospc(sol105_ctx *ctx, sol105_model *M, std::vector<uint8_t> &fm,
                   const std::vector<int32_t> &zero_piv) {
    double mes = 0.0;
    int rc = max_elem_stiff(ctx, M, mes);
    if (rc) return rc;
    bool found = false;
    int herr[4];
    for (int e : {-12, -10, -8, -6}) {
        const double shift = std::max(mes, 1.0) * std::pow(10.0, (double)e);
        rc = factor_free_set(ctx, M, fm, shift, herr);
        if (rc) return rc;
        if (herr[0]) continue;
        const int nf = M->sym.n_free;
        std::vector<double> piv(nf), L(nf);
        HIPCHK(hipMemcpy(piv.data(), M->d_piv.p, nf * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < nf; i++) L[i] = std::sqrt(std::fabs(piv[i]));
        std::vector<double> srt(L);
        std::sort(srt.begin(), srt.end());
        const double med = (nf % 2) ? srt[nf / 2] : 0.5 * (srt[nf / 2 - 1] + srt[nf / 2]);
        const double thr = std::min(3.0 * std::sqrt(shift), 1e-4 * med);
        int n_mech = 0;
        for (int i = 0; i < nf; i++) n_mech += L[i] < thr;
        found = true;
        if (n_mech > 0.5 * nf) break;  // threshold too aggressive for this shift: nothing fixed
        for (int i = 0; i < nf; i++)
            if (L[i] < thr) fm[M->sym.dof_of_elim[i]] = 0;
        break;
    }
    if (found) {
        rc = factor_free_set(ctx, M, fm, 0.0, herr);
        if (rc) return rc;
        if (herr[0] == 0) {
            M->factor_valid = true;
            return SOL105_OK;
        }
    }
    int n_fix = 0, n_free = 0;
    for (int32_t d : zero_piv)
        if (fm[d]) { fm[d] = 0; n_fix++; }
    for (uint8_t f : fm) n_free += f;
    if (n_fix > 0 && n_fix <= n_free) {
        rc = factor_free_set(ctx, M, fm, 0.0, herr);
        if (rc) return rc;
        if (herr[0] == 0) {
            M->factor_valid = true;
            return SOL105_OK;
        }
    }
    return set_err(ctx, SOL105_E_SINGULAR,
                   "stiffness matrix is not positive definite on the free set after the factorisation AUTOSPC "
                   "(no zero pivots to fix: K_ff is indefinite)");
}

}  // namespace

int factorize(sol105_ctx *ctx, sol105_model *M, bool shifted_fallback) {
    hipStream_t s = ctx->stream;
    int ndof = M->ndof();
    // AUTOSPC from the assembled diagonal
    extract_kdiag(M->n_nodes, M->d_diag.p, M->d_K.p, M->d_kdiag.p, s);
    std::vector<double> kd(ndof);
    HIPCHK(hipMemcpyAsync(kd.data(), M->d_kdiag.p, ndof * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<uint8_t> fm;
    autospc_mask(M, kd, fm);
    int herr[4];
    int rc = factor_free_set(ctx, M, fm, 0.0, herr);
    if (rc) return rc;
    if (herr[0] == 0) {
        if (herr[1] == 0) {
            M->factor_valid = true;
            return SOL105_OK;
        }
        // singular-pivot refix (boundary_conditions.jl:742-782): fix the
        // flagged DOFs and factor once more; the ratios are not checked again
        std::vector<uint8_t> flag(M->sym.n_free);
        HIPCHK(hipMemcpy(flag.data(), M->d_flag.p, flag.size(), hipMemcpyDeviceToHost));
        for (int e = 0; e < M->sym.n_free; e++)
            if (flag[e]) fm[M->sym.dof_of_elim[e]] = 0;
        rc = factor_free_set(ctx, M, fm, 0.0, herr);
        if (rc) return rc;
        if (herr[0] == 0) {
            M->factor_valid = true;
            return SOL105_OK;
        }
    }
    if (!shifted_fallback)
        return set_err(ctx, SOL105_E_SINGULAR, "stiffness matrix is not positive definite on the free set");
    std::vector<int32_t> zero_piv;
    rc = near_zero_pivots(ctx, M, zero_piv);
    if (rc) return rc;
    return factor_autospc(ctx, M, fm, zero_piv);
}

}  // namespace s105

extern "C" {

int sol105_static_solve(sol105_ctx *ctx, sol105_model *M, double *u, uint8_t *free_mask) {
    if (!ctx || !M) return set_err(ctx, SOL105_E_INVALID, "null argument");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    if (!M->K_valid) {
        int rc = sol105_assemble_K(ctx, M, nullptr);
        if (rc) return rc;
    }
    if (!M->factor_valid) {
        int rc = factorize(ctx, M, true);
        if (rc) return rc;
    }
    int nfree = M->sym.n_free, ndof = M->ndof();
    double *X = nullptr;
    HIPCHK(hipMallocAsync((void **)&X, sizeof(double) * nfree, s));
    launch_gather_elim(nfree, ndof, 1, M->d_dof_of_elim.p, M->d_force.p, X, s);
    const hipError_t te = tri_solve(M, X, 1, s);
    if (te == hipSuccess) launch_scatter_elim(nfree, ndof, 1, M->d_elim_of_dof.p, X, M->d_u.p, s);
    if (te == hipSuccess && M->has_mpc())
        launch_mpc_recover((int)M->mpc.dep.size(), M->d_mpc_dep.p, M->d_mpc_ptr.p, M->d_mpc_ind.p, M->d_mpc_coef.p,
                           M->d_u.p, ndof, 1, s);
    HIPCHK(hipFreeAsync(X, s));
    HIPCHK(te);
    HIPCHK(hipGetLastError());
    M->u_valid = true;
    M->KG_valid = false;
    if (u) HIPCHK(hipMemcpyAsync(u, M->d_u.p, ndof * sizeof(double), hipMemcpyDeviceToHost, s));
    if (free_mask) std::memcpy(free_mask, M->free_mask.data(), ndof);
    HIPCHK(hipStreamSynchronize(s));
    return SOL105_OK;
}

int sol105_assemble_KG(sol105_ctx *ctx, sol105_model *M, const double *u, double *vals) {
    if (!ctx || !M) return set_err(ctx, SOL105_E_INVALID, "null argument");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    if (u) {
        HIPCHK(hipMemcpyAsync(M->d_u.p, u, M->ndof() * sizeof(double), hipMemcpyHostToDevice, s));
        M->u_valid = true;
    }
    if (!M->u_valid) return set_err(ctx, SOL105_E_STATE, "no static displacement: run sol105_static_solve first");
    ElemArgs a = elem_args(M);
    launch_q4_kg(a, M->d_u.p, M->d_eb.p, s);
    launch_t3_kg(a, M->d_u.p, M->d_eb.p, s);
    launch_bar_kg(a, M->d_u.p, M->d_eb.p, s);
    launch_block_gather(M->pat.n_blocks(), M->d_cptr.p, M->d_cidx.p, M->d_eb.p, M->d_KG.p, s);
    // KG takes the same constraint redistribution as K (assembly.jl:6617-6620)
    if (M->has_mpc())
        launch_mpc_apply(M->mpc_nt, M->d_mpc_tgt.p, M->d_mpc_tptr.p, M->d_mpc_src.p, M->d_mpc_tcoef.p, M->mpc_nz,
                         M->d_mpc_zero.p, M->d_KG.p, s);
    HIPCHK(hipGetLastError());
    M->KG_valid = true;
    if (vals) {
        HIPCHK(hipMemcpyAsync(vals, M->d_KG.p, M->pat.n_blocks() * 36 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return SOL105_OK;
}

namespace {

// Block Lanczos on C = L^-1 (-KG) L^-T with full (CGS2) reorthogonalisation.
// Returns the Ritz values / vectors of the `want` largest |theta|.
// Eigen-decomposition of a small symmetric matrix (b <= 16, column-major) on
// the host by cyclic Jacobi; w descending, U column-major with matching
// columns.  Used only on the rare rank-deficient Lanczos step.
void small_sym_eig_host(int b, std::vector<double> A, std::vector<double> &w, std::vector<double> &U) {
    U.assign((size_t)b * b, 0.0);
    for (int i = 0; i < b; i++) U[(size_t)i * b + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, dia = 0.0;
        for (int q = 0; q < b; q++)
            for (int p = 0; p < b; p++) (p == q ? dia : off) += A[(size_t)q * b + p] * A[(size_t)q * b + p];
        if (off <= 1e-32 * dia || off == 0.0) break;
        for (int p = 0; p < b - 1; p++)
            for (int q = p + 1; q < b; q++) {
                double apq = A[(size_t)q * b + p];
                if (apq == 0.0) continue;
                double app = A[(size_t)p * b + p], aqq = A[(size_t)q * b + q];
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
                for (int k = 0; k < b; k++) {  // A <- J^T A J
                    double akp = A[(size_t)p * b + k], akq = A[(size_t)q * b + k];
                    A[(size_t)p * b + k] = c * akp - sn * akq;
                    A[(size_t)q * b + k] = sn * akp + c * akq;
                }
                for (int k = 0; k < b; k++) {
                    double apk = A[(size_t)k * b + p], aqk = A[(size_t)k * b + q];
                    A[(size_t)k * b + p] = c * apk - sn * aqk;
                    A[(size_t)k * b + q] = sn * apk + c * aqk;
                }
                for (int k = 0; k < b; k++) {
                    double ukp = U[(size_t)p * b + k], ukq = U[(size_t)q * b + k];
                    U[(size_t)p * b + k] = c * ukp - sn * ukq;
                    U[(size_t)q * b + k] = sn * ukp + c * ukq;
                }
            }
    }
    std::vector<int> idx(b);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return A[(size_t)x * b + x] > A[(size_t)y * b + y]; });
    std::vector<double> U2((size_t)b * b);
    w.resize(b);
    for (int c = 0; c < b; c++) {
        w[c] = A[(size_t)idx[c] * b + idx[c]];
        for (int k = 0; k < b; k++) U2[(size_t)c * b + k] = U[(size_t)idx[c] * b + k];
    }
    U.swap(U2);
}

struct LanczosOut {
    std::vector<double> theta;  // selected Ritz values (descending |theta|)
    std::vector<double> resid;
    int m = 0, steps = 0, injected = 0, generated = 0, restarts = 0, seeded = 0;
};

// Block Lanczos on C = L^-1 (-KG) L^-T with full CGS2 reorthogonalisation.
// Each new block is orthonormalised by CholQR2 whose first Cholesky also
// checks every pivot against the block's norm after the first Gram-Schmidt
// pass: a direction left below 1e-7 of it is numerically inside the basis
// (the Krylov space is exhausted or a block multiplicity exceeded b).  Such a
// block is detected at the next Rayleigh-Ritz check; the iteration then
// returns to it and rebuilds it with an eigen-decomposition of its Gram
// matrix, keeping the independent directions and replacing the others by
// seeded random vectors orthogonalised against the whole basis.  The
// Lanczos relation C V = V T + Q R E^T stays exact (the new directions enter
// with zero coefficients), so Rayleigh-Ritz and the residual bound hold.
// With algebraic set the wanted pairs are those of largest theta (smallest
// positive lambda) instead of largest |theta|.
int lanczos(sol105_ctx *ctx, sol105_model *M, int want, double *d_Y /* n x want, elimination order */,
            LanczosOut &out, bool algebraic = false) {
    hipStream_t s = ctx->stream;
    int n = M->sym.n_free, ndof = M->ndof();
    const bool debug = std::getenv("SOL105_DEBUG") != nullptr;
    int b = want <= 8 ? 8 : 16;
    // the basis and its residual block stay inside the free space (a basis
    // of more than n - b vectors leaves the residual block rank-deficient)
    int mmax = std::min((ctx->lanczos_max_vectors / b) * b, ((n - b) / b) * b);
    if (mmax < b) return set_err(ctx, SOL105_E_NOCONV, "free set too small for the eigensolver (%d)", n);
    int guard = 2;
    int kneed = std::min(want + guard, mmax);
    const double tol = 1e-10;
    const double dep_tol2 = 1e-14;
    // thick restart at mmax vectors: keep the kk Ritz vectors of largest
    // |theta| and the next block
    const int kk = mmax >= 128 + b ? (SYMEIG_MAX_K / b) * b : std::max(b, (mmax / 2 / b) * b);
    const int max_restarts = ctx->lanczos_max_restarts;
    const int kz = std::max(kneed, kk);
    LanczosWork &W = M->lw;
    DevBuf<double> &V = W.V, &Wb = W.Wb, &xf = W.xf, &yf = W.yf, &T = W.T, &A = W.A, &Zs = W.Zs, &thd = W.thd,
                   &ework = W.ework, &part = W.part, &small = W.small, &H2 = W.H2, &Ytmp = W.Ytmp;
    DevBuf<int> &ierr = W.ierr, &sw = W.sw;
    HIPCHK(V.alloc((size_t)n * (mmax + b)));
    HIPCHK(Wb.alloc((size_t)n * b));
    HIPCHK(xf.alloc((size_t)ndof * b));
    HIPCHK(yf.alloc((size_t)ndof * b));
    HIPCHK(T.alloc((size_t)(mmax + b) * (mmax + b)));
    HIPCHK(A.alloc((size_t)mmax * mmax));
    HIPCHK(part.alloc(std::max<int64_t>(
        std::max(ts_gemm_tn_part_size(n, mmax + b, b), ts_gemm_nn_tn_part_size(n, mmax + b, b)), 1)));
    HIPCHK(small.alloc(6 * 16 * 16));
    HIPCHK(ierr.alloc(4));
    HIPCHK(sw.alloc(1));
    if (kz > SYMEIG_MAX_K || mmax > SYMEIG_MAX_M)
        return set_err(ctx, SOL105_E_UNSUPPORTED, "Rayleigh-Ritz size out of range (k=%d, m=%d)", kz, mmax);
    HIPCHK(Zs.alloc((size_t)mmax * kz));
    HIPCHK(thd.alloc(kz));
    HIPCHK(ework.alloc(symeig_work_size(mmax, kz)));
    HIPCHK(H2.alloc((size_t)(mmax + b) * b));
    int ldT = mmax + b;
    dev_fill(T.p, (int64_t)ldT * ldT, 0.0, s);
    HIPCHK(hipMemsetAsync(ierr.p, 0, sizeof(int) * 4, s));
    double *G1 = small.p, *G2 = small.p + 256, *Rt = small.p + 512, *G0 = small.p + 768, *Hs = small.p + 1024,
           *R1 = small.p + 1280;
    auto blk = [&](int jj) { return V.p + (size_t)jj * b * n; };
    auto Tsub = [&](int jj) { return T.p + (size_t)(jj + 1) * b + (size_t)jj * b * ldT; };  // R of block jj+1
    // CholQR2 of Q (n x b); with g0 the first Cholesky checks for dependent
    // directions.  g1_ready: G1 = Q^T Q was formed by the kernel that wrote Q.
    // Each triangular solve factors its Gram block itself; the first one
    // forms the second pass's Gram matrix and keeps R1, the second one forms
    // R = R2 R1.
    auto cholqr2 = [&](double *Q, double *Rout, int ldr, const double *g0, int bidx, bool g1_ready) {
        if (!g1_ready) ts_gemm_tn(n, b, b, Q, n, Q, n, G1, 16, 1.0, 0.0, part.p, s);
        chol_trsm_right_upper_gram(n, b, Q, n, G1, 16, g0, 16, g0 ? dep_tol2 : 0.0, bidx, ierr.p, R1, part.p, G2, 16,
                                   s);
        chol_trsm_right_upper(n, b, Q, n, G2, 16, ierr.p, Rout ? R1 : nullptr, 16, Rout, ldr, s);
    };
    // dst = C Q_jj, orthogonalised (CGS2) against V[:, 0:(jj+1)b]; coefficients
    // into T's column block jj; G0 = Gram matrix after the first pass.  The
    // backward solve goes from Q_jj into Wb, the KG product back into Wb and
    // the forward solve from there into dst.
    auto expand = [&](int jj, double *dst) {
        HIPCHK(solve_bwd(M->plan, M->dev, M->dsolve, M->splan, M->d_PT.p, M->d_W.p, M->d_part.p, blk(jj), Wb.p, b, s,
                         ctx->timer()));
        launch_scatter_elim_il(n, ndof, b, M->d_elim_of_dof.p, Wb.p, xf.p, s);
        launch_bsr_spmv(M->n_nodes, M->d_rp.p, M->d_col.p, M->d_KG.p, xf.p, yf.p, b, -1.0, s);
        launch_gather_elim_il(n, b, M->d_dof_of_elim.p, yf.p, Wb.p, s);
        HIPCHK(solve_fwd(M->plan, M->dev, M->dsolve, M->splan, M->d_P.p, M->d_W.p, M->d_part.p, Wb.p, dst, b, s,
                         ctx->timer()));
        int kcols = (jj + 1) * b;
        double *Tcol = T.p + (size_t)jj * b * ldT;
        // first pass: coefficients into T, update; G0 = Gram matrix after it
        ts_gemm_tn(n, kcols, b, V.p, n, dst, n, Tcol, ldT, 1.0, 0.0, part.p, s);
        // ... and the second pass's coefficients (also added into T), in
        // the same pass over V as the update
        ts_gemm_nn_gram_tn(n, kcols, b, V.p, n, Tcol, ldT, dst, n, part.p, G0, 16, H2.p, kcols, Tcol, ldT, s);
        // second pass update; G1 = Gram matrix of the result (the first
        // Cholesky-QR pass of the new block)
        ts_gemm_nn_gram(n, kcols, b, V.p, n, H2.p, kcols, dst, n, -1.0, 1.0, part.p, G1, 16, s);
        return SOL105_OK;
    };
    // the new block is built in place: column block jj+1 of V is not part of
    // the basis that the products of step jj read
    auto fast_step = [&](int jj) {
        int rc = expand(jj, blk(jj + 1));
        if (rc) return rc;
        cholqr2(blk(jj + 1), Tsub(jj), ldT, G0, jj + 1, true);
        return SOL105_OK;
    };
    // rank-revealing rebuild of block jj+1 (host-assisted; rare)
    auto robust_step = [&](int jj, int &rank) {
        int rc = expand(jj, blk(jj + 1));
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(Wb.p, blk(jj + 1), sizeof(double) * n * b, hipMemcpyDeviceToDevice, s));
        ts_gemm_tn(n, b, b, Wb.p, n, Wb.p, n, G1, 16, 1.0, 0.0, part.p, s);
        std::vector<double> hG0(256), hG1(256);
        HIPCHK(hipMemcpyAsync(hG0.data(), G0, 256 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(hG1.data(), G1, 256 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<double> g((size_t)b * b), w, U;
        double gmax = 0.0;
        for (int c = 0; c < b; c++) {
            gmax = std::max(gmax, hG0[(size_t)c * 16 + c]);
            for (int r = 0; r < b; r++) g[(size_t)c * b + r] = 0.5 * (hG1[(size_t)c * 16 + r] + hG1[(size_t)r * 16 + c]);
        }
        small_sym_eig_host(b, g, w, U);
        int r = 0;
        while (r < b && w[r] > dep_tol2 * gmax && w[r] > 0.0) r++;
        rank = r;
        std::vector<double> hH(256, 0.0), hR(256, 0.0);
        for (int c = 0; c < r; c++) {
            double sq = std::sqrt(w[c]);
            for (int k = 0; k < b; k++) {
                hH[(size_t)c * 16 + k] = U[(size_t)c * b + k] / sq;   // H = U_r S^-1/2
                hR[(size_t)k * 16 + c] = sq * U[(size_t)c * b + k];   // R rows = S^1/2 U_r^T
            }
        }
        HIPCHK(hipMemcpyAsync(Hs, hH.data(), 256 * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(Rt, hR.data(), 256 * sizeof(double), hipMemcpyHostToDevice, s));
        double *Qn = blk(jj + 1);
        if (r > 0) ts_gemm_nn(n, b, r, Wb.p, n, Hs, 16, Qn, n, 1.0, 0.0, s);
        if (r < b) {
            double *Z = Qn + (size_t)r * n;
            int nz = b - r, kc = (jj + 1) * b + r;
            hash_fill(n, nz, Z, n, 0x5EED00ull + (uint64_t)jj, s);
            for (int pass = 0; pass < 2; pass++) {
                ts_gemm_tn(n, kc, nz, V.p, n, Z, n, H2.p, kc, 1.0, 0.0, part.p, s);
                ts_gemm_nn(n, kc, nz, V.p, n, H2.p, kc, Z, n, -1.0, 1.0, s);
            }
            for (int pass = 0; pass < 2; pass++) {
                ts_gemm_tn(n, nz, nz, Z, n, Z, n, G1, 16, 1.0, 0.0, part.p, s);
                small_chol_upper(nz, G1, 16, ierr.p, s);
                trsm_right_upper(n, nz, Z, n, G1, 16, s);
            }
        }
        // refinement of the whole block: Qn = Qn R2^-1, R = R2 R
        ts_gemm_tn(n, b, b, Qn, n, Qn, n, G2, 16, 1.0, 0.0, part.p, s);
        small_chol_upper(b, G2, 16, ierr.p, s);
        trsm_right_upper(n, b, Qn, n, G2, 16, s);
        upper_mul(b, G2, 16, Rt, 16, Tsub(jj), ldT, s);
        return SOL105_OK;
    };
    // step graphs: valid while the key (addresses, sizes, launch shapes) holds
    StepGraphs &G = W.graphs;
    const bool use_graphs = ctx->graphs && !ctx->profiling;
    if (use_graphs) {
        KeyHash k;
        hash_solve_state(k, M);
        for (const void *p : {(const void *)V.p, (const void *)Wb.p, (const void *)xf.p, (const void *)yf.p,
                              (const void *)T.p, (const void *)part.p, (const void *)small.p, (const void *)H2.p,
                              (const void *)ierr.p})
            k.add(p);
        k.add(n);
        k.add(ndof);
        k.add(b);
        k.add(ldT);
        if (G.key != k.h) {
            G.clear();
            G.key = k.h;
        } else {
            G.armed = true;
        }
    }
    auto run_step = [&](int jj) {
        if (!use_graphs || !G.armed) return fast_step(jj);
        if ((size_t)jj >= G.exec.size()) G.exec.resize(jj + 1, nullptr);
        if (!G.exec[jj]) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = fast_step(jj);
            hipError_t e = hipStreamEndCapture(s, &g);
            hipGraphExec_t ex = nullptr;
            if (rc == SOL105_OK && e == hipSuccess) e = hipGraphInstantiateWithFlags(&ex, g, 0);
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
            HIPCHK(e);
            G.exec[jj] = ex;
        }
        HIPCHK(hipGraphLaunch(G.exec[jj], s));
        return SOL105_OK;
    };
    start_block(n, b, V.p, n, 0, s);
    // warm start: the Ritz vectors of the last solve of this model (same
    // free set and elimination order) replace the first columns of the start
    // block; design iterations move them little, so fewer steps are needed
    int seeded = 0;
    if (ctx->warm_start && M->y_cols > 0 && M->y_sym == M->sym_gen && d_Y) {
        seeded = std::min(b, M->y_cols);
        HIPCHK(hipMemcpyAsync(V.p, d_Y, sizeof(double) * n * seeded, hipMemcpyDeviceToDevice, s));
    }
    M->y_cols = 0;
    cholqr2(V.p, nullptr, 0, nullptr, 0, false);
    int j = 0;
    // first Rayleigh-Ritz check: where the last solve of this model converged
    // (design loops and repeated solves; one block earlier when warm-started,
    // since a check costs about two steps), else at about 3 vectors per
    // wanted pair
    const bool warm = M->warm_m > 0 && M->warm_b == b;
    // a warm-started solve tries one block earlier only while that keeps
    // working (the previous warm-started solve converged at its first check)
    const bool early = seeded && warm && M->warm_first_ok;
    int next_check = early  ? std::min(mmax, std::max(2 * b, M->warm_m - b))
                     : warm ? std::min(mmax, std::max(b, M->warm_m))
                            : std::min(mmax, std::max(2 * b, ((kneed * 3 + b - 1) / b) * b));
    const int first_check = next_check;
    out.seeded = seeded;
    std::vector<double> hR;
    bool converged = false;
    int m = 0, injected = 0, restarts = 0, generated = 0;
    // Thick restart (the basis is full and the wanted pairs have not
    // converged): V <- [V Z_kk, Q], T <- [diag(theta_kk) 0; R Z_kk(last b rows, :) ...].
    // With Z the exact Ritz vectors of T, C V Z = V Z diag(theta) + Q R E^T Z,
    // so the Lanczos relation holds for the new basis with that arrowhead T;
    // the next expansion fills T's new column block through CGS2.
    auto restart = [&]() {
        symmetrize_copy(m, ldT, T.p, A.p, s);
        // block tridiagonal (half-bandwidth b) until the first restart
        sym_eig_top(m, A.p, kk, thd.p, Zs.p, ework.p, (int64_t)ework.n, sw.p, s, restarts ? 0 : b, algebraic);
        std::vector<double> th(kk), zl((size_t)b * kk), R((size_t)b * b);
        HIPCHK(hipMemcpyAsync(th.data(), thd.p, sizeof(double) * kk, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpy2DAsync(zl.data(), b * sizeof(double), Zs.p + (m - b), (size_t)m * sizeof(double),
                                b * sizeof(double), kk, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpy2DAsync(R.data(), b * sizeof(double), T.p + (size_t)m + (size_t)(m - b) * ldT,
                                ldT * sizeof(double), b * sizeof(double), b, hipMemcpyDeviceToHost, s));
        int eig_info = 0;
        HIPCHK(hipMemcpyAsync(&eig_info, sw.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(Ytmp.alloc((size_t)n * kk));
        for (int c0 = 0; c0 < kk; c0 += 16)
            ts_gemm_nn(n, m, std::min(16, kk - c0), V.p, n, Zs.p + (size_t)c0 * m, m, Ytmp.p + (size_t)c0 * n, n,
                       1.0, 0.0, s);
        HIPCHK(hipStreamSynchronize(s));
        if (eig_info != 0) return set_err(ctx, SOL105_E_NOCONV, "restart: inverse iteration failed (m=%d)", m);
        HIPCHK(hipMemcpyAsync(V.p, Ytmp.p, sizeof(double) * n * kk, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(V.p + (size_t)kk * n, blk(j), sizeof(double) * n * b, hipMemcpyDeviceToDevice, s));
        // arrowhead: (kk + b) x kk leading block of T, column-major
        std::vector<double> H((size_t)(kk + b) * kk, 0.0);
        for (int i = 0; i < kk; i++) {
            H[(size_t)i * (kk + b) + i] = th[i];
            for (int p = 0; p < b; p++) {
                double v = 0.0;
                for (int q = 0; q < b; q++) v += R[(size_t)q * b + p] * zl[(size_t)i * b + q];
                H[(size_t)i * (kk + b) + kk + p] = v;
            }
        }
        dev_fill(T.p, (int64_t)ldT * ldT, 0.0, s);
        HIPCHK(hipMemcpy2DAsync(T.p, ldT * sizeof(double), H.data(), (kk + b) * sizeof(double),
                                (kk + b) * sizeof(double), kk, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(ierr.p, 0, sizeof(int) * 4, s));
        HIPCHK(hipStreamSynchronize(s));
        j = kk / b;
        m = kk;
        restarts++;
        return SOL105_OK;
    };
    bool step = true;
    while (true) {
        if (step) {
            int rc = run_step(j);
            if (rc) return rc;
            j++;
            generated += b;
            m = j * b;
        }
        step = true;
        bool last = (m >= mmax) && restarts >= max_restarts;
        bool full = (m >= mmax);
        if (!(m >= next_check || full)) continue;
        symmetrize_copy(m, ldT, T.p, A.p, s);
        // the kchk Ritz pairs of largest |theta| (A is overwritten)
        const int kchk = std::min(kneed, m);
        sym_eig_top(m, A.p, kchk, thd.p, Zs.p, ework.p, (int64_t)ework.n, sw.p, s, restarts ? 0 : b, algebraic);
        HIPCHK(hipGetLastError());
        std::vector<double> th(kchk), hZl((size_t)b * kchk);
        HIPCHK(hipMemcpyAsync(th.data(), thd.p, sizeof(double) * kchk, hipMemcpyDeviceToHost, s));
        int eig_info = 0;
        HIPCHK(hipMemcpyAsync(&eig_info, sw.p, sizeof(int), hipMemcpyDeviceToHost, s));
        // last block row of the Ritz vectors in the Lanczos basis
        HIPCHK(hipMemcpy2DAsync(hZl.data(), b * sizeof(double), Zs.p + (m - b), (size_t)m * sizeof(double),
                                b * sizeof(double), kchk, hipMemcpyDeviceToHost, s));
        hR.assign((size_t)b * b, 0.0);
        if (m < b || m > mmax || (size_t)ldT * ldT > T.n)
            return set_err(ctx, SOL105_E_DEVICE, "block Lanczos: basis size %d outside [%d, %d] (T %zu, ld %d)", m,
                           b, mmax, T.n, ldT);
        if (debug)
            fprintf(stderr, "[sol105] check m=%d j=%d mmax=%d kchk=%d restarts=%d T.n=%zu ldT=%d\n", m, j, mmax, kchk,
                    restarts, T.n, ldT);
        HIPCHK(hipMemcpy2DAsync(hR.data(), b * sizeof(double), T.p + (size_t)m + (size_t)(m - b) * ldT,
                                ldT * sizeof(double), b * sizeof(double), b, hipMemcpyDeviceToHost, s));
        int herr[4];
        HIPCHK(hipMemcpyAsync(herr, ierr.p, sizeof(herr), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (herr[0] & 4) {
            // a block became rank-deficient: rebuild it and continue from there
            int jb = herr[1] - 1;
            HIPCHK(hipMemsetAsync(ierr.p, 0, sizeof(int) * 4, s));
            int rank = b;
            int rc = robust_step(jb - 1, rank);
            if (rc) return rc;
            injected += b - rank;
            if (debug) fprintf(stderr, "[sol105] block %d rank-deficient: rank %d, %d directions injected\n", jb, rank,
                               b - rank);
            j = jb;
            m = j * b;
            // the rebuilt block was the residual block of a full basis: that
            // basis is checked again (and restarted), not extended past V
            if (m >= mmax) step = false;
            continue;
        }
        if (herr[0] & 2)
            return set_err(ctx, SOL105_E_NOCONV, "block Lanczos: Cholesky-QR breakdown at %d vectors", m);
        if (eig_info != 0)
            return set_err(ctx, SOL105_E_NOCONV, "projected eigenproblem: inverse iteration failed (m=%d)", m);
        // residual of Ritz pair t: || R_j * Z[m-b:m, t] ||
        auto resid = [&](int t) {
            double r2 = 0.0;
            for (int p = 0; p < b; p++) {
                double v = 0.0;
                for (int q = 0; q < b; q++) v += hR[(size_t)q * b + p] * hZl[(size_t)t * b + q];
                r2 += v * v;
            }
            return std::sqrt(r2);
        };
        bool ok = true;
        if (debug) {
            fprintf(stderr, "[sol105] RR m=%d theta:", m);
            for (int t = 0; t < std::min(kchk, 6); t++) fprintf(stderr, " %.6e(r=%.2e)", th[t], resid(t));
            fprintf(stderr, "\n");
        }
        for (int t = 0; t < kchk; t++) {
            double r = resid(t);
            if (!std::isfinite(th[t]) || !std::isfinite(r)) { ok = false; break; }
            if (std::fabs(th[t]) < 1e-300) continue;
            if (r > tol * std::fabs(th[t])) { ok = false; break; }
        }
        if (ok || last) {
            converged = ok;
            int k = std::min(want, kchk);
            out.theta.assign(th.begin(), th.begin() + k);
            out.resid.clear();
            for (int t = 0; t < k; t++) out.resid.push_back(resid(t));
            // Ritz vectors Y = V_m Z[:, 0:k], 16 columns per product
            for (int c0 = 0; c0 < k; c0 += 16)
                ts_gemm_nn(n, m, std::min(16, k - c0), V.p, n, Zs.p + (size_t)c0 * m, m, d_Y + (size_t)c0 * n, n,
                           1.0, 0.0, s);
            HIPCHK(hipStreamSynchronize(s));
            break;
        }
        if (full) {
            int rc = restart();
            if (rc) return rc;
            if (debug) fprintf(stderr, "[sol105] thick restart %d: kept %d Ritz vectors\n", restarts, kk);
            next_check = std::min(mmax, m + 4 * b);
            continue;
        }
        next_check = std::min(mmax, m + (warm ? b : 4 * b));
    }
    out.m = m;
    out.generated = generated;
    out.restarts = restarts;
    out.steps = j;
    out.injected = injected;
    if (!converged) return set_err(ctx, SOL105_E_NOCONV, "block Lanczos did not converge in %d vectors", m);
    M->warm_first_ok = seeded && warm && restarts == 0 && m == first_check;
    M->warm_m = restarts ? mmax : m;
    M->warm_b = b;
    M->y_cols = (int)out.theta.size();
    M->y_sym = M->sym_gen;
    return SOL105_OK;
}

// The k eigenpairs of largest |theta| (algebraic: of largest theta) of
// C = L^-1 (-KG) L^-T for a free set of at most SYMEIG_MAX_M DOFs, k <=
// SYMEIG_MAX_K: C is formed column block by column block from the operator
// applied to the identity (the same products as a Lanczos step),
// symmetrised, and handed whole to the projected eigensolver; Y = the
// eigenvectors (the basis is the identity).
int dense_pairs(sol105_ctx *ctx, sol105_model *M, int k, bool algebraic, double *d_Y, LanczosOut &out) {
    hipStream_t s = ctx->stream;
    const int n = M->sym.n_free, ndof = M->ndof();
    LanczosWork &W = M->lw;
    HIPCHK(W.V.alloc((size_t)n * n));
    HIPCHK(W.Wb.alloc((size_t)n * MAXR));
    HIPCHK(W.xf.alloc((size_t)ndof * MAXR));
    HIPCHK(W.yf.alloc((size_t)ndof * MAXR));
    HIPCHK(W.A.alloc((size_t)n * n));
    HIPCHK(W.Zs.alloc((size_t)n * k));
    HIPCHK(W.thd.alloc(k));
    HIPCHK(W.ework.alloc(symeig_work_size(n, k)));
    HIPCHK(W.sw.alloc(1));
    HIPCHK(W.Ytmp.alloc((size_t)n * MAXR));
    std::vector<double> eye((size_t)n * MAXR);
    for (int c0 = 0; c0 < n; c0 += MAXR) {
        const int b = std::min(MAXR, n - c0);
        std::fill(eye.begin(), eye.end(), 0.0);
        for (int j = 0; j < b; j++) eye[(size_t)j * n + c0 + j] = 1.0;
        HIPCHK(hipMemcpyAsync(W.Ytmp.p, eye.data(), sizeof(double) * n * b, hipMemcpyHostToDevice, s));
        HIPCHK(solve_bwd(M->plan, M->dev, M->dsolve, M->splan, M->d_PT.p, M->d_W.p, M->d_part.p, W.Ytmp.p, W.Wb.p,
                         b, s, ctx->timer()));
        launch_scatter_elim_il(n, ndof, b, M->d_elim_of_dof.p, W.Wb.p, W.xf.p, s);
        launch_bsr_spmv(M->n_nodes, M->d_rp.p, M->d_col.p, M->d_KG.p, W.xf.p, W.yf.p, b, -1.0, s);
        launch_gather_elim_il(n, b, M->d_dof_of_elim.p, W.yf.p, W.Wb.p, s);
        HIPCHK(solve_fwd(M->plan, M->dev, M->dsolve, M->splan, M->d_P.p, M->d_W.p, M->d_part.p, W.Wb.p,
                         W.V.p + (size_t)c0 * n, b, s, ctx->timer()));
        HIPCHK(hipStreamSynchronize(s));  // eye is reused by the next block
    }
    symmetrize_copy(n, n, W.V.p, W.A.p, s);
    sym_eig_top(n, W.A.p, k, W.thd.p, W.Zs.p, W.ework.p, (int64_t)W.ework.n, W.sw.p, s, 0, algebraic ? 1 : 0);
    HIPCHK(hipGetLastError());
    std::vector<double> th(k);
    int info = 0;
    HIPCHK(hipMemcpyAsync(th.data(), W.thd.p, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&info, W.sw.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(d_Y, W.Zs.p, sizeof(double) * n * k, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    if (info != 0) return set_err(ctx, SOL105_E_NOCONV, "dense eigensolve failed (n=%d)", n);
    out.theta = th;
    out.m = n;
    out.generated = n;
    M->y_cols = 0;
    return SOL105_OK;
}

}  // namespace

int sol105_buckling(sol105_ctx *ctx, sol105_model *M, int32_t max_modes, double *lambda_out, double *modes_out,
                    int32_t *n_out) {
    if (!ctx || !M || !n_out) return set_err(ctx, SOL105_E_INVALID, "null argument");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    if (!M->factor_valid) return set_err(ctx, SOL105_E_STATE, "no factorization: run sol105_static_solve first");
    if (!M->KG_valid) return set_err(ctx, SOL105_E_STATE, "no KG: run sol105_assemble_KG first");
    const int n = M->sym.n_free, ndof = M->ndof();
    ModeSelect sel;
    sel.nd = M->nd;
    sel.n_free = n;
    sel.v1 = M->v1;
    sel.v2 = M->v2;
    sel.n_shell_cards = M->n_shell_cards;
    // a free set of at most SYMEIG_MAX_K DOFs: the dense path on the device,
    // every pair of C = L^-1 (-KG) L^-T, as the reference's dense eigensolve
    sel.all_pairs = n <= SYMEIG_MAX_K;
    // the dense path under a range keeps the whole spectrum, of which only
    // the positive roots in [V1, V2] can be output: the eigensolve targets
    // the largest theta (the smallest positive lambda) there.  Up to
    // SYMEIG_MAX_M free DOFs C is formed whole and its SYMEIG_MAX_K largest
    // pairs taken at once (a Lanczos basis would span most of the space);
    // above that block Lanczos, ceil(3 ND) pairs first and, when those do not
    // decide the output, as many as the projected eigensolver takes
    const bool grow = sel.dense() && sel.has_range() && !sel.all_pairs;
    const bool whole = sel.all_pairs || (grow && n <= SYMEIG_MAX_M);
    const int k_whole = sel.all_pairs ? n : std::min(n, SYMEIG_MAX_K);
    const int req_max = std::min(SYMEIG_MAX_K - 2, std::max(n - 2, 1));
    int req = sel.requested();
    if (!sel.all_pairs && req > SYMEIG_MAX_K - 2) {
        if (!grow) return set_err(ctx, SOL105_E_UNSUPPORTED, "more than %d requested modes", SYMEIG_MAX_K - 2);
        req = req_max;
    }
    DevBuf<double> &Y = M->lw.Y, &Xf = M->lw.Xf;
    std::vector<double> lam;
    std::vector<int> pos, surv;
    bool decided = true;
    for (;;) {
        const int ycols = whole ? k_whole : req;
        if (Y.n < (size_t)n * ycols) M->y_cols = 0;  // reallocated: no kept Ritz vectors
        HIPCHK(Y.alloc((size_t)n * ycols));
        LanczosOut lo;
        int rc = whole ? dense_pairs(ctx, M, k_whole, grow, Y.p, lo) : lanczos(ctx, M, req, Y.p, lo, grow);
        if (rc) return rc;
        const int k = (int)lo.theta.size();
        M->lanczos_m = lo.generated;
        M->lanczos_steps = lo.steps;
        // x = L^-T y, expand to all DOFs (16 right-hand sides per triangular solve)
        HIPCHK(Xf.alloc((size_t)ndof * std::max(k, 1)));
        if (k > 0) {
            DevBuf<double> &Ys = M->lw.Ys;
            HIPCHK(Ys.alloc((size_t)n * k));
            for (int c0 = 0; c0 < k; c0 += MAXR) {
                const int nc = std::min(MAXR, k - c0);
                HIPCHK(solve_bwd(M->plan, M->dev, M->dsolve, M->splan, M->d_PT.p, M->d_W.p, M->d_part.p,
                                 Y.p + (size_t)c0 * n, Ys.p + (size_t)c0 * n, nc, s, ctx->timer()));
            }
            launch_scatter_elim(n, ndof, k, M->d_elim_of_dof.p, Ys.p, Xf.p, s);
        }
        // candidates: lambda = 1/theta, the requested count, positive, in range, ascending
        pos = select_candidates(sel, lo.theta.data(), k, lam);
        // localization filter (>= 100 shell cards, share 0.10, scan ND + 16)
        surv.clear();
        const int limit = localization_scan_limit(sel, (int)pos.size());
        if (limit > 0) {
            const int n_shell = M->n_q4 + M->n_t3;
            DevBuf<double> &scr = M->lw.scr, &sh = M->lw.sh;
            HIPCHK(scr.alloc(2 * ((n_shell + 255) / 256) + 2));
            HIPCHK(sh.alloc(2 * (size_t)limit));
            for (int t = 0; t < limit; t++)
                elem_energy_share(M->n_q4, M->d_q4.p, M->n_t3, M->d_t3.p, ndof, Xf.p, pos[t], scr.p, sh.p + 2 * t, s);
            std::vector<double> hsh(2 * (size_t)limit);
            HIPCHK(hipMemcpyAsync(hsh.data(), sh.p, hsh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (int t = 0; t < limit; t++) {
                double mx = hsh[2 * t], tot = hsh[2 * t + 1];
                if (tot > 0 && mx / tot > 0.10) continue;
                surv.push_back(pos[t]);
            }
            for (int t = limit; t < (int)pos.size(); t++) surv.push_back(pos[t]);
        } else {
            surv = pos;
        }
        // every pair with |lambda| below the smallest computed |theta|'s
        // inverse is known (for the largest theta: every positive root below
        // the smallest computed theta's inverse, all of them once a computed
        // theta is not positive)
        double cover = INFINITY;
        if (!sel.all_pairs)
            for (int i = 0; i < k; i++) {
                if (grow && !(lo.theta[i] > 0.0)) { cover = INFINITY; break; }
                cover = std::min(cover, 1.0 / std::fabs(lo.theta[i]));
            }
        decided = selection_decided(sel, surv, lam, cover);
        M->sel_info[0] = k;
        M->sel_info[1] = (double)pos.size();
        if (decided || !grow || whole || req >= req_max) break;
        req = req_max;
    }
    if (!decided && grow)
        return set_err(ctx, SOL105_E_UNSUPPORTED,
                       "EIGRL range: the %d smallest positive roots do not decide the output (%zu in range)",
                       whole ? k_whole : req, pos.size());
    M->sel_info[2] = cluster_skip(sel, surv, lam);
    M->sel_info[3] = decided ? 1.0 : 0.0;
    // mode count and cluster filter
    std::vector<int> outp = select_output(sel, surv, lam);
    const int nout = std::min((int)outp.size(), (int)max_modes);
    HIPCHK(M->d_modes.alloc((size_t)ndof * std::max(nout, 1)));
    M->lambda.assign(nout, 0.0);
    DevBuf<double> &scr = M->lw.scr;
    HIPCHK(scr.alloc(std::max<size_t>(1024, scr.n)));
    for (int t = 0; t < nout; t++) {
        int i = outp[t];
        M->lambda[t] = lam[i];
        HIPCHK(hipMemcpyAsync(M->d_modes.p + (size_t)t * ndof, Xf.p + (size_t)i * ndof, ndof * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
        // dependent DOFs of the mode before its normalisation (solve_case.jl:3153-3162)
        if (M->has_mpc())
            launch_mpc_recover((int)M->mpc.dep.size(), M->d_mpc_dep.p, M->d_mpc_ptr.p, M->d_mpc_ind.p,
                               M->d_mpc_coef.p, M->d_modes.p + (size_t)t * ndof, ndof, 1, s);
        normalize_maxabs(ndof, M->d_modes.p + (size_t)t * ndof, scr.p, s);
    }
    M->n_modes = nout;
    *n_out = nout;
    if (lambda_out)
        for (int t = 0; t < nout; t++) lambda_out[t] = M->lambda[t];
    if (modes_out && nout > 0)
        HIPCHK(hipMemcpyAsync(modes_out, M->d_modes.p, (size_t)ndof * nout * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    return SOL105_OK;
}

int sol105_buckling_info(const sol105_ctx *ctx, const sol105_model *M, double *out) {
    (void)ctx;
    if (!M || !out) return SOL105_E_INVALID;
    std::memcpy(out, M->sel_info, sizeof(M->sel_info));
    return SOL105_OK;
}

int sol105_solve(sol105_ctx *ctx, sol105_model *M, int32_t max_modes, double *lambda_out, double *modes_out,
                 int32_t *n_out) {
    if (!ctx || !M || !n_out) return set_err(ctx, SOL105_E_INVALID, "null argument");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    int rc = sol105_assemble_K(ctx, M, nullptr);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    rc = factorize(ctx, M, true);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[3], s));
    rc = sol105_static_solve(ctx, M, nullptr, nullptr);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[4], s));
    rc = sol105_assemble_KG(ctx, M, nullptr, nullptr);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[5], s));
    rc = sol105_buckling(ctx, M, max_modes, lambda_out, modes_out, n_out);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[6], s));
    HIPCHK(hipEventSynchronize(ctx->ev[6]));
    ctx->timings[0] = ev_ms(ctx->ev[0], ctx->ev[1]);
    ctx->timings[1] = 0.0;
    ctx->timings[2] = ev_ms(ctx->ev[1], ctx->ev[3]);
    ctx->timings[3] = ev_ms(ctx->ev[3], ctx->ev[4]);
    ctx->timings[4] = ev_ms(ctx->ev[4], ctx->ev[5]);
    ctx->timings[5] = ev_ms(ctx->ev[5], ctx->ev[6]);
    ctx->timings[6] = 0.0;
    ctx->timings[7] = ev_ms(ctx->ev[0], ctx->ev[6]);
    ctx->stats[0] = M->sym.n_free;
    ctx->stats[1] = (double)M->sym.nnz_L;
    ctx->stats[2] = M->sym.flops;
    ctx->stats[3] = (double)M->sym.fronts.size();
    ctx->stats[4] = M->sym.height;
    ctx->stats[5] = M->lanczos_m;
    ctx->stats[6] = M->lanczos_steps;
    ctx->stats[7] = (double)M->sym.front_storage * 8.0;
    return SOL105_OK;
}

}  // extern "C"

namespace {

// Shared part of the sensitivity entry points: adjoint load g (ndof, global
// frame, zero on fixed DOFs), psi = K_ff^-1 g, the three per-element terms
// and phi' KG phi.
int sens_terms(sol105_ctx *ctx, sol105_model *M, int32_t mode, DevBuf<double> &terms, DevBuf<double> &g,
               double &pkp) {
    if (!ctx || !M) return set_err(ctx, SOL105_E_INVALID, "null argument");
    if (mode < 0 || mode >= M->n_modes) return set_err(ctx, SOL105_E_STATE, "mode %d not available", mode);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    int ndof = M->ndof(), n = M->sym.n_free;
    const double *phi = M->d_modes.p + (size_t)mode * ndof;
    double lam = M->lambda[mode];
    ElemArgs a = elem_args(M);
    const int n_shell = M->n_q4 + M->n_t3;
    DevBuf<double> egq, egt, egb, ge, psi, kgphi, scr, dots, phif;
    HIPCHK(egq.alloc((size_t)std::max(M->n_q4, 1) * 24));
    HIPCHK(egt.alloc((size_t)std::max(M->n_t3, 1) * 18));
    HIPCHK(egb.alloc((size_t)std::max(M->n_bar, 1) * 12));
    HIPCHK(g.alloc(ndof));
    HIPCHK(ge.alloc(std::max(n, 1)));
    HIPCHK(psi.alloc(ndof));
    HIPCHK(kgphi.alloc(ndof));
    HIPCHK(scr.alloc(512));
    HIPCHK(dots.alloc(2));
    HIPCHK(phif.alloc(ndof));
    HIPCHK(terms.alloc((size_t)std::max(n_shell, 1) * 3));
    // reference mode: six condensed membrane modes in a model without line
    // elements (adjoint_buckling.jl:950-955), else the four of the forward KG
    const bool ref = ctx->sens_mode == 1, six = M->n_bar == 0;
    if (ref) launch_q4_adjoint_rhs_ref(a, M->d_u.p, phi, lam, six, egq.p, s);
    else launch_q4_adjoint_rhs(a, M->d_u.p, phi, lam, egq.p, s);
    launch_t3_kg_sens(a, M->d_u.p, phi, lam, terms.p, egt.p, s);
    launch_bar_adjoint_rhs(a, phi, lam, egb.p, s);
    // with RBE2 / RBE3 constraints the exact adjoint load is T' g: the
    // entries on dependent DOFs fold onto their independent DOFs, and psi
    // gets its dependent DOFs back (psi = T psi_ind).  The reference's path
    // drops those entries and keeps psi zero there (adjoint_buckling.jl:
    // 2490-2493 restrict g to the free DOFs, which exclude dependent ones).
    const bool fold = M->has_mpc() && !ref;
    node_gather_vec(M->n_nodes, M->n_q4, M->n_t3, M->d_diag.p, M->d_cptr.p, M->d_cidx.p, egq.p, egt.p, egb.p,
                    fold ? nullptr : M->d_free.p, g.p, s);
    if (fold)
        launch_mpc_fold(M->mpc_nrecv, M->d_mpc_frecv.p, M->d_mpc_fptr.p, M->d_mpc_fdep.p, M->d_mpc_fcoef.p,
                        M->d_free.p, ndof, g.p, s);
    launch_gather_elim(n, ndof, 1, M->d_dof_of_elim.p, g.p, ge.p, s);
    HIPCHK(tri_solve(M, ge.p, 1, s));
    launch_scatter_elim(n, ndof, 1, M->d_elim_of_dof.p, ge.p, psi.p, s);
    if (fold)
        launch_mpc_recover((int)M->mpc.dep.size(), M->d_mpc_dep.p, M->d_mpc_ptr.p, M->d_mpc_ind.p, M->d_mpc_coef.p,
                           psi.p, ndof, 1, s);
    // phi' KG phi on the free set; phi on the free set plus its dependent
    // DOFs (the element terms take the whole mode, adjoint_buckling.jl:2472)
    launch_gather_elim(n, ndof, 1, M->d_dof_of_elim.p, phi, ge.p, s);
    launch_scatter_elim(n, ndof, 1, M->d_elim_of_dof.p, ge.p, phif.p, s);
    if (M->has_mpc())
        launch_mpc_recover((int)M->mpc.dep.size(), M->d_mpc_dep.p, M->d_mpc_ptr.p, M->d_mpc_ind.p, M->d_mpc_coef.p,
                           phif.p, ndof, 1, s);
    launch_bsr_spmv(M->n_nodes, M->d_rp.p, M->d_col.p, M->d_KG.p, phif.p, kgphi.p, 1, 1.0, s);
    dot_det(ndof, phif.p, kgphi.p, scr.p, dots.p, s);
    if (ref) launch_q4_sens_terms_ref(a, M->d_u.p, phif.p, psi.p, six, terms.p, s);
    else launch_q4_sens_terms(a, M->d_u.p, phif.p, psi.p, terms.p, s);
    launch_t3_sens_terms(a, M->d_u.p, phif.p, psi.p, terms.p, s);
    HIPCHK(hipMemcpyAsync(&pkp, dots.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    return SOL105_OK;
}

}  // namespace

extern "C" {

int sol105_sensitivity_thickness(sol105_ctx *ctx, sol105_model *M, int32_t mode, double *dlam_dt) {
    if (!dlam_dt) return set_err(ctx, SOL105_E_INVALID, "null argument");
    DevBuf<double> terms, g, dl;
    double pkp = 0.0;
    int rc = sens_terms(ctx, M, mode, terms, g, pkp);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    const int n_shell = M->n_q4 + M->n_t3;
    HIPCHK(dl.alloc(std::max(n_shell, 1)));
    launch_sens_combine(n_shell, terms.p, M->lambda[mode], pkp, dl.p, s);
    HIPCHK(hipMemcpyAsync(dlam_dt, dl.p, sizeof(double) * n_shell, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SOL105_OK;
}

int sol105_sensitivity_terms(sol105_ctx *ctx, sol105_model *M, int32_t mode, double *terms_out,
                             double *adjoint_load, double *phi_kg_phi) {
    DevBuf<double> terms, g;
    double pkp = 0.0;
    int rc = sens_terms(ctx, M, mode, terms, g, pkp);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    if (terms_out)
        HIPCHK(hipMemcpyAsync(terms_out, terms.p, sizeof(double) * 3 * (M->n_q4 + M->n_t3), hipMemcpyDeviceToHost, s));
    if (adjoint_load) HIPCHK(hipMemcpyAsync(adjoint_load, g.p, sizeof(double) * M->ndof(), hipMemcpyDeviceToHost, s));
    if (phi_kg_phi) *phi_kg_phi = pkp;
    HIPCHK(hipStreamSynchronize(s));
    return SOL105_OK;
}

int sol105_profile_enable(sol105_ctx *ctx, int32_t on) {
    if (!ctx) return SOL105_E_INVALID;
    (void)hipSetDevice(ctx->device);
    ctx->ptimer.reset();
    ctx->profiling = on != 0;
    return SOL105_OK;
}

int sol105_symeig_top(sol105_ctx *ctx, int32_t m, const double *a, int32_t k, int32_t bw, int32_t algebraic,
                      double *theta_out, double *z_out) {
    if (!ctx || !a || !theta_out || !z_out) return set_err(ctx, SOL105_E_INVALID, "null argument");
    if (m < 1 || m > SYMEIG_MAX_M) return set_err(ctx, SOL105_E_INVALID, "m = %d out of range 1..%d", m, SYMEIG_MAX_M);
    if (k < 1 || k > std::min<int>(m, SYMEIG_MAX_K))
        return set_err(ctx, SOL105_E_INVALID, "k = %d out of range 1..%d", k, std::min<int>(m, SYMEIG_MAX_K));
    if (bw != 0 && (bw < 4 || bw > SYMEIG_BAND_MAX_W))
        return set_err(ctx, SOL105_E_INVALID, "bw = %d: 0 or 4..%d", bw, SYMEIG_BAND_MAX_W);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    DevBuf<double> A, th, Z, work;
    DevBuf<int> info;
    HIPCHK(A.upload(a, (size_t)m * m, s));
    HIPCHK(th.alloc(k));
    HIPCHK(Z.alloc((size_t)m * k));
    const int64_t wn = symeig_work_size(m, k);
    HIPCHK(work.alloc(wn));
    HIPCHK(info.alloc(1));
    HIPCHK(hipMemsetAsync(info.p, 0, sizeof(int), s));
    sym_eig_top(m, A.p, k, th.p, Z.p, work.p, wn, info.p, s, bw, algebraic);
    HIPCHK(hipGetLastError());
    int hinfo = 0;
    HIPCHK(hipMemcpyAsync(theta_out, th.p, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(z_out, Z.p, sizeof(double) * m * k, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hinfo != 0) return set_err(ctx, SOL105_E_NOCONV, "projected eigensolver failed (info %d)", hinfo);
    return SOL105_OK;
}

int sol105_profile_read(sol105_ctx *ctx, double *out) {
    if (!ctx || !out) return SOL105_E_INVALID;
    (void)hipSetDevice(ctx->device);
    ctx->ptimer.collect();
    const ProdTimer &T = ctx->ptimer;
    out[0] = (double)(T.launches[0] + T.launches[1]);
    out[1] = T.total_ms[0] + T.total_ms[1];
    out[2] = T.total_bytes[0] + T.total_bytes[1];
    out[3] = (double)T.launches[0];
    out[4] = T.total_ms[0];
    out[5] = T.total_bytes[0];
    out[6] = T.total_ms[1];
    out[7] = T.total_bytes[1];
    return SOL105_OK;
}

int sol105_last_timings(const sol105_ctx *ctx, double *ms) {
    if (!ctx || !ms) return SOL105_E_INVALID;
    std::memcpy(ms, ctx->timings, sizeof(ctx->timings));
    return SOL105_OK;
}

int sol105_last_stats(const sol105_ctx *ctx, double *st) {
    if (!ctx || !st) return SOL105_E_INVALID;
    std::memcpy(st, ctx->stats, sizeof(ctx->stats));
    return SOL105_OK;
}

}  // extern "C"
