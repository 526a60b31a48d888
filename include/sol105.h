/*
 * sol105.h -- C ABI of the MI355X SOL 105 hot path (libsol105_hip.so).
 *
 * The functions replace, for the supported decks, the reference's Julia
 * entry points listed beside each declaration (paths relative to the
 * reference root).  Conventions:
 *   - every function returns SOL105_OK (0) or a negative SOL105_E_* code;
 *     sol105_last_error(ctx) gives the message of the last failure;
 *   - host buffers are caller-owned; device memory is owned by the context;
 *   - one context per host thread (a context serialises its own work on its
 *     own HIP stream);
 *   - node indices are 0-based ranks of the GRID ids sorted ascending
 *     (JFEM/src/solver/assembly.jl:1958-1961); DOF index = 6*rank + component;
 *   - matrices are returned in 6x6 block-sparse-row form (BSR) over all
 *     6*n_nodes DOFs: block row pointers over nodes, block column indices,
 *     36 values per block stored row-major.
 */
#ifndef SOL105_H
#define SOL105_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOL105_OK 0
#define SOL105_E_INVALID (-1)      /* bad argument / inconsistent model arrays  */
#define SOL105_E_UNSUPPORTED (-2)  /* feature outside the implemented branch    */
#define SOL105_E_DEVICE (-3)       /* HIP runtime error or no device            */
#define SOL105_E_SINGULAR (-4)     /* stiffness not positive definite           */
#define SOL105_E_NOCONV (-5)       /* eigensolver did not converge              */
#define SOL105_E_STATE (-6)        /* call order violated (e.g. KG before u)    */
#define SOL105_E_PENDING (-7)      /* server request not finished yet           */

typedef struct sol105_ctx sol105_ctx;
typedef struct sol105_model sol105_model;

/* Model arrays (structure of arrays).  All pointers are host pointers and are
 * copied during sol105_model_upload.  Replaces the model Dict consumed by
 * Solver.assemble_stiffness (JFEM/src/solver/assembly.jl:1953-2530). */
typedef struct sol105_model_desc {
    int32_t n_nodes;
    const double *xyz;          /* n_nodes x 3 basic-frame coordinates            */
    int32_t n_q4;
    const int32_t *q4_nodes;    /* n_q4 x 4 node ranks (CQUAD4 G1..G4)            */
    const double *q4_prop;      /* n_q4 x 6: E, nu, T, 12I/T^3, TS/T, G_shear     */
    int32_t n_t3;
    const int32_t *t3_nodes;    /* n_t3 x 3 node ranks (CTRIA3 G1..G3)            */
    const double *t3_prop;      /* n_t3 x 6, columns as q4_prop                   */
    int32_t n_bar;
    const int32_t *bar_nodes;   /* n_bar x 2 node ranks (GA, GB)                  */
    const double *bar_vref;     /* n_bar x 3 orientation vector (V, or G0-GA)     */
    const double *bar_prop;     /* n_bar x 8: A, I1, I2, J, K1, K2, E, G          */
    const uint8_t *spc_mask;    /* 6*n_nodes, 1 = constrained by SPC1/SPC/PS      */
    const double *force;        /* 6*n_nodes applied load (analysis frame)        */
    const uint8_t *node_has_shell; /* n_nodes (AUTOSPC rotational topology)       */
    const uint8_t *node_has_bar;   /* n_nodes                                     */
    double k6rot;               /* PARAM,K6ROT (default 100)                      */
    int32_t autospc;            /* PARAM,AUTOSPC (default 1)                      */
    int32_t nd;                 /* EIGRL ND                                       */
    double eigrl_v1;            /* EIGRL V1, V2: a range when V2 > V1 and either  */
    double eigrl_v2;            /*   is non-zero (blank fields: 0)                */
    int32_t n_shell_cards;      /* shell cards of the deck (0: n_q4 + n_t3); the  */
                                /*   mode filters act on decks with >= 100        */
    /* Rigid and interpolation constraints, eliminated as dependent DOFs
     * (JFEM/src/solver/constraints.jl:37-364): RBE2 (and RBAR, converted
     * to RBE2 as ModelBuilder.jl:708-725 does) as master-slave pairs, RBE3
     * as a reference grid with weight groups.  Component sets are the
     * Nastran digit strings read as integers (123456).  All counts may be 0. */
    int32_t n_rbe2;             /* master-slave pairs                             */
    const int32_t *rbe2;        /* n_rbe2 x 3: master rank, CM, slave rank        */
    int32_t n_rbe3;             /* RBE3 elements                                  */
    const int32_t *rbe3;        /* n_rbe3 x 4: REFGRID rank, REFC, first group,   */
                                /*   group count                                  */
    int32_t n_rbe3_grp;         /* weight groups                                  */
    const int32_t *rbe3_grp;    /* n_rbe3_grp x 3: components, first node, count  */
    const double *rbe3_wt;      /* n_rbe3_grp weights                             */
    int32_t n_rbe3_node;
    const int32_t *rbe3_node;   /* node ranks of the groups                       */
    /* Explicit MPC equations sum_i A_i u(G_i, C_i) = 0, the first term the
     * dependent DOF (constraints.jl:290-316): u_dep = sum_{i>1} -A_i/A_1
     * u_i.  An MPC replaces RBE2 / RBE3 entries of its dependent DOF, and
     * a later MPC of the same DOF replaces an earlier one (card order). */
    int32_t n_mpc;              /* equations                                      */
    const int32_t *mpc_eq;      /* n_mpc x 2: first term, term count (>= 2)       */
    int32_t n_mpc_term;
    const int32_t *mpc_term;    /* n_mpc_term x 2: node rank, component 1..6      */
    const double *mpc_coef;     /* n_mpc_term coefficients A_i                    */
    /* RSPLINE (constraints.jl:146-213): each dependent (GRID, component) is
     * projected on the polyline of the independent GRIDs and follows the
     * two ends of the nearest segment linearly.  Replaces RBE2 / RBE3
     * entries; MPC equations replace it. */
    int32_t n_rspline;
    const int32_t *rspline;     /* n_rspline x 4: first independent, count,       */
                                /*   first dependent, count                       */
    int32_t n_rspline_indep;
    const int32_t *rspline_indep; /* independent node ranks, polyline order       */
    int32_t n_rspline_dep;
    const int32_t *rspline_dep; /* n_rspline_dep x 2: node rank, component 1..6   */
    /* RBE1 (constraints.jl:76-143): the rigid motion about the centroid of
     * the independent GRIDs fitted to the independent DOFs (pseudo-inverse,
     * 1 to 6 DOFs) drives the dependent DOFs.  Replaces RBE2 / RBE3
     * entries; RSPLINE and MPC entries replace it. */
    int32_t n_rbe1;
    const int32_t *rbe1;        /* n_rbe1 x 4: first independent, count,          */
                                /*   first dependent, count                       */
    int32_t n_rbe1_indep;
    const int32_t *rbe1_indep;  /* n_rbe1_indep x 2: node rank, component 1..6    */
    int32_t n_rbe1_dep;
    const int32_t *rbe1_dep;    /* n_rbe1_dep x 2: node rank, component 1..6      */
} sol105_model_desc;

/* Native deck reader: Nastran deck text (executive and case control, bulk
 * data) to the model arrays of this path, host code only.  Replaces the
 * reference's parse and model build for the supported card subset
 * (JFEM/src/parsing/card_processor.jl:51-212, JFEM/src/ModelBuilder.jl:
 * 421-823) and the per-element array derivation of assemble_stiffness
 * (JFEM/src/solver/assembly.jl:1953-2530).  buckling_index picks the
 * buckling subcase (METHOD) when there are several.  Decks outside the
 * supported branch return SOL105_E_UNSUPPORTED, malformed ones
 * SOL105_E_INVALID, with the message in err (err_len bytes, may be NULL). */
typedef struct sol105_deck sol105_deck;
int sol105_deck_read(const char *text, int64_t len, int32_t buckling_index, sol105_deck **out, char *err,
                     int32_t err_len);
int sol105_deck_free(sol105_deck *deck);
/* out: n_nodes, n_q4, n_t3, n_bar, EIGRL ND, buckling subcase id, static
 * subcase id, shell cards of the deck, SOL (101 when the deck names none) */
int sol105_deck_sizes(const sol105_deck *deck, int64_t *out /* 9 */);
/* A model description pointing into the deck's arrays (valid until
 * sol105_deck_free), ready for sol105_model_upload. */
int sol105_deck_desc(const sol105_deck *deck, sol105_model_desc *desc);
/* GRID ids (ascending, the node ranks' ids), element ids and PSHELL ids in
 * array order; any pointer may be NULL. */
int sol105_deck_ids(const sol105_deck *deck, int64_t *node_ids, int64_t *q4_eid, int64_t *q4_pid, int64_t *t3_eid,
                    int64_t *t3_pid, int64_t *bar_eid);
/* Stress recovery points of each CBAR's PBARL section (n_bar x 8: C1 C2 D1 D2
 * E1 E2 F1 F2, JFEM/src/parsing/extract_properties.jl:167-187), the
 * bar_points argument of sol105_static_recover. */
int sol105_deck_bar_points(const sol105_deck *deck, double *out);

/* Device selection and context.  No reference counterpart (the reference is
 * CPU-only); one context owns one HIP stream and its scratch. */
int sol105_ctx_create(int device, sol105_ctx **out);
int sol105_ctx_destroy(sol105_ctx *ctx);
const char *sol105_last_error(const sol105_ctx *ctx);

/* Eigensolver limits of a context, the counterparts of the reference's
 * KrylovKit krylovdim and maxiter (JFEM/src/solver/solve_case.jl:1847-1861):
 *   SOL105_OPT_LANCZOS_MAX_VECTORS   basis vectors before a thick restart,
 *                                    32 .. 1024 (default 1024; rounded down
 *                                    to a multiple of the block size);
 *   SOL105_OPT_LANCZOS_MAX_RESTARTS  thick restarts before SOL105_E_NOCONV,
 *                                    0 .. 1000 (default 40);
 *   SOL105_OPT_WARM_START            0 or 1 (default 0): start the block
 *                                    Lanczos iteration of a model from the
 *                                    Ritz vectors of its previous solve (same
 *                                    free set), as a design loop or a served
 *                                    stream moves them little.  The results
 *                                    then depend on that history within the
 *                                    convergence tolerance, no longer bitwise.
 *                                    No counterpart in the reference.
 *   SOL105_OPT_STEP_GRAPHS           0 or 1 (default 1): replay the block
 *                                    Lanczos expansion steps (both triangular
 *                                    solves, the KG product, CGS2 and
 *                                    Cholesky-QR) as HIP graphs, captured per
 *                                    block index from the second solve of a
 *                                    model (or batch) with unchanged buffers
 *                                    and launch shapes.  Results are bitwise
 *                                    those of direct launches.  The
 *                                    environment variable SOL105_STEP_GRAPHS=0
 *                                    makes 0 the default of new contexts
 *                                    (kernel-trace profilers that do not
 *                                    follow stream capture).  No
 *                                    counterpart in the reference.
 *   SOL105_OPT_SENS_MODE             0 or 1 (default 0): the CQUAD4 terms of
 *                                    sol105_sensitivity_thickness /
 *                                    _terms.  0: the exact derivative of the
 *                                    forward path (dKe/dt of the MacNeal
 *                                    element, d(phi' KG(u) phi)/du through
 *                                    the principal-stress split).  1: the
 *                                    reference's own analytic path
 *                                    (JFEM/src/solver/adjoint_buckling.jl:
 *                                    2406-2637 with its defaults): dKe/dt of
 *                                    the legacy MITC4 element
 *                                    stiffness_quad4_generic (dKdx.jl:219-
 *                                    247), the unit-basis adjoint load
 *                                    phi' KG(sigma(e_j)) phi (adjoint_
 *                                    buckling.jl:1024-1042) and phi' KG phi
 *                                    / t, with six condensed membrane modes
 *                                    when the model has no CBAR, else four.
 *                                    CTRIA3 and CBAR terms are the same in
 *                                    both modes.
 * SOL105_E_INVALID for an unknown option or a value out of range. */
#define SOL105_OPT_LANCZOS_MAX_VECTORS 1
#define SOL105_OPT_LANCZOS_MAX_RESTARTS 2
#define SOL105_OPT_WARM_START 3
#define SOL105_OPT_STEP_GRAPHS 4
#define SOL105_OPT_SENS_MODE 5
int sol105_ctx_set_option(sol105_ctx *ctx, int32_t option, double value);

/* Upload a model, build the BSR pattern, the deterministic element->block
 * scatter map and the fill-reducing ordering.  Replaces the indexing part of
 * Solver.assemble_stiffness (JFEM/src/solver/assembly.jl:1957-1989). */
int sol105_model_upload(sol105_ctx *ctx, const sol105_model_desc *desc, sol105_model **out);
int sol105_model_free(sol105_ctx *ctx, sol105_model *model);

/* Replace the CQUAD4 thickness column (design update) without re-analysis. */
int sol105_model_set_q4_thickness(sol105_ctx *ctx, sol105_model *model, const double *t /* n_q4 */);
/* Replace the thickness of every shell, CQUAD4 first, then CTRIA3 (design
 * update; the pattern, maps and ordering are kept). */
int sol105_model_set_shell_thickness(sol105_ctx *ctx, sol105_model *model, const double *t /* n_q4 + n_t3 */);

/* Pattern query: n_blocks, then block row pointers (n_nodes+1) and block
 * columns (n_blocks).  Either array may be NULL. */
int sol105_pattern(sol105_ctx *ctx, const sol105_model *model, int64_t *n_blocks,
                   int32_t *block_row_ptr, int32_t *block_col);

/* K assembly: element Ke kernels + deterministic block gather.  Replaces
 * Solver.assemble_stiffness (JFEM/src/solver/assembly.jl:1953-4375).
 * vals (n_blocks*36) may be NULL (keeps K on the device). */
int sol105_assemble_K(sol105_ctx *ctx, sol105_model *model, double *vals);

/* No counterpart in the reference.  K assembly as sol105_assemble_K for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the call allocates
 * nothing.  The element kernels stage their 6x6 node blocks in a
 * library-owned static device array of SOL105_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends) instead of the model's own buffer; a model
 * whose staging ((16 n_q4 + 9 n_t3 + 4 n_bar) x 36 doubles) exceeds it is
 * rejected with SOL105_E_INVALID.  K lands in the model's device buffer (as
 * sol105_assemble_K with vals = NULL) and, when d_vals is not NULL, is copied
 * to d_vals, a DEVICE pointer to n_blocks x 36 doubles; the call then
 * returns after the copy.  The values are those of sol105_assemble_K, bit for
 * bit.  All calls on one GPU share its array: calls that overlap on one GPU
 * race. */
#define SOL105_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int sol105_assemble_K_noalloc(sol105_ctx *ctx, sol105_model *model, double *d_vals);

/* Static prestress solve: AUTOSPC, sparse Cholesky (device), solve.
 * Replaces Solver.solve_case / apply_bc_and_solve
 * (JFEM/src/solver/solve_case.jl:845-871, boundary_conditions.jl:530-1012).
 * u (6*n_nodes) may be NULL; free_mask (6*n_nodes) may be NULL. */
int sol105_static_solve(sol105_ctx *ctx, sol105_model *model, double *u, uint8_t *free_mask);

/* SOL 101 result recovery on the device-resident static state (after
 * sol105_static_solve): the reactions r = K u - F (6*n_nodes, analysis
 * frame) and the element results of the current displacement.  Replaces the
 * result build of Solver.solve_case (JFEM/src/solver/solve_case.jl:126-186,
 * 866-868) and recover_shell_stresses! / recover_bar_stresses!
 * (JFEM/src/solver/stress_recovery.jl:163-513).  Element outputs are
 * component-major, column c of element e at out[c * n + e]:
 *   q4_out  (44 x n_q4): N (fx fy fxy), M (mx my mxy), Q (qx qy) at the
 *           centroid, stresses at z1 = -T/2 and z2 = +T/2 (normal_x,
 *           normal_y, shear_xy), strains at z1 and z2, corner N (4 x 3) and
 *           corner M (4 x 3, mxy held at the centroid value), element frame;
 *   t3_out  (20 x n_t3): the first 20 columns of the CQUAD4 layout;
 *   bar_out (18 x n_bar): axial, shear_1, shear_2, torque, moment_a1,
 *           moment_a2, moment_b1, moment_b2, axial stress, end A stresses at
 *           C D E F, end B stresses at C D E F, 1 (0: bar shorter than 1e-9,
 *           which the reference leaves out).
 * bar_points (n_bar x 8: C1 C2 D1 D2 E1 E2 F1 F2 of the PBARL section) may be
 * NULL (all zero).  Any output may be NULL; kernel_ms (4: reactions, CQUAD4,
 * CTRIA3, CBAR device times in ms) may be NULL. */
int sol105_static_recover(sol105_ctx *ctx, sol105_model *model, const double *bar_points, double *residual,
                          double *q4_out, double *t3_out, double *bar_out, double *kernel_ms);

/* KG assembly from the device-resident static displacement (or from u if
 * non-NULL).  Replaces Solver.assemble_geometric_stiffness
 * (JFEM/src/solver/assembly.jl:4382-6635).  vals may be NULL. */
int sol105_assemble_KG(sol105_ctx *ctx, sol105_model *model, const double *u, double *vals);

/* Buckling eigensolve with the reference's output semantics (mode count,
 * positive, EIGRL range, localization and cluster filters, max-abs
 * normalisation).  Replaces Solver.solve_buckling
 * (JFEM/src/solver/solve_case.jl:1863-3205).  lambda_out has room for
 * max_modes values, modes_out (may be NULL) for max_modes x 6*n_nodes
 * doubles (mode-major).  *n_out receives the count.
 * EIGRL range: the reference keeps the whole spectrum of a free set of at
 * most 4000 DOFs, and above that takes the ceil(3 ND) roots of smallest
 * |lambda| plus a shift-invert search near V2 (solve_case.jl:2131-2196,
 * 2422-2560).  Here a free set of at most 64 DOFs gets every pair; up to 4000
 * DOFs the eigensolve returns the smallest positive roots (ceil(3 ND), then
 * 62 when those do not decide the output, SOL105_E_UNSUPPORTED when 62 do
 * not either); above 4000 DOFs the ceil(3 ND) roots of smallest |lambda| and
 * no shift-invert search.  sol105_buckling_info tells whether the computed
 * roots decided the output. */
int sol105_buckling(sol105_ctx *ctx, sol105_model *model, int32_t max_modes,
                    double *lambda_out, double *modes_out, int32_t *n_out);

/* Selection record of the model's last buckling eigensolve:
 * [0] eigenpairs computed, [1] positive candidates in the EIGRL range,
 * [2] modes skipped by the cluster filter, [3] 1 when the computed pairs
 * decide the output as the complete in-range spectrum would (always without
 * a range), 0 when further in-range eigenvalues above the computed ones
 * could change it (the output then is the reference's with
 * JFEM_SOL105_SKIP_RANGE_AUGMENTATION=true on a free set above 4000 DOFs).
 * No counterpart in the reference (its diagnostics dictionary). */
int sol105_buckling_info(const sol105_ctx *ctx, const sol105_model *model, double *out /* 4 */);

/* End-to-end SOL 105 solve (K -> u -> KG -> buckling) on the device.
 * Replaces the SOL 105 part of solve_model/_solve_sol105
 * (JFEM/src/JFEMSolver.jl:789-872, 1296-1568). */
int sol105_solve(sol105_ctx *ctx, sol105_model *model, int32_t max_modes,
                 double *lambda_out, double *modes_out, int32_t *n_out);

/* Per-element thickness sensitivity d(lambda_m)/dt_e of buckling mode m
 * (0-based, after sol105_buckling) for every shell (CQUAD4 first, then
 * CTRIA3), including the prestress (adjoint) term.  Replaces
 * solve_adjoint_buckling with one design variable per element
 * (JFEM/src/solver/adjoint_buckling.jl:2406-2725). */
int sol105_sensitivity_thickness(sol105_ctx *ctx, sol105_model *model, int32_t mode,
                                 double *dlam_dt /* n_q4 + n_t3 */);

/* The terms of d(lambda_m)/dt_e separately, for checking the adjoint formula
 * term by term (JFEM/src/solver/adjoint_buckling.jl:2611):
 *   terms ((n_q4 + n_t3) x 3, may be NULL) = [phi' dKe/dt phi,
 *                 phi' KGe phi / t, psi' dKe/dt u] per shell (CQUAD4 first),
 *   adjoint_load (6*n_nodes, may be NULL) = lambda d(phi' KG(u) phi)/du on
 *                                    the free DOFs (zero elsewhere),
 *   *phi_kg_phi (may be NULL) = phi' KG phi,
 * so that d(lambda)/dt_e = -(terms[0] + lambda terms[1] - terms[2]) / phi_kg_phi
 * with psi = K_ff^-1 adjoint_load. */
int sol105_sensitivity_terms(sol105_ctx *ctx, sol105_model *model, int32_t mode, double *terms,
                             double *adjoint_load, double *phi_kg_phi);

/* Diagnostics, no counterpart in the reference (which has no device code):
 * HIP-event timing of the block triangular-solve product kernel, the
 * dominant kernel of the eigensolve (one launch per tree height and solve
 * direction).  sol105_profile_enable resets the counters and turns timing
 * on (on != 0) or off; sol105_profile_read returns
 *   out[0] launches timed, out[1] their summed duration in ms,
 *   out[2] their algorithmic bytes (8 x (factor entries + nrhs x vector
 *          rows read and written)),
 *   out[3], out[4], out[5]: launches, ms and bytes of the leaf height alone,
 *   out[6], out[7]: ms and bytes of the heights above it. */
int sol105_profile_enable(sol105_ctx *ctx, int32_t on);
int sol105_profile_read(sol105_ctx *ctx, double *out /* 8 */);

/* Diagnostics: the projected (Rayleigh-Ritz) eigensolver of the block
 * Lanczos iteration on a caller's symmetric matrix.  No reference
 * counterpart (the reference's projected problems are KrylovKit's); for
 * tests of the device eigensolver.  A: m x m, column-major, host memory
 * (1 <= m <= 1024); the k (1 <= k <= min(m, 64)) pairs of largest |theta|
 * (algebraic != 0: largest theta) into theta_out (k, descending) and z_out
 * (m x k, column-major).  bw: 0 for a dense matrix, 4..16 to declare a
 * half-bandwidth (the band-reduction path, which falls back to the dense one
 * when the band claim fails).  SOL105_E_NOCONV when the solver reports
 * failure. */
int sol105_symeig_top(sol105_ctx *ctx, int32_t m, const double *a, int32_t k, int32_t bw, int32_t algebraic,
                      double *theta_out, double *z_out);

/* Batched solves.  A batch stacks independent models (design variants,
 * queued requests) into one block-diagonal system: element kernels,
 * assembly, AUTOSPC (per member), the multifrontal factorisation and the
 * triangular solves run once per elimination-tree height for all members,
 * and the block Lanczos iteration runs on every member's own rows
 * (per-member inner products, orthonormalisation and Rayleigh-Ritz) until
 * each has converged.  Per member, results follow sol105_buckling's
 * semantics; a member whose iteration needs the rank-revealing rebuild, or
 * whose stacked factorisation fails, is re-solved alone.  Replaces the loop
 * of independent solves of the reference's batch runner
 * (JFEM/tools/manifest_batch_core.jl:409-478).  Members must share PARAM
 * K6ROT and AUTOSPC. */
typedef struct sol105_batch sol105_batch;
int sol105_batch_create(sol105_ctx *ctx, const sol105_model_desc *members, int32_t n_members, sol105_batch **out);
int sol105_batch_free(sol105_ctx *ctx, sol105_batch *batch);
/* K -> u -> KG -> buckling for every member.  lambda_out: n_members x
 * max_modes (row-major, zero-padded), n_out and status (SOL105_OK or the
 * member's own error code): n_members each; any may be NULL.  Returns
 * SOL105_OK when the batch ran, also if single members failed. */
int sol105_batch_solve(sol105_ctx *ctx, sol105_batch *batch, int32_t max_modes, double *lambda_out, int32_t *n_out,
                       int32_t *status);
/* Mode shapes of one member after sol105_batch_solve: n_out[member] modes of
 * 6*n_nodes[member] values each (mode-major, max |component| = 1). */
int sol105_batch_modes(sol105_ctx *ctx, const sol105_batch *batch, int32_t member, double *modes_out);
/* out[8]: device ms of the last batch solve, Lanczos vectors per member at
 * the end, block steps, members re-solved alone, stacked free DOFs, front
 * storage bytes, factor flops, tree height.  No reference counterpart. */
int sol105_batch_stats(const sol105_ctx *ctx, const sol105_batch *batch, double *out /* 8 */);

/* Request server for design loops and request streams.  The reference
 * serves them with a JSONL worker that runs one manifest case after another
 * (JFEM/tools/jfem_worker_jsonl.jl, JFEM/tools/manifest_batch_core.jl:
 * 409-478).  Here each mesh is registered once (uploaded, pattern and
 * ordering built once) and designs of it are submitted as new shell
 * thicknesses, so analysis and factor buffers stay on the device between
 * design iterations.  n_workers threads, each with its own HIP stream, take
 * requests as soon as a worker and device memory are free, so requests
 * overlap on the device; pending requests of small meshes are stacked into
 * one batched solve with SOL105_SERVER_STACKING.  A request is admitted when its
 * estimated device footprint fits hbm_budget_bytes beside the work in flight
 * (<= 0: three quarters of the free device memory).  Requests of one mesh
 * run in submission order.  With SOL105_SERVER_WARM_START a worker starts a
 * mesh's eigensolve from the Ritz vectors of that mesh's previous design
 * (SOL105_OPT_WARM_START), and a
 * request repeating its mesh's last solved design is answered from that
 * solve.  No reference counterpart for the entry points themselves. */
typedef struct sol105_server sol105_server;
/* flags: SOL105_SERVER_STACKING (stack small pending requests into batched
 * solves) | SOL105_SERVER_WARM_START (start each eigensolve of a mesh from
 * the Ritz vectors of its previous design) */
#define SOL105_SERVER_STACKING 1
#define SOL105_SERVER_WARM_START 2
int sol105_server_create(int device, int32_t n_workers, double hbm_budget_bytes, int32_t flags,
                         sol105_server **out);
int sol105_server_destroy(sol105_server *server);  /* finishes queued requests first */
const char *sol105_server_last_error(const sol105_server *server);
int sol105_server_add_model(sol105_server *server, const sol105_model_desc *desc, int32_t *model_id);
/* shell_thickness (n_q4 + n_t3, CQUAD4 first) may be NULL: the mesh's
 * current design.  max_modes above 64 is taken as 64. */
int sol105_server_submit(sol105_server *server, int32_t model_id, const double *shell_thickness,
                         int32_t max_modes, int64_t *ticket);
/* As sol105_server_submit; flags SOL105_REQ_MODES also keeps the request's
 * mode shapes for sol105_server_wait_modes. */
#define SOL105_REQ_MODES 1
int sol105_server_submit_ex(sol105_server *server, int32_t model_id, const double *shell_thickness,
                            int32_t max_modes, int32_t flags, int64_t *ticket);
/* SOL105_OK with the results (then the ticket is released) or
 * SOL105_E_PENDING.  lambda_out: max_modes values; status: the request's
 * own result code; info (may be NULL): latency ms (submit -> done), queue
 * ms, service ms, requests in its stacked solve, start and finish ms since
 * the server was created, Lanczos basis vectors its eigensolve generated,
 * 1 if it repeated its mesh's last solved design and was answered from that
 * solve (else 0). */
int sol105_server_poll(sol105_server *server, int64_t ticket, double *lambda_out, int32_t *n_out,
                       int32_t *status, double *info /* 8 */);
/* as sol105_server_poll, blocking until the request is done; SOL105_E_INVALID
 * when another caller released the ticket meanwhile */
int sol105_server_wait(sol105_server *server, int64_t ticket, double *lambda_out, int32_t *n_out,
                       int32_t *status, double *info /* 8 */);
/* as sol105_server_wait, also copying the mode shapes of a request submitted
 * with SOL105_REQ_MODES (n_out x 6 n_nodes doubles, mode-major, max |component|
 * = 1) into modes_out (room for max_modes x 6 n_nodes; may be NULL) */
int sol105_server_wait_modes(sol105_server *server, int64_t ticket, double *lambda_out, int32_t *n_out,
                             int32_t *status, double *info /* 8 */, double *modes_out);
/* out[10]: requests done, stacked solves, requests in them, queued, bytes in
 * flight, budget, workers, registered meshes, requests answered from a
 * repeated design, largest bytes in flight so far */
int sol105_server_stats(sol105_server *server, double *out /* 10 */);

/* Launch shapes of the triangular-solve products per tree height (no
 * reference counterpart: the reference's solves are CHOLMOD's).  cfg holds,
 * per height, forward then backward (kmax, ks, pch, kg): fronts whose
 * inner dimension is at most kmax form one k-slice, wider ones are cut in
 * ks-wide slices (kmax and ks at most 384), pch four-k steps of factor loads
 * per chunk (8, 16, 32; 0 = chosen by the launch size), kg groups of four
 * wavefronts per workgroup that split a tile's k-steps and sum their parts
 * in a fixed order (1, 2, 4; 0 = 1; pch 32 only with kg 1).  _get needs a
 * symbolic analysis (a first solve); cfg may be NULL to query n_heights.
 * _set with values <= 0 for kmax or ks keeps the default; the shapes persist
 * over re-analyses of the model. */
int sol105_solve_cfg_get(sol105_ctx *ctx, const sol105_model *model, int32_t *cfg, int32_t *n_heights);
/* Shape of each height (heights x 4: fronts, largest m, largest ns, nonzero
 * factor entries), the key of a persisted table of tuned shapes. */
int sol105_solve_heights(sol105_ctx *ctx, const sol105_model *model, int64_t *shape, int32_t *n_heights);
int sol105_solve_cfg_set(sol105_ctx *ctx, sol105_model *model, const int32_t *cfg, int32_t n_heights);
/* Autotuner: on the model's current factor, times every height of both
 * directions for candidate shapes (median of reps, 3..50) and keeps the
 * fastest (a candidate must win by 3 %, and again in an alternating
 * re-timing against the starting shape); if the whole block solve is not
 * faster with the new shapes, the starting shapes are restored.  report (may
 * be NULL): block solve ms before and after, launches timed, heights. */
int sol105_solve_autotune(sol105_ctx *ctx, sol105_model *model, int32_t reps, double *report);
/* The same for the stacked model of a batch (after a batch solve). */
int sol105_batch_autotune(sol105_ctx *ctx, sol105_batch *batch, int32_t reps, double *report);
/* The stacked model of a batch, for sol105_solve_heights / _cfg_get /
 * _cfg_set only (the batch owns it; other model calls on it are not
 * supported). */
sol105_model *sol105_batch_model(sol105_batch *batch);

/* Phase timings of the last sol105_solve in milliseconds (device events):
 * [0] K assembly, [1] AUTOSPC+symbolic, [2] factorization, [3] static solve,
 * [4] KG assembly, [5] eigensolve, [6] post-processing, [7] total. */
int sol105_last_timings(const sol105_ctx *ctx, double *ms /* 8 */);

/* Statistics of the factorization of the last solve: [0] n_free, [1] nnz(L),
 * [2] factor flops, [3] number of fronts, [4] tree height, [5] Lanczos basis
 * size, [6] Lanczos block steps, [7] bytes of front storage. */
int sol105_last_stats(const sol105_ctx *ctx, double *stats /* 8 */);

#ifdef __cplusplus
}
#endif
#endif /* SOL105_H */
