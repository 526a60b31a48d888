#!/usr/bin/env python3
"""SOL 105 solves/s on one MI355X (stringer-stiffened panel, ~122k DOF).

Workload: 200 x 100 CQUAD4 skin (1000 x 500 mm, t = 1.5 mm) with CBAR
stringers (PBARL BAR 3 x 20 mm) on every 10th longitudinal mesh line, clamped
loaded edges, axial compression, EIGRL ND = 10, FP64.  One step = one full
SOL 105 solve on the device: K assembly -> AUTOSPC + Cholesky -> static
solve -> KG assembly -> block-Lanczos buckling eigensolve (10 modes, eigen-
vectors, reference filters).  The model arrays are resident in HBM before
the timed region (deck parse and upload excluded, reported separately).

The line also carries the batched-vs-plain comparison of the metric
("batched"): a sweep of 256 seeded plate variants solved one model at a time
and as one stacked batch; the served request stream ("served": p50/p99
latency of a Poisson stream of design requests, served vs a plain queue);
and the large cylinder case with its sensitivities ("cylinder").

Prints ONE JSON line (rank 0).  --gpus N with N > 1 exits: this build runs
on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SOL105 solves/sec + p50/p99 latency at 1e-8 eigen rel-err, batched vs plain GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--ny", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-only", action="store_true", help="only time the CPU oracle sample")
    ap.add_argument("--batch-variants", type=int, default=256,
                    help="seeded plate variants of the batched-vs-plain comparison (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--batch-cpu-sample", type=int, default=32,
                    help="variants of the sweep also solved by the CPU oracle on all host cores (0: none)")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-vs-plain sweep (for profiling)")
    ap.add_argument("--no-batch-plain", action="store_true", help="skip the one-at-a-time leg of the sweep")
    ap.add_argument("--cylinder-fd", type=int, default=8,
                    help="seeded shells of the cylinder's finite-difference check (0: no cylinder leg)")
    ap.add_argument("--cylinder-only", action="store_true", help="only the cylinder leg")
    ap.add_argument("--served-requests", type=int, default=120,
                    help="requests of the served-stream leg (0: no served leg)")
    ap.add_argument("--served-only", action="store_true", help="only the served-stream leg")
    ap.add_argument("--no-static", action="store_true", help="skip the SOL 101 result-recovery leg")
    ap.add_argument("--no-autotune", action="store_true", help="time the default solve-product shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from jf_2026_05_19_amd import bdf, decks, model as M

    if args.served_only:
        print(json.dumps({"metric": METRIC, "served": served_stream(max(args.served_requests, 1))}))
        return
    if args.cylinder_only:
        from jf_2026_05_19_amd import _native as N
        print(json.dumps({"metric": METRIC, "cylinder": cylinder_case(N.Context(0), max(args.cylinder_fd, 1))}))
        return
    if args.batch_only:
        from jf_2026_05_19_amd import _native as N
        print(json.dumps({"metric": METRIC, "batched": batched_sweep(N.Context(0), args.batch_variants,
                                                                     args.batch_steps,
                                                                     plain=not args.no_batch_plain,
                                                                     cpu_sample=0)}))
        return
    from jf_2026_05_19_amd import _native as N
    txt = decks.panel_deck(nx=args.nx, ny=args.ny)
    # deck text to model arrays: the native reader (the product path) and,
    # for comparison, the host builder it is tested against
    t0 = time.perf_counter()
    arr = N.read_deck(txt)
    t_parse = time.perf_counter() - t0
    t0 = time.perf_counter()
    M.build_arrays(bdf.bdf_to_model_text(txt))
    t_parse_host = time.perf_counter() - t0
    log("model: nodes=%d q4=%d bars=%d ndof=%d (deck read %.3fs native, %.3fs host builder)" %
        (arr.n_nodes, arr.n_q4, arr.n_bar, arr.ndof, t_parse, t_parse_host))
    workload = ("stiffened panel %dx%d CQUAD4 + %d CBAR stringers, %d DOF, ND=%d, FP64" %
                (args.nx, args.ny, arr.n_bar, arr.ndof, arr.nd))
    result = {"metric": METRIC, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
              "data": "synthetic (seeded deck generator)",
              "config": {"workload": workload, "nx": args.nx, "ny": args.ny, "ndof": arr.ndof, "nd": arr.nd,
                         "parallelism": "single GPU"}}

    cpu, cpu_ref = None, None
    if not args.no_cpu_baseline:
        cpu, cpu_ref = cpu_baseline(arr)
        result["cpu_baseline"] = cpu
    if args.cpu_only:
        print(json.dumps(result))
        return

    from jf_2026_05_19_amd import _native as N
    ctx = N.Context(0)
    t0 = time.perf_counter()
    dm = N.DeviceModel(ctx, arr)
    t_upload = time.perf_counter() - t0
    log("upload+pattern %.3fs" % t_upload)
    lam = None
    for w in range(args.warmup):
        t1 = time.perf_counter()
        lam, _ = dm.solve(want_modes=False)
        log("warmup %d: %.2f ms lambda1=%.6f" % (w, 1e3 * (time.perf_counter() - t1), lam[0] if len(lam) else -1))
    # default product shapes, then the autotuner (setup, untimed like the
    # symbolic analysis): the timed steps below run the tuned shapes
    t1 = time.perf_counter()
    for k in range(max(args.steps // 2, 1)):
        dm.solve(want_modes=False)
    t_default = (time.perf_counter() - t1) / max(args.steps // 2, 1)
    t1 = time.perf_counter()
    tune = dm.autotune(7) if not args.no_autotune else None
    if tune:
        log("autotune: block solve %.3f -> %.3f ms (%d launches timed, %.2fs)" % (tune[0], tune[1], tune[2],
                                                                             time.perf_counter() - t1))
    # untimed: the first solve under the tuned shapes launches directly, the
    # second captures the Lanczos steps as HIP graphs; the timed ones replay
    dm.solve(want_modes=False)
    dm.solve(want_modes=False)
    lat = []
    phase = np.zeros(8)
    t_start = time.perf_counter()
    for k in range(args.steps):
        t1 = time.perf_counter()
        lam, _ = dm.solve(want_modes=False)
        lat.append(time.perf_counter() - t1)
        phase += ctx.timings()
    t_total = time.perf_counter() - t_start
    # the same solves with every kernel launched directly (no step graphs)
    ctx.set_option(N.Context.OPT_STEP_GRAPHS, 0)
    t1 = time.perf_counter()
    for k in range(args.steps):
        dm.solve(want_modes=False)
    t_direct = (time.perf_counter() - t1) / max(args.steps, 1)
    ctx.set_option(N.Context.OPT_STEP_GRAPHS, 1)
    log("steps: %.2f ms/solve with step graphs, %.2f ms with direct launches" % (1e3 * t_total / args.steps,
                                                                                 1e3 * t_direct))
    # the roofline pass: the same solves again with HIP events around every
    # launch of the dominant kernel (outside the timed region: the events
    # themselves cost time between launches)
    ctx.profile_enable(True)
    for k in range(args.steps):
        dm.solve(want_modes=False)
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    phase /= max(args.steps, 1)
    stats = ctx.stats()
    lat_ms = np.array(lat) * 1e3
    result.update({
        "value": args.steps / t_total,
        "ms_per_step": 1e3 * t_total / args.steps,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "lambda": [float(x) for x in lam],
        "phase_ms": {"K_assembly": phase[0], "autospc_symbolic_factor": phase[2], "static_solve": phase[3],
                     "KG_assembly": phase[4], "eigensolve": phase[5], "total": phase[7]},
        "factor_stats": {"n_free": stats[0], "nnz_L": stats[1], "factor_flops": stats[2], "fronts": stats[3],
                         "tree_height": stats[4], "lanczos_vectors": stats[5], "lanczos_block_steps": stats[6]},
        "parse_s": t_parse, "parse_s_host_builder": t_parse_host, "upload_s": t_upload,
        "ms_per_step_default_shapes": 1e3 * t_default,
        "ms_per_step_direct_launches": 1e3 * t_direct,
        "autotune": None if not tune else {"block_solve_ms_before": tune[0], "block_solve_ms_after": tune[1],
                                           "launches_timed": tune[2], "heights": tune[3]},
    })
    result["roofline"] = roofline(prof)
    result["factor_rate"] = factor_rate(ctx, stats)
    if cpu_ref is not None:
        result["accuracy_vs_cpu"] = accuracy_vs_cpu(dm, cpu_ref)
    if not args.no_static:
        result["static_recovery"] = static_recovery(dm, arr, cpu_sample=0 if args.no_cpu_baseline else 300)
    dm.free()
    if args.batch_variants > 0:
        result["batched"] = batched_sweep(ctx, args.batch_variants, args.batch_steps,
                                          cpu_sample=0 if args.no_cpu_baseline else args.batch_cpu_sample)
    if args.cylinder_fd > 0:
        result["cylinder"] = cylinder_case(ctx, args.cylinder_fd)
    if args.served_requests > 0:
        result["served"] = served_stream(args.served_requests)
    if cpu is not None:
        result["gpu_over_cpu"] = result["value"] / cpu["value"]
    print(json.dumps(result))


# HBM bytes per launch of the dominant kernel, averaged over all its launches
# in a solve, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by
# the same counter's reading for a 256 MiB device copy (factors 2.0 and 1.0);
# the runs and the per-height figures are in
# profiles/prod_mfma_kernel_c2_panel_pmc.txt.
PROD_TRAFFIC_BYTES = None


def _rate(launches, ms, nbytes):
    if launches == 0 or ms <= 0:
        return None
    avg_s = ms * 1e-3 / launches
    achieved = nbytes / launches / avg_s / 1e9
    return {"launches": launches, "avg_launch_us": avg_s * 1e6, "algorithmic_bytes_per_launch": nbytes / launches,
            "achieved": achieved, "frac": achieved / 8000.0}


def roofline(prof):
    """Roofline of the dominant kernel, the block triangular-solve product
    (prod_mfma_kernel), over ALL of its launches in the timed region (every
    tree height, both solve directions, every block Lanczos step):
    achieved = sum of algorithmic bytes (factor entries streamed once + the
    right-hand-side rows read and written) / sum of launch durations, both
    from HIP events recorded on the solve stream around each launch.  The
    leaf height (the two largest grids) and the heights above it are also
    given separately."""
    la, ms, nb = prof["all"]
    r = _rate(la, ms, nb)
    if r is None:
        return None
    peak = 8000.0  # MI355X HBM3E GB/s
    return {"kernel": "prod_mfma_kernel (block triangular solve: FP64 MFMA product with the explicit factor inverse)",
            "bound": "hbm", "achieved": r["achieved"], "peak": peak, "unit": "GB/s", "frac": r["achieved"] / peak,
            "traffic": PROD_TRAFFIC_BYTES, "algorithmic_bytes_per_launch": r["algorithmic_bytes_per_launch"],
            "avg_launch_us": r["avg_launch_us"], "launches": la, "total_ms": ms,
            "leaf_height": _rate(*prof["leaf"]), "upper_heights": _rate(*prof["upper"])}


def recovery_bytes(arr, n_blocks):
    """Algorithmic HBM bytes of the result-recovery launches: every input
    read once and every output written once (node coordinates and
    displacements once per node).  Per CQUAD4 4 node indices + 4 property
    words + 44 output words; per CTRIA3 3 + 4 + 20 + the macro-blend weight;
    per CBAR 2 indices + 3 + 8 + 8 words in and 18 out; per node 3 + 6 words;
    reactions: the K blocks (36 words + a column index each), u, F and r."""
    nodes = 72 * arr.n_nodes
    return {"q4": (16 * arr.n_q4 + 8 * (4 + 44) * arr.n_q4 + nodes) if arr.n_q4 else 0,
            "t3": (12 * arr.n_t3 + 8 * (4 + 20 + 1) * arr.n_t3 + nodes) if arr.n_t3 else 0,
            "bar": (8 * arr.n_bar + 8 * (3 + 8 + 8 + 18) * arr.n_bar + nodes) if arr.n_bar else 0,
            "residual": n_blocks * (36 * 8 + 4) + 4 * (arr.n_nodes + 1) + 3 * 8 * arr.ndof}


def static_recovery(dm, arr, reps=5, cpu_sample=300):
    """SOL 101 leg on the solved static state: the device's result recovery
    (reactions, every CQUAD4/CTRIA3/CBAR row; sol105_static_recover) timed by
    its own HIP events per kernel, against the algorithmic bytes; and the CPU
    restatement's per-element rate on a sample of elements (the oracle, one
    core) for scale."""
    dm.static_recover()
    ms = np.zeros(4)
    wall = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = dm.static_recover(timings=True)
        wall.append(time.perf_counter() - t0)
        ms += out[4]
    ms /= reps
    nb = recovery_bytes(arr, dm.n_blocks)
    kern = {}
    for i, k in enumerate(("residual", "q4", "t3", "bar")):
        if ms[i] > 0 and nb[k] > 0:
            kern[k] = {"us": 1e3 * ms[i], "algorithmic_bytes": nb[k],
                       "achieved_GBps": nb[k] / (ms[i] * 1e-3) / 1e9,
                       "frac_hbm": nb[k] / (ms[i] * 1e-3) / 1e9 / 8000.0}
    res = {"kernels": kern, "wall_ms_median": 1e3 * float(np.median(wall)),
           "elements": {"q4": arr.n_q4, "t3": arr.n_t3, "bar": arr.n_bar}}
    if cpu_sample > 0 and arr.n_q4 > 0:
        from oracle import oracle as O
        from oracle import recovery as RC
        u = dm.static_solve()[0]
        xyz = np.asarray(arr.xyz)
        n = min(cpu_sample, arr.n_q4)
        t0 = time.perf_counter()
        for e in range(n):
            nodes = [int(v) for v in arr.q4[e]]
            R, lc = O.q4_frame(xyz[nodes].ravel())
            E, nu, h, br = (float(v) for v in arr.q4prop[e][:4])
            RC.q4_recover(lc, RC._local_u(R, u, nodes), E, nu, h, br)
        dt = time.perf_counter() - t0
        res["cpu_oracle_q4_per_s"] = n / dt
        if "q4" in kern:
            res["q4_gpu_over_cpu_oracle"] = arr.n_q4 / (kern["q4"]["us"] * 1e-6) / (n / dt)
    log("static recovery: %s, wall %.2f ms" % (", ".join("%s %.1f us" % (k, v["us"]) for k, v in kern.items()),
                                               res["wall_ms_median"]))
    return res


def accuracy_vs_cpu(dm, ref):
    """Eigenvalues and mode shapes of the device solve against the CPU
    oracle's solve of the same deck (the cpu_baseline leg): largest relative
    eigenvalue error and smallest MAC over modes of non-degenerate pairs."""
    lam, modes = dm.solve()
    lr, mr = ref["eigenvalues"], ref["modes"]
    out = {"n_modes": [len(lam), len(lr)]}
    if len(lam) != len(lr) or len(lam) == 0:
        out["max_eig_rel_err_vs_cpu"] = float("inf")
        return out
    out["max_eig_rel_err_vs_cpu"] = float(np.max(np.abs(lam - lr) / np.abs(lr)))
    macs = []
    for k in range(len(lam)):
        gap = min([abs(lr[k] - lr[j]) / abs(lr[k]) for j in range(len(lr)) if j != k] + [1.0])
        if gap >= 1e-6:
            a, b = modes[:, k], mr[:, k]
            macs.append(float((a @ b) ** 2 / ((a @ a) * (b @ b))))
    out["min_mac_vs_cpu"] = min(macs) if macs else None
    log("accuracy vs CPU oracle: max eig rel err %.2e, min MAC %s" % (out["max_eig_rel_err_vs_cpu"],
                                                                      out["min_mac_vs_cpu"]))
    return out


def factor_rate(ctx, stats):
    """Phase-level FP64 rate of the numeric factorization (informational)."""
    t = ctx.timings()
    fact_ms = t[2]
    flops = stats[2]
    return {"tflops": flops / (fact_ms * 1e-3) / 1e12 if fact_ms > 0 else 0.0, "factor_flops": flops,
            "phase_ms": fact_ms}


def _parse_variant(txt):
    from jf_2026_05_19_amd import _native as N
    return N.read_deck(txt)


def batched_sweep(ctx, n, steps, plain=True, cpu_sample=32):
    """Design sweep of n seeded plate variants (mesh 10x10..80x80, t in
    U(1, 4) mm, E in {70e3, 210e3} MPa, ND = 5): every variant solved one
    after another (plain: one model per solve) and all of them as one batch
    (sol105_batch_solve).  Both keep the models resident and re-use their
    symbolic analysis; the first solve of each is an untimed warmup."""
    from concurrent.futures import ProcessPoolExecutor
    from jf_2026_05_19_amd import decks
    from jf_2026_05_19_amd import _native as N
    t0 = time.perf_counter()
    texts = [t for t, _ in decks.variant_sweep(n=n)]
    with ProcessPoolExecutor(max_workers=8) as ex:
        arrs = list(ex.map(_parse_variant, texts, chunksize=8))
    ndofs = [a.ndof for a in arrs]
    log("batched sweep: %d variants, %d..%d DOF, %d total (parse %.1fs)" % (n, min(ndofs), max(ndofs), sum(ndofs),
                                                                            time.perf_counter() - t0))
    # plain: one solve per variant
    lat, lam_plain, t_plain = [float("nan")], None, float("nan")
    if plain:
        models, lam_plain = [], []
        for i, a in enumerate(arrs):
            models.append(N.DeviceModel(ctx, a))
            lam_plain.append(models[-1].solve(16, want_modes=False)[0])   # warmup (symbolic analysis)
            if (i + 1) % 64 == 0:
                log("plain warmup: %d/%d models" % (i + 1, n))
        lat = []
        t1 = time.perf_counter()
        for m in models:
            ts = time.perf_counter()
            m.solve(16, want_modes=False)
            lat.append(time.perf_counter() - ts)
        t_plain = time.perf_counter() - t1
        for m in models:
            m.free()
        del models
        log("plain: %d solves in %.3fs" % (n, t_plain))
    # batched: all variants in one stacked solve
    B = N.Batch(ctx, arrs)
    t2 = time.perf_counter()
    lam_b, st = B.solve(16)
    log("batch warmup (incl. symbolic analysis): %.2fs" % (time.perf_counter() - t2))
    bt = []
    for k in range(steps):
        ts = time.perf_counter()
        lam_b, st = B.solve(16)
        bt.append(time.perf_counter() - ts)
    t_batch_default = float(np.mean(bt))
    # the solve-product autotuner on the stacked model, then the same steps
    tune = B.autotune(5)
    log("batch autotune: block solve %.3f -> %.3f ms (%d launches timed)" % tune[:3])
    lam_b, st = B.solve(16)   # direct launches under the tuned shapes
    lam_b, st = B.solve(16)   # captures the step graphs
    bt = []
    for k in range(steps):
        ts = time.perf_counter()
        lam_b, st = B.solve(16)
        bt.append(time.perf_counter() - ts)
    t_batch = float(np.mean(bt))
    ctx.set_option(N.Context.OPT_STEP_GRAPHS, 0)
    bt = []
    for k in range(steps):
        ts = time.perf_counter()
        B.solve(16)
        bt.append(time.perf_counter() - ts)
    t_batch_direct = float(np.mean(bt))
    ctx.set_option(N.Context.OPT_STEP_GRAPHS, 1)
    stats = B.stats()
    B.free()
    err = 0.0 if plain else float("nan")
    for la, lb in zip(lam_plain or [], lam_b):
        if len(la) != len(lb):
            err = float("inf")
            break
        if len(la):
            err = max(err, float(np.max(np.abs(lb - la) / np.abs(la))))
    lat_ms = np.array(lat) * 1e3
    log("batched: %d solves in %.3fs (%.1fx plain), max rel err vs plain %.2e" % (n, t_batch, t_plain / t_batch, err))
    cpu = batched_cpu_sample(arrs, lam_b, cpu_sample) if cpu_sample > 0 else None
    return {"workload": "%d seeded plate variants 500x500 mm, mesh 10x10..80x80 (%d..%d DOF, %d total), ND=5, "
                        "t U(1,4) mm, E in {70e3, 210e3} MPa, FP64" % (n, min(ndofs), max(ndofs), sum(ndofs)),
            "batched_solves_per_s": n / t_batch, "plain_solves_per_s": n / t_plain,
            "batched_solves_per_s_default_shapes": n / t_batch_default,
            "batched_solves_per_s_direct_launches": n / t_batch_direct,
            "autotune": {"block_solve_ms_before": tune[0], "block_solve_ms_after": tune[1],
                         "launches_timed": tune[2], "heights": tune[3]},
            "speedup_vs_plain": t_plain / t_batch, "batch_ms": 1e3 * t_batch, "batch_steps": steps,
            "plain_p50_ms": float(np.percentile(lat_ms, 50)), "plain_p99_ms": float(np.percentile(lat_ms, 99)),
            "batched_latency_ms": 1e3 * t_batch,
            "max_eig_rel_err_vs_plain": err, "members_ok": int(np.sum(st == 0)),
            "members_resolved_alone": int(stats["fallbacks"]), "stacked_free_dofs": int(stats["n_free"]),
            "lanczos_vectors": int(stats["lanczos_vectors"]), "cpu": cpu,
            "batched_over_cpu": (n / t_batch) / cpu["solves_per_s"] if cpu else None,
            "mean_dof_sweep": float(np.mean(ndofs))}


def batched_cpu_sample(arrs, lam_b, k):
    """The CPU oracle on the first k variants of the sweep, spread over all
    host cores (one process, one BLAS thread each): its throughput, and the
    batched eigenvalues of those variants against it."""
    from oracle import oracle as O
    O.lib()
    cores = host_cores()
    k = min(k, len(arrs))
    t0 = time.perf_counter()
    ref = O.eigenvalues_many(arrs[:k], workers=cores)
    wall = time.perf_counter() - t0
    err = 0.0
    for la, lr in zip(lam_b[:k], ref):
        if len(la) != len(lr):
            err = float("inf")
            break
        if len(lr):
            err = max(err, float(np.max(np.abs(la - lr) / np.abs(lr))))
    log("batched CPU sample: %d variants on %d cores in %.1fs, max eig rel err batched vs CPU %.2e" %
        (k, cores, wall, err))
    return {"solves_per_s": k / wall, "variants": k, "cores": cores, "wall_s": wall, "kind": "port",
            "mean_dof_sample": float(np.mean([a.ndof for a in arrs[:k]])),
            "max_eig_rel_err_batched_vs_cpu": err,
            "sample": "the first %d variants of the sweep, one oracle solve per process on %d cores "
                      "(one BLAS thread each)" % (k, cores)}


def _replay(arrs, trace, workers, batching, warm=False):
    """Serves a request trace [(arrival s, mesh, thickness)]: submits each
    request at its arrival time, returns per-request (eigenvalues, info) and
    the wall time from the first arrival to the last result."""
    from jf_2026_05_19_amd import _native as N
    srv = N.Server(workers=workers, batching=batching, warm_start=warm)
    ids = [srv.add_model(a) for a in arrs]
    for i in ids:                      # first design of every mesh: symbolic analysis, untimed
        srv.wait(srv.submit(i))
    t0 = time.perf_counter()
    tickets = []
    for arrival, f, t in trace:
        dt = arrival - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        tickets.append(srv.submit(ids[f], t))
    out = [srv.wait(tk) for tk in tickets]
    wall = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    return out, wall, stats


def _trace(arrs, n_req, rate, rng):
    """Poisson arrivals at `rate`, each a +-5 % per-shell perturbation of the
    previous design of a uniformly drawn mesh."""
    cur = [a.shell_thickness() for a in arrs]
    trace, t = [], 0.0
    for r in range(n_req):
        t += float(rng.exponential(1.0 / rate))
        f = int(rng.integers(len(arrs)))
        cur[f] = cur[f] * rng.uniform(0.95, 1.05, size=cur[f].shape)
        trace.append((t, f, cur[f].copy()))
    return trace


def _stream_summary(res, wall, stats, n_req):
    lat = np.array([i["latency_ms"] for _, _, i in res])
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "requests_per_s": n_req / wall, "stacked_solves": int(stats["stacked_solves"]),
            "stacked_requests": int(stats["stacked_requests"]), "cached": int(stats["cached"]),
            "peak_inflight_bytes": stats["peak_inflight_bytes"],
            "lanczos_vectors_mean": float(np.mean([i["lanczos_vectors"] for _, _, i in res]))}


def served_stream(n_req, workers=4, seed=20260520, oracle_max_dof=30000):
    """Served optimisation stream: ten seeded panel meshes (2k..500k DOF
    log-uniform, 10 % CTRIA3, CBAR stringers, ND = 5); each request a +-5 %
    per-shell thickness perturbation of the previous design of its mesh.
    Served (4 worker streams, stacking of small pending requests,
    HBM-footprint admission) and, as the reference, the same trace through
    one worker without stacking (a plain queue).  Two legs: Poisson arrivals at 80 % of the
    one-at-a-time device capacity (latency), and at twice that capacity
    (saturation: throughput).  Latency = arrival -> result.  The served
    results of the meshes up to `oracle_max_dof` DOF are checked against the
    CPU oracle."""
    from concurrent.futures import ProcessPoolExecutor
    from jf_2026_05_19_amd import decks
    from jf_2026_05_19_amd import _native as N
    texts = [t for t, _ in decks.served_families(seed=seed)]
    with ProcessPoolExecutor(max_workers=8) as ex:
        arrs = list(ex.map(_parse_variant, texts))
    ndofs = [a.ndof for a in arrs]
    # one-at-a-time service time of each mesh (warm)
    ctx = N.Context(0)
    svc = []
    for a in arrs:
        dm = N.DeviceModel(ctx, a)
        dm.solve(16, want_modes=False)
        ts = time.perf_counter()
        dm.solve(16, want_modes=False)
        svc.append(time.perf_counter() - ts)
        dm.free()
    ctx.close()
    capacity = 1.0 / float(np.mean(svc))
    rng = np.random.default_rng(seed)
    out = {"workload": "served stream: %d requests per leg over 10 panel meshes %d..%d DOF (10%% CTRIA3, CBAR), "
                       "+-5%% thickness steps per mesh, ND=5, FP64; Poisson arrivals at 0.8x and 2x the "
                       "one-at-a-time capacity (%.1f req/s)" % (n_req, min(ndofs), max(ndofs), capacity),
           "workers": workers, "mean_solo_service_ms": 1e3 * float(np.mean(svc))}
    for leg, load in (("load_0.8", 0.8), ("saturated_2x", 2.0)):
        trace = _trace(arrs, n_req, load * capacity, rng)
        res_s, wall_s, st_s = _replay(arrs, trace, workers, True)
        res_q, wall_q, st_q = _replay(arrs, trace, 1, False)
        err = 0.0
        for (la, sa, _), (lb, sb, _) in zip(res_s, res_q):
            if sa != 0 or sb != 0 or len(la) != len(lb):
                err = float("inf")
                break
            if len(la):
                err = max(err, float(np.max(np.abs(la - lb) / np.abs(lb))))
        S, Q = _stream_summary(res_s, wall_s, st_s, n_req), _stream_summary(res_q, wall_q, st_q, n_req)
        log("served %s: p50 %.1f ms p99 %.1f ms %.1f req/s (stacked %d/%d, cached %d, %.0f vectors/solve); "
            "plain queue: p50 %.1f ms p99 %.1f ms %.1f req/s (%.0f vectors/solve)" %
            (leg, S["p50_ms"], S["p99_ms"], S["requests_per_s"], S["stacked_requests"], S["stacked_solves"],
             S["cached"], S["lanczos_vectors_mean"], Q["p50_ms"], Q["p99_ms"], Q["requests_per_s"],
             Q["lanczos_vectors_mean"]))
        out[leg] = {"arrival_rate_per_s": load * capacity, "served": S, "plain_queue": Q,
                    "throughput_served_over_plain": S["requests_per_s"] / Q["requests_per_s"],
                    "max_eig_rel_err_served_vs_queue": err}
        if leg == "load_0.8":
            out[leg]["vs_cpu_oracle"] = _served_vs_oracle(arrs, trace, res_s, oracle_max_dof)
    return out


def _served_vs_oracle(arrs, trace, res, max_dof):
    """The served eigenvalues of every request on a mesh of at most max_dof
    DOF against the CPU oracle's solve of the same design (on all host cores)."""
    from oracle import oracle as O
    idx = [k for k, (_, f, _) in enumerate(trace) if arrs[f].ndof <= max_dof]
    if not idx:
        return None
    ref = O.eigenvalues_many([arrs[trace[k][1]].with_shell_thickness(trace[k][2]) for k in idx], workers=host_cores())
    err = 0.0
    for k, lr in zip(idx, ref):
        la = res[k][0]
        if len(la) != len(lr):
            return {"requests": len(idx), "max_eig_rel_err": float("inf")}
        if len(lr):
            err = max(err, float(np.max(np.abs(la - lr) / np.abs(lr))))
    log("served vs CPU oracle: %d requests on meshes <= %d DOF, max eig rel err %.2e" % (len(idx), max_dof, err))
    return {"requests": len(idx), "max_dof": max_dof, "max_eig_rel_err": err}


def cylinder_case(ctx, n_fd, steps=3, seed=20260521):
    """Large single case: cylinder R = 500, L = 1000, t = 1 mm, 416 x 400
    GRIDs (998,400 DOF) with a 10 %-by-area CTRIA3 band, clamped at z = 0,
    axial ring load at z = L, ND = 10.  Timed: warm solves (symbolic analysis
    in the untimed first solve), then d(lambda)/dt of every shell for all
    modes.  Checked: mode-1 sensitivities of n_fd seeded shells against
    central differences of the device's own forward path (delta = 1e-4 t,
    mode tracked by MAC; lambda is resolved to ~1e-13 relative, so smaller
    steps drown the difference in that noise)."""
    from jf_2026_05_19_amd import bdf, decks, model as M
    from jf_2026_05_19_amd import _native as N
    t0 = time.perf_counter()
    txt = decks.cylinder_deck()
    t0 = time.perf_counter()
    arr = N.read_deck(txt)
    log("cylinder: nodes=%d q4=%d t3=%d ndof=%d (deck read %.2fs)" % (arr.n_nodes, arr.n_q4, arr.n_t3, arr.ndof,
                                                                      time.perf_counter() - t0))
    dm = N.DeviceModel(ctx, arr)
    t1 = time.perf_counter()
    lam, modes = dm.solve(16)
    log("cylinder first solve (incl. symbolic analysis) %.2fs, lambda1 %.6f" % (time.perf_counter() - t1, lam[0]))
    tune = dm.autotune(5)
    log("cylinder autotune: block solve %.3f -> %.3f ms" % tune[:2])
    lat = []
    for k in range(steps):
        ts = time.perf_counter()
        lam, modes = dm.solve(16, want_modes=(k == steps - 1))
        lat.append(time.perf_counter() - ts)
    tim, st = ctx.timings(), ctx.stats()
    rec = static_recovery(dm, arr, cpu_sample=0)
    ts = time.perf_counter()
    sens = [dm.sensitivity_thickness(k) for k in range(len(lam))]
    t_sens = time.perf_counter() - ts
    log("cylinder: solve %.3fs, sensitivities of %d shells x %d modes %.3fs" % (np.median(lat), arr.n_shell,
                                                                               len(lam), t_sens))
    # finite-difference check of mode 1 on seeded shells
    rng = np.random.default_rng(seed)
    elems = np.sort(rng.choice(arr.n_shell, size=n_fd, replace=False))
    t_sh = arr.shell_thickness()
    phi0 = modes[:, 0]
    errs = []
    for e in elems:
        vals = []
        for sgn in (1.0, -1.0):
            t = t_sh.copy()
            t[e] *= 1.0 + sgn * 1e-4
            dm.set_shell_thickness(t)
            l2, m2 = dm.solve(16)
            macs = [float((phi0 @ m2[:, j]) ** 2 / ((phi0 @ phi0) * (m2[:, j] @ m2[:, j]))) for j in range(len(l2))]
            vals.append(l2[int(np.argmax(macs))])
        fd = (vals[0] - vals[1]) / (2e-4 * t_sh[e])
        errs.append(abs(sens[0][e] - fd))
        log("cylinder FD shell %d: adjoint %.6e fd %.6e" % (e, sens[0][e], fd))
    dm.set_shell_thickness(t_sh)
    scale = float(np.abs(sens[0]).max())
    dm.free()
    return {"workload": "cylinder R=500 L=1000 t=1 mm, 416x400 GRIDs, %d CQUAD4 + %d CTRIA3 (10%% band), %d DOF, "
                        "ND=10, FP64" % (arr.n_q4, arr.n_t3, arr.ndof),
            "solve_ms": 1e3 * float(np.median(lat)), "phase_ms": {"K_assembly": tim[0], "factor": tim[2],
                                                               "static_solve": tim[3], "KG_assembly": tim[4],
                                                               "eigensolve": tim[5]},
            "lambda": [float(x) for x in lam], "lanczos_vectors": int(st[5]), "factor_flops": float(st[2]),
            "sensitivity_ms_all_shells_all_modes": 1e3 * t_sens, "shells": arr.n_shell, "modes": len(lam),
            "static_recovery": rec,
            "autotune": {"block_solve_ms_before": tune[0], "block_solve_ms_after": tune[1],
                         "launches_timed": tune[2], "heights": tune[3]},
            "fd_check": {"shells": [int(e) for e in elems], "max_abs_err": float(max(errs)),
                         "max_err_rel_to_max_sensitivity": float(max(errs)) / scale, "step": "1e-4 t, central",
                         "mode": 1}}


def host_cores():
    """Host cores this process may use, capped at 16 (the box's CPU share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(arr):
    """The CPU oracle's throughput on the host: one full SOL 105 solve of
    the same deck per core, all concurrently (one process and one BLAS
    thread each: C element kernels, SciPy SuperLU static solve, ARPACK
    shift-invert eigensolve), solves/s = cores / wall time.  Also returns the
    first process's eigenpairs for the accuracy check."""
    from oracle import oracle as O
    O.lib()
    cores = host_cores()
    wall, outs = O.concurrent_solves(arr, cores)
    single = [o["times"]["total"] for o in outs]
    t0 = outs[0]["times"]
    log("cpu oracle: %d concurrent solves in %.1fs (single solve %.1f..%.1f s), lambda1=%.6f" %
        (cores, wall, min(single), max(single), outs[0]["eigenvalues"][0]))
    out = {"value": cores / wall, "unit": "solves/s", "cores": cores, "kind": "port",
           "sample": "%d concurrent full solves of the same deck, one per core (one process and one BLAS thread each: "
                     "C element kernels, SciPy SuperLU static solve, ARPACK shift-invert eigsh), %.1f s wall"
                     % (cores, wall),
           "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "single_solve_s_median": float(np.median(single)),
           "phase_s": {k: float(v) for k, v in t0.items() if k != "superlu_fill_nnz"},
           "superlu_factor_s": float(t0.get("superlu_factor", float("nan"))),
           "superlu_fill_nnz": int(t0.get("superlu_fill_nnz", 0))}
    return out, outs[0]


if __name__ == "__main__":
    main()
