"""Benchmark: tet-iterations/s of forward + adjoint, batched mixed precision
vs the plain FP64 GPU path, on one MI355X.

Default workload (the batched-rollout configuration): 40x12x12 box
(34,560 tets, 6,929 vertices), 64 straight fibres, ground plane with friction,
R = 256 rollouts sharing one Cholesky factor (run as two concurrent
sub-batches of 128 on their own streams), seeded sinusoidal activations,
100 steps x 10 iterations, forward + adjoint with the whole rollout taped by
the forward pass (about 232 GB of the 309 GB of HBM, so the adjoint recomputes
nothing; --ckpt k keeps checkpoints every k steps and recomputes segments).

Precision (--precision): mixed_adjoint (default) runs the forward in FP64, so
the trajectory, its iteration counts and contact branches are the FP64 path's,
and the reverse sweep's local-step VJP in FP32; mixed runs the local step in
FP32 in both directions. Factor, solves, contact and accumulation are FP64 in
both. The line reports the errors of both forms against an FP64 pass of the
same batch (every rollout, full horizon) and the all-FP32 form's throughput.

One bench "step" = one full forward + adjoint pass over the batch (inputs
already resident in HBM). value = sum over rollouts of T * executed PD
iterations * steps / (t_forward + t_adjoint).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(args):
    from diff_fem_amd import scene
    if args.workload == "c5":
        return scene.config5(seed=5)
    d = scene.config4(seed=4)
    d.contact = bool(args.contact)
    if not d.contact:
        d.planes = np.zeros((0, 6))
    return d


def host_cpu_info():
    """nproc, the CPU model (lscpu) and the threads this process may use."""
    import platform
    import subprocess
    model = platform.processor() or ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=20).stdout.strip())
    except (OSError, subprocess.SubprocessError, ValueError):
        nproc = os.cpu_count() or 1
    return dict(nproc=nproc, model=model)


def cpu_baseline(d, x_s0, v_s0, acts, s0, steps, threads):
    """Oracle (C++ FP64 restatement) timed on a bounded sample of the workload:
    steps [s0, s0 + steps) of the rollouts, starting from their GPU FP64 state
    at step s0 (the body has reached the floor, so the sample includes contact
    work), forward + reverse sweep. Single-threaded on one rollout, and OpenMP
    over `threads` rollouts on that many cores. Setup (mesh, muscle selection,
    factorisation) is excluded."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    o = O.OracleScene(d)
    setup_s = time.perf_counter() - t0
    n3 = 3 * d.N
    res = {}
    for name, R in (("single", 1), ("multi", threads)):
        xs = np.ascontiguousarray(np.tile(x_s0, (R, 1)))
        vs = np.ascontiguousarray(np.tile(v_s0, (R, 1)))
        a = np.ascontiguousarray(acts[:R, s0:s0 + steps])
        gx = np.full((R, n3), 1.0 / n3)
        t0 = time.perf_counter()
        _, _, its = o.fwd_adj_batch(xs, vs, steps, gx, np.zeros_like(gx), act=a, nthreads=R)
        dt = time.perf_counter() - t0
        res[name] = (float(d.T) * float(its.sum()) / dt, dt, R)
    v1, dt1, _ = res["single"]
    vm, dtm, Rm = res["multi"]
    info = host_cpu_info()
    smp = "oracle FP64 C++, steps %d-%d (ground contact) of the %s rollouts, forward + reverse sweep" % (
        s0, s0 + steps - 1, d.name)
    return dict(value=vm, unit="tet-iterations/s", cores=Rm, kind="port",
                sample=smp + ": %d rollouts over %d threads (%.1f s)" % (Rm, Rm, dtm),
                single_thread={"value": v1, "cores": 1, "sample": smp + ": 1 rollout (%.1f s)" % dt1},
                nproc=info["nproc"], cpu_model=info["model"], setup_s=setup_s)


def cpu_baseline_packed(d, meshes, x_s0, v_s0, acts, s0, steps, threads, n_meshes):
    """Oracle timed on a bounded sample of the packed-mesh world. Its bodies
    share no constraint and each keeps its own obstacles, so the sample is
    the first `n_meshes` bodies, each as a scene of its own (moved back to
    where its mesh was built, which changes nothing but the coordinates),
    steps [s0, s0 + steps) from the world's GPU FP64 state at step s0,
    forward + reverse sweep; one body per thread on `threads` threads. Setup
    (mesh, muscle selection, factorisation) is excluded; each body's
    iterations are its own (the world iterates until every body converges)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    jobs = []
    nv = ng = 0
    for b, m in enumerate(meshes):
        if b < n_meshes:
            off = d.x_rest[nv] - m.x_rest[0]
            xs = (x_s0.reshape(-1, 3)[nv:nv + m.N] - off).reshape(1, -1)
            vs = v_s0.reshape(-1, 3)[nv:nv + m.N].reshape(1, -1)
            a = acts[:1, s0:s0 + steps, ng:ng + m.n_groups] if m.seg_group.size else None
            jobs.append((m, np.ascontiguousarray(xs), np.ascontiguousarray(vs),
                         None if a is None else np.ascontiguousarray(a)))
        nv += m.N
        if m.seg_group.size:
            ng += m.n_groups
    t0 = time.perf_counter()
    scenes = [O.OracleScene(m) for m, _, _, _ in jobs]
    setup_s = time.perf_counter() - t0

    def run(k):
        m, xs, vs, a = jobs[k]
        n3 = 3 * m.N
        gx = np.full((1, n3), 1.0 / n3)
        t1 = time.perf_counter()
        _, _, its = scenes[k].fwd_adj_batch(xs, vs, steps, gx, np.zeros_like(gx), act=a, nthreads=1)
        dt = time.perf_counter() - t1
        log("  CPU sample: body %d (%d tets) done in %.1f s" % (k, m.T, dt))
        return float(m.T) * float(its.sum()), dt

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(run, range(len(jobs))))
    wall = time.perf_counter() - t0
    work = sum(w for w, _ in res)
    info = host_cpu_info()
    smp = ("oracle FP64 C++, bodies 0-%d of the packed world (%d tets of %d), each a scene of its own, "
           "steps %d-%d (sphere and ground contact), forward + reverse sweep, one body per thread (%.1f s)"
           % (len(jobs) - 1, sum(m.T for m, _, _, _ in jobs), d.T, s0, s0 + steps - 1, wall))
    return dict(value=work / wall, unit="tet-iterations/s", cores=min(threads, len(jobs)), kind="port", sample=smp,
                per_thread={"value": work / sum(t for _, t in res), "cores": 1,
                            "sample": "the same work over the summed per-body thread times"},
                nproc=info["nproc"], cpu_model=info["model"], setup_s=setup_s)


def oracle_check(d, x0, v0, acts, steps, gx, gv, rows):
    """FP64 oracle forward + reverse sweep of the listed rollouts at the full
    horizon (OpenMP over them): positions and gradients to compare with."""
    from oracle import oracle as O
    o = O.OracleScene(d)
    R = len(rows)
    xs = np.ascontiguousarray(x0[rows])
    vs = np.ascontiguousarray(v0[rows])
    a = np.ascontiguousarray(acts[rows])
    xT, _ = o.rollout_batch(xs, vs, steps, act=a, nthreads=R)
    gx0, _, _ = o.fwd_adj_batch(xs, vs, steps, np.ascontiguousarray(gx[rows]), np.ascontiguousarray(gv[rows]),
                                act=a, nthreads=R)
    return xT, gx0


def rel_rows(a, b):
    """Per row: max |a - b| / max |b|."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.abs(a - b).max(1) / np.maximum(np.abs(b).max(1), 1e-300)


def solve_traffic(rhs, nnz_lower, top_rows, leaf_blocks):
    """(HBM bytes per solve, source) from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE runs, scaled by a copy kernel of known size),
    when it was measured on this same factor, dense top block, dense leaf
    subtrees and right-hand-side count; else None. It is not measured in this run: the
    source names the counter run."""
    path = os.path.join(ROOT, "profiles", "solve_c4_R256_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    vals, section = {}, False
    for line in open(path):
        if line.startswith("Values read by bench.py"):
            section = True
        elif section and line.startswith("- ") and ":" in line:
            k, v = line[2:].split(":", 1)
            if v.strip():
                vals[k.strip()] = v.strip() if k.strip() == "measured_by" else v.strip().split()[0]
    try:
        if (int(vals["rhs"]) == rhs and int(vals["nnz_lower"]) == nnz_lower and int(vals["top_rows"]) == top_rows
                and int(vals["leaf_blocks"]) == leaf_blocks):
            return float(vals["hbm_bytes_per_solve"]), "profiles/solve_c4_R256_hbm_traffic.md (%s)" % (
                vals.get("measured_by", "rocprofv3 FETCH_SIZE/WRITE_SIZE counter runs"))
    except (KeyError, ValueError):
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["c4", "c5"], default="c4",
                    help="c4: 40x12x12 muscle box, R rollouts; c5: 64 packed uneven meshes (one world), R = 1")
    ap.add_argument("--steps", type=int, default=2, help="timed forward+adjoint passes")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["mixed", "mixed_adjoint"], default="mixed_adjoint",
                    help="mixed: FP32 local step forward and in the reverse sweep; mixed_adjoint: FP64 forward "
                         "(the FP64 path's trajectory and branches), FP32 local-step VJP in the reverse sweep")
    ap.add_argument("--rollouts", type=int, default=256)
    ap.add_argument("--parts", type=int, default=2,
                    help="independent sub-batches run concurrently on their own streams (engine.SplitScene); "
                         "one part's latency-bound solve levels overlap the other's element kernels")
    ap.add_argument("--sim-steps", type=int, default=0,
                    help="simulated steps per pass (0: 100 for c4, 150 for c5)")
    ap.add_argument("--ckpt", type=int, default=0,
                    help="checkpoint interval; 0: tape the whole rollout in the forward pass (no recompute in the "
                         "adjoint; about 2.3 GB of HBM per step for the 40x12x12 box at R = 256)")
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="sim steps of the plain FP64 comparator (0: same as --sim-steps)")
    ap.add_argument("--contact", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=24, help="steps of the CPU baseline sample")
    ap.add_argument("--cpu-from", type=int, default=14,
                    help="first step of the CPU sample (the dropped body reaches the floor at about step 14)")
    ap.add_argument("--cpu-threads", type=int, default=16, help="threads of the multi-core CPU baseline (at most nproc)")
    ap.add_argument("--cpu-bodies", type=int, default=16,
                    help="packed-mesh workload: bodies in the CPU baseline sample (one per thread)")
    ap.add_argument("--cpu-steps-c5", type=int, default=10,
                    help="packed-mesh workload: steps of the CPU baseline sample")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 batch pass (throughput and error figures)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: replay each step's PD block as a captured hipGraph (device-side early exit)")
    ap.add_argument("--solve-only", action="store_true",
                    help="only time back-to-back solves (short program for counter runs)")
    ap.add_argument("--plan", default="",
                    help="CUT,CHAINS: fix the whole-solve plan instead of the autotuned one (counter runs)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the rollout batch is not sharded across GPUs in this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from diff_fem_amd import engine, scene

    d = build_workload(args)
    prec = engine.PDM_MIXED if args.precision == "mixed" else engine.PDM_MIXED_ADJOINT
    if args.sim_steps <= 0:
        args.sim_steps = 150 if args.workload == "c5" else 100
    R, S = args.rollouts, args.sim_steps
    if args.ckpt <= 0:
        args.ckpt = S
    if args.workload == "c5":
        R = 1 if args.rollouts == 256 else args.rollouts
    parts = max(1, min(args.parts, R))
    Rp = R // parts  # rollouts (and 3 Rp right-hand sides) of the first part's solves
    log("setup: T=%d N=%d R=%d (%d part(s)) steps=%d" % (d.T, d.N, R, parts, S))
    t0 = time.perf_counter()
    # graph replays cannot carry per-solve events: with graphs the solves are
    # timed in one extra eager pass after the timed ones
    flags = engine.PDM_HOST_EARLY_EXIT | (engine.PDM_USE_GRAPH if args.graph else engine.PDM_TIME_SOLVES)
    if args.solve_only:  # one part's solve: the shape the pipeline launches
        sc = engine.Scene(d, R=Rp, precision=prec, flags=flags, checkpoint_every=args.ckpt)
        parts_list = [sc]
    elif parts > 1:
        sc = engine.SplitScene(d, R=R, parts=parts, precision=prec, flags=flags,
                               checkpoint_every=args.ckpt)
        parts_list = sc.parts
    else:
        sc = engine.Scene(d, R=R, precision=prec, flags=flags, checkpoint_every=args.ckpt)
        parts_list = [sc]
    p0 = parts_list[0]
    info = sc.info()
    log("setup done in %.1fs: %s" % (time.perf_counter() - t0, info))
    for k, p in enumerate(parts_list):
        log("part %d: solve launch shapes per level ((Ct, threads) fwd, bwd): %s" % (k, p.solve_shapes(),))
        log("part %d: solve column chains (streams): %d, subtree cut level: %d" % (k, p.solve_chains, p.solve_cut))
    dev = torch.device("cuda:0")
    flops = 2.0 * 2.0 * info["nnz_lower"] * 3 * Rp
    if args.plan:
        cut, chains = (int(v) for v in args.plan.split(","))
        for p in parts_list:
            p.set_solve_plan(cut, chains)
            p.solve_shapes()
        log("solve plan fixed: subtree cut %d, %d column chain(s)" % (p0.solve_cut, p0.solve_chains))
    if args.solve_only:
        ms = sc.time_solve(20)
        if args.copy_calib:
            a = torch.ones(args.copy_calib << 17, dtype=torch.float64, device=dev)
            b = torch.empty_like(a)
            for _ in range(3):
                b.copy_(a)
            torch.cuda.synchronize()
        print(json.dumps({"solve_ms": ms, "flops_per_solve": flops, "rhs": 3 * Rp, "nnz_lower": info["nnz_lower"],
                          "copy_calib_bytes": args.copy_calib << 20}), flush=True)
        return
    acts = torch.from_numpy(scene.activations_sinusoid(4, S, d.n_groups, R, d.h)).to(dev).contiguous()
    x0 = torch.from_numpy(np.tile(d.x_rest.ravel(), (R, 1))).to(dev).contiguous()
    v0 = torch.zeros_like(x0)
    gx = torch.ones_like(x0) / x0.numel()
    gv = torch.zeros_like(x0)
    xT = torch.empty_like(x0)
    vT = torch.empty_like(x0)

    def one_pass():
        sc.forward_device(x0, v0, S, act=acts, xT=xT, vT=vT, record=True)
        g = sc.backward_device(gx, gv)
        return g

    for w in range(args.warmup):
        one_pass()
        torch.cuda.synchronize()
        free_b, total_b = torch.cuda.mem_get_info(dev)
        log("warmup %d done; device memory in use %.1f of %.1f GB" % (w, (total_b - free_b) / 1e9, total_b / 1e9))
    sc.solve_timing()  # restart the per-solve event counters
    g0 = sc.graph_launches()
    n_solve, solve_ms = 0, 0.0
    times, tet_its, fwd_ms, bwd_ms = [], [], [], []
    g = None
    for k in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = one_pass()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        its = sc.iterations()
        tet_its.append(float(d.T) * float(its.sum()))
        times.append(dt)
        tm = sc.timing()
        ns, sm = sc.solve_timing()
        n_solve += ns
        solve_ms += sm
        fwd_ms.append(tm["forward_ms"])
        bwd_ms.append(tm["backward_ms"])
        cps, cjt = sc.contact_stats()
        log("pass %d: %.3fs fwd %.1fms bwd %.1fms mean iters %.2f contact passes/step %.2f jacobi/step %.2f" % (
            k, dt, tm["forward_ms"], tm["backward_ms"], its.mean(), cps.mean(), cjt.mean()))
    total_t = sum(times)
    value = sum(tet_its) / total_t
    replays = sc.graph_launches() - g0
    out = {
        "metric": "tet-iterations/sec fwd+adjoint, batched mixed vs plain FP64 on GPU at stated err",
        "value": value, "unit": "tet-iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * total_t / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-local/fp64-solve" if prec == engine.PDM_MIXED else "fp64-forward/fp32-adjoint-local/fp64-solve", "data": "synthetic (seeded box mesh, seeded activations)",
        "config": {"workload": ("c4_muscle_40x12x12_R%d_%dsteps_fwd+adjoint%s" % (R, S, "" if d.contact else "_nocontact")
                                if args.workload == "c4" else "c5_packed_64_meshes_R%d_%dsteps_fwd+adjoint" % (R, S)),
                   "tets": d.T, "vertices": d.N, "rollouts": R, "sim_steps": S, "max_iter": d.max_iter,
                   "checkpoint_every": args.ckpt, "contact": bool(d.contact), "nnz_factor": info["nnz_factor"],
                   "hipgraph": replays > 0, "graph_replays_per_pass": replays / max(args.steps, 1),
                   "sub_batches": parts,
                   "forward_ms": float(np.mean(fwd_ms)), "adjoint_ms": float(np.mean(bwd_ms))},
    }
    if args.graph:
        # Solves inside a replayed graph cannot be bracketed by events: one
        # more pass, eager with the same concurrency, times every solve.
        for p in parts_list:
            p.set_batch(p.R, prec, engine.PDM_HOST_EARLY_EXIT | engine.PDM_TIME_SOLVES, args.ckpt)
        sc.solve_timing()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one_pass()
        torch.cuda.synchronize()
        t_eager = time.perf_counter() - t0
        n_solve, solve_ms = sc.solve_timing()
        out["config"]["eager_pass_s"] = t_eager
        log("instrumented eager pass: %.3fs, %d solves" % (t_eager, n_solve))
    # Roofline of the dominant kernel: the supernodal triangular solve
    # (k_fwd_level + k_bwd_level over all levels, one part's 3R/parts
    # right-hand sides per launch sequence). Its mean duration comes from HIP
    # events around every solve of a forward+adjoint pass, on each part's
    # stream (so it includes the other part's kernels sharing the GPU
    # meanwhile). At 3R/parts = 384 it is FP64-MFMA bound by arithmetic
    # intensity: algorithmic flops = 2 sweeps x 2 x nnz(L, lower trapezoid) x
    # RHS; algorithmic bytes = factor read twice + right-hand sides read and
    # written once per sweep.
    ms_live = solve_ms / max(n_solve, 1)
    bytes_alg = 2.0 * info["nnz_lower"] * 8 + 2.0 * 2 * 3 * Rp * d.N * 8
    if flops / bytes_alg >= 78.6e12 / 8.0e12:  # above the FP64 ridge: matrix-core bound
        achieved = flops / (ms_live * 1e-3) / 1e12
        rl = {"bound": "mfma", "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s", "frac": achieved / 78.6}
    else:  # few right-hand sides: the factor stream bounds it
        achieved = bytes_alg / (ms_live * 1e-3) / 1e9
        rl = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0}
    tr = solve_traffic(3 * Rp, info["nnz_lower"], info["top_rows"], info["leaf_blocks"])
    rl.update({"traffic": tr[0] if tr else None, "traffic_source": tr[1] if tr else None,
               "kernel": "supernodal solve (dense leaf subtrees, level sweeps, dense top separators), %d RHS" % (3 * Rp),
               "column_chains": p0.solve_chains,
               "ms_per_solve": ms_live, "solves_timed": n_solve,
               "ms_per_solve_back_to_back": p0.time_solve(20), "flops_per_solve": flops,
               "bytes_alg_per_solve": bytes_alg})
    # the same solve alone on the GPU (no other part's kernels beside it)
    rl["frac_back_to_back"] = flops / (rl["ms_per_solve_back_to_back"] * 1e-3) / 1e12 / 78.6 \
        if rl["bound"] == "mfma" else bytes_alg / (rl["ms_per_solve_back_to_back"] * 1e-3) / 1e9 / 8000.0
    out["roofline"] = rl
    # Whole-pass HBM fraction of the metric's own definition: algorithmic
    # bytes per local/global iteration for R rollouts (local step with FP32
    # local data, s = 4, plus both triangular sweeps), per executed
    # tet-iteration, counted for the forward and again for the adjoint, plus
    # the checkpoint traffic; over the timed pass time and 8 TB/s.
    if args.workload == "c4":
        def per_tet_it(s_b):
            b_local = d.T * (16 + 10 * s_b) + R * 3 * d.N * s_b + R * 3 * d.N * 8
            b_solve = 2.0 * info["nnz_lower"] * 8 + 2.0 * R * 3 * d.N * 8
            return (b_local + b_solve) / (R * d.T)
        s_fwd = 4.0 if prec == engine.PDM_MIXED else 8.0
        b_pass = (per_tet_it(s_fwd) + per_tet_it(4.0)) * (sum(tet_its) / len(tet_its)) \
            + (S / args.ckpt) * 2.0 * R * 3 * d.N * 8
        t_pass = total_t / args.steps
        out["pass_roofline"] = {"bound": "hbm", "achieved": b_pass / t_pass / 1e9, "peak": 8000.0, "unit": "GB/s",
                                "frac": b_pass / t_pass / 8.0e12, "bytes_alg_per_pass": b_pass,
                                "bytes_alg_per_tet_iteration": 0.5 * (per_tet_it(s_fwd) + per_tet_it(4.0))}
    tet_its_pass = sum(tet_its) / len(tet_its)
    x_mixed = xT.cpu().numpy()
    g_mixed = {k: g[k].cpu().numpy() for k in ("x0", "v0", "act", "E")}
    if args.workload == "c4" and not args.no_fp64:
        def batch_pass(p_, label):
            """Warm-up + one timed pass of the whole batch at precision p_
            (same scene, buffers and flags); returns its result."""
            log("%s batch (R = %d) ..." % (label, R))
            fl = engine.PDM_HOST_EARLY_EXIT | (engine.PDM_USE_GRAPH if args.graph else 0)
            for p in parts_list:
                p.set_batch(p.R, p_, fl, args.ckpt)
            one_pass()  # warm-up: first launches of this precision's kernels
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gg = one_pass()
            torch.cuda.synchronize()
            tp = time.perf_counter() - t0
            it = sc.iterations()
            return dict(t=tp, its=it, value=float(d.T) * float(it.sum()) / tp, x=xT.cpu().numpy(),
                        g={k: gg[k].cpu().numpy() for k in ("x0", "v0", "act", "E")})

        def errors(x_m, g_m, its_m, ref):
            ep = rel_rows(x_m, ref["x"])
            eg = np.maximum.reduce([rel_rows(g_m[k], ref["g"][k]) for k in ("x0", "v0", "act")])
            return {"pos_max": float(ep.max()), "grad_max": float(eg.max()),
                    "grad_E": float(np.abs(g_m["E"] - ref["g"]["E"]).max() / np.abs(ref["g"]["E"]).max()),
                    "rollouts_pos_above_1e-5": int((ep > 1e-5).sum()),
                    "rollouts_grad_above_1e-4": int((eg > 1e-4).sum()),
                    "same_iterations": bool(np.array_equal(its_m, ref["its"]))}
        # The same batch in FP64 everywhere: its throughput, and the reference
        # the mixed errors are measured on (all rollouts, full horizon).
        ref = batch_pass(engine.PDM_FP64, "FP64")
        out["fp64_batch"] = {"value": ref["value"], "unit": "tet-iterations/s", "rollouts": R, "s_per_pass": ref["t"]}
        out["err"] = {"vs": "FP64 GPU batch, same inputs, all rollouts, full horizon",
                      "grads": "x0, v0, activations per rollout; E summed over rollouts"}
        out["err"].update(errors(x_mixed, g_mixed, its, ref))
        # the other mixed form, for comparison
        alt = engine.PDM_MIXED if prec == engine.PDM_MIXED_ADJOINT else engine.PDM_MIXED_ADJOINT
        oth = batch_pass(alt, "alternative mixed")
        out["other_mixed"] = {"precision": "mixed" if alt == engine.PDM_MIXED else "mixed_adjoint",
                              "dtype": "fp32-local/fp64-solve" if alt == engine.PDM_MIXED
                              else "fp64-forward/fp32-adjoint-local/fp64-solve",
                              "value": oth["value"], "unit": "tet-iterations/s", "s_per_pass": oth["t"],
                              "err": errors(oth["x"], oth["g"], oth["its"], ref)}
        if not args.no_cpu:
            rows = [0, R - 1]
            log("oracle check of rollouts %s at %d steps ..." % (rows, S))
            xo, go = oracle_check(d, x0.cpu().numpy(), v0.cpu().numpy(), acts.cpu().numpy(), S,
                                  gx.cpu().numpy(), gv.cpu().numpy(), rows)
            out["err"]["oracle"] = {"rollouts": rows,
                                    "fp64_pos": float(rel_rows(ref["x"][rows], xo).max()),
                                    "fp64_grad_x0": float(rel_rows(ref["g"]["x0"][rows], go).max()),
                                    "mixed_pos": float(rel_rows(x_mixed[rows], xo).max()),
                                    "mixed_grad_x0": float(rel_rows(g_mixed["x0"][rows], go).max())}
        log("errors: %s; other mixed form: %s" % (out["err"], out["other_mixed"]))
    # one re-factorisation (what a material update costs; outside the timed
    # region): L's values, the multifrontal factor and the W blocks on the
    # device; the same factorisation in host code (one thread) beside it
    t0 = time.perf_counter()
    p0.set_material(d.E)
    f_dev = p0.timing()["factor_ms"]
    f_wall = 1000.0 * (time.perf_counter() - t0)
    p0.refactor(on_host=True)
    f_host = p0.timing()["factor_ms"]
    p0.refactor()
    out["factor"] = {"ms": f_dev, "wall_ms": f_wall, "where": "device (per-level extend-add, panel, update kernels)",
                     "host_code_ms": f_host, "flops": info["factor_flops"],
                     "gflops_per_s": info["factor_flops"] / (f_dev * 1e6)}
    x_s0 = v_s0 = None
    if args.workload == "c5" and not args.no_cpu:
        # the world's state at the start of the CPU sample (FP64 forward)
        for p in parts_list:
            p.set_batch(p.R, engine.PDM_FP64, engine.PDM_HOST_EARLY_EXIT, args.ckpt)
        xs, vs = torch.empty_like(x0), torch.empty_like(v0)
        sc.forward_device(x0, v0, args.cpu_from, act=acts[:, :args.cpu_from].contiguous(), xT=xs, vT=vs)
        torch.cuda.synchronize()
        x_s0, v_s0 = xs.cpu().numpy()[0], vs.cpu().numpy()[0]
    sc.close()
    del sc
    if not args.no_plain:
        log("plain FP64 comparator (R = 1) ...")
        ps = args.plain_steps or S
        pl = engine.Scene(d, R=1, precision=engine.PDM_FP64, flags=engine.PDM_HOST_EARLY_EXIT,
                          checkpoint_every=args.ckpt)
        px0 = x0[:1].contiguous()
        pv0 = v0[:1].contiguous()
        pa = acts[:1, :ps].contiguous()
        pg = gx[:1].contiguous()
        pgv = gv[:1].contiguous()

        def plain_pass():
            pl.forward_device(px0, pv0, ps, act=pa, record=True)
            pl.backward_device(pg, pgv)
        plain_pass()  # warm-up: buffers, solve autotuning, first launches
        torch.cuda.synchronize()
        pt, pit = [], []
        for _ in range(6):  # a pass is about 0.2 s and its host round trips jitter
            t0 = time.perf_counter()
            plain_pass()
            torch.cuda.synchronize()
            pt.append(time.perf_counter() - t0)
            pit.append(d.T * float(pl.iterations().sum()))
        pv = sum(pit) / sum(pt)
        out["plain_fp64"] = {"value": pv, "unit": "tet-iterations/s", "rollouts": 1, "sim_steps": ps,
                             "host_early_exit": True, "passes_timed": len(pt), "warmup": 1,
                             "pass_s_min": min(pt), "pass_s_max": max(pt)}
        out["speedup_vs_plain_fp64"] = value / pv
        if "fp64_batch" in out:
            out["fp64_batch"]["speedup_vs_plain_fp64"] = out["fp64_batch"]["value"] / pv
        if args.workload == "c4" and not args.no_cpu:
            # state at the start of the CPU sample (after the drop onto the floor)
            xs, vs = torch.empty_like(px0), torch.empty_like(pv0)
            pl.forward_device(px0, pv0, args.cpu_from, act=acts[:1, :args.cpu_from].contiguous(), xT=xs, vT=vs)
            torch.cuda.synchronize()
            x_s0, v_s0 = xs.cpu().numpy()[0], vs.cpu().numpy()[0]
        pl.close()
    if not args.no_cpu and x_s0 is not None:
        threads = min(host_cpu_info()["nproc"], args.cpu_threads)
        if args.workload == "c4":
            log("CPU baseline (oracle, 1 and %d threads) ..." % threads)
            out["cpu_baseline"] = cpu_baseline(d, x_s0, v_s0, acts.cpu().numpy(), args.cpu_from, args.cpu_steps,
                                               threads)
        else:
            # the whole world in one oracle scene would spend minutes in its
            # muscle selection (every tet against every segment): bodies apart
            log("CPU baseline (oracle, first %d bodies on %d threads, %d steps) ..." % (
                args.cpu_bodies, threads, args.cpu_steps_c5))
            out["cpu_baseline"] = cpu_baseline_packed(d, scene.config5_meshes(seed=5), x_s0, v_s0,
                                                      acts.cpu().numpy(), args.cpu_from, args.cpu_steps_c5, threads,
                                                      args.cpu_bodies)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
