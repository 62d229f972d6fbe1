"""Benchmark: tet-iterations/s of forward + adjoint, batched mixed precision
vs the plain FP64 GPU path, on one MI355X.

Default workload (the batched-rollout configuration): 40x12x12 box
(34,560 tets, 6,929 vertices), 64 straight fibres, ground plane with friction,
R = 256 rollouts sharing one Cholesky factor (run as two concurrent
sub-batches of 128 on their own streams), seeded sinusoidal activations,
100 steps x 10 iterations, forward + adjoint with the whole rollout taped by
the forward pass (about 232 GB of the 309 GB of HBM, so the adjoint recomputes
nothing; --ckpt k keeps checkpoints every k steps and recomputes segments),
FP32 local step with FP64 factor/solves/adjoint accumulation.

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


def cpu_baseline(d, steps, nthreads):
    """Oracle (C++ FP64 restatement, single thread) on a bounded sample: one
    rollout, `steps` steps forward + adjoint."""
    from oracle import oracle as O
    from diff_fem_amd import scene
    o = O.OracleScene(d)
    acts = scene.activations_sinusoid(4, steps, d.n_groups, 1, d.h)[0]
    x0 = d.x_rest.ravel().copy()
    v0 = np.zeros_like(x0)
    t0 = time.perf_counter()
    xT, vT, its = o.rollout(x0, v0, steps, act=acts, record=True)
    g, _ = o.backward(np.ones_like(xT), np.zeros_like(vT))
    dt = time.perf_counter() - t0
    tet_it = float(d.T * its[:, 0].sum())
    return dict(value=tet_it / dt, unit="tet-iterations/s", cores=1, kind="port",
                sample="oracle FP64 C++, 1 rollout x %d steps fwd+adjoint of %s (%.1f s)" % (steps, d.name, dt))


def solve_traffic(rhs, nnz_lower):
    """HBM bytes per solve from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE runs, scaled by a copy kernel of known size), when it was
    measured on this same factor and right-hand-side count; else None."""
    path = os.path.join(ROOT, "profiles", "solve_c4_R256_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if line.startswith("- ") and ":" in line:
            k, v = line[2:].split(":", 1)
            vals[k.strip()] = v.strip().split()[0]
    try:
        if int(vals["rhs"]) == rhs and int(vals["nnz_lower"]) == nnz_lower:
            return float(vals["hbm_bytes_per_solve"])
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
    ap.add_argument("--rollouts", type=int, default=256)
    ap.add_argument("--parts", type=int, default=2,
                    help="independent sub-batches run concurrently on their own streams (engine.SplitScene); "
                         "one part's latency-bound solve levels overlap the other's element kernels")
    ap.add_argument("--sim-steps", type=int, default=100)
    ap.add_argument("--ckpt", type=int, default=0,
                    help="checkpoint interval; 0: tape the whole rollout in the forward pass (no recompute in the "
                         "adjoint; about 2.3 GB of HBM per step for the 40x12x12 box at R = 256)")
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="sim steps of the plain FP64 comparator (0: same as --sim-steps)")
    ap.add_argument("--contact", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=4)
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
    R, S = args.rollouts, args.sim_steps
    if args.ckpt <= 0:
        args.ckpt = S
    if args.workload == "c5":
        R = 1 if args.rollouts == 256 else args.rollouts
    parts = max(1, min(args.parts, R))
    Rp = R // parts  # rollouts (and 3 Rp right-hand sides) of the first part's solves
    log("setup: T=%d N=%d R=%d (%d part(s)) steps=%d" % (d.T, d.N, R, parts, S))
    t0 = time.perf_counter()
    flags = engine.PDM_HOST_EARLY_EXIT | engine.PDM_TIME_SOLVES
    if args.graph:
        flags |= engine.PDM_USE_GRAPH
    if args.solve_only:  # one part's solve: the shape the pipeline launches
        sc = engine.Scene(d, R=Rp, precision=engine.PDM_MIXED, flags=flags, checkpoint_every=args.ckpt)
        parts_list = [sc]
    elif parts > 1:
        sc = engine.SplitScene(d, R=R, parts=parts, precision=engine.PDM_MIXED, flags=flags,
                               checkpoint_every=args.ckpt)
        parts_list = sc.parts
    else:
        sc = engine.Scene(d, R=R, precision=engine.PDM_MIXED, flags=flags, checkpoint_every=args.ckpt)
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
    n_solve, solve_ms = 0, 0.0
    times, tet_its, fwd_ms, bwd_ms = [], [], [], []
    for k in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one_pass()
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
    out = {
        "metric": "tet-iterations/sec fwd+adjoint, batched mixed vs plain FP64 on GPU at stated err",
        "value": value, "unit": "tet-iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * total_t / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32-local/fp64-solve", "data": "synthetic (seeded box mesh, seeded activations)",
        "config": {"workload": ("c4_muscle_40x12x12_R%d_%dsteps_fwd+adjoint%s" % (R, S, "" if d.contact else "_nocontact")
                                if args.workload == "c4" else "c5_packed_64_meshes_R%d_%dsteps_fwd+adjoint" % (R, S)),
                   "tets": d.T, "vertices": d.N, "rollouts": R, "sim_steps": S, "max_iter": d.max_iter,
                   "checkpoint_every": args.ckpt, "contact": bool(d.contact), "nnz_factor": info["nnz_factor"],
                   "hipgraph": bool(args.graph), "sub_batches": parts,
                   "forward_ms": float(np.mean(fwd_ms)), "adjoint_ms": float(np.mean(bwd_ms))},
    }
    # Roofline of the dominant kernel: the supernodal triangular solve
    # (k_fwd_level + k_bwd_level over all levels, one part's 3R/parts
    # right-hand sides per launch sequence), the largest share of kernel time.
    # Its mean duration comes from HIP events around every solve of the timed
    # passes, on each part's stream (so it includes the other part's kernels
    # sharing the GPU meanwhile). At 3R/parts = 384 it is FP64-MFMA bound by
    # arithmetic intensity: algorithmic flops = 2 sweeps x 2 x nnz(L, lower
    # trapezoid) x RHS; algorithmic bytes = factor read twice + right-hand
    # sides read and written once per sweep.
    ms_live = solve_ms / max(n_solve, 1)
    bytes_alg = 2.0 * info["nnz_lower"] * 8 + 2.0 * 2 * 3 * Rp * d.N * 8
    if flops / bytes_alg >= 78.6e12 / 8.0e12:  # above the FP64 ridge: matrix-core bound
        achieved = flops / (ms_live * 1e-3) / 1e12
        rl = {"bound": "mfma", "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s", "frac": achieved / 78.6}
    else:  # few right-hand sides: the factor stream bounds it
        achieved = bytes_alg / (ms_live * 1e-3) / 1e9
        rl = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0}
    rl.update({"traffic": solve_traffic(3 * Rp, info["nnz_lower"]),
               "kernel": "supernodal solve k_fwd_level+k_bwd_level, %d RHS" % (3 * Rp),
               "column_chains": p0.solve_chains,
               "ms_per_solve": ms_live, "solves_timed": n_solve,
               "ms_per_solve_back_to_back": p0.time_solve(20), "flops_per_solve": flops,
               "bytes_alg_per_solve": bytes_alg})
    # the same solve alone on the GPU (no other part's kernels beside it)
    rl["frac_back_to_back"] = flops / (rl["ms_per_solve_back_to_back"] * 1e-3) / 1e12 / 78.6 \
        if rl["bound"] == "mfma" else bytes_alg / (rl["ms_per_solve_back_to_back"] * 1e-3) / 1e9 / 8000.0
    out["roofline"] = rl
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
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pl.forward_device(px0, pv0, ps, act=pa, record=True)
        pl.backward_device(gx[:1].contiguous(), gv[:1].contiguous())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        pv = d.T * float(pl.iterations().sum()) / dt
        out["plain_fp64"] = {"value": pv, "unit": "tet-iterations/s", "rollouts": 1, "sim_steps": ps,
                             "host_early_exit": True}
        out["speedup_vs_plain_fp64"] = value / pv
        pl.close()
    if not args.no_cpu and args.workload == "c4":
        # (the packed 64-mesh world is out of reach of the single-threaded
        # oracle: its muscle selection scans every tet for every segment)
        log("CPU baseline (oracle, 1 thread) ...")
        out["cpu_baseline"] = cpu_baseline(d, args.cpu_steps, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
