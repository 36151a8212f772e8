#!/usr/bin/env python3
"""Benchmark: tet updates/s of the deformable step on one MI355X.

Workload (default): the 1 m cube of 40x40x40 hexes split into 384,000 Kuhn
tets (68,921 nodes), nodes jittered +-2% of the edge, per-tet neo-Hookean
material with E = 5e6 exp(0.3 N(0,1)) Pa and nu ~ U[0.25, 0.4], bottom face
pinned, gravity, dt = 1/60 s, 1 substep, PCG tolerance 1e-3 with at most 64
iterations (the reference Settings defaults).  A step is one complete
deformable step: element kernels, assembly, three implicit solves, velocity
and position updates and the plastic commit.

value  = tets x substeps x steps / wall time of the timed steps, mixed
         precision (FP32 element kernels and PCG vectors, FP64 rest inverse,
         node state, dot products and pivots).
fp64   = the same for the plain FP64 path.
Inputs are resident in HBM before the timed region; every step returns after
its stream has synchronized, so the wall clock brackets finished work.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tet updates/sec: mixed-precision batched vs plain FP64, 1 GPU, ≤1e-5 rel force err"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(api, mesh, mats, precision, steps, warmup, params):
    ctx = api.Context(0, precision)
    b = ctx.add_body(mesh, mats)
    for _ in range(warmup):
        ctx.step(params)
    reps = []
    t0 = time.perf_counter()
    for _ in range(steps):
        reps.append(ctx.step(params))
    wall = time.perf_counter() - t0
    return ctx, b, wall, reps


def bytes_per_tet(kind, s, nt):
    """Algorithmic HBM bytes per tet for one launch of a kernel, s = bytes of
    the element precision, nt = nodes per tet.  Dm^-1 is FP64 (72 B) in both
    paths; node vectors are xyz of s bytes; conn is 16 B."""
    if kind == "assemble":  # conn, Dm^-1, V/lambda/mu, R write, 12 slot writes, u read
        return 16 + 72 + 3 * s + 9 * s + 12 * s + nt * 3 * s
    if kind == "op_element":  # conn, Dm^-1, V/lambda/mu, R read, slot writes, d read
        return 16 + 72 + 3 * s + 9 * s + 12 * s + nt * 3 * s
    if kind == "op_gather":  # slot reads, incidence codes, p read, q write, mass, anchored
        return 12 * s + 16 + nt * (3 * s + 3 * s + 8 + 1 + 4)
    raise ValueError(kind)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=40)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--preconditioner", choices=["factor", "jacobi"], default="factor")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        return 2

    from talus_amd import api, scenes

    scene = scenes.cube(args.cells)
    body = scene.bodies[0]
    mesh, mats = scenes.build_mesh(body)
    T = int(mesh["tet_nodes"].shape[0])
    N = int(mesh["anchored"].shape[0])
    pre = api.FACTOR if args.preconditioner == "factor" else api.BLOCK_JACOBI
    params = api.default_params(substeps=scene.substeps, preconditioner=pre)
    log(f"[bench] workload cube{args.cells}: {T} tets {N} nodes")

    log("[bench] mixed precision")
    ctx_m, bm, wall_m, reps_m = run_gpu(api, mesh, mats, api.MIXED, args.steps, args.warmup, params)
    value = T * scene.substeps * args.steps / wall_m
    cg_m = sum(r["cg_iterations"] for r in reps_m) / args.steps
    capped_m = sum(r["cg_capped"] for r in reps_m)
    log(f"[bench] mixed {1e3 * wall_m / args.steps:.3f} ms/step, {value:.4g} tet/s, cg/step {cg_m:.1f}")
    finfo = ctx_m.factor_info()
    if finfo["fronts"]:
        finfo["tflops"] = finfo["flops"] / max(finfo["factor_seconds"], 1e-12) / 1e12
        log(f"[bench] factor {json.dumps(finfo)}")

    fp64 = None
    ctx_d = None
    if not args.skip_fp64:
        log("[bench] fp64")
        ctx_d, bd, wall_d, reps_d = run_gpu(api, mesh, mats, api.FP64, args.steps, args.warmup, params)
        v64 = T * scene.substeps * args.steps / wall_d
        fp64 = {"value": v64, "ms_per_step": 1e3 * wall_d / args.steps,
                "cg_per_step": sum(r["cg_iterations"] for r in reps_d) / args.steps,
                "cg_capped": sum(r["cg_capped"] for r in reps_d)}
        log(f"[bench] fp64 {fp64['ms_per_step']:.3f} ms/step, {v64:.4g} tet/s")

    # relative force error of the mixed element kernel on an FP64 state
    rel_err = None
    if ctx_d is not None:
        x, _, _ = ctx_d.state(bd)
        tf_d, *_ = ctx_d.eval_forces(bd, x)
        tf_m, *_ = ctx_m.eval_forces(bm, x)
        num = np.linalg.norm((tf_m - tf_d).reshape(T, -1), axis=1).max()
        den = np.linalg.norm(tf_d.reshape(T, -1), axis=1).max()
        rel_err = float(num / den)
        log(f"[bench] rel force error mixed vs fp64: {rel_err:.3e}")

    # per-kernel launch durations (HIP events on the context stream)
    nt = N / T
    kinds = {"assemble": 0, "op_element": 1, "op_gather": 2}
    per_step_launches = {"assemble": 3 * scene.substeps,
                         "op_element": cg_m + 1 * scene.substeps,
                         "op_gather": cg_m}
    kernels = {}
    for name, kind in kinds.items():
        ms = ctx_m.time_kernel(kind, 50)
        B = bytes_per_tet(name, 4, nt) * T
        kernels[name] = {"avg_ms": ms, "bytes": B, "gbs": B / (ms * 1e-3) / 1e9,
                         "ms_per_step": ms * per_step_launches[name]}
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    kd = kernels[dominant]
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": kd["gbs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": kd["gbs"] / HBM_PEAK_GBS, "traffic": None}
    log(f"[bench] kernels {json.dumps(kernels)}")

    cpu = None
    if not args.skip_cpu:
        import oracle

        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from helpers import oracle_body

        threads = min(os.cpu_count() or 1, 64)
        orc, _ = oracle_body(body)
        log(f"[bench] cpu baseline, {threads} threads")
        t0 = time.perf_counter()
        orc.step(threads=threads)
        one = time.perf_counter() - t0
        n = max(1, min(20, int(args.cpu_seconds / max(one, 1e-3))))
        t0 = time.perf_counter()
        for _ in range(n):
            orc.step(threads=threads)
        wall = time.perf_counter() - t0
        cpu = {"value": T * n / wall, "unit": "tet updates/s", "cores": threads, "kind": "port",
               "sample": f"{n} FP64 steps of the same {T}-tet workload (after 1 warm-up) with the "
                         f"oracle's threaded C restatement"}
        log(f"[bench] cpu {cpu['value']:.4g} tet/s")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "tet updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 elements + f64 state",
        "data": "synthetic",
        "config": {
            "workload": f"cube {args.cells}^3 hexes -> {T} tets, {N} nodes, jittered 2%, per-tet "
                        f"lognormal E, bottom pinned, dt 1/60, {scene.substeps} substep, tol 1e-3, "
                        f"cg<=64 (<=32 with the factor), {args.preconditioner}-preconditioned PCG",
            "tets": T, "nodes": N, "precision": "mixed",
            "cg_per_step": cg_m, "cg_capped": capped_m,
            "preconditioner": args.preconditioner,
            "refactors": sum(r["refactors"] for r in reps_m),
        },
        "factor": finfo if finfo["fronts"] else None,
        "fp64": fp64,
        "mixed_over_fp64": (value / fp64["value"]) if fp64 else None,
        "rel_force_error": rel_err,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
