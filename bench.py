#!/usr/bin/env python3
"""Benchmark: tet updates/s of the deformable step on one MI355X.

Workload (default, "bodies"): 512 independent Kuhn-cuboid bodies with tet
counts log-uniform in [200, 20000] (2,228,550 tets in total), per-body
neo-Hookean material E log-uniform [1e5, 1e8] Pa, nu ~ U[0.2, 0.45], yield on
for half of them (yield stress 1e-3 E, hardening 1e-2 E), x=0 face pinned,
gravity, dt = 1/60 s, 4 substeps, PCG tolerance 1e-3 preconditioned with the
supernodal block Cholesky (at most 32 iterations), every body in one batched
launch sequence per pass.  "cube" is the 40^3-hex cube (384,000 tets, per-tet
lognormal E, 1 substep).

A step is one complete deformable step: element kernels, assembly, three
implicit solves per substep, velocity and position updates, plastic commit.

Each batched run takes the workload's whole step count (120 for the bodies,
240 for the cube): `--warmup` untimed steps, the `--steps` timed steps (the
window), then the rest untimed, so that divergence anywhere in the run is
seen.

value   = sum(tets) x substeps x steps / wall time of the timed steps for the
          mixed-precision batched path (FP32 element kernels and vectors, FP64
          rest inverse, node state, dot products and Cholesky pivots), with
          the solve plan chosen by the autotuner.
fp64    = the same batched path in FP64; fp64_plain = FP64 with every body in
          its own context, stepped one after another (no batching), over the
          first --plain-steps steps of the same window from the same state.
Inputs are resident in HBM before the timed region; every step returns after
its stream has synchronized, so the wall clock brackets finished work.

The line is valid only when both batched runs stay finite with no capped
solve over every step and the mixed element forces are within 1e-5 of FP64
(at the window end and the run end); otherwise "value" is null, "failed"
says why and the exit status is 1.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tet updates/sec: mixed-precision batched vs plain FP64, 1 GPU, ≤1e-5 rel force err"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6  # MI355X dense FP64 matrix peak
FORCE_BAR = 1.0e-5  # the metric's relative force error bar, mixed against FP64
LEAF_CANDIDATES = (16, 24, 32, 48, 64)  # supernode leaf sizes the leaf tuner tries
KINDS = {"assemble": 0, "op_element": 1, "op_gather": 2, "solve_forward_lanes": 3,
         "solve_backward_lanes": 4, "solve_forward_rows": 5, "solve_backward_rows": 6,
         "solve_forward_split": 7, "solve_backward_split": 8}
KERNEL_NAMES = {"assemble": "k_assemble", "op_element": "k_op_element", "op_gather": "k_op_gather",
                "solve_forward_lanes": "k_solve_forward_lanes",
                "solve_backward_lanes": "k_solve_backward_lanes",
                "solve_forward_rows": "k_solve_forward", "solve_backward_rows": "k_solve_backward",
                "solve_forward_split": "k_solve_forward_split",
                "solve_backward_split": "k_solve_backward_split"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name):
    from talus_amd import scenes

    if name == "cube":
        sc = scenes.cube(40)
        desc = ("cube 40^3 hexes -> 384,000 tets, 68,921 nodes, jittered 2%, per-tet lognormal E "
                "(median 5e6 Pa, sigma 0.3), nu U[0.25,0.4], neo-Hookean, bottom pinned, dt 1/60, "
                "1 substep, PCG tol 1e-3 with the block Cholesky preconditioner")
    else:
        sc = scenes.bodies()
        desc = ("512 bodies, tets log-uniform [200, 20000] (Kuhn cuboids, 0.05 m cells), E log-uniform "
                "[1e5, 1e8] Pa, nu U[0.2,0.45], yield on p=0.5 (1e-3 E, hardening 1e-2 E), neo-Hookean, "
                "x=0 pinned, dt 1/60, 4 substeps, PCG tol 1e-3 with the block Cholesky preconditioner, "
                "all bodies batched")
    meshes = [scenes.build_mesh(b) for b in sc.bodies]
    return sc, meshes, desc


def run_batch(api, meshes, precision, params, steps, warmup, tune, plan=None, streams=1,
              snapshot=False, leaf=None, tune_leaf=False):
    """Every body in one context (supernode leaf size `leaf`), stepped
    `warmup` untimed steps, then `steps` timed steps (the window).  Returns
    the context, body ids, the window's wall time, the report of every step
    so far and, if asked, the state of every body at the start of the
    window."""
    ctx = api.Context(0, precision)
    ids = [ctx.add_body(m, mats) for m, mats in meshes]
    if leaf:
        ctx.set_leaf(leaf)
    ctx.set_streams(streams)
    reps = [ctx.step(params)]
    tuned = None
    if tune_leaf:
        tuned = ctx.tune_leaf(LEAF_CANDIDATES, 3)  # leaf size, then the per-level kernels
    elif plan and len(plan) == len(ctx.plan()):
        ctx.set_plan([int(c) for c in plan])
    elif tune or plan:
        tuned = ctx.tune(3)  # no plan, or one made for another level structure
    for _ in range(max(warmup - 1, 0)):
        reps.append(ctx.step(params))
    snap = body_snapshot(ctx, ids) if snapshot else None
    t0 = time.perf_counter()
    timed = [ctx.step(params) for _ in range(steps)]
    wall = time.perf_counter() - t0
    return ctx, ids, wall, reps + timed, tuned, snap


def read_plan(path):
    """A saved solve plan: its digits and, when the file names it
    ("leaf=N"), the supernode leaf size it was tuned for."""
    tokens = open(path).read().split()
    leaf = next((int(t[5:]) for t in tokens[1:] if t.startswith("leaf=")), None)
    return tokens[0], leaf


def body_snapshot(ctx, ids):
    """x, v, a, plastic points, rotations and volume frames of every body."""
    out = []
    for b in ids:
        x, v, a = ctx.state(b)
        out.append({"x": x, "v": v, "a": a, "plastic": ctx.plastic(b), "R": ctx.rotations(b),
                    "W": ctx.frames(b)})
    return out


def restore(ctx, b, snap):
    ctx.set_state(b, snap["x"], snap["v"], snap["a"])
    ctx.set_plastic(b, snap["plastic"])
    ctx.set_rotations(b, snap["R"])
    ctx.set_frames(b, snap["W"])


def health(reps):
    """Non-finite steps and capped solves over the given step reports."""
    return {"steps": len(reps), "non_finite_steps": sum(1 for r in reps if r["non_finite"]),
            "cg_capped": sum(r["cg_capped"] for r in reps)}


def plain(api, meshes, params, steps, snaps):
    """Every body in its own FP64 context, one after another, over the same
    window as the batch: each body starts from the batch's FP64 state at the
    window start, takes one untimed step (its factorization), is put back to
    that state and then steps `steps` timed steps."""
    wall = 0.0
    bad = capped = 0
    for (m, mats), snap in zip(meshes, snaps):
        with api.Context(0, api.FP64) as ctx:
            b = ctx.add_body(m, mats)
            restore(ctx, b, snap)
            ctx.step(params)
            restore(ctx, b, snap)
            t0 = time.perf_counter()
            reps = [ctx.step(params) for _ in range(steps)]
            wall += time.perf_counter() - t0
            bad += sum(1 for r in reps if r["non_finite"])
            capped += sum(r["cg_capped"] for r in reps)
    return wall, bad, capped


def force_error(ctx_m, ids_m, ctx_d, ids_d):
    """Worst body of max_e |f_e^mixed - f_e^FP64|_2 / max_e |f_e^FP64|_2 on the
    FP64 state: positions, committed plastic points, rotations and volume
    frames of the FP64 run are given to the mixed kernel, so the difference
    is the element kernel's precision alone (the two runs' trajectories
    differ within the PCG tolerance).  A body whose forces are not finite
    counts as an infinite error.  Returns (error, body)."""
    worst, at = 0.0, None
    for bm, bd in zip(ids_m, ids_d):
        x, _, _ = ctx_d.state(bd)
        own = (ctx_m.plastic(bm), ctx_m.rotations(bm), ctx_m.frames(bm))
        ctx_m.set_plastic(bm, ctx_d.plastic(bd))
        ctx_m.set_rotations(bm, ctx_d.rotations(bd))
        ctx_m.set_frames(bm, ctx_d.frames(bd))
        fd, *_ = ctx_d.eval_forces(bd, x)
        fm, *_ = ctx_m.eval_forces(bm, x)
        # the mixed run continues from its own plastic state
        ctx_m.set_plastic(bm, own[0])
        ctx_m.set_rotations(bm, own[1])
        ctx_m.set_frames(bm, own[2])
        T = fd.shape[0]
        with np.errstate(over="ignore", invalid="ignore"):
            num = float(np.linalg.norm((fm - fd).reshape(T, -1), axis=1).max())
            den = float(np.linalg.norm(fd.reshape(T, -1), axis=1).max())
        if not (np.isfinite(num) and np.isfinite(den)):
            err = float("inf")
        else:
            err = num / den if den > 0.0 else (0.0 if num == 0.0 else float("inf"))
        if not err <= worst:
            worst, at = err, bd
    return worst, at


def trajectory_drift(ctx_m, ids_m, ctx_d, ids_d, meshes):
    """Worst body of max |x^mixed - x^FP64| / max |x^FP64 - X| after both runs
    took the same steps: the mixed path's error in positions, FP64 as the
    reference (the PCG tolerance bounds it, not the element precision)."""
    worst, at = 0.0, None
    for (m, _), bm, bd in zip(meshes, ids_m, ids_d):
        xm, _, _ = ctx_m.state(bm)
        xd, _, _ = ctx_d.state(bd)
        X = m["rest_positions"].astype(np.float64)
        with np.errstate(over="ignore", invalid="ignore"):
            err = float(np.abs(xm - xd).max() / max(np.abs(xd - X).max(), 1e-300))
        if not err <= worst:
            worst, at = err, bd
    return worst, at


def traffic_for(kernel, work, plan):
    """HBM bytes per launch measured by rocprofv3 PMC counters for this kernel,
    workload and solve plan (profiles/traffic_<kernel>_<workload>.json), else
    None: a different plan launches the kernel on other levels."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{work}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec.get("bytes_per_launch") if rec.get("plan") == plan else None


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(sc, snaps, budget_s, threads, tolerance):
    """The oracle's FP64 step (its C restatement, block-Jacobi PCG to the same
    relative tolerance as the GPU, no iteration cap) on a bounded sample of
    the same bodies, each from the GPU FP64 batch's state at the window start,
    one body per worker thread; only the steps are timed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_body

    order = np.random.default_rng(99).permutation(len(sc.bodies))

    def prepared(i):
        b = oracle_body(sc.bodies[i])[0]
        s = snaps[i]
        b.set_state(s["x"], s["v"], s["a"])
        b.set_plastic(s["plastic"])
        b.set_rotations(s["R"])
        b.set_frames(s["W"])
        return b

    done = updates = tets = iters = 0
    wall = 0.0
    with cf.ThreadPoolExecutor(threads) as pool:
        while done < len(order) and wall < budget_s:
            chunk = [prepared(i) for i in order[done:done + threads]]
            t0 = time.perf_counter()
            reps = list(pool.map(lambda b: b.step(substeps=sc.substeps, threads=1, cg_iterations=1 << 20,
                                                  tolerance=tolerance), chunk))
            wall += time.perf_counter() - t0
            tets += sum(b.tets for b in chunk)
            updates += sum(b.tets * r["substeps"] for b, r in zip(chunk, reps))
            iters += sum(r["cg_iterations"] for r in reps)
            done += len(chunk)
    return {"value": updates / wall, "unit": "tet updates/s", "cores": threads, "kind": "port",
            "sample": f"one FP64 step ({sc.substeps} substeps) of {done} of the {len(order)} bodies "
                      f"(seeded order, {tets} tets) from the GPU FP64 batch's state at the window start, "
                      f"{threads} worker thread(s), one body per thread, with the oracle's C restatement "
                      f"(block-Jacobi PCG to relative tolerance {tolerance:g}, uncapped: "
                      f"{iters} CG iterations)",
            "host": host_cpu()}


def copy_probe(nbytes):
    """One device-to-device hipMemcpy of nbytes through the HIP runtime (the
    reference transfer for scaling PMC byte counters)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemcpy failed")  # 3 = hipMemcpyDeviceToDevice
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    log(f"[bench] copy probe: {nbytes} bytes read, {nbytes} bytes written")


def slab_run(steps):
    """The crushed slab (100,860 tets, mixed precision, the factor, tears
    detected in the step and the topology updated on the host): tet updates/s
    of the first `steps` steps with the footprint driven at 1 m/s, and the
    events of each step.  The tears start in the first step; the run stops
    at the tear onset window, past which the three-solve substep diverges
    on pinned fragments."""
    from talus_amd import api, scenes

    sc = scenes.slab()
    mesh, mats = scenes.build_mesh(sc.bodies[0])
    rest = mesh["rest_positions"].astype(np.float64)
    driven = scenes.slab_driven_nodes(sc, rest, mesh["anchored"])
    z0 = rest[driven, 2].copy()
    params = api.default_params(preconditioner=api.FACTOR, gravity=sc.gravity,
                                flags=api.DETECT_TEARS | api.RECORD_EVENTS)
    per_step = []
    with api.Context(0, api.MIXED) as ctx:
        b = ctx.add_body(mesh, mats)
        T = ctx.info(b)[1]
        wall = 0.0
        for s in range(steps):
            x, v, a = ctx.state(b)
            x[driven, 2] = scenes.slab_drive(sc, z0, s + 1)
            ctx.set_state(b, x, v, a)
            t0 = time.perf_counter()
            r = ctx.step(params)
            dt = time.perf_counter() - t0
            wall += dt
            ctx.events()
            per_step.append({"ms": 1e3 * dt, "yield_events": r["yield_events"], "tear_events": r["tear_events"],
                             "split_nodes": r["split_nodes"], "cg_iterations": r["cg_iterations"],
                             "refactors": r["refactors"], "non_finite": r["non_finite"]})
        nodes = ctx.info(b)[0]
    return {"metric": "tet updates/sec, crushed slab with plasticity and tearing, 1 GPU",
            "value": T * steps / wall, "unit": "tet updates/s", "n_gpus": 1, "steps": steps,
            "ms_per_step": 1e3 * wall / steps, "higher_is_better": True, "dtype": "f32 elements + f64 node state",
            "data": "synthetic",
            "config": {"workload": "slab 2 x 1 x 0.1 m, 82x41x5 hexes -> 100,860 tets, E 1 MPa, nu 0.3, "
                                   "rho 1000, yield 2e5 Pa, hardening 1e5 Pa, tear strain 0.35, ends held, "
                                   "0.4 x 0.4 m footprint driven down at 1 m/s, dt 1/60, 1 substep, factor",
                       "tets": T, "nodes_after": nodes, "per_step": per_step}}


# Algorithmic bytes of one element solve and one element update of the
# strainstep hex lattice (FP64, element-major rows, lattice.hip): everything
# the kernel must read and write once for that element.  The 18x18 inverse
# (its upper triangle, 171 doubles) counts in full when every element has its
# own, and in proportion to the distinct inverses when equal ones are shared.
LAT_SOLVE_BYTES = (72 + 72 + 32 + 8 + 4 + 4 + 64     # shape^-1, turn, nodes, bend, body, inverse id, 1/m
                   + 8 * 3 * 8 * 2                   # corner velocities read and written
                   + 192 + 144 + 144 + 144 * 2       # volume gradient, bias, softness, impulse r/w
                   + 144 + 144 + 24 + 16)            # coupling, image, gain/rhs/schur, volume impulse r/w
LAT_UPDATE_BYTES = (8 * 3 * 8 + 72 + 32 + 4 * 8 * 2  # corner positions, shape^-1, nodes, rotation r/w
                    + 96 + 144 + 144 + 64 + 24 + 1 + 4 + 4   # rest hourglass, rate, cap, 1/m, volume, flags
                    + 144 + 8 + 8 * 3 * 8 * 2        # impulses, corner velocities r/w (warm start)
                    + 72 + 144 + 192 + 144 + 144 + 24)  # turn, bias, volume gradient, coupling, image, scalars
LAT_INVERSE_BYTES = 18 * 19 // 2 * 8


def lattice_run(workload, steps, warmup, cpu_seconds, threads, direct=False):
    """Strainstep hex lattices on the GPU, FP64: "cube" is the 40^3 cube
    (64,000 hexes, every hex its own material), "bodies" the batch of 512
    uneven cuboid bodies in one launch sequence.  Per step islandSubsteps
    substeps of integrate, update + warm start, 4 colour-swept Gauss-Seidel
    passes of the 18-row element solve, advance."""
    from talus_amd import api, scenes

    if workload == "cube":
        specs = [scenes.hex_cube(40)]
        desc = ("40^3 hex cube, unit edge, nodes jittered 2 %, per-hex E = 5e6 exp(0.3 N) Pa, nu U[0.25, 0.4], "
                "rho 1000, bottom face held, dt 1/60, islandSubsteps, 4 passes")
    else:
        specs = scenes.hex_bodies()
        desc = ("512 elastic cuboid hex lattices, 0.05 m cells, about n/6 cells for n log-uniform in [200, 20000], "
                "E log-uniform [1e5, 1e8] Pa, nu U[0.2, 0.45], rho 1000, x = 0 face held, dt 1/60, each body "
                "its own islandSubsteps, all bodies in one launch sequence")
    bodies = []
    for kw in specs:
        kw = dict(kw)
        mats = kw.pop("materials")
        bodies.append((api.build_hex_lattice(materials=mats, **kw), mats))
    H = sum(int(m["hex_nodes"].shape[0]) for m, _ in bodies)
    params = api.lattice_params(flags=api.DIRECT_LAUNCH if direct else 0)
    with api.Context() as ctx:
        ids = [ctx.add_lattice(m, mats) for m, mats in bodies]
        for _ in range(max(warmup, 1)):
            r = ctx.lattice_step(params)
        S = r["substeps"]
        t0 = time.perf_counter()
        reps = [ctx.lattice_step(params) for _ in range(steps)]
        wall = time.perf_counter() - t0
        # kernel durations: one more step with an event pair around every
        # update and solve launch (the same stream, direct launches)
        timed = ctx.lattice_step(api.lattice_params(flags=api.KERNEL_TIMING))
        finite = all(bool(np.isfinite(ctx.lattice_state(i)[0]).all()) for i in ids)
    solves = reps[-1]["element_solves"]
    E = reps[-1]["elements"]
    share = min(1.0, timed["inverses"] / max(E, 1))
    solve_bytes = (LAT_SOLVE_BYTES + share * LAT_INVERSE_BYTES) * solves / timed["solve_launches"]
    updates = solves // 4  # every element updated once per substep it takes
    update_bytes = (LAT_UPDATE_BYTES + share * LAT_INVERSE_BYTES) * updates / max(timed["update_launches"], 1)
    solve_ms = timed["solve_ms"] / timed["solve_launches"]
    update_ms = timed["update_ms"] / max(timed["update_launches"], 1)
    kernels = {
        "k_lat_solve": {"avg_ms": solve_ms, "bytes": solve_bytes, "gbs": solve_bytes / (solve_ms * 1e-3) / 1e9,
                        "launches_per_step": timed["solve_launches"], "ms_per_step": timed["solve_ms"]},
        "k_lat_frame+warm": {"avg_ms": update_ms, "bytes": update_bytes,
                             "gbs": update_bytes / (update_ms * 1e-3) / 1e9,
                             "launches_per_step": timed["update_launches"], "ms_per_step": timed["update_ms"]},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    value = solves * steps / wall
    line = {
        "metric": "hex element solves/sec, strainstep lattice (FP64), 1 GPU",
        "value": value, "unit": "element solves/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * wall / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc, "bodies": len(bodies), "hexes": H, "elements": E, "substeps_max": S,
                   "element_solves_per_step": solves, "hex_updates_per_s": updates * steps / wall,
                   "inverses": timed["inverses"], "finite": finite,
                   "device_ms_per_step": sum(r["device_ms"] for r in reps) / steps},
        "kernels": kernels,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kernels[dom]["gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": kernels[dom]["bytes"], "avg_ms": kernels[dom]["avg_ms"]},
    }
    if cpu_seconds > 0:
        from oracle import hex as OH

        def orc(m, mats):
            om = {"rest": m["rest_positions"].astype(np.float64), "mass": m["node_mass"], "anchored": m["anchored"],
                  "hex": m["hex_nodes"], "volume": m["volume"], "lane": m["hex_material"], "voxel": m["voxel_size"]}
            return OH.LatticeBody(om, mats)

        obodies = [orc(m, mats) for m, mats in bodies]
        used = min(threads, len(obodies))
        with cf.ThreadPoolExecutor(used) as ex:
            list(ex.map(lambda b: b.step(), obodies))  # builds the elements
            t0 = time.perf_counter()
            n = 0
            while True:
                list(ex.map(lambda b: b.step(), obodies))
                n += 1
                if time.perf_counter() - t0 > cpu_seconds:
                    break
            cw = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": solves * n / cw, "unit": "element solves/s", "cores": used,
                                "kind": "port",
                                "sample": f"{n} step(s) of the whole workload with the oracle's C restatement "
                                          f"(serial cell-order sweep within a body, as the reference; "
                                          f"{used} thread(s), bodies spread over them)"}
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    return line


def contact_run(queries, warmup, cpu_seconds):
    """Contact generation (findContacts) of the 40^3 cube's surface (19,200
    triangles) against five sampled-field partners (scenes.contact_cube) on
    the GPU: queries/s over `queries` timed queries (each launches its
    kernels, synchronizes and copies the contacts to the host).  The CPU
    baseline runs the oracle's FP32 C restatement of the same query on one
    host thread and compares its contacts with the GPU's byte for byte."""
    from talus_amd import api, scenes

    w = scenes.contact_cube(40)
    q = dict(w["query"])
    dt = q.pop("dt")
    with api.Context(0, api.FP64) as ctx:
        b = ctx.add_body(w["mesh"], w["materials"])
        ctx.set_state(b, w["rest"], w["velocity"], np.zeros_like(w["rest"]))
        ids = [ctx.field_add(f) for f in w["fields"]]
        shapes = [dict(s, field=ids[s["field"]]) for s in w["shapes"]]
        for _ in range(max(warmup, 1)):
            got, found = ctx.find_contacts(b, shapes, dt, **q)
        t0 = time.perf_counter()
        for _ in range(queries):
            got, found = ctx.find_contacts(b, shapes, dt, **q)
        wall = time.perf_counter() - t0
        x, v, _ = ctx.state(b)
        srf = ctx.surface(b)
    T = int(srf["triangles"].shape[0])
    S = len(shapes)
    value = queries / wall
    line = {
        "metric": "contact queries/sec (findContacts: tet surface vs sampled-field partners), 1 GPU",
        "value": value, "unit": "queries/s", "n_gpus": 1, "steps": queries, "warmup": warmup,
        "ms_per_step": 1e3 * wall / queries, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": w["name"] + ": 40^3 cube (384,000 tets) falling onto a plane, three moving "
                   "0.25 m spheres (one spinning) and one far away, bias 2 mm, threshold 2 cm, culling on",
                   "triangles": T, "shapes": S, "contacts": int(found),
                   "triangle_shape_pairs_per_s": T * S * value},
    }
    if cpu_seconds > 0:
        from oracle import contact as OC

        oshapes = [dict(s, field=w["fields"][s["field"]]) for s in w["shapes"]]
        okw = dict(anchored=w["mesh"]["anchored"], friction=w["materials"]["friction"], **q)
        xs, vs = x.astype(np.float32), v.astype(np.float32)
        t0 = time.perf_counter()
        n = 0
        while True:
            want = OC.find_contacts(srf, xs, vs, oshapes, dt, **okw)
            n += 1
            if time.perf_counter() - t0 > cpu_seconds:
                break
        cw = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": n / cw, "unit": "queries/s", "cores": 1, "kind": "port",
                                "sample": f"{n} query(ies) of the same state with the oracle's FP32 C "
                                          f"restatement on one host thread"}
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
        line["parity"] = {"contacts_bit_equal": bool(got.tobytes() == want.tobytes()),
                          "gpu_contacts": int(found), "oracle_contacts": int(len(want))}
    return line


def adjoint_check(steps, every, elements, seed=5, rest_coords=8):
    """Differentiation workload: the 1 x 0.1 x 0.1 m beam (103x9x9 hexes,
    50,058 tets, per-tet lognormal E, nu U[0.25,0.4]), a converged 60-step
    rollout with checkpoints every 10 steps, loss |x_tip - target|^2; the
    autograd gradients in E and nu of `elements` random tets and in 8 rest
    coordinates (`rest_coords` of them) against central finite differences
    (h = 1e-6 relative for E, nu; 1e-7 m for rest positions) of the same FP64
    rollout."""
    import torch

    from talus_amd import adjoint, api, scenes

    sc = scenes.beam()
    body = sc.bodies[0]
    mesh, mats = scenes.build_mesh(body)
    E0 = mats["youngs_modulus"].astype(np.float64)
    nu0 = mats["poissons_ratio"].astype(np.float64)
    params = api.default_params(preconditioner=api.FACTOR, tolerance=1e-8, cg_iterations=500)
    out = {"workload": f"beam 103x9x9 hexes, {E0.size} tets, {steps} converged steps, checkpoints "
                       f"every {every}, loss |x_tip - target|^2", "tets": int(E0.size)}
    with api.Context(0, api.FP64) as ctx:
        b = ctx.add_body(mesh, mats)
        rest0 = torch.tensor(mesh["rest_positions"].astype(np.float64))
        pos = mesh["rest_positions"]
        tip = int(np.argmax(pos[:, 0] * 1e3 + pos[:, 1] + 1e-3 * pos[:, 2]))
        target = torch.tensor(sc.extra["target"], dtype=torch.float64)

        def loss(E, nu, rest):
            x, _ = adjoint.rollout(ctx, b, E, nu, rest, steps=steps, every=every, params=params)
            return ((x[tip] - target) ** 2).sum()

        E = torch.tensor(E0, requires_grad=True)
        nu = torch.tensor(nu0, requires_grad=True)
        rest = rest0.clone().requires_grad_(True)
        t0 = time.perf_counter()
        L = loss(E, nu, rest)
        t1 = time.perf_counter()
        L.backward()
        t2 = time.perf_counter()
        out.update(loss=float(L.detach()), forward_s=t1 - t0, backward_s=t2 - t1)
        log(f"[bench] adjoint: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s")
        grads = {"E": E.grad.numpy(), "nu": nu.grad.numpy(), "rest": rest.grad.numpy().reshape(-1)}
        rng = np.random.default_rng(seed)
        picks = rng.choice(E0.size, elements, replace=False)
        free = np.flatnonzero(np.repeat(mesh["anchored"] == 0, 3))
        rest_picks = rng.choice(free, rest_coords, replace=False)
        for kind, idxs in (("E", picks), ("nu", picks), ("rest", rest_picks)):
            errs, rels = [], []
            for idx in idxs:
                vals = []
                with torch.no_grad():
                    for sgn in (1.0, -1.0):
                        Ep, nup, rp = torch.tensor(E0), torch.tensor(nu0), rest0.clone().reshape(-1)
                        if kind == "E":
                            h = 1e-6 * E0[idx]
                            Ep[idx] += sgn * h
                        elif kind == "nu":
                            h = 1e-6 * nu0[idx]
                            nup[idx] += sgn * h
                        else:
                            h = 1e-7
                            rp[idx] += sgn * h
                        vals.append(float(loss(Ep, nup, rp.reshape(-1, 3))))
                fd = (vals[0] - vals[1]) / (2 * h)
                g = float(grads[kind][idx])
                log(f"[bench] adjoint {kind}[{int(idx)}]: gradient {g:.6e}, central difference {fd:.6e}")
                errs.append(abs(g - fd))
                rels.append(abs(g - fd) / max(abs(fd), 1e-300))
            scale = float(np.abs(grads[kind]).max())
            out[kind] = {"checked": len(idxs), "max_abs_err_over_max_grad": max(errs) / scale,
                         "median_rel_err": float(np.median(rels)), "max_grad": scale}
            log(f"[bench] adjoint {kind}: {json.dumps(out[kind])}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["bodies", "cube"], default="bodies")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--run-steps", type=int, default=None,
                    help="steps per batched run (default: the workload's, 120 or 240); a shorter run "
                         "reports no metric value (for profiling)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-refine", action="store_true",
                    help="skip the FP64 residual measurement and the refined step after the run")
    ap.add_argument("--refine", type=float, default=1e-9,
                    help="refine_tolerance of the refined step after the run")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--streams", type=int, default=3,
                    help="engines the batch is split over, stepped concurrently (talus_mi_set_streams)")
    ap.add_argument("--direct-launch", action="store_true",
                    help="launch the PCG iteration kernels directly instead of replaying a "
                         "captured HIP graph (for tracers that do not follow graph launches)")
    ap.add_argument("--plan", default=None,
                    help="fixed solve plan (digits, 2 per level) instead of autotuning the mixed context")
    ap.add_argument("--retune", action="store_true",
                    help="autotune the mixed context even where profiles/solve_plan_<workload>.txt "
                         "holds a plan tuned earlier")
    ap.add_argument("--tune-leaf", action="store_true",
                    help="autotune the supernode leaf size too (talus_mi_tune_leaf), then the plan")
    ap.add_argument("--leaf", type=int, default=None,
                    help="supernode leaf size (default: the saved plan's, else the library's 64)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-seconds-single", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--plain-steps", type=int, default=10,
                    help="timed steps of the unbatched FP64 run, from the start of the timed window")
    ap.add_argument("--slab", type=int, default=0, metavar="STEPS",
                    help="instead of the batch: the crushed slab's first STEPS steps with tearing")
    ap.add_argument("--adjoint", action="store_true",
                    help="run the differentiation check (50k-tet beam) instead of the throughput bench")
    ap.add_argument("--adjoint-steps", type=int, default=60)
    ap.add_argument("--adjoint-elements", type=int, default=32)
    ap.add_argument("--lattice", choices=["cube", "bodies"], default=None,
                    help="strainstep hex lattices (FP64): element solves/s of the 40^3 cube or the 512-body batch")
    ap.add_argument("--skip-contacts", action="store_true",
                    help="leave the contact generation measurement out of the default line")
    ap.add_argument("--contacts", action="store_true",
                    help="contact generation queries/s on the 40^3 cube against five sampled-field partners")
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with one 256 MiB device-to-device copy (calibrates PMC byte counters)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not built)", file=sys.stderr)
        return 2

    if args.slab:
        print(json.dumps(slab_run(args.slab)), flush=True)
        return 0

    if args.lattice:
        print(json.dumps(lattice_run(args.lattice, args.steps, args.warmup,
                                     0.0 if args.skip_cpu else args.cpu_seconds, args.cpu_threads,
                                     args.direct_launch)), flush=True)
        if args.copy_probe:
            copy_probe(256 << 20)
        return 0

    if args.contacts:
        print(json.dumps(contact_run(max(args.steps, 1), args.warmup,
                                     0.0 if args.skip_cpu else min(args.cpu_seconds, 3.0))), flush=True)
        return 0

    if args.adjoint:
        res = adjoint_check(args.adjoint_steps, 10, args.adjoint_elements)
        print(json.dumps({"metric": "adjoint gradient vs central finite differences (max |g - fd| / max |g|)",
                          "value": max(res[k]["max_abs_err_over_max_grad"] for k in ("E", "nu", "rest")),
                          "unit": "relative", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
                          "data": "synthetic", "config": res}), flush=True)
        return 0

    from talus_amd import api

    sc, meshes, desc = workload(args.workload)
    args.streams = max(1, min(args.streams, len(meshes)))
    # a plan tuned once (talus_mi_tune) and re-applied, as a host would do,
    # so runs are reproducible and match the committed profiles of that plan
    plan_source = "fixed" if args.plan else "tuned"
    suffix = f"_s{args.streams}" if args.streams > 1 else ""
    saved = os.path.join(ROOT, "profiles", f"solve_plan_{args.workload}{suffix}.txt")
    saved_leaf = None
    if args.plan is None and not args.retune and not args.tune_leaf and not args.no_tune and os.path.exists(saved):
        args.plan, saved_leaf = read_plan(saved)
        plan_source = f"profiles/solve_plan_{args.workload}{suffix}.txt (tuned earlier, re-applied)"
    if args.leaf is None:
        args.leaf = saved_leaf

    T = sum(int(m["tet_nodes"].shape[0]) for m, _ in meshes)
    N = sum(int(m["anchored"].shape[0]) for m, _ in meshes)
    sub = sc.substeps
    params = api.default_params(substeps=sub, preconditioner=api.FACTOR,
                                flags=api.DIRECT_LAUNCH if args.direct_launch else 0)
    log(f"[bench] workload {args.workload}: {len(meshes)} bodies, {T} tets, {N} nodes")

    w0 = max(args.warmup, 1)  # the first step (factorization, plan) is never timed
    total = max(sc.steps if args.run_steps is None else args.run_steps, w0 + args.steps)
    window = [w0, w0 + args.steps]
    log(f"[bench] {total} steps per run, timed window [{window[0]}, {window[1]})")

    log("[bench] mixed batched")
    ctx_m, ids_m, wall_m, reps_m, tuned, _ = run_batch(api, meshes, api.MIXED, params, args.steps,
                                                      args.warmup, not args.no_tune, args.plan,
                                                      args.streams, leaf=args.leaf, tune_leaf=args.tune_leaf)
    leaf = ctx_m.leaf()
    value = T * sub * args.steps / wall_m
    timed_m = reps_m[-args.steps:]
    plan_steps = "".join(map(str, ctx_m.plan()))
    apps = sum(r["preconditioner_applications"] for r in timed_m) / args.steps
    log(f"[bench] mixed {1e3 * wall_m / args.steps:.3f} ms/step, {value:.4g} tet/s, "
        f"{apps:.1f} preconditioner applications/step, tuned {tuned}")
    finfo = ctx_m.factor_info()

    fp64 = None
    ctx_d = ids_d = snaps = None
    reps_d = []
    if not args.skip_fp64:
        log("[bench] fp64 batched")
        ctx_d, ids_d, wall_d, reps_d, tuned_d, snaps = run_batch(api, meshes, api.FP64, params, args.steps,
                                                                 args.warmup, not args.no_tune, None,
                                                                 args.streams, snapshot=True, leaf=leaf)
        timed_d = reps_d[-args.steps:]
        fp64 = {"value": T * sub * args.steps / wall_d, "ms_per_step": 1e3 * wall_d / args.steps,
                "cg_iterations_per_step": sum(r["cg_iterations"] for r in timed_d) / args.steps,
                "cg_capped": sum(r["cg_capped"] for r in timed_d)}
        log(f"[bench] fp64 batched {fp64['ms_per_step']:.3f} ms/step, {fp64['value']:.4g} tet/s")

    # precision, at the end of the timed window and after the whole run
    precision = {}
    if ctx_d is not None:
        err, body = force_error(ctx_m, ids_m, ctx_d, ids_d)
        drift, dbody = trajectory_drift(ctx_m, ids_m, ctx_d, ids_d, meshes)
        precision[f"step_{window[1]}"] = {"rel_force_error": err, "worst_body": body,
                                          "trajectory_drift": drift, "drift_body": dbody}
        log(f"[bench] after step {window[1]}: rel force error {err:.3e} (body {body}), "
            f"trajectory drift {drift:.3e} (body {dbody})")
    while len(reps_m) < total:
        reps_m.append(ctx_m.step(params))
    if ctx_d is not None:
        while len(reps_d) < total:
            reps_d.append(ctx_d.step(params))
        err, body = force_error(ctx_m, ids_m, ctx_d, ids_d)
        drift, dbody = trajectory_drift(ctx_m, ids_m, ctx_d, ids_d, meshes)
        precision[f"step_{total}"] = {"rel_force_error": err, "worst_body": body,
                                      "trajectory_drift": drift, "drift_body": dbody}
        log(f"[bench] after step {total}: rel force error {err:.3e} (body {body}), "
            f"trajectory drift {drift:.3e} (body {dbody})")
    rel_err = max((p["rel_force_error"] for p in precision.values()), default=None,
                  key=lambda e: e if e == e else float("inf"))
    runs = {"mixed": health(reps_m)}
    if reps_d:
        runs["fp64"] = health(reps_d)
    log(f"[bench] run health {json.dumps(runs)}")
    # the FP64 residual |right - A a| / |right| of the mixed solves (the FP32
    # PCG meets its tolerance on its own FP32 residual), measured on one more
    # step after the run, then a step with FP64 iterative refinement
    solve_precision = None
    if not args.skip_refine:
        measured = ctx_m.step(api.default_params(substeps=sub, preconditioner=api.FACTOR, refine_tolerance=1e-30,
                                                 refine_rounds=0))
        t0 = time.perf_counter()
        refined = ctx_m.step(api.default_params(substeps=sub, preconditioner=api.FACTOR,
                                                refine_tolerance=args.refine, refine_rounds=8))
        refined_ms = 1e3 * (time.perf_counter() - t0)
        solve_precision = {
            "fp64_residual": measured["refine_worst_relative"],
            "refined": {"tolerance": args.refine, "fp64_residual": refined["refine_worst_relative"],
                        "unmet": refined["refine_unmet"], "corrections": refined["refinements"],
                        "ms_per_step": refined_ms, "non_finite": refined["non_finite"]},
            "steps": [total, total + 2]}
        precision["solve"] = solve_precision
        log(f"[bench] mixed solves' FP64 residual {measured['refine_worst_relative']:.3e}; refined to "
            f"{refined['refine_worst_relative']:.3e} with {refined['refinements']} corrections, "
            f"{refined_ms:.1f} ms/step")

    fp64_plain = None
    if not args.skip_plain and snaps is not None:
        k = max(1, min(args.plain_steps, args.steps))
        log(f"[bench] fp64 plain (one context per body, sequential), steps [{window[0]}, {window[0] + k})")
        wall_p, bad_p, capped_p = plain(api, meshes, params, k, snaps)
        fp64_plain = {"value": T * sub * k / wall_p, "ms_per_step": 1e3 * wall_p / k,
                      "steps": k, "window": [window[0], window[0] + k],
                      "non_finite_steps": bad_p, "cg_capped": capped_p}
        log(f"[bench] fp64 plain {fp64_plain['ms_per_step']:.1f} ms/step, {fp64_plain['value']:.4g} tet/s")

    # per-kernel launch durations (HIP events on the context stream) and the
    # launches of each kernel per step, on the whole batch in one engine (the
    # kernels' own efficiency; the timed steps overlap S engines' kernels),
    # with the one-engine plan the committed profiles were taken with
    if args.streams > 1:
        ctx_m.set_streams(1)
        one = os.path.join(ROOT, "profiles", f"solve_plan_{args.workload}.txt")
        plan1 = None
        if os.path.exists(one) and not args.retune and not args.tune_leaf:
            plan1, leaf1 = read_plan(one)
            plan1 = plan1 if leaf1 in (None, leaf) else None
        if plan1 and len(plan1) == len(ctx_m.plan()):
            ctx_m.set_plan([int(c) for c in plan1])
        else:
            ctx_m.tune(3)
    # with the factor every solve applies the preconditioner once before its
    # first iteration and once per iteration; a step makes 3 solves per pass
    # (the kernel table below is an estimate of one engine's step)
    S = args.streams
    apps1 = apps / S  # preconditioner applications per step of one engine (estimate for S > 1)
    passes = sub
    iters = max(apps1 - 3.0 * passes, 0.0)
    kernels = {}
    for name, kind in KINDS.items():
        t = ctx_m.time_kernel(kind, 20)
        if not t["launches"]:
            continue
        per_step = {"assemble": 3 * passes, "op_element": iters + passes, "op_gather": iters + passes}
        launches = per_step.get(name, apps1 * t["launches"])
        kernels[name] = {"avg_ms": t["avg_ms"], "bytes": t["bytes"],
                         "gbs": t["bytes"] / (t["avg_ms"] * 1e-3) / 1e9,
                         "launches_per_step": launches, "ms_per_step": t["avg_ms"] * launches}
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    kd = kernels[dominant]
    roofline = {"bound": "hbm", "kernel": KERNEL_NAMES[dominant], "achieved": kd["gbs"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kd["gbs"] / HBM_PEAK_GBS,
                "traffic": traffic_for(KERNEL_NAMES[dominant], args.workload,
                                       "".join(map(str, ctx_m.plan()))),
                "algorithmic_bytes": kd["bytes"], "avg_ms": kd["avg_ms"]}
    log(f"[bench] kernels {json.dumps(kernels)}")
    # the supernodal factorization's kernels: FP64 MFMA tiles issued and the
    # flops on the fronts' own entries, per kernel over one factorization
    factor_kernels = {}
    for name, k in ctx_m.factor_profile().items():
        sec = k["ms"] * 1e-3
        entry = {"launches": k["launches"], "workgroups": k["workgroups"], "ms": k["ms"]}
        if k["mfma_flops"] > 0.0:
            entry.update(mfma_tflops=k["mfma_flops"] / sec / 1e12, useful_tflops=k["useful_flops"] / sec / 1e12,
                         mfma_frac=k["mfma_flops"] / sec / 1e12 / FP64_MFMA_PEAK_TFS,
                         useful_frac=k["useful_flops"] / sec / 1e12 / FP64_MFMA_PEAK_TFS)
        factor_kernels[name] = entry
    log(f"[bench] factor kernels {json.dumps(factor_kernels)}")

    cpu = None
    if not args.skip_cpu and snaps is not None:
        threads = max(1, min(args.cpu_threads, len(os.sched_getaffinity(0))))
        log(f"[bench] cpu baseline, {threads} threads, ~{args.cpu_seconds:.0f} s")
        cpu = cpu_baseline(sc, snaps, args.cpu_seconds, threads, params.tolerance)
        log(f"[bench] cpu {cpu['value']:.4g} tet/s")
        log(f"[bench] cpu baseline, 1 thread, ~{args.cpu_seconds_single:.0f} s")
        one = cpu_baseline(sc, snaps, args.cpu_seconds_single, 1, params.tolerance)
        cpu["single_thread"] = {"value": one["value"], "sample": one["sample"]}
        log(f"[bench] cpu 1 thread {one['value']:.4g} tet/s")

    contacts = None
    if not args.skip_contacts:
        log("[bench] contact generation (contact_cube40)")
        c = contact_run(50, 3, 0.0 if args.skip_cpu else 2.0)
        contacts = {"workload": c["config"]["workload"], "queries_per_s": c["value"],
                    "ms_per_query": c["ms_per_step"], "triangles": c["config"]["triangles"],
                    "shapes": c["config"]["shapes"], "contacts": c["config"]["contacts"],
                    "cpu_queries_per_s": c.get("cpu_baseline", {}).get("value"),
                    "gpu_over_cpu": c.get("gpu_over_cpu"), "parity": c.get("parity")}
        log(f"[bench] contacts {c['ms_per_step']:.3f} ms/query, {json.dumps(c.get('parity'))}")

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
        "dtype": "f32 elements and vectors + f64 rest inverse, node state, dots, pivots",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "bodies": len(meshes), "tets": T, "nodes": N, "substeps": sub, "precision": "mixed",
            "streams": args.streams,
            "steps_run": total, "timed_window": window,
            "cg_iterations_per_step": sum(r["cg_iterations"] for r in timed_m) / args.steps,
            "cg_capped": sum(r["cg_capped"] for r in timed_m),
            "preconditioner_applications_per_step": apps,
            "refactors": sum(r["refactors"] for r in timed_m),
            "solve_plan": plan_steps, "solve_plan_source": plan_source, "leaf": leaf,
            "kernel_table_plan": "".join(map(str, ctx_m.plan())),
            "tune": tuned,
        },
        "factor": finfo,
        "fp64": fp64,
        "fp64_plain": fp64_plain,
        "mixed_over_fp64": (value / fp64["value"]) if fp64 else None,
        "mixed_over_fp64_plain": (value / fp64_plain["value"]) if fp64_plain else None,
        "rel_force_error": rel_err,
        "precision": precision,
        "runs": runs,
        "kernels": kernels,
        "factor_kernels": factor_kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        "contacts": contacts,
    }
    # the metric holds only for a finite run with converged solves and the
    # mixed kernel within 1e-5 of FP64: otherwise the line carries no value;
    # checks skipped on request (profiling runs) leave it without a value too
    failed, unchecked = [], []
    if total < sc.steps:
        unchecked.append(f"run of {total} steps, the workload has {sc.steps}")
    for name, h in runs.items():
        if h["non_finite_steps"]:
            failed.append(f"{name}: {h['non_finite_steps']} non-finite step(s)")
        if h["cg_capped"]:
            failed.append(f"{name}: {h['cg_capped']} capped solve(s)")
    if fp64_plain and (fp64_plain["non_finite_steps"] or fp64_plain["cg_capped"]):
        failed.append("fp64 plain: non-finite steps or capped solves")
    if rel_err is None:
        unchecked.append("no FP64 run to check the force error against (--skip-fp64)")
    elif not rel_err <= FORCE_BAR:
        failed.append(f"relative force error {rel_err} above {FORCE_BAR}")
    if failed or unchecked:
        if failed:
            line["failed"] = failed
        if unchecked:
            line["unchecked"] = unchecked
        line["value_if_valid"] = line["value"]
        line["value"] = None
    print(json.dumps(line), flush=True)
    if args.copy_probe:
        copy_probe(256 << 20)
    if failed:
        log("[bench] INVALID: " + "; ".join(failed))
        return 1
    if unchecked:
        log("[bench] no metric value: " + "; ".join(unchecked))
    return 0


if __name__ == "__main__":
    sys.exit(main())
