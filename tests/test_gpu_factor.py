"""GPU block Cholesky preconditioner (SystemFactor replacement) against a
dense numpy solve of the system the oracle assembles column by column
(textbook Cholesky pin: residual of the solve), and FACTOR-preconditioned
steps against the oracle's block-Jacobi steps at tight tolerance (the
preconditioner changes the iterates, never the converged solution)."""
import math

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import deformed, oracle_body, product_body

pytestmark = pytest.mark.gpu


def dense_system(orc, m, inertia, step):
    n = m["rest"].shape[0]
    free = np.repeat(m["anchored"] == 0, 3)
    A = np.zeros((3 * n, 3 * n))
    for j in np.flatnonzero(free):
        e = np.zeros(3 * n)
        e[j] = 1.0
        A[:, j] = orc.apply_system(inertia, step, e).reshape(-1)
    return A[np.ix_(free, free)], free


def _check_solve(body, x=None, bar=1e-10, precision=api.FP64, plan=None, leaf=None):
    orc, m = oracle_body(body)
    if x is not None:
        orc.eval_forces(x)  # commits rotations and volume frames
    R, W = orc.rotations(), orc.frames()
    inertia, step = 1.0, math.sqrt(8.0 / 27.0) / 60.0
    A, free = dense_system(orc, m, inertia, step)
    rng = np.random.default_rng(21)
    r = rng.standard_normal(m["rest"].shape)
    r[m["anchored"] != 0] = 0.0
    z_ref = np.zeros(r.size)
    z_ref[free] = np.linalg.solve(A, r.reshape(-1)[free])
    with api.Context(0, precision) as ctx:
        b, _ = product_body(ctx, body)
        if leaf is not None:
            ctx.set_leaf(leaf)
            assert ctx.leaf() == leaf
        ctx.set_rotations(b, R)
        ctx.set_frames(b, W)
        if plan is not None:
            ctx.set_plan(np.full(ctx.plan().size, plan, dtype=np.uint8))
        z = ctx.factor_solve(b, inertia, step, r).reshape(-1)
        info = ctx.factor_info()
    assert info["ok"] == 1 and info["fronts"] > 0
    err = np.abs(z - z_ref).max() / np.abs(z_ref).max()
    assert err <= bar, err
    assert np.all(z[~free] == 0.0)
    return info


def test_factor_solve_cantilever_rest():
    info = _check_solve(scenes.cantilever().bodies[0])
    assert info["levels"] >= 1


def test_factor_solve_rotated_state():
    body = scenes.cantilever().bodies[0]
    _, m = oracle_body(body)
    _check_solve(body, x=deformed(m["rest"], seed=3, scale=0.002, twist=0.8))


def test_factor_solve_jittered_cube_with_many_fronts():
    body = scenes.cube(9, seed=6).bodies[0]
    info = _check_solve(body)
    assert info["fronts"] > 8 and info["levels"] >= 3


@pytest.mark.parametrize("plan", [0, 1, 2])
def test_factor_solve_either_kernel_shape(plan):
    """Every solve kernel shape (rows per wavefront, one row/column per lane,
    64 rows/columns per workgroup split over its wavefronts) gives the dense
    solution on every level."""
    body = scenes.cube(9, seed=6).bodies[0]
    _check_solve(body, plan=plan)


@pytest.mark.parametrize("leaf", [8, 16, 32, 128])
def test_factor_solve_any_leaf_size(leaf):
    """The supernode leaf size changes the fronts, never the solution."""
    body = scenes.cube(9, seed=6).bodies[0]
    info = _check_solve(body, leaf=leaf)
    ref = _check_solve(body)
    assert (info["fronts"] > ref["fronts"]) == (leaf < 64)


def test_leaf_tuner_picks_a_candidate_and_its_plan_reproduces():
    sc = scenes.bodies(count=6, seed=13, lo=300, hi=3000)
    meshes = [scenes.build_mesh(b) for b in sc.bodies]
    runs = []
    chosen = None
    for _ in range(2):
        with api.Context(0, api.MIXED) as ctx:
            ids = [ctx.add_body(m, mats) for m, mats in meshes]
            if chosen is None:
                rep = ctx.tune_leaf(candidates=(16, 32, 64), repeats=2)
                assert rep["leaf"] in (16, 32, 64) and ctx.leaf() == rep["leaf"]
                assert all(t > 0.0 for t in rep["candidates"].values())
                assert rep["after_seconds"] <= min(rep["candidates"].values()) * 1.5
                chosen = (rep["leaf"], ctx.plan())
                assert chosen[1].size == 2 * rep["levels"]
            else:
                ctx.set_leaf(chosen[0])
                ctx.set_plan(chosen[1])
                assert np.array_equal(ctx.plan(), chosen[1])
            for _ in range(3):
                r = ctx.step(substeps=2, preconditioner=api.FACTOR)
                assert not r["non_finite"] and r["cg_capped"] == 0
            runs.append([ctx.state(b)[0] for b in ids])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_set_leaf_keeps_the_state():
    sc = scenes.bodies(count=3, seed=5, lo=300, hi=2000)
    meshes = [scenes.build_mesh(b) for b in sc.bodies]
    with api.Context(0, api.FP64) as ctx:
        ids = [ctx.add_body(m, mats) for m, mats in meshes]
        for _ in range(2):
            ctx.step(substeps=1, preconditioner=api.FACTOR)
        before = [ctx.state(b) for b in ids]
        ctx.set_leaf(24)
        after = [ctx.state(b) for b in ids]
        for p, q in zip(before, after):
            for u, v in zip(p, q):
                assert np.array_equal(u, v)
        r = ctx.step(substeps=1, preconditioner=api.FACTOR)
        assert not r["non_finite"] and r["cg_capped"] == 0
        with pytest.raises(api.TalusError):
            ctx.set_leaf(4)


def test_tuned_plan_is_valid_and_reproducible():
    sc = scenes.bodies(count=6, seed=13, lo=300, hi=3000)
    meshes = [scenes.build_mesh(b) for b in sc.bodies]
    runs = []
    plan = None
    for _ in range(2):
        with api.Context(0, api.MIXED) as ctx:
            ids = [ctx.add_body(m, mats) for m, mats in meshes]
            if plan is None:
                rep = ctx.tune(2)
                assert rep["levels"] > 0 and rep["after_seconds"] > 0.0
                plan = ctx.plan()
                assert plan.size == 2 * rep["levels"] and set(np.unique(plan)) <= {0, 1, 2}
            else:
                ctx.set_plan(plan)
                assert np.array_equal(ctx.plan(), plan)
            for _ in range(3):
                r = ctx.step(substeps=2)
                assert not r["non_finite"]
            runs.append([ctx.state(b)[0] for b in ids])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_factor_solve_mixed_storage_is_a_good_preconditioner():
    # the mixed path stores G in FP16 (unit roundoff 2^-11 = 4.9e-4) with
    # exact power-of-two column scales and accumulates in FP64: the solve is
    # within a few roundoffs of the dense solve
    _check_solve(scenes.cube(6, seed=8).bodies[0], bar=1e-3, precision=api.MIXED)
    _check_solve(scenes.cantilever().bodies[0], bar=1e-3, precision=api.MIXED)


def test_factor_solve_batched_equals_single():
    sc = scenes.bodies(count=3, seed=12, lo=300, hi=3000)
    r_list, z_single = [], []
    for body in sc.bodies:
        with api.Context(0, api.FP64) as ctx:
            b, mesh = product_body(ctx, body)
            rng = np.random.default_rng(len(r_list))
            r = rng.standard_normal(mesh["rest_positions"].shape)
            z_single.append(ctx.factor_solve(b, 1.0, 0.005, r))
            r_list.append(r)
    with api.Context(0, api.FP64) as ctx:
        ids = [product_body(ctx, body)[0] for body in sc.bodies]
        for i, b in enumerate(ids):
            z = ctx.factor_solve(b, 1.0, 0.005, r_list[i])
            assert np.array_equal(z, z_single[i])


def test_factor_preconditioned_steps_match_oracle():
    body = scenes.cantilever().bodies[0]
    orc, m = oracle_body(body)
    steps = 20
    for _ in range(steps):
        orc.step(tolerance=1e-10, cg_iterations=4000)
    xo, vo, _ = orc.state()
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        reps = [ctx.step(tolerance=1e-10, cg_iterations=200, factor_budget=200,
                         preconditioner=api.FACTOR) for _ in range(steps)]
        x, v, _ = ctx.state(b)
    disp = np.abs(xo - m["rest"]).max()
    assert np.abs(x - xo).max() <= 1e-8 * disp
    assert sum(r["cg_capped"] for r in reps) == 0
    # the factor of the exact first-step system converges almost at once
    assert reps[0]["cg_iterations"] <= 3 * 12


def test_factor_converges_where_block_jacobi_caps():
    body = scenes.cube(12, seed=13).bodies[0]
    with api.Context(0, api.MIXED) as ctx:
        product_body(ctx, body)
        bj = [ctx.step() for _ in range(3)]
    with api.Context(0, api.MIXED) as ctx:
        product_body(ctx, body)
        fac = [ctx.step(preconditioner=api.FACTOR) for _ in range(3)]
        info = ctx.factor_info()
    assert sum(r["cg_capped"] for r in bj) > 0
    assert sum(r["cg_capped"] for r in fac) == 0
    assert all(r["cg_worst_relative"] <= 1e-3 for r in fac)
    assert info["ok"] == 1 and info["factor_seconds"] > 0


def test_factor_steps_are_bit_identical():
    body = scenes.cube(8, seed=14).bodies[0]
    outs = []
    for _ in range(2):
        with api.Context(0, api.MIXED) as ctx:
            b, _ = product_body(ctx, body)
            for _ in range(3):
                ctx.step(preconditioner=api.FACTOR)
            outs.append(ctx.state(b)[0])
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("precision", [api.FP64, api.MIXED])
def test_graph_replay_equals_direct_launch(precision):
    """The captured HIP graph of a PCG iteration runs the same kernels as the
    direct launches: trajectories are bit-identical, block Jacobi and factor,
    across several graph replays with host work in between."""
    body = scenes.cube(6, seed=12).bodies[0]
    for pre in (api.BLOCK_JACOBI, api.FACTOR):
        runs = []
        for flags in (0, api.DIRECT_LAUNCH):
            with api.Context(0, precision) as ctx:
                b, _ = product_body(ctx, body)
                for _ in range(4):
                    r = ctx.step(preconditioner=pre, flags=flags)
                    assert not r["non_finite"]
                    ctx.state(b)  # host synchronisation between replays
                runs.append(ctx.state(b))
        for a, c in zip(*runs):
            assert np.array_equal(a, c)


def test_partial_refresh_keeps_batched_equal_to_single_bodies():
    """Bodies whose solves slow down get their own fronts refactored (the
    other bodies' fronts are skipped); a body's trajectory is the same
    whether it steps alone or in the batch."""
    sc = scenes.bodies(count=5, seed=21, lo=300, hi=3000)
    for b in sc.bodies:
        b.materials["youngs_modulus"] = 3e4  # soft: large rotations, drifting factors
    p = api.default_params(substeps=2, preconditioner=api.FACTOR)
    singles = []
    for body in sc.bodies:
        with api.Context(0, api.MIXED) as ctx:
            i, _ = product_body(ctx, body)
            for _ in range(12):
                ctx.step(p)
            singles.append(ctx.state(i)[0])
    with api.Context(0, api.MIXED) as ctx:
        ids = [product_body(ctx, body)[0] for body in sc.bodies]
        refactors = sum(ctx.step(p)["refactors"] for _ in range(12))
        for i, x in zip(ids, singles):
            assert np.array_equal(ctx.state(i)[0], x)
    assert refactors > 1  # the first factorization and at least one partial refresh


def test_streams_split_the_batch_without_changing_any_body():
    """Bodies split over three concurrently stepped engines give every body
    the same state, events and report sums as one engine, with the same
    solve-kernel shape on every level (lanes)."""
    sc = scenes.bodies(count=7, seed=21, lo=300, hi=4000)
    params = api.default_params(substeps=2, preconditioner=api.FACTOR, flags=api.RECORD_EVENTS)
    runs = []
    for streams in (1, 3):
        with api.Context(0, api.MIXED) as ctx:
            ids = [product_body(ctx, body)[0] for body in sc.bodies]
            ctx.set_streams(streams)
            ctx.set_plan(np.ones(ctx.plan().size, dtype=np.uint8))
            reps = [ctx.step(params) for _ in range(3)]
            runs.append(([ctx.state(b) for b in ids], reps, ctx.events(), ctx.factor_info()))
    (s1, r1, e1, f1), (s3, r3, e3, f3) = runs
    for a, b in zip(s1, s3):
        for u, w in zip(a, b):
            assert np.array_equal(u, w)
    for a, b in zip(r1, r3):
        for key in ("cg_iterations", "cg_capped", "solves", "yield_events", "bodies", "nodes", "elements"):
            assert a[key] == b[key], key
    assert e1 == e3
    assert f1["fronts"] == f3["fronts"] and f1["entries"] == f3["entries"]


def test_large_fronts_solve_the_system_with_every_plan():
    """A 14^3 cube has separators with more than 512 pivots, so the lane and
    split plans also run the rows shape (with its pre-gathered front vectors)
    on part of a level: every plan solves A z = r (A applied matrix-free by
    the operator) to 1e-9 in FP64."""
    body = scenes.cube(14, seed=5).bodies[0]
    inertia, step = 1.0, math.sqrt(8.0 / 27.0) / 60.0
    with api.Context(0, api.FP64) as ctx:
        b, mesh = product_body(ctx, body)
        free = np.repeat(mesh["anchored"] == 0, 3)
        r = np.random.default_rng(3).standard_normal(mesh["rest_positions"].shape)
        r[mesh["anchored"] != 0] = 0.0
        levels = ctx.plan().size
        for plan in (0, 1, 2):
            ctx.set_plan(np.full(levels, plan, dtype=np.uint8))
            z = ctx.factor_solve(b, inertia, step, r).reshape(-1)
            Az = ctx.apply_system(b, inertia, step, z).reshape(-1)
            err = np.abs(Az[free] - r.reshape(-1)[free]).max() / np.abs(r).max()
            assert err <= 1e-9, (plan, err)
        assert ctx.factor_info()["ok"] == 1
