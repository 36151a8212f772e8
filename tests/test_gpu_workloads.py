"""The benchmark workloads on the GPU against the oracle and against what the
mathematics fixes.

* A reduced batch of the 512-body workload (13 of its bodies, soft and stiff,
  yielding and elastic, one of them a body that poured and once diverged)
  stepped together in one context with the block Cholesky preconditioner,
  against the oracle stepping each body alone: per-step yield events equal
  and positions within 1e-7 of the displacement for 30 steps (4 substeps).
* predictorInverts (solver.cc:568-623, 3574-3583): a predictor that inverts
  a tet raises the step to 4 substeps, on the GPU as in the oracle, and the
  step lands where the oracle's does.
* The jittered, per-tet lognormal cube: 12^3 hexes for 60 steps against the
  oracle at tight tolerance, and the full 40^3 cube (384,000 tets) for its
  240 steps in FP64 and mixed precision: finite, no capped solve, sagging as
  a supported elastic block does, mixed within 1e-4 of FP64.
"""
from __future__ import annotations

import concurrent.futures as cf

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import oracle_body, product_body

pytestmark = pytest.mark.gpu

BATCH = [6, 120, 159, 179, 218, 221, 32, 424, 74, 24, 235, 96, 55]


def test_workload_batch_matches_the_oracle():
    sc = scenes.bodies()
    bodies = [sc.bodies[i] for i in BATCH]
    assert sum(np.isfinite(b.materials["yield_stress"][0]) for b in bodies) >= 6
    steps = 30
    oracles = [oracle_body(b) for b in bodies]

    def run(pair):
        orc, _ = pair
        events = []
        for _ in range(steps):
            orc.step(substeps=4, tolerance=1e-10, cg_iterations=100000)
            events.append(np.flatnonzero(orc.yield_flags()))
        return events

    with cf.ThreadPoolExecutor(8) as ex:
        oracle_events = list(ex.map(run, oracles))
    params = api.default_params(substeps=4, preconditioner=api.FACTOR, tolerance=1e-10,
                                cg_iterations=4000, factor_budget=4000, flags=api.RECORD_EVENTS)
    with api.Context(0, api.FP64) as ctx:
        ids = [product_body(ctx, b)[0] for b in bodies]
        reports = [ctx.step(params) for _ in range(steps)]
        events = ctx.events()
        states = [ctx.state(b) for b in ids]
    assert sum(r["cg_capped"] for r in reports) == 0
    assert not any(r["non_finite"] for r in reports)
    yielded = 0
    for k, ((orc, om), b) in enumerate(zip(oracles, ids)):
        for s in range(steps):
            got = np.array(sorted(e[2] for e in events if e[0] == api.EVENT_YIELD and e[1] == b and e[3] == s))
            assert np.array_equal(got, oracle_events[k][s]), (BATCH[k], s)
            yielded += got.size
        xo, vo, _ = orc.state()
        x, v, _ = states[k]
        disp = np.abs(xo - om["rest"]).max()
        assert np.abs(x - xo).max() <= 1e-7 * disp, BATCH[k]
        # a stiff body at rest has |v| ~ 1e-13 m/s: 1e-12 m/s absolute floor
        assert np.abs(v - vo).max() <= 1e-6 * max(np.abs(vo).max(), 1e-6), BATCH[k]
    assert yielded > 0


def _tiny_body(model=api.NEO_HOOKEAN):
    mats = scenes.material_array(1, youngs_modulus=1e5, poissons_ratio=0.3, density=1000.0,
                                 elastic_model=model)
    return scenes.Body(resolution=(3, 2, 2), voxel=(0.05, 0.05, 0.05),
                       supports=[((0.0, 0.0, 0.0), (0.01, 1e3, 1e3))], materials=mats)


@pytest.mark.parametrize("model", [api.NEO_HOOKEAN, api.COROTATED])
def test_inverting_predictor_takes_four_substeps_like_the_oracle(model):
    body = _tiny_body(model)
    orc, om = oracle_body(body)
    rest = om["rest"]
    v = np.zeros_like(rest)
    tip = rest[:, 0] > 0.09
    v[tip, 0] = -20.0  # the tip face travels 0.22 m in the predictor, past the middle layer
    orc.set_state(rest, v, np.zeros_like(rest))
    r_o = orc.step(tolerance=1e-12, cg_iterations=5000)
    assert r_o["substeps"] == 4
    xo, vo, _ = orc.state()
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        ctx.set_state(b, rest, v, np.zeros_like(rest))
        r = ctx.step(tolerance=1e-12, cg_iterations=5000)
        x, vg, _ = ctx.state(b)
        # the same body at rest keeps the configured substep
        ctx.set_state(b, rest, np.zeros_like(rest), np.zeros_like(rest))
        assert ctx.step(tolerance=1e-12, cg_iterations=5000)["substeps"] == 1
    assert r["substeps"] == 4
    disp = np.abs(xo - rest).max()
    assert np.abs(x - xo).max() <= 1e-8 * disp
    assert np.abs(vg - vo).max() <= 1e-7 * np.abs(vo).max()


def test_cube12_trajectory_matches_the_oracle():
    body = scenes.cube(12, seed=2024).bodies[0]
    orc, om = oracle_body(body)
    steps = 60
    for _ in range(steps):
        orc.step(tolerance=1e-10, cg_iterations=100000, threads=8)
    xo, vo, _ = orc.state()
    params = api.default_params(preconditioner=api.FACTOR, tolerance=1e-10, cg_iterations=4000,
                                factor_budget=4000)
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        reports = [ctx.step(params) for _ in range(steps)]
        x, v, _ = ctx.state(b)
    assert sum(r["cg_capped"] for r in reports) == 0
    disp = np.abs(xo - om["rest"]).max()
    assert disp > 1e-4
    assert np.abs(x - xo).max() <= 1e-8 * disp
    assert np.abs(v - vo).max() <= 1e-6 * np.abs(vo).max()


def test_cube40_runs_its_240_steps():
    """The 384,000-tet cube for its 240 steps with the default step settings
    (PCG tolerance 1e-3, block Cholesky): finite, no capped solve, the top
    face sagging by the order of rho g H^2 / E_median (a few mm) and at rest
    at the end (generalized-alpha with rho_inf = 0.5 damps the oscillation),
    the mixed path within 1e-4 of FP64 in positions."""
    body = scenes.cube(40).bodies[0]
    mesh, mats = scenes.build_mesh(body)
    rest = mesh["rest_positions"].astype(np.float64)
    top = rest[:, 1] > 1.0 - 1e-3
    params = api.default_params(preconditioner=api.FACTOR)
    out = {}
    for precision in (api.FP64, api.MIXED):
        with api.Context(0, precision) as ctx:
            b = ctx.add_body(mesh, mats)
            reports = [ctx.step(params) for _ in range(240)]
            x, v, _ = ctx.state(b)
        assert sum(r["cg_capped"] for r in reports) == 0, precision
        assert not any(r["non_finite"] for r in reports), precision
        out[precision] = (x, v, reports)
    x, v, reports = out[api.FP64]
    sag = -(x[top, 1] - rest[top, 1]).mean()
    assert 1e-4 < sag < 1e-2, sag
    assert np.abs(v).max() < 0.05 * sag * 60.0  # nearly at rest after 4 s
    xm = out[api.MIXED][0]
    assert np.abs(xm - x).max() <= 1e-4 * np.abs(x - rest).max()
