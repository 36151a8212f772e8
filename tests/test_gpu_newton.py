"""The plastic Newton on the GPU (bodies with yielding corotated lanes,
deformable::step's plastic branch, solver.cc:3951-4250) against the oracle.

* A soft corotated cantilever that yields at its root and flows, stepped in
  FP64 with either preconditioner: every step's yield events and plastic
  counters (Newton solves, backtracks, linear reads, fallbacks, unconverged
  substeps) equal the oracle's, and positions stay within 1e-8 of the
  displacement over 30 steps.
* The same body batched with a neo-Hookean yielding body and an elastic one
  (which take the three-solve scheme in the same launches) gives every body
  the bits it gets when stepped alone.
* With plastic_passes = 1 the body takes the elastic scheme, as the
  reference does when plasticPasses <= 1.
"""
from __future__ import annotations

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import oracle_body, product_body

pytestmark = pytest.mark.gpu

COUNTERS = ("plastic_solves", "plastic_backtracks", "plastic_linear_reads", "plastic_handoffs",
            "plastic_fallbacks", "plastic_rejections", "plastic_capped")


def _corotated(cells=(8, 2, 2), yield_stress=500.0, hardening=1e4):
    mats = scenes.material_array(1, youngs_modulus=1e6, poissons_ratio=0.3, density=1000.0,
                                 elastic_model=api.COROTATED, yield_stress=yield_stress,
                                 hardening_modulus=hardening)
    return scenes.Body(resolution=tuple(c + 1 for c in cells), voxel=(0.1, 0.05, 0.05),
                       supports=[((0.0, 0.0, 0.0), (0.01, 1e3, 1e3))], materials=mats)


@pytest.mark.parametrize("preconditioner", [api.BLOCK_JACOBI, api.FACTOR])
def test_newton_matches_the_oracle_step_by_step(preconditioner):
    body = _corotated()
    orc, om = oracle_body(body)
    steps = 30
    params = api.default_params(tolerance=1e-10, cg_iterations=5000, factor_budget=5000,
                                preconditioner=preconditioner, flags=api.RECORD_EVENTS)
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        totals = dict.fromkeys(COUNTERS, 0)
        for s in range(steps):
            ro = orc.step(tolerance=1e-10, cg_iterations=5000)
            r = ctx.step(params)
            for k in COUNTERS:
                assert r[k] == ro[k], (s, k, r[k], ro[k])
                totals[k] += r[k]
            ev = ctx.events()
            got = np.array(sorted(e[2] for e in ev if e[0] == api.EVENT_YIELD))
            assert np.array_equal(got, np.flatnonzero(orc.yield_flags())), s
        x, v, _ = ctx.state(b)
        plastic = ctx.plastic(b)
    xo, vo, _ = orc.state()
    disp = np.abs(xo - om["rest"]).max()
    assert disp > 0.05
    assert np.abs(x - xo).max() <= 1e-8 * disp
    assert np.abs(v - vo).max() <= 1e-7 * np.abs(vo).max()
    assert np.abs(plastic - orc.plastic()).max() <= 1e-8 * max(np.abs(orc.plastic()).max(), 1e-3)
    assert totals["plastic_solves"] > 3 * 2 * steps and totals["plastic_capped"] == 0


def _batch():
    snh = scenes.material_array(1, youngs_modulus=2e5, poissons_ratio=0.3, density=1000.0,
                                elastic_model=api.NEO_HOOKEAN, yield_stress=300.0, hardening_modulus=2e3)
    ela = scenes.material_array(1, youngs_modulus=3e6, poissons_ratio=0.35, density=1000.0,
                                elastic_model=api.NEO_HOOKEAN)
    sup = [((0.0, 0.0, 0.0), (0.01, 1e3, 1e3))]
    return [_corotated(),
            scenes.Body(resolution=(7, 3, 3), voxel=(0.05, 0.05, 0.05), supports=sup, materials=snh),
            _corotated(cells=(5, 3, 2), yield_stress=2e3),
            scenes.Body(resolution=(5, 3, 3), voxel=(0.05, 0.05, 0.05), supports=sup, materials=ela)]


@pytest.mark.parametrize("precision", [api.FP64, api.MIXED])
def test_batched_newton_bodies_equal_each_body_alone(precision):
    bodies = _batch()
    params = api.default_params(substeps=2, preconditioner=api.FACTOR, tolerance=1e-6,
                                cg_iterations=500, factor_budget=500)
    alone = []
    for body in bodies:
        with api.Context(0, precision) as ctx:
            b, _ = product_body(ctx, body)
            reps = [ctx.step(params) for _ in range(6)]
            alone.append((ctx.state(b)[0], sum(r["plastic_solves"] for r in reps)))
    with api.Context(0, precision) as ctx:
        ids = [product_body(ctx, body)[0] for body in bodies]
        reps = [ctx.step(params) for _ in range(6)]
        for i, b in enumerate(ids):
            assert np.array_equal(ctx.state(b)[0], alone[i][0]), i
    assert sum(r["plastic_solves"] for r in reps) == sum(a[1] for a in alone) > 0


def test_one_plastic_pass_takes_the_elastic_scheme():
    body = _corotated()
    orc, _ = oracle_body(body)
    ro = orc.step(tolerance=1e-10, cg_iterations=5000, plastic_passes=1)
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        r = ctx.step(tolerance=1e-10, cg_iterations=5000, plastic_passes=1)
        x, _, _ = ctx.state(b)
    assert r["plastic_solves"] == ro["plastic_solves"] == 0 and r["solves"] == 3
    assert np.abs(x - orc.state()[0]).max() <= 1e-10


@pytest.mark.parametrize("precision", [api.FP64, api.MIXED])
def test_refinement_beside_newton_bodies(precision):
    """FP64 refinement refines the three-solve bodies' solves and leaves the
    Newton bodies' own solves alone: batched bodies still equal each body
    stepped alone, and the yielding corotated body's state is the one it
    reaches without refinement."""
    bodies = _batch()
    base = dict(substeps=2, preconditioner=api.FACTOR, tolerance=1e-6, cg_iterations=500, factor_budget=500)
    refined = api.default_params(refine_tolerance=1e-10, **base)
    plain = api.default_params(**base)
    alone = []
    for body in bodies:
        with api.Context(0, precision) as ctx:
            b, _ = product_body(ctx, body)
            reps = [ctx.step(refined) for _ in range(4)]
            alone.append(ctx.state(b)[0])
    with api.Context(0, precision) as ctx:
        ids = [product_body(ctx, body)[0] for body in bodies]
        reps = [ctx.step(refined) for _ in range(4)]
        for i, b in enumerate(ids):
            assert np.array_equal(ctx.state(b)[0], alone[i]), i
        assert all(r["refine_unmet"] == 0 for r in reps)
        assert sum(r["refinements"] for r in reps) > 0 and sum(r["plastic_solves"] for r in reps) > 0
    with api.Context(0, precision) as ctx:
        b, _ = product_body(ctx, bodies[0])
        for _ in range(4):
            ctx.step(plain)
        assert np.array_equal(ctx.state(b)[0], alone[0])
