"""The oracle's plastic Newton (deformable::step for bodies with yielding
corotated lanes, solver.cc:3951-4250) and the consistent J2 tangent it
solves with (tangentOf, material.cc:636-647; tangent blocks 1033-1054).

Pinned by what the mathematics fixes, not by retyping the formulas:
* the tangent is the derivative of the return-mapped stress (central
  differences of the return map itself), at a loading point and in the
  elastic range; a dropped thetaBar term, a wrong sign or the elastic
  tangent for a loading point fails it;
* the body's tangent operator is the derivative of the return-mapped node
  forces, under uniform symmetric strains (where the corotated frame stays
  the identity, so the rotation derivative the corotated tangent omits is
  zero);
* a Newton step that yields is a root of the substep's nonlinear residual
  r(a) = f(x_w + h_eff^2 a) + m g - alpha_m m a_n - (1 - alpha_m) m a with
  the return map from the committed state, far closer to one than the
  elastic three-solve scheme's answer, and the step report's counters add
  up.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from talus_amd import api, scenes

from helpers import oracle_body


def _sym(rng, scale):
    A = scale * rng.standard_normal((3, 3))
    return 0.5 * (A + A.T)


@pytest.mark.parametrize("hardening", [0.0, 2e4])
def test_j2_tangent_is_the_derivative_of_the_return_map(hardening):
    lam, mu = O.lame(1e6, 0.3)
    rng = np.random.default_rng(3)
    plastic = np.zeros(10)
    plastic[:9] = _sym(rng, 2e-3).reshape(9)
    plastic[9] = 0.01
    for trial in range(6):
        strain = _sym(rng, 2e-2)
        sigma, lin, returned = O.j2_point(strain, lam, mu, 1e3, hardening, plastic)
        assert returned  # 2e-2 strain at E = 1e6 is far past 1e3 Pa
        assert 0.0 < lin[0] < 1.0 and lin[1] > 0.0
        for _ in range(3):
            e = _sym(rng, 1.0)
            h = 1e-7
            fd = (O.j2_point(strain + h * e, lam, mu, 1e3, hardening, plastic)[0] -
                  O.j2_point(strain - h * e, lam, mu, 1e3, hardening, plastic)[0]) / (2 * h)
            tan = O.j2_tangent(lin, lam, mu, e)
            assert np.abs(tan - fd).max() <= 1e-6 * np.abs(fd).max()
            # the elastic tangent differs from the derivative at a loading point
            elastic = lam * np.trace(e) * np.eye(3) + 2 * mu * e
            assert np.abs(elastic - fd).max() > 1e-2 * np.abs(fd).max()
    # in the elastic range the tangent is Hooke's law
    small = _sym(rng, 1e-5)
    _, lin, returned = O.j2_point(small, lam, mu, 1e3, hardening)
    assert not returned and lin[0] == 1.0 and lin[1] == 0.0
    e = _sym(rng, 1.0)
    assert np.allclose(O.j2_tangent(lin, lam, mu, e), lam * np.trace(e) * np.eye(3) + 2 * mu * e,
                       rtol=1e-12, atol=1e-12 * mu)


def _plastic_body(yield_stress=500.0, hardening=1e4, anchored=True, cells=(3, 2, 2)):
    mats = scenes.material_array(1, youngs_modulus=1e6, poissons_ratio=0.3, density=1000.0,
                                 elastic_model=api.COROTATED, yield_stress=yield_stress,
                                 hardening_modulus=hardening)
    sups = [((0.0, 0.0, 0.0), (0.01, 1e3, 1e3))] if anchored else []
    return scenes.Body(resolution=tuple(c + 1 for c in cells), voxel=(0.1, 0.05, 0.05), supports=sups,
                       materials=mats)


def test_tangent_operator_is_the_derivative_of_the_return_mapped_forces():
    orc, m = oracle_body(_plastic_body(anchored=False))
    X = m["rest"]
    rng = np.random.default_rng(8)
    S = np.eye(3) + _sym(rng, 1e-2)
    x = X @ S.T
    E = _sym(rng, 1.0)
    d = X @ E.T
    y, loading = orc.tangent_apply(x, d)
    assert loading == m["tet"].shape[0]  # every tet past the yield stress
    h = 1e-7
    fp = orc.eval_forces(x + h * d)[2]
    fm = orc.eval_forces(x - h * d)[2]
    fd = -(fp - fm) / (2 * h)
    assert np.abs(y - fd).max() <= 1e-5 * np.abs(fd).max()
    # the elastic operator at the same state is not the derivative
    ye = orc.apply_system(0.0, 1.0, d)
    assert np.abs(ye - fd).max() > 1e-2 * np.abs(fd).max()


def _residual(orc, m, x0, v0, a0, a, plastic0, dt=1.0 / 60.0, g=(0.0, -9.81, 0.0)):
    """r(a) of one substep of the generalized-alpha rule (rho_inf = 0.5) with
    the return map from the committed plastic state plastic0."""
    s = 0.5
    am, af = (2 * s - 1) / (s + 1), s / (s + 1)
    gamma = 0.5 - am + af
    beta = 0.25 * (1 - am + af) ** 2
    carry = (0.5 - beta) * dt * dt
    weighted = (1 - af) * (x0 + dt * v0 + carry * a0) + af * x0
    reach = (1 - af) * beta * dt * dt
    orc.set_plastic(plastic0)
    f = orc.eval_forces(weighted + reach * a)[2]
    mass = m["mass"][:, None]
    r = f + mass * np.asarray(g) - am * mass * a0 - (1 - am) * mass * a
    r[m["anchored"] != 0] = 0.0
    return r


def test_newton_step_is_a_root_of_the_plastic_residual():
    body = _plastic_body()
    out = {}
    for passes in (40, 1):
        orc, m = oracle_body(body)
        n = m["rest"].shape[0]
        x0, v0, a0 = m["rest"].copy(), np.zeros((n, 3)), np.zeros((n, 3))
        rep = orc.step(tolerance=1e-12, cg_iterations=20000, plastic_passes=passes)
        _, _, a = orc.state()
        p0 = np.zeros((m["tet"].shape[0], 10))
        r = _residual(orc, m, x0, v0, a0, a, p0)
        r0 = _residual(orc, m, x0, v0, a0, np.zeros_like(a), p0)
        out[passes] = (np.linalg.norm(r) / np.linalg.norm(r0), rep)
    newton, rep = out[40]
    elastic, rep1 = out[1]
    assert rep["yield_events"] > 0 and rep1["plastic_solves"] == 0
    assert rep["plastic_capped"] == 0 and rep["plastic_solves"] >= 3
    assert rep["solves"] == rep["plastic_solves"]
    assert newton <= 1e-6, newton
    assert elastic >= 1e3 * newton, (elastic, newton)


def test_newton_steps_a_yielding_cantilever_without_capping():
    """40 steps of a soft corotated cantilever that yields at its root and
    flows: every substep converges (no capped Newton, no rejection) and the
    body stays finite."""
    orc, m = oracle_body(_plastic_body(cells=(8, 2, 2)))
    totals = {}
    for _ in range(40):
        rep = orc.step(tolerance=1e-8, cg_iterations=5000)
        assert not rep["non_finite"]
        for k, v in rep.items():
            if k.startswith("plastic_") or k == "yield_events":
                totals[k] = totals.get(k, 0) + v
    assert totals["yield_events"] > 0
    assert totals["plastic_capped"] == 0 and totals["plastic_rejections"] == 0


def _corotated_strip():
    mats = scenes.material_array(1, youngs_modulus=1e6, poissons_ratio=0.3, density=1000.0,
                                 elastic_model=api.COROTATED)
    return scenes.Body(resolution=(4, 2, 2), voxel=(0.1, 0.05, 0.05), materials=mats)


def test_initial_stress_is_the_symmetric_rotation_derivative_of_the_forces():
    """initialStress (material.cc:1077-1174): the corotated forces f depend on
    x through the rotation too; -df/dx = R K R^T + A, and the initial-stress
    operator is the symmetric part of A (central differences of the forces,
    minus the operator with the rotation held).  It vanishes at rest and
    under a pure rotation (no force, no offset)."""
    from helpers import deformed

    orc, m = oracle_body(_corotated_strip())
    X = m["rest"]
    x = deformed(X, seed=3, scale=0.02, twist=0.7)
    rng = np.random.default_rng(2)
    d1, d2 = rng.standard_normal((2,) + X.shape)
    h = 1e-6

    def jd(d):
        return -(orc.eval_forces(x + h * d)[2] - orc.eval_forces(x - h * d)[2]) / (2 * h)

    j1, j2 = jd(d1), jd(d2)
    orc.eval_forces(x)
    k1, k2 = orc.apply_system(0.0, 1.0, d1), orc.apply_system(0.0, 1.0, d2)
    expect = 0.5 * ((d2 * (j1 - k1)).sum() + (d1 * (j2 - k2)).sum())
    s1 = orc.stressed_apply(x, d1) - k1
    s2 = orc.stressed_apply(x, d2) - k2
    got = (d2 * s1).sum()
    assert abs(got - (d1 * s2).sum()) <= 1e-9 * abs(got)  # symmetric
    assert abs(got - expect) <= 1e-4 * abs(expect), (got, expect)
    assert abs(expect) > 1e-4 * abs((d2 * k1).sum())  # the term is not negligible here
    c = X.mean(axis=0)
    Rz = np.array([[np.cos(0.9), -np.sin(0.9), 0.0], [np.sin(0.9), np.cos(0.9), 0.0], [0.0, 0.0, 1.0]])
    for pos in (X, (X - c) @ Rz.T + c):
        orc.eval_forces(pos)
        k = orc.apply_system(0.0, 1.0, d1)
        assert np.abs(orc.stressed_apply(pos, d1) - k).max() <= 1e-9 * np.abs(k).max()
