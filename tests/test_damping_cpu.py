"""Stiffness damping in the oracle, pinned by what the mathematics fixes.

The reference's damping (solver/solver.cc:625-905, 3586-3592, 3726-3783;
factor/factor.cc:755-857) adds, per element with stiffnessDamping tau > 0,
the force tau P R K R^T P v: K the element's elastic stiffness, R its
committed rotation and P the removal of the element's rigid modes fitted at
the assembly point.  These tests check

  * rigid velocities produce no damping force (P removes them),
  * the damping operator is symmetric positive semi-definite,
  * at rest it is exactly tau times the elastic stiffness (K annihilates the
    rigid modes, so P K P = K),
  * a rigid rotation Q of the body conjugates it, Q D Q^T (the order of R and
    R^T),
  * with the energy-conserving rule (spectral radius 1: alpha_m = alpha_f =
    gamma = 1/2, beta = 1/4) a small linear vibration loses, per step,
    exactly h vbar^T C vbar of kinetic + elastic energy, vbar the step's mean
    velocity and C = tau K: this fixes the alpha-point velocity of the
    viscous force and the damping coefficient of the operator.
"""
import math

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import deformed, oracle_body


def _body(tau, model=api.COROTATED, res=(5, 3, 3), E=1e5):
    sc = scenes.cantilever()
    b = sc.bodies[0]
    b.resolution = res
    b.materials = scenes.material_array(1, youngs_modulus=E, poissons_ratio=0.3, density=1000.0,
                                        elastic_model=model, stiffness_damping=tau)
    return b


def _rot(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def _damping(orc, x, d):
    """C d alone: no inertia, no stiffness, damping coefficient 1."""
    return orc.damped_apply(x, 0.0, 0.0, 1.0, d)


@pytest.mark.parametrize("model", [api.COROTATED, api.NEO_HOOKEAN])
def test_rigid_velocity_is_not_damped(model):
    orc, m = oracle_body(_body(0.01, model=model))
    x = deformed(m["rest"], seed=3, scale=0.01, twist=0.5)
    c = x.mean(axis=0)
    d = np.cross([0.3, -1.0, 0.7], x - c) + np.array([1.0, 2.0, 3.0])
    free = m["anchored"] == 0
    y = _damping(orc, x, d)
    scale = np.abs(_damping(orc, x, np.random.default_rng(1).standard_normal(x.shape))).max()
    assert np.abs(y[free]).max() <= 1e-12 * scale


def test_damping_operator_is_symmetric_positive_semidefinite():
    orc, m = oracle_body(_body(0.01))
    x = deformed(m["rest"], seed=4, scale=0.01, twist=0.4)
    free = m["anchored"] == 0
    n = x.shape[0]
    C = np.zeros((3 * n, 3 * n))
    for j in range(3 * n):
        e = np.zeros(3 * n)
        e[j] = 1.0
        C[:, j] = _damping(orc, x, e.reshape(-1, 3)).reshape(-1)
    keep = np.repeat(free, 3)
    Cf = C[np.ix_(keep, keep)]
    assert np.abs(Cf - Cf.T).max() <= 1e-12 * np.abs(Cf).max()
    assert np.linalg.eigvalsh(Cf).min() >= -1e-10 * np.abs(Cf).max()


def test_damping_at_rest_is_tau_times_elastic_stiffness():
    tau = float(np.float32(0.004))       # stiffnessDamping is a float field
    orc, m = oracle_body(_body(tau))
    X = m["rest"]
    rng = np.random.default_rng(5)
    d = rng.standard_normal(X.shape)
    d[m["anchored"] != 0] = 0
    stiff = orc.damped_apply(X, 0.0, 1.0, 0.0, d)     # K d from the elastic operator
    assert np.abs(_damping(orc, X, d) - tau * stiff).max() <= 1e-12 * np.abs(tau * stiff).max()


def test_rigid_rotation_conjugates_the_damping_operator():
    tau = 0.004
    orc, m = oracle_body(_body(tau))
    X = m["rest"]
    c = X.mean(axis=0)
    Q = _rot([1.0, 2.0, -0.5], 0.9)
    x = (X - c) @ Q.T + c
    rng = np.random.default_rng(6)
    d = rng.standard_normal(X.shape)
    d[m["anchored"] != 0] = 0
    at_rest = _damping(orc, X, d)
    turned = _damping(orc, x, d @ Q.T)
    assert np.abs(turned - at_rest @ Q.T).max() <= 1e-10 * np.abs(at_rest).max()


def test_damped_vibration_loses_exactly_the_damping_work_per_step():
    tau = 0.01
    orc, m = oracle_body(_body(tau))
    X = m["rest"]
    free = m["anchored"] == 0
    n = X.shape[0]
    v0 = np.zeros((n, 3))
    v0[free] = 1e-5 * np.random.default_rng(7).standard_normal((free.sum(), 3))
    orc.set_state(v=v0)
    h = 1 / 60.0

    def energy():
        x, v, _ = orc.state()
        _, te, _, _ = orc.eval_forces(x)
        return 0.5 * (m["mass"][:, None] * v * v).sum() + te.sum()

    e_prev = energy()
    losses, works = [], []
    for _ in range(12):
        _, v_prev, _ = orc.state()
        orc.step(dt=h, gravity=(0, 0, 0), tolerance=1e-13, cg_iterations=5000, spectral_radius=1.0)
        _, v, _ = orc.state()
        e = energy()
        vbar = 0.5 * (v + v_prev)
        hf = np.float32(h).astype(np.float64)
        works.append(hf * (vbar * _damping(orc, X, vbar)).sum())
        losses.append(e_prev - e)
        e_prev = e
    losses, works = np.array(losses), np.array(works)
    assert (works > 0).all()
    # the loss is the damping work to the (small) nonlinearity of corotation
    assert np.abs(losses - works).max() <= 2e-3 * works.max()
    # and not the undamped rule's (zero) loss, nor a half or a double of it
    assert losses.sum() == pytest.approx(works.sum(), rel=1e-3)


def test_damping_shortens_the_ring_down_of_a_released_cantilever():
    """Released from rest under gravity, the tip swing of the damped beam
    decays more, and more with more tau."""
    swings = []
    for tau in (0.0, 0.002, 0.006):
        orc, m = oracle_body(_body(tau, res=(9, 3, 3), E=1e6))
        tip = int(np.argmax(m["rest"][:, 0] + 1e-3 * m["rest"][:, 1]))
        ys = []
        for _ in range(90):
            orc.step(substeps=2, tolerance=1e-10, cg_iterations=5000)
            ys.append(orc.state()[0][tip, 1])
        swings.append(np.ptp(ys[-30:]))
    assert swings[0] > swings[1] > swings[2]
