"""bodyRotation (factor.cc:615-646) on the GPU: the block Cholesky, factored
at one configuration, is applied to a body that has turned since as
F (LL^T)^-1 F^T with F the polar factor of the mass-weighted correlation
sum m (x - c)(r - c_r)^T of the positions now (x) and at the factorization
(r).

* For a rigidly turned and moved body F is the turn itself (closed form).
* For a deformed body F equals the rotation of numpy's SVD polar
  decomposition of the same correlation.
* A free body spinning at 3 rad/s keeps its factor: with the frame the PCG
  needs fewer iterations and refactorizations than with the factor held in
  its own frame (TALUS_MI_FIXED_FRAME), and both land on the same
  trajectory (the preconditioner shapes the iterates, not the solution).
"""
from __future__ import annotations

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import product_body

pytestmark = pytest.mark.gpu


def _free_body():
    mats = scenes.material_array(1, youngs_modulus=5e5, poissons_ratio=0.3, density=1000.0,
                                 elastic_model=api.NEO_HOOKEAN)
    return scenes.Body(resolution=(7, 4, 3), voxel=(0.05, 0.05, 0.05), materials=mats)


def _turn(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def _factored(ctx, body):
    b, mesh = product_body(ctx, body)
    ctx.step(preconditioner=api.FACTOR, gravity=(0.0, 0.0, 0.0))  # factors at rest
    X = mesh["rest_positions"].astype(np.float64)
    return b, mesh, X


@pytest.mark.parametrize("precision", [api.FP64, api.MIXED])
def test_frame_of_a_turned_body_is_the_turn(precision):
    with api.Context(0, precision) as ctx:
        b, mesh, X = _factored(ctx, _free_body())
        x0, _, _ = ctx.state(b)
        R = _turn([1.0, 2.0, 0.5], 0.8)
        c = x0.mean(axis=0)
        ctx.set_state(b, (x0 - c) @ R.T + c + np.array([0.3, -0.2, 1.0]))
        F = ctx.body_frame(b)
    assert np.abs(F - R).max() <= 1e-10


def test_frame_of_a_deformed_body_is_the_polar_factor():
    with api.Context(0, api.FP64) as ctx:
        b, mesh, X = _factored(ctx, _free_body())
        x0, _, _ = ctx.state(b)
        rng = np.random.default_rng(4)
        A = _turn([0.0, 0.3, 1.0], 2.5) @ (np.eye(3) + 0.1 * rng.standard_normal((3, 3)))
        x = x0 @ A.T + 0.01 * rng.standard_normal(x0.shape)
        ctx.set_state(b, x)
        F = ctx.body_frame(b)
    m = np.zeros(len(x0))
    tets = mesh["tet_nodes"].reshape(-1, 4)
    # masses rho V / 4 per incident tet, as the library assigns them
    vol = np.abs(np.linalg.det(np.stack([X[tets[:, k]] - X[tets[:, 0]] for k in (1, 2, 3)], axis=2))) / 6.0
    np.add.at(m, tets.reshape(-1), np.repeat(1000.0 * vol / 4.0, 4))
    c, cr = (m[:, None] * x).sum(0) / m.sum(), (m[:, None] * x0).sum(0) / m.sum()
    C = ((x - c) * m[:, None]).T @ (x0 - cr)
    U, _, Vt = np.linalg.svd(C)
    assert np.linalg.det(U @ Vt) > 0
    assert np.abs(F - U @ Vt).max() <= 1e-8


def test_factor_follows_a_spinning_body():
    body = _free_body()
    out = {}
    for flags in (0, api.FIXED_FRAME):
        with api.Context(0, api.FP64) as ctx:
            b, mesh = product_body(ctx, body)
            X = mesh["rest_positions"].astype(np.float64)
            c = X.mean(axis=0)
            w = np.array([0.0, 0.0, 3.0])
            ctx.set_state(b, X, np.cross(w, X - c))
            params = api.default_params(preconditioner=api.FACTOR, gravity=(0.0, 0.0, 0.0), tolerance=1e-8,
                                        cg_iterations=500, factor_budget=500, flags=flags)
            reps = [ctx.step(params) for _ in range(60)]
            x, _, _ = ctx.state(b)
        out[flags] = (x, sum(r["cg_iterations"] for r in reps), sum(r["refactors"] for r in reps))
    x_f, it_f, ref_f = out[0]
    x_0, it_0, ref_0 = out[api.FIXED_FRAME]
    assert np.abs(x_f - x_0).max() <= 1e-6 * np.abs(x_0).max()
    assert it_f < it_0 and ref_f <= ref_0, (it_f, it_0, ref_f, ref_0)
