"""GPU adjoint of the converged implicit rollout: against the oracle's FP64
CPU gradient (tests/test_adjoint_cpu.py pins that one by finite
differences), and end to end through the torch.autograd harness against
central finite differences of the GPU rollout itself (h = 1e-6 relative;
bar |g_i - fd_i| <= 1e-4 max_j |g_j| over the gradient of the same kind,
since a random element's gradient can sit below the finite difference's
own rounding)."""
import numpy as np
import pytest
import torch

from talus_amd import adjoint, api, scenes

from helpers import oracle_body, product_body
from test_adjoint_cpu import EVERY, STEPS, loss_parts, small_beam

pytestmark = pytest.mark.gpu


def test_gpu_gradient_matches_the_oracle():
    body = small_beam()
    orc, m = oracle_body(body)
    rest = m["rest"]
    zero = np.zeros_like(rest)
    orc.set_state(rest, zero, zero)
    orc.rollout(STEPS, EVERY)
    L, xb, w = loss_parts(orc, rest)
    go = orc.rollout_backward(xb, w)
    xo, vo, _ = orc.state()
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        ctx.set_state(b, rest, zero, zero)
        ctx.rollout(STEPS, EVERY, preconditioner=api.FACTOR, tolerance=1e-10, cg_iterations=500)
        x, v, _ = ctx.state(b)
        g = ctx.rollout_backward(b, xb, w, 1e-12)
        x_after, _, _ = ctx.state(b)
    disp = np.abs(xo - rest).max()
    assert np.abs(x - xo).max() <= 1e-9 * disp
    assert np.array_equal(x_after, x)  # the state is left at the rollout's end
    for k, ko in (("lambda", "lambda"), ("mu", "mu"), ("rest", "X"), ("x0", "x0"), ("v0", "v0")):
        scale = np.abs(go[ko]).max()
        assert np.abs(g[k] - go[ko]).max() <= 1e-6 * scale, k


def _beam(cells=(12, 3, 3), seed=4):
    rng = np.random.default_rng(seed)
    h = 0.04
    tets = 6 * cells[0] * cells[1] * cells[2]
    E = 2e5 * np.exp(0.3 * rng.standard_normal(tets))
    nu = rng.uniform(0.25, 0.4, tets)
    mats = scenes.material_array(tets, youngs_modulus=E, poissons_ratio=nu, density=1000.0,
                                 elastic_model=api.NEO_HOOKEAN)
    body = scenes.Body(resolution=tuple(c + 1 for c in cells), voxel=(h, h, h),
                       supports=[((0.0, 0.06, 0.06), (0.2 * h, 1.0, 1.0))], materials=mats,
                       per_tet=True)
    return body, E, nu


def test_autograd_gradients_match_gpu_finite_differences():
    body, E0, nu0 = _beam()
    steps, every = 10, 4
    params = api.default_params(preconditioner=api.FACTOR, tolerance=1e-10, cg_iterations=500)
    with api.Context(0, api.FP64) as ctx:
        b, mesh = product_body(ctx, body)
        rest0 = torch.tensor(mesh["rest_positions"].astype(np.float64))
        tip = int(np.argmax(mesh["rest_positions"][:, 0] * 1e3 + mesh["rest_positions"][:, 1]))
        target = rest0[tip] + torch.tensor([0.0, -0.02, 0.01], dtype=torch.float64)
        wv = torch.tensor(np.random.default_rng(9).standard_normal(rest0.shape) * 1e-3)

        def loss(E, nu, rest):
            x, v = adjoint.rollout(ctx, b, E, nu, rest, steps=steps, every=every, params=params)
            return ((x[tip] - target) ** 2).sum() + (wv * v).sum()

        E = torch.tensor(E0, requires_grad=True)
        nu = torch.tensor(nu0, requires_grad=True)
        rest = rest0.clone().requires_grad_(True)
        L = loss(E, nu, rest)
        L.backward()
        gE, gnu, grest = E.grad.numpy(), nu.grad.numpy(), rest.grad.numpy()
        rng = np.random.default_rng(17)
        with torch.no_grad():
            for idx in rng.choice(E0.size, 3, replace=False):
                for which, base, grad in (("E", E0, gE), ("nu", nu0, gnu)):
                    h = 1e-6 * base[idx]
                    vals = []
                    for sgn in (1.0, -1.0):
                        p = base.copy()
                        p[idx] += sgn * h
                        Ep = torch.tensor(p) if which == "E" else torch.tensor(E0)
                        nup = torch.tensor(p) if which == "nu" else torch.tensor(nu0)
                        vals.append(float(loss(Ep, nup, rest0)))
                    fd = (vals[0] - vals[1]) / (2 * h)
                    assert abs(grad[idx] - fd) <= 1e-4 * np.abs(grad).max(), (which, idx, grad[idx], fd)
            free = np.flatnonzero(mesh["anchored"] == 0)
            for node in rng.choice(free, 2, replace=False):
                c = int(rng.integers(3))
                h = 1e-7
                vals = []
                for sgn in (1.0, -1.0):
                    r = rest0.clone()
                    r[node, c] += sgn * h
                    vals.append(float(loss(torch.tensor(E0), torch.tensor(nu0), r)))
                fd = (vals[0] - vals[1]) / (2 * h)
                assert abs(grest[node, c] - fd) <= 1e-4 * np.abs(grest).max(), (node, c, grest[node, c], fd)
