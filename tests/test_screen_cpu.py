"""The FP64 neo-Hookean hex, the lattice energy ledger and the screen factor
in the oracle (oracle/screen_oracle.c), pinned by what the mathematics and
the reference's own tests fix.

Reference tests mirrored (paths relative to the Talus tree):
  rest and rigidly turned cell carry no force  tests/dynamics_tests.cc:249-296
  node Hessian vs FD of its gradient           dynamics_tests.cc:191-247
  inverted cell keeps finite terms             dynamics_tests.cc:298-336
  rest stiffness vs FD of the corner forces    dynamics_tests.cc:2755-2800
  squeezed free block shortens by F L / (E A)  dynamics_tests.cc:2802-2881
  a rigid load makes no dent                   dynamics_tests.cc:2883-2910
plus: the gradient is the FD of the cell energy, an affine deformation's
energy equals the tet oracle's stable neo-Hookean on the cell's six Kuhn
tets (Gauss quadrature is exact for a constant F), the volume ratio is
det F, the rest body's point-mass sums, and the held screen solve equals a
dense numpy solve of the assembled system.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import hex as H
from oracle import screen as S

HALF = np.array([0.0125, 0.01, 0.015])
VOLUME = 8.0 * HALF.prod()
SHEAR, SHIFTED = S.moduli(2.0e5, 0.4)


def rest_corners(half=HALF):
    return np.array([[h if (c >> k) & 1 else -h for k, h in enumerate(half)] for c in range(8)])


def deformed_corners(rest):
    shape = np.array([[1.08, 0.05, -0.03], [-0.04, 0.93, 0.06], [0.02, -0.01, 1.11]])
    c = np.arange(8)
    noise = 1e-3 * np.stack([np.sin(1.0 + c), np.cos(2.0 * c), np.sin(3.0 * c)], -1)
    return rest @ shape.T + noise


def rot(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def terms(x, rest, corner):
    return S.node_terms(x, rest, HALF, VOLUME, SHEAR, SHIFTED, corner)


def test_moduli_closed_form_and_clamp():
    E, nu = np.float32(2.0e5), np.float32(0.4)
    mu = float(E) / (2 * (1 + float(nu)))
    lam = float(E) * float(nu) / ((1 + float(nu)) * (1 - 2 * float(nu)))
    assert S.moduli(E, nu) == pytest.approx((mu, lam + mu), rel=1e-15)
    assert S.moduli(1e5, 0.6) == S.moduli(1e5, 0.49)
    assert S.moduli(1e5, -0.2) == pytest.approx((5e4, 5e4), rel=1e-15)


def test_rest_and_turned_cells_carry_no_force():
    rest = rest_corners()
    turned = rest @ rot([1.0, 2.0, -0.5], 0.7).T + np.array([0.3, -0.1, 0.2])
    unit = SHEAR * VOLUME / HALF[0]
    for c in range(8):
        assert np.linalg.norm(terms(rest, rest, c)[0]) <= 1e-12 * unit
        assert np.linalg.norm(terms(turned, rest, c)[0]) <= 1e-9 * unit
    assert abs(S.cell_energy(turned, rest, HALF, VOLUME, SHEAR, SHIFTED)) <= 1e-12 * unit * HALF[0]


def test_gradient_is_the_energy_derivative():
    rest = rest_corners()
    x = deformed_corners(rest)
    step = 1e-7 * HALF[0]
    scale = max(np.linalg.norm(terms(x, rest, c)[0]) for c in range(8))
    for c in range(8):
        g, _ = terms(x, rest, c)
        for k in range(3):
            p, m = x.copy(), x.copy()
            p[c, k] += step
            m[c, k] -= step
            fd = (S.cell_energy(p, rest, HALF, VOLUME, SHEAR, SHIFTED)
                  - S.cell_energy(m, rest, HALF, VOLUME, SHEAR, SHIFTED)) / (2 * step)
            assert abs(fd - g[k]) <= 1e-6 * scale


def test_node_hessian_is_the_derivative_of_its_gradient():
    rest = rest_corners()
    x = deformed_corners(rest)
    step = 1e-7
    for c in range(8):
        _, Hc = terms(x, rest, c)
        scale = np.abs(Hc).max()
        for k in range(3):
            p, m = x.copy(), x.copy()
            p[c, k] += step
            m[c, k] -= step
            fd = (terms(p, rest, c)[0] - terms(m, rest, c)[0]) / (2 * step)
            assert np.abs(fd - Hc[:, k]).max() <= 1e-6 * scale


def test_inverted_cell_keeps_finite_terms():
    rest = rest_corners()
    inverted = rest * np.array([1.0, -1.0, 1.0])
    assert S.volume_ratio(inverted, rest, HALF) < 0.0
    assert np.isfinite(S.cell_energy(inverted, rest, HALF, VOLUME, SHEAR, SHIFTED))
    for c in range(8):
        g, Hc = terms(inverted, rest, c)
        assert np.isfinite(g).all() and np.isfinite(Hc).all()


def test_rest_stiffness_is_the_derivative_of_the_corner_forces():
    rest = rest_corners()
    K = S.rest_stiffness(HALF, VOLUME, SHEAR, SHIFTED)
    step = 1e-7 * HALF[0]
    worst = 0.0
    for moved in range(8):
        for k in range(3):
            p, m = rest.copy(), rest.copy()
            p[moved, k] += step
            m[moved, k] -= step
            for c in range(8):
                fd = (terms(p, rest, c)[0] - terms(m, rest, c)[0]) / (2 * step)
                worst = max(worst, np.abs(fd - K[3 * c:3 * c + 3, 3 * moved + k]).max())
    assert worst <= 1e-6 * np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    # translations are in its null space
    for k in range(3):
        t = np.zeros(24)
        t[k::3] = 1.0
        assert np.abs(K @ t).max() <= 1e-10 * np.abs(K).max()


KUHN = [(0, 1, 3, 7), (0, 1, 7, 5), (0, 2, 7, 3), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 7, 6)]


def test_affine_energy_equals_the_tet_oracle_on_the_kuhn_tets():
    """A constant F is integrated exactly by the 8-point rule, so the cell's
    energy is the stable neo-Hookean density times V, which the tet oracle
    (pinned by tests/test_oracle.py) gives on the cell's six tets."""
    rest = rest_corners()
    A = np.array([[1.05, 0.08, -0.02], [0.03, 0.9, 0.04], [-0.06, 0.01, 1.12]])
    x = rest @ A.T
    total = 0.0
    for tet in KUHN:
        X = rest[list(tet)]
        Ds = (X[1:] - X[0]).T
        if np.linalg.det(Ds) < 0:
            tet = (tet[0], tet[2], tet[1], tet[3])
            X = rest[list(tet)]
            Ds = (X[1:] - X[0]).T
        V = np.linalg.det(Ds) / 6.0
        u = (x[list(tet)] - X).reshape(-1)
        _, e = oracle.snh_tet(u, np.linalg.inv(Ds), V, SHIFTED - SHEAR, SHEAR)
        total += e
    energy = S.cell_energy(x, rest, HALF, VOLUME, SHEAR, SHIFTED)
    assert energy == pytest.approx(total, rel=1e-12)
    assert S.volume_ratio(x, rest, HALF) == pytest.approx(np.linalg.det(A), rel=1e-13)


def jelly(E=2.0e5, poisson=0.4, density=1000.0):
    m = oracle.Material()
    m.friction, m.density, m.strength = 0.7, density, math.inf
    m.youngs_modulus, m.poissons_ratio = E, poisson
    m.yield_stress, m.damping_ratio, m.tear_strain = math.inf, math.inf, math.inf
    m.plastic_max, m.crush_stress, m.crush_limit = math.inf, math.inf, 1.0
    return m


def block(cells=(4, 1, 4), cell=0.025, mats=None, supports=()):
    mats = mats or [jelly()]
    return H.build_hex_lattice([c + 1 for c in cells], [cell] * 3, (0.0, 0.0, 0.0), mats, list(supports)), mats


def test_lattice_energy_terms():
    mesh, mats = block((3, 2, 2))
    shear, shifted = S.cell_moduli(mesh, mats)
    rng = np.random.default_rng(3)
    n = mesh["rest"].shape[0]
    offset = np.array([1.0, -2.0, 0.5])
    x = mesh["rest"] + offset + 1e-3 * rng.standard_normal((n, 3))
    v = rng.standard_normal((n, 3))
    g = np.array([0.0, -9.81, 0.5])
    ke, pe, el = S.lattice_energy(mesh, shear, shifted, x, v, g)
    m = mesh["mass"].astype(np.float64)
    free = mesh["anchored"] == 0
    assert ke == pytest.approx(0.5 * (m[free, None] * v[free] ** 2).sum(), rel=1e-13)
    assert pe == pytest.approx(-(m[:, None] * g * x).sum(), rel=1e-13)
    # translation invariance: the elastic part ignores the offset
    cells = [S.cell_energy(x[h] - offset, mesh["rest"][h], mesh["voxel"].astype(np.float64) / 2,
                           float(mesh["volume"][i]), shear[i], shifted[i]) for i, h in enumerate(mesh["hex"])]
    assert el == pytest.approx(sum(cells), rel=1e-9)


def test_rest_body_point_mass_sums():
    mesh, _ = block((4, 2, 3), cell=0.02)
    rb = S.rest_body(mesh)
    m = mesh["mass"].astype(np.float64)
    X = mesh["rest"]
    c = (m[:, None] * X).sum(0) / m.sum()
    a = X - c
    I = (m[:, None, None] * ((a * a).sum(1)[:, None, None] * np.eye(3) - a[:, :, None] * a[:, None, :])).sum(0)
    assert rb["total"] == pytest.approx(m.sum(), rel=1e-14)
    assert np.allclose(rb["centre"], c, rtol=0, atol=1e-14)
    assert np.allclose(rb["inertia"], I, rtol=1e-12, atol=0)
    R = rb["axes"]
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-12) and np.linalg.det(R) > 0
    assert np.allclose(I @ R, R * rb["moments"], rtol=0, atol=1e-10 * np.abs(I).max())
    assert np.all(np.diff(rb["moments"]) >= 0)
    assert rb["thickness"] == pytest.approx(np.ptp(X, axis=0).min(), rel=1e-14)


def _edge_share(p, lo, hi):
    return 0.5 if abs(p - lo) < 1e-9 or abs(p - hi) < 1e-9 else 1.0


def test_squeezed_free_block_shortens_by_load_length_over_stiffness_area():
    E = 2.0e5
    mesh, mats = block((4, 1, 4), mats=[jelly(E=E)])
    shear, shifted = S.cell_moduli(mesh, mats)
    X = mesh["rest"]
    lo, hi = X.min(0), X.max(0)
    load = 4.905
    f = np.zeros_like(X)
    top = np.abs(X[:, 1] - hi[1]) < 1e-9
    bottom = np.abs(X[:, 1] - lo[1]) < 1e-9
    share = np.array([_edge_share(p[0], lo[0], hi[0]) * _edge_share(p[2], lo[2], hi[2]) for p in X])
    f[top, 1] = -load * share[top] / share[top].sum()
    f[bottom, 1] = load * share[bottom] / share[bottom].sum()
    u, dent = S.screen_load(mesh, shear, shifted, f)
    shortening = u[bottom, 1].mean() - u[top, 1].mean()
    expected = load * (hi[1] - lo[1]) / (np.float32(E) * (hi[0] - lo[0]) * (hi[2] - lo[2]))
    assert shortening == pytest.approx(expected, rel=1e-3)
    assert dent == pytest.approx(np.linalg.norm(u, axis=1).max() / np.ptp(X, axis=0).min(), rel=1e-14)


def test_a_rigid_load_makes_no_dent():
    mesh, mats = block((4, 1, 4))
    shear, shifted = S.cell_moduli(mesh, mats)
    rb = S.rest_body(mesh)
    arm = mesh["rest"] - rb["centre"]
    f = mesh["mass"].astype(np.float64)[:, None] * (np.array([1.0, 2.0, -3.0]) + np.cross([4.0, -1.0, 2.0], arm))
    _, dent = S.screen_load(mesh, shear, shifted, f)
    assert dent < 1e-9


def test_held_screen_equals_a_dense_solve_of_the_assembled_system():
    mesh, mats = block((3, 2, 2), mats=[jelly(), jelly(E=6e5, poisson=0.3)])
    rng = np.random.default_rng(8)
    mesh["lane"] = rng.integers(0, 2, mesh["lane"].size).astype(np.uint32)
    shear, shifted = S.cell_moduli(mesh, mats)
    n = mesh["rest"].shape[0]
    held = np.flatnonzero(mesh["rest"][:, 0] < 1e-9)
    f = rng.standard_normal((n, 3))
    u, dent = S.screen_load(mesh, shear, shifted, f, held)
    half = mesh["voxel"].astype(np.float64) / 2
    A = np.zeros((3 * n, 3 * n))
    for i, h in enumerate(mesh["hex"]):
        K = S.rest_stiffness(half, float(mesh["volume"][i]), shear[i], shifted[i])
        idx = (3 * h[:, None] + np.arange(3)).reshape(-1)
        A[np.ix_(idx, idx)] += K
    largest = A.diagonal().max()
    hold = np.zeros(n, bool)
    hold[held] = True
    shift = np.where(hold, 1e6 * largest, 1e-9 * largest / mesh["mass"].max() * mesh["mass"])
    A += np.diag(np.repeat(shift, 3))
    b = np.where(hold[:, None], 0.0, f).reshape(-1)
    ref = np.linalg.solve(A, b).reshape(n, 3)
    assert np.abs(u - ref).max() <= 1e-9 * np.abs(ref).max()
    assert np.abs(u[hold]).max() <= 1e-5 * np.abs(u).max()
    assert dent == pytest.approx(np.linalg.norm(ref, axis=1).max() / np.ptp(mesh["rest"], axis=0).min(), rel=1e-8)
