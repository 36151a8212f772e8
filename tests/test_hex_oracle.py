"""The hex-lattice oracle (oracle/hex_oracle.c) against the closed forms the
reference's own strainstep tests hold (tests/dynamics_strainstep_tests.cc):
rigid motion gives no strain and no impulse, rotationOf finds a turn, the
solve keeps momentum and is exact in one block solve, row stiffness, strain
rows, bending energy of one beam cell, the volume gradient against finite
differences, the rows' soft springs against the soft-step formula, free fall
as semi-implicit Euler and a cantilever against beam theory.  No GPU."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import hex as H


def unit_cell(half, mass):
    """testCell (dynamics_strainstep_tests.cc:187-201): corners at half *
    signs, inverse masses 1/(m (1 + 0.1 c))."""
    rest = np.array([half * H.corner_signs(c) for c in range(8)])
    inv = np.array([1.0 / (mass * (1.0 + 0.1 * c)) for c in range(8)])
    return rest, inv


def unit_element(rest, inv, youngs, poisson, h):
    side = rest[7, 0] - rest[0, 0]
    e = H.make(rest, inv, side ** 3, np.float32(youngs), np.float32(poisson), h)
    assert e is not None
    return e


def rot(axis, angle):
    return H.quat_from_axis_angle(axis, angle)


def qmat(q):
    return H.quat_matrix(q)


def test_turned_and_moved_hex_has_no_strain_and_no_impulse():
    """dynamics_strainstep_tests.cc:1115-1138"""
    rest, inv = unit_cell(0.1, 1.0)
    e = unit_element(rest, inv, 1.0e5, 0.3, 1.0 / 240.0)
    q = rot([1.0, 2.0, 3.0], 2.0)
    pos = np.array([0.5, -1.0, 2.0]) + rest @ qmat(q).T
    vel = np.zeros((8, 3))
    H.set_rotation(e, q)
    H.update(e, pos)
    assert np.abs(e.arr("strain")).max() <= 1e-12
    H.solve(e, vel)
    assert np.linalg.norm(vel, axis=1).max() <= 1e-12


def test_rotation_of_recovers_a_turn_from_half_a_radian_away():
    """dynamics_strainstep_tests.cc:1140-1154"""
    q = rot([1.0, -1.0, 2.0], 0.5)
    stretch = np.array([[1.1, 0.02, 0.0], [0.02, 0.95, 0.01], [0.0, 0.01, 1.05]]).T
    found = H.rotation_of(qmat(q) @ stretch)
    diff = qmat(found) - qmat(q)
    assert np.linalg.norm(diff, axis=0).max() <= 1e-6


def test_solves_keep_a_deformed_hex_momentum():
    """dynamics_strainstep_tests.cc:1156-1190"""
    rest, inv = unit_cell(0.1, 2.0)
    e = unit_element(rest, inv, 1.0e5, 0.45, 1.0 / 240.0)
    # glm column-major dmat3(c0, c1, c2): columns given
    shape = np.array([[1.08, 0.03, -0.02], [-0.05, 0.93, 0.04], [0.02, -0.01, 1.11]]).T
    pos = np.zeros((8, 3))
    vel = np.zeros((8, 3))
    before = np.zeros(3)
    scale = 0.0
    for c in range(8):
        t = float(c)
        pos[c] = shape @ rest[c] + 1e-3 * np.array([math.sin(1 + t), math.cos(2 * t), math.sin(3 * t)])
        vel[c] = [math.cos(t), 0.5 * math.sin(2 * t), -0.3 * t]
        before += vel[c] / inv[c]
        scale += np.linalg.norm(vel[c]) / inv[c]
    H.update(e, pos)
    H.solve(e, vel)
    H.solve(e, vel)
    after = (vel / inv[:, None]).sum(axis=0)
    assert np.linalg.norm(after - before) <= 1e-12 * scale


def test_one_block_solve_is_exact():
    """dynamics_strainstep_tests.cc:1192-1222: a second solve changes no row."""
    rest, inv = unit_cell(0.1, 1.0)
    e = unit_element(rest, inv, 1.0e5, 0.3, 1.0 / 240.0)
    q = rot([0.0, 1.0, 1.0], 0.3)
    shape = np.array([[1.05, 0.0, 0.0], [0.02, 0.97, 0.0], [0.0, 0.01, 1.02]]).T
    pos = (rest @ shape.T) @ qmat(q).T
    vel = np.zeros((8, 3))
    H.set_rotation(e, q)
    H.update(e, pos)
    H.solve(e, vel)
    first, fv = e.arr("impulse"), e.vol_impulse
    H.solve(e, vel)
    assert np.all(np.abs(e.arr("impulse") - first) <= 1e-9 * (1 + np.abs(first)))
    assert abs(e.vol_impulse - fv) <= 1e-9 * (1 + abs(fv))
    assert abs(first[0]) > 0.0


def test_bent_cell_energy_equals_beam_theory():
    """dynamics_strainstep_tests.cc:1224-1248"""
    half, side, curv, E = 0.1, 0.2, 0.05, 1.0e5
    rest, inv = unit_cell(half, 1.0)
    e = unit_element(rest, inv, E, 0.3, 1.0 / 240.0)
    pos = rest + np.stack([-curv * rest[:, 0] * rest[:, 1], np.zeros(8), np.zeros(8)], axis=1)
    H.update(e, pos)
    strain = e.arr("strain")
    assert np.abs(np.delete(strain, 6)).max() <= 1e-12
    energy = 0.5 * e.stiffness[6] * strain[6] ** 2
    theory = E * curv ** 2 * side ** 5 / 24.0
    assert abs(energy - theory) <= 1e-9 * theory


def expected_stiffness(row, E, nu, V):
    mu = E / (2 * (1 + nu))
    return 2 * mu * V if row < 3 else (4 * mu * V if row < 6 else E * V / 3.0)


def test_row_stiffness_from_youngs_modulus_and_poisson():
    """dynamics_strainstep_tests.cc:1250-1268"""
    half, E, nu = 0.1, 1.0e5, 0.375
    V = (2 * half) ** 3
    rest, inv = unit_cell(half, 1.0)
    e = unit_element(rest, inv, E, nu, 1.0 / 240.0)
    for r in range(18):
        x = expected_stiffness(r, E, nu, V)
        assert abs(e.stiffness[r] - x) <= 1e-9 * x
    xv = E * nu / ((1 + nu) * (1 - 2 * nu)) * V
    assert abs(e.vol_stiffness - xv) <= 1e-9 * xv


def test_six_strain_rows_of_a_stretched_and_sheared_hex():
    """dynamics_strainstep_tests.cc:1270-1290"""
    rest, inv = unit_cell(0.1, 1.0)
    e = unit_element(rest, inv, 1.0e5, 0.3, 1.0 / 240.0)
    S = np.array([[1.02, 0.03, 0.01], [0.03, 0.99, 0.02], [0.01, 0.02, 1.04]])
    H.update(e, rest @ S.T)
    expected = np.array([0.02, -0.01, 0.04, 0.03, 0.02, 0.01] + [0.0] * 12)
    assert np.abs(e.arr("strain") - expected).max() <= 1e-12


def test_sheared_or_warped_rest_cells_that_only_turn_have_no_strain():
    """dynamics_strainstep_tests.cc:1292-1331"""
    half = 0.1
    shear = np.array([[1.0, 0.3, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])  # columns (1,0,0),(0.3,1,0),(0,0,1)
    q = rot([-1.0, 0.5, 2.0], 1.1)
    sheared = np.array([half * (shear @ H.corner_signs(c)) for c in range(8)])
    warped = sheared.copy()
    warped[7] += [0.02, 0.01, -0.015]
    warped[3] += [-0.01, 0.0, 0.02]
    s_shape, _ = H.corner_rest(sheared)
    w_shape, w_hg = H.corner_rest(warped)
    assert abs(s_shape[0, 1] - s_shape[1, 0]) > 1e-3
    assert abs(w_shape[0, 1] - w_shape[1, 0]) > 1e-3
    assert np.linalg.norm(w_hg[0]) > 1e-3
    _, inv = unit_cell(half, 1.0)
    for cell in (sheared, warped):
        side = cell[7, 0] - cell[0, 0]
        e = H.make(cell, inv, side ** 3, np.float32(1e5), np.float32(0.3), 1.0 / 240.0)
        pos = np.array([1.0, 2.0, 3.0]) + cell @ qmat(q).T
        H.set_rotation(e, q)
        H.update(e, pos)
        assert np.abs(e.arr("strain")).max() <= 1e-12
        assert abs(e.rhs / e.vol_rate) <= 1e-12


def test_volume_gradient_is_the_derivative_of_the_volume():
    """dynamics_strainstep_tests.cc:1333-1374"""
    delta = 1e-6
    rest, inv = unit_cell(0.1, 1.0)
    e = unit_element(rest, inv, 1.0e5, 0.3, 1.0 / 240.0)
    q = rot([1.0, 2.0, 3.0], 0.3)
    S = np.array([[1.02, 0.03, 0.01], [0.03, 0.99, 0.02], [0.01, 0.02, 1.04]])
    pos = (rest @ S.T) @ qmat(q).T
    H.update(e, pos)
    grad = e.arr("vol_gradient").reshape(8, 3)

    def volume_at(moved):
        c = H.Hex.from_buffer_copy(e)
        H.update(c, moved)
        return c.rhs / c.vol_rate + 1.0

    assert abs(volume_at(pos) - np.linalg.det(S)) <= 1e-12
    for c in range(8):
        for k in range(3):
            f, b = pos.copy(), pos.copy()
            f[c, k] += delta
            b[c, k] -= delta
            slope = (volume_at(f) - volume_at(b)) / (2 * delta)
            assert abs(slope - grad[c, k]) <= 1e-7


def make_soft(hertz, zeta, h):
    """Soft step of a spring of `hertz` and damping ratio `zeta` over h
    (E. Catto, "Solver2D", 2024: omega = 2 pi f, a1 = 2 zeta + h omega,
    a2 = h omega a1, a3 = 1 / (1 + a2); bias rate omega / a1, mass scale
    a2 a3, impulse scale a3)."""
    omega = 2 * math.pi * hertz
    a1 = 2 * zeta + h * omega
    a2 = h * omega * a1
    a3 = 1 / (1 + a2)
    return omega / a1, a2 * a3, a3


def test_row_springs_are_critically_damped_soft_steps():
    """dynamics_strainstep_tests.cc:1376-1420"""
    half, E, nu, h = 0.1, 1.0e5, 0.375, 1.0 / 240.0
    V = (2 * half) ** 3
    grad, bend = 1.0 / (8 * half), 1.0 / (4 * 2 * half)
    rest, inv = unit_cell(half, 1.0)
    e = unit_element(rest, inv, E, nu, h)
    sum_inv = inv.sum()

    def expect(stiffness, weight, rate, softness):
        hertz = math.sqrt(stiffness * weight) / (2 * math.pi)
        bias, mass_scale, impulse_scale = make_soft(hertz, 1.0, h)
        assert abs(rate - bias) <= 1e-9 * bias
        assert abs(softness - impulse_scale / mass_scale) <= 1e-9 * softness

    for r in range(18):
        jsq = grad * grad if r < 3 else (0.5 * grad * grad if r < 6 else bend * bend)
        w = jsq * sum_inv
        expect(expected_stiffness(r, E, nu, V), w, e.rate[r], e.softness[r] / w)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu)) * V
    expect(lam, 3 * grad * grad * sum_inv, e.vol_rate, e.vol_softness)


# ---- lattice bodies --------------------------------------------------------

def jelly(E, density=1000.0, poisson=0.4):
    m = oracle.Material()
    m.friction, m.density, m.strength = 0.7, density, math.inf
    m.youngs_modulus, m.poissons_ratio = E, poisson
    m.yield_stress, m.damping_ratio, m.tear_strain = math.inf, math.inf, math.inf
    m.plastic_max, m.crush_stress, m.crush_limit = math.inf, math.inf, 1.0
    return m


def hold_at(centre, half):
    s = oracle.Support()
    M = [1, 0, 0, -centre[0], 0, 1, 0, -centre[1], 0, 0, 1, -centre[2]]
    s.world_to_support[:] = [float(v) for v in M]
    s.half[:] = [float(v) for v in half]
    return s


def box(half, cell, material, supports=(), lanes=2):
    """A solid box lattice of hexes centred on the origin (softBox,
    dynamics_strainstep_tests.cc:103-123, for boxes the voxel divides)."""
    cells = [int(round(2 * h / cell)) for h in half]
    mesh = H.build_hex_lattice([c + 1 for c in cells], [cell] * 3, [-h for h in half],
                               [material] * lanes, list(supports))
    return mesh


@pytest.mark.parametrize("requested,expected", [(0, 4), (8, 8)])
def test_free_fall_is_semi_implicit_euler(requested, expected):
    """checkFreeFall (dynamics_strainstep_tests.cc:156-183, 1071-1077)"""
    mat = jelly(8.0e3)
    mesh = box((0.3, 0.3, 0.3), 0.15, mat)
    body = H.LatticeBody(mesh, [mat, mat], offset=(0.0, 50.0, 0.0))
    body.set_substeps(requested)
    x0, _ = body.state()
    steps = 30
    for _ in range(steps):
        assert body.step() == expected
    x, v = body.state()
    h = (1.0 / 60.0) / expected
    n = expected * steps
    assert np.abs(v[:, 1] + 9.81 * h * n).max() <= 1e-9
    assert np.abs(x[:, 1] - x0[:, 1] + 9.81 * h * h * n * (n + 1) / 2).max() <= 1e-9
    assert np.abs(x[:, [0, 2]] - x0[:, [0, 2]]).max() <= 1e-12


def test_hex_mesh_counts_masses_and_colours():
    mat = jelly(8.0e3)
    mesh = box((0.3, 0.3, 0.3), 0.15, mat)
    assert mesh["hex"].shape == (64, 8) and mesh["rest"].shape == (125, 3)
    # rho V / 8 per corner, summed in float: total mass rho * box volume
    assert abs(mesh["mass"].astype(np.float64).sum() - 1000.0 * 0.6 ** 3) <= 1e-4
    body = H.LatticeBody(mesh, [mat])
    ncol, hcol = body.colours()
    assert set(hcol.tolist()) == set(range(8))
    # no two hexes of one colour share a node
    for c in range(8):
        nodes = mesh["hex"][hcol == c].reshape(-1)
        assert np.unique(nodes).size == nodes.size
    # an interior node gathers 8 corner shares, a box corner node one
    assert mesh["mass"].max() == 8 * mesh["mass"].min()


def cantilever(thickness, cell, poisson, substeps, steps, colored=False, damping=math.inf, change=None):
    """cantilever (dynamics_strainstep_tests.cc:305-360): a 4 m beam held at
    x = -2, tip sag averaged over the last half of the frames; change =
    (frame, substeps) asks for other substeps from that frame on."""
    length = 4.0
    mat = jelly(5.0e5, density=5.0, poisson=poisson)
    mat.friction = 0.5
    mat.damping_ratio = damping
    clamp = [hold_at((-length / 2, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((length / 2, thickness / 2, thickness / 2), cell, mat, clamp)
    body = H.LatticeBody(mesh, [mat, mat])
    body.set_substeps(substeps)
    tip = int(np.argmin(np.linalg.norm(mesh["rest"] - [length / 2, thickness / 2, thickness / 2], axis=1)))
    window = steps // 2
    sag = 0.0
    for frame in range(1, steps + 1):
        if change is not None and frame == change[0]:
            body.set_substeps(change[1])
        body.step(colored=colored)
        if frame > steps - window:
            x, _ = body.state()
            sag += (mesh["rest"][tip, 1] - x[tip, 1]) / window
    inertia = thickness ** 4 / 12.0
    theory = 5.0 * 9.81 * thickness ** 2 * length ** 4 / (8 * 5.0e5 * inertia)
    return sag, theory, mesh["hex"].shape[0]


@pytest.mark.parametrize("colored", [False, True])
@pytest.mark.parametrize("poisson", [0.0, 0.45])
def test_one_cell_thick_cantilever_bends_as_beam_theory(poisson, colored):
    """dynamics_strainstep_tests.cc:1806-1813"""
    sag, theory, cells = cantilever(0.4, 0.4, poisson, 48, 600, colored)
    assert cells == 10
    assert abs(sag - theory) / theory < 0.1


@pytest.mark.parametrize("colored", [False, True])
def test_two_cell_thick_cantilever_bends_as_beam_theory(colored):
    """dynamics_strainstep_tests.cc:1815-1820"""
    sag, theory, cells = cantilever(0.4, 0.2, 0.3, 16, 600, colored)
    assert cells == 80
    assert abs(sag - theory) / theory < 0.1


def test_colour_order_and_cell_order_land_together():
    """The colour sweep is another Gauss-Seidel order of the same element
    solves: on a sagging block both orders settle to the same shape."""
    mat = jelly(2.0e5, poisson=0.3)
    clamp = [hold_at((-0.4, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((0.4, 0.1, 0.1), 0.1, mat, clamp)
    a, b = H.LatticeBody(mesh, [mat]), H.LatticeBody(mesh, [mat])
    for _ in range(120):
        a.step(colored=False)
        b.step(colored=True)
    xa, _ = a.state()
    xb, _ = b.state()
    sag = np.abs(xa - mesh["rest"]).max()
    assert sag > 1e-3
    assert np.abs(xa - xb).max() <= 0.02 * sag


# ---- lattice plasticity (plasticflow.cc) -----------------------------------

def von_mises(elastic, shear):
    """vonMises (plasticflow.cc:96-108) from numpy's eigenvalues of F F^T."""
    w = np.linalg.eigvalsh(elastic @ elastic.T)
    e = 0.5 * np.log(w)
    return math.sqrt(1.5) * 2.0 * shear * np.linalg.norm(e - e.mean())


def clay(E, yield_stress, **kw):
    m = jelly(E, density=1000.0, poisson=0.35)
    m.yield_stress, m.hardening_modulus, m.plastic_max = yield_stress, 0.0, 1.0
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def one_cell(mat, half=0.1, voxel=0.2):
    mesh = box((half, half, half), voxel, mat, lanes=1)
    assert mesh["hex"].shape[0] == 1
    return mesh, H.LatticeBody(mesh, [mat])


def sheared_cell(mat):
    mesh, body = one_cell(mat)
    rest = mesh["rest"]
    body.set_state(rest + np.stack([0.2 * rest[:, 1], np.zeros(len(rest)), np.zeros(len(rest))], axis=1))
    return mesh, body


def test_shape_gradients_sum_to_zero_and_reproduce_linear_fields():
    """trilinear gradients: a linear field x -> A x has gradient A at every point"""
    half = np.array([0.1, 0.2, 0.15])
    corners = np.array([half * H.corner_signs(c) for c in range(8)])
    A = np.array([[1.0, 0.2, -0.3], [0.4, 0.9, 0.1], [0.0, 0.5, 1.2]])
    for p in range(8):
        g = H.shape_gradients(0.5773502691896258 * H.corner_signs(p), half)
        assert np.abs(g.sum(axis=0)).max() <= 1e-14
        F = (corners @ A.T).T @ g
        np.testing.assert_allclose(F, A, atol=1e-13)


def test_point_inside_its_yield_surface_stays_elastic():
    """dynamics_strainstep_tests.cc:3090-3094"""
    E = np.array([[1.0, 0.001, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])  # columns (1,0,0), (0.001,1,0), (0,0,1)
    assert H.return_map(E, 1.0e5, 1.0e3) is None


def test_point_past_yield_returns_to_its_surface_and_keeps_volume():
    """dynamics_strainstep_tests.cc:3096-3104"""
    E = np.array([[1.05, 0.02, 0.0], [0.2, 0.98, 0.01], [0.0, 0.03, 1.01]]).T
    mapped, trial, strain = H.return_map(E, 1.0e5, 2.0e3)
    assert abs(von_mises(mapped, 1.0e5) - 2.0e3) <= 1e-9 * 2.0e3
    assert abs(np.linalg.det(mapped) - np.linalg.det(E)) <= 1e-12
    assert abs(strain - (trial - 2.0e3) / 3.0e5) <= 1e-12


def test_hardening_follows_the_hardening_modulus():
    """dynamics_strainstep_tests.cc:3106-3138"""
    mat = clay(1.0e5, 2.0e3, hardening_modulus=5.0e4)
    mesh, body = sheared_cell(mat)
    body.flow()
    P, a = body.plastic()
    shear, _ = H.material(np.float32(1e5), np.float32(0.35))
    half = 0.5 * mesh["voxel"].astype(np.float64)
    g = H.shape_gradients(0.5773502691896258 * H.corner_signs(0), half)
    x, _ = body.state()
    F = x[mesh["hex"][0]].T @ g
    y = 2.0e3 + 5.0e4 * a[0]
    assert a[0] > 0.0
    assert abs(von_mises(F @ P[0], shear) - y) <= 1e-6 * y


def test_plastic_stretch_stays_within_the_cap():
    """dynamics_strainstep_tests.cc:3140-3167"""
    mat = clay(5.0e3, 50.0)
    mesh, body = one_cell(mat, half=0.25, voxel=0.5)
    rest = mesh["rest"]
    for pull in range(1, 7):
        s = 1.0 + pull
        body.set_state(np.stack([rest[:, 0] * s, rest[:, 1] / s, rest[:, 2]], axis=1))
        body.flow()
    P, _ = body.plastic()
    largest = max(np.linalg.norm(np.linalg.inv(p) - np.eye(3)) for p in P)
    assert largest > 0.5
    assert largest <= 1.0 + 1e-9


def test_crush_stops_at_the_crush_limit():
    """dynamics_strainstep_tests.cc:3169-3196"""
    mat = clay(1.0e5, 1.0e9, crush_stress=1.0e3, crush_limit=0.7)
    mesh, body = one_cell(mat)
    for _ in range(10):
        body.set_state(0.6 * mesh["rest"])
        body.flow()
    P, _ = body.plastic()
    smallest = min(np.linalg.det(np.linalg.inv(p)) for p in P)
    assert smallest < 0.99
    assert smallest >= float(np.float32(0.7)) - 1e-9


def test_rest_fit_of_a_uniform_stretch_writes_the_same_stretch_back():
    """dynamics_strainstep_tests.cc:3198-3223"""
    mesh, body = one_cell(clay(1.0e5, 2.0e3))
    inv = np.linalg.inv(np.array([[1.1, 0.05, 0.0], [0.0, 0.9, 0.02], [0.01, 0.0, 1.0]]).T)
    body.set_plastic(np.tile(inv, (8, 1, 1)), np.zeros(8))
    shape, hg, _ = body.plastic_rest(0)
    body.write_fitted(0, shape, hg)
    P, _ = body.plastic()
    assert max(np.linalg.norm(p - inv) for p in P) <= 1e-12
    assert np.linalg.norm(hg, axis=1).max() <= 1e-12


def test_rest_fit_keeps_the_plastic_volume_and_is_idempotent():
    """dynamics_strainstep_tests.cc:3225-3263"""
    mesh, body = one_cell(clay(1.0e5, 2.0e3))
    P = []
    for p in range(8):
        side = H.corner_signs(p)[2]
        P.append(np.linalg.inv(np.diag([math.exp(0.2 * side), math.exp(-0.2 * side), 1.0])))
    body.set_plastic(np.array(P), np.zeros(8))
    body.write_fitted(0, *body.plastic_rest(0)[:2])
    W, _ = body.plastic()
    assert abs(np.mean([np.linalg.det(np.linalg.inv(p)) for p in W]) - 1.0) <= 1e-12
    for _ in range(5):
        body.write_fitted(0, *body.plastic_rest(0)[:2])
    W2, _ = body.plastic()
    assert max(np.linalg.norm(a - b) for a, b in zip(W, W2)) <= 1e-12


def test_resting_plastic_state_is_untouched_when_nothing_moves():
    """dynamics_strainstep_tests.cc:3265-3285"""
    _, body = sheared_cell(clay(1.0e5, 2.0e3))
    assert body.flow() > 0.0
    P0, a0 = body.plastic()
    assert body.flow() <= 1e-12
    P1, a1 = body.plastic()
    np.testing.assert_array_equal(P0, P1)
    np.testing.assert_array_equal(a0, a1)


def bent_strip(release_after=None, colored=False, frames=120):
    """A strip one cell thick, both ends held, its tip end driven down 0.5 m
    over a second (stripShape / 3329-3397, with the default damping ratio:
    the reference's strip uses damping ratio 1, whose lowest-mode damping is
    not restated)."""
    mat = jelly(1.0e5, density=10.0, poisson=0.3)
    mat.yield_stress, mat.plastic_max = 1.0e5 * 0.002, 1.0
    ends = [hold_at((-1.2, 0.0, 0.0), (0.01, 1.0, 1.0)), hold_at((1.2, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((1.2, 0.1, 0.1), 0.2, mat, ends, lanes=1)
    body = H.LatticeBody(mesh, [mat], offset=(0.0, 2.0, 0.0))
    tip = np.nonzero(mesh["rest"][:, 0] > 1.19)[0]
    y0 = body.state()[0][tip, 1].copy()
    held = None
    for frame in range(1, frames + 1):
        x, v = body.state()
        x[tip, 1] = y0 - 0.5 * min(frame / 60.0, 1.0)
        body.set_state(x, v)
        body.step(gravity=(0.0, 0.0, 0.0), colored=colored)
        if frame == 90:
            held = body.plastic()[1].sum() / 8.0
    return body, mesh, tip, y0, held


def test_strip_held_bent_past_yield_stops_flowing():
    """dynamics_strainstep_tests.cc:3329-3359"""
    body, _, _, _, held = bent_strip()
    final = body.plastic()[1].sum() / 8.0
    assert held > 0.0
    assert final - held < 0.01 * held


# ---- cell tearing ----------------------------------------------------------

def tearable():
    """tearable (dynamics_strainstep_tests.cc:905-911)"""
    m = jelly(5.0e3, density=20.0, poisson=0.3)
    m.tear_strain = 0.4
    return m


def orphans(body):
    hx, _, _ = body.topology()
    return body.nodes - np.unique(hx).size


def test_sym_eigen3_matches_numpy():
    rng = np.random.default_rng(3)
    for _ in range(50):
        A = rng.standard_normal((3, 3))
        A = A @ A.T if rng.random() < 0.5 else A + A.T
        w, V = H.sym_eigen3(A)
        ref = np.linalg.eigvalsh(A)
        np.testing.assert_allclose(np.sort(w), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        np.testing.assert_allclose(A @ V, V * w, atol=1e-12 * np.abs(ref).max())
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-13)
    w, V = H.sym_eigen3(np.diag([2.0, 5.0, 1.0]))  # diagonal: no rotation
    assert w.tolist() == [2.0, 5.0, 1.0] and V.tolist() == np.eye(3).tolist()


def test_cell_stretch_is_the_largest_singular_value_of_a_uniform_deformation():
    """stretchOf (tearing.cc:258-300): a uniform x -> F x gives every corner
    frame F, so the stretch is F's largest singular value and the direction
    its right singular vector; a rotation stretches nothing"""
    mat = tearable()
    mesh = box((0.4, 0.2, 0.2), 0.2, mat)
    rng = np.random.default_rng(8)
    for trial in range(6):
        F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        if np.linalg.det(F) < 0:
            F[:, 0] *= -1
        body = H.LatticeBody(mesh, [mat, mat])
        x, v = body.state()
        body.set_state(x @ F.T + 0.3, v)
        _, s, Vt = np.linalg.svd(F)
        for cell in (0, 3, 7):
            stretch, d = body.cell_stretch(cell)
            assert abs(stretch - s[0]) <= 1e-12 * s[0]
            assert abs(abs(d @ Vt[0]) - 1.0) <= 1e-9
            assert abs(d).max() == d.max()  # sign: largest component positive
    c, s_ = math.cos(0.7), math.sin(0.7)
    body = H.LatticeBody(mesh, [mat, mat])
    x, v = body.state()
    body.set_state(x @ np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]]).T, v)
    assert abs(body.cell_stretch(2)[0] - 1.0) <= 1e-12
    assert body.overstretched().size == 0


def test_overstretched_uses_one_plus_the_lane_tear_strain():
    """overstretched (tearing.cc:488-503): strictly past 1 + tearStrain, per
    lane; lanes without a finite tear strain never tear"""
    mat = tearable()
    mesh = box((0.4, 0.2, 0.2), 0.2, mat)
    for scale, torn in ((1.39, 0), (1.41, 16)):
        body = H.LatticeBody(mesh, [mat, mat])
        x, v = body.state()
        x[:, 0] *= scale
        body.set_state(x, v)
        assert body.overstretched().size == torn
    plain = jelly(5.0e3)
    body = H.LatticeBody(mesh, [plain, plain])
    x, v = body.state()
    x[:, 0] *= 3.0
    body.set_state(x, v)
    assert body.overstretched().size == 0


def test_tear_splits_the_nodes_two_halves_share():
    """dynamics_strainstep_tests.cc:3759-3796: a crack through the second
    cell of a strip one cell thick splits the four nodes the halves share"""
    mat = tearable()
    mesh = box((1.0, 0.25, 0.25), 0.5, mat)
    body = H.LatticeBody(mesh, [mat, mat], offset=(0.0, 1.0, 0.0))
    nodes = body.nodes
    mass = mesh["mass"].astype(np.float64).sum()
    assert mesh["hex"].shape[0] == 4 and body.pieces() == 1
    x, v = body.state()
    x[:, 0] *= 1.01
    body.set_state(x, v)
    assert body.tear([1]) == 4
    assert body.nodes == nodes + 4
    assert body.pieces() == 2
    hx, m, crack = body.topology()
    assert hx.shape == (4, 8)
    assert abs(m.astype(np.float64).sum() - mass) <= 1e-6 * mass
    # the crack runs between cells 1 and 2: cell 1's +x face, cell 2's -x face
    assert crack.tolist() == [0, 2, 1, 0]
    # the copies sit where their originals are
    xs, _ = body.state()
    assert xs.shape == (nodes + 4, 3)
    for _ in range(120):
        body.step()
    x, v = body.state()
    assert np.isfinite(x).all() and np.isfinite(v).all()
    assert orphans(body) == 0


def test_tear_never_chips_a_single_cell():
    """MIN_PIECE_CELLS (tearing.h:13): a crack through an end cell would cut
    off one cell alone, so nothing splits; a listed cell whose corners are
    all anchored cannot split either"""
    mat = tearable()
    mesh = box((1.0, 0.25, 0.25), 0.5, mat)
    body = H.LatticeBody(mesh, [mat, mat])
    x, v = body.state()
    x[:, 0] *= 1.01
    body.set_state(x, v)
    assert body.tear([0]) == 0 and body.tear([3]) == 0 and body.tear([]) == 0
    assert body.nodes == 20 and body.pieces() == 1
    crack = body.topology()[2]
    assert not crack.any()


def test_pieces_count_separate_blocks():
    """latticePieces (resolution.cc:198-228) of a lattice with a gap"""
    mat = tearable()
    solid = np.ones(4 * 1 * 1, dtype=np.uint8)
    solid[1] = 0
    mesh = H.build_hex_lattice([5, 2, 2], [0.5] * 3, [0.0] * 3, [mat], [], solid=solid)
    body = H.LatticeBody(mesh, [mat])
    assert body.pieces() == 2


def pulled_strip(frames=240):
    """dynamics_strainstep_tests.cc:3719-3757: a strip held at both ends, the
    right end drawn out 2.4 m over three seconds"""
    mat = tearable()
    ends = [hold_at((-1.2, 0.0, 0.0), (0.01, 1.0, 1.0)), hold_at((1.2, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((1.2, 0.2, 0.2), 0.2, mat, ends)
    body = H.LatticeBody(mesh, [mat, mat], offset=(0.0, 2.0, 0.0))
    right = np.nonzero(mesh["rest"][:, 0] > 1.19)[0]
    start = body.state()[0][right, 0].copy()
    first = None
    for frame in range(frames):
        x, v = body.state()
        x[right, 0] = start + 2.4 * min(frame / 180.0, 1.0)
        body.set_state(x, v)
        body.step()
        if first is None and body.nodes > mesh["rest"].shape[0]:
            first = frame
    return body, mesh, first


def longest_edge(body, cell):
    hx, _, _ = body.topology()
    x, _ = body.state()
    longest = 0.0
    for nd in hx:
        for c in range(8):
            for a in range(3):
                o = c ^ (1 << a)
                if o > c:
                    longest = max(longest, np.linalg.norm(x[nd[c]] - x[nd[o]]) / cell)
    return longest


def test_strip_pulled_apart_tears_into_pieces():
    """dynamics_strainstep_tests.cc:3719-3757"""
    body, mesh, first = pulled_strip()
    mass = mesh["mass"].astype(np.float64).sum()
    hx, m, crack = body.topology()
    assert first is not None
    assert body.pieces() >= 2
    assert hx.shape[0] == mesh["hex"].shape[0]
    assert abs(m.astype(np.float64).sum() - mass) <= 1e-6 * mass
    assert body.nodes > mesh["rest"].shape[0]
    assert orphans(body) == 0
    assert longest_edge(body, 0.2) < 1.8
    assert crack.any()
    x, v = body.state()
    assert np.isfinite(x).all() and np.isfinite(v).all()


# ---- lowest mode and damping ----------------------------------------------

def beam_lowest(poisson=0.0, held=True):
    mat = jelly(1.0e5, density=5.0, poisson=poisson)
    clamp = [hold_at((-2.0, 0.0, 0.0), (0.01, 1.0, 1.0))] if held else []
    mesh = box((2.0, 0.2, 0.2), 0.2, mat, clamp)
    body = H.LatticeBody(mesh, [mat, mat])
    inertia, area = 0.4 ** 4 / 12.0, 0.4 * 0.4
    root = 1.8751 if held else 4.7300
    theory = root * root * math.sqrt(1.0e5 * inertia / (5.0 * area * 4.0 ** 4))
    return body.lowest_angular(), theory, mesh["hex"].shape[0]


@pytest.mark.parametrize("poisson,held", [(0.0, True), (0.0, False), (0.4, True)])
def test_lowest_mode_of_a_beam_is_beam_theory(poisson, held):
    """dynamics_strainstep_tests.cc:1536-1563 (clamped), 1565-1592 (free),
    1614-1650 (clamped, volume row on): within 10 % of Euler-Bernoulli"""
    omega, theory, cells = beam_lowest(poisson, held)
    assert cells == 80
    assert abs(omega - theory) / theory < 0.1


def test_lowest_mode_of_a_free_cell_is_its_hourglass_mode():
    """dynamics_strainstep_tests.cc:1594-1612: 2 / side sqrt(E / 3 rho)"""
    mat = jelly(1.0e5, density=5.0, poisson=0.0)
    mesh = box((0.1, 0.1, 0.1), 0.2, mat)
    body = H.LatticeBody(mesh, [mat])
    expected = 2.0 / 0.2 * math.sqrt(1.0e5 / (3.0 * 5.0))
    assert abs(body.lowest_angular() - expected) <= 1e-6 * expected


def test_lowest_mode_of_a_pinned_lattice_is_infinite():
    """dynamics_strainstep_tests.cc:1652-1666"""
    mat = jelly(1.0e5, density=5.0)
    mesh = box((0.1, 0.1, 0.1), 0.2, mat, [hold_at((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))])
    assert mesh["anchored"].all()
    assert H.LatticeBody(mesh, [mat]).lowest_angular() == math.inf


@pytest.mark.parametrize("change", [None, (100, 32)])
def test_damped_cantilever_bends_as_beam_theory(change):
    """dynamics_strainstep_tests.cc:1852-1870: damping ratio 0.2, 16
    substeps, and 32 from frame 100 (the lowest frequency found at the first
    build is kept through the rebuild)"""
    sag, theory, cells = cantilever(0.4, 0.2, 0.3, 16, 600, damping=0.2, change=change)
    assert cells == 80
    assert abs(sag - theory) / theory < 0.1


def test_a_damped_lattice_finds_its_lowest_frequency_once():
    """buildElements (strainstep.cc:300-331): the first build of a damped
    lattice finds the lowest frequency of its undamped elements; rebuilds for
    other substeps keep it; an undamped lattice never looks for it"""
    mat = jelly(5.0e5, density=5.0, poisson=0.3)
    mat.damping_ratio = 0.2
    clamp = [hold_at((-2.0, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((2.0, 0.2, 0.2), 0.2, mat, clamp)
    body = H.LatticeBody(mesh, [mat, mat])
    assert body.lowest() == 0.0
    first = body.lowest_angular()
    body.step()
    assert body.lowest() == first > 0.0
    for _ in range(5):
        body.step()
    body.set_substeps(32)
    body.step()
    assert body.lowest() == first
    plain = H.LatticeBody(mesh, [jelly(5.0e5, density=5.0, poisson=0.3)] * 2)
    plain.step()
    assert plain.lowest() == 0.0


FACE_CORNERS = [(0, 2, 4, 6), (1, 3, 5, 7), (0, 1, 4, 5), (2, 3, 6, 7), (0, 1, 2, 3), (4, 5, 6, 7)]


def unshared_face_nodes(hx):
    """nodes of the cell faces no other cell shares, from the corner lists"""
    count = {}
    for nd in hx:
        for f in FACE_CORNERS:
            key = tuple(sorted(int(nd[c]) for c in f))
            count[key] = count.get(key, 0) + 1
    return sorted({n for key, k in count.items() if k == 1 for n in key})


def test_surface_nodes_of_a_lattice_with_holes():
    """buildMesh's surface nodes (solver.cc:2369-2413): the corners of every
    face whose neighbour cell is missing or off the grid"""
    mat = tearable()
    res = (6, 5, 4)
    cells = [r - 1 for r in res]
    rng = np.random.default_rng(2)
    solid = (rng.random(int(np.prod(cells))) > 0.25).astype(np.uint8)
    mesh = H.build_hex_lattice(list(res), [0.1] * 3, [0.0] * 3, [mat], [], solid=solid)
    body = H.LatticeBody(mesh, [mat])
    # the grid rule, from the solid mask: node ids by position
    pos = {tuple(np.round(p / 0.1).astype(int)): i for i, p in enumerate(mesh["rest"])}
    expect = set()
    for z in range(cells[2]):
        for y in range(cells[1]):
            for x in range(cells[0]):
                if not solid[x + cells[0] * (y + cells[1] * z)]:
                    continue
                for axis in range(3):
                    for side in (0, 1):
                        at = [x, y, z]
                        at[axis] += 1 if side else -1
                        off = not (0 <= at[axis] < cells[axis])
                        if off or not solid[at[0] + cells[0] * (at[1] + cells[1] * at[2])]:
                            for c in range(8):
                                if ((c >> axis) & 1) == side:
                                    corner = (x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1))
                                    expect.add(pos[corner])
    assert body.surface().tolist() == sorted(expect)
    full = H.LatticeBody(box((0.4, 0.3, 0.2), 0.2, mat), [mat])
    assert full.surface().size == 5 * 4 * 3 - 3 * 2 * 1  # every node but the interior ones


@pytest.mark.parametrize("cells", [[1], [5, 6, 9]])
def test_tear_puts_crack_faces_and_copies_on_the_surface(cells):
    """markCracks (tearing.cc:426-484; dynamics_strainstep_tests.cc:3786-3789:
    every copy of the split strip is a surface node): after a tear the surface
    is exactly the nodes of the faces no cell shares any more"""
    mat = tearable()
    mesh = box((1.0, 0.25, 0.25), 0.5, mat) if cells == [1] else box((0.6, 0.3, 0.2), 0.2, mat)
    body = H.LatticeBody(mesh, [mat, mat])
    nodes = body.nodes
    x, v = body.state()
    x[:, 0] *= 1.3
    body.set_state(x, v)
    made = body.tear(cells)
    assert made > 0
    surface = body.surface().tolist()
    assert all(c in surface for c in range(nodes, nodes + made))
    assert surface == unshared_face_nodes(body.topology()[0])
