"""Hex-cell tearing of strainstep lattices (talus_mi_lattice_tear, the
overstretched flags of talus_mi_lattice_step) against the hex oracle's
restatement of src/talus/dynamics/strainstep/tearing.cc: the split of the
same cells at the same positions is bit-exact (nodes, corner lists, masses,
crack faces, copied states), and the reference's strip tests
(tests/dynamics_strainstep_tests.cc:3719-3796) hold on the GPU with the
oracle tearing at the same step into the same pieces.  FP64."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import hex as H
from talus_amd import api

pytestmark = pytest.mark.gpu


def tearable():
    """tearable (dynamics_strainstep_tests.cc:905-911)"""
    return api.PhysicalMaterial(youngs_modulus=5.0e3, density=20.0, poissons_ratio=0.3, friction=0.7,
                                tear_strain=0.4)


def orc_material(m):
    o = oracle.Material()
    for name, _ in oracle.Material._fields_:
        setattr(o, name, getattr(m, name))
    return o


def box(half, cell, mat, supports=(), lanes=2):
    cells = [int(round(2 * h / cell)) for h in half]
    return api.build_hex_lattice([c + 1 for c in cells], [cell] * 3, [-h for h in half], [mat] * lanes,
                                 list(supports))


def orc_mesh(mesh):
    return {"rest": mesh["rest_positions"].astype(np.float64), "mass": mesh["node_mass"],
            "anchored": mesh["anchored"], "hex": mesh["hex_nodes"], "volume": mesh["volume"],
            "lane": mesh["hex_material"], "voxel": mesh["voxel_size"]}


def same_topology(ctx, b, orc):
    hx, m, crack = ctx.lattice_topology(b)
    ohx, om, ocrack = orc.topology()
    np.testing.assert_array_equal(hx, ohx)
    np.testing.assert_array_equal(m, om)
    np.testing.assert_array_equal(crack, ocrack)
    np.testing.assert_array_equal(ctx.lattice_surface(b), orc.surface())
    assert ctx.lattice_pieces(b) == orc.pieces()


def test_tear_through_a_strip_one_cell_thick_matches_the_oracle():
    """dynamics_strainstep_tests.cc:3759-3796 on the library, bit-exact with
    the oracle, then 120 steps of the torn strip on the GPU"""
    mat = tearable()
    mesh = box((1.0, 0.25, 0.25), 0.5, mat)
    mass = mesh["node_mass"].astype(np.float64).sum()
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)] * 2, offset=(0.0, 1.0, 0.0))
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat, mat], offset=(0.0, 1.0, 0.0))
        nodes = ctx.lattice_nodes(b)
        x, v = ctx.lattice_state(b)
        x[:, 0] *= 1.01
        ctx.set_lattice_state(b, x, v)
        orc.set_state(x, v)
        s, d = ctx.lattice_cell_stretch(b, 1)
        os_, od = orc.cell_stretch(1)
        assert s == os_ and d.tolist() == od.tolist()
        assert abs(s - 1.01) <= 1e-12 and d.tolist() == [1.0, 0.0, 0.0]
        assert ctx.lattice_tear(b, [1]) == 4 == orc.tear([1])
        assert ctx.lattice_nodes(b) == nodes + 4
        assert ctx.lattice_pieces(b) == 2
        same_topology(ctx, b, orc)
        hx, m, crack = ctx.lattice_topology(b)
        assert abs(m.astype(np.float64).sum() - mass) <= 1e-6 * mass
        assert crack[1] != 0 and crack[2] != 0
        xs, vs = ctx.lattice_state(b)
        ox, ov = orc.state()
        np.testing.assert_array_equal(xs, ox)
        np.testing.assert_array_equal(vs, ov)
        for frame in range(120):
            r = ctx.lattice_step()
            orc.step(colored=True)
            if frame == 0:
                assert r["rebuilds"] == 1
        x, v = ctx.lattice_state(b)
        ox, ov = orc.state()
        assert np.isfinite(x).all() and np.isfinite(v).all()
        assert np.unique(ctx.lattice_topology(b)[0]).size == ctx.lattice_nodes(b)  # no orphans
        assert float(np.abs(x - ox).max()) <= 1e-8
        assert float(np.abs(v - ov).max()) <= 1e-6


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tear_of_random_cells_matches_the_oracle_bitwise(seed):
    """tearCells over several cells of a deformed block: cracks, pinched
    splits, masses and crack faces equal the oracle's, node for node"""
    rng = np.random.default_rng(seed)
    mat = tearable()
    mesh = box((0.6, 0.3, 0.2), 0.1, mat)
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)] * 2)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat, mat])
        x, v = ctx.lattice_state(b)
        A = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
        if np.linalg.det(A) < 0:
            A[:, 0] *= -1.0
        x = x @ A.T + 0.02 * np.sin(7.0 * x[:, [1, 2, 0]])
        v = rng.standard_normal(v.shape)
        ctx.set_lattice_state(b, x, v)
        orc.set_state(x, v)
        hexes = mesh["hex_nodes"].shape[0]
        for c in rng.choice(hexes, 5, replace=False):
            s, d = ctx.lattice_cell_stretch(b, int(c))
            os_, od = orc.cell_stretch(int(c))
            assert s == os_ and d.tolist() == od.tolist()
        cells = rng.choice(hexes, 12, replace=False).astype(np.uint32)
        made = ctx.lattice_tear(b, cells)
        assert made == orc.tear(cells)
        assert made > 0
        assert ctx.lattice_nodes(b) == orc.nodes
        same_topology(ctx, b, orc)
        xs, vs = ctx.lattice_state(b)
        ox, ov = orc.state()
        np.testing.assert_array_equal(xs, ox)
        np.testing.assert_array_equal(vs, ov)
        # the torn block steps on: impulses and states stay close to the oracle's
        for _ in range(3):
            ctx.lattice_step(gravity=(0.0, 0.0, 0.0))
            orc.step(gravity=(0.0, 0.0, 0.0), colored=True)
        x, _ = ctx.lattice_state(b)
        ox, _ = orc.state()
        assert float(np.abs(x - ox).max()) <= 1e-8


def test_strip_pulled_apart_tears_like_the_oracle():
    """dynamics_strainstep_tests.cc:3719-3757: both ends held, the right end
    drawn out 2.4 m over three seconds; the GPU's overstretched flags tear the
    same cells at the same step as the oracle, into the same pieces"""
    mat = tearable()
    ends = [api.Support.box((-1.2, 0.0, 0.0), (0.01, 1.0, 1.0)), api.Support.box((1.2, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((1.2, 0.2, 0.2), 0.2, mat, ends)
    cells, nodes = mesh["hex_nodes"].shape[0], mesh["node_mass"].size
    mass = mesh["node_mass"].astype(np.float64).sum()
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)] * 2, offset=(0.0, 2.0, 0.0))
    right = np.nonzero(mesh["rest_positions"][:, 0] > 1.19)[0]
    torn_steps = []
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat, mat], offset=(0.0, 2.0, 0.0))
        start = ctx.lattice_state(b)[0][right, 0].copy()
        for frame in range(240):
            x, v = ctx.lattice_state(b)
            x[right, 0] = start + 2.4 * min(frame / 180.0, 1.0)
            ctx.set_lattice_state(b, x, v)
            ox, ov = orc.state()
            ox[right, 0] = x[right, 0]
            orc.set_state(ox, ov)
            r = ctx.lattice_step()
            before = orc.nodes
            orc.step(colored=True)
            assert (r["node_copies"] > 0) == (orc.nodes > before), frame
            if r["node_copies"]:
                torn_steps.append(frame)
                assert ctx.lattice_nodes(b) == orc.nodes, frame
                same_topology(ctx, b, orc)
        x, v = ctx.lattice_state(b)
        hx, m, crack = ctx.lattice_topology(b)
        assert torn_steps
        assert ctx.lattice_pieces(b) >= 2
        assert hx.shape[0] == cells
        assert abs(m.astype(np.float64).sum() - mass) <= 1e-6 * mass
        assert ctx.lattice_nodes(b) > nodes
        assert np.unique(hx).size == ctx.lattice_nodes(b)
        longest = 0.0
        for nd in hx:
            for c in range(8):
                for a in range(3):
                    o = c ^ (1 << a)
                    if o > c:
                        longest = max(longest, np.linalg.norm(x[nd[c]] - x[nd[o]]) / 0.2)
        assert longest < 1.8
        assert np.isfinite(x).all() and np.isfinite(v).all()
        ox, _ = orc.state()
        assert float(np.abs(x - ox).max()) <= 1e-6


def test_a_tear_keeps_the_plastic_refresh_of_the_other_bodies():
    """A tear lays the device arrays out again; a plastic body stepped in the
    same context whose cells flowed in that step still has their elements
    rebuilt (refreshFlowed) at the next step, bit for bit as when it is
    stepped alone"""
    from test_gpu_lattice import clay_block
    clay_mesh, clay = clay_block()
    mat = tearable()
    strip = box((1.0, 0.25, 0.25), 0.5, mat)
    states = []
    for with_strip in (False, True):
        with api.Context() as ctx:
            s = ctx.add_lattice(strip, [mat, mat], offset=(0.0, 5.0, 0.0)) if with_strip else None
            b = ctx.add_lattice(clay_mesh, [clay])
            refreshed = 0
            for frame in range(30):
                r = ctx.lattice_step()
                refreshed += r["refreshed"]
                if with_strip and frame == 10:
                    assert ctx.lattice_plastic(b) is not None
                    x, v = ctx.lattice_state(s)
                    x[:, 0] *= 1.01  # a stretch along the strip sets the crack across it
                    ctx.set_lattice_state(s, x, v)
                    assert ctx.lattice_tear(s, [1]) == 4
            states.append((ctx.lattice_state(b), ctx.lattice_plastic(b), refreshed))
    (x0, v0), (P0, a0), n0 = states[0]
    (x1, v1), (P1, a1), n1 = states[1]
    assert n0 > 0 and n0 == n1
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(v1, v0)
    np.testing.assert_array_equal(P1, P0)
    np.testing.assert_array_equal(a1, a0)


def test_damped_plastic_strip_tears_like_the_oracle():
    """Damping, plastic flow and tearing together: a damped (ratio 0.3)
    yielding tearable strip drawn out as in dynamics_strainstep_tests.cc:
    3719-3757; the GPU yields, rebuilds and tears as the oracle's colour
    sweep does, and keeps the lowest frequency of its first build"""
    mat = tearable()
    mat.damping_ratio, mat.yield_stress, mat.plastic_max = 0.3, 200.0, 1.0
    ends = [api.Support.box((-1.2, 0.0, 0.0), (0.01, 1.0, 1.0)), api.Support.box((1.2, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((1.2, 0.2, 0.2), 0.2, mat, ends)
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)] * 2, offset=(0.0, 2.0, 0.0))
    right = np.nonzero(mesh["rest_positions"][:, 0] > 1.19)[0]
    tears = refreshed = 0
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat, mat], offset=(0.0, 2.0, 0.0))
        start = ctx.lattice_state(b)[0][right, 0].copy()
        for frame in range(120):
            x, v = ctx.lattice_state(b)
            x[right, 0] = start + 2.4 * min(frame / 90.0, 1.0)
            ctx.set_lattice_state(b, x, v)
            ox, ov = orc.state()
            ox[right, 0] = x[right, 0]
            orc.set_state(ox, ov)
            r = ctx.lattice_step()
            before = orc.nodes
            orc.step(colored=True)
            refreshed += r["refreshed"]
            assert (r["node_copies"] > 0) == (orc.nodes > before), frame
            if r["node_copies"]:
                tears += 1
                same_topology(ctx, b, orc)
        assert ctx.lattice_lowest(b) == orc.lowest() > 0.0
        assert tears > 0 and refreshed > 0 and ctx.lattice_pieces(b) >= 2
        x, _ = ctx.lattice_state(b)
        ox, _ = orc.state()
        P, a = ctx.lattice_plastic(b)
        oP, oa = orc.plastic()
    assert float(np.abs(x - ox).max()) <= 1e-6
    assert float(np.abs(a - oa).max()) <= 1e-6 * max(1.0, float(np.abs(oa).max()))
    assert P.shape == oP.shape
    assert float(np.abs(P - oP).max()) <= 1e-6 * max(1.0, float(np.abs(oP).max()))
