"""Strainstep hex lattices on the GPU (talus_mi_lattice_*), against the hex
oracle stepping the same bodies in the same (colour, cell) order, and against
the closed forms the reference's strainstep tests hold
(tests/dynamics_strainstep_tests.cc): free fall as semi-implicit Euler
(156-183, 1071-1077) and cantilevers within 10 % of beam theory (1806-1820).
Bodies batched in one launch sequence equal each body stepped alone, bit for
bit (the reference's parallel = serial identity, 4998-5050).  FP64."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import hex as H
from talus_amd import api

pytestmark = pytest.mark.gpu


def material(E, density=1000.0, poisson=0.4, friction=0.7):
    return api.PhysicalMaterial(youngs_modulus=E, density=density, poissons_ratio=poisson,
                                friction=friction)


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


def rel(a, b, scale):
    return float(np.abs(a - b).max() / scale)


@pytest.mark.parametrize("requested,expected", [(0, 4), (8, 8)])
def test_free_fall_is_semi_implicit_euler(requested, expected):
    mat = material(8.0e3)
    mesh = box((0.3, 0.3, 0.3), 0.15, mat)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat, mat], offset=(0.0, 50.0, 0.0))
        ctx.set_lattice_substeps(b, requested)
        x0, _ = ctx.lattice_state(b)
        for _ in range(30):
            assert ctx.lattice_step()["substeps"] == expected
        x, v = ctx.lattice_state(b)
    h = (1.0 / 60.0) / expected
    n = expected * 30
    assert np.abs(v[:, 1] + 9.81 * h * n).max() <= 1e-9
    assert np.abs(x[:, 1] - x0[:, 1] + 9.81 * h * h * n * (n + 1) / 2).max() <= 1e-9
    assert np.abs(x[:, [0, 2]] - x0[:, [0, 2]]).max() <= 1e-12


def sagging_block(E=2.0e5, poisson=0.3, cells=(8, 2, 2), cell=0.1):
    mat = material(E, poisson=poisson)
    half = [c * cell / 2 for c in cells]
    clamp = [api.Support.box((-half[0], 0.0, 0.0), (0.01, 1.0, 1.0))]
    return box(half, cell, mat, clamp, lanes=1), mat


@pytest.mark.parametrize("steps", [1, 40])
def test_gpu_matches_the_oracle_colour_sweep(steps):
    mesh, mat = sagging_block()
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)])
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat])
        nc, hc = ctx.lattice_colours(b)
        onc, ohc = orc.colours()
        np.testing.assert_array_equal(nc, onc)
        np.testing.assert_array_equal(hc, ohc)
        for _ in range(steps):
            r = ctx.lattice_step()
            assert orc.step(colored=True) == r["substeps"]
        x, v = ctx.lattice_state(b)
        cells, imp = ctx.lattice_impulses(b)
    ox, ov = orc.state()
    ocells, oimp = orc.impulses()
    np.testing.assert_array_equal(cells, ocells)
    motion = np.abs(ox - mesh["rest_positions"]).max()
    assert motion > 1e-5
    assert rel(x, ox, motion) <= 1e-9
    assert rel(v, ov, np.abs(ov).max()) <= 1e-9
    assert rel(imp, oimp, np.abs(oimp).max()) <= 1e-9


def test_gpu_matches_the_oracle_on_a_turned_moving_jelly():
    """A block thrown with a spin: large rotations through rotationOf."""
    mat = material(5.0e4, poisson=0.45)
    mesh = box((0.3, 0.2, 0.2), 0.1, mat)
    rest = mesh["rest_positions"].astype(np.float64)
    w = np.array([0.0, 0.0, 6.0])
    v0 = np.cross(w, rest) + [1.0, 3.0, 0.0]
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)] * 2)
    orc.set_state(rest, v0)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat, mat])
        ctx.set_lattice_state(b, rest, v0)
        for _ in range(30):
            ctx.lattice_step()
            orc.step(colored=True)
        x, v = ctx.lattice_state(b)
    ox, ov = orc.state()
    assert rel(x, ox, np.abs(ox - rest).max()) <= 1e-9
    assert rel(v, ov, np.abs(ov).max()) <= 1e-9


def cantilever(ctx, thickness, cell, poisson, substeps, steps):
    length = 4.0
    mat = material(5.0e5, density=5.0, poisson=poisson, friction=0.5)
    clamp = [api.Support.box((-length / 2, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((length / 2, thickness / 2, thickness / 2), cell, mat, clamp)
    b = ctx.add_lattice(mesh, [mat, mat])
    ctx.set_lattice_substeps(b, substeps)
    rest = mesh["rest_positions"]
    tip = int(np.argmin(np.linalg.norm(rest - [length / 2, thickness / 2, thickness / 2], axis=1)))
    return b, mesh, tip, 5.0 * 9.81 * thickness ** 2 * length ** 4 / (8 * 5.0e5 * thickness ** 4 / 12.0)


@pytest.mark.parametrize("thickness,cell,poisson,substeps,cells",
                         [(0.4, 0.4, 0.0, 48, 10), (0.4, 0.4, 0.45, 48, 10), (0.4, 0.2, 0.3, 16, 80)])
def test_cantilever_bends_as_beam_theory(thickness, cell, poisson, substeps, cells):
    """dynamics_strainstep_tests.cc:1806-1820; sag averaged over the last
    300 of 600 frames."""
    steps = 600
    with api.Context() as ctx:
        b, mesh, tip, theory = cantilever(ctx, thickness, cell, poisson, substeps, steps)
        assert mesh["hex_nodes"].shape[0] == cells
        sag = 0.0
        for frame in range(1, steps + 1):
            ctx.lattice_step()
            if frame > steps // 2:
                x, _ = ctx.lattice_state(b)
                sag += (mesh["rest_positions"][tip, 1] - x[tip, 1]) / (steps // 2)
    assert abs(sag - theory) / theory < 0.1


def test_batched_bodies_equal_each_body_alone_bitwise():
    """Uneven bodies with different substep counts in one launch sequence."""
    specs = [((0.3, 0.2, 0.2), 0.1, 2.0e5, 0.3, 0), ((0.2, 0.2, 0.2), 0.1, 8.0e3, 0.4, 8),
             ((0.5, 0.1, 0.1), 0.1, 5.0e6, 0.45, 0)]
    bodies = []
    for half, cell, E, nu, sub in specs:
        mat = material(E, poisson=nu)
        clamp = [api.Support.box((-half[0], 0.0, 0.0), (0.01, 1.0, 1.0))]
        bodies.append((box(half, cell, mat, clamp, lanes=1), mat, sub))

    def run(which):
        with api.Context() as ctx:
            ids = []
            for k in which:
                mesh, mat, sub = bodies[k]
                i = ctx.add_lattice(mesh, [mat], offset=(0.0, 1.0 * k, 0.0))
                ctx.set_lattice_substeps(i, sub)
                ids.append(i)
            subs = [ctx.lattice_step()["substeps"] for _ in range(12)]
            return [ctx.lattice_state(i) for i in ids], subs

    together, subs = run([0, 1, 2])
    assert max(subs) >= 8
    for k in range(3):
        alone, _ = run([k])
        np.testing.assert_array_equal(together[k][0], alone[0][0])
        np.testing.assert_array_equal(together[k][1], alone[0][1])


def test_state_round_trips_and_unsupported_materials_are_refused():
    mesh, mat = sagging_block()
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat])
        x, v = ctx.lattice_state(b)
        np.testing.assert_array_equal(x, mesh["rest_positions"].astype(np.float64))
        ctx.lattice_step()
        x1, v1 = ctx.lattice_state(b)
        ctx.set_lattice_state(b, x1 + 0.5, v1)
        x2, _ = ctx.lattice_state(b)
        np.testing.assert_array_equal(x2, x1 + 0.5)
        with pytest.raises(api.TalusError):
            ctx.add_lattice(mesh, [])


# ---- plasticity (flowLattice, refreshFlowed) --------------------------------

def clay_block(E=2.0e5, yield_stress=1.5e3, cells=(8, 2, 2), cell=0.1, hardening=0.0, cap=1.0):
    mat = material(E, poisson=0.35)
    mat.yield_stress, mat.hardening_modulus, mat.plastic_max = yield_stress, hardening, cap
    half = [c * cell / 2 for c in cells]
    clamp = [api.Support.box((-half[0], 0.0, 0.0), (0.01, 1.0, 1.0))]
    return box(half, cell, mat, clamp, lanes=1), mat


@pytest.mark.parametrize("hardening", [0.0, 5.0e4])
def test_plastic_sag_matches_the_oracle(hardening):
    """A clay beam held at one end sags past its yield: positions, velocities,
    impulses, per-point P and plastic strain follow the oracle's colour sweep,
    through the flow after each step and the element rebuilds before the next."""
    mesh, mat = clay_block(hardening=hardening)
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)])
    steps = 40
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat])
        refreshed = 0
        for _ in range(steps):
            r = ctx.lattice_step()
            refreshed += r["refreshed"]
            orc.step(colored=True)
        x, v = ctx.lattice_state(b)
        cells, imp = ctx.lattice_impulses(b)
        P, a = ctx.lattice_plastic(b)
    ox, ov = orc.state()
    oP, oa = orc.plastic()
    _, oimp = orc.impulses()
    assert a.sum() > 0.0 and refreshed > 0  # it flowed and elements were rebuilt
    motion = np.abs(ox - mesh["rest_positions"]).max()
    assert rel(x, ox, motion) <= 1e-8
    assert rel(v, ov, np.abs(ov).max()) <= 1e-8
    assert rel(imp, oimp, np.abs(oimp).max()) <= 1e-8
    assert np.abs(P - oP).max() <= 1e-9
    assert rel(a, oa, oa.max()) <= 1e-8


def test_plastic_state_round_trips_and_drives_a_refresh():
    mesh, mat = clay_block()
    rng = np.random.default_rng(3)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat])
        assert ctx.lattice_plastic(b) is None
        ctx.lattice_step()
        P, a = ctx.lattice_plastic(b)
        P2 = P + 0.01 * rng.standard_normal(P.shape)
        ctx.set_lattice_plastic(b, P2, a + 0.5)
        Q, a2 = ctx.lattice_plastic(b)
        np.testing.assert_array_equal(Q, P2)
        np.testing.assert_array_equal(a2, a + 0.5)
        assert ctx.lattice_step()["refreshed"] > 0


def test_strip_held_bent_past_yield_stops_flowing_on_the_gpu():
    """dynamics_strainstep_tests.cc:3329-3359 (the strip with the default
    damping ratio), GPU against the oracle's plastic strain"""
    mat = material(1.0e5, density=10.0, poisson=0.3, friction=0.5)
    mat.yield_stress, mat.plastic_max = 1.0e5 * 0.002, 1.0
    ends = [api.Support.box((-1.2, 0.0, 0.0), (0.01, 1.0, 1.0)), api.Support.box((1.2, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((1.2, 0.1, 0.1), 0.2, mat, ends, lanes=1)
    tip = np.nonzero(mesh["rest_positions"][:, 0] > 1.19)[0]
    orc = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)], offset=(0.0, 2.0, 0.0))
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat], offset=(0.0, 2.0, 0.0))
        y0 = ctx.lattice_state(b)[0][tip, 1].copy()
        held = None
        for frame in range(1, 121):
            x, v = ctx.lattice_state(b)
            x[tip, 1] = y0 - 0.5 * min(frame / 60.0, 1.0)
            ctx.set_lattice_state(b, x, v)
            ox, ov = orc.state()
            ox[tip, 1] = y0 - 0.5 * min(frame / 60.0, 1.0)
            orc.set_state(ox, ov)
            ctx.lattice_step(gravity=(0.0, 0.0, 0.0))
            orc.step(gravity=(0.0, 0.0, 0.0), colored=True)
            if frame == 90:
                held = ctx.lattice_plastic(b)[1].sum() / 8.0
        final = ctx.lattice_plastic(b)[1].sum() / 8.0
        x, _ = ctx.lattice_state(b)
    assert held > 0.0
    assert final - held < 0.01 * held
    oa = orc.plastic()[1].sum() / 8.0
    assert abs(final - oa) <= 1e-6 * oa
    ox, _ = orc.state()
    assert np.abs(x - ox).max() <= 1e-6


def damped_cantilever(substeps, steps, change=None, ctx=None, orc=None):
    """cantilever (dynamics_strainstep_tests.cc:305-360) with damping ratio
    0.2 (1852-1870): a 4 m beam 0.4 thick in 0.2 m cells held at x = -2; tip
    sag averaged over the last half of the frames"""
    mat = material(5.0e5, density=5.0, poisson=0.3, friction=0.5)
    mat.damping_ratio = 0.2
    clamp = [api.Support.box((-2.0, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh = box((2.0, 0.2, 0.2), 0.2, mat, clamp)
    b = ctx.add_lattice(mesh, [mat, mat])
    ctx.set_lattice_substeps(b, substeps)
    o = None
    if orc:
        o = H.LatticeBody(orc_mesh(mesh), [orc_material(mat)] * 2)
        o.set_substeps(substeps)
    rest = mesh["rest_positions"].astype(np.float64)
    tip = int(np.argmin(np.linalg.norm(rest - [2.0, 0.2, 0.2], axis=1)))
    window, sag = steps // 2, 0.0
    for frame in range(1, steps + 1):
        if change is not None and frame == change[0]:
            ctx.set_lattice_substeps(b, change[1])
            if o is not None:
                o.set_substeps(change[1])
        r = ctx.lattice_step()
        if o is not None:
            o.step(colored=True)
        if frame > steps - window:
            sag += (rest[tip, 1] - ctx.lattice_state(b)[0][tip, 1]) / window
    inertia = 0.4 ** 4 / 12.0
    theory = 5.0 * 9.81 * 0.4 ** 2 * 4.0 ** 4 / (8 * 5.0e5 * inertia)
    return b, o, sag, theory, r


def test_damped_lattice_matches_the_oracle():
    """cellDamping at the lowest frequency (strainstep.cc:151-160,
    lowestmode.cc): the library finds the oracle's lowest frequency bit for
    bit and steps the damped cantilever as the oracle's colour sweep does"""
    with api.Context() as ctx:
        b, o, _, _, r = damped_cantilever(16, 40, change=(20, 32), ctx=ctx, orc=True)
        assert r["substeps"] == 32
        assert ctx.lattice_lowest(b) == o.lowest() > 0.0
        x, v = ctx.lattice_state(b)
        ox, ov = o.state()
        cells, imp = ctx.lattice_impulses(b)
        ocells, oimp = o.impulses()
    np.testing.assert_array_equal(cells, ocells)
    assert rel(x, ox, 1.0) <= 1e-9
    assert rel(v, ov, max(1e-3, float(np.abs(ov).max()))) <= 1e-8
    assert rel(imp, oimp, float(np.abs(oimp).max())) <= 1e-8


@pytest.mark.parametrize("change", [None, (100, 32)])
def test_damped_cantilever_bends_as_beam_theory(change):
    """dynamics_strainstep_tests.cc:1852-1870 on the GPU"""
    with api.Context() as ctx:
        b, _, sag, theory, r = damped_cantilever(16, 600, change=change, ctx=ctx)
        assert r["substeps"] == (32 if change else 16)
    assert abs(sag - theory) / theory < 0.1
