"""The FP64 neo-Hookean hex terms, the energy ledger and the screen solve of
a hex lattice on the GPU (talus_mi_lattice_snh_terms / _energy / _rest_body
/ _screen) against the oracle (oracle/screen_oracle.c, pinned by
tests/test_screen_cpu.py): terms and energies to 1e-12, the screen
displacement to 1e-8 of the oracle's dense factor, and the reference's own
screen checks (dynamics_tests.cc:2802-2910) on the device result."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import screen as S
from talus_amd import api

pytestmark = pytest.mark.gpu


def material(E, poisson=0.4, density=1000.0):
    return api.PhysicalMaterial(youngs_modulus=E, density=density, poissons_ratio=poisson, friction=0.7)


def orc_material(m):
    o = oracle.Material()
    for name, _ in oracle.Material._fields_:
        setattr(o, name, getattr(m, name))
    return o


def orc_mesh(mesh):
    return {"rest": mesh["rest_positions"].astype(np.float64), "mass": mesh["node_mass"],
            "anchored": mesh["anchored"], "hex": mesh["hex_nodes"], "volume": mesh["volume"],
            "lane": mesh["hex_material"], "voxel": mesh["voxel_size"]}


def lattice(cells=(4, 3, 2), cell=0.025, mats=None, supports=(), lanes=None):
    mats = mats or [material(2.0e5)]
    mesh = api.build_hex_lattice([c + 1 for c in cells], [cell] * 3, (0.0, 0.0, 0.0), mats, list(supports))
    if lanes is not None:
        mesh["hex_material"] = lanes.astype(np.uint32)
    return mesh, mats


def oracle_terms(om, shear, shifted, x):
    """per node sums of nodeTerms and per cell ledger energies"""
    half = om["voxel"].astype(np.float64) / 2
    n = om["rest"].shape[0]
    g, Hs = np.zeros((n, 3)), np.zeros((n, 3, 3))
    energy = np.empty(len(om["hex"]))
    for i, h in enumerate(om["hex"]):
        d, r = x[h], om["rest"][h]
        V = float(om["volume"][i])
        for c in range(8):
            gc, hc = S.node_terms(d, r, half, V, shear[i], shifted[i], c)
            g[h[c]] += gc
            Hs[h[c]] += hc
        shift = d[0] - r[0]
        energy[i] = S.cell_energy(d - shift, r, half, V, shear[i], shifted[i])
    return energy, g, Hs


def test_terms_match_the_oracle():
    rng = np.random.default_rng(4)
    mats = [material(2.0e5), material(9.0e5, poisson=0.3)]
    mesh, _ = lattice(mats=mats, lanes=rng.integers(0, 3, 24))  # one lane past the table
    om = orc_mesh(mesh)
    shear, shifted = S.cell_moduli(om, [orc_material(m) for m in mats])
    X = om["rest"]
    turn = np.array([[np.cos(0.4), -np.sin(0.4), 0.0], [np.sin(0.4), np.cos(0.4), 0.0], [0.0, 0.0, 1.0]])
    x = X @ turn.T + np.array([0.5, 1.0, -0.2]) + 2e-3 * rng.standard_normal(X.shape)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, mats)
        e, g, Hs = ctx.lattice_snh_terms(b, x)
    eo, go, Ho = oracle_terms(om, shear, shifted, x)
    assert np.abs(e - eo).max() <= 1e-12 * np.abs(eo).max()
    assert np.abs(g - go).max() <= 1e-12 * np.abs(go).max()
    assert np.abs(Hs - Ho).max() <= 1e-12 * np.abs(Ho).max()
    # a translation-invariant energy: the node gradients sum to zero
    assert np.abs(g.sum(axis=0)).max() <= 1e-10 * np.abs(g).max()


def test_ledger_matches_the_oracle_after_steps():
    mat = material(3.0e5)
    sup = [api.Support.box((0.0, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh, mats = lattice(cells=(6, 2, 2), mats=[mat], supports=sup)
    om = orc_mesh(mesh)
    shear, shifted = S.cell_moduli(om, [orc_material(mat)])
    g = (0.0, -9.81, 0.0)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, mats, offset=(0.0, 2.0, 0.0))
        for _ in range(10):
            ctx.lattice_step()
        x, v = ctx.lattice_state(b)
        led = ctx.lattice_energy(b, g)
    ke, pe, el = S.lattice_energy(om, shear, shifted, x, v, np.array(g))
    assert el > 0 and ke > 0
    assert led["kinetic"] == pytest.approx(ke, rel=1e-12)
    assert led["gravity"] == pytest.approx(pe, rel=1e-12)
    assert led["elastic"] == pytest.approx(el, rel=1e-12)


def test_rest_body_matches_the_oracle():
    mesh, mats = lattice(cells=(5, 2, 3), cell=0.02)
    om = orc_mesh(mesh)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, mats)
        rb = ctx.lattice_rest_body(b)
    ro = S.rest_body(om)
    for k in ("total", "thickness"):
        assert rb[k] == pytest.approx(ro[k], rel=1e-14)
    for k in ("centre", "inertia", "moments", "axes"):
        assert np.abs(rb[k] - ro[k]).max() <= 1e-12 * max(np.abs(ro[k]).max(), 1.0)


def _edge_share(p, lo, hi):
    return 0.5 if abs(p - lo) < 1e-9 or abs(p - hi) < 1e-9 else 1.0


def test_squeezed_free_block_matches_the_oracle_and_f_l_over_e_a():
    E = 2.0e5
    mesh, mats = lattice(cells=(4, 1, 4), mats=[material(E)])
    om = orc_mesh(mesh)
    shear, shifted = S.cell_moduli(om, [orc_material(m) for m in mats])
    X = om["rest"]
    lo, hi = X.min(0), X.max(0)
    load = 4.905
    f = np.zeros_like(X)
    top = np.abs(X[:, 1] - hi[1]) < 1e-9
    bottom = np.abs(X[:, 1] - lo[1]) < 1e-9
    share = np.array([_edge_share(p[0], lo[0], hi[0]) * _edge_share(p[2], lo[2], hi[2]) for p in X])
    f[top, 1] = -load * share[top] / share[top].sum()
    f[bottom, 1] = load * share[bottom] / share[bottom].sum()
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, mats)
        u, dent, its = ctx.lattice_screen(b, f)
    uo, dento = S.screen_load(om, shear, shifted, f)
    assert its > 0
    assert np.abs(u - uo).max() <= 1e-8 * np.abs(uo).max()
    assert dent == pytest.approx(dento, rel=1e-8)
    shortening = u[bottom, 1].mean() - u[top, 1].mean()
    expected = load * (hi[1] - lo[1]) / (np.float32(E) * (hi[0] - lo[0]) * (hi[2] - lo[2]))
    assert shortening == pytest.approx(expected, rel=1e-3)


def test_rigid_load_makes_no_dent_on_the_gpu():
    mesh, mats = lattice(cells=(4, 1, 4))
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, mats)
        rb = ctx.lattice_rest_body(b)
        arm = mesh["rest_positions"].astype(np.float64) - rb["centre"]
        m = mesh["node_mass"].astype(np.float64)[:, None]
        f = m * (np.array([1.0, 2.0, -3.0]) + np.cross([4.0, -1.0, 2.0], arm))
        _, dent, _ = ctx.lattice_screen(b, f)
    assert dent < 1e-9


def test_held_and_anchored_screen_matches_the_oracle():
    rng = np.random.default_rng(12)
    mats = [material(2.0e5), material(6.0e5, poisson=0.3)]
    sup = [api.Support.box((0.0, 0.0, 0.0), (0.01, 1.0, 1.0))]
    mesh, _ = lattice(cells=(5, 2, 2), mats=mats, supports=sup, lanes=rng.integers(0, 2, 20))
    om = orc_mesh(mesh)
    shear, shifted = S.cell_moduli(om, [orc_material(m) for m in mats])
    n = om["rest"].shape[0]
    held = np.flatnonzero(om["rest"][:, 0] > om["rest"][:, 0].max() - 1e-9)[:3]
    f = rng.standard_normal((n, 3))
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, mats)
        u, dent, _ = ctx.lattice_screen(b, f, held)
        u2, dent2, _ = ctx.lattice_screen(b, f)  # the anchored nodes alone
    uo, dento = S.screen_load(om, shear, shifted, f, held)
    uo2, dento2 = S.screen_load(om, shear, shifted, f)
    assert np.abs(u - uo).max() <= 1e-8 * np.abs(uo).max()
    assert np.abs(u2 - uo2).max() <= 1e-8 * np.abs(uo2).max()
    assert dent == pytest.approx(dento, rel=1e-8) and dent2 == pytest.approx(dento2, rel=1e-8)
    assert not np.allclose(u, u2)


def test_terms_follow_a_tear():
    """After tearCells adds node copies, the terms cover the new nodes and
    still sum to zero (the screen rebinds to the torn mesh)."""
    mat = api.PhysicalMaterial(youngs_modulus=2.0e5, density=1000.0, poissons_ratio=0.4, friction=0.7,
                               tear_strain=0.05)
    mesh, mats = lattice(cells=(6, 1, 1), mats=[mat])
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, mats)
        n0 = ctx.lattice_nodes(b)
        stretched = mesh["rest_positions"].astype(np.float64) * np.array([1.2, 1.0, 1.0])
        ctx.set_lattice_state(b, x=stretched)
        _, g0, _ = ctx.lattice_snh_terms(b)
        made = ctx.lattice_tear(b, [2])
        n1 = ctx.lattice_nodes(b)
        x, _ = ctx.lattice_state(b)
        x = x + 1e-3 * np.random.default_rng(2).standard_normal(x.shape)
        e, g, _ = ctx.lattice_snh_terms(b, x)
    assert made > 0 and n1 == n0 + made
    assert g0.shape == (n0, 3) and g.shape == (n1, 3)
    assert np.isfinite(e).all() and np.abs(g.sum(axis=0)).max() <= 1e-10 * np.abs(g).max()
