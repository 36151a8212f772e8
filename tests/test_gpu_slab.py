"""The crushed slab (plasticity and tearing together): a 2 x 1 x 0.1 m slab,
E 1 MPa, nu 0.3, rho 1000 kg/m^3, yield stress 2e5 Pa, linear hardening
1e5 Pa, tear strain 0.35, both ends held, the top-face nodes under the
0.4 x 0.4 m box footprint held and driven down at 1 m/s by setting their
positions before each step (the reference's tests drive held nodes that way:
tests/dynamics_strainstep_tests.cc:3740-3744).

* Event lists: a 20 x 10 x 3 cell slab stepped on the GPU (FP64, block
  Jacobi, the oracle's solver) and by the oracle in lockstep, through yield
  onset and the first tears: per step the yield events, the tear events and
  the connectivity after the tears are identical and the states agree.
* Full size (82 x 41 x 5 cells, 100,860 tets, mixed precision, block Cholesky
  preconditioner): per step the tear events equal the oracle's stretch test on
  the GPU's positions (no stretch within 1e-9 of the threshold) and the
  connectivity after the tears is bit-exact with the oracle's restatement
  applied to the GPU's pre-tear mesh; tets are never removed, mass is kept, no
  node is orphaned and the state is finite; the torn body's surface and crack
  faces equal the oracle's and still enclose the slab's volume.
"""
import numpy as np
import pytest

import oracle as O
from talus_amd import api, scenes

from helpers import oracle_body, product_body

pytestmark = pytest.mark.gpu

THRESHOLD = 1.0 + np.float32(0.35)


def _drive(ctx, b, sc, driven, z0, step):
    x, v, a = ctx.state(b)
    x[driven, 2] = scenes.slab_drive(sc, z0, step + 1)
    ctx.set_state(b, x, v, a)


def _tears(ev, step):
    return np.array(sorted(e[2] for e in ev if e[0] == api.EVENT_TEAR and e[3] == step), dtype=np.uint32)


def _yields(ev, step):
    return np.array(sorted(e[2] for e in ev if e[0] == api.EVENT_YIELD and e[3] == step), dtype=np.int64)


def _check_tear_step(topo_before, x_after, n_before, topo_after, torn):
    """The step's host topology update against the oracle's restatement."""
    ot, orest, _, oan, _ = O.tear_topology(topo_before["tet_nodes"], topo_before["rest"], x_after[:n_before],
                                           topo_before["anchored"], torn)
    np.testing.assert_array_equal(topo_after["tet_nodes"], ot)
    np.testing.assert_array_equal(topo_after["rest"], orest)
    np.testing.assert_array_equal(topo_after["anchored"], oan)


def test_slab_event_lists_match_the_oracle_in_lockstep():
    sc = scenes.slab(cells=(20, 10, 3))
    orc, m = oracle_body(sc.bodies[0])
    rest = m["rest"]
    driven = scenes.slab_driven_nodes(sc, rest, m["anchored"])
    assert driven.size == 25
    z0 = rest[driven, 2].copy()
    steps = 33
    yields = tears = 0
    with api.Context(0, api.FP64) as ctx:
        b, mesh = product_body(ctx, sc.bodies[0])
        for s in range(steps):
            _drive(ctx, b, sc, driven, z0, s)
            xo, vo, ao = orc.state()
            xo[driven, 2] = scenes.slab_drive(sc, z0, s + 1)
            orc.set_state(xo, vo, ao)
            before = ctx.topology(b)
            n_before = before["rest"].shape[0]
            r = ctx.step(preconditioner=api.BLOCK_JACOBI, gravity=sc.gravity,
                         flags=api.DETECT_TEARS | api.RECORD_EVENTS)
            ev = ctx.events()
            orc.step(gravity=sc.gravity)
            yo = np.flatnonzero(orc.yield_flags())
            stretch = orc.stretch()
            to = np.flatnonzero(stretch > THRESHOLD).astype(np.uint32)
            assert np.abs(stretch - THRESHOLD).min() > 1e-9, s
            assert not r["non_finite"], s
            np.testing.assert_array_equal(_yields(ev, s), yo, err_msg=f"yield events at step {s}")
            np.testing.assert_array_equal(_tears(ev, s), to, err_msg=f"tear events at step {s}")
            assert r["tear_events"] == to.size
            x, _, _ = ctx.state(b)
            xo, _, _ = orc.state()
            disp = np.abs(xo[: rest.shape[0]] - rest).max()
            assert np.abs(x[:n_before] - xo).max() <= 1e-8 * disp, s
            if to.size:
                orc.tear(to)
                after = ctx.topology(b)
                _check_tear_step(before, x, n_before, after, to)
                np.testing.assert_array_equal(after["tet_nodes"], orc.topology()["tet"])
            yields += yo.size
            tears += to.size
    assert yields > 100 and tears > 10


def test_full_slab_tear_onset_connectivity_and_invariants():
    sc = scenes.slab()
    mesh, mats = scenes.build_mesh(sc.bodies[0])
    rest = mesh["rest_positions"].astype(np.float64)
    driven = scenes.slab_driven_nodes(sc, rest, mesh["anchored"])
    z0 = rest[driven, 2].copy()
    with api.Context(0, api.MIXED) as ctx:
        b = ctx.add_body(mesh, mats)
        n0, t0 = ctx.info(b)
        assert t0 == 100860
        mass0 = ctx.topology(b)["mass"].sum()
        params = api.default_params(preconditioner=api.FACTOR, gravity=sc.gravity,
                                    flags=api.DETECT_TEARS | api.RECORD_EVENTS)
        total = 0
        for s in range(3):
            _drive(ctx, b, sc, driven, z0, s)
            before = ctx.topology(b)
            n_before = before["rest"].shape[0]
            r = ctx.step(params)
            ev = ctx.events()
            assert not r["non_finite"], s
            x, v, a = ctx.state(b)
            assert np.isfinite(x).all() and np.isfinite(v).all()
            stretch = O.stretch_tets(before["tet_nodes"], before["rest"], x[:n_before])
            torn = np.flatnonzero(stretch > THRESHOLD).astype(np.uint32)
            assert np.abs(stretch - THRESHOLD).min() > 1e-9, s
            np.testing.assert_array_equal(_tears(ev, s), torn, err_msg=f"tear events at step {s}")
            after = ctx.topology(b)
            if torn.size:
                _check_tear_step(before, x, n_before, after, torn)
            n, t = ctx.info(b)
            assert t == t0
            assert r["split_nodes"] == n - n_before
            assert np.unique(after["tet_nodes"]).size == n  # no orphan node
            assert abs(after["mass"].sum() - mass0) <= 1e-6 * mass0
            total += torn.size
        assert total > 1000
        # the torn body's surface: the library's equals the oracle's on the
        # same mesh, and it still encloses the body's volume
        srf = ctx.surface(b)
        topo = ctx.topology(b)
        ref = O.surface(topo["tet_nodes"], topo["rest"], np.zeros(t0, dtype=np.uint32))
        for key in ("triangles", "triangle_tet", "edges", "nodes", "crack_faces"):
            np.testing.assert_array_equal(srf[key], ref[key], err_msg=key)
        tri = srf["triangles"]
        r = topo["rest"]
        enclosed = np.einsum("ij,ij->i", r[tri[:, 0]], np.cross(r[tri[:, 1]], r[tri[:, 2]])).sum() / 6.0
        assert abs(enclosed - 2.0 * 1.0 * 0.1) <= 1e-6
        assert srf["crack_faces"].sum() > 0
