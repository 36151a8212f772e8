"""Tear topology (host side, no GPU): the library's talus_mi_tear_mesh against
the oracle's restatement, and the invariants the reference's strip tests hold
(tests/dynamics_strainstep_tests.cc:3720-3796: no cell removed, mass kept to
1e-6, at least two pieces, no orphan node, split nodes start where their
origin is)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from talus_amd import api, scenes

from helpers import oracle_body


def tearable(**kw):
    return scenes.material_array(1, youngs_modulus=1e6, tear_strain=0.4,
                                 elastic_model=api.NEO_HOOKEAN, **kw)


def strip(cells=(6, 1, 1), voxel=0.2, supports=()):
    body = scenes.Body(resolution=tuple(c + 1 for c in cells), voxel=(voxel,) * 3,
                       materials=tearable(), supports=list(supports))
    mesh, mats = scenes.build_mesh(body)
    return body, mesh, mats


def pieces(tet, nodes):
    """Node-connected components over the tets."""
    parent = np.arange(nodes)

    def root(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for t in tet:
        for c in t[1:]:
            a, b = root(int(t[0])), root(int(c))
            if a != b:
                parent[max(a, b)] = min(a, b)
    return len({root(int(n)) for n in np.unique(tet)})


def check_tear(mesh, mats, x, torn):
    """Tears with the library and the oracle; asserts identical topology and
    the tear invariants; returns the library's mesh and origins."""
    new, origin = api.tear_mesh(mesh, scenes.materials_list(mats), x, torn)
    rest = mesh["rest_positions"].astype(np.float64)
    ot, orest, ox, oan, oorg = O.tear_topology(mesh["tet_nodes"], rest, x, mesh["anchored"], torn)
    n0 = mesh["anchored"].size
    # identical connectivity, origins, anchoring and rest positions
    np.testing.assert_array_equal(new["tet_nodes"], ot)
    np.testing.assert_array_equal(origin, oorg)
    np.testing.assert_array_equal(new["anchored"], oan)
    np.testing.assert_array_equal(new["rest_positions"], orest.astype(np.float32))
    # no tet removed, every tet keeps its corners up to the split copies
    assert new["tet_nodes"].shape == mesh["tet_nodes"].shape
    back = np.concatenate([np.arange(n0), origin]).astype(np.uint32)
    while True:  # copies of copies map back in several hops
        nxt = back[back]
        if np.array_equal(nxt, back):
            break
        back = nxt
    np.testing.assert_array_equal(back[new["tet_nodes"]], mesh["tet_nodes"])
    # new nodes copy the rest position of their origin; none is anchored
    np.testing.assert_array_equal(new["rest_positions"][n0:], mesh["rest_positions"][back[n0:]])
    assert not new["anchored"][n0:].any()
    # no orphan node; mass kept
    assert np.unique(new["tet_nodes"]).size == new["anchored"].size
    m0, m1 = float(mesh["node_mass"].astype(np.float64).sum()), float(new["node_mass"].astype(np.float64).sum())
    assert abs(m1 - m0) <= 1e-6 * m0
    # masses are rho V / 4 per incident tet on the torn mesh
    rho = float(mats["density"][0]) if mats.shape[0] == 1 else None
    if rho is not None:
        expect = np.zeros(new["anchored"].size)
        np.add.at(expect, new["tet_nodes"].reshape(-1),
                  np.repeat(rho * new["volume"].astype(np.float64) / 4.0, 4))
        np.testing.assert_allclose(new["node_mass"], expect, rtol=1e-5)
    return new, origin


def test_strip_cut_through_one_hex_column_falls_apart():
    _, mesh, mats = strip()
    rest = mesh["rest_positions"].astype(np.float64)
    x = rest.copy()
    x[:, 0] *= 1.01
    cent = rest[mesh["tet_nodes"]].mean(axis=1)
    torn = np.nonzero((cent[:, 0] > 0.4) & (cent[:, 0] < 0.6))[0]
    new, origin = check_tear(mesh, mats, x, torn)
    assert origin.size > 0
    assert pieces(new["tet_nodes"], new["anchored"].size) >= 2


def test_pinned_nodes_are_never_split():
    sups = [((0.6, 0.0, 0.0), (1.0, 1.0, 1.0))]  # every node anchored
    _, mesh, mats = strip(supports=sups)
    assert mesh["anchored"].all()
    x = mesh["rest_positions"].astype(np.float64) * 1.05
    new, origin = check_tear(mesh, mats, x, np.arange(mesh["tet_nodes"].shape[0]))
    assert origin.size == 0
    np.testing.assert_array_equal(new["tet_nodes"], mesh["tet_nodes"])


def test_empty_and_out_of_range_torn_lists_change_nothing():
    _, mesh, mats = strip()
    x = mesh["rest_positions"].astype(np.float64)
    for torn in (np.zeros(0, np.uint32), np.array([10 ** 6], np.uint32)):
        new, origin = check_tear(mesh, mats, x, torn)
        assert origin.size == 0
        np.testing.assert_array_equal(new["tet_nodes"], mesh["tet_nodes"])


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_tears_of_a_jittered_block_match_the_oracle(seed):
    rng = np.random.default_rng(seed)
    cells = (5, 4, 3)
    h = 0.1
    res = tuple(c + 1 for c in cells)
    iz, iy, ix = np.meshgrid(*(np.arange(r) for r in res[::-1]), indexing="ij")
    pos = (np.stack([ix, iy, iz], -1).reshape(-1, 3) * h
           + rng.uniform(-0.05, 0.05, (ix.size, 3)) * h).astype(np.float32)
    body = scenes.Body(resolution=res, voxel=(h, h, h), sample_positions=pos, materials=tearable(),
                       supports=[((0.0, 0.0, 0.0), (0.01, 1.0, 1.0))])
    mesh, mats = scenes.build_mesh(body)
    rest = mesh["rest_positions"].astype(np.float64)
    # anisotropic stretch plus noise so every tet has a distinct crack normal
    A = np.diag([1.0, 1.0, 1.0]) + 0.3 * rng.standard_normal((3, 3))
    x = rest @ A.T + 0.01 * h * rng.standard_normal(rest.shape)
    T = mesh["tet_nodes"].shape[0]
    torn = np.sort(rng.choice(T, size=T // 6, replace=False)).astype(np.uint32)
    new, origin = check_tear(mesh, mats, x, torn)
    assert origin.size > 0
    # tear again on the torn mesh: repeated tears compose
    x2 = list(x)
    for o in origin:  # a copy may come from an earlier copy
        x2.append(x2[o])
    x2 = np.asarray(x2)
    torn2 = np.sort(rng.choice(T, size=T // 8, replace=False)).astype(np.uint32)
    check_tear(new, mats, x2, torn2)


def test_oracle_body_tear_matches_the_library_masses_and_keeps_state():
    body, mesh, mats = strip(cells=(8, 2, 2), voxel=0.1)
    orc, _ = oracle_body(body)
    x, v, a = orc.state()
    rng = np.random.default_rng(4)
    x = x * np.array([1.3, 1.0, 1.0]) + 1e-3 * rng.standard_normal(x.shape)
    v = rng.standard_normal(x.shape)
    orc.set_state(x, v, a)
    cent = mesh["rest_positions"][mesh["tet_nodes"]].mean(axis=1)
    torn = np.nonzero(np.abs(cent[:, 0] - 0.4) < 0.05)[0].astype(np.uint32)
    new, origin = check_tear(mesh, mats, x, torn)
    added = orc.tear(torn)
    assert added == origin.size
    topo = orc.topology()
    np.testing.assert_array_equal(topo["tet"], new["tet_nodes"])
    np.testing.assert_allclose(topo["mass"], new["node_mass"], rtol=1e-6)
    x2, v2, _ = orc.state()
    n0 = x.shape[0]
    np.testing.assert_array_equal(x2[:n0], x)
    back = np.concatenate([np.arange(n0), origin])
    for i in range(n0, back.size):
        back[i] = back[back[i]]
    np.testing.assert_array_equal(x2[n0:], x[back[n0:]])
    np.testing.assert_array_equal(v2[n0:], v[back[n0:]])


def test_stretch_on_plain_arrays_equals_the_body_stretch_and_numpy():
    """O.stretch_tets (the tear test on a mesh given as arrays) equals the
    oracle body's own stretch test bit for bit, and the per-F stretch of a tet
    whose F is known."""
    sc = scenes.slab(cells=(6, 4, 2))
    orc, m = oracle_body(sc.bodies[0])
    rng = np.random.default_rng(3)
    x = m["rest"] + 0.004 * rng.standard_normal(m["rest"].shape)
    z = np.zeros_like(x)
    orc.set_state(x, z, z)
    t = orc.topology()
    np.testing.assert_array_equal(O.stretch_tets(t["tet"], t["rest"], x), orc.stretch())
    rest = np.array([[0.0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, 0.1]])
    F = np.array([[1.3, 0.2, 0.0], [0.0, 0.9, 0.1], [0.05, 0.0, 1.1]])
    s = O.stretch_tets([[0, 1, 2, 3]], rest, rest @ F.T)
    assert abs(s[0] - O.stretch(F)[0]) <= 1e-12
    with pytest.raises(ValueError):
        O.stretch_tets([[0, 1, 2, 9]], rest, rest)


def _column(mesh, rest, hexes_x, which):
    """Tets of the cells with x index 1 (the middle column of a 3-cell-long
    strip): all of them, or those with three corners on the plane x = 2h
    (whose crack plane is therefore x = 2h)."""
    tet = mesh["tet_nodes"].reshape(-1, 4)
    h = 0.1
    cent = rest[tet].mean(axis=1)
    col = np.flatnonzero(np.floor(cent[:, 0] / h).astype(int) == 1)
    if which == "all":
        return col
    return np.array([e for e in col if np.isclose(rest[tet[e], 0], 2 * h).sum() == 3])


@pytest.mark.parametrize("section", [(1, 1), (2, 1), (2, 2)])
def test_tearing_a_cross_section_splits_its_plane_into_two_pieces(section):
    """Hand-derived outcome of tearCells' rules (tearing.cc:292-513 on tets) on
    a 3-cell strip stretched 1.5x along x, with the x = 0 face anchored: the
    tets of the middle column whose three nearest corners lie on x = 2h
    crack along that plane.  Corners closest to the crack plane split first
    and the pinched ones follow, so exactly the nodes of the cross-section
    x = 2h ((ny+1)(nz+1) of them) are duplicated, the strip falls into 2
    pieces, the copies sit on the plane and are not anchored, and the
    original ids stay on the side of negative rest offset (side 0 keeps the
    node, the copy goes to the other side's cells)."""
    ny, nz = section
    body, mesh, mats = strip(cells=(3, ny, nz), voxel=0.1,
                             supports=[((0.0, 0.5, 0.5), (0.01, 1.0, 1.0))])
    rest = mesh["rest_positions"].astype(np.float64)
    x = rest * np.array([1.5, 1.0, 1.0])
    torn = _column(mesh, rest, 3, "plane")
    assert torn.size == 2 * ny * nz  # two Kuhn tets per cell have three corners on x = 2h
    new, origin = check_tear(mesh, mats, x, torn)
    n0 = rest.shape[0]
    plane = np.flatnonzero(np.isclose(rest[:, 0], 0.2))
    assert plane.size == (ny + 1) * (nz + 1)
    assert new["anchored"].size - n0 == plane.size
    np.testing.assert_array_equal(np.sort(origin), plane)
    assert np.allclose(new["rest_positions"][n0:, 0], 0.2)
    assert pieces(new["tet_nodes"].reshape(-1, 4), new["anchored"].size) == 2
    tet = new["tet_nodes"].reshape(-1, 4)
    cent = rest[mesh["tet_nodes"].reshape(-1, 4)].mean(axis=1)[:, 0]
    for e in range(tet.shape[0]):
        on_plane = np.isin(mesh["tet_nodes"].reshape(-1, 4)[e], plane)
        if not on_plane.any():
            continue
        if cent[e] < 0.2:
            assert (tet[e][on_plane] < n0).all(), e  # originals on the low side
        else:
            assert (tet[e][on_plane] >= n0).all(), e  # copies on the high side


@pytest.mark.parametrize("section", [(1, 1), (2, 1), (2, 2)])
def test_tearing_a_whole_column_cuts_both_of_its_planes(section):
    """Every tet of the middle column torn: cells whose nearest corners lie on
    x = h crack there, the others on x = 2h, so both cross-sections are
    duplicated (2 (ny+1)(nz+1) new nodes) and the strip falls into 3 pieces,
    the middle column (6 ny nz tets, at least MIN_PIECE_CELLS) one of them."""
    ny, nz = section
    body, mesh, mats = strip(cells=(3, ny, nz), voxel=0.1,
                             supports=[((0.0, 0.5, 0.5), (0.01, 1.0, 1.0))])
    rest = mesh["rest_positions"].astype(np.float64)
    x = rest * np.array([1.5, 1.0, 1.0])
    new, origin = check_tear(mesh, mats, x, _column(mesh, rest, 3, "all"))
    n0 = rest.shape[0]
    assert new["anchored"].size - n0 == 2 * (ny + 1) * (nz + 1)
    assert sorted(np.unique(np.round(new["rest_positions"][n0:, 0], 6))) == [0.1, 0.2]
    assert pieces(new["tet_nodes"].reshape(-1, 4), new["anchored"].size) == 3


def test_one_torn_tet_of_a_one_cell_strip_adds_four_nodes():
    """The tet form of the reference's strip test (tests/dynamics_strainstep_
    tests.cc:3760-3796: a one-cell-thick strip split adds exactly 4 nodes and
    leaves 2 pieces): any single tet of the middle cell cracks the strip
    through the whole 2 x 2-node cross-section nearest its corners."""
    body, mesh, mats = strip(cells=(3, 1, 1), voxel=0.1)
    rest = mesh["rest_positions"].astype(np.float64)
    x = rest * np.array([1.5, 1.0, 1.0])
    for e in _column(mesh, rest, 3, "all"):
        new, origin = check_tear(mesh, mats, x, [int(e)])
        assert new["anchored"].size - rest.shape[0] == 4
        assert np.unique(np.round(rest[origin, 0], 6)).size == 1
        assert pieces(new["tet_nodes"].reshape(-1, 4), new["anchored"].size) == 2
