"""The deformed-field re-bake of a hex lattice on the GPU
(talus_mi_lattice_bake_field) against the oracle (oracle/field_oracle.c,
pinned by tests/test_field_cpu.py with the reference's deformed-field
tests): at rest, moved and turned, a sunk top, a stretched and bent
lattice, a torn lattice's crack faces and carving; the baked distances equal
the oracle's to 1e-6 and the changed boxes match."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import field as FD
from talus_amd import api

from test_field_cpu import HALF, SINK, VOXEL, box_field, sample_points

pytestmark = pytest.mark.gpu


def material(tear=math.inf):
    return api.PhysicalMaterial(youngs_modulus=2e8, density=1000.0, poissons_ratio=0.3, friction=0.7,
                                tear_strain=tear)


def lattice(tear=math.inf):
    cells = np.round(2 * HALF / VOXEL).astype(int)
    mat = material(tear)
    return api.build_hex_lattice([c + 1 for c in cells], [VOXEL] * 3, tuple(-HALF), [mat], []), mat


def orc_mesh(mesh, rest=None):
    return {"rest": mesh["rest_positions"].astype(np.float64) if rest is None else rest,
            "mass": mesh["node_mass"], "anchored": mesh["anchored"], "hex": mesh["hex_nodes"],
            "volume": mesh["volume"], "lane": mesh["hex_material"], "voxel": mesh["voxel_size"]}


def quat_matrix(q):
    w, x, y, z = (float(np.float32(v)) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def run(mesh, mat, pos, position=(0.0, 0.0, 0.0), rotation=(1.0, 0.0, 0.0, 0.0), carve=False):
    field = box_field()
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat])
        ctx.set_lattice_state(b, x=pos)
        baked, box = ctx.lattice_bake_field(b, field, position=position, rotation=rotation, carve=carve)
    model = np.hstack([quat_matrix(rotation), np.asarray(position, np.float32).astype(np.float64)[:, None]])
    ob, obox = FD.bake_deformed(orc_mesh(mesh), field, pos, model, carve=carve)
    return field, baked, box, ob, obox


def check(baked, box, ob, obox):
    assert np.abs(baked - ob).max() <= 1e-6
    assert np.array_equal(box[1], obox[1]) and (box[1] == 0).all() == (obox[1] == 0).all()
    if (obox[1] > 0).all():
        assert np.array_equal(box[0], obox[0])


def test_bake_at_rest_and_moved_rigidly():
    mesh, mat = lattice()
    X = mesh["rest_positions"].astype(np.float64)
    field, baked, box, ob, obox = run(mesh, mat, X)
    check(baked, box, ob, obox)
    assert (box[1] == 0).all() and np.abs(baked - field["distances"]).max() <= 1e-5
    axis = np.array([0.2, 1.0, 0.1]) / np.linalg.norm([0.2, 1.0, 0.1])
    q = np.concatenate([[math.cos(0.25)], math.sin(0.25) * axis])
    t = (0.3, 0.2, -0.1)
    R = quat_matrix(q)
    pos = X @ R.T + np.asarray(t, np.float32).astype(np.float64)
    field, baked, box, ob, obox = run(mesh, mat, pos, position=t, rotation=q)
    check(baked, box, ob, obox)
    assert np.abs(baked - field["distances"]).max() <= 1e-5


def test_bake_sunk_top_and_bent_lattice():
    mesh, mat = lattice()
    X = mesh["rest_positions"].astype(np.float64)
    top = X[:, 1].max()
    pos = X.copy()
    pos[np.abs(X[:, 1] - top) < 1e-5, 1] -= SINK
    field, baked, box, ob, obox = run(mesh, mat, pos)
    check(baked, box, ob, obox)
    cur = dict(field, distances=baked)
    assert abs(FD.sdf_probe(cur, [0.0, top - SINK, 0.0])) <= 2e-4
    assert 0 < box[1][1] < field["res"][1]
    # a smooth bend and stretch
    bent = X * np.array([1.15, 1.0, 0.95])
    bent[:, 1] += 0.3 * X[:, 0] ** 2
    field, baked, box, ob, obox = run(mesh, mat, bent)
    check(baked, box, ob, obox)
    assert (box[1] > 0).all()


def test_bake_reads_crack_faces_and_carves():
    mesh, mat = lattice(tear=0.05)
    X = mesh["rest_positions"].astype(np.float64)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat])
        ctx.set_lattice_state(b, x=X * np.array([1.3, 1.0, 1.0]))
        made = ctx.lattice_tear(b, [20, 21])
        x, _ = ctx.lattice_state(b)
        hexes, masses, crack = ctx.lattice_topology(b)
        field = box_field()
        baked, box = ctx.lattice_bake_field(b, field)
        carved, _ = ctx.lattice_bake_field(b, field, carve=True)
    assert made > 0 and crack.any()
    # node copies keep their origin's rest position: recover them from the
    # cells (cell i at x-fastest lattice index), originals from the mesh
    n = x.shape[0]
    rest = np.zeros((n, 3))
    cells = np.round(2 * HALF / VOXEL).astype(int)
    for i, h in enumerate(hexes):
        cz, rem = divmod(i, cells[0] * cells[1])
        cy, cx = divmod(rem, cells[0])
        for c in range(8):
            rest[h[c]] = -HALF + VOXEL * np.array([cx + (c & 1), cy + ((c >> 1) & 1), cz + ((c >> 2) & 1)])
    for c in range(X.shape[0], n):
        rest[c] = X[np.argmin(np.linalg.norm(X - rest[c], axis=1))]
    rest[:X.shape[0]] = X
    om = orc_mesh(dict(mesh, hex_nodes=hexes, node_mass=masses, anchored=np.zeros(x.shape[0], np.uint8)), rest)
    ob, obox = FD.bake_deformed(om, field, x, crack=crack)
    oc, _ = FD.bake_deformed(om, field, x, crack=crack, carve=True)
    check(baked, box, ob, obox)
    assert np.abs(carved - oc).max() <= 1e-6


def test_torn_lattice_surface_mesh_matches_the_oracle_and_stays_closed():
    """collectSurface's hex faces of a torn lattice: the library's surface of
    the current mesh equals the oracle's on the same topology, encloses the
    lattice's volume (crack faces come in opposite pairs), uses every edge
    an even number of times, and carries the lattice's crack faces."""
    from test_surface_cpu import area, edge_uses, enclosed_volume
    mesh, mat = lattice(tear=0.05)
    X = mesh["rest_positions"].astype(np.float64)
    with api.Context() as ctx:
        b = ctx.add_lattice(mesh, [mat])
        ctx.set_lattice_state(b, x=X * np.array([1.3, 1.0, 1.0]))
        made = ctx.lattice_tear(b, [20, 21, 52])
        srf = ctx.lattice_surface_mesh(b)
        hexes, masses, crack = ctx.lattice_topology(b)
        n = ctx.lattice_nodes(b)
    assert made > 0 and crack.any()
    np.testing.assert_array_equal(srf["crack_faces"], crack)
    cells = np.round(2 * HALF / VOXEL).astype(int)
    rest = np.zeros((n, 3))
    for i, h in enumerate(hexes):
        cz, rem = divmod(i, cells[0] * cells[1])
        cy, cx = divmod(rem, cells[0])
        for c in range(8):
            rest[h[c]] = -HALF + VOXEL * np.array([cx + (c & 1), cy + ((c >> 1) & 1), cz + ((c >> 2) & 1)])
    # a copy keeps its origin's rest position: take the original node there
    for c in range(X.shape[0], n):
        rest[c] = X[np.argmin(np.linalg.norm(X - rest[c], axis=1))]
    rest[:X.shape[0]] = X
    om = orc_mesh(dict(mesh, hex_nodes=hexes, node_mass=masses, anchored=np.zeros(n, np.uint8)), rest)
    orc = FD.hex_surface(om)
    for key in orc:
        np.testing.assert_array_equal(srf[key], orc[key], err_msg=key)
    box = 8.0 * HALF.prod()
    assert abs(enclosed_volume(srf, rest) - box) <= 1e-6 * box
    assert area(srf["triangles"], rest) > 2 * 4 * (HALF[0] * HALF[1] + HALF[1] * HALF[2] + HALF[0] * HALF[2])
    assert np.all(edge_uses(srf) % 2 == 0)
