"""Host side of the strainstep hex lattices (no GPU): the library's hex mesh
builder and its element build (talus_mi_build_hex_lattice,
talus_mi_hex_element_make) equal the oracle's bit for bit; bodies the
library does not model are refused."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import hex as H
from talus_amd import api


def lib_material(**kw):
    return api.PhysicalMaterial(**kw)


def orc_material(m):
    o = oracle.Material()
    for name, _ in oracle.Material._fields_:
        setattr(o, name, getattr(m, name))
    return o


def support(center, half):
    return api.Support.box(center, half)


def orc_support(s):
    o = oracle.Support()
    o.world_to_support[:] = list(s.world_to_support)
    o.half[:] = list(s.half_extents)
    return o


@pytest.mark.parametrize("jitter,lanes,holes", [(False, 1, False), (True, 3, True)])
def test_hex_mesh_builder_matches_oracle(jitter, lanes, holes):
    rng = np.random.default_rng(5)
    res = (6, 4, 5)
    voxel = (0.1, 0.12, 0.08)
    cells = (res[0] - 1) * (res[1] - 1) * (res[2] - 1)
    mats = [lib_material(youngs_modulus=1e5 * (k + 1), density=900.0 + 50 * k) for k in range(lanes)]
    kw = {}
    if jitter:
        iz, iy, ix = np.meshgrid(*(np.arange(r) for r in res[::-1]), indexing="ij")
        pos = np.stack([ix * voxel[0], iy * voxel[1], iz * voxel[2]], -1).reshape(-1, 3)
        kw["sample_positions"] = (pos + rng.uniform(-0.01, 0.01, pos.shape)).astype(np.float32)
    if lanes > 1:
        kw["cell_lane"] = rng.integers(0, lanes + 1, cells).astype(np.uint32)  # one lane past the table
    if holes:
        solid = np.ones(cells, dtype=np.uint8)
        solid[rng.choice(cells, cells // 5, replace=False)] = 0
        kw["solid"] = solid
    sups = [support((0.0, 0.2, 0.2), (0.01, 1.0, 1.0))]
    lib = api.build_hex_lattice(res, voxel, (0.0, 0.0, 0.0), mats, sups, **kw)
    orc = H.build_hex_lattice(res, voxel, (0.0, 0.0, 0.0), [orc_material(m) for m in mats],
                              [orc_support(s) for s in sups],
                              sample_pos=kw.get("sample_positions"), cell_lane=kw.get("cell_lane"),
                              solid=kw.get("solid"))
    np.testing.assert_array_equal(lib["rest_positions"].astype(np.float64), orc["rest"])
    np.testing.assert_array_equal(lib["node_mass"], orc["mass"])
    np.testing.assert_array_equal(lib["anchored"], orc["anchored"])
    np.testing.assert_array_equal(lib["hex_nodes"], orc["hex"])
    np.testing.assert_array_equal(lib["volume"], orc["volume"])
    np.testing.assert_array_equal(lib["hex_material"], orc["lane"])
    assert lib["anchored"].any() and not lib["anchored"].all()
    # every node belongs to a hex; masses are positive
    assert np.unique(lib["hex_nodes"]).size == lib["node_mass"].size
    assert (lib["node_mass"] > 0).all()


def random_hex(rng, scale=0.1):
    base = np.array([[scale * s for s in H.corner_signs(c)] for c in range(8)])
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    return base @ A.T + 0.01 * scale * rng.standard_normal((8, 3))


FIELDS = [("gradient", "gradient"), ("inverse_mass", "inv_mass"), ("rest_hourglass", "rest_hg"),
          ("stiffness", "stiffness"), ("rate", "rate"), ("softness", "softness"), ("cap", "cap"),
          ("inverse", "inverse")]


@pytest.mark.parametrize("seed", range(6))
def test_element_build_matches_oracle_bitwise(seed):
    rng = np.random.default_rng(seed)
    corners = random_hex(rng)
    inv = 1.0 / rng.uniform(0.5, 2.0, 8)
    inv[rng.random(8) < 0.3] = 0.0  # held corners
    if not inv.any():
        inv[0] = 1.0
    E = float(np.float32(10 ** rng.uniform(4, 8)))
    nu = float(np.float32([0.0, 0.3, 0.45, 0.499, 0.6][seed % 5]))
    h = 1.0 / (60.0 * rng.integers(1, 17))
    vol = abs(np.linalg.det(H.corner_rest(corners)[0])) * 8.0
    lib = api.hex_element_make(corners, inv, vol, E, nu, h)
    orc = H.make(corners, inv, vol, np.float32(E), np.float32(nu), h)
    assert orc is not None
    for a, b in FIELDS:
        np.testing.assert_array_equal(lib.arr(a), orc.arr(b), err_msg=a)
    for a, b in (("bend", "bend"), ("size", "size"), ("volume_stiffness", "vol_stiffness"),
                 ("volume_rate", "vol_rate"), ("volume_softness", "vol_softness"),
                 ("volume_cap", "vol_cap")):
        assert getattr(lib, a) == getattr(orc, b), a
    assert lib.volumetric == orc.volumetric


def test_element_build_refuses_what_makeelement_refuses():
    corners = random_hex(np.random.default_rng(1))
    with pytest.raises(api.TalusError):  # no free corner
        api.hex_element_make(corners, np.zeros(8), 1e-3, 1e5, 0.3, 1 / 240)
    with pytest.raises(api.TalusError):  # no shear modulus
        api.hex_element_make(corners, np.ones(8), 1e-3, 0.0, 0.3, 1 / 240)
    with pytest.raises(api.TalusError):  # no volume
        api.hex_element_make(corners, np.ones(8), 0.0, 1e5, 0.3, 1 / 240)
    flat = corners.copy()
    flat[:, 2] = 0.0
    with pytest.raises(api.TalusError):  # degenerate rest shape
        api.hex_element_make(flat, np.ones(8), 1e-3, 1e5, 0.3, 1 / 240)
    for _ in (0,):
        assert H.make(corners, np.zeros(8), 1e-3, np.float32(1e5), np.float32(0.3), 1 / 240) is None


def test_hex_row_stiffness_closed_form_through_the_library():
    """dynamics_strainstep_tests.cc:1250-1268, on the library's build."""
    half, E, nu = 0.1, 1.0e5, 0.375
    V = (2 * half) ** 3
    corners = np.array([half * H.corner_signs(c) for c in range(8)])
    e = api.hex_element_make(corners, np.ones(8), V, E, nu, 1 / 240)
    mu = E / (2 * (1 + nu))
    for r in range(18):
        x = 2 * mu * V if r < 3 else (4 * mu * V if r < 6 else E * V / 3.0)
        assert abs(e.stiffness[r] - x) <= 1e-9 * x
    lam = E * nu / ((1 + nu) * (1 - 2 * nu)) * V
    assert abs(e.volume_stiffness - lam) <= 1e-9 * lam
    assert math.isclose(e.bend, 1.0 / (4 * 2 * half))
