"""Shared builders: the same seeded scene on the product and on the oracle."""
from __future__ import annotations

import numpy as np

import oracle as O
from talus_amd import api, scenes


def oracle_support(center, half):
    s = O.Support()
    s.world_to_support[:] = [1, 0, 0, -center[0], 0, 1, 0, -center[1], 0, 0, 1, -center[2]]
    s.half[:] = list(half)
    return s


def oracle_mesh(body: scenes.Body):
    sups = [oracle_support(c, h) for c, h in body.supports]
    mats = body.materials if not body.per_tet else body.materials[:1]
    m = O.build_lattice(body.resolution, body.voxel, body.origin, materials=mats, supports=sups,
                        sample_pos=body.sample_positions)
    if body.per_tet:
        m["lane"] = np.arange(m["tet"].shape[0], dtype=np.uint32)
    return m


def oracle_body(body: scenes.Body):
    m = oracle_mesh(body)
    return O.Body(m, body.materials), m


def product_body(ctx: api.Context, body: scenes.Body):
    mesh, mats = scenes.build_mesh(body)
    return ctx.add_body(mesh, mats), mesh


def deformed(rest, seed=0, scale=0.01, twist=0.3):
    """A smooth large rotation plus seeded noise of the rest positions."""
    rng = np.random.default_rng(seed)
    c = rest.mean(axis=0)
    a = twist
    R = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
    ext = np.ptp(rest, axis=0).max()
    x = (rest - c) @ R.T + c + scale * ext * rng.standard_normal(rest.shape)
    return x


def rel_force_error(f, ref):
    """max_e |f_e - ref_e|_2 / max_e |ref_e|_2 over per-tet corner forces."""
    d = np.linalg.norm((f - ref).reshape(f.shape[0], -1), axis=1).max()
    return d / np.linalg.norm(ref.reshape(ref.shape[0], -1), axis=1).max()
