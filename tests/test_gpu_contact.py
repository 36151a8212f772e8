"""GPU parity of contact generation (talus_mi_field_*, talus_mi_find_contacts)
against the FP32 oracle (oracle/contact_oracle.c) on the same float state:
every contact bit-equal (nodes, weights, partner, normal, local point, gap,
bias, friction, triangle, in the same order) for shapes that do not spin;
a spinning shape's turn uses correctly rounded sine and cosine where the C
library's are within 0.56 ulp, so there the contact set is equal and the
values agree to 1e-5."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import contact as OC
from talus_amd import api

from test_contact_cpu import cube_body, jittered_scene, plane_field, sphere_field

pytestmark = pytest.mark.gpu


def gpu_contacts(ctx, b, fields, shapes, dt, **kw):
    ids = [ctx.field_add(f) for f in fields]
    sh = []
    for s in shapes:
        s = dict(s)
        s["field"] = ids[s["field"]]
        sh.append(s)
    return ctx.find_contacts(b, sh, dt, **kw)


def oracle_contacts(ctx, b, mesh, mats, fields, shapes, dt, thickness=None, **kw):
    x, v, _ = ctx.state(b)
    srf = ctx.surface(b)
    osh = [dict(s, field=fields[s["field"]]) for s in shapes]
    return OC.find_contacts(srf, x.astype(np.float32), v.astype(np.float32), osh, dt, anchored=mesh["anchored"],
                            friction=mats["friction"], thickness=thickness, **kw)


def add_body(ctx, mesh, mats, x=None, v=None):
    b = ctx.add_body(mesh, mats)
    if x is not None:
        ctx.set_state(b, x, v, np.zeros_like(x))
    return b


def test_field_lipschitz_matches_the_oracle():
    with api.Context(0, api.FP64) as ctx:
        for f in (sphere_field(), plane_field(normal=(1.0, 2.0, 2.0)),
                  dict(origin=(0, 0, 0), voxel=(0.5, 0.25, 1.0), res=(7, 5, 6),
                       distances=np.random.default_rng(4).normal(size=210).astype(np.float32))):
            fid = ctx.field_add(f)
            assert np.float32(ctx.field_lipschitz(fid)) == np.float32(OC.field_lipschitz(f))


@pytest.mark.parametrize("cull", [True, False])
def test_resting_cube_bit_equal(cull):
    _, mesh, mats = cube_body()
    fields = [plane_field()]
    shapes = [dict(field=0)]
    with api.Context(0, api.FP64) as ctx:
        b = add_body(ctx, mesh, mats)
        got, n = gpu_contacts(ctx, b, fields, shapes, 1 / 60, bias=0.01, threshold=0.1, cull=cull,
                              voxel=(0.125,) * 3)
        want = oracle_contacts(ctx, b, mesh, mats, fields, shapes, 1 / 60, bias=0.01, threshold=0.1, cull=cull,
                               voxel=(0.125,) * 3)
    assert n == len(want) == 25
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("cull", [True, False])
def test_moving_scene_bit_equal(seed, cull):
    mesh, mats, pos, vel, shapes = jittered_scene(seed)
    fields = [s["field"] for s in shapes]
    shp = [dict(s, field=i) for i, s in enumerate(shapes)]
    kw = dict(bias=0.005, threshold=0.03, cull=cull, gravity=(0.0, -9.81, 0.0), voxel=(0.1,) * 3)
    with api.Context(0, api.FP64) as ctx:
        b = add_body(ctx, mesh, mats, pos.astype(np.float64), vel.astype(np.float64))
        got, n = gpu_contacts(ctx, b, fields, shp, 1 / 60, **kw)
        want = oracle_contacts(ctx, b, mesh, mats, fields, shp, 1 / 60, **kw)
    assert n == len(want) > 0
    assert got.tobytes() == want.tobytes()


def test_skin_capacity_and_second_body():
    mesh, mats, pos, vel, shapes = jittered_scene(3)
    fields = [s["field"] for s in shapes]
    shp = [dict(s, field=i) for i, s in enumerate(shapes)]
    kw = dict(bias=0.005, threshold=0.03, voxel=(0.1,) * 3)
    with api.Context(0, api.FP64) as ctx:
        b = add_body(ctx, mesh, mats, pos.astype(np.float64), vel.astype(np.float64))
        T = ctx.surface(b)["triangles"].shape[0]
        thick = np.random.default_rng(3).uniform(0.0, 0.01, T).astype(np.float32)
        ids = [ctx.field_add(f) for f in fields]
        sh = [dict(s, field=ids[s["field"]]) for s in shp]
        got, n = ctx.find_contacts(b, sh, 1 / 60, thickness=thick, **kw)
        want = oracle_contacts(ctx, b, mesh, mats, fields, shp, 1 / 60, thickness=thick, **kw)
        assert n == len(want) > 3 and got.tobytes() == want.tobytes()
        part, n2 = ctx.find_contacts(b, sh, 1 / 60, thickness=thick, capacity=3, **kw)
        assert n2 == n and part.tobytes() == want[:3].tobytes()
        # a second body changes the context (its surfaces are rebuilt); the
        # first body's answer stays, the second's is its own
        _, mesh2, mats2 = cube_body()
        b2 = add_body(ctx, mesh2, mats2)
        again, _ = ctx.find_contacts(b, sh, 1 / 60, thickness=thick, **kw)
        assert again.tobytes() == want.tobytes()
        other, n3 = ctx.find_contacts(b2, [dict(field=ids[0])], 1 / 60, bias=0.01, threshold=0.1,
                                      voxel=(0.125,) * 3)
        want2 = oracle_contacts(ctx, b2, mesh2, mats2, [fields[0]], [dict(field=0)], 1 / 60, bias=0.01,
                                threshold=0.1, voxel=(0.125,) * 3)
        assert n3 == 25 and other.tobytes() == want2.tobytes()
        # a field rewritten in place: the plane raised by 0.02
        raised = plane_field(height=0.02)
        ctx.field_set(ids[0], raised)
        moved, _ = ctx.find_contacts(b2, [dict(field=ids[0])], 1 / 60, bias=0.01, threshold=0.1,
                                     voxel=(0.125,) * 3)
        want3 = oracle_contacts(ctx, b2, mesh2, mats2, [raised], [dict(field=0)], 1 / 60, bias=0.01,
                                threshold=0.1, voxel=(0.125,) * 3)
        assert moved.tobytes() == want3.tobytes() and np.allclose(moved["gap"], 0.0625 - 0.02 - 0.01, atol=1e-6)


def test_spinning_shape_agrees():
    mesh, mats, pos, vel, shapes = jittered_scene(1)
    fields = [sphere_field()]
    shp = [dict(field=0, center=(0.1, 0.6, 0.0), linear=(0.0, -2.0, 0.5), angular=(3.0, -1.0, 2.0),
                rotation=(0.9, 0.1, -0.3, 0.3))]
    kw = dict(bias=0.005, threshold=0.03, voxel=(0.1,) * 3)
    with api.Context(0, api.FP64) as ctx:
        b = add_body(ctx, mesh, mats, pos.astype(np.float64), vel.astype(np.float64))
        got, n = gpu_contacts(ctx, b, fields, shp, 1 / 60, **kw)
        want = oracle_contacts(ctx, b, mesh, mats, fields, shp, 1 / 60, **kw)
    assert n == len(want) > 0
    assert np.array_equal(got["triangle"], want["triangle"]) and np.array_equal(got["nodes"], want["nodes"])
    for key in ("weights", "normal", "local_point", "gap", "bias", "friction"):
        assert np.allclose(got[key], want[key], rtol=1e-5, atol=1e-5), key


def test_contacts_after_steps_follow_the_state():
    """The query reads the body's device state after its steps."""
    _, mesh, mats = cube_body(origin=(-0.25, 0.02, -0.25))
    fields = [plane_field()]
    with api.Context(0, api.FP64) as ctx:
        b = add_body(ctx, mesh, mats)
        fid = ctx.field_add(fields[0])
        for _ in range(3):
            ctx.step(substeps=1, gravity=(0.0, -9.81, 0.0))
        got, n = ctx.find_contacts(b, [dict(field=fid)], 1 / 60, bias=0.0, threshold=0.05, gravity=(0, -9.81, 0),
                                   voxel=(0.125,) * 3)
        want = oracle_contacts(ctx, b, mesh, mats, fields, [dict(field=0)], 1 / 60, bias=0.0, threshold=0.05,
                               gravity=(0, -9.81, 0), voxel=(0.125,) * 3)
        x, _, _ = ctx.state(b)
    assert x[:, 1].min() < 0.02 - 1e-4  # it fell
    assert n == len(want) > 0 and got.tobytes() == want.tobytes()


def test_bad_arguments_are_refused():
    with api.Context(0, api.FP64) as ctx:
        _, mesh, mats = cube_body()
        b = add_body(ctx, mesh, mats)
        with pytest.raises(api.TalusError):
            ctx.find_contacts(b, [dict(field=7)], 1 / 60)
        with pytest.raises(api.TalusError):
            ctx.field_add(dict(origin=(0, 0, 0), voxel=(1, 1, 1), res=(1, 4, 4), distances=np.zeros(16, np.float32)))
        got, n = ctx.find_contacts(b, [], 1 / 60)
        assert n == 0 and len(got) == 0
