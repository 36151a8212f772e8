"""The allocation-free operator entry point (talus_mi_apply_system_noalloc)
against talus_mi_apply_system through the C ABI: the same values bit for
bit, on two calls in a row (the second on a smaller body, so the shared
scratch array holds the first call's slots), in both precisions, plus its
refusals (a body beyond the scratch array, a context not built since its
last edit)."""
from __future__ import annotations

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import deformed


def _body(cells, plastic):
    kw = dict(youngs_modulus=2e6, poissons_ratio=0.35, density=1000.0, elastic_model=api.NEO_HOOKEAN)
    if plastic:
        kw.update(yield_stress=1e3, hardening_modulus=1e4)
    return scenes.Body(resolution=tuple(c + 1 for c in cells), voxel=(0.05, 0.05, 0.05),
                       supports=[((0.0, 0.0, 0.0), (0.01, 1e3, 1e3))],
                       materials=scenes.material_array(1, **kw))


def _prepare(ctx, body, seed):
    """Adds a body, puts it in a turned and strained state (committed
    rotations and volume frames at that state) with seeded plastic offsets."""
    mesh, mats = scenes.build_mesh(body)
    b = ctx.add_body(mesh, mats)
    rest = mesh["rest_positions"].astype(np.float64)
    x = deformed(rest, seed=seed, scale=0.004, twist=0.6)
    ctx.set_state(b, x)
    if np.isfinite(mats["yield_stress"][0]):
        rng = np.random.default_rng(seed)
        pl = np.zeros((mesh["tet_nodes"].shape[0], 10))
        pl[:, :9] = 0.05 * rng.standard_normal((pl.shape[0], 9))
        ctx.set_plastic(b, pl)
    return b, x


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [api.FP64, api.MIXED])
def test_noalloc_operator_equals_apply_system_on_two_calls(precision):
    import torch

    with api.Context(0, precision) as ctx:
        big, xb = _prepare(ctx, _body((20, 5, 5), plastic=True), seed=3)  # 3,000 tets
        small, xs = _prepare(ctx, _body((3, 2, 2), plastic=False), seed=4)  # 72 tets
        for b, x in ((big, xb), (small, xs)):
            ctx.eval_forces(b, x)  # rotations and volume frames at x
        rng = np.random.default_rng(11)
        for b, inertia, step in ((big, 1.0, 1.0 / 240.0), (small, 0.75, 1.0 / 60.0)):
            n, _ = ctx.info(b)
            d = rng.standard_normal((n, 3))
            expect = ctx.apply_system(b, inertia, step, d)
            dv = torch.tensor(d.reshape(-1), dtype=torch.float64, device="cuda")
            y = torch.full_like(dv, float("nan"))
            torch.cuda.synchronize()
            ctx.apply_system_noalloc(b, inertia, step, dv.data_ptr(), y.data_ptr())
            torch.cuda.synchronize()
            got = y.cpu().numpy().reshape(n, 3)
            assert np.isfinite(expect).all() and np.abs(expect).max() > 0.0
            assert np.array_equal(got, expect)


@pytest.mark.gpu
def test_noalloc_operator_refuses_oversized_and_unbuilt_bodies():
    import torch

    dv = torch.zeros(3, dtype=torch.float64, device="cuda")
    with api.Context(0, api.FP64) as ctx:
        # 50 x 50 x 35 hexes = 525,000 tets: 12 doubles each exceed 48 MiB
        mesh, mats = scenes.build_mesh(_body((50, 50, 35), plastic=False))
        b = ctx.add_body(mesh, mats)
        with pytest.raises(api.TalusError, match="scratch bytes"):
            ctx.apply_system_noalloc(b, 1.0, 0.01, dv.data_ptr(), dv.data_ptr())
    with api.Context(0, api.FP64) as ctx:
        b, _ = _prepare(ctx, _body((3, 2, 2), plastic=False), seed=5)
        n, _ = ctx.info(b)
        ctx.apply_system(b, 1.0, 0.01, np.zeros((n, 3)))  # builds the device arrays
        ctx.add_body(*scenes.build_mesh(_body((2, 2, 2), plastic=False)))  # an edit, not built yet
        with pytest.raises(api.TalusError, match="not built"):
            ctx.apply_system_noalloc(b, 1.0, 0.01, dv.data_ptr(), dv.data_ptr())
