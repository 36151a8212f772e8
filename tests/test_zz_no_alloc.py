"""The allocation-free operator entry point (talus_mi_apply_system_noalloc)
against talus_mi_apply_system through the C ABI: the same values bit for
bit, on two calls in a row (the second on a smaller body, so the shared
scratch array holds the first call's slots), in both precisions, plus its
refusals (a body beyond the scratch array, a context not built since its
last edit)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import deformed


class DeviceBuffers:
    """Device buffers for the entry point's inputs and outputs, from the HIP
    runtime the library itself runs on (libamdhip64.so.7, loaded with it), so
    the test holds no second runtime's state."""

    H2D, D2H = 1, 2

    def __init__(self):
        api.lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} returned HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        self._ok(self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.held.append(p.value)
        self._ok(self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy")
        return p.value

    def download(self, ptr, n):
        out = np.empty(n, dtype=np.float64)
        self._ok(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self._ok(self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, self.D2H), "hipMemcpy")
        return out

    def close(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held.clear()


@pytest.fixture
def buffers():
    held = DeviceBuffers()
    yield held
    held.close()


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
def test_noalloc_operator_equals_apply_system_on_two_calls(precision, buffers):
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
            dv = buffers.upload(d.reshape(-1))
            y = buffers.upload(np.full(3 * n, np.nan))
            ctx.apply_system_noalloc(b, inertia, step, dv, y)
            got = buffers.download(y, 3 * n).reshape(n, 3)
            assert np.isfinite(expect).all() and np.abs(expect).max() > 0.0
            assert np.array_equal(got, expect)


@pytest.mark.gpu
def test_noalloc_operator_refuses_oversized_and_unbuilt_bodies(buffers):
    dv = buffers.upload(np.zeros(3))
    with api.Context(0, api.FP64) as ctx:
        # 50 x 50 x 35 hexes = 525,000 tets: 12 doubles each exceed 48 MiB
        mesh, mats = scenes.build_mesh(_body((50, 50, 35), plastic=False))
        b = ctx.add_body(mesh, mats)
        with pytest.raises(api.TalusError, match="scratch bytes"):
            ctx.apply_system_noalloc(b, 1.0, 0.01, dv, dv)
    with api.Context(0, api.FP64) as ctx:
        b, _ = _prepare(ctx, _body((3, 2, 2), plastic=False), seed=5)
        n, _ = ctx.info(b)
        ctx.apply_system(b, 1.0, 0.01, np.zeros((n, 3)))  # builds the device arrays
        ctx.add_body(*scenes.build_mesh(_body((2, 2, 2), plastic=False)))  # an edit, not built yet
        with pytest.raises(api.TalusError, match="not built"):
            ctx.apply_system_noalloc(b, 1.0, 0.01, dv, dv)
