"""The operator with the initial-stress term (initialStress /
applyInitialStress, material.cc:1077-1174) through the C ABI against the
oracle, whose term is pinned by central differences of the corotated forces
(tests/test_plastic_newton_cpu.py): FP64 to 1e-12, mixed precision to 1e-5,
for an elastic and a plastic corotated body and a neo-Hookean one (which
adds no term)."""
from __future__ import annotations

import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import deformed, oracle_body, product_body

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model,plastic", [(api.COROTATED, False), (api.COROTATED, True), (api.NEO_HOOKEAN, False)])
@pytest.mark.parametrize("precision,bar", [(api.FP64, 1e-12), (api.MIXED, 1e-5)])
def test_stressed_operator_matches_the_oracle(model, plastic, precision, bar):
    kw = dict(yield_stress=2e3, hardening_modulus=1e4) if plastic else {}
    mats = scenes.material_array(1, youngs_modulus=1e6, poissons_ratio=0.3, density=1000.0,
                                 elastic_model=model, **kw)
    body = scenes.Body(resolution=(6, 3, 3), voxel=(0.1, 0.05, 0.05), materials=mats,
                       supports=[((0.0, 0.0, 0.0), (0.01, 1e3, 1e3))])
    orc, om = oracle_body(body)
    x = deformed(om["rest"], seed=6, scale=0.01, twist=0.5)
    d = np.random.default_rng(7).standard_normal(x.shape)
    inertia, step = 0.75, 1.0 / 120.0
    orc.eval_forces(x)
    ref = inertia * om["mass"][:, None] * d + step * step * orc.stressed_apply(x, d)
    ref[om["anchored"] != 0] = 0.0
    with api.Context(0, precision) as ctx:
        b, _ = product_body(ctx, body)
        y = ctx.apply_system_stressed(b, x, inertia, step, d)
    assert np.abs(y - ref).max() <= bar * np.abs(ref).max()
