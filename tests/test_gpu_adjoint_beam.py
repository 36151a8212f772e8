"""GPU: the differentiation workload at its full size.

The 1 x 0.1 x 0.1 m beam (103 x 9 x 9 hexes, 50,058 tets, per-tet
lognormal E and uniform nu), a converged 60-step FP64 rollout with
checkpoints every 10 steps, loss |x_tip - target|^2: the autograd gradients
(adjoint of the implicit solve through talus_mi_rollout_backward) in E and
nu of two seeded tets and in two rest coordinates against central finite
differences of the same rollout (h = 1e-6 relative, 1e-7 m), bar 1e-4 of
the largest gradient.  The small-beam tests pin the adjoint against the
FP64 oracle's gradient; this one checks the full-size workload.
"""
import pytest

import bench

pytestmark = pytest.mark.gpu


def test_full_size_beam_gradients_match_finite_differences():
    out = bench.adjoint_check(steps=60, every=10, elements=2, rest_coords=2)
    assert out["tets"] == 50058
    for kind in ("E", "nu", "rest"):
        assert out[kind]["checked"] == 2
        assert out[kind]["max_abs_err_over_max_grad"] <= 1e-4, (kind, out[kind])
