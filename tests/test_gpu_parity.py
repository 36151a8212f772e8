"""GPU parity: libtalus_mi (through the C ABI) against the FP64 oracle.

Bars: FP64 element forces within 1e-12 relative of the oracle; mixed
(FP32 element kernel) within 1e-5 relative (max_e |df_e| / max_e |f_e|);
operator within 1e-12; trajectories within the stated tolerances; runs
bit-identical to each other.
"""
import numpy as np
import pytest

from talus_amd import api, scenes

from helpers import deformed, oracle_body, product_body, rel_force_error

pytestmark = pytest.mark.gpu


def _scene_body(model=api.NEO_HOOKEAN, **mat):
    sc = scenes.cantilever()
    b = sc.bodies[0]
    b.materials = scenes.material_array(1, youngs_modulus=1e6, poissons_ratio=0.3,
                                        density=1000.0, elastic_model=model, **mat)
    return b


@pytest.mark.parametrize("model", [api.NEO_HOOKEAN, api.COROTATED])
@pytest.mark.parametrize("precision,bar", [(api.FP64, 1e-12), (api.MIXED, 1e-5)])
def test_element_forces_match_oracle(model, precision, bar):
    body = _scene_body(model)
    orc, om = oracle_body(body)
    x = deformed(om["rest"], seed=1, scale=0.003)
    tf_o, te_o, nf_o, R_o = orc.eval_forces(x)
    with api.Context(0, precision) as ctx:
        b, _ = product_body(ctx, body)
        tf, te, nf, R = ctx.eval_forces(b, x)
    err = rel_force_error(tf, tf_o)
    assert err <= bar, err
    assert np.abs(R - R_o).max() <= (1e-9 if precision == api.FP64 else 1e-5)
    scale = np.abs(nf_o).max()
    assert np.abs(nf - nf_o).max() / scale <= 10 * bar
    assert abs(te.sum() - te_o.sum()) <= 10 * bar * abs(te_o).sum()


@pytest.mark.parametrize("model", [api.NEO_HOOKEAN, api.COROTATED])
@pytest.mark.parametrize("plastic", [False, True])
def test_mixed_forces_under_large_rotation_and_small_strain(model, plastic):
    """A 1 rad turn with 1e-4 strain: FP32 F - I would lose the strain to
    rounding; the mixed kernel forms R^T F - I in FP64 and stays within the
    1e-5 bar."""
    kw = dict(yield_stress=1e12, hardening_modulus=0.0) if plastic else {}
    body = _scene_body(model, **kw)
    orc, om = oracle_body(body)
    rest = om["rest"]
    rng = np.random.default_rng(31)
    c = rest.mean(axis=0)
    a = 1.0
    Rz = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
    strain = 1e-4 * rng.standard_normal((3, 3))
    x = (rest - c) @ (Rz @ (np.eye(3) + strain)).T + c + 1e-6 * rng.standard_normal(rest.shape)
    T = om["tet"].shape[0]
    pl = None
    if plastic:
        pl = np.zeros((T, 10))
        pl[:, :9] = (2e-4 * rng.standard_normal((T, 3, 3))).reshape(T, 9)
        orc.set_plastic(pl)
    tf_o, *_ = orc.eval_forces(x)
    with api.Context(0, api.MIXED) as ctx:
        b, _ = product_body(ctx, body)
        if plastic:
            ctx.set_plastic(b, pl)
        tf, *_ = ctx.eval_forces(b, x)
    assert rel_force_error(tf, tf_o) <= 1e-5


def test_element_forces_jittered_per_tet_materials():
    sc = scenes.cube(6, seed=3)
    body = sc.bodies[0]
    orc, om = oracle_body(body)
    x = deformed(om["rest"], seed=2, scale=0.002, twist=1.0)
    tf_o, *_ = orc.eval_forces(x)
    for precision, bar in [(api.FP64, 1e-12), (api.MIXED, 1e-5)]:
        with api.Context(0, precision) as ctx:
            b, _ = product_body(ctx, body)
            tf, *_ = ctx.eval_forces(b, x)
        assert rel_force_error(tf, tf_o) <= bar


def test_inverted_elements_stay_finite_and_match():
    body = _scene_body(api.COROTATED)
    orc, om = oracle_body(body)
    x = om["rest"].copy()
    # flip a layer of nodes through its neighbours to invert tets
    sel = np.abs(x[:, 0] - 0.4) < 1e-6
    x[sel, 0] -= 0.12
    tf_o, _, _, R_o = orc.eval_forces(x)
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        tf, _, _, R = ctx.eval_forces(b, x)
    assert np.isfinite(tf).all()
    assert rel_force_error(tf, tf_o) <= 1e-10
    assert np.abs(R - R_o).max() <= 1e-8


def test_operator_matches_oracle():
    body = _scene_body()
    orc, om = oracle_body(body)
    x = deformed(om["rest"], seed=4, scale=0.002)
    orc.eval_forces(x)  # commits rotations and volume frames
    R, W = orc.rotations(), orc.frames()
    d = np.random.default_rng(5).standard_normal(om["rest"].shape)
    inertia, step = 1.0, np.sqrt(8.0 / 27.0) / 60.0
    y_o = orc.apply_system(inertia, step, d)
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        ctx.set_rotations(b, R)
        ctx.set_frames(b, W)
        y = ctx.apply_system(b, inertia, step, d)
    assert np.abs(y - y_o).max() <= 1e-12 * np.abs(y_o).max()


def _run_product(body, precision, steps, **kw):
    with api.Context(0, precision) as ctx:
        b, _ = product_body(ctx, body)
        reports = [ctx.step(**kw) for _ in range(steps)]
        x, v, a = ctx.state(b)
    return x, v, a, reports


def _run_oracle(body, steps, tolerance, cg_iterations, substeps=1):
    orc, om = oracle_body(body)
    reps = [orc.step(tolerance=tolerance, cg_iterations=cg_iterations, substeps=substeps)
            for _ in range(steps)]
    x, v, a = orc.state()
    return x, v, a, reps, om


def test_cantilever_trajectory_fp64_tight():
    body = _scene_body()
    steps = 60
    xo, vo, ao, reps, om = _run_oracle(body, steps, 1e-10, 2000)
    x, v, a, reports = _run_product(body, api.FP64, steps, tolerance=1e-10, cg_iterations=2000)
    disp = np.abs(xo - om["rest"]).max()
    assert np.abs(x - xo).max() <= 1e-8 * disp
    assert np.abs(v - vo).max() <= 1e-7 * np.abs(vo).max()
    assert sum(r["cg_capped"] for r in reports) == 0


def test_cantilever_trajectory_default_tolerance():
    body = _scene_body()
    steps = 60
    xo, vo, _, reps, om = _run_oracle(body, steps, 1e-3, 64)
    x, v, _, reports = _run_product(body, api.FP64, steps)
    disp = np.abs(xo - om["rest"]).max()
    assert np.abs(x - xo).max() <= 1e-6 * disp
    its_o = sum(r["cg_iterations"] for r in reps)
    its = sum(r["cg_iterations"] for r in reports)
    assert abs(its - its_o) <= 0.02 * its_o + 3


def test_mixed_trajectory_close_to_fp64():
    body = _scene_body()
    steps = 60
    xo, *_, om = _run_oracle(body, steps, 1e-6, 2000)
    x, *_ = _run_product(body, api.MIXED, steps, tolerance=1e-6, cg_iterations=2000)
    disp = np.abs(xo - om["rest"]).max()
    assert np.abs(x - xo).max() <= 1e-3 * disp


def test_runs_are_bit_identical():
    body = scenes.cube(6, seed=9).bodies[0]
    for precision in (api.FP64, api.MIXED):
        a = _run_product(body, precision, 5)
        b = _run_product(body, precision, 5)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_batched_bodies_equal_single_bodies():
    sc = scenes.bodies(count=4, seed=11, lo=200, hi=2000)
    singles = []
    for body in sc.bodies:
        x, v, a, _ = _run_product(body, api.MIXED, 3, substeps=2)
        singles.append(x)
    with api.Context(0, api.MIXED) as ctx:
        ids = [product_body(ctx, body)[0] for body in sc.bodies]
        for _ in range(3):
            ctx.step(substeps=2)
        for i, b in enumerate(ids):
            x, _, _ = ctx.state(b)
            assert np.array_equal(x, singles[i])


def test_plastic_yield_events_match_oracle():
    body = _scene_body(api.NEO_HOOKEAN, yield_stress=2e3, hardening_modulus=1e3)
    orc, om = oracle_body(body)
    steps = 20
    yo = []
    for s in range(steps):
        orc.step(tolerance=1e-10, cg_iterations=2000)
        yo.append(np.flatnonzero(orc.yield_flags()))
    with api.Context(0, api.FP64) as ctx:
        b, _ = product_body(ctx, body)
        for s in range(steps):
            ctx.step(tolerance=1e-10, cg_iterations=2000, flags=api.RECORD_EVENTS)
        ev = ctx.events()
        po = ctx.plastic(b)
        x, _, _ = ctx.state(b)
    yg = [np.array(sorted(e[2] for e in ev if e[0] == api.EVENT_YIELD and e[3] == s)) for s in range(steps)]
    total = sum(len(y) for y in yo)
    assert total > 0
    for s in range(steps):
        assert np.array_equal(yg[s], yo[s]), s
    assert np.abs(po - orc.plastic()).max() <= 1e-8


@pytest.mark.parametrize("streams", [1, 2])
def test_body_reports_split_the_step_report(streams):
    """talus_mi_body_report: each body's own StepReport (deformable::step
    reports per body); the batch's counters are their sums, a body stepped
    alone reports what it reports batched."""
    sc = scenes.bodies(count=5, seed=12, lo=200, hi=2000)
    params = api.default_params(substeps=2, preconditioner=api.FACTOR)
    alone = []
    for body in sc.bodies:
        with api.Context(0, api.FP64) as ctx:
            b, _ = product_body(ctx, body)
            ctx.step(params)
            alone.append(ctx.body_report(b))
    with api.Context(0, api.FP64) as ctx:
        ids = [product_body(ctx, body)[0] for body in sc.bodies]
        ctx.set_streams(streams)
        assert ctx.body_report(ids[0])["solves"] == 0  # before the first step
        rep = ctx.step(params)
        own = [ctx.body_report(b) for b in ids]
    for key in ("cg_iterations", "cg_capped", "solves", "yield_events"):
        assert rep[key] == sum(r[key] for r in own), key
    for key in ("kinetic_energy", "elastic_energy"):
        assert rep[key] == pytest.approx(sum(r[key] for r in own), rel=1e-12)
    assert rep["cg_worst_relative"] == max(r["cg_worst_relative"] for r in own)
    for r, a, body in zip(own, alone, sc.bodies):
        assert r["bodies"] == 1 and r["elements"] == a["elements"] and r["substeps"] == 2
        assert r["solves"] == a["solves"] == 6
        assert r["kinetic_energy"] == pytest.approx(a["kinetic_energy"], rel=1e-9)
