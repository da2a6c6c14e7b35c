"""The force model on space-weather and EOP tables whose values change every
day and 3-hour slot: F10.7 and F10.7A distinct per day, eight distinct ap per
day, a drifting pole and UT1-UTC, and one TAI-UTC step (37 -> 38 s) at the
first UTC midnight of the run (69 s after its start). With constant tables a
wrong day, slot, lag or interpolation weight gives the right answer; here it
does not.

Tolerances as in the constant-table tests: each force family within 1e-9 of
its magnitude; propagation within 1 mm and 1 um/s of the oracle; epochs
identical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DAY = 86400.0
EPOCH0_ET = 4035.0 * DAY - 43200.0


def _ctx(cfg, **opt):
    import parahpop_amd
    return parahpop_amd.Context(cfg, **opt)


def _oracle(cfg):
    import oracle
    return oracle.Oracle(cfg)


def _compare(s_g, s_o, pos_mm=1.0, vel_um_s=1.0):
    dp = np.linalg.norm(s_g[:3] - s_o[:3], axis=0) * 1e6
    dv = np.linalg.norm(s_g[3:] - s_o[3:], axis=0) * 1e9
    assert np.max(dp) <= pos_mm, np.max(dp)
    assert np.max(dv) <= vel_um_s, np.max(dv)


def _subset(t, idx):
    idx = np.asarray(idx)
    out = {k: (v[idx] if hasattr(v, "shape") and v.ndim == 1 else v) for k, v in t.items()}
    out["state"] = np.ascontiguousarray(t["state"][:, idx])
    out["n"] = len(idx)
    return out


@pytest.fixture(scope="module")
def varying_workload(tmp_path_factory):
    """Full force model, 70x70 field, 128 objects (seed 2 population), varying tables."""
    from parahpop_amd import synth
    d = tmp_path_factory.mktemp("varying")
    cfg, tg = synth.make_workload(str(d), 2, n=128, duration=3600.0, varying=True)
    return cfg, synth.load_targets(tg)


# before and after the TAI-UTC step, in later 3-hour slots, on the next UTC day
ETS = [EPOCH0_ET + x for x in (10.0, 150.0, 10800.0 + 700.0, 0.55 * DAY, 1.2 * DAY)]


@pytest.mark.parametrize("which,name", [(1, "sh"), (4, "drag"), (5, "corrections"), (-1, "all")])
def test_force_families_on_varying_tables(varying_workload, which, name):
    cfg, t = varying_workload
    ctx = _ctx(cfg)
    orc = _oracle(cfg)
    area = t["drag_area"] * 1e-6
    cr_eff = t["cr"] * t["srp_area"] / t["drag_area"]
    for et in ETS:
        acc = ctx.force(t["state"][:3], t["state"][3:], t["mass"], area, cr_eff, t["cd"], et, which)
        ref = np.stack([orc.force(t["state"][:3, i], t["state"][3:, i], et, t["mass"][i], area[i], cr_eff[i], t["cd"][i], which)
                        for i in range(t["n"])], axis=1)
        scale = np.max(np.linalg.norm(ref, axis=0))
        assert scale > 0
        err = np.max(np.linalg.norm(acc - ref, axis=0))
        assert err <= 1e-9 * scale, (name, et - EPOCH0_ET, err, scale)


def test_drag_sees_the_tables_change(varying_workload):
    """The drag family differs between slots and days (the tables are not
    constant in effect), and the GPU follows each change."""
    cfg, t = varying_workload
    ctx = _ctx(cfg)
    area = t["drag_area"] * 1e-6
    cr_eff = t["cr"] * t["srp_area"] / t["drag_area"]
    accs = [ctx.force(t["state"][:3], t["state"][3:], t["mass"], area, cr_eff, t["cd"], et, 4) for et in ETS[1:]]
    for a, b in zip(accs, accs[1:]):
        assert np.max(np.abs(a - b)) > 1e-3 * np.max(np.abs(a))


def test_full_force_one_hour_on_varying_tables(varying_workload):
    """Shared-epoch batch (split stages, exact records) across the TAI-UTC step
    and the UTC day boundary of the weather table."""
    cfg, t = varying_workload
    e_g, s_g, st = _ctx(cfg).propagate(t, 3600.0)
    e_o, s_o, _ = _oracle(cfg).propagate(t, 3600.0, threads=8)
    np.testing.assert_array_equal(e_g, e_o)
    assert st["kernels_per_step"] == 27
    _compare(s_g, s_o)


@pytest.mark.parametrize("waves", [16, 8])
def test_per_lane_headline_kernels_on_varying_tables(varying_workload, waves):
    """Per-lane (interpolated) epoch records with epochs spread over 20 hours,
    so that lanes sit in different weather days and slots and on both sides of
    the TAI-UTC step: the one-workgroup SH kernel with the drag tile kernel."""
    cfg, t = varying_workload
    t = dict(t)
    t["epoch"] = t["epoch"] + np.arange(t["n"]) * (20.0 / 24.0) / t["n"]
    e_g, s_g, st = _ctx(cfg, split_waves=waves, force_grid=1).propagate(t, 1800.0)
    assert st["kernels_per_step"] == 27
    e_o, s_o, _ = _oracle(cfg).propagate(t, 1800.0, threads=8)
    np.testing.assert_array_equal(e_g, e_o)
    _compare(s_g, s_o)


def test_uneven_low_altitude_drag_on_varying_tables(varying_workload):
    """64 objects between 150 and 300 km, per-object epochs over 20 hours and
    per-object durations (600 s and 1830 s) in one uneven batch, on the split
    stages and on the fused kernel: each against the oracle."""
    cfg, t = varying_workload
    n = 64
    rng = np.random.default_rng(57)
    u = rng.normal(size=(3, n))
    u /= np.linalg.norm(u, axis=0)
    w = rng.normal(size=(3, n))
    w -= u * np.sum(u * w, axis=0)
    w /= np.linalg.norm(w, axis=0)
    r = 6378.137 + np.linspace(150.0, 300.0, n)
    low = _subset(t, np.arange(n))
    low["state"] = np.concatenate([u * r, w * np.sqrt(398600.4415 / r)])
    low["epoch"] = low["epoch"] + rng.permutation(n) * (20.0 / 24.0) / n
    dur = np.where(np.arange(n) % 2 == 0, 600.0, 1830.0)
    orc = _oracle(cfg)
    ref = {}
    for d in (600.0, 1830.0):
        idx = np.nonzero(dur == d)[0]
        ref[d] = (idx, orc.propagate(_subset(low, idx), d, threads=8))
    for opt in ({}, {"split": 0}):
        ctx = _ctx(cfg, **opt)
        b = ctx.batch(low, dur)
        st = b.run()
        assert b.done()
        e, s = b.download()
        if not opt:
            assert st["kernels_per_step"] == 27
        for d, (idx, (e_o, s_o, _)) in ref.items():
            np.testing.assert_array_equal(e[idx], e_o)
            _compare(s[:, idx], s_o)
