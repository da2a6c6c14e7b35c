"""The kernels behind the headline numbers.

Full-force batches with drag run split stages; their SH kernel is the
one-workgroup kernel (splitShKernel: 16 or 8 waves per 64-object tile, shared
or per-lane epoch records) or, with the split_groups option while few tiles are
active, the grouped kernel (splitShGroupKernel) that spreads each tile's SH
orders over 2 to 4 workgroups.

Tolerances: one object alone vs in a 16,384-object batch, bit for bit (the
default one-workgroup kernel sums each object's terms in a fixed order); grouped
vs one workgroup 1e-12 relative (only the SH summation order differs); against
the CPU oracle 1 mm and 1 um/s (1 h, or the steps of the large batch);
per-lane (interpolated) records against the oracle's exact epochs also 1 mm and
1 um/s.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _per_lane(t):
    t = dict(t)
    t["epoch"] = t["epoch"] + np.arange(t["n"]) * 0.37 / 86400.0   # distinct epochs
    return t


@pytest.mark.parametrize("waves", [16, 8])
@pytest.mark.parametrize("per_lane", [False, True])
def test_grouped_sh_matches_the_one_workgroup_kernel(full_workload, waves, per_lane):
    """128 objects (2 tiles): split_groups 1 (the default) runs the
    one-workgroup kernel, 2 and 4 the grouped kernel; only the SH summation
    order differs (1e-12 relative after 30 minutes)."""
    cfg, t = full_workload
    opt = {"split_waves": waves, "force_grid": 1} if per_lane else {"split_waves": waves}
    if per_lane:
        t = _per_lane(t)
    e1, s1, st1 = _ctx(cfg, **opt).propagate(t, 1800.0)
    assert st1["sh_groups_min"] == st1["sh_groups_max"] == 1 and st1["kernels_per_step"] == 27
    for g in (2, 4):
        e, s, st = _ctx(cfg, split_groups=g, **opt).propagate(t, 1800.0)
        assert st["sh_groups_min"] == st["sh_groups_max"] == g
        np.testing.assert_array_equal(e, e1)
        rel = np.max(np.linalg.norm(s[:3] - s1[:3], axis=0) / np.linalg.norm(s1[:3], axis=0))
        assert rel < 1e-12


@pytest.mark.parametrize("waves", [16, 8])
@pytest.mark.parametrize("per_lane", [False, True])
def test_one_workgroup_kernel_matches_oracle(full_workload, waves, per_lane):
    """The one-workgroup SH kernel with the drag tile kernel against the oracle
    over one hour."""
    cfg, t = full_workload
    opt = {"split_waves": waves, "force_grid": 1} if per_lane else {"split_waves": waves}
    if per_lane:
        t = _per_lane(t)
    e_g, s_g, st = _ctx(cfg, split_groups=1, **opt).propagate(t, 3600.0)
    assert st["sh_groups_max"] == 1
    e_o, s_o, _ = _oracle(cfg).propagate(t, 3600.0, threads=8)
    np.testing.assert_array_equal(e_g, e_o)
    _compare(s_g, s_o)


@pytest.fixture(scope="module")
def large_workload(workdir):
    """16,384 objects of the full-force population (256 tiles: every CU holds a
    tile, so the default options run the one-workgroup kernel)."""
    from parahpop_amd import synth
    cfg, tg = synth.make_workload(str(workdir / "w2big"), 2, n=16384, duration=600.0)
    return cfg, synth.load_targets(tg)


def test_large_batch_runs_one_workgroup_kernel_and_keeps_each_objects_bits(large_workload):
    """The headline shape for ten steps: 64 objects of the batch equal the same
    objects run alone (one tile) bit for bit, so an object's result does not
    depend on the rest of its batch, and the oracle within 1 mm / 1 um/s."""
    cfg, t = large_workload
    ctx = _ctx(cfg)
    e_b, s_b, st = ctx.propagate(t, 600.0)
    assert st["sh_groups_min"] == st["sh_groups_max"] == 1 and st["waves"] == 16
    idx = np.linspace(0, t["n"] - 1, 64).astype(int)
    sub = _subset(t, idx)
    e_a, s_a, st_a = ctx.propagate(sub, 600.0)
    assert st_a["sh_groups_max"] == 1
    np.testing.assert_array_equal(e_a, e_b[idx])
    np.testing.assert_array_equal(s_a, s_b[:, idx])
    e_o, s_o, _ = _oracle(cfg).propagate(sub, 600.0, threads=8)
    np.testing.assert_array_equal(e_b[idx], e_o)
    _compare(s_b[:, idx], s_o)


def test_large_per_lane_batch_matches_oracle(large_workload):
    """The uneven-batch shape (per-lane epoch records, one-workgroup kernel at
    256 tiles) for five steps: 64 objects against the oracle."""
    cfg, t = large_workload
    t = _per_lane(t)
    ctx = _ctx(cfg, force_grid=1)
    b = ctx.batch(t, 300.0)
    st = b.run()
    assert b.done() and st["sh_groups_max"] == 1 and st["kernels_per_step"] == 27
    e_b, s_b = b.download()
    idx = np.linspace(0, t["n"] - 1, 64).astype(int)
    e_o, s_o, _ = _oracle(cfg).propagate(_subset(t, idx), 300.0, threads=8)
    np.testing.assert_array_equal(e_b[idx], e_o)
    _compare(s_b[:, idx], s_o)


def test_batch_state_reads_a_partial_run(full_workload):
    """phx_batch_state after part of a run equals a run of that length."""
    cfg, t = full_workload
    ctx = _ctx(cfg)
    b = ctx.batch(t, 3600.0)
    b.run(10)
    e, s = b.state()
    e10, s10, _ = ctx.propagate(t, 600.0)
    np.testing.assert_array_equal(e, e10)
    np.testing.assert_array_equal(s, s10)
