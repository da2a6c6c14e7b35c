"""Split stages replayed as graphs.

The split path captures `graph_steps` steps (13 x (SH kernel, drag kernel) +
the combine kernel per step) as one graph and replays it while the active tiles
stay the same; exact records take their step index from device memory, which
the combine kernel advances. The results must be bit for bit those of
launching every kernel (graph_steps = 0): for shared epochs (exact records),
for uneven epochs and durations (grid records, tiles ending at different
steps), and across calls that run part of a batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(cfg, **opt):
    import parahpop_amd
    return parahpop_amd.Context(cfg, **opt)


def test_exact_records_graph_replay_equals_launches(full_workload):
    cfg, t = full_workload
    e0, s0, st0 = _ctx(cfg, graph_steps=0).propagate(t, 1800.0)
    assert st0["graph_replays"] == 0 and st0["kernels_per_step"] == 27
    e1, s1, st1 = _ctx(cfg, graph_steps=8).propagate(t, 1800.0)
    assert st1["graph_replays"] == 3 and st1["graph_captures"] == 1   # 24 of the 30 steps
    np.testing.assert_array_equal(e1, e0)
    np.testing.assert_array_equal(s1, s0)


def test_graph_replay_across_partial_runs(full_workload):
    """A batch run as 5 + 20 + 5 steps (graphs inside the calls) equals one run."""
    cfg, t = full_workload
    ctx = _ctx(cfg, graph_steps=4)
    b = ctx.batch(t, 1800.0)
    reps = 0
    for k in (5, 20, 5):
        reps += b.run(k)["graph_replays"]
    assert b.done() and reps == 1 + 5 + 1
    e, s = b.download()
    e0, s0, _ = _ctx(cfg, graph_steps=0).propagate(t, 1800.0)
    np.testing.assert_array_equal(e, e0)
    np.testing.assert_array_equal(s, s0)


def test_uneven_batch_graph_replay_equals_launches(workdir):
    """Per-object epochs and durations (grid records): tiles end at different
    steps, so the graphs are captured per active-tile set and single steps fill
    in where a tile ends within the next graph_steps steps."""
    from parahpop_amd import synth
    cfg, tg = synth.make_workload(str(workdir / "w4g"), 4, n=640, duration=7200.0)
    t = synth.load_targets(tg)
    rng = np.random.default_rng(9)
    dur = np.exp(rng.uniform(np.log(600.0), np.log(7200.0), t["n"]))
    out = {}
    for k in (0, 4):
        ctx = _ctx(cfg, graph_steps=k)
        b = ctx.batch(t, dur)
        st = b.run()
        assert b.done() and st["kernels_per_step"] == 27
        out[k] = (b.download(), st)
    (e0, s0), st0 = out[0]
    (e1, s1), st1 = out[4]
    assert st0["graph_replays"] == 0
    assert st1["graph_replays"] >= 4 and st1["graph_captures"] >= 2
    assert st1["object_steps"] == st0["object_steps"]
    np.testing.assert_array_equal(e1, e0)
    np.testing.assert_array_equal(s1, s0)
