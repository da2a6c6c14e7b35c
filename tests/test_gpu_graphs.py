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
    # 30 steps = 3 x 8 + 4 + 2; the 8-step graph is captured with the 4- and 2-step ones
    assert st1["graph_replays"] == 5 and st1["graph_captures"] == 3
    np.testing.assert_array_equal(e1, e0)
    np.testing.assert_array_equal(s1, s0)


def test_graph_replay_across_partial_runs(full_workload):
    """A batch run as 5 + 20 + 5 steps (graphs inside the calls) equals one run;
    the graphs captured by the first call are replayed by the later ones."""
    cfg, t = full_workload
    ctx = _ctx(cfg, graph_steps=4)
    b = ctx.batch(t, 1800.0)
    reps, caps = [], []
    for k in (5, 20, 5):
        st = b.run(k)
        reps.append(st["graph_replays"])
        caps.append(st["graph_captures"])
    assert b.done() and reps == [1, 5, 1] and caps == [2, 0, 0]
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


def test_option_change_between_calls_recaptures(full_workload):
    """A graph holds the kernel arguments of the options it was captured under:
    after an option change the next call captures again, and the batch equals
    one that launches every kernel under the same sequence of options."""
    cfg, t = full_workload
    out = {}
    for k in (0, 4):
        ctx = _ctx(cfg, graph_steps=k)
        b = ctx.batch(t, 1800.0)
        b.run(10)
        ctx.set_option("split_cost_pg", 40.0)   # moves the SH order blocks between waves
        st = b.run(10)
        out[k] = (b.state(), st)
        if k:
            assert st["graph_captures"] >= 1
    (e0, s0), _ = out[0]
    (e1, s1), _ = out[4]
    np.testing.assert_array_equal(e1, e0)
    np.testing.assert_array_equal(s1, s0)
