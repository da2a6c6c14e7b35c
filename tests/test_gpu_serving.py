"""Serving (continuous batching) through phx_submit / phx_poll.

* A request's result is the same alone and in a mix of concurrent requests
  (bit-identical: every lane reads only its own state and its own request's
  epoch records).
* Results agree with the CPU oracle within 1 mm / 1 um/s (grid epoch records,
  durations up to two hours).
* Admission by footprint: with a pool smaller than the sum of the requests,
  later requests wait for slots and still complete with the same results;
  a request larger than the whole pool is refused with PHX_ENOMEM.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _subset(t, idx):
    idx = np.asarray(idx)
    out = {k: (v[idx] if hasattr(v, "shape") and v.ndim == 1 else v) for k, v in t.items()}
    out["state"] = np.ascontiguousarray(t["state"][:, idx])
    out["n"] = len(idx)
    return out


@pytest.fixture(scope="module")
def uneven(tmp_path_factory):
    from parahpop_amd import synth
    cfg, tg = synth.make_workload(str(tmp_path_factory.mktemp("serve")), 4, n=300, duration=7200.0)
    return cfg, synth.load_targets(tg)


def _compare(s_g, s_o, pos_mm=1.0, vel_um_s=1.0):
    dp = np.linalg.norm(s_g[:3] - s_o[:3], axis=0) * 1e6
    dv = np.linalg.norm(s_g[3:] - s_o[3:], axis=0) * 1e9
    assert np.max(dp) <= pos_mm, np.max(dp)
    assert np.max(dv) <= vel_um_s, np.max(dv)


def test_serve_alone_equals_mix_and_oracle(uneven):
    import oracle
    import parahpop_amd
    cfg, t = uneven
    reqs = [(_subset(t, range(0, 37)), 1800.0), (_subset(t, range(37, 137)), 3600.0),
            (_subset(t, range(137, 138)), 600.0), (_subset(t, range(138, 300)), 7200.0),
            (_subset(t, range(0, 37)), 1800.0)]
    with parahpop_amd.Context(cfg, serve_steps_per_tick=4) as ctx:
        alone = ctx.wait(ctx.submit(*reqs[0]))
        tickets = [ctx.submit(r, d) for r, d in reqs]
        mixed = [ctx.wait(k) for k in tickets]
        assert ctx.option("serve_graph_captures") >= 1
        assert ctx.option("serve_completed") == 6
        assert ctx.option("serve_object_steps") > 0
    np.testing.assert_array_equal(mixed[0][1], alone[1])
    np.testing.assert_array_equal(mixed[4][1], alone[1])
    np.testing.assert_array_equal(mixed[0][0], alone[0])
    orc = oracle.Oracle(cfg)
    for (r, d), (e, s, t_sub, t_fin) in zip(reqs[:4], mixed[:4]):
        e_o, s_o, _ = orc.propagate(r, d, threads=8)
        np.testing.assert_array_equal(e, e_o)
        _compare(s, s_o)
        assert t_fin >= t_sub


def test_serve_admission_waits_for_footprint(uneven):
    import parahpop_amd
    cfg, t = uneven
    a, b, c = _subset(t, range(0, 100)), _subset(t, range(100, 200)), _subset(t, range(200, 250))
    with parahpop_amd.Context(cfg) as ref:
        ra = ref.propagate(a, 1200.0)
        rb = ref.propagate(b, 600.0)
        rc = ref.propagate(c, 900.0)
    # 128 slots: a (100) runs alone; b (100) waits for a; c (50) waits behind b
    with parahpop_amd.Context(cfg, serve_slots=128, serve_steps_per_tick=2) as ctx:
        ka, kb, kc = ctx.submit(a, 1200.0), ctx.submit(b, 600.0), ctx.submit(c, 900.0)
        out = [ctx.wait(k) for k in (ka, kb, kc)]
        with pytest.raises(parahpop_amd.PhxError) as ex:
            ctx.submit(_subset(t, range(0, 200)), 600.0)
        assert ex.value.code == parahpop_amd.PHX_ENOMEM
    # b could only start after a finished
    assert out[1][3] >= out[0][3]
    for (e, s, _, _), (e_r, s_r, _) in zip(out, (ra, rb, rc)):
        np.testing.assert_array_equal(e, e_r)
        # serving uses grid epoch records, the batch path exact ones
        _compare(s, s_r, pos_mm=1e-2, vel_um_s=1e-2)


def test_serve_rejections(uneven):
    import parahpop_amd
    cfg, t = uneven
    with parahpop_amd.Context(cfg) as ctx:
        bad = _subset(t, range(4))
        bad["mass"] = bad["mass"].copy()
        bad["mass"][2] = -1.0
        with pytest.raises(parahpop_amd.PhxError) as ex:
            ctx.submit(bad, 600.0)
        assert ex.value.code == parahpop_amd.PHX_EINVAL
        with pytest.raises(parahpop_amd.PhxError):
            ctx.submit(_subset(t, range(4)), -5.0)
        ok = ctx.submit(_subset(t, range(4)), 300.0)
        ctx.wait(ok)
        ctx._pending[12345] = 1
        with pytest.raises(parahpop_amd.PhxError):
            ctx.poll(12345)
