"""The Chebyshev (.brie) ephemeris on the GPU against the oracle reading the
same file: the full force model (third bodies, SRP with shadow, tides take the
Sun and Moon from the file). The file's Sun and Moon are displaced from the
analytic series (1,000 km and 100 km), so agreement shows both sides read it.

Tolerances (FP64): force families within 1e-9 of their magnitude; one hour
within 1 mm and 1 um/s, on shared-epoch records and on the interpolated grid."""
import json
import os

import numpy as np
import pytest

from parahpop_amd import brie_io as brie_writer

pytestmark = pytest.mark.gpu

DAY = 86400.0
EPOCH0_ET = 4035.0 * DAY - 43200.0


def _shift(naif, et):
    if naif == 10:
        return np.array([1000.0, -500.0, 200.0])
    if naif == 301:
        return np.array([100.0, 0.0, -50.0])
    return 0.0


@pytest.fixture(scope="module")
def brie_workload(workdir):
    import oracle
    from parahpop_amd import synth
    d = str(workdir / "w2brie")
    cfg, tg = synth.make_workload(d, 2, n=128, duration=3600.0)
    units = brie_writer.ssb_units(oracle.body_position, EPOCH0_ET - 2 * DAY, EPOCH0_ET + 4 * DAY, shift=_shift)
    brie_writer.write_brie(os.path.join(d, "eph.brie"), units)
    doc = json.load(open(cfg))
    doc["model"]["environment"]["ephemeris"] = ["eph.brie"]
    out = os.path.join(d, "CONFIG_brie.json")
    json.dump(doc, open(out, "w"))
    return cfg, out, synth.load_targets(tg)


def _ctx(cfg, **opt):
    import parahpop_amd
    return parahpop_amd.Context(cfg, **opt)


@pytest.mark.parametrize("which,name", [(2, "third"), (3, "srp"), (5, "corrections"), (-1, "all")])
def test_brie_force_families_match_oracle(brie_workload, which, name):
    import oracle
    plain, cfg, t = brie_workload
    ctx, orc = _ctx(cfg), oracle.Oracle(cfg)
    area = t["drag_area"] * 1e-6
    cr_eff = t["cr"] * t["srp_area"] / t["drag_area"]
    et = EPOCH0_ET + 1234.5
    acc = ctx.force(t["state"][:3], t["state"][3:], t["mass"], area, cr_eff, t["cd"], et, which)
    ref = np.stack([orc.force(t["state"][:3, i], t["state"][3:, i], et, t["mass"][i], area[i], cr_eff[i], t["cd"][i], which)
                    for i in range(t["n"])], axis=1)
    scale = np.max(np.linalg.norm(ref, axis=0))
    assert np.max(np.linalg.norm(acc - ref, axis=0)) <= 1e-9 * scale, name
    if which == 2:
        # the file's bodies, not the analytic series, reach the kernel
        ana = _ctx(plain).force(t["state"][:3], t["state"][3:], t["mass"], area, cr_eff, t["cd"], et, which)
        assert np.max(np.linalg.norm(acc - ana, axis=0)) > 1e-5 * scale


@pytest.mark.parametrize("opts", [{}, {"force_grid": 1}])
def test_brie_one_hour_matches_oracle(brie_workload, opts):
    import oracle
    _, cfg, t = brie_workload
    e_g, s_g, _ = _ctx(cfg, **opts).propagate(t, 3600.0)
    e_o, s_o, _ = oracle.Oracle(cfg).propagate(t, 3600.0, threads=8)
    np.testing.assert_array_equal(e_g, e_o)
    dp = np.linalg.norm(s_g[:3] - s_o[:3], axis=0) * 1e6
    dv = np.linalg.norm(s_g[3:] - s_o[3:], axis=0) * 1e9
    assert np.max(dp) <= 1.0 and np.max(dv) <= 1.0, (np.max(dp), np.max(dv))


def test_brie_coverage_is_checked(brie_workload):
    import parahpop_amd
    _, cfg, t = brie_workload
    with pytest.raises(parahpop_amd.PhxError, match="ephemeris range"):
        _ctx(cfg).propagate(t, 10 * DAY)
