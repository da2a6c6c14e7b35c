"""The IERS2000 Earth orientation of a .brot2 file on the GPU against the oracle
reading the same file: the full force model (SH, NRLMSISE-00 drag and tides
rotate with it). The file carries the IAU 2006 channels with a 0.3 arcsec
polar-motion wobble, so agreement shows both sides use the file.

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


@pytest.fixture(scope="module")
def brot_workload(workdir):
    from parahpop_amd import synth
    d = str(workdir / "w2brot")
    cfg, tg = synth.make_workload(d, 2, n=128, duration=3600.0)
    brie_writer.write_brot2(os.path.join(d, "itrf.brot2"), 4035.0 - 4.0, 4035.0 + 6.0,
                            brie_writer.iau2006_channels(xp_arcsec=0.3, yp_arcsec=0.3, wobble=True))
    doc = json.load(open(cfg))
    doc["model"]["environment"]["orientations"] = ["itrf.brot2"]
    out = os.path.join(d, "CONFIG_brot.json")
    json.dump(doc, open(out, "w"))
    return cfg, out, synth.load_targets(tg)


def _ctx(cfg, **opt):
    import parahpop_amd
    return parahpop_amd.Context(cfg, **opt)


@pytest.mark.parametrize("which,name", [(1, "sh"), (4, "drag"), (5, "corrections"), (-1, "all")])
def test_brot_force_families_match_oracle(brot_workload, which, name):
    import oracle
    plain, cfg, t = brot_workload
    ctx, orc = _ctx(cfg), oracle.Oracle(cfg)
    area = t["drag_area"] * 1e-6
    cr_eff = t["cr"] * t["srp_area"] / t["drag_area"]
    et = EPOCH0_ET + 1234.5
    acc = ctx.force(t["state"][:3], t["state"][3:], t["mass"], area, cr_eff, t["cd"], et, which)
    ref = np.stack([orc.force(t["state"][:3, i], t["state"][3:, i], et, t["mass"][i], area[i], cr_eff[i], t["cd"][i], which)
                    for i in range(t["n"])], axis=1)
    scale = np.max(np.linalg.norm(ref, axis=0))
    assert scale > 0 or name == "drag"
    assert np.max(np.linalg.norm(acc - ref, axis=0)) <= 1e-9 * max(scale, 1e-30), name
    if which == 1:
        # the file's rotation, not the analytic model, reaches the kernel
        ana = _ctx(plain).force(t["state"][:3], t["state"][3:], t["mass"], area, cr_eff, t["cd"], et, which)
        assert np.max(np.linalg.norm(acc - ana, axis=0)) > 1e-7 * scale


@pytest.mark.parametrize("opts", [{}, {"force_grid": 1}])
def test_brot_one_hour_matches_oracle(brot_workload, opts):
    import oracle
    _, cfg, t = brot_workload
    e_g, s_g, _ = _ctx(cfg, **opts).propagate(t, 3600.0)
    e_o, s_o, _ = oracle.Oracle(cfg).propagate(t, 3600.0, threads=8)
    np.testing.assert_array_equal(e_g, e_o)
    dp = np.linalg.norm(s_g[:3] - s_o[:3], axis=0) * 1e6
    dv = np.linalg.norm(s_g[3:] - s_o[3:], axis=0) * 1e9
    assert np.max(dp) <= 1.0 and np.max(dv) <= 1.0, (np.max(dp), np.max(dv))


def test_brot_envelope_is_checked(brot_workload):
    import parahpop_amd
    _, cfg, t = brot_workload
    with pytest.raises(parahpop_amd.PhxError, match="orientation range"):
        _ctx(cfg).propagate(t, 5 * DAY)
