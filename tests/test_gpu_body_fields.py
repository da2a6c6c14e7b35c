"""Fields of a body other than the Earth (the Moon's spherical harmonics or J2)
in its own axes from .brot2 orientation units, on the GPU against the oracle.

* The third-body force family (point masses plus the Moon's field) at 128
  lunar-orbit positions: within 1e-9 of the family and within 1e-6 of the
  Moon-field part alone, for the IAU body model (types 1 and 2), a MOON_PA
  binary frame on ICRF and on ECLIPJ2000 (type 3), and an explicitly named
  IAU_MOON frame.
* Propagation of lunar orbiters (50-1000 km above the Moon, any inclination)
  for two hours at 30 s: positions within 1 mm and velocities within 1 um/s of
  the oracle, on shared epochs (exact records) and on uneven epochs (grid
  records interpolated per object).
* The full force model (drag, tides, 70x70 Earth field) with the Moon's field
  added, LEO objects and lunar orbiters in one batch, so the split-stage
  kernels carry the Moon's field: one hour within 1 mm / 1 um/s.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DAY = 86400.0


def _lunar_targets(orc, n, seed=11, epochs=None):
    """n lunar orbiters (circular, 50-1000 km altitude, random planes) as
    geocentric ICRF states at their epochs (MJD2000 TDB)."""
    from parahpop_amd import synth
    rng = np.random.default_rng(seed)
    epochs = np.full(n, synth.EPOCH0) if epochs is None else np.asarray(epochs, dtype=float)
    gm = synth.GM_MOON
    state = np.zeros((6, n))
    for i in range(n):
        et = epochs[i] * DAY - 43200.0
        moon = orc.body_position(301, et)
        vmoon = (orc.body_position(301, et + 1.0) - orc.body_position(301, et - 1.0)) / 2.0
        r = synth.R_MOON_FIELD + rng.uniform(50.0, 1000.0)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        w = np.cross(u, rng.normal(size=3))
        w /= np.linalg.norm(w)
        state[:3, i] = moon + r * u
        state[3:, i] = vmoon + np.sqrt(gm / r) * w
    return {"id": np.arange(1, n + 1, dtype=np.int64), "epoch": epochs, "state": state,
            "mass": rng.uniform(100.0, 2000.0, n), "drag_area": rng.uniform(1.0, 10.0, n),
            "srp_area": rng.uniform(1.0, 10.0, n), "cd": np.full(n, 2.2), "cr": np.full(n, 1.3), "n": n}


@pytest.mark.parametrize("field,frame", [("sh", "iau"), ("sh", "moon_pa"), ("j2", "moon_pa_ecl"), ("sh", "iau_named")])
def test_body_field_family_matches_oracle(tmp_path, field, frame):
    import oracle
    import parahpop_amd
    from parahpop_amd import synth
    cfg = synth.make_lunar_workload(str(tmp_path), field=field, frame=frame)
    orc = oracle.Oracle(cfg)
    t = _lunar_targets(orc, 128)
    et = synth.EPOCH0 * DAY - 43200.0 + 1234.5
    area = t["drag_area"] * 1e-6
    cr_eff = t["cr"] * t["srp_area"] / t["drag_area"]
    with parahpop_amd.Context(cfg) as ctx:
        acc = ctx.force(t["state"][:3], t["state"][3:], t["mass"], area, cr_eff, t["cd"], et, 2)
    ref = np.stack([orc.force(t["state"][:3, i], t["state"][3:, i], et, t["mass"][i], area[i], cr_eff[i], t["cd"][i], 2)
                    for i in range(t["n"])], axis=1)
    # the Moon-field part: the family less its point masses (Moon and Sun)
    gms = {10: 132712440041.27942, 301: synth.GM_MOON}
    point = np.zeros_like(ref)
    for naif, gm in gms.items():
        b = orc.body_position(naif, et)
        d = t["state"][:3] - b[:, None]
        point += -gm * d / np.linalg.norm(d, axis=0) ** 3 - (gm * b / np.linalg.norm(b) ** 3)[:, None]
    fieldpart = np.linalg.norm(ref - point, axis=0)
    err = np.linalg.norm(acc - ref, axis=0)
    assert np.all(fieldpart > 1e-12 * np.linalg.norm(ref, axis=0))
    assert np.max(err / np.linalg.norm(ref, axis=0)) <= 1e-9
    assert np.max(err / fieldpart) <= 1e-6, np.max(err / fieldpart)


@pytest.mark.parametrize("frame,uneven", [("iau", False), ("moon_pa_ecl", False), ("moon_pa", True)])
def test_lunar_orbiters_match_oracle(tmp_path, frame, uneven):
    import oracle
    import parahpop_amd
    from parahpop_amd import synth
    cfg = synth.make_lunar_workload(str(tmp_path), field="sh", frame=frame, duration=7200.0, step=30.0)
    orc = oracle.Oracle(cfg)
    epochs = None
    if uneven:
        epochs = synth.EPOCH0 + np.random.default_rng(5).uniform(0.0, 1.0, 96)
    t = _lunar_targets(orc, 96, epochs=epochs)
    with parahpop_amd.Context(cfg) as ctx:
        e_g, s_g, st = ctx.propagate(t, 7200.0)
    e_o, s_o, _ = orc.propagate(t, 7200.0, threads=8)
    np.testing.assert_array_equal(e_g, e_o)
    dp = np.linalg.norm(s_g[:3] - s_o[:3], axis=0) * 1e6
    dv = np.linalg.norm(s_g[3:] - s_o[3:], axis=0) * 1e9
    print("lunar orbiters (%s, uneven %s): max %.3g mm, %.3g um/s" % (frame, uneven, dp.max(), dv.max()))
    assert dp.max() <= 1.0 and dv.max() <= 1.0


def test_full_force_with_the_moon_field_split_stages(tmp_path):
    import oracle
    import parahpop_amd
    from parahpop_amd import brie_io, synth
    cfg, tg = synth.make_workload(str(tmp_path), 2, n=64, duration=3600.0)
    d = os.path.dirname(cfg)
    synth.write_gravity(os.path.join(d, "moon.txt"), 8, 4, radius=synth.R_MOON_FIELD, gm=synth.GM_MOON, c20=-9.09e-5,
                        scale=2e-5)
    brie_io.write_brot2_units(os.path.join(d, "moon.brot2"), brie_io.moon_iau_units())
    doc = json.load(open(cfg))
    env = doc["model"]["environment"]
    env["bodies"]["Moon"] = {"gravity": {"shape": {"body": "Moon", "file": "moon.txt", "delimiter": ",",
                                                   "headerLines": 0, "degree": 8, "order": 8}}}
    env["orientations"] = ["moon.brot2"]
    json.dump(doc, open(cfg, "w"))
    orc = oracle.Oracle(cfg)
    leo = synth.load_targets(tg)
    lunar = _lunar_targets(orc, 64)
    t = {k: (np.concatenate([leo[k], lunar[k]], axis=-1) if hasattr(leo[k], "shape") else leo[k]) for k in leo}
    t["id"] = np.arange(1, 129, dtype=np.int64)
    t["n"] = 128
    with parahpop_amd.Context(cfg) as ctx:
        assert ctx.option("split") == 1
        e_g, s_g, st = ctx.propagate(t, 3600.0)
        assert st["kernels_per_step"] > 0
    e_o, s_o, _ = orc.propagate(t, 3600.0, threads=8)
    np.testing.assert_array_equal(e_g, e_o)
    dp = np.linalg.norm(s_g[:3] - s_o[:3], axis=0) * 1e6
    dv = np.linalg.norm(s_g[3:] - s_o[3:], axis=0) * 1e9
    print("full force + Moon field: LEO max %.3g mm, lunar max %.3g mm" % (dp[:64].max(), dp[64:].max()))
    assert dp.max() <= 1.0 and dv.max() <= 1.0
