"""The reference's build checks (scripts/validate.py:61-131, example/compare.py)
run against the `phx_gpu` process: every case writes CONFIG/TARGETS files,
runs the executable from another working directory, and compares
final_states.csv with the CPU oracle at compare.py's tolerances (1 mm, 1e-6 m/s,
final epochs within 1e-4 s). Failure cases must exit non-zero without writing
summary.json and must leave existing outputs untouched.
"""
import csv
import json
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "parahpop_amd", "phx_gpu")


def _launch(cfg, targets, out, expect_success=True):
    p = subprocess.run([EXE, cfg, targets, out], cwd=tempfile.gettempdir(), stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert (p.returncode == 0) == expect_success, (p.returncode, p.stderr[-2000:])
    if expect_success:
        assert json.load(open(os.path.join(out, "summary.json")))["validation_passed"]
    return p


def _read(out):
    with open(os.path.join(out, "final_states.csv")) as f:
        rows = list(csv.DictReader(f))
    return {int(r["id"]): r for r in rows}


def _targets_json(path, t):
    doc = {"targets": [{"id": int(t["id"][i]), "epoch_mjd2000_tdb": float(t["epoch"][i]),
                        "state_km_km_s": [float(t["state"][d, i]) for d in range(6)],
                        "mass_kg": float(t["mass"][i]), "drag_area_m2": float(t["drag_area"][i]),
                        "srp_area_m2": float(t["srp_area"][i]), "cd": float(t["cd"][i]), "cr": float(t["cr"][i])}
                       for i in range(t["n"])]}
    json.dump(doc, open(path, "w"))


def _compare_with_oracle(cfg, t, out):
    import oracle
    duration = json.load(open(cfg))["run"]["duration_seconds"]
    e_o, s_o, _ = oracle.Oracle(cfg).propagate(t, duration, threads=8)
    rows = _read(out)
    assert sorted(rows) == sorted(int(x) for x in t["id"])
    cols = ["x_km", "y_km", "z_km", "vx_km_s", "vy_km_s", "vz_km_s"]
    dp = dv = 0.0
    for i in range(t["n"]):
        r = rows[int(t["id"][i])]
        g = [float(r[c]) for c in cols]
        assert all(math.isfinite(x) for x in g)
        assert abs(float(r["epoch_mjd2000_tdb"]) - e_o[i]) * 86400 <= 1e-4
        dp = max(dp, 1000 * math.sqrt(sum((g[k] - s_o[k, i]) ** 2 for k in range(3))))
        dv = max(dv, 1000 * math.sqrt(sum((g[k] - s_o[k, i]) ** 2 for k in range(3, 6))))
    assert dp <= 1e-3 and dv <= 1e-6, (dp, dv)
    return dp, dv


@pytest.fixture(scope="module")
def envisat_like(tmp_path_factory):
    """One full-force LEO object in the ENVISAT-like regime (synthetic population)."""
    from parahpop_amd import synth
    d = tmp_path_factory.mktemp("val")
    cfg, tg = synth.make_workload(str(d), 2, n=1, duration=600.0)
    return str(d), cfg, synth.load_targets(tg)


def _variant(d, cfg, name, mutate):
    doc = json.load(open(cfg))
    mutate(doc)
    p = os.path.join(d, name + ".json")
    json.dump(doc, open(p, "w"))
    return p


def _batch(t, count):
    """count copies of target 0 with epochs +i s and masses +10 i kg (validate.py:73-92)."""
    out = {k: (np.repeat(v[:1], count) if hasattr(v, "shape") and v.ndim == 1 else v) for k, v in t.items()}
    out["state"] = np.repeat(t["state"][:, :1], count, axis=1)
    out["id"] = np.arange(1, count + 1)
    out["epoch"] = t["epoch"][0] + np.arange(count) / 86400.0
    out["mass"] = t["mass"][0] + 10.0 * np.arange(count)
    out["n"] = count
    return out


@pytest.mark.parametrize("case,duration,count", [("full_600s", 600.0, 1), ("full_65s_endpoint", 65.0, 1),
                                                 ("full_5s_endpoint", 5.0, 1), ("batch_4", 65.0, 4),
                                                 ("batch_33", 65.0, 33)])
def test_validate_cases(envisat_like, case, duration, count):
    d, cfg, t = envisat_like

    def mut(doc):
        doc["run"]["duration_seconds"] = duration
        if count > 1:
            doc["execution"]["maxThreads"] = 2
    c = _variant(d, cfg, case, mut)
    tt = _batch(t, count) if count > 1 else t
    tp = os.path.join(d, case + "_targets.json")
    _targets_json(tp, tt)
    out = os.path.join(d, case + "_out")
    _launch(c, tp, out)
    _compare_with_oracle(c, tt, out)
    text = open(os.path.join(out, "summary.json")).read()
    summary = json.loads(text)
    assert summary["max_end_epoch_error_seconds"] <= 1e-4
    assert summary["targets"] == count and summary["duration_seconds"] == duration
    # driver.cpp:166-175: nine keys, written by nlohmann::json (sorted, indent 2)
    keys = ["backend", "duration_seconds", "environment_seconds", "initial_step_seconds", "integrator",
            "max_end_epoch_error_seconds", "propagation_seconds", "targets", "validation_passed"]
    assert list(summary) == keys
    assert text.startswith('{\n  "backend": "device",\n') and text.endswith("}\n")
    assert '  "targets": %d,\n' % count in text and '  "validation_passed": true\n' in text


def test_central_invariants(tmp_path):
    """validate.py:95-107: energy and angular momentum within 1e-9 over 600 s."""
    from parahpop_amd import synth
    cfg, tg = synth.make_workload(str(tmp_path), 1, n=8, duration=600.0)
    out = str(tmp_path / "out")
    _launch(cfg, tg, out)
    t = synth.load_targets(tg)
    rows = _read(out)
    cols = ["x_km", "y_km", "z_km", "vx_km_s", "vy_km_s", "vz_km_s"]
    for i in range(t["n"]):
        s0 = t["state"][:, i]
        s1 = np.array([float(rows[int(t["id"][i])][c]) for c in cols])

        def inv(s):
            r, v = s[:3], s[3:]
            return v @ v / 2 - 398600.4415 / np.linalg.norm(r), np.cross(r, v)
        e0, h0 = inv(s0)
        e1, h1 = inv(s1)
        assert abs(e1 - e0) / abs(e0) <= 1e-9
        assert np.linalg.norm(h1 - h0) / np.linalg.norm(h0) <= 1e-9


def test_failure_cases(envisat_like):
    """validate.py:109-131: insufficient steps, existing outputs, removed options, bad threads."""
    d, cfg, t = envisat_like
    tp = os.path.join(d, "fail_targets.json")
    _targets_json(tp, t)
    bad = _variant(d, cfg, "insufficient", lambda doc: doc["execution"]["integration"].update(maxNumSteps=1))
    out = os.path.join(d, "insufficient_out")
    p = _launch(bad, tp, out, expect_success=False)
    assert p.stderr.startswith("ParaHPOP: ") and not os.path.exists(os.path.join(out, "summary.json"))
    good = os.path.join(d, "good_out")
    _launch(cfg, tp, good)
    before = open(os.path.join(good, "final_states.csv")).read()
    _launch(cfg, tp, good, expect_success=False)
    assert open(os.path.join(good, "final_states.csv")).read() == before
    for field in ("absTol", "maxStepSize"):
        c = _variant(d, cfg, "unsupported_" + field, lambda doc: doc["execution"]["integration"].update({field: 1.0}))
        _launch(c, tp, os.path.join(d, "unsupported_" + field), expect_success=False)
    for value in (0, -1):
        c = _variant(d, cfg, "threads_%d" % value, lambda doc: doc["execution"].update(maxThreads=value))
        _launch(c, tp, os.path.join(d, "threads_%d" % value), expect_success=False)


def test_process_with_reference_data_formats(tmp_path):
    """The process reads the reference's data formats end to end: a .brie
    Chebyshev ephemeris, a .brot2 IERS2000 Earth orientation, J2-only gravity and
    a piecewise exponential atmosphere with Earth flattening; 32 objects, 2 h,
    final_states.csv against the oracle at compare.py's tolerances."""
    import oracle
    from parahpop_amd import brie_io as brie_writer
    from parahpop_amd import synth
    d = str(tmp_path)
    cfg, tg = synth.make_alternate_workload(d, n=32, duration=7200.0)
    et0 = 4035.0 * 86400.0 - 43200.0
    brie_writer.write_brie(os.path.join(d, "eph.brie"),
                           brie_writer.ssb_units(oracle.body_position, et0 - 2 * 86400.0, et0 + 3 * 86400.0, bodies=(10, 301)))
    brie_writer.write_brot2(os.path.join(d, "itrf.brot2"), 4031.0, 4041.0,
                            brie_writer.iau2006_channels(xp_arcsec=0.2, yp_arcsec=0.3, wobble=True))
    doc = json.load(open(cfg))
    doc["model"]["environment"]["ephemeris"] = ["eph.brie"]
    doc["model"]["environment"]["orientations"] = ["itrf.brot2"]
    json.dump(doc, open(cfg, "w"))
    t = synth.load_targets(tg)
    out = os.path.join(d, "out")
    _launch(cfg, tg, out)
    _compare_with_oracle(cfg, t, out)
