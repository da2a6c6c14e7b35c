"""IERS2000 Earth orientation from the reference's .brot2 format: the file
built from the analytic model's channels reproduces that model; the nutation
and ERP Lagrange stencils against numpy interpolation over the same nodes;
the envelope and missing-unit errors (CPU: oracle)."""
import json
import os

import numpy as np
import pytest

from parahpop_amd import brie_io as brie_writer

DAY = 86400.0
EPOCH0_ET = 4035.0 * DAY - 43200.0
TT0 = 4035.0 - 3.0   # TT MJD2000 days of the first node


def _config(out, orient, files=()):
    from parahpop_amd import synth
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "constants.json"), "w") as f:
        json.dump(synth.CONSTANTS, f)
    model = {"environment": {"bodies": {"Earth": {"gravity": True}}, "ephemeris": "analytic", "orientations": orient,
                             "constants": "constants.json"}}
    p = os.path.join(out, "CONFIG.json")
    synth.write_config(p, model, 3600.0, 60.0, 100)
    return p


def test_brot2_reproduces_the_analytic_orientation(tmp_path):
    """A file of the analytic model's channels (cubic X, Y, s; constant UT1 - TT;
    no polar motion) gives the analytic rotation and Earth rate: the IERS2000
    composition R = PM ER PN with PN = R3(-e-s) R2(d) R3(e) against the
    independently written analytic model (two TDB-TT series: a few 1e-9 rad;
    the rates differ by the precession rate)."""
    import oracle
    brie_writer.write_brot2(str(tmp_path / "o.brot2"), TT0, TT0 + 12.0, brie_writer.iau2006_channels())
    b = oracle.Oracle(_config(str(tmp_path / "b"), ["../o.brot2"]))
    a = oracle.Oracle(_config(str(tmp_path / "a"), "analytic"))
    rng = np.random.default_rng(4)
    for et in EPOCH0_ET + rng.uniform(-1.5, 7.5, 30) * DAY:
        Rb, wb = b.earth_orientation(et)
        Ra, wa = a.earth_orientation(et)
        np.testing.assert_allclose(Rb, Ra, atol=6e-9, rtol=0)
        np.testing.assert_allclose(Rb @ Rb.T, np.eye(3), atol=1e-14)
        # the IERS2000 rate includes the precession rate (2004" per century,
        # 3.1e-12 rad/s), which the analytic model's rigid rate leaves out
        np.testing.assert_allclose(wb, wa, atol=5e-12, rtol=0)


def _tt(et):
    g = np.radians(357.53 + 0.9856003 * (et / DAY))
    return et / DAY + 0.5 - (0.001658 * np.sin(g) + 0.000014 * np.sin(2 * g)) / DAY


def test_lagrange_stencils_match_numpy(tmp_path):
    """Channel X = 1e-3 sin(tt/2.3) alone (Y = s = 0): the pole row of R is the
    interpolated CIP (X, 0, sqrt(1 - X^2)), X from the six nodes centred on the
    bracketing 0.5-day interval. UT1 - TT = 1e-4 sin(tt/1.7) alone on an uneven
    key grid: R = R3(theta), theta from the four keys centred on the bracketing
    interval. numpy's degree-5 / degree-3 fits through those nodes are exact."""
    import oracle
    keys = TT0 + np.cumsum(np.r_[0.0, np.where(np.arange(15) % 4 == 3, 1.0000115741, 1.0)])
    brie_writer.write_brot2(str(tmp_path / "x.brot2"), TT0, TT0 + 14.0,
                            lambda t: (1e-3 * np.sin(t / 2.3), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    brie_writer.write_brot2(str(tmp_path / "u.brot2"), TT0, TT0 + 14.0,
                            lambda t: (0.0, 0.0, 0.0, 0.0, 0.0, 1e-4 * np.sin(t / 1.7), 0.0, 0.0), erp_keys=keys)
    ox = oracle.Oracle(_config(str(tmp_path / "x"), ["../x.brot2"]))
    ou = oracle.Oracle(_config(str(tmp_path / "u"), ["../u.brot2"]))
    nodes = TT0 + 0.5 * np.arange(29)
    rng = np.random.default_rng(9)
    for et in EPOCH0_ET + rng.uniform(-1.0, 8.0, 25) * DAY:
        tt = _tt(et)
        k = int(np.floor((tt - TT0) / 0.5))
        lo = min(max(k - 2, 0), len(nodes) - 6)
        xs = nodes[lo:lo + 6]
        X = np.polyval(np.polyfit(xs - tt, 1e-3 * np.sin(xs / 2.3), 5), 0.0)
        R, _ = ox.earth_orientation(et)
        assert abs(R[2, 0] - X) < 1e-15 and abs(R[2, 1]) < 1e-18
        last = np.searchsorted(keys, tt, side="right") - 1
        st = min(max(last - 1, 0), len(keys) - 4)
        ks = keys[st:st + 4]
        dut1 = np.polyval(np.polyfit(ks - tt, 1e-4 * np.sin(ks / 1.7), 3), 0.0)
        frac = (np.fmod(et, DAY) / DAY + 0.5) % 1.0 + (tt - (et / DAY + 0.5))   # TT day fraction
        theta = 2 * np.pi * (np.fmod(frac + dut1 - 0.5, 1.0) + 0.7790572732640 + 0.00273781191135448 * (tt + dut1 - 0.5))
        R, w = ou.earth_orientation(et)
        got = np.arctan2(R[0, 1], R[0, 0]) - (-6.24e-15 * (tt - 0.5))   # less the TIO locator s' in PM
        assert abs(np.angle(np.exp(1j * (got - theta)))) < 2e-12, (et, got, theta)


def test_orientation_envelope_and_unit_errors(tmp_path):
    import oracle
    brie_writer.write_brot2(str(tmp_path / "o.brot2"), TT0, TT0 + 6.0, brie_writer.iau2006_channels())
    o = oracle.Oracle(_config(str(tmp_path / "c"), ["../o.brot2"]))
    t = {"id": np.array([1]), "epoch": np.array([4035.0]), "state": np.array([[7000.0], [0], [0], [0], [7.5], [0]]),
         "mass": np.array([100.0]), "drag_area": np.array([1.0]), "srp_area": np.array([1.0]), "cd": np.array([2.2]),
         "cr": np.array([1.3]), "n": 1}
    o.propagate(t, 3600.0, threads=1)
    with pytest.raises(RuntimeError, match="orientation range"):
        o.propagate(t, 5 * DAY, threads=1)
    with open(tmp_path / "empty.brot2", "wb") as f:
        f.write(brie_writer.cbor({"layout": 2, "core": {"metadata": {"nBodyUnits": 1, "intMetadata": [5] + [0] * 11,
                                                                     "doubleMetadata": [0.0, 1.0]}, "data": [0.0]}}))
    with pytest.raises(RuntimeError, match="no IERS2000 or body orientation unit"):
        oracle.Oracle(_config(str(tmp_path / "e"), ["../empty.brot2"]))
    # a body unit whose sizes exceed its data
    with open(tmp_path / "short.brot2", "wb") as f:
        f.write(brie_writer.cbor({"layout": 2, "core": {"metadata": {"nBodyUnits": 1, "intMetadata": [1] + [0] * 11,
                                                                     "doubleMetadata": [0.0, 1.0]}, "data": [0.0]}}))
    with pytest.raises(RuntimeError, match="inconsistent sizes"):
        oracle.Oracle(_config(str(tmp_path / "s"), ["../short.brot2"]))
