"""J2-only gravity and (piecewise) exponential drag: the oracle pinned by an
independent field evaluation, the textbook ellipsoid formula and the band rules,
and the configuration rules of these blocks (CPU only)."""
import json
import os

import numpy as np
import pytest

GM = 398600.4415
EPOCH0_ET = 4035.0 * 86400.0 - 43200.0
FLAT = 0.0033528106647474805


def _write(out, model):
    from parahpop_amd import synth
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "constants.json"), "w") as f:
        json.dump(synth.CONSTANTS, f)
    p = os.path.join(out, "CONFIG.json")
    synth.write_config(p, model, 600.0, 60.0, 20)
    return p


def _earth_only(earth):
    return {"environment": {"bodies": {"Earth": earth}, "ephemeris": "analytic", "constants": "constants.json"}}


def _random_positions(rng, n, lo=6600.0, hi=9000.0):
    r = rng.normal(size=(n, 3))
    return r * (lo + (hi - lo) * rng.random((n, 1))) / np.linalg.norm(r, axis=1, keepdims=True)


def test_j2_equals_degree2_zonal_field(tmp_path):
    """J2-only gravity (J2.h:17-92) equals the spherical-harmonic evaluation of a
    degree-2 field whose only coefficient is Cbar20 = -J2/sqrt(5), at the same
    GM and radius: two independent code paths, agreement to rounding."""
    import oracle
    from parahpop_amd import synth
    j2 = synth.CONSTANTS[3]["J2"]
    out = str(tmp_path)
    with open(os.path.join(out, "g.tab"), "w") as f:
        f.write("%.10f, %.10f, 0.0, 2, 2, 1, 0.0, 0.0\n" % (synth.R_EARTH, GM))
        f.write("1, 0, 0.0, 0.0\n1, 1, 0.0, 0.0\n2, 0, %.17e, 0.0\n2, 1, 0.0, 0.0\n2, 2, 0.0, 0.0\n" % (-j2 / np.sqrt(5.0)))
    sh = oracle.Oracle(_write(os.path.join(out, "sh"), _earth_only(
        {"gravity": {"shape": {"body": "Earth", "file": "../g.tab", "degree": 2, "order": 2}}})))
    ob = oracle.Oracle(_write(os.path.join(out, "j2"), _earth_only({"gravity": {"oblateness": True}})))
    rng = np.random.default_rng(11)
    for r in _random_positions(rng, 32):
        et = EPOCH0_ET + 3600.0 * rng.random()
        a = ob.force(r, np.zeros(3), et, 1000.0, 1e-6, 1.0, 2.2, which=1)
        b = sh.force(r, np.zeros(3), et, 1000.0, 1e-6, 1.0, 2.2, which=1)
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), (a, b)


def test_j2_soi_and_inline_constant(tmp_path):
    """Zero beyond the sphere of influence; an inline body "j2" replaces the
    constants-file value and the acceleration is linear in it (exactly 2x)."""
    import oracle
    from parahpop_amd import synth
    j2 = synth.CONSTANTS[3]["J2"]
    a1 = oracle.Oracle(_write(str(tmp_path / "a"), _earth_only({"gravity": ["oblateness"]})))
    a2 = oracle.Oracle(_write(str(tmp_path / "b"), _earth_only({"gravity": {"oblateness": True}, "j2": 2.0 * j2})))
    r = np.array([5000.0, -3000.0, 4000.0])
    f1 = a1.force(r, np.zeros(3), EPOCH0_ET, 1000.0, 1e-6, 1.0, 2.2, which=1)
    f2 = a2.force(r, np.zeros(3), EPOCH0_ET, 1000.0, 1e-6, 1.0, 2.2, which=1)
    assert np.linalg.norm(f1) > 0
    np.testing.assert_array_equal(f2, 2.0 * f1)
    far = np.array([0.0, 0.0, 924700.0])
    np.testing.assert_array_equal(a1.force(far, np.zeros(3), EPOCH0_ET, 1000.0, 1e-6, 1.0, 2.2, which=1), np.zeros(3))


def test_geodetic_altitude_inverts_the_ellipsoid_formula():
    """Geodetic altitude (Geodetic.h:40-72) recovers h from the textbook forward
    map x = (N + h) cos(lat) (cos lon, sin lon), z = (N (1 - e^2) + h) sin(lat),
    N = a / sqrt(1 - e^2 sin^2 lat), in any frame (pole rotated), to 1e-8 km;
    with zero flattening it is ||r|| - a exactly."""
    import oracle
    a = 6378.137
    e2 = (2.0 - FLAT) * FLAT
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    pole = q @ np.array([0.0, 0.0, 1.0])
    for _ in range(400):
        lat = np.radians(rng.uniform(-89.99, 89.99))
        lon = np.radians(rng.uniform(-180, 180))
        h = rng.uniform(-20.0, 40000.0)
        n = a / np.sqrt(1.0 - e2 * np.sin(lat) ** 2)
        x = np.array([(n + h) * np.cos(lat) * np.cos(lon), (n + h) * np.cos(lat) * np.sin(lon),
                      (n * (1.0 - e2) + h) * np.sin(lat)])
        got = oracle.geodetic_altitude(q @ x, pole, a, FLAT)
        assert abs(got - h) <= 1e-8, (np.degrees(lat), h, got)
        assert oracle.geodetic_altitude(x, [0, 0, 1.0], a, 0.0) == np.sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) - a


def test_exponential_bands_follow_the_band_rules(tmp_path):
    """Bands are sorted by base altitude whatever the config order; the band of h
    is the last one with h0 <= h (band 0 below them all); density is rho0 at a
    band's base, continues to its cutoff inclusive, and is zero in a gap and
    above the last cutoff (PiecewiseExponentialAtmosphere.h:100-170)."""
    import oracle
    blocks = [{"rho0": 5.0, "h0": 300.0, "scaleHeight": 50.0, "hCutoff": 400.0},
              {"rho0": 100.0, "h0": 100.0, "scaleHeight": 10.0, "hCutoff": 200.0},
              {"rho0": 20.0, "h0": 200.0, "scaleHeight": 30.0, "hCutoff": 250.0}]
    o = oracle.Oracle(_write(str(tmp_path), _earth_only({"gravity": True, "atmosphere": {"exponential": blocks}})))
    assert o.exp_density(100.0) == 100.0
    assert o.exp_density(200.0) == 20.0             # the upper band owns its base
    assert o.exp_density(300.0) == 5.0
    assert o.exp_density(250.0) == pytest.approx(20.0 * np.exp(-50.0 / 30.0), rel=1e-15)   # cutoff inclusive
    assert o.exp_density(np.nextafter(250.0, 1e9)) == 0.0  # gap 250..300 reads as vacuum
    assert o.exp_density(299.0) == 0.0
    assert o.exp_density(400.0) > 0.0 and o.exp_density(400.5) == 0.0
    assert o.exp_density(50.0) == pytest.approx(100.0 * np.exp(5.0), rel=1e-15)   # below all bands: band 0
    assert o.exp_density(150.0) == pytest.approx(100.0 * np.exp(-5.0), rel=1e-15)
    single = oracle.Oracle(_write(str(tmp_path / "one"), _earth_only(
        {"gravity": True, "atmosphere": {"exponential": blocks[1]}})))
    assert single.exp_density(150.0) == o.exp_density(150.0)
    assert single.exp_density(250.0) == 0.0


def test_exponential_drag_composition(tmp_path):
    """Drag = -1/2 rho (cd A/m) |v_rel| v_rel with v_rel = v - omega x r and rho
    at the geodetic altitude of the Earth-fixed position (pole = ITRF z axis),
    linear in cd, zero above the cutoff (PrivateAccelerations.h:167-202)."""
    import oracle
    from parahpop_amd import synth
    o = oracle.Oracle(_write(str(tmp_path), _earth_only(
        {"gravity": True, "flattening": FLAT, "atmosphere": {"exponential": synth.exponential_blocks()}})))
    rng = np.random.default_rng(8)
    for r in _random_positions(rng, 16, 6500.0, 7800.0):
        v = np.cross([0, 0, 1.0], r)
        v *= 7.5 / np.linalg.norm(v)
        R, w = oracle.earth_orientation(EPOCH0_ET)
        h = oracle.geodetic_altitude(R @ r, [0, 0, 1.0], synth.R_EARTH, FLAT)
        rho = o.exp_density(h)
        vrel = v - np.cross(w, r)
        expect = -0.5 * rho * (2.2 * 1e-6 / 100.0) * np.linalg.norm(vrel) * vrel
        got = o.force(r, v, EPOCH0_ET, 100.0, 1e-6, 1.0, 2.2, which=4)
        np.testing.assert_allclose(got, expect, rtol=1e-9, atol=0)
        got2 = o.force(r, v, EPOCH0_ET, 100.0, 1e-6, 1.0, 4.4, which=4)
        np.testing.assert_array_equal(got2, 2.0 * got)
    high = np.array([0.0, 6378.137 + 1600.0, 0.0])
    np.testing.assert_array_equal(o.force(high, [7.0, 0, 0], EPOCH0_ET, 100.0, 1e-6, 1.0, 2.2, which=4), np.zeros(3))


def test_alternate_workload_propagates_on_the_oracle(tmp_path):
    """The alternate force model runs end to end on the oracle (16 objects, 10 min)
    and differs from a drag-free run only for the low objects."""
    import oracle
    from parahpop_amd import synth
    cfg, tg = synth.make_alternate_workload(str(tmp_path), n=16, duration=600.0)
    t = synth.load_targets(tg)
    e, s, st = oracle.Oracle(cfg).propagate(t, 600.0, threads=2)
    assert st["object_steps"] == 160 and np.all(np.isfinite(s))
    assert np.all(e == e[0])
