"""Orientation units of bodies other than the Earth (.brot2 types 1-3) and the
fields that act in their axes, on the CPU oracle.

Pins, each against something the restatement cannot share a mistake with:
* IAU body model: the third row of ICRF -> body is the pole (cos d cos a,
  cos d sin a, sin d) and the first row is the prime meridian, W along the body
  equator from the node z x pole (the IAU definition), with RA, Dec and W
  evaluated here from the published Moon model; the rotation rate is the
  13.176 deg/day of the model's W.
* Binary frames: the passive 3-1-3 rotation of the fitted Euler angles, built
  here from scipy's Rotation, within the Chebyshev fit error; an ECLIPJ2000-based
  frame is that rotation times R1(eps), eps = 84381.448 arcsec.
* Frame resolution: an explicit name, the unique loaded binary frame, the IAU
  model, and the reference's rejections.
* J2 of the Moon: the textbook J2 acceleration in the body axes from scipy's
  rotation; nothing beyond the sphere of influence.
"""
import json
import os

import numpy as np
import pytest

from parahpop_amd import brie_io, synth

DAY = 86400.0
ET0 = synth.EPOCH0 * DAY - 43200.0


def _moon_angles(et, terms=4):
    """RA, Dec, W (degrees) of the IAU Moon model truncated as moon_iau_units."""
    T = et / 3155760000.0
    d = et / DAY
    E = np.radians([125.045 - 1935.5364525 * T, 250.089 - 3871.072905 * T, 260.008 + 475263.3328725 * T,
                    176.625 + 487269.629985 * T][:terms])
    ra = 269.9949 + 0.0031 * T - 3.8787 * np.sin(E[0]) - 0.1204 * np.sin(E[1]) + 0.0700 * np.sin(E[2]) \
        - 0.0172 * np.sin(E[3])
    dec = 66.5392 + 0.0130 * T + 1.5419 * np.cos(E[0]) + 0.0239 * np.cos(E[1]) - 0.0278 * np.cos(E[2]) \
        + 0.0068 * np.cos(E[3])
    W = 38.3213 + 13.17635815 * d - 1.4e-12 * d * d + 3.5610 * np.sin(E[0]) + 0.1208 * np.sin(E[1]) \
        - 0.0642 * np.sin(E[2]) + 0.0158 * np.sin(E[3])
    return ra, dec, W


def _passive313(angles):
    """ICRF -> frame matrix of the passive z-x-z rotation by (phi, theta, psi):
    the inverse of scipy's intrinsic 'ZXZ' rotation."""
    from scipy.spatial.transform import Rotation
    return Rotation.from_euler("ZXZ", angles).as_matrix().T


def _oracle(cfg):
    import oracle
    return oracle.Oracle(cfg)


def test_iau_body_model_pole_and_prime_meridian(tmp_path):
    o = _oracle(synth.make_lunar_workload(str(tmp_path), field="j2", frame="iau"))
    for et in ET0 + np.array([0.0, 3.7e3, 2.5 * DAY, -40 * DAY, 3650 * DAY]):
        R, target = o.body_axes(301, et)
        assert target == 301
        ra, dec, W = np.radians(_moon_angles(et))
        pole = np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)])
        node = np.cross([0.0, 0.0, 1.0], pole)
        node /= np.linalg.norm(node)
        xb = np.cos(W) * node + np.sin(W) * np.cross(pole, node)
        np.testing.assert_allclose(R[2], pole, atol=2e-15)
        np.testing.assert_allclose(R[0], xb, atol=2e-14)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-15)
        assert abs(np.linalg.det(R) - 1.0) < 1e-14
    # rotation rate: the body x axis turns at the model's 13.176 deg/day
    R0, _ = o.body_axes(301, ET0)
    R1, _ = o.body_axes(301, ET0 + 3600.0)
    ang = np.degrees(np.arccos(np.clip(R0[0] @ R1[0], -1, 1)))
    assert abs(ang - 13.17635815 / 24.0) < 2e-3 * 13.17635815 / 24.0


@pytest.mark.parametrize("frame", ["moon_pa", "moon_pa_ecl"])
def test_binary_frame_is_the_passive_313_rotation(tmp_path, frame):
    o = _oracle(synth.make_lunar_workload(str(tmp_path), field="j2", frame=frame))
    euler = synth.moon_pa_euler if frame == "moon_pa" else synth.moon_ecliptic_euler
    eps = np.radians(84381.448 / 3600.0)
    E = np.array([[1, 0, 0], [0, np.cos(eps), np.sin(eps)], [0, -np.sin(eps), np.cos(eps)]])
    for et in ET0 + np.array([0.0, 1234.5, 0.37 * DAY, 1.5 * DAY, 2.9 * DAY]):
        R, target = o.body_axes(301, et)
        assert target == 31008
        want = _passive313(euler(et))
        if frame == "moon_pa_ecl":
            want = want @ E
        np.testing.assert_allclose(R, want, atol=1e-12)


def test_frame_resolution_and_rejections(tmp_path):
    import oracle
    cfg = synth.make_lunar_workload(str(tmp_path / "pa"), field="sh", frame="moon_pa")
    # the unique loaded binary frame wins over the IAU model; a named frame wins over both
    assert _oracle(cfg).body_axes(301, ET0)[1] == 31008
    doc = json.load(open(cfg))
    moon = doc["model"]["environment"]["bodies"]["Moon"]
    for name, want in [("IAU_MOON", 301), ("moon_body_fixed", 301), ("MOON_PA", 31008)]:
        moon["orientation"] = name
        json.dump(doc, open(cfg, "w"))
        assert _oracle(cfg).body_axes(301, ET0)[1] == want
    for name, msg in [("ITRF93", "describes body 399"), ("ICRF", "inertial frame"), ("FOO_FRAME", "not a known"),
                      ("IAU_EARTH", "describes body 399")]:
        moon["orientation"] = name
        json.dump(doc, open(cfg, "w"))
        with pytest.raises(RuntimeError, match=msg):
            oracle.Oracle(cfg)
    del moon["orientation"]
    # rotation data missing: no file, or a file without the Moon
    doc["model"]["environment"]["orientations"] = "analytic"
    json.dump(doc, open(cfg, "w"))
    with pytest.raises(RuntimeError, match="needs its rotation data"):
        oracle.Oracle(cfg)
    brie_io.write_brot2_units(str(tmp_path / "pa" / "mars.brot2"),
                              [brie_io.iau_body_unit(499, [317.68, -0.1061], [52.89, -0.0609], [176.63, 350.89198226])])
    doc["model"]["environment"]["orientations"] = ["mars.brot2"]
    json.dump(doc, open(cfg, "w"))
    with pytest.raises(RuntimeError, match="needs its rotation data loaded"):
        oracle.Oracle(cfg)


def test_moon_j2_in_body_axes_and_sphere_of_influence(tmp_path):
    o = _oracle(synth.make_lunar_workload(str(tmp_path), field="j2", frame="iau"))
    consts = {c["NAME"]: c for c in synth.CONSTANTS if "NAME" in c}
    gm, j2, rm, soi = (consts["Moon"][k] for k in ("GM", "J2", "R", "SOIOld"))
    et = ET0 + 1800.0
    moon = o.body_position(301, et)
    sun = o.body_position(10, et)
    R, _ = o.body_axes(301, et)
    gm_sun = consts["Sun"]["GM"]

    def third(r, b, g):
        d = r - b
        return -g * d / np.linalg.norm(d) ** 3 - g * b / np.linalg.norm(b) ** 3

    for rel in ([1900.0, 300.0, -250.0], [-200.0, 1800.0, 900.0], [60000.0, 0.0, 20000.0]):
        r = moon + np.array(rel)
        got = o.force(r, np.zeros(3), et, 100.0, 1e-6, 0.0, 2.2, which=2)
        point = third(r, moon, gm) + third(r, sun, gm_sun)
        f = R @ np.array(rel)
        rr = np.linalg.norm(f)
        # textbook J2 acceleration (Wakker eq. 20.6) in the body axes
        k = -1.5 * gm * j2 * rm ** 2 / rr ** 5
        a_body = k * np.array([f[0] * (1 - 5 * f[2] ** 2 / rr ** 2), f[1] * (1 - 5 * f[2] ** 2 / rr ** 2),
                               f[2] * (3 - 5 * f[2] ** 2 / rr ** 2)])
        want = point + (R.T @ a_body if np.linalg.norm(rel) <= soi else 0.0)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-22)
    # beyond the sphere of influence (a LEO object): point masses only
    r = np.array([7000.0, 0.0, 0.0])
    got = o.force(r, np.zeros(3), et, 100.0, 1e-6, 0.0, 2.2, which=2)
    # (the Moon's J2 there would be 3e-10 of the tidal term)
    np.testing.assert_allclose(got, third(r, moon, gm) + third(r, sun, gm_sun), rtol=1e-12)
