"""Pins of the oracle's environment functions (the EOP time model, the
NRLMSISE-00 input preparation, solid and ocean tides, the third-body term, the
dual-cone shadow and the polar limit of the SH sum) against what fixes them
independently of the restatement: tables whose values change every day and
3-hour slot (so a wrong day, slot or lag fails), the published leap seconds,
calendar arithmetic of the Python standard library, the IERS 2010 Earth
rotation angle and GMST, the published NRLMSISE-00 input definitions, IERS 2010
eq. 6.6 with scipy's Legendre functions, the textbook Doodson variables, the
disk-overlap geometry by brute force, and the SH sum's own limit at the pole.

The weather and EOP checks read the synthetic files with the parsers written
here, so they also pin the table loaders (model.cpp) that the oracle and the
library share."""
import datetime
import json
import math
import os

import numpy as np
import pytest
from scipy import special

DAY = 86400.0
EPOCH0_ET = 4035.0 * DAY - 43200.0


@pytest.fixture(scope="module")
def varying(tmp_path_factory):
    """Full-force configuration with varying space weather and EOP (one day)."""
    import oracle
    from parahpop_amd import synth
    d = str(tmp_path_factory.mktemp("varying"))
    cfg, _ = synth.make_workload(d, 2, n=4, duration=DAY, varying=True)
    return d, cfg, oracle.Oracle(cfg)


def _eop_rows(path):
    rows = np.loadtxt(path)
    return {"mjd": rows[:, 3], "xp": rows[:, 4], "yp": rows[:, 5], "dut": rows[:, 6], "dat": rows[:, 12]}


def _tdb_minus_tt(et):
    """Explanatory Supplement (1992) approximation of TDB - TT (s), good to ~30 us."""
    g = math.radians(357.53 + 0.98560028 * (et / DAY))
    return 0.001657 * math.sin(g) + 0.000014 * math.sin(2 * g)


def test_eop_time_model_interpolates_and_steps(varying):
    """EarthCorrections.h:67-105: xp, yp and UT1-UTC are linear in the UTC day
    fraction (the table is linear, so the result is the linear function itself,
    whatever the day index), TT - UTC = 32.184 s + TAI-UTC of the UTC day (one
    1 s step in the table), and the ET -> UTC inversion lands on that UTC."""
    d, _, orc = varying
    e = _eop_rows(os.path.join(d, "EOP-All.txt"))
    first = e["mjd"][0]
    step_mjd = e["mjd"][np.argmax(e["dat"] > e["dat"][0])]
    step_et = (step_mjd - 51544.5) * DAY
    ets = [EPOCH0_ET + x for x in (0.0, 1234.5, 21600.0, 43199.0)]
    # around the UTC midnight of the step (not inside the leap second itself,
    # whose UTC is ambiguous)
    ets += [step_et + 69.184 + x for x in (-2.0, -0.6, 1.6, 2.0, 3600.0)]
    seen = set()
    for et in ets:
        t = orc.time_at(et)
        assert t["valid"] == 1.0
        u = t["mjd_utc"] - first
        assert abs(t["xp"] - (0.05 + 0.0013 * u)) < 1e-12
        assert abs(t["yp"] - (0.35 - 0.0009 * u)) < 1e-12
        assert abs(t["ut1_utc"] - (-0.2 + 0.0011 * u)) < 1e-12
        dat = e["dat"][np.where(e["mjd"] == math.floor(t["mjd_utc"]))[0][0]]
        assert t["tt_utc"] == 32.184 + dat
        seen.add(dat)
        # ET - UTC = (TT - UTC) + (TDB - TT)
        utc_et = (t["mjd_utc"] - 51544.5) * DAY
        assert abs((et - utc_et) - (t["tt_utc"] + _tdb_minus_tt(et))) < 1e-4
    assert seen == {37.0, 38.0}
    assert orc.time_at((first - 51544.5 - 5.0) * DAY)["valid"] == 0.0   # before the table


# IERS Bulletin C: TAI - UTC from these UTC dates (MJD), 1999-2017
LEAP_SECONDS = [(51179, 32.0), (53736, 33.0), (54832, 34.0), (56109, 35.0), (57204, 36.0), (57754, 37.0)]


def _tai_utc(mjd):
    return [v for m, v in LEAP_SECONDS if mjd >= m][-1]


def _weather(path):
    rows = np.array([[float(x) for x in ln.split()] for ln in open(path)])
    dates = [datetime.date(int(r[0]), int(r[1]), int(r[2])) for r in rows]
    mjd = np.array([(dd - datetime.date(1858, 11, 17)).days for dd in dates], dtype=float)
    # columns: y m d bsrn nd kp*8 sum ap*8 Ap cp c9 isn F10.7adj q F10.7A-adj-c81 ...
    return {"mjd": mjd, "ap": rows[:, 14:22], "Ap": rows[:, 22], "f107": rows[:, 26], "f107a": rows[:, 28]}


def _gast_hours_iau2006(mjd_utc):
    """GAST from the IERS 2010 Earth rotation angle (eq. 5.15) and GMST (eq.
    5.32) with UT1 = UTC, plus the four leading nutation terms of the equation
    of the equinoxes."""
    du = mjd_utc - 51544.5
    era = 2 * math.pi * ((0.7790572732640 + 0.00273781191135448 * du) % 1.0 + du % 1.0)
    t = (du + (32.184 + _tai_utc(mjd_utc)) / DAY) / 36525.0
    gmst = era + math.radians((0.014506 + 4612.156534 * t + 1.3915817 * t * t) / 3600.0)
    om = math.radians(125.04452 - 1934.136261 * t)
    ls = math.radians(280.4665 + 36000.7698 * t)
    lm = math.radians(218.3165 + 481267.8813 * t)
    dpsi = -17.20 * math.sin(om) - 1.32 * math.sin(2 * ls) - 0.23 * math.sin(2 * lm) + 0.21 * math.sin(2 * om)
    eps = math.radians(23.43929111 - 46.8150 * t / 3600.0)
    return (math.degrees(gmst + math.radians(dpsi / 3600.0) * math.cos(eps)) / 15.0) % 24.0


def test_nrlmsise_inputs_follow_the_published_definitions(varying):
    """Nrlmsise00.cu:91-241: UTC from ET with the leap seconds of Bulletin C;
    day of year and seconds of day (calendar arithmetic); GAST hours (IERS 2010
    ERA + GMST); and the weather record and slot: F10.7 of the previous day,
    the 81-day mean F10.7A and the daily Ap of the current day, the 3-hour ap of
    the current slot and 3, 6 and 9 hours before, and the means of the eight
    3-hour values 12-33 h and 36-57 h before (NRLMSISE-00 input definitions)."""
    d, cfg, orc = varying
    w = _weather(os.path.join(d, "SW-All.txt"))
    series = w["ap"].ravel()   # 3-hour ap, day-major
    fixed = np.array([6778.137, 0.0, 0.0])
    for dt in (0.0, 3.0, 10799.0, 10801.0, 21601.0, 43199.5, 43201.0, 50000.0, 86399.0, 2 * DAY + 7777.0):
        et = EPOCH0_ET + dt
        x = orc.nrl_inputs(et, fixed)
        assert x["valid"]
        mjd = 51544.5 + (et - _tdb_minus_tt(et) - 32.184 - _tai_utc(51544.5 + et / DAY)) / DAY
        assert abs(x["mjd_utc"] - mjd) < 1e-9   # 86 us
        when = datetime.datetime(1858, 11, 17) + datetime.timedelta(days=x["mjd_utc"])
        assert x["doy"] == when.timetuple().tm_yday
        sod = (x["mjd_utc"] - math.floor(x["mjd_utc"])) * DAY
        assert abs(x["sec"] - sod) < 1e-6
        assert abs(((x["gast_h"] - _gast_hours_iau2006(x["mjd_utc"]) + 12) % 24) - 12) < 2e-5   # 0.07 s
        assert abs(((x["lst_h"] - x["gast_h"] - x["lon"] / 15.0 + 12) % 24) - 12) < 1e-12
        day = int(np.where(w["mjd"] == math.floor(x["mjd_utc"]))[0][0])
        slot = int((x["mjd_utc"] % 1.0) * 8)
        i = day * 8 + slot
        assert x["f107"] == w["f107"][day - 1]
        assert x["f107a"] == w["f107a"][day]
        assert x["ap_daily"] == w["Ap"][day] == x["ap"][0]
        np.testing.assert_array_equal(x["ap"][1:5], series[[i, i - 1, i - 2, i - 3]])
        assert x["ap"][5] == pytest.approx(series[i - 11:i - 3].mean(), abs=1e-12)
        assert x["ap"][6] == pytest.approx(series[i - 19:i - 11].mean(), abs=1e-12)
    # the table keeps the days from floor(coverage start) - 4, and a record
    # needs three days of history (Nrlmsise00.cu:220; loader model.cpp)
    start = math.floor(_find_key(json.load(open(cfg)), "coverageStartMjdUtc")) - 4
    for days, ok in ((2.5, False), (2.99, False), (3.01, True), (3.5, True)):
        et = (start + days - 51544.5) * DAY + 32.184 + _tai_utc(start)
        assert orc.nrl_inputs(et, fixed)["valid"] == ok, days


def _find_key(tree, key):
    if isinstance(tree, dict):
        if key in tree:
            return tree[key]
        tree = list(tree.values())
    if isinstance(tree, list):
        for v in tree:
            found = _find_key(v, key)
            if found is not None:
                return found
    return None


def test_nrlmsise_calendar_and_sidereal_time_over_three_decades(varying):
    """Nrlmsise00.cu:91-187 over 2000-2030 (outside the weather table, so only
    the time part): UTC with the published leap seconds, day of year and
    seconds of day from the standard library's calendar, GAST from the IERS
    2010 Earth rotation angle and GMST; the dates avoid the leap seconds."""
    _, _, orc = varying
    rng = np.random.default_rng(11)
    fixed = np.array([6778.137, 0.0, 0.0])
    for day in rng.uniform(0.6, 30 * 365.25, 400):
        et = day * DAY
        x = orc.nrl_inputs(et, fixed)
        mjd_tt = 51544.5 + (et - _tdb_minus_tt(et)) / DAY
        if any(abs(mjd_tt - m) < 2e-3 for m, _ in LEAP_SECONDS):
            continue
        mjd = mjd_tt - (32.184 + _tai_utc(mjd_tt - 40.0 / DAY)) / DAY
        assert abs(x["mjd_utc"] - mjd) < 1e-9
        when = datetime.datetime(1858, 11, 17) + datetime.timedelta(days=x["mjd_utc"])
        assert x["doy"] == when.timetuple().tm_yday, (et, when)
        assert abs(x["sec"] - (x["mjd_utc"] % 1.0) * DAY) < 1e-5
        assert abs(((x["gast_h"] - _gast_hours_iau2006(x["mjd_utc"]) + 12) % 24) - 12) < 2e-5


def _pbar(n, m, x):
    """Fully normalized associated Legendre function without the Condon-Shortley phase."""
    norm = math.sqrt((2 - (m == 0)) * (2 * n + 1) * math.factorial(n - m) / math.factorial(n + m))
    return norm * (-1) ** m * special.lpmv(m, n, x)


def test_solid_tide_step1_is_iers_eq_6_6(varying):
    """HpopCorrections.h:201-255: the first step of the solid tide for one
    point Moon is IERS 2010 eq. 6.6 (degrees 2 and 3) and eq. 6.7 (degree 4)
    with the Love numbers of Table 6.3 and normalized Legendre functions
    (scipy), for Moon directions off every axis; the Sun, the pole tide and the
    frequency-dependent terms cancel in the difference with a distant Moon."""
    _, _, orc = varying
    R, mu_ratio = 6378.1363e3, 1.0 / 81.3005682214972154
    E = np.eye(3)
    time5 = [55580.2, -0.19, 69.184, 0.05, 0.35]
    sun = [1.4e11, 3.0e10, 1.0e10]
    k2 = {0: 0.30190 + 0j, 1: 0.29830 - 0.00144j, 2: 0.30102 - 0.00130j}
    k3 = {0: 0.093, 1: 0.093, 2: 0.093, 3: 0.094}
    k4 = {0: -0.00089, 1: -0.00080, 2: -0.00057}
    far = [1e15, 0.0, 0.0]
    C0, S0 = orc.solid_tide(far, sun, E, time5, frequency_terms=False, zero_tide=False)
    for lat, lon, r in ((20.0, 35.0, 3.844e8), (-27.0, 200.0, 3.7e8), (5.0, -80.0, 4.05e8)):
        phi, lam = math.radians(lat), math.radians(lon)
        moon = r * np.array([math.cos(phi) * math.cos(lam), math.cos(phi) * math.sin(lam), math.sin(phi)])
        C, S = orc.solid_tide(moon, sun, E, time5, frequency_terms=False, zero_tide=False)
        dC, dS = C - C0, S - S0
        x = math.sin(phi)
        for n, ks in ((2, k2), (3, k3)):
            for m, k in ks.items():
                z = k / (2 * n + 1) * mu_ratio * (R / r) ** (n + 1) * _pbar(n, m, x) * np.exp(-1j * m * lam)
                assert dC[n, m] == pytest.approx(z.real, rel=1e-9, abs=1e-22)
                assert dS[n, m] == pytest.approx(-z.imag, rel=1e-9, abs=1e-22)
        for m, k in k4.items():
            z = k / 5 * mu_ratio * (R / r) ** 3 * _pbar(2, m, x) * np.exp(-1j * m * lam)
            assert dC[4, m] == pytest.approx(z.real, rel=1e-9, abs=1e-22)
            assert dS[4, m] == pytest.approx(-z.imag, rel=1e-9, abs=1e-22)


def test_solid_tide_zero_tide_term(varying):
    """HpopCorrections.h:256-259: the zero-tide convention removes the permanent
    tide A0 H0 k20 = 4.4228e-8 x (-0.31460) x 0.30190 from C20 (IERS 2010
    eq. 6.13), so C20 rises by 4.2007e-9 and nothing else changes."""
    _, _, orc = varying
    args = ([3.8e8, 1e7, 2e7], [1.4e11, 3.0e10, 1.0e10], np.eye(3), [55580.2, -0.19, 69.184, 0.05, 0.35])
    Cz, Sz = orc.solid_tide(*args, zero_tide=True)
    Cf, Sf = orc.solid_tide(*args, zero_tide=False)
    diff = Cz - Cf
    assert diff[2, 0] == pytest.approx(4.4228e-8 * 0.31460 * 0.30190, rel=1e-12)
    assert diff[2, 0] == pytest.approx(4.2007e-9, rel=1e-4)
    diff[2, 0] = 0.0
    assert not diff.any() and np.array_equal(Sz, Sf)


def test_pole_tides_follow_the_iers_mean_pole(varying):
    """HpopCorrections.h:187-199,260-263,318-321: the solid pole tide (IERS 2010
    eq. 6.22) and the ocean pole tide (eq. 6.24) are driven by m1 = xp - x_mean,
    m2 = -(yp - y_mean) with the IERS 2010 mean pole (Table 7.7): the mean pole
    read back from the coefficients is 55.974 / 346.346 mas at J2000, continuous
    at 2010.0 where the cubic hands over to the linear model (a slip in any
    coefficient breaks it), linear after it with the published rates, and the
    two pole tides agree on (m1, m2)."""
    _, _, orc = varying
    import oracle
    far = [1e18, 0.0, 0.0]

    def mean_pole(years, xp=0.12, yp=0.31):
        mjd = 51544.5 + 365.25 * years - 69.184 / DAY
        time5 = [mjd, 0.0, 69.184, xp, yp]
        C, S = orc.solid_tide(far, far, np.eye(3), time5, frequency_terms=False, zero_tide=False)
        m1, m2 = np.linalg.solve([[1.0, 0.0115], [-0.0115, 1.0]], np.array([C[2, 1], S[2, 1]]) / -1.333e-9)
        args, gmst, _ = oracle.tide_arguments(time5)
        Co, So = orc.ocean_tide(time5)
        Cw, Sw = orc.ocean_waves(args, gmst)
        assert Co[2, 1] - Cw[2, 1] == pytest.approx(-2.1778e-10 * (m1 - 0.01724 * m2), rel=1e-6)
        assert So[2, 1] - Sw[2, 1] == pytest.approx(-1.7232e-10 * (m2 - 0.03365 * m1), rel=1e-6)
        d = Co - Cw
        d[2, 1] = 0.0
        assert np.max(np.abs(d)) < 1e-25 and np.max(np.abs(np.delete(So - Sw, 7 * 2 + 1))) < 1e-25
        return 1e3 * (xp - m1), 1e3 * (yp + m2)

    x0, y0 = mean_pole(0.0)
    assert abs(x0 - 55.974) < 1e-4 and abs(y0 - 346.346) < 1e-4
    assert mean_pole(0.0, xp=-0.3, yp=0.7) == pytest.approx((x0, y0), abs=1e-6)
    xa, ya = mean_pole(10.0 - 1e-6)
    xb, yb = mean_pole(10.0 + 1e-6)
    assert abs(xa - xb) < 2e-3 and abs(ya - yb) < 2e-3
    x1, y1 = mean_pole(12.0)
    x2, y2 = mean_pole(16.0)
    x3, y3 = mean_pole(20.0)
    assert (x2 - x1) / 4 == pytest.approx(7.6141, abs=1e-5) and (y2 - y1) / 4 == pytest.approx(-0.6287, abs=1e-5)
    assert abs((x3 - x2) - (x2 - x1)) < 1e-5 and abs((y3 - y2) - (y2 - y1)) < 1e-5


def _tide_potential(rb, C, S, nmax, mu=398600.4415e9, R=6378.1363e3):
    r = np.linalg.norm(rb)
    x, lam = rb[2] / r, math.atan2(rb[1], rb[0])
    u = 0.0
    for n in range(2, nmax + 1):
        for m in range(n + 1):
            u += (R / r) ** n * _pbar(n, m, x) * (C[n, m] * math.cos(m * lam) + S[n, m] * math.sin(m * lam))
    return mu / r * u


@pytest.mark.parametrize("nmax", [4, 6])
def test_tide_acceleration_is_the_gradient_of_the_potential(varying, nmax):
    """HpopCorrections.h:339-367: the tide acceleration is the ICRF gradient of
    U = mu/r sum_n (R/r)^n sum_m Pbar_nm(sin phi) (C cos m lambda + S sin m
    lambda) evaluated in the rotated (Earth-fixed) frame, against central
    differences of that potential (scipy Legendre functions) for random
    coefficients, rotations and positions."""
    _, _, orc = varying
    rng = np.random.default_rng(nmax)
    for _ in range(4):
        C = np.zeros((7, 7))
        S = np.zeros((7, 7))
        for n in range(2, nmax + 1):
            C[n, :n + 1] = rng.normal(size=n + 1) * 1e-9
            S[n, 1:n + 1] = rng.normal(size=n) * 1e-9
        E, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        r = rng.normal(size=3)
        r *= rng.uniform(6.7e6, 4.2e7) / np.linalg.norm(r)
        a = orc.tide_acceleration(r, E, C, S, nmax)
        h = 1e-6 * np.linalg.norm(r)
        g = np.array([(_tide_potential(E @ (r + h * e), C, S, nmax) - _tide_potential(E @ (r - h * e), C, S, nmax)) / (2 * h)
                      for e in np.eye(3)])
        assert np.linalg.norm(a - g) < 1e-7 * np.linalg.norm(g), (a, g)


DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "parahpop_amd", "data")


def _frequency_table():
    rows = [ln.split() for ln in open(os.path.join(DATA, "iers2010_solid_tide_frequency.txt")) if ln.strip() and ln[0] != "#"]
    return [(int(r[0]), np.array([int(v) for v in r[1:6]]), float(r[6]), float(r[7])) for r in rows]


def test_solid_tide_frequency_table_and_terms(varying):
    """HpopCorrections.h:148-169,229-251 with HpopFrequencyTerms.inc: the table
    holds IERS 2010 Tables 6.5a/b/c (21, 48 and 2 rows; known rows 055.565,
    P1, K1, N2 and M2 at their published amplitudes), and the frequency-
    dependent corrections are eqs. 6.8a-c summed in complex form, with theta_f
    = m (theta_g + pi) - N.F."""
    _, _, orc = varying
    import oracle
    table = _frequency_table()
    assert [sum(1 for t in table if t[0] == m) for m in range(3)] == [21, 48, 2]
    known = {(0, (0, 0, 0, 0, 1)): (16.6, -6.7), (1, (0, 0, 2, -2, 2)): (-43.4, 2.9), (1, (0, 0, 0, 0, 0)): (470.9, -30.2),
             (2, (1, 0, 2, 0, 2)): (-0.3, 0.0), (2, (0, 0, 2, 0, 2)): (-1.2, 0.0)}
    got = {(m, tuple(mult)): (ip, op) for m, mult, ip, op in table}
    for k, v in known.items():
        assert got[k] == v, k
    for mjd in (55580.2, 57000.9):
        time5 = [mjd, -0.19, 69.184, 0.05, 0.35]
        args, gmst, _ = oracle.tide_arguments(time5)
        moon, sun = [3.8e8, 1e7, 2e7], [1.4e11, 3.0e10, 1.0e10]
        C1, S1 = orc.solid_tide(moon, sun, np.eye(3), time5, frequency_terms=True, zero_tide=False)
        C0, S0 = orc.solid_tide(moon, sun, np.eye(3), time5, frequency_terms=False, zero_tide=False)
        z = np.zeros(3, dtype=complex)   # dC2m - i dS2m
        for m, mult, ip, op in table:
            e = np.exp(1j * (m * (gmst + math.pi) - mult @ args))
            if m == 0:
                z[0] += ip * e.real - op * e.imag
            elif m == 1:
                z[1] += -1j * (ip + 1j * op) * e
            else:
                z[2] += ip * e
        z *= 1e-12
        for m in range(3):
            assert C1[2, m] - C0[2, m] == pytest.approx(z[m].real, rel=1e-9, abs=1e-22)
            assert S1[2, m] - S0[2, m] == pytest.approx(-z[m].imag if m else 0.0, rel=1e-9, abs=1e-22)
        d = C1 - C0
        d[2, :3] = 0.0
        assert not d.any()


def _fes_table():
    waves, terms = [], []
    for ln in open(os.path.join(DATA, "fes2004_ocean_tide.txt")):
        r = ln.split()
        if r and r[0] == "W":
            waves.append((np.array([int(v) for v in r[1:7]]), int(r[7]), int(r[8]), r[9]))
        elif r and r[0] == "T":
            terms.append((int(r[1]), int(r[2]), *[float(v) for v in r[3:7]]))
    return waves, terms


def test_fes2004_table_follows_its_doodson_numbers():
    """Fes2004Terms.inc: each wave's argument multipliers are its Doodson
    number (digits minus 5 after the first), and the waves' term ranges tile
    the 402 terms in order."""
    waves, terms = _fes_table()
    assert len(waves) == 18 and len(terms) == 402
    for mult, _, _, number in waves:
        digits = number.replace(".", "").zfill(6)
        assert list(mult) == [int(digits[0])] + [int(c) - 5 for c in digits[1:]], number
    assert waves[0][1] == 0 and waves[-1][2] == 402
    assert all(waves[k][2] == waves[k + 1][1] for k in range(17))


def test_ocean_tide_sum_is_iers_eq_6_15(varying):
    """HpopCorrections.h:306-317: dC - i dS = sum over waves of (C+ - i S+)
    e^(i theta) + (C- + i S-) e^(-i theta) (IERS 2010 eq. 6.15), with S[n][0]
    kept zero (FES convention)."""
    _, _, orc = varying
    import oracle
    waves, terms = _fes_table()
    for mjd in (55580.2, 56000.7):
        args, gmst, beta = oracle.tide_arguments([mjd, -0.19, 69.184, 0.05, 0.35])
        C, S = orc.ocean_waves(args, gmst)
        z = np.zeros((7, 7), dtype=complex)
        for mult, b, e, _ in waves:
            theta = float(mult @ beta)
            for n, m, cp, sp, cm, sm in terms[b:e]:
                z[n, m] += (cp - 1j * sp) * np.exp(1j * theta) + (cm + 1j * sm) * np.exp(-1j * theta)
        scale = np.abs(z).max()
        assert np.max(np.abs(C - z.real)) < 1e-12 * scale
        zs = -z.imag
        zs[:, 0] = 0.0
        assert np.max(np.abs(S - zs)) < 1e-12 * scale


def _doodson_textbook(mjd_tt, gmst):
    """Doodson variables from the mean longitudes (Meeus, Astronomical
    Algorithms, ch. 22 and 47): tau = GMST + pi - s, s, h, p, N' = -Omega, p_s."""
    t = (mjd_tt - 51544.5) / 36525.0
    s = math.radians(218.3164477 + 481267.88123421 * t)
    h = math.radians(280.46646 + 36000.76983 * t)
    p = math.radians(83.3532465 + 4069.0137287 * t)
    node = math.radians(125.04452 - 1934.136261 * t)
    ps = math.radians(282.93735 + 1.71946 * t)
    return np.array([gmst + math.pi - s, s, h, p, -node, ps])


def test_ocean_tide_arguments_and_invariants(varying):
    """HpopCorrections.h:298-317: the Doodson arguments built from the Delaunay
    arguments equal the textbook mean longitudes (an index slip moves one by
    degrees); the FES2004 sum has S[n][0] = 0 and is unchanged when every
    argument moves by a multiple of 2 pi."""
    _, _, orc = varying
    import oracle
    rng = np.random.default_rng(3)
    for mjd in (55580.2, 56000.7, 57500.1):
        time5 = [mjd, -0.19, 69.184, 0.05, 0.35]
        args, gmst, beta = oracle.tide_arguments(time5)
        ref = _doodson_textbook(mjd + 69.184 / DAY, gmst)
        dev = (beta - ref + math.pi) % (2 * math.pi) - math.pi
        assert np.max(np.abs(dev)) < 2e-4, dev
        C, S = orc.ocean_waves(args, gmst)
        assert not S[:, 0].any()
        assert np.abs(C).max() > 1e-11
        shift = 2 * math.pi * rng.integers(-3, 4, size=5)
        C2, S2 = orc.ocean_waves(args + shift, gmst + 2 * math.pi * 2)
        scale = np.abs(C).max()
        assert np.max(np.abs(C2 - C)) < 1e-9 * scale and np.max(np.abs(S2 - S)) < 1e-9 * scale


def test_third_body_direct_minus_indirect():
    """PointGravity.h:76-81: the non-barycentric third-body term vanishes at the
    centre, equals gm (1/(D-x)^2 - 1/D^2) on the body's axis, and tends to the
    tidal 2 gm x / D^3 (along the axis) and -gm y / D^3 (across it) near the
    centre (the first term of its expansion in x / D)."""
    import oracle
    gm, D = 4902.800192171393, 384400.0
    body = [D, 0.0, 0.0]
    assert not oracle.third_body([0.0, 0.0, 0.0], body, gm).any()
    for x in (7000.0, -42164.0, 200000.0):
        a = oracle.third_body([x, 0.0, 0.0], body, gm)
        assert a[0] == pytest.approx(gm * (1 / (D - x) ** 2 - 1 / D ** 2), rel=1e-12)
        assert a[1] == 0.0 and a[2] == 0.0
    for x in (10.0, 100.0):
        a = oracle.third_body([x, 0.0, 0.0], body, gm)
        assert a[0] == pytest.approx(2 * gm * x / D ** 3, rel=3 * x / D)
        a = oracle.third_body([0.0, x, 0.0], body, gm)
        assert a[1] == pytest.approx(-gm * x / D ** 3, rel=3 * x / D)


def _overlap_fraction(a, b, c, n=1400):
    """Fraction of a disk of radius a covered by a disk of radius b whose centre
    is c away, by counting grid points."""
    g = (np.arange(n) + 0.5) / n * 2 - 1
    X, Y = np.meshgrid(g * a, g * a)
    inside = X ** 2 + Y ** 2 <= a * a
    covered = inside & ((X - c) ** 2 + Y ** 2 <= b * b)
    return covered.sum() / inside.sum()


def test_shadow_partial_and_annular_phases():
    """Shadow.h:84-142: the lit fraction of the Sun is 1 - (covered area)/(pi
    a^2) for apparent radii a (Sun) and b (occulter) at angular separation c:
    checked against a brute-force area count in the partial phase, equal to
    1 - b^2/a^2 in the annular phase, 0 in the total phase, and continuous at
    the edges of the partial zone."""
    import oracle
    rS, rB = 696000.0, 1738.0
    ds, db = 1.496e8, 378000.0          # distances: the Moon's disk a little smaller than the Sun's
    a, b = math.asin(rS / ds), math.asin(rB / db)
    assert b < a

    def factor(c, db_=db):
        sun = [-ds, 0.0, 0.0]   # object minus Sun: the Sun lies along +x from the object
        body = [-db_ * math.cos(c), -db_ * math.sin(c), 0.0]
        return oracle.shadow(sun, body, rS, rB)

    assert factor(0.0) == pytest.approx(1 - b * b / (a * a), rel=1e-12)      # annular, centred
    assert factor(0.6 * (a - b)) == pytest.approx(1 - b * b / (a * a), rel=1e-12)
    for frac in (0.15, 0.5, 0.85):
        c = (a - b) + frac * 2 * b
        assert factor(c) == pytest.approx(1 - _overlap_fraction(a, b, c), abs=2e-3)
    eps = 1e-9
    assert factor(a + b - eps) == pytest.approx(1.0, abs=1e-4) and factor(a + b + eps) == 1.0
    assert factor(a - b + eps) == pytest.approx(1 - b * b / (a * a), abs=1e-4)
    # total phase: an occulter whose disk exceeds the Sun's
    big = math.asin(rB / 300000.0)
    assert big > a

    def factor_big(c):
        return oracle.shadow([-ds, 0.0, 0.0], [-300000.0 * math.cos(c), -300000.0 * math.sin(c), 0.0], rS, rB)

    assert factor_big(0.0) == 0.0 and factor_big(0.5 * (big - a)) == 0.0
    assert factor_big(big - a + eps) == pytest.approx(0.0, abs=1e-4)


def test_sh_polar_limit_is_the_limit_of_the_sum(full_workload):
    """Coefficients.h:109-132 / DegreeLoop.h:98-103: at the literal pole
    (x = y = 0) the order-1 terms use the analytic limit of P_n1 / cos(phi); the
    result must be the limit of the ordinary sum as the position approaches the
    pole from any longitude (both poles)."""
    import oracle
    cfg, _ = full_workload
    orc = oracle.Oracle(cfg)
    for z in (6778.0, -7100.0):
        at_pole = orc.sh_body_fixed([0.0, 0.0, z])
        assert np.all(np.isfinite(at_pole)) and np.linalg.norm(at_pole) > 0
        for lam in (0.0, 1.0, 2.5, -2.0):
            eps = 1e-7
            near = orc.sh_body_fixed([eps * math.cos(lam), eps * math.sin(lam), z])
            assert np.max(np.abs(near - at_pole)) < 1e-9 * np.linalg.norm(at_pole), (z, lam, near, at_pole)
