"""Seeded synthetic inputs for the five propagation workloads.

Writes the reference's own file formats so the same directory drives either
implementation: CONFIG.json / TARGETS.json (docs/configuration.md), a PDS-SHA
gravity file (interface/config/model/sphericalharmonics/Coefficients.h:188-362),
a CelesTrak SW-All file (model/environment/atmosphere/Nrlmsise00Atmosphere.h:146-188),
an HPOP EOP-All file (model/environment/earthcorrections/EarthCorrections.h:141-177)
and constants.json (data/constants.json).

Workloads (object populations and force models):
  1  1,024 LEO objects, central gravity only, 60 s step, 600 s
  2  16,384 LEO/MEO objects, 70x70 field + NRLMSISE-00 + SRP/shadow + Sun,
     Moon, planets + solid/ocean tides + relativity, 60 s step, 1 day
  3  as 2 with a 120x120 Kaula-rule field, 30 s step, mixed precision
  4  131,072 objects, uneven epochs and durations, 70% LEO/20% MEO/10% GEO
  5  serving stream (request generator in serve.py)
"""
import json
import math
import os

import numpy as np

GM_EARTH = 398600.4415
R_EARTH = 6378.137
R_FIELD = 6378.1363
EPOCH0 = 4035.0

CONSTANTS = [
    {"CONSTANTS": "synthetic set (HPOP 4.2.2 / GGM03C / DE440 compatible values)",
     "AU": 149597870.7, "CLIGHT": 299792.458, "EMRAT": 81.30056822149722, "PS": 4.559821181358739e-09},
    {"NAME": "Sun", "NAIFID": 10, "GM": 132712440041.27942, "SOI": 10000000000.0, "R": 696000.0,
     "A_MEAN": 0.0, "J2": 0.0},
    {"NAME": "Moon", "NAIFID": 301, "MAIN_BODY": "Earth", "GM": 4902.800192171393, "SOIOld": 66194.3772804927,
     "R": 1738.0, "A_MEAN": 384466.5276989999, "J2": 0.00020343},
    {"NAME": "Earth", "NAIFID": 399, "MAIN_BODY": "Sun", "GM": 398600.4415, "SOIOld": 924646.788599425,
     "R": 6378.137, "A_MEAN": 149597870.7, "J2": 0.0010826355254902923},
]

PLANET_GM = {
    "MercuryBarycenter": 22031.868551, "VenusBarycenter": 324858.592, "MarsBarycenter": 42828.375816,
    "JupiterBarycenter": 126712764.1, "SaturnBarycenter": 37940584.8418, "UranusBarycenter": 5794556.4,
    "NeptuneBarycenter": 6836527.10058, "PlutoBarycenter": 975.5,
}


def kepler_to_cartesian(a, e, inc, raan, argp, M, gm=GM_EARTH):
    """Classical elements (km, rad) -> ICRF position/velocity (km, km/s), vectorised."""
    E = M.copy()
    for _ in range(12):
        E = E - (E - e * np.sin(E) - M) / (1.0 - e * np.cos(E))
    cE, sE = np.cos(E), np.sin(E)
    sq = np.sqrt(1.0 - e * e)
    x_p = a * (cE - e)
    y_p = a * sq * sE
    r = a * (1.0 - e * cE)
    n = np.sqrt(gm / a ** 3)
    vx_p = -a * n * sE / (1.0 - e * cE)
    vy_p = a * n * sq * cE / (1.0 - e * cE)
    cO, sO, cw, sw, ci, si = np.cos(raan), np.sin(raan), np.cos(argp), np.sin(argp), np.cos(inc), np.sin(inc)
    P = np.stack([cO * cw - sO * sw * ci, sO * cw + cO * sw * ci, sw * si])
    Q = np.stack([-cO * sw - sO * cw * ci, -sO * sw + cO * cw * ci, cw * si])
    pos = P * x_p + Q * y_p
    vel = P * vx_p + Q * vy_p
    del r
    return pos, vel


def population(n, seed, mix=(1.0, 0.0, 0.0), leo=(300.0, 2000.0), meo=(19000.0, 24000.0), physical=True):
    """Objects drawn from seeded distributions; mix = fractions (LEO, MEO, GEO)."""
    rng = np.random.default_rng(seed)
    kind = rng.choice(3, size=n, p=np.array(mix) / sum(mix))
    alt = np.where(kind == 0, rng.uniform(*leo, n), rng.uniform(*meo, n))
    a = R_EARTH + alt
    a = np.where(kind == 2, 42164.0 + rng.uniform(-50.0, 50.0, n), a)
    e = rng.uniform(0.0, 0.02, n)
    inc = np.radians(rng.uniform(0.0, 180.0, n))
    raan = np.radians(rng.uniform(0.0, 360.0, n))
    argp = np.radians(rng.uniform(0.0, 360.0, n))
    M = np.radians(rng.uniform(0.0, 360.0, n))
    # keep perigee above 150 km for the LEO shell
    a = np.maximum(a, (R_EARTH + 150.0) / (1.0 - e))
    pos, vel = kepler_to_cartesian(a, e, inc, raan, argp, M)
    out = {"state": np.vstack([pos, vel]), "kind": kind}
    if physical:
        mass = np.exp(rng.uniform(math.log(10.0), math.log(5000.0), n))
        am = np.exp(rng.normal(math.log(0.01), 1.0, n))  # m^2/kg
        area = am * mass
        out.update(mass=mass, drag_area=area, srp_area=area.copy(), cd=np.full(n, 2.2), cr=np.full(n, 1.3))
    else:
        out.update(mass=np.full(n, 1000.0), drag_area=np.full(n, 10.0), srp_area=np.full(n, 10.0),
                   cd=np.full(n, 2.2), cr=np.full(n, 1.3))
    return out


def write_targets(path, pop, epochs, ids=None):
    n = pop["state"].shape[1]
    ids = np.arange(1, n + 1) if ids is None else ids
    st = pop["state"]
    targets = []
    for i in range(n):
        targets.append({
            "id": int(ids[i]),
            "epoch_mjd2000_tdb": float(epochs[i]),
            "state_km_km_s": [float(st[d, i]) for d in range(6)],
            "mass_kg": float(pop["mass"][i]),
            "drag_area_m2": float(pop["drag_area"][i]),
            "srp_area_m2": float(pop["srp_area"][i]),
            "cd": float(pop["cd"][i]),
            "cr": float(pop["cr"][i]),
        })
    with open(path, "w") as f:
        json.dump({"description": "synthetic population", "targets": targets}, f, separators=(",", ":"))


def write_gravity(path, degree, seed, kaula=False, radius=R_FIELD, gm=GM_EARTH, c20=-4.8417e-4, scale=1e-5):
    """Fully normalized PDS-SHA field: C20 = c20, other C/S ~ N(0, (scale/n^2)^2), S_n0 = 0."""
    rng = np.random.default_rng(seed)
    lines = ["%.10f, %.10f, 0.0, %d, %d, 1, 0.0, 0.0" % (radius, gm, degree, degree)]
    for n in range(1, degree + 1):
        for m in range(n + 1):
            if n == 1:
                c = s = 0.0
            else:
                sig = scale / n ** 2
                c = rng.normal(0.0, sig)
                s = rng.normal(0.0, sig) if m > 0 else 0.0
                if n == 2 and m == 0:
                    c = c20
            lines.append("%d, %d, %.17e, %.17e" % (n, m, c, s))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def mjd_to_calendar(mjd):
    jd = int(math.floor(mjd)) + 2400001
    a = jd + 32044
    b = (4 * a + 3) // 146097
    c = a - 146097 * b // 4
    d = (4 * c + 3) // 1461
    e = c - 1461 * d // 4
    m = (5 * e + 2) // 153
    day = e - (153 * m + 2) // 5 + 1
    month = m + 3 - 12 * (m // 10)
    year = 100 * b + d - 4800 + m // 10
    return year, month, day


def varying_weather_day(k):
    """Space weather of day k of a varying table: adjusted F10.7 and its
    81-day centred mean, daily Ap and the eight 3-hour ap, every value distinct
    from the neighbouring days' and slots' (so a wrong day, slot or lag shows)."""
    f107 = 90.0 + 5.0 * k + 17.0 * math.sin(0.9 * k)
    f107a = 120.0 + 0.75 * k
    aps = [3 + (11 * k + 7 * j + (j * j) % 5) % 61 for j in range(8)]
    ap_daily = 4 + (13 * k) % 47
    return f107, f107a, ap_daily, aps


def write_space_weather(path, mjd_first, days, f107=150.0, ap=15, varying=False):
    """CelesTrak SW-All layout: y m d bsrn nd kp*8 sum ap*8 avg cp c9 isn adjF q adjC81 adjL81 obsF obsC81 obsL81.
    Constant (F10.7 = F10.7A = f107, ap = Ap = ap) or varying (varying_weather_day)."""
    lines = []
    for k in range(days):
        y, mo, d = mjd_to_calendar(mjd_first + k)
        if varying:
            fk, fa, apd, aps = varying_weather_day(k)
        else:
            fk, fa, apd, aps = f107, f107, ap, [ap] * 8
        kp = " ".join(["27"] * 8)
        lines.append("%d %02d %02d 2420 %d %s 216 %s %d 0.8 4 50 %.1f 0 %.1f %.1f %.1f %.1f %.1f" % (
            y, mo, d, k + 1, kp, " ".join(str(a) for a in aps), apd, fk, fa, fa, fk, fa, fa))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def varying_eop_day(k, step_day):
    """EOP of day k of a varying table: xp, yp (arcsec) and UT1-UTC (s) drifting
    linearly, TAI-UTC 37 s before day step_day and 38 s from it."""
    return 0.05 + 0.0013 * k, 0.35 - 0.0009 * k, -0.2 + 0.0011 * k, 37.0 if k < step_day else 38.0


def write_eop(path, mjd_first, days, varying=False, step_day=None):
    """HPOP EOP-All rows: y m d mjd xp yp ut1-utc lod dpsi deps dx dy tai-utc.
    Constant (zero pole, UT1 = UTC, TAI-UTC 37 s) or varying (varying_eop_day,
    with the TAI-UTC step on day step_day, by default mid-table)."""
    if step_day is None:
        step_day = days // 2
    lines = []
    for k in range(days):
        y, mo, d = mjd_to_calendar(mjd_first + k)
        xp, yp, dut, dat = varying_eop_day(k, step_day) if varying else (0.0, 0.0, 0.0, 37.0)
        lines.append("%d %2d %2d %d %.6f %.6f %.7f 0.0000000 0.000000 0.000000 0.000000 0.000000 %d"
                     % (y, mo, d, mjd_first + k, xp, yp, dut, int(dat)))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def full_model(degree, gravity_file, start_mjd_utc, end_mjd_utc):
    bodies = {
        "Earth": {
            "occulting": True,
            "flattening": 0.0033528106647474805,
            "gravity": {"shape": {"body": "Earth", "file": gravity_file, "delimiter": ",", "headerLines": 0,
                                  "degree": degree, "order": degree}},
            "atmosphere": {"nrlmsise00": {"spaceWeatherFile": "SW-All.txt", "coverageStartMjdUtc": start_mjd_utc,
                                          "coverageEndMjdUtc": end_mjd_utc, "hCutoff": 1000.0, "densityScale": 1.0}},
        },
        "Sun": {"gravity": True, "radiation": True},
        "Moon": {"gravity": True},
    }
    for name, gm in PLANET_GM.items():
        bodies[name] = {"gm": gm, "gravity": True}
    return {"environment": {
        "bodies": bodies, "ephemeris": "analytic", "orientations": "analytic", "constants": "constants.json",
        "preferSSB": True,
        "earthCorrections": {"model": "iers2010_fes2004", "tideSystem": "zero_tide", "solidEarthTides": True,
                             "oceanTides": True, "relativity": True, "eopFile": "EOP-All.txt"}}}


# Exponential atmosphere bands (base altitude km, base density kg/m^3, scale
# height km): Vallado, Fundamentals of Astrodynamics and Applications, Table 8-4.
EXPONENTIAL_BANDS = [
    (0, 1.225, 7.249), (25, 3.899e-2, 6.349), (30, 1.774e-2, 6.682), (40, 3.972e-3, 7.554),
    (50, 1.057e-3, 8.382), (60, 3.206e-4, 7.714), (70, 8.770e-5, 6.549), (80, 1.905e-5, 5.799),
    (90, 3.396e-6, 5.382), (100, 5.297e-7, 5.877), (110, 9.661e-8, 7.263), (120, 2.438e-8, 9.473),
    (130, 8.484e-9, 12.636), (140, 3.845e-9, 16.149), (150, 2.070e-9, 22.523), (180, 5.464e-10, 29.740),
    (200, 2.789e-10, 37.105), (250, 7.248e-11, 45.546), (300, 2.418e-11, 53.628), (350, 9.518e-12, 53.298),
    (400, 3.725e-12, 58.515), (450, 1.585e-12, 60.828), (500, 6.967e-13, 63.822), (600, 1.454e-13, 71.835),
    (700, 3.614e-14, 88.667), (800, 1.170e-14, 124.64), (900, 5.245e-15, 181.05), (1000, 3.019e-15, 268.00),
]


def exponential_blocks(bands=EXPONENTIAL_BANDS, top=1500.0):
    """Atmosphere blocks of the config schema: rho0 in kg/km^3, each band's cutoff
    at the next band's base (the last at `top`)."""
    out = []
    for k, (h0, rho, scale) in enumerate(bands):
        cut = bands[k + 1][0] if k + 1 < len(bands) else top
        out.append({"rho0": rho * 1e9, "h0": float(h0), "scaleHeight": float(scale), "hCutoff": float(cut)})
    return out


def alternate_model(atmosphere=None, flattening=0.0033528106647474805, third_bodies=True):
    """J2-only Earth gravity with an exponential atmosphere (piecewise by default),
    Sun and Moon point masses, SRP with Earth shadow and relativity."""
    earth = {"gravity": {"oblateness": True}, "flattening": flattening, "occulting": True,
             "atmosphere": {"exponential": exponential_blocks() if atmosphere is None else atmosphere}}
    bodies = {"Earth": earth}
    if third_bodies:
        bodies["Sun"] = {"gravity": True, "radiation": True}
        bodies["Moon"] = {"gravity": True}
    return {"environment": {
        "bodies": bodies, "ephemeris": "analytic", "orientations": "analytic", "constants": "constants.json",
        "preferSSB": True, "earthCorrections": {"model": "iers2010_fes2004", "relativity": True}}}


def make_alternate_workload(out, n=16384, duration=86400.0, step=60.0, seed=2, model=None):
    """Workload-2 population (80 % LEO 300-2000 km, 20 % MEO) under the alternate
    force model; returns the CONFIG.json and TARGETS.json paths."""
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "constants.json"), "w") as f:
        json.dump(CONSTANTS, f, indent=1)
    pop = population(n, seed, mix=(0.8, 0.2, 0.0))
    epochs = np.full(n, EPOCH0)
    max_steps = int(math.ceil(duration / step)) + 2
    write_config(os.path.join(out, "CONFIG.json"), model or alternate_model(), duration, step, max_steps)
    write_targets(os.path.join(out, "TARGETS.json"), pop, epochs)
    return os.path.join(out, "CONFIG.json"), os.path.join(out, "TARGETS.json")


GM_MOON = 4902.800192171393
R_MOON_FIELD = 1738.0


def moon_pa_euler(et, libration=0.01):
    """3-1-3 Euler angles (phi, theta, psi; radians) of a stand-in lunar
    principal-axis frame: the IAU model's (pi/2 + RA, pi/2 - Dec, W) without
    nutation-precession terms, plus a 27.3-day libration of `libration` rad in
    each angle."""
    T = et / 3155760000.0
    d = et / 86400.0
    a = np.radians(269.9949 + 0.0031 * T)
    dl = np.radians(66.5392 + 0.0130 * T)
    w = np.radians(38.3213 + 13.17635815 * d)
    lib = libration * np.sin(2.0 * np.pi * d / 27.321661)
    return np.array([np.pi / 2 + a + lib, np.pi / 2 - dl + 0.5 * lib, w - lib])


def moon_ecliptic_euler(et):
    """3-1-3 Euler angles from ECLIPJ2000 of a stand-in lunar frame: the node of
    the lunar equator on the ecliptic (regressing in 18.6 years), its 1.5424 deg
    inclination and the rotation angle."""
    d = et / 86400.0
    node = np.radians(125.045 - 0.0529921 * d)
    return np.array([node, np.radians(1.5424), np.radians(180.0 + 38.3213 + 13.17635815 * d) - node])


def make_lunar_workload(out, field="sh", frame="iau", degree=8, duration=7200.0, step=30.0, seed=3):
    """A model with a field of the Moon in its own axes: Earth (point mass with
    a degree-4 field), Sun and Moon point masses, the Moon's degree-`degree`
    coefficient field ("sh") or J2 ("j2"), its axes from a .brot2 file holding
    the IAU Moon model (types 1 and 2) and, for frame "moon_pa" / "moon_pa_ecl",
    a MOON_PA binary frame (type 3, ICRF or ECLIPJ2000 base) that the body
    resolves to (frame "iau_named" names IAU_MOON explicitly). Returns the
    CONFIG.json path; the targets are the caller's."""
    from . import brie_io
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "constants.json"), "w") as f:
        json.dump(CONSTANTS, f, indent=1)
    write_gravity(os.path.join(out, "earth4.txt"), 4, seed)
    write_gravity(os.path.join(out, "moon.txt"), degree, seed + 1, radius=R_MOON_FIELD, gm=GM_MOON, c20=-9.09e-5,
                  scale=2e-5)
    units = brie_io.moon_iau_units()
    et0 = (EPOCH0 - 2.0) * 86400.0 - 43200.0
    if frame in ("moon_pa", "moon_pa_ecl"):
        euler = moon_pa_euler if frame == "moon_pa" else moon_ecliptic_euler
        units.append(brie_io.binary_frame_unit(31008, euler, et0, et0 + 6 * 86400.0, 86400.0 / 4, 12,
                                               inert_frame=1 if frame == "moon_pa" else 17))
    brie_io.write_brot2_units(os.path.join(out, "moon.brot2"), units)
    moon = {"gravity": {"shape": {"body": "Moon", "file": "moon.txt", "delimiter": ",", "headerLines": 0,
                                  "degree": degree, "order": degree}} if field == "sh" else {"oblateness": True}}
    if frame == "iau_named":
        moon["orientation"] = "IAU_MOON"
    bodies = {"Earth": {"gravity": {"shape": {"body": "Earth", "file": "earth4.txt", "delimiter": ",", "headerLines": 0,
                                              "degree": 4, "order": 4}}},
              "Sun": {"gravity": True}, "Moon": moon}
    model = {"environment": {"bodies": bodies, "ephemeris": "analytic", "orientations": ["moon.brot2"],
                             "constants": "constants.json", "preferSSB": True}}
    cfg = os.path.join(out, "CONFIG.json")
    write_config(cfg, model, duration, step, int(math.ceil(duration / step)) + 2)
    return cfg


def central_model():
    return {"environment": {"bodies": {"Earth": {"gravity": True}}, "ephemeris": "analytic",
                            "constants": "constants.json", "preferSSB": True}}


def write_config(path, model, duration, step, max_steps, precision=None):
    cfg = {"run": {"duration_seconds": float(duration)},
           "execution": {"maxThreads": 1, "integration": {"initialStepSize": float(step), "maxNumSteps": int(max_steps)}},
           "model": model}
    if precision:
        cfg["execution"]["precision"] = precision
    with open(path, "w") as f:
        json.dump(cfg, f, indent=2)


def environment_files(out, degree, seed, mjd_lo, mjd_hi, kaula=False, varying=False):
    with open(os.path.join(out, "constants.json"), "w") as f:
        json.dump(CONSTANTS, f, indent=1)
    gfile = "gravity_%d.tab" % degree
    write_gravity(os.path.join(out, gfile), degree, seed, kaula)
    first = int(math.floor(mjd_lo)) - 10
    days = int(math.ceil(mjd_hi)) - first + 10
    write_space_weather(os.path.join(out, "SW-All.txt"), first, days, varying=varying)
    # the TAI-UTC step falls at the first UTC midnight after the workload starts
    # (MJD 55579, 69 s after an epoch of ET 4034.5 days)
    write_eop(os.path.join(out, "EOP-All.txt"), first, days, varying=varying, step_day=10)
    return gfile


def make_workload(out, which, n=None, duration=None, step=None, degree=None, seed=None, precision=None, varying=False):
    """Write CONFIG.json, TARGETS.json and data files for workload 1..4; returns the paths.
    varying: space weather and EOP tables whose values change every day and
    3-hour slot (drifting pole and UT1, one TAI-UTC step) instead of constant ones."""
    os.makedirs(out, exist_ok=True)
    mjd_utc0 = EPOCH0 + 51544.5
    if which == 1:
        n = n or 1024
        pop = population(n, seed or 1, mix=(1, 0, 0), leo=(400.0, 1200.0), physical=False)
        epochs = np.full(n, EPOCH0)
        with open(os.path.join(out, "constants.json"), "w") as f:
            json.dump(CONSTANTS, f, indent=1)
        model = central_model()
        duration = duration or 600.0
        step = step or 60.0
    elif which in (2, 3):
        n = n or 16384
        pop = population(n, seed or 2, mix=(0.8, 0.2, 0.0))
        epochs = np.full(n, EPOCH0)
        degree = degree or (70 if which == 2 else 120)
        duration = duration or 86400.0
        step = step or (60.0 if which == 2 else 30.0)
        gfile = environment_files(out, degree, 70 if which == 2 else 120, mjd_utc0, mjd_utc0 + duration / 86400.0,
                                  varying=varying)
        model = full_model(degree, gfile, mjd_utc0 - 1, mjd_utc0 + duration / 86400.0 + 1)
        if which == 3 and precision is None:
            precision = {"mode": "mixed", "fp64Degree": 20}
    elif which == 4:
        n = n or 131072
        pop = population(n, seed or 4, mix=(0.7, 0.2, 0.1))
        rng = np.random.default_rng((seed or 4) + 1000)
        epochs = EPOCH0 + rng.uniform(0.0, 30.0, n)
        degree = degree or 70
        step = step or 60.0
        duration = duration or 86400.0
        gfile = environment_files(out, degree, 70, mjd_utc0, mjd_utc0 + 30.0 + 8.0, varying=varying)
        model = full_model(degree, gfile, mjd_utc0 - 1, mjd_utc0 + 39.0)
    else:
        raise ValueError("workload must be 1..4")
    max_steps = int(math.ceil(duration / step)) + 2
    write_config(os.path.join(out, "CONFIG.json"), model, duration, step, max_steps, precision)
    write_targets(os.path.join(out, "TARGETS.json"), pop, epochs)
    return os.path.join(out, "CONFIG.json"), os.path.join(out, "TARGETS.json")


def load_targets(path):
    """TARGETS.json -> dict of numpy arrays in the C-ABI layout (state SoA [6][n])."""
    with open(path) as f:
        t = json.load(f)["targets"]
    n = len(t)
    return {
        "id": np.array([x["id"] for x in t], dtype=np.int64),
        "epoch": np.array([x["epoch_mjd2000_tdb"] for x in t], dtype=np.float64),
        "state": np.ascontiguousarray(np.array([x["state_km_km_s"] for x in t], dtype=np.float64).T),
        "mass": np.array([x["mass_kg"] for x in t], dtype=np.float64),
        "drag_area": np.array([x["drag_area_m2"] for x in t], dtype=np.float64),
        "srp_area": np.array([x["srp_area_m2"] for x in t], dtype=np.float64),
        "cd": np.array([x["cd"] for x in t], dtype=np.float64),
        "cr": np.array([x["cr"] for x in t], dtype=np.float64),
        "n": n,
    }


def targets_from_population(pop, epochs):
    n = pop["state"].shape[1]
    return {"id": np.arange(1, n + 1, dtype=np.int64), "epoch": np.asarray(epochs, dtype=np.float64),
            "state": np.ascontiguousarray(pop["state"], dtype=np.float64), "mass": pop["mass"],
            "drag_area": pop["drag_area"], "srp_area": pop["srp_area"], "cd": pop["cd"], "cr": pop["cr"], "n": n}


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workload", type=int)
    ap.add_argument("out")
    ap.add_argument("--n", type=int)
    ap.add_argument("--duration", type=float)
    args = ap.parse_args()
    print(make_workload(args.out, args.workload, n=args.n, duration=args.duration))
