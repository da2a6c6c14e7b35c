"""Writes Chebyshev ephemerides in the reference's .brie format (CBOR; layout 2
as third_party/brie/brie/util/cbor.cpp:200-300 reads it, layout 1 as
third_party/brie/brie/states/metadata/EphUnit.h:50-85 reads it), fitted to given
position functions, with the body-unit data layout of
third_party/brie/brie/core/BodyInterpolator.h:55-100 (radius, interval starts,
coefficients [component][degree][interval]).

`ssb_units` builds the reference's chain structure (planetary barycenters, the
Sun and the Earth-Moon barycenter about the solar-system barycenter, the Earth
and the Moon about the Earth-Moon barycenter) from geocentric position
functions; `circular_geocentric` is a smooth stand-in ephemeris (circular
orbits) for synthetic workloads."""
import math
import struct

import numpy as np


def cbor(obj):
    """Minimal CBOR encoder: maps with text keys, arrays, ints, float64, text."""
    def head(major, n):
        if n < 24:
            return bytes([(major << 5) | n])
        for info, fmt in ((24, ">B"), (25, ">H"), (26, ">I"), (27, ">Q")):
            if n < (1 << (8 * struct.calcsize(fmt))):
                return bytes([(major << 5) | info]) + struct.pack(fmt, n)
        raise ValueError(n)
    if isinstance(obj, bool):
        return bytes([0xf5 if obj else 0xf4])
    if isinstance(obj, (int, np.integer)):
        return head(0, int(obj)) if obj >= 0 else head(1, -1 - int(obj))
    if isinstance(obj, (float, np.floating)):
        return b"\xfb" + struct.pack(">d", float(obj))
    if isinstance(obj, str):
        b = obj.encode()
        return head(3, len(b)) + b
    if isinstance(obj, (list, tuple, np.ndarray)):
        return head(4, len(obj)) + b"".join(cbor(x) for x in obj)
    if isinstance(obj, dict):
        return head(5, len(obj)) + b"".join(cbor(k) + cbor(v) for k, v in obj.items())
    raise TypeError(type(obj))


def fit_unit(func, t0, t1, interval, pdeg):
    """Chebyshev fit of func(et) -> (3,) on intervals [t0 + k h, t0 + (k+1) h).
    Returns the body-unit data array and the number of intervals."""
    n = int(math.ceil((t1 - t0) / interval))
    r = 0.5 * interval
    starts = t0 + interval * np.arange(n)
    x = np.cos(np.pi * (np.arange(pdeg + 1) + 0.5) / (pdeg + 1))   # Chebyshev nodes
    coef = np.zeros((3, pdeg + 1, n))
    for k in range(n):
        vals = np.array([func(starts[k] + r * (xi + 1.0)) for xi in x])   # (pdeg+1, 3)
        for d in range(3):
            coef[d, :, k] = np.polynomial.chebyshev.chebfit(x, vals[:, d], pdeg)
    data = np.concatenate([[r], starts, coef.reshape(-1)])
    return data, n


def write_brie(path, units, layout=2):
    """units: dicts with target, center, t0, t1, interval, pdeg, func."""
    built = []
    for u in units:
        data, n = fit_unit(u["func"], u["t0"], u["t1"], u["interval"], u["pdeg"])
        built.append((u, data, n))
    if layout == 2:
        N = len(built)
        fields = [[], [], [], [], [], [], [], []]
        reals = [[], []]
        offset = 0
        for u, data, n in built:
            for f, v in enumerate((1, 2, u["target"], u["center"], n, u["pdeg"], offset, len(data))):
                fields[f].append(int(v))
            reals[0].append(float(data[1]))
            reals[1].append(float(data[1] + 2 * data[0] * n))
            offset += len(data)
        doc = {"layout": 2, "core": {
            "metadata": {"nBodyUnits": N, "intMetadata": [v for f in fields for v in f],
                         "doubleMetadata": reals[0] + reals[1]},
            "data": np.concatenate([d for _, d, _ in built])}}
    else:
        doc = {"layout": 1, "core": [
            {"metadata": {"frame": 1, "dtype": 2, "target": u["target"], "center": u["center"], "nintervals": n,
                          "pdeg": u["pdeg"], "startEpoch": float(d[1]), "finalEpoch": float(d[1] + 2 * d[0] * n)},
             "data": d} for u, d, n in built]}
    with open(path, "wb") as f:
        f.write(cbor(doc))
    return built


def ssb_units(geocentric, t0, t1, emrat=81.30056822149722, bodies=(10, 301, 1, 2, 4, 5, 6, 7, 8, 9),
              shift=None):
    """Body units in the reference's chain structure (planetary barycenters,
    the Sun and the Earth-Moon barycenter about the solar-system barycenter, the
    Earth and the Moon about the Earth-Moon barycenter) whose geocentric
    differences reproduce geocentric(naif, et). shift(naif, et) adds a
    displacement to a body's geocentric position."""
    mu = 1.0 / (1.0 + emrat)
    day = 86400.0

    def g(naif, et):
        p = np.asarray(geocentric(naif, et), dtype=float)
        return p + (shift(naif, et) if shift else 0.0)

    def wobble(et):   # the Sun about the barycenter: an arbitrary smooth motion
        a = 2 * np.pi * et / (11.86 * 365.25 * day)
        return 7.4e5 * np.array([np.cos(a), np.sin(a), 0.02 * np.sin(2 * a)])

    def earth_emb(et):
        return -mu * g(301, et)

    units = [
        {"target": 3, "center": 0, "interval": 8 * day, "pdeg": 13,
         "func": lambda et: wobble(et) - g(10, et) - earth_emb(et)},
        {"target": 399, "center": 3, "interval": 4 * day, "pdeg": 13, "func": earth_emb},
        {"target": 10, "center": 0, "interval": 16 * day, "pdeg": 11, "func": wobble},
    ]
    if 301 in bodies:
        units.append({"target": 301, "center": 3, "interval": 4 * day, "pdeg": 13,
                      "func": lambda et: (1.0 - mu) * g(301, et)})
    for b in bodies:
        if b in (10, 301):
            continue
        units.append({"target": b, "center": 0, "interval": {1: 8, 2: 16, 4: 16}.get(b, 32) * day, "pdeg": 13,
                      "func": (lambda b: lambda et: wobble(et) - g(10, et) + g(b, et))(b)})
    for u in units:
        u["t0"], u["t1"] = t0, t1
    return units


def circular_geocentric(naif, et):
    """Geocentric km of a body on circular orbits: the Sun at 1 au (1 year), the
    Moon at 384,400 km (27.32 days, 5.1 deg inclined), barycenters on circular
    heliocentric orbits of their mean distances and periods."""
    day = 86400.0
    t = et / day

    def circ(r, period, phase, inc=0.0):
        a = 2 * np.pi * t / period + phase
        return np.array([r * np.cos(a), r * np.sin(a) * np.cos(inc), r * np.sin(a) * np.sin(inc)])
    sun = circ(1.495978707e8, 365.25636, 1.75, np.radians(23.44))
    if naif == 10:
        return sun
    if naif == 301:
        return circ(384400.0, 27.321661, 0.4, np.radians(23.44 + 5.14))
    orbits = {1: (0.387, 87.97, 0.2), 2: (0.723, 224.70, 1.1), 4: (1.524, 686.98, 2.3), 5: (5.203, 4332.59, 0.6),
              6: (9.537, 10759.2, 4.0), 7: (19.19, 30688.5, 5.1), 8: (30.07, 60182.0, 2.9), 9: (39.48, 90560.0, 3.7)}
    au, period, phase = orbits[naif]
    return sun + circ(au * 1.495978707e8, period, phase, np.radians(23.44))


def write_brot2(path, tt0, tt1, channels, h_nut=0.5, h_erp=1.0, erp_keys=None, target=3000):
    """IERS2000 Earth orientation (.brot2, layout 2, one unit_type-4 unit) with
    the dual-table payload of third_party/brie/brie/orientations/detail/LagrangeInterp.h:25-40:
    channels(tt) -> (X, Y, s, xp, yp, UT1-TT [days], dX, dY) at TT MJD2000 days;
    nutation nodes every h_nut days from tt0, ERP nodes at erp_keys (default
    every h_erp days). The envelope (ET seconds) leaves three nodes of margin."""
    nodes_nut = tt0 + h_nut * np.arange(int(math.ceil((tt1 - tt0) / h_nut)) + 1)
    keys = np.asarray(erp_keys, dtype=float) if erp_keys is not None else \
        tt0 + h_erp * np.arange(int(math.ceil((tt1 - tt0) / h_erp)) + 1)
    nut = np.array([channels(t)[:3] for t in nodes_nut]).T          # (3, n1)
    erp = np.array([channels(t)[3:] for t in keys]).T               # (5, n2)
    n1, n2 = nut.shape[1], erp.shape[1]
    data = np.concatenate([nut.reshape(-1), erp.reshape(-1), [tt0, h_nut, keys[0], h_erp], keys])
    lo = max(nodes_nut[2], keys[1])
    hi = min(nodes_nut[-4], keys[-3])
    ints = [4, target, 0, 1, n1, n2, 5, 3, 0, 0, 0, len(data)]
    doc = {"layout": 2, "core": {"metadata": {"nBodyUnits": 1, "intMetadata": ints,
                                              "doubleMetadata": [(lo - 0.5) * 86400.0, (hi - 0.5) * 86400.0]},
                                 "data": data}}
    with open(path, "wb") as f:
        f.write(cbor(doc))
    return data


def iau2006_channels(xp_arcsec=0.0, yp_arcsec=0.0, wobble=False):
    """Channels of the analytic Earth orientation: IAU 2006 X, Y, s polynomials
    (IERS 2010 Table 5.2a), UT1 = TT - 69.184 s, optional polar motion (a
    constant, or a one-year wobble of that amplitude) and zero CIP corrections."""
    asec = np.pi / 648000.0

    def ch(tt):
        t = (tt - 0.5) / 36525.0
        X = (-0.016617 + t * (2004.191898 + t * (-0.4297829 + t * (-0.19861834)))) * asec
        Y = (-0.006951 + t * (-0.025896 + t * (-22.4072747 + t * 0.00190059))) * asec
        s = -0.5 * X * Y + (0.000094 + 0.00380865 * t) * asec
        a = 2 * np.pi * tt / 433.0 if wobble else 0.0
        return (X, Y, s, xp_arcsec * asec * np.cos(a), yp_arcsec * asec * np.sin(a) if wobble else yp_arcsec * asec,
                -69.184 / 86400.0, 0.0, 0.0)
    return ch


# ---------------------------------------------------------------------------
# Body-orientation units of .brot2 files (third_party/brie/brie/orientations/
# metadata/Layout.h:19-60): twelve int fields per unit, field-major (unit type,
# target, eight unit fields, data offset, data size) and the envelope (start,
# final ET seconds) as the real fields.

def iau_body_unit(target, ra, dec, pm, np_ra=(), np_dec=(), np_pm=(), barycenter=0, radii=()):
    """Type-1 unit: RA and Dec polynomials in Julian centuries, the prime
    meridian polynomial in days past J2000 TDB (degrees, ascending powers), then
    the nutation-precession amplitudes of RA (sine), Dec (cosine) and PM (sine)
    over the angles of the `barycenter` type-2 unit."""
    data = np.concatenate([np.asarray(x, dtype=float) for x in (ra, dec, pm, np_ra, np_dec, np_pm, radii)])
    ints = [len(ra) - 1, len(dec) - 1, len(pm) - 1, len(radii), barycenter, len(np_ra), len(np_dec), len(np_pm)]
    return {"type": 1, "target": target, "ints": ints, "data": data, "start": -1e12, "final": 1e12}


def np_angles_unit(target, angles):
    """Type-2 unit: nutation-precession angle polynomials (degrees, Julian
    centuries, ascending powers), all of one degree."""
    a = np.asarray(angles, dtype=float)
    return {"type": 2, "target": target, "ints": [a.shape[1] - 1, a.shape[0], 0, 0, 0, 0, 0, 0],
            "data": a.reshape(-1), "start": -1e12, "final": 1e12}


def binary_frame_unit(target, euler, t0, t1, interval, pdeg, inert_frame=1):
    """Type-3 unit: Chebyshev fit of euler(et) -> (phi, theta, psi) (radians,
    3-1-3 angles from the base frame: 1 ICRF, 17 ECLIPJ2000) in the ephemeris
    body-unit layout; the envelope is the fitted span."""
    data, n = fit_unit(euler, t0, t1, interval, pdeg)
    return {"type": 3, "target": target, "ints": [inert_frame, 2, n, pdeg, 0, 0, 0, 0], "data": data,
            "start": float(data[1]), "final": float(data[1] + 2 * data[0] * n)}


def write_brot2_units(path, units):
    """One .brot2 file (layout 2) holding the given units (dicts from the unit
    builders above)."""
    N = len(units)
    fields = [[] for _ in range(12)]
    offset = 0
    for u in units:
        vals = [u["type"], u["target"]] + list(u["ints"]) + [offset, len(u["data"])]
        for f, v in enumerate(vals):
            fields[f].append(int(v))
        offset += len(u["data"])
    doc = {"layout": 2, "core": {"metadata": {"nBodyUnits": N, "intMetadata": [v for f in fields for v in f],
                                              "doubleMetadata": [float(u["start"]) for u in units] +
                                                                [float(u["final"]) for u in units]},
                                 "data": np.concatenate([np.asarray(u["data"], dtype=float) for u in units])}}
    with open(path, "wb") as f:
        f.write(cbor(doc))


def moon_iau_units(terms=4):
    """The Moon's IAU rotation model (WGCCRE 2009 / NAIF pck00010 Moon block,
    truncated to the first `terms` nutation-precession angles E1..E13 of the
    Earth barycenter, in degrees and centuries): a type-1 unit for body 301 and
    the type-2 unit of barycenter 3."""
    E = np.array([[125.045, -1935.5364525], [250.089, -3871.072905], [260.008, 475263.3328725],
                  [176.625, 487269.629985], [357.529, 35999.0509575], [311.589, 964468.49931],
                  [134.963, 477198.869325], [276.617, 12006.300765], [34.226, 63863.5132425],
                  [15.134, -5806.6093575], [119.743, 131.84064], [239.961, 6003.1503825],
                  [25.053, 473327.79642]])[:terms]
    amp_ra = [-3.8787, -0.1204, 0.0700, -0.0172, 0.0, 0.0072, 0.0, 0.0, 0.0, -0.0052, 0.0, 0.0, 0.0043][:terms]
    amp_dec = [1.5419, 0.0239, -0.0278, 0.0068, 0.0, -0.0029, 0.0009, 0.0, 0.0, 0.0008, 0.0, 0.0, -0.0009][:terms]
    amp_pm = [3.5610, 0.1208, -0.0642, 0.0158, 0.0252, -0.0066, -0.0047, -0.0046, 0.0028, 0.0052, 0.0040, 0.0019,
              -0.0044][:terms]
    body = iau_body_unit(301, [269.9949, 0.0031], [66.5392, 0.0130], [38.3213, 13.17635815, -1.4e-12],
                         amp_ra, amp_dec, amp_pm, barycenter=3, radii=(1737.4, 1737.4, 1737.4))
    return [body, np_angles_unit(3, E)]
