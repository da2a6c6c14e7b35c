"""ORACLE — test infrastructure only (parity checker and timed CPU baseline).

ctypes binding of liboracle.so, the CPU restatement of the reference host
propagation path. Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
DATA_DIR = os.path.join(os.path.dirname(HERE), "parahpop_amd", "data")


class OrcStats(ctypes.Structure):
    _fields_ = [("object_steps", ctypes.c_double), ("rhs_evaluations", ctypes.c_double),
                ("seconds", ctypes.c_double), ("threads", ctypes.c_int)]


FLOPS_LIB_PATH = os.path.join(HERE, "liboracle_flops.so")
_lib = None
_flops_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liboracle.so not built: run `python oracle/build.py`")
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def flops_library():
    """The FLOP-counting build of the oracle (same sources, counting scalar)."""
    global _flops_lib
    if _flops_lib is None:
        if not os.path.exists(FLOPS_LIB_PATH):
            raise ImportError("liboracle_flops.so not built: run `make -C oracle flops`")
        _flops_lib = _bind(ctypes.CDLL(FLOPS_LIB_PATH))
    return _flops_lib


def _bind(lib):
    vp, sz, d, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int
    lib.orc_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
    lib.orc_propagate.argtypes = [vp, sz] + [vp] * 8 + [d, i, vp, vp, vp]
    lib.orc_acceleration.argtypes = [vp, vp, vp, d, d, d, d, d, vp]
    lib.orc_force_term.argtypes = [vp, i, vp, vp, d, d, d, d, d, vp]
    lib.orc_nrlmsise.argtypes = [ctypes.c_char_p, i] + [d] * 8 + [vp, i, vp, vp]
    lib.orc_body_position.argtypes = [i, d, vp]
    lib.orc_earth_orientation.argtypes = [d, vp, vp]
    lib.orc_ctx_body_position.argtypes = [vp, i, d, vp]
    lib.orc_ctx_earth_orientation.argtypes = [vp, d, vp, vp]
    lib.orc_body_axes.argtypes = [vp, i, d, vp]
    lib.orc_geodetic_altitude.argtypes = [vp, vp, d, d]
    lib.orc_geodetic_altitude.restype = d
    lib.orc_exp_density.argtypes = [vp, d]
    lib.orc_exp_density.restype = d
    lib.orc_time_at.argtypes = [vp, d, vp]
    lib.orc_nrl_inputs.argtypes = [vp, d, vp, vp]
    lib.orc_solid_tide.argtypes = [vp, vp, vp, vp, vp, i, i, vp, vp]
    lib.orc_ocean_waves.argtypes = [vp, vp, d, vp, vp]
    lib.orc_tide_arguments.argtypes = [vp, vp, vp, vp]
    lib.orc_third_body.argtypes = [vp, vp, d, vp]
    lib.orc_shadow.argtypes = [vp, vp, d, d]
    lib.orc_shadow.restype = d
    lib.orc_sh_body_fixed.argtypes = [vp, vp, vp]
    lib.orc_ocean_tide.argtypes = [vp, vp, vp, vp]
    lib.orc_tide_acceleration.argtypes = [vp, vp, vp, vp, vp, i, vp]
    lib.orc_last_error.restype = ctypes.c_char_p
    lib.orc_destroy.argtypes = [vp]
    lib.orc_destroy.restype = None
    lib.orc_flop_counts.argtypes = [vp]
    lib.orc_flop_reset.argtypes = []
    lib.orc_flop_reset.restype = None
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    def __init__(self, config_path, counting=False):
        self.lib = flops_library() if counting else library()
        h = ctypes.c_void_p()
        if self.lib.orc_create(os.fsencode(config_path), os.fsencode(DATA_DIR), ctypes.byref(h)) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    __del__ = close

    def propagate(self, t, duration, threads=0):
        n = len(t["id"])
        arrs = [np.ascontiguousarray(t["id"], dtype=np.int64)] + [
            np.ascontiguousarray(t[k], dtype=np.float64) for k in ("epoch", "state", "mass", "drag_area", "srp_area", "cd", "cr")]
        oe, os_ = np.zeros(n), np.zeros((6, n))
        st = OrcStats()
        rc = self.lib.orc_propagate(self.h, n, *[_p(a) for a in arrs], float(duration), int(threads), _p(oe), _p(os_),
                                    ctypes.byref(st))
        if rc != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return oe, os_, {k: getattr(st, k) for k, _ in OrcStats._fields_}

    def flop_counts(self):
        """Counts of the calling thread since flop_reset() (counting build only):
        arithmetic FLOPs and transcendental calls of the per-object force model
        and step, of the epoch-only inputs and of the NRLMSISE-00 density, and
        the entries into each category's scopes (NRLMSISE-00: evaluations)."""
        out = np.zeros(9)
        if self.lib.orc_flop_counts(_p(out)) != 0:
            raise RuntimeError("not the FLOP-counting build of the oracle")
        keys = ("per_object", "epoch", "nrlmsise")
        return {"flops": dict(zip(keys, out[:3])), "transcendental": dict(zip(keys, out[3:6])),
                "scope_entries": dict(zip(keys, out[6:]))}

    def flop_reset(self):
        self.lib.orc_flop_reset()

    def force(self, r, v, et, mass, area_km2, cr_eff, cd, which=-1):
        out = np.zeros(3)
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        rc = self.lib.orc_force_term(self.h, int(which), _p(r), _p(v), float(et), float(mass), float(area_km2),
                                     float(cr_eff), float(cd), _p(out))
        if rc != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return out

    def body_position(self, naif, et):
        p = np.zeros(3)
        if self.lib.orc_ctx_body_position(self.h, int(naif), float(et), _p(p)) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return p

    def earth_orientation(self, et):
        R, w = np.zeros(9), np.zeros(3)
        self.lib.orc_ctx_earth_orientation(self.h, float(et), _p(R), _p(w))
        return R.reshape(3, 3), w

    def body_axes(self, naif, et):
        """ICRF -> axes of the configured field of body `naif`; (R, rotation target)."""
        R = np.zeros(9)
        target = self.lib.orc_body_axes(self.h, int(naif), float(et), _p(R))
        if target < 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return R.reshape(3, 3), target

    def exp_density(self, h):
        return self.lib.orc_exp_density(self.h, float(h))

    def time_at(self, et):
        """EOP time model at ET: dict(mjd_utc, ut1_utc, tt_utc, xp, yp, valid)."""
        o = np.zeros(6)
        self.lib.orc_time_at(self.h, float(et), _p(o))
        return dict(zip(("mjd_utc", "ut1_utc", "tt_utc", "xp", "yp", "valid"), o))

    def nrl_inputs(self, et, fixed):
        """NRLMSISE-00 inputs at ET for an Earth-fixed position (km)."""
        o = np.zeros(19)
        self.lib.orc_nrl_inputs(self.h, float(et), _p(np.ascontiguousarray(fixed, dtype=np.float64)), _p(o))
        keys = ("mjd_utc", "doy", "sec", "gast_h", "lst_h", "alt", "lat", "lon", "f107", "f107a", "ap_daily")
        out = dict(zip(keys, o[:11]))
        out["ap"] = o[11:18].copy()
        out["valid"] = bool(o[18] == 1.0)
        return out

    def solid_tide(self, moon, sun, E, time5, frequency_terms=True, zero_tide=True):
        C, S = np.zeros(49), np.zeros(49)
        a = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (moon, sun, E, time5)]
        self.lib.orc_solid_tide(self.h, *[_p(x) for x in a], int(frequency_terms), int(zero_tide), _p(C), _p(S))
        return C.reshape(7, 7), S.reshape(7, 7)

    def ocean_waves(self, args5, gmst):
        C, S = np.zeros(49), np.zeros(49)
        self.lib.orc_ocean_waves(self.h, _p(np.ascontiguousarray(args5, dtype=np.float64)), float(gmst), _p(C), _p(S))
        return C.reshape(7, 7), S.reshape(7, 7)

    def ocean_tide(self, time5):
        """Whole ocean-tide coefficient set (waves plus the ocean pole tide)."""
        C, S = np.zeros(49), np.zeros(49)
        self.lib.orc_ocean_tide(self.h, _p(np.ascontiguousarray(time5, dtype=np.float64)), _p(C), _p(S))
        return C.reshape(7, 7), S.reshape(7, 7)

    def tide_acceleration(self, r, E, C, S, max_degree):
        """Acceleration (m/s^2, ICRF) of a degree 2..max_degree potential at r (m, ICRF)."""
        out = np.zeros(3)
        a = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (r, E, C, S)]
        if self.lib.orc_tide_acceleration(self.h, *[_p(x) for x in a], int(max_degree), _p(out)) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return out

    def sh_body_fixed(self, p):
        out = np.zeros(3)
        if self.lib.orc_sh_body_fixed(self.h, _p(np.ascontiguousarray(p, dtype=np.float64)), _p(out)) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        return out


def tide_arguments(time5):
    """Delaunay arguments (l, l', F, D, Omega), GMST and Doodson arguments for time
    data (mjdUtc, UT1-UTC, TT-UTC, xp, yp)."""
    a, g, b = np.zeros(5), np.zeros(1), np.zeros(6)
    library().orc_tide_arguments(_p(np.ascontiguousarray(time5, dtype=np.float64)), _p(a), _p(g), _p(b))
    return a, g[0], b


def third_body(r, b, gm):
    out = np.zeros(3)
    library().orc_third_body(_p(np.ascontiguousarray(r, dtype=np.float64)), _p(np.ascontiguousarray(b, dtype=np.float64)),
                             float(gm), _p(out))
    return out


def shadow(sun, body, r_sun, r_body):
    return library().orc_shadow(_p(np.ascontiguousarray(sun, dtype=np.float64)), _p(np.ascontiguousarray(body, dtype=np.float64)),
                                float(r_sun), float(r_body))


def geodetic_altitude(r, pole, req, flattening):
    r = np.ascontiguousarray(r, dtype=np.float64)
    pole = np.ascontiguousarray(pole, dtype=np.float64)
    return library().orc_geodetic_altitude(_p(r), _p(pole), float(req), float(flattening))


def nrlmsise(doy, sec, alt, lat, lon, lst, f107a, f107, ap, apa=None):
    lib = library()
    d, t = np.zeros(9), np.zeros(2)
    a = None if apa is None else np.ascontiguousarray(apa, dtype=np.float64)
    rc = lib.orc_nrlmsise(os.fsencode(DATA_DIR), int(doy), sec, alt, lat, lon, lst, f107a, f107, ap,
                          _p(a) if a is not None else None, 1 if a is not None else 0, _p(d), _p(t))
    if rc != 0:
        raise RuntimeError(lib.orc_last_error().decode())
    return d, t


def body_position(naif, et):
    p = np.zeros(3)
    rc = library().orc_body_position(int(naif), float(et), _p(p))
    return p if rc == 0 else None


def earth_orientation(et):
    R, w = np.zeros(9), np.zeros(3)
    library().orc_earth_orientation(float(et), _p(R), _p(w))
    return R.reshape(3, 3), w
