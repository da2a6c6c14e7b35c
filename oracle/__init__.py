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


_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liboracle.so not built: run `python oracle/build.py`")
        lib = ctypes.CDLL(LIB_PATH)
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
        lib.orc_geodetic_altitude.argtypes = [vp, vp, d, d]
        lib.orc_geodetic_altitude.restype = d
        lib.orc_exp_density.argtypes = [vp, d]
        lib.orc_exp_density.restype = d
        lib.orc_last_error.restype = ctypes.c_char_p
        lib.orc_destroy.argtypes = [vp]
        lib.orc_destroy.restype = None
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    def __init__(self, config_path):
        self.lib = library()
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

    def exp_density(self, h):
        return self.lib.orc_exp_density(self.h, float(h))


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
