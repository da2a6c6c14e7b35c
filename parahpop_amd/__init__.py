"""phx: MI355X-native batched fixed-step RKF78 orbit propagation.

Python binding of libphx.so (the C-ABI in include/phx.h) through ctypes.
The library is built in-tree by ``parahpop_amd/build.py``; importing this
module never falls back to another implementation: :func:`library` raises
when libphx.so is missing or cannot be loaded.
"""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PHX_LIB_PATH names another build of the library (the tables are read from its data/ folder)
LIB_PATH = os.environ.get("PHX_LIB_PATH") or os.path.join(HERE, "libphx.so")
REC_SIZE = 384   # doubles per epoch record (phx_common.h REC_SIZE)

PHX_OK, PHX_EINVAL, PHX_EDEVICE, PHX_ENOTDONE, PHX_ENOMEM, PHX_EAGAIN = 0, -1, -2, -3, -4, -5

EXPORTED = [
    "phx_create", "phx_propagate", "phx_propagate_noalloc", "phx_batch_create", "phx_batch_create_uneven", "phx_batch_autotune", "phx_batch_run", "phx_batch_done", "phx_batch_download",
    "phx_batch_state", "phx_batch_reset", "phx_batch_destroy", "phx_set_option", "phx_get_option", "phx_submit", "phx_poll",
    "phx_last_error", "phx_destroy", "phx_force_probe", "phx_epoch_record", "phx_nrlmsise_probe", "phx_version",
]


class PhxError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("%s (code %d)" % (message, code))
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [("object_steps", ctypes.c_double), ("kernel_seconds", ctypes.c_double),
                ("epoch_seconds", ctypes.c_double), ("wall_seconds", ctypes.c_double),
                ("epoch_records", ctypes.c_double), ("launches", ctypes.c_double),
                ("waves", ctypes.c_int), ("steps_per_launch", ctypes.c_int),
                ("kernels_per_step", ctypes.c_int), ("sh_groups_min", ctypes.c_int), ("sh_groups_max", ctypes.c_int),
                ("graph_replays", ctypes.c_double), ("graph_captures", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("id", ctypes.c_void_p), ("epoch_mjd2000_tdb", ctypes.c_void_p),
                ("state6_soa", ctypes.c_void_p), ("mass_kg", ctypes.c_void_p), ("drag_area_m2", ctypes.c_void_p),
                ("srp_area_m2", ctypes.c_void_p), ("cd", ctypes.c_void_p), ("cr", ctypes.c_void_p),
                ("duration_s", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("out_epoch", ctypes.c_void_p), ("out_state6_soa", ctypes.c_void_p),
                ("submit_time_s", ctypes.c_double), ("finish_time_s", ctypes.c_double)]


_lib = None


def library():
    """Load libphx.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libphx.so not built: run `python parahpop_amd/build.py` (expected %s)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp, sz, d, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int
    lib.phx_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    lib.phx_propagate.argtypes = [vp, sz] + [vp] * 8 + [d, vp, vp, vp]
    lib.phx_propagate_noalloc.argtypes = [vp, sz] + [vp] * 7 + [d, vp]
    lib.phx_batch_create.argtypes = [vp, sz] + [vp] * 8 + [d, ctypes.POINTER(vp)]
    lib.phx_batch_create_uneven.argtypes = [vp, sz] + [vp] * 9 + [ctypes.POINTER(vp)]
    lib.phx_batch_autotune.argtypes = [vp, i, vp]
    lib.phx_batch_run.argtypes = [vp, i, vp]
    lib.phx_batch_done.argtypes = [vp]
    lib.phx_batch_download.argtypes = [vp, vp, vp]
    lib.phx_batch_state.argtypes = [vp, vp, vp]
    lib.phx_batch_reset.argtypes = [vp]
    lib.phx_batch_destroy.argtypes = [vp]
    lib.phx_batch_destroy.restype = None
    lib.phx_set_option.argtypes = [vp, ctypes.c_char_p, d]
    lib.phx_get_option.argtypes = [vp, ctypes.c_char_p]
    lib.phx_get_option.restype = d
    lib.phx_last_error.argtypes = [vp]
    lib.phx_last_error.restype = ctypes.c_char_p
    lib.phx_destroy.argtypes = [vp]
    lib.phx_destroy.restype = None
    lib.phx_force_probe.argtypes = [vp, sz] + [vp] * 6 + [d, i, vp]
    lib.phx_epoch_record.argtypes = [vp, d, vp, sz]
    lib.phx_nrlmsise_probe.argtypes = [vp, sz, vp, i, vp]
    lib.phx_version.restype = ctypes.c_char_p
    lib.phx_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    lib.phx_poll.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(Result)]
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device context created from a CONFIG.json (owns device tables)."""

    def __init__(self, config_path, **options):
        self.lib = library()
        h = ctypes.c_void_p()
        rc = self.lib.phx_create(os.fsencode(config_path), ctypes.byref(h))
        if rc != PHX_OK:
            raise PhxError(rc, self.lib.phx_last_error(None).decode())
        self.h = h
        self._batches = weakref.WeakSet()
        for k, v in options.items():
            self.set_option(k, v)

    def close(self):
        if getattr(self, "h", None):
            # batches hold the context: release them first
            for b in list(getattr(self, "_batches", ())):
                b.close()
            self.lib.phx_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def error(self):
        return self.lib.phx_last_error(self.h).decode()

    def _check(self, rc):
        if rc != PHX_OK:
            raise PhxError(rc, self.error())

    def set_option(self, key, value):
        self._check(self.lib.phx_set_option(self.h, key.encode(), float(value)))

    def option(self, key):
        return self.lib.phx_get_option(self.h, key.encode())

    @staticmethod
    def _inputs(t):
        n = int(t["n"]) if "n" in t else len(t["id"])
        arrs = [np.ascontiguousarray(t["id"], dtype=np.int64), _f64(t["epoch"]), _f64(t["state"]).reshape(6, n),
                _f64(t["mass"]), _f64(t["drag_area"]), _f64(t["srp_area"]), _f64(t["cd"]), _f64(t["cr"])]
        arrs[2] = np.ascontiguousarray(arrs[2])
        return n, arrs

    def propagate(self, targets, duration):
        """Propagate a target dict (see synth.load_targets); returns (epochs, states[6][n], stats)."""
        n, arrs = self._inputs(targets)
        out_e = np.zeros(n)
        out_s = np.zeros((6, n))
        st = Stats()
        rc = self.lib.phx_propagate(self.h, n, *[_ptr(a) for a in arrs], float(duration), _ptr(out_e), _ptr(out_s),
                                    ctypes.byref(st))
        self._check(rc)
        return out_e, out_s, st.as_dict()

    def propagate_noalloc(self, n, d_epoch, d_state, d_mass, d_drag_area, d_srp_area, d_cd, d_cr, duration):
        """phx_propagate_noalloc on device buffers (integer device addresses; the
        epoch and state buffers are overwritten); returns the stats."""
        st = Stats()
        ptrs = [ctypes.c_void_p(int(p)) for p in (d_epoch, d_state, d_mass, d_drag_area, d_srp_area, d_cd, d_cr)]
        self._check(self.lib.phx_propagate_noalloc(self.h, int(n), *ptrs, float(duration), ctypes.byref(st)))
        return st.as_dict()

    def batch(self, targets, duration):
        """Device-resident batch; duration is a scalar or one value per object."""
        return Batch(self, targets, duration)

    def submit(self, targets, duration):
        """Queue a propagation request on the context's server; returns a ticket."""
        n, arrs = self._inputs(targets)
        req = Request(n, *[_ptr(a) for a in arrs], float(duration))
        ticket = ctypes.c_uint64()
        self._check(self.lib.phx_submit(self.h, ctypes.byref(req), ctypes.byref(ticket)))
        self._pending = getattr(self, "_pending", {})
        self._pending[ticket.value] = n
        return ticket.value

    def poll(self, ticket):
        """None while the request is queued or running; else (epochs, states[6][n], submit_s, finish_s)."""
        n = self._pending[ticket]
        e = np.zeros(n)
        s = np.zeros((6, n))
        res = Result(n, _ptr(e), _ptr(s), 0.0, 0.0)
        rc = self.lib.phx_poll(self.h, ctypes.c_uint64(ticket), ctypes.byref(res))
        if rc == PHX_EAGAIN:
            return None
        del self._pending[ticket]
        self._check(rc)
        return e, s, res.submit_time_s, res.finish_time_s

    def wait(self, ticket, timeout=600.0, interval=0.002):
        """Poll until the request is done (raises TimeoutError)."""
        import time
        t_end = time.monotonic() + timeout
        while True:
            r = self.poll(ticket)
            if r is not None:
                return r
            if time.monotonic() > t_end:
                raise TimeoutError("request %d not done after %.0f s" % (ticket, timeout))
            time.sleep(interval)

    def force(self, pos, vel, mass, area_km2, cr_eff, cd, et, which=-1):
        pos, vel = _f64(pos).reshape(3, -1), _f64(vel).reshape(3, -1)
        n = pos.shape[1]
        acc = np.zeros((3, n))
        rc = self.lib.phx_force_probe(self.h, n, _ptr(np.ascontiguousarray(pos)), _ptr(np.ascontiguousarray(vel)),
                                      _ptr(_f64(mass)), _ptr(_f64(area_km2)), _ptr(_f64(cr_eff)), _ptr(_f64(cd)),
                                      float(et), int(which), _ptr(acc))
        self._check(rc)
        return acc

    def epoch_record(self, et):
        rec = np.zeros(REC_SIZE)
        self._check(self.lib.phx_epoch_record(self.h, float(et), _ptr(rec), REC_SIZE))
        return rec

    def nrlmsise(self, rows, use_ap_array=False):
        rows = _f64(rows).reshape(-1, 16)
        out = np.zeros((rows.shape[0], 11))
        self._check(self.lib.phx_nrlmsise_probe(self.h, rows.shape[0], _ptr(rows), int(use_ap_array), _ptr(out)))
        return out


class Batch:
    """Device-resident batch: inputs uploaded once, run (and re-run) on the device."""

    def __init__(self, ctx, targets, duration):
        self.ctx = ctx
        n, arrs = ctx._inputs(targets)
        self.n = n
        self._keep = arrs
        h = ctypes.c_void_p()
        if np.ndim(duration) == 0:
            rc = ctx.lib.phx_batch_create(ctx.h, n, *[_ptr(a) for a in arrs], float(duration), ctypes.byref(h))
        else:
            self._dur = _f64(duration)
            assert self._dur.shape == (n,)
            rc = ctx.lib.phx_batch_create_uneven(ctx.h, n, *[_ptr(a) for a in arrs], _ptr(self._dur), ctypes.byref(h))
        ctx._check(rc)
        self.h = h
        ctx._batches.add(self)

    def run(self, max_steps=0):
        st = Stats()
        self.ctx._check(self.ctx.lib.phx_batch_run(self.h, int(max_steps), ctypes.byref(st)))
        return st.as_dict()

    def done(self):
        return bool(self.ctx.lib.phx_batch_done(self.h))

    def autotune(self, trial_steps=4):
        """Pick the fastest tile shape for this batch (sets the context's options)."""
        st = Stats()
        self.ctx._check(self.ctx.lib.phx_batch_autotune(self.h, int(trial_steps), ctypes.byref(st)))
        return st.as_dict()

    def reset(self):
        self.ctx._check(self.ctx.lib.phx_batch_reset(self.h))

    def download(self):
        e = np.zeros(self.n)
        s = np.zeros((6, self.n))
        self.ctx._check(self.ctx.lib.phx_batch_download(self.h, _ptr(e), _ptr(s)))
        return e, s

    def state(self):
        """Current (epochs, states[6][n]) of every object, finished or not."""
        e = np.zeros(self.n)
        s = np.zeros((6, self.n))
        self.ctx._check(self.ctx.lib.phx_batch_state(self.h, _ptr(e), _ptr(s)))
        return e, s

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                self.ctx.lib.phx_batch_destroy(self.h)
            self.h = None

    __del__ = close
