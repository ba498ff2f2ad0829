"""ctypes binding of the engine's C-ABI (include/j1j2_engine.h).

Used by the tests and the benchmark.  There is no fallback: if
libj1j2_engine.so is missing or a call fails, an exception is raised.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libj1j2_engine.so")

SAMPLER_PT, SAMPLER_LOCAL = 0, 1
FP64_PLAIN, FP32_BATCHED = 0, 1
LADDER_LINEAR, LADDER_SPLIT, LADDER_LOCAL = 0, 1, 2
PHASE_EQUIL, PHASE_MEAS = 0, 1
FLAG_WRITELOCNEM, FLAG_WRITEHIST, FLAG_WRITENUMFLIPS = 1, 2, 4
FLAG_SHADOW_FP64, FLAG_NO_GRAPH, FLAG_FORCE_GLOBAL = 8, 16, 32
START_HOT, START_FER, START_STR = 0, 1, 2
PATH_NAMES = {0: "fp64_plain", 1: "fp32_resident", 2: "fp32_global", 3: "fp32_resident_bits"}

# every function declared in include/j1j2_engine.h
EXPORTS = [
    "j1j2_ladder", "j1j2_make_disorder", "j1j2_make_disorder_batch", "j1j2_initial_spins", "j1j2_create", "j1j2_create_batch",
    "j1j2_destroy", "j1j2_last_error", "j1j2_num_configs", "j1j2_config_shape",
    "j1j2_upload_disorder", "j1j2_set_spins", "j1j2_set_temperatures", "j1j2_run",
    "j1j2_run_cycles", "j1j2_fetch", "j1j2_get_spins", "j1j2_get_maps", "j1j2_get_observables",
    "j1j2_get_sums", "j1j2_get_series", "j1j2_get_flips", "j1j2_get_mismatch", "j1j2_set_timing",
    "j1j2_get_timing", "j1j2_sweep_path", "j1j2_launch_plan", "j1j2_synchronize",
    "j1j2_run_cycles_noalloc",
]
NOALLOC_MAX_LATTICES = 16384


class Params(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int32), ("H", ctypes.c_int32), ("nreplica", ctypes.c_int32),
                ("ptswapever", ctypes.c_int32), ("ptmeasindex", ctypes.c_int32),
                ("neq", ctypes.c_int32), ("nmess", ctypes.c_int32), ("nconf", ctypes.c_int32),
                ("iconf0", ctypes.c_int32), ("bins", ctypes.c_int32), ("sampler", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("ladder", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("J1", ctypes.c_double), ("deltaJ1", ctypes.c_double), ("J2", ctypes.c_double),
                ("J3", ctypes.c_double), ("tmin", ctypes.c_double), ("tmax", ctypes.c_double),
                ("dt", ctypes.c_double), ("tladdermid", ctypes.c_double),
                ("tladderfrac", ctypes.c_double), ("seed", ctypes.c_uint64)]


def make_params(L, H=None, nreplica=1, ptswapever=5, ptmeasindex=1, neq=0, nmess=0, nconf=1,
                iconf0=1, bins=100, sampler=SAMPLER_PT, precision=FP32_BATCHED,
                ladder=LADDER_LINEAR, flags=0, J1=1.0, deltaJ1=0.0, J2=-1.0, J3=1.0, tmin=8.0,
                tmax=11.5, dt=1.0, tladdermid=9.5, tladderfrac=0.7, seed=12345):
    return Params(L, L if H is None else H, nreplica, ptswapever, ptmeasindex, neq, nmess, nconf,
                  iconf0, bins, sampler, precision, ladder, flags, J1, deltaJ1, J2, J3, tmin, tmax,
                  dt, tladdermid, tladderfrac, seed)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"engine library missing: {LIB_PATH} (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "j1j2_ladder": [ctypes.POINTER(Params), P],
        "j1j2_make_disorder": [i32, i32, d, i32, i32, d, P],
        "j1j2_make_disorder_batch": [i32, i32, d, i32, i32, i32, d, i32, P],
        "j1j2_initial_spins": [i32, i32, i32, i32, i32, i64, P],
        "j1j2_create": [ctypes.POINTER(Params), i32, ctypes.POINTER(P)],
        "j1j2_create_batch": [P, i32, i32, ctypes.POINTER(P)],
        "j1j2_last_error": [ctypes.c_char_p, i32],
        "j1j2_num_configs": [P, P],
        "j1j2_config_shape": [P, i32, P, P, P],
        "j1j2_upload_disorder": [P, i32, P, i64],
        "j1j2_set_spins": [P, i32, P, i32],
        "j1j2_set_temperatures": [P, i32, P, i32],
        "j1j2_run": [P, i32, i32],
        "j1j2_run_cycles": [P, i32, i32, i32],
        "j1j2_fetch": [P, i32, P, P, P, P, P, P, P, P, P, P],
        "j1j2_get_spins": [P, i32, P, i32],
        "j1j2_get_maps": [P, i32, P, P],
        "j1j2_get_observables": [P, i32, P],
        "j1j2_get_sums": [P, i32, P, P, P, P],
        "j1j2_get_series": [P, i32, P, i64, P],
        "j1j2_get_flips": [P, i32, P],
        "j1j2_get_mismatch": [P, i32, P, P],
        "j1j2_set_timing": [P, i32],
        "j1j2_get_timing": [P, i32, P, P],
        "j1j2_sweep_path": [P, P, P],
        "j1j2_launch_plan": [P, ctypes.c_char_p, i32],
        "j1j2_synchronize": [P],
        "j1j2_run_cycles_noalloc": [P, i32, i32, i32, P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.j1j2_destroy.argtypes = [P]
    L.j1j2_destroy.restype = None


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"j1j2 error {code}: {msg}")
        self.code = code


def last_error():
    buf = ctypes.create_string_buffer(512)
    lib().j1j2_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != 0:
        raise EngineError(rc, last_error())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- host helpers
def ladder(params):
    t = np.zeros(params.nreplica, np.float64)
    check(lib().j1j2_ladder(ctypes.byref(params), _p(t)))
    return t


def make_disorder(L, H, xi=3.0, seed=1, map_to_uniform=True, phi=2.0):
    out = np.zeros(L * L * H, np.float64)
    check(lib().j1j2_make_disorder(L, H, xi, seed, 1 if map_to_uniform else 0, phi, _p(out)))
    return out


def make_disorder_batch(L, H, xi=3.0, seed0=1, nconf=1, map_to_uniform=True, phi=2.0, nthreads=0):
    out = np.zeros((nconf, L * L * H), np.float64)
    check(lib().j1j2_make_disorder_batch(L, H, xi, seed0, nconf, 1 if map_to_uniform else 0, phi, nthreads,
                                         _p(out)))
    return out


def initial_spins(L, H, start, straxis=0, seed=1, skip=0):
    s = np.zeros(L * L * H, np.int8)
    check(lib().j1j2_initial_spins(L, H, start, straxis, seed, skip, _p(s)))
    return s


# ---------------------------------------------------------------- engine
class Engine:
    def __init__(self, groups, device=0):
        if isinstance(groups, Params):
            groups = [groups]
        arr = (Params * len(groups))(*groups)
        self._h = ctypes.c_void_p()
        check(lib().j1j2_create_batch(ctypes.cast(arr, ctypes.c_void_p), len(groups), device,
                                      ctypes.byref(self._h)))
        self.groups = list(groups)
        n = ctypes.c_int32()
        check(lib().j1j2_num_configs(self._h, ctypes.byref(n)))
        self.nconf = n.value

    def close(self):
        if self._h:
            lib().j1j2_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, conf):
        L, H, R = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().j1j2_config_shape(self._h, conf, ctypes.byref(L), ctypes.byref(H), ctypes.byref(R)))
        return L.value, H.value, R.value

    def upload_disorder(self, conf, p):
        p = np.ascontiguousarray(p, np.float64)
        check(lib().j1j2_upload_disorder(self._h, conf, _p(p), p.size))

    def set_spins(self, conf, spins):
        s = np.ascontiguousarray(spins, np.int8)
        L, H, R = self.shape(conf)
        check(lib().j1j2_set_spins(self._h, conf, _p(s), R))

    def set_temperatures(self, conf, temps):
        t = np.ascontiguousarray(temps, np.float64)
        check(lib().j1j2_set_temperatures(self._h, conf, _p(t), t.size))

    def run(self, phase, ncycles=-1):
        check(lib().j1j2_run(self._h, phase, ncycles))

    def run_cycles(self, phase, ncycles, begin=False):
        check(lib().j1j2_run_cycles(self._h, phase, ncycles, 1 if begin else 0))

    def run_cycles_noalloc(self, phase, ncycles, begin=False, stream=None):
        """Enqueue without allocating or waiting; synchronize() before reading."""
        check(lib().j1j2_run_cycles_noalloc(self._h, phase, ncycles, 1 if begin else 0, stream))

    def spins(self, conf):
        L, H, R = self.shape(conf)
        s = np.zeros(R * L * L * H, np.int8)
        check(lib().j1j2_get_spins(self._h, conf, _p(s), R))
        return s.reshape(R, L * L * H)

    def maps(self, conf):
        R = self.shape(conf)[2]
        a, b = np.zeros(R, np.int32), np.zeros(R, np.int32)
        check(lib().j1j2_get_maps(self._h, conf, _p(a), _p(b)))
        return a, b

    def observables(self, conf):
        R = self.shape(conf)[2]
        o = np.zeros(5 * R, np.float64)
        check(lib().j1j2_get_observables(self._h, conf, _p(o)))
        return o.reshape(5, R)

    def sums(self, conf):
        R = self.shape(conf)[2]
        s = np.zeros(R * 20, np.float64)
        nm = np.zeros(1, np.int64)
        best, fin = np.zeros(R), np.zeros(R)
        check(lib().j1j2_get_sums(self._h, conf, _p(s), _p(nm), _p(best), _p(fin)))
        return s.reshape(R, 20), int(nm[0]), best, fin

    def series(self, conf, cap):
        R = self.shape(conf)[2]
        n = np.zeros(1, np.int64)
        check(lib().j1j2_get_series(self._h, conf, None, 0, _p(n)))
        buf = np.zeros(R * 5 * max(int(n[0]), 1), np.float64)
        check(lib().j1j2_get_series(self._h, conf, _p(buf), int(n[0]), _p(n)))
        return buf[: R * 5 * int(n[0])].reshape(R, 5, int(n[0]))

    def fetch(self, conf):
        L, H, R = self.shape(conf)
        N = L * L * H
        bins = self.groups[0].bins
        out = dict(basefields=np.zeros(49 * R), hist=np.zeros(R * bins), hist_min=np.zeros(R),
                   hist_bin=np.zeros(R), swap_att=np.zeros(max(R - 1, 1), np.int32),
                   swap_acc=np.zeros(max(R - 1, 1), np.int32), visitboth=np.zeros(1, np.int32),
                   locspin=np.zeros(N), locnem=np.zeros(N), locnem_samples=np.zeros(1, np.int32))
        check(lib().j1j2_fetch(self._h, conf, *[_p(out[k]) for k in
                                                ("basefields", "hist", "hist_min", "hist_bin",
                                                 "swap_att", "swap_acc", "visitboth", "locspin",
                                                 "locnem", "locnem_samples")]))
        out["basefields"] = out["basefields"].reshape(49, R)
        out["hist"] = out["hist"].reshape(bins, R)
        out["swap_att"] = out["swap_att"][: R - 1]
        out["swap_acc"] = out["swap_acc"][: R - 1]
        out["visitboth"] = int(out["visitboth"][0])
        out["locnem_samples"] = int(out["locnem_samples"][0])
        return out

    def flips(self, conf):
        L, H, R = self.shape(conf)
        f = np.zeros(R * L * L * H, np.int32)
        check(lib().j1j2_get_flips(self._h, conf, _p(f)))
        return f.reshape(R, L * L * H)

    def mismatch(self, conf):
        a, b = np.zeros(1, np.int64), np.zeros(1, np.int64)
        check(lib().j1j2_get_mismatch(self._h, conf, _p(a), _p(b)))
        return int(a[0]), int(b[0])

    def set_timing(self, on):
        check(lib().j1j2_set_timing(self._h, 1 if on else 0))

    def timing(self, kind):
        ms, n = np.zeros(1), np.zeros(1, np.int64)
        check(lib().j1j2_get_timing(self._h, kind, _p(ms), _p(n)))
        return float(ms[0]), int(n[0])

    def sweep_path(self):
        p, b = np.zeros(1, np.int32), np.zeros(1)
        check(lib().j1j2_sweep_path(self._h, _p(p), _p(b)))
        return int(p[0]), float(b[0])

    def launch_plan(self):
        buf = ctypes.create_string_buffer(512)
        check(lib().j1j2_launch_plan(self._h, buf, 512))
        return buf.value.decode()

    def synchronize(self):
        check(lib().j1j2_synchronize(self._h))
