"""ctypes binding of the CPU oracle (oracle/j1j2_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline -- never by the engine.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# J1J2_ORACLE_LIB names another build of the same sources (the sanitizer test's)
LIB_PATH = os.environ.get("J1J2_ORACLE_LIB") or os.path.join(_HERE, "_build", "libj1j2_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_i32, _u32, _i64, _u64, _d, _f = (ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64,
                                  ctypes.c_uint64, ctypes.c_double, ctypes.c_float)


class Kiss(ctypes.Structure):
    _fields_ = [("x", _i32), ("y", _i32), ("z", _i32), ("w", _i32), ("iset", _i32), ("gset", _d)]


class Lattice(ctypes.Structure):
    _fields_ = [("L", _i32), ("H", _i32), ("N", _i32), ("J1h", _d), ("J1v", _d), ("J2", _d),
                ("J3", _d), ("p", _P), ("kc", _P), ("k2xp", _P), ("k2xm", _P), ("k2yp", _P),
                ("k2ym", _P), ("j2x2", _f), ("j3x2", _f), ("q", _P)]


NSUMS = 20


class PT(ctypes.Structure):
    _fields_ = [("lat", Lattice), ("nrep", _i32), ("ptswapever", _i32), ("ptmeasindex", _i32),
                ("bins", _i32), ("precision", _i32), ("local", _i32), ("seed", _u32),
                ("iconf", _u32), ("spins", _P), ("rep_at_temp", _P), ("temp_of_rep", _P),
                ("beta", _P), ("cycle", _u64), ("sweep", _u64), ("swap_phase", _i32),
                ("att", _P), ("acc", _P), ("sums", _P), ("best", _P), ("final_en", _P),
                ("hist", _P), ("hmin", _P), ("hmax", _P), ("hbin", _P), ("nmeasure", _i64),
                ("vcold", _P), ("vhot", _P), ("locspin", _P), ("locnem", _P),
                ("locsamples", _i64), ("flips", _P), ("shadow", _i64 * 2), ("do_shadow", ctypes.c_int),
                ("obs", _P), ("etot", _P), ("series", _P), ("series_len", _i64),
                ("series_cap", _i64)]


def _declare(L):
    L.ora_kiss_init.argtypes = [ctypes.POINTER(Kiss), _i32]
    L.ora_kiss_uniform.argtypes = [ctypes.POINTER(Kiss)]
    L.ora_kiss_uniform.restype = _d
    L.ora_kiss_gauss.argtypes = [ctypes.POINTER(Kiss)]
    L.ora_kiss_gauss.restype = _d
    L.ora_kiss_fill.argtypes = [_i32, _i64, _P]
    L.ora_philox4x32_10.argtypes = [_P, _P, _P]
    L.ora_call_words.argtypes = [ctypes.c_int32, _u32, _u32, _u32, ctypes.c_uint64, ctypes.c_int32,
                                 ctypes.c_int32, _P, _P]
    L.ora_call_words.restype = ctypes.c_int
    L.ora_accept_threshold_f32.argtypes = [_f]
    L.ora_accept_threshold_f32.restype = _f
    L.ora_exp64.argtypes = [_d]
    L.ora_exp64.restype = _d
    L.ora_lattice_init.argtypes = [ctypes.POINTER(Lattice), _i32, _i32, _d, _d, _d, _d, _P]
    L.ora_lattice_free.argtypes = [ctypes.POINTER(Lattice)]
    L.ora_field_f64.argtypes = [ctypes.POINTER(Lattice), _P, _i32, _P]
    L.ora_field_f64.restype = _d
    L.ora_field_f32.argtypes = [ctypes.POINTER(Lattice), _P, _i32]
    L.ora_field_f32.restype = _f
    L.ora_colour.argtypes = [_i32, _i32]
    L.ora_colour_index.argtypes = [_i32, _i32]
    L.ora_sweep.argtypes = [ctypes.POINTER(Lattice), _P, ctypes.c_int, _d, _u32, _u32, _u32,
                            _u64, _P, _P]
    L.ora_observables.argtypes = [ctypes.POINTER(Lattice), _P, _P]
    L.ora_energy_total.argtypes = [ctypes.POINTER(Lattice), _P]
    L.ora_energy_total.restype = _d
    L.ora_measurement_ref.argtypes = [ctypes.POINTER(Lattice), _P, _P]
    L.ora_nbar_ref.argtypes = [ctypes.POINTER(Lattice), _P]
    L.ora_nbar_ref.restype = _d
    L.ora_initial_spins.argtypes = [_i32, _i32, _i32, _i32, _i32, _P]
    L.ora_ladder.argtypes = [_i32, _i32, _d, _d, _d, _d, _d, _P]
    L.ora_pt_init.argtypes = [ctypes.POINTER(PT), _i32, _i32, _d, _d, _d, _d, _P, _i32, _P, _i32,
                              _i32, _i32, _i32, _i32, _u32, _u32, _P, _i64, _i32, _i32]
    L.ora_pt_free.argtypes = [ctypes.POINTER(PT)]
    L.ora_pt_begin_phase.argtypes = [ctypes.POINTER(PT)]
    L.ora_pt_reset_measurement.argtypes = [ctypes.POINTER(PT)]
    L.ora_pt_cycle.argtypes = [ctypes.POINTER(PT), ctypes.c_int]
    L.ora_pt_basefields.argtypes = [ctypes.POINTER(PT), _P]
    L.ora_pt_hist.argtypes = [ctypes.POINTER(PT), _P]
    L.ora_derive.argtypes = [_P, _i64, _d, _i32, _d, _d, _P]
    L.ora_process_local.argtypes = [_P, _P, _P, _P, _P, _i64, _d, _i32, _i32, _P, _P, _P, _P]
    L.ora_enumerate.argtypes = [ctypes.POINTER(Lattice), _d, _P]
    L.ora_corr_field.argtypes = [_i32, _i32, _i32, _d, _i32, _i32, _d, _P]
    L.ora_erfcc.argtypes = [_d]
    L.ora_erfcc.restype = _d


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- RNGs
def kiss_stream(seed, n):
    out = np.zeros(n, np.float64)
    lib().ora_kiss_fill(seed, n, _ptr(out))
    return out


def kiss_gauss_stream(seed, n):
    k = Kiss()
    lib().ora_kiss_init(ctypes.byref(k), seed)
    return np.array([lib().ora_kiss_gauss(ctypes.byref(k)) for _ in range(n)])


def philox(ctr, key):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().ora_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return o


def call_words(L, seed, iconf, replica, sweep, c, m):
    """Colour indices and 32-bit words of the sites of Philox call m."""
    qi = np.zeros(8, np.int32)
    r = np.zeros(8, np.uint32)
    n = lib().ora_call_words(L, seed, iconf, replica, sweep, c, m, _ptr(qi), _ptr(r))
    return qi[:n], r[:n]


def accept_threshold_f32(y):
    return lib().ora_accept_threshold_f32(float(y))


def exp64(x):
    return lib().ora_exp64(float(x))


# ---------------------------------------------------------------- lattice
class OracleLattice:
    def __init__(self, L, H, p, J1=1.0, deltaJ1=0.0, J2=-1.0, J3=1.0):
        self.p = np.ascontiguousarray(p, np.float64)
        self.L, self.H, self.N = L, H, L * L * H
        assert self.p.size == self.N
        self.s = Lattice()
        rc = lib().ora_lattice_init(ctypes.byref(self.s), L, H, J1, deltaJ1, J2, J3, _ptr(self.p))
        if rc != 0:
            raise ValueError("invalid lattice")

    def __del__(self):
        try:
            lib().ora_lattice_free(ctypes.byref(self.s))
        except Exception:
            pass

    def field_f64(self, spins, site):
        s = np.ascontiguousarray(spins, np.int8)
        return lib().ora_field_f64(ctypes.byref(self.s), _ptr(s), site, None)

    def field_f32(self, spins, site):
        s = np.ascontiguousarray(spins, np.int8)
        return lib().ora_field_f32(ctypes.byref(self.s), _ptr(s), site)

    def sweep(self, spins, precision, beta, seed, iconf, replica, sweep, flips=None, shadow=None):
        assert spins.dtype == np.int8 and spins.flags.c_contiguous
        fp = _ptr(flips) if flips is not None else None
        sp = _ptr(shadow) if shadow is not None else None
        lib().ora_sweep(ctypes.byref(self.s), _ptr(spins), precision, beta, seed, iconf,
                        replica, sweep, fp, sp)

    def observables(self, spins):
        s = np.ascontiguousarray(spins, np.int8)
        o = np.zeros(8, np.int64)
        lib().ora_observables(ctypes.byref(self.s), _ptr(s), _ptr(o))
        return o

    def energy_total(self, obs):
        o = np.ascontiguousarray(obs, np.int64)
        return lib().ora_energy_total(ctypes.byref(self.s), _ptr(o))

    def measurement_ref(self, spins):
        s = np.ascontiguousarray(spins, np.int8)
        o = np.zeros(4, np.float64)
        lib().ora_measurement_ref(ctypes.byref(self.s), _ptr(s), _ptr(o))
        return o

    def nbar_ref(self, spins):
        s = np.ascontiguousarray(spins, np.int8)
        return lib().ora_nbar_ref(ctypes.byref(self.s), _ptr(s))

    def enumerate(self, beta):
        o = np.zeros(8, np.float64)
        if lib().ora_enumerate(ctypes.byref(self.s), beta, _ptr(o)) != 0:
            raise ValueError("lattice too large to enumerate")
        return o


def colour(L, site):
    return lib().ora_colour(L, site)


def initial_spins(L, H, start, straxis, seed):
    s = np.zeros(L * L * H, np.int8)
    lib().ora_initial_spins(L, H, start, straxis, seed, _ptr(s))
    return s


def ladder(mode, ntemp, tmin, tmax, dt=0.0, tmid=0.0, tfrac=0.0):
    t = np.zeros(ntemp, np.float64)
    lib().ora_ladder(mode, ntemp, tmin, tmax, dt, tmid, tfrac, _ptr(t))
    return t


def corr_field(n1, n2, n3, xi, seed, do_uniform=True, max_value=2.0):
    out = np.zeros(n1 * n2 * n3, np.float64)
    if lib().ora_corr_field(n1, n2, n3, xi, seed, 1 if do_uniform else 0, max_value, _ptr(out)):
        raise MemoryError("corr_field")
    return out


def erfcc(x):
    return lib().ora_erfcc(float(x))


def derive(sums, nmeasure, beta, N, best, final):
    s = np.ascontiguousarray(sums, np.float64)
    o = np.zeros(49, np.float64)
    lib().ora_derive(_ptr(s), nmeasure, beta, N, best, final, _ptr(o))
    return o


def process_local(mag, en, strx, stry, nbar, beta, N, bins):
    arrs = [np.ascontiguousarray(a, np.float64) for a in (mag, en, strx, stry, nbar)]
    o = np.zeros(49, np.float64)
    h = np.zeros(bins, np.float64)
    hmin = np.zeros(1)
    hbin = np.zeros(1)
    lib().ora_process_local(*[_ptr(a) for a in arrs], len(arrs[0]), beta, N, bins, _ptr(o),
                            _ptr(h), _ptr(hmin), _ptr(hbin))
    return o, h, hmin[0], hbin[0]


# ---------------------------------------------------------------- PT engine
class OraclePT:
    """Restatement of one disorder config's PT ladder under the engine's rule."""

    def __init__(self, L, H, p, temps, spins0, ptswapever=5, ptmeasindex=1, bins=100,
                 precision=1, local=False, seed=12345, iconf=1, J1=1.0, deltaJ1=0.0,
                 J2=-1.0, J3=1.0, series_cap=0, count_flips=False, shadow=False):
        self.N = L * L * H
        self.nrep = len(temps)
        self.bins = bins
        self._p = np.ascontiguousarray(p, np.float64)
        self._t = np.ascontiguousarray(temps, np.float64)
        s0 = np.ascontiguousarray(spins0, np.int8).reshape(-1)
        assert s0.size == self.nrep * self.N
        self.s = PT()
        rc = lib().ora_pt_init(ctypes.byref(self.s), L, H, J1, deltaJ1, J2, J3, _ptr(self._p),
                               self.nrep, _ptr(self._t), ptswapever, ptmeasindex, bins,
                               precision, 1 if local else 0, seed, iconf, _ptr(s0), series_cap,
                               1 if count_flips else 0, 1 if shadow else 0)
        if rc != 0:
            raise ValueError("invalid PT setup")

    def __del__(self):
        try:
            lib().ora_pt_free(ctypes.byref(self.s))
        except Exception:
            pass

    def _arr(self, ptr, dtype, n):
        ct = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(n,)).copy()

    def begin_phase(self):
        lib().ora_pt_begin_phase(ctypes.byref(self.s))

    def reset_measurement(self):
        lib().ora_pt_reset_measurement(ctypes.byref(self.s))

    def set_spins(self, spins):
        """Overwrite every replica's spins (a fresh start; the sweep clock runs on)."""
        s = np.ascontiguousarray(spins, np.int8).reshape(-1)
        assert s.size == self.nrep * self.N
        ctypes.memmove(self.s.spins, s.ctypes.data, s.size)

    def set_temperatures(self, temps):
        """New temperature per slot; the chain carries on from its current state."""
        b = 1.0 / np.ascontiguousarray(temps, np.float64)
        assert b.size == self.nrep
        ctypes.memmove(self.s.beta, b.ctypes.data, b.nbytes)

    def cycle(self, measure):
        lib().ora_pt_cycle(ctypes.byref(self.s), 1 if measure else 0)

    def run(self, ncycles, measure):
        self.begin_phase()
        for _ in range(ncycles):
            self.cycle(measure)

    @property
    def spins(self):
        return self._arr(self.s.spins, np.int8, self.nrep * self.N).reshape(self.nrep, self.N)

    @property
    def rep_at_temp(self):
        return self._arr(self.s.rep_at_temp, np.int32, self.nrep)

    @property
    def temp_of_rep(self):
        return self._arr(self.s.temp_of_rep, np.int32, self.nrep)

    @property
    def obs(self):
        return self._arr(self.s.obs, np.int64, self.nrep * 8).reshape(self.nrep, 8)

    @property
    def etot(self):
        return self._arr(self.s.etot, np.float64, self.nrep)

    @property
    def attempts(self):
        return self._arr(self.s.att, np.int64, max(1, self.nrep - 1))

    @property
    def accepts(self):
        return self._arr(self.s.acc, np.int64, max(1, self.nrep - 1))

    @property
    def sums(self):
        return self._arr(self.s.sums, np.float64, self.nrep * NSUMS).reshape(self.nrep, NSUMS)

    @property
    def best(self):
        return self._arr(self.s.best, np.float64, self.nrep)

    @property
    def final_en(self):
        return self._arr(self.s.final_en, np.float64, self.nrep)

    @property
    def hist_counts(self):
        return self._arr(self.s.hist, np.int64, self.nrep * self.bins).reshape(self.nrep, self.bins)

    @property
    def hist_min(self):
        return self._arr(self.s.hmin, np.float64, self.nrep)

    @property
    def hist_bin(self):
        return self._arr(self.s.hbin, np.float64, self.nrep)

    @property
    def visitboth(self):
        c = self._arr(self.s.vcold, np.uint8, self.nrep)
        h = self._arr(self.s.vhot, np.uint8, self.nrep)
        return int(np.sum((c > 0) & (h > 0)))

    @property
    def locspin(self):
        return self._arr(self.s.locspin, np.int64, self.N)

    @property
    def locnem(self):
        return self._arr(self.s.locnem, np.int64, self.N)

    @property
    def flips(self):
        return self._arr(self.s.flips, np.int32, self.nrep * self.N).reshape(self.nrep, self.N)

    @property
    def shadow(self):
        return (int(self.s.shadow[0]), int(self.s.shadow[1]))

    @property
    def series(self):
        cap = self.s.series_cap
        a = self._arr(self.s.series, np.float64, self.nrep * 5 * cap).reshape(self.nrep, 5, cap)
        return a[:, :, : self.s.series_len]

    def basefields(self):
        o = np.zeros(49 * self.nrep, np.float64)
        lib().ora_pt_basefields(ctypes.byref(self.s), _ptr(o))
        return o.reshape(49, self.nrep)

    def hist(self):
        o = np.zeros(self.nrep * self.bins, np.float64)
        lib().ora_pt_hist(ctypes.byref(self.s), _ptr(o))
        return o.reshape(self.nrep, self.bins)
