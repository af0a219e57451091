"""ctypes binding of libspk_hip.so (include/spk_hip.h).

There is no fallback: if the library is missing or fails to load, every entry
point raises.  Build it with ``python -m spparks_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libspk_hip.so")

APPS = {"potts": 0, "ising": 1, "potts/neigh": 2, "potts/neighonly": 3, "ising/single": 4}
LATTICES = {"sq/4n": 0, "sq/8n": 1, "sc/6n": 2, "sc/26n": 3}
MAXNEIGH = {"sq/4n": 4, "sq/8n": 8, "sc/6n": 6, "sc/26n": 26}

# every symbol include/spk_hip.h declares
ABI_SYMBOLS = (
    "spk_hip_last_error", "spk_hip_ncolors", "spk_hip_metropolis_table",
    "spk_hip_master_draws",
    "spk_hip_create", "spk_hip_destroy", "spk_hip_set_temperature",
    "spk_hip_upload", "spk_hip_download", "spk_hip_set_range",
    "spk_hip_set_value", "spk_hip_init_seeds", "spk_hip_sweeps",
    "spk_hip_energy", "spk_hip_sync", "spk_hip_set_launch", "spk_hip_set_bitplanes", "spk_hip_set_resident",
    "spk_hip_last_timing",
    "spk_hip_batch_create", "spk_hip_batch_destroy", "spk_hip_replica_bytes",
    "spk_hip_batch_add", "spk_hip_batch_remove", "spk_hip_batch_upload",
    "spk_hip_batch_download", "spk_hip_batch_run", "spk_hip_batch_status",
    "spk_hip_batch_trace", "spk_hip_batch_set_launch", "spk_hip_batch_last_timing",
    "spk_hip_batch_pool", "spk_hip_checksum", "spk_hip_batch_checksum", "spk_hip_serve",
    "spk_hip_set_graph", "spk_hip_prepare", "spk_hip_sweeps_noalloc",
)


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class Desc(C.Structure):
    _fields_ = [("app", C.c_int), ("lattice", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("nz", C.c_int), ("nspins", C.c_int),
                ("temperature", C.c_double)]


_lib = None
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def load_library(path=None):
    """Load libspk_hip.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError("libspk_hip.so not built (%s); run python -m spparks_amd.build" % p)
    L = C.CDLL(p)
    vp = C.c_void_p
    L.spk_hip_last_error.restype = C.c_char_p
    L.spk_hip_ncolors.argtypes = [C.POINTER(Desc)]
    L.spk_hip_metropolis_table.argtypes = [C.c_double, C.c_int, vp, vp]
    L.spk_hip_master_draws.argtypes = [C.c_int, C.c_int, vp]
    L.spk_hip_create.argtypes = [C.POINTER(Desc), C.POINTER(vp)]
    L.spk_hip_destroy.argtypes = [vp]
    L.spk_hip_destroy.restype = None
    L.spk_hip_set_temperature.argtypes = [vp, C.c_double]
    L.spk_hip_upload.argtypes = [vp, vp, vp]
    L.spk_hip_download.argtypes = [vp, vp, vp]
    L.spk_hip_set_range.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.spk_hip_set_value.argtypes = [vp, C.c_int]
    L.spk_hip_init_seeds.argtypes = [vp, C.c_double]
    L.spk_hip_sweeps.argtypes = [vp, C.c_int, vp, vp, vp]
    L.spk_hip_energy.argtypes = [vp, C.POINTER(C.c_double)]
    L.spk_hip_sync.argtypes = [vp]
    L.spk_hip_set_launch.argtypes = [vp, C.c_int]
    L.spk_hip_set_bitplanes.argtypes = [vp, C.c_int]
    L.spk_hip_set_resident.argtypes = [vp, C.c_int]
    L.spk_hip_checksum.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.spk_hip_set_graph.argtypes = [vp, C.c_int]
    L.spk_hip_prepare.argtypes = [vp, C.c_int]
    L.spk_hip_sweeps_noalloc.argtypes = [vp, C.c_int, vp]
    L.spk_hip_last_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    if path is None:
        _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = load_library().spk_hip_last_error()
        raise HipError(rc, msg.decode() if msg else "error")
    return rc


def make_desc(app, lattice, nx, ny, nz=1, nspins=2, temperature=0.0):
    return Desc(APPS[app], LATTICES[lattice], int(nx), int(ny), int(nz), int(nspins),
                float(temperature))


def ncolors(app, lattice, nx, ny, nz=1):
    d = make_desc(app, lattice, nx, ny, nz, 2, 1.0)
    return _check_nonneg(load_library().spk_hip_ncolors(C.byref(d)))


def _check_nonneg(v):
    if v < 0:
        _check(v)
    return v


def metropolis_table(temperature, maxd):
    tab = np.zeros(maxd + 1, np.float64)
    thr = np.zeros(maxd + 1, np.int32)
    _check(load_library().spk_hip_metropolis_table(float(temperature), int(maxd),
                                                   tab.ctypes.data, thr.ctypes.data))
    return tab, thr


def checksum(spin, seed):
    """spk_hip_checksum of host arrays in reference site order: the wrapping
    sum over sites g of mix64(g*C1 + spin*C2 + seed), mix64 = splitmix64's
    finaliser."""
    spin, seed = np.asarray(spin), np.asarray(seed)
    total = np.uint64(0)
    step = 1 << 24   # bounded temporaries for lattices of up to 2^31 sites
    with np.errstate(over="ignore"):
        for lo in range(0, len(spin), step):
            hi = min(lo + step, len(spin))
            g = np.arange(lo, hi, dtype=np.uint64)
            z = (g * np.uint64(0xD1B54A32D192ED03)
                 + spin[lo:hi].astype(np.uint32).astype(np.uint64) * np.uint64(0xABC98388FB8FAC03)
                 + seed[lo:hi].astype(np.uint32).astype(np.uint64))
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            total += z.sum(dtype=np.uint64)
    return int(total)


def master_draws(seed, n):
    """First n draws of the `seed` command's generator (host only)."""
    out = np.zeros(n, np.float64)
    _check(load_library().spk_hip_master_draws(int(seed), int(n), out.ctypes.data))
    return out


def strict_init_params(seed):
    """(set-range generator seed, site-seed base rseed) of a script that runs
    `seed S`, `set site range ...`, then its first strict `run`."""
    d = master_draws(seed, 5)
    s0 = int(d[0] * 2147483647.0)
    return (s0 if s0 != 0 else 1), float(d[4])


class Lattice:
    """One periodic lattice resident on the GPU (a libspk_hip.so handle)."""

    def __init__(self, app, lattice, nx, ny, nz=1, nspins=2, temperature=0.0):
        self.L = load_library()
        self.app, self.lattice = app, lattice
        self.shape = (int(nx), int(ny), int(nz))
        self.n = int(nx) * int(ny) * int(nz)
        self.nspins = int(nspins)
        self.desc = make_desc(app, lattice, nx, ny, nz, nspins, temperature)
        h = C.c_void_p()
        _check(self.L.spk_hip_create(C.byref(self.desc), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.spk_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def ncolors(self):
        return _check_nonneg(self.L.spk_hip_ncolors(C.byref(self.desc)))

    def set_temperature(self, t):
        _check(self.L.spk_hip_set_temperature(self.h, float(t)))

    def upload(self, spin=None, seed=None):
        sp = None if spin is None else np.ascontiguousarray(spin, np.int32)
        se = None if seed is None else np.ascontiguousarray(seed, np.int32)
        for a in (sp, se):
            if a is not None and a.size != self.n:
                raise ValueError("array of %d values for %d sites" % (a.size, self.n))
        _check(self.L.spk_hip_upload(self.h, None if sp is None else sp.ctypes.data,
                                     None if se is None else se.ctypes.data))

    def download(self):
        spin = np.empty(self.n, np.int32)
        seed = np.empty(self.n, np.int32)
        _check(self.L.spk_hip_download(self.h, spin.ctypes.data, seed.ctypes.data))
        return spin, seed

    def set_range(self, seed0, lo, hi):
        _check(self.L.spk_hip_set_range(self.h, int(seed0), int(lo), int(hi)))

    def set_value(self, v):
        _check(self.L.spk_hip_set_value(self.h, int(v)))

    def init_seeds(self, rseed):
        _check(self.L.spk_hip_init_seeds(self.h, float(rseed)))

    def sweeps(self, nsweeps, per_sweep=True):
        """Run nsweeps; returns (naccept_total, cumulative naccept[], energy[])."""
        tot = C.c_int64(0)
        if per_sweep and nsweeps > 0:
            nacc = np.zeros(nsweeps, np.int64)
            eng = np.zeros(nsweeps, np.float64)
            _check(self.L.spk_hip_sweeps(self.h, int(nsweeps), C.addressof(tot),
                                         nacc.ctypes.data, eng.ctypes.data))
            return tot.value, nacc, eng
        _check(self.L.spk_hip_sweeps(self.h, int(nsweeps), C.addressof(tot), None, None))
        return tot.value, None, None

    def energy(self):
        e = C.c_double()
        _check(self.L.spk_hip_energy(self.h, C.byref(e)))
        return e.value

    def sync(self):
        _check(self.L.spk_hip_sync(self.h))

    def set_launch(self, block):
        _check(self.L.spk_hip_set_launch(self.h, int(block)))

    def set_resident(self, enable):
        _check(self.L.spk_hip_set_resident(self.h, int(bool(enable))))

    def set_bitplanes(self, enable):
        """0/False: off; 1/True: on (two sweeps per launch where the rows
        hold 33..64 plane words); 8, 16 or 32: the per-colour kernel with
        that many sites per thread."""
        _check(self.L.spk_hip_set_bitplanes(self.h, int(enable)))

    def set_graph(self, chunk_sweeps):
        _check(self.L.spk_hip_set_graph(self.h, int(chunk_sweeps)))

    def sweeps_noalloc(self, nsweeps, rows_ptr):
        """Queue nsweeps sweeps that allocate nothing; per-sweep (Naccept,
        dE) pairs go to the caller's device buffer at rows_ptr (2*nsweeps
        uint64).  Asynchronous: sync() before reading rows."""
        _check(self.L.spk_hip_sweeps_noalloc(self.h, int(nsweeps), C.c_void_p(int(rows_ptr))))

    def prepare(self, nsweeps):
        """Capture the graphs (and pack the bit planes) a sweeps(nsweeps)
        call will use; runs no sweep."""
        _check(self.L.spk_hip_prepare(self.h, int(nsweeps)))

    def checksum(self):
        v = C.c_uint64()
        _check(self.L.spk_hip_checksum(self.h, C.byref(v)))
        return v.value

    def last_timing(self):
        ms = C.c_double()
        nl = C.c_int64()
        _check(self.L.spk_hip_last_timing(self.h, C.byref(ms), C.byref(nl)))
        return ms.value, nl.value
