"""ctypes view of the CPU oracle (oracle/liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing being
measured or shipped.  The product path (spparks_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_random_park.so")

APP = {"potts": 0, "ising": 1, "potts/neigh": 2, "potts/neighonly": 3, "ising/single": 4}
LATTICE = {"sq/4n": 0, "sq/8n": 1, "sc/6n": 2, "sc/26n": 3}


class Desc(C.Structure):
    _fields_ = [("app", C.c_int), ("lattice", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("nz", C.c_int), ("nspins", C.c_int),
                ("temperature", C.c_double)]


def make_desc(app, lattice, nx, ny, nz=1, nspins=2, temperature=0.0):
    return Desc(APP[app], LATTICE[lattice], nx, ny, nz, nspins, float(temperature))


_lib = None
_ref = None


def build():
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_pm_uniform.argtypes = [C.POINTER(C.c_int)]
        L.orc_pm_uniform.restype = C.c_double
        L.orc_pm_irandom.argtypes = [C.POINTER(C.c_int), C.c_int]
        L.orc_pm_from_double.argtypes = [C.c_double]
        L.orc_pm_tagreset.argtypes = [C.c_double, C.c_int64, C.c_int]
        L.orc_ranmars_uniform.restype = C.c_double
        L.orc_maxneigh.argtypes = [C.c_int]
        L.orc_ncolors.argtypes = [C.POINTER(Desc)]
        L.orc_color.argtypes = [C.POINTER(Desc), C.c_int64]
        L.orc_stencil.argtypes = [C.c_int, i32p]
        L.orc_neighbors.argtypes = [C.POINTER(Desc), i32p]
        L.orc_set_range.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, i32p]
        L.orc_site_seeds.argtypes = [C.c_double, C.c_int64, i32p]
        L.orc_script_init.argtypes = [C.POINTER(Desc), C.c_int, C.c_int, C.c_int, i32p, i32p]
        L.orc_energy.argtypes = [C.POINTER(Desc), i32p]
        L.orc_energy.restype = C.c_double
        L.orc_sweeps.argtypes = [C.POINTER(Desc), i32p, i32p, C.c_int,
                                 C.c_void_p, C.c_void_p, C.c_int]
        L.orc_site_event.argtypes = [C.POINTER(Desc), i32p, C.c_int64, C.POINTER(C.c_int)]
        L.orc_run_nsweeps.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.orc_run_nsweeps.restype = C.c_int64
        _lib = L
    return _lib


def ref():
    """The reference's own RandomPark (oracle/_ref), or None where not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        R.ref_uniform_seq.argtypes = [C.c_int, C.c_int, f64p, i32p]
        R.ref_irandom_seq.argtypes = [C.c_int, C.c_int, C.c_int, i32p]
        R.ref_from_double.argtypes = [C.c_double]
        R.ref_tagreset.argtypes = [C.c_double, i64p, C.c_int, C.c_int, i32p]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers

class RanMars(C.Structure):
    _fields_ = [("u", C.c_double * 98), ("c", C.c_double), ("cd", C.c_double),
                ("cm", C.c_double), ("i97", C.c_int), ("j97", C.c_int)]

    def __init__(self, seed):
        super().__init__()
        lib().orc_ranmars_init(C.byref(self), int(seed))

    def uniform(self):
        return lib().orc_ranmars_uniform(C.byref(self))


def pm_uniform_seq(seed, n):
    s = C.c_int(int(seed))
    u = np.empty(n, np.float64)
    st = np.empty(n, np.int32)
    L = lib()
    for i in range(n):
        u[i] = L.orc_pm_uniform(C.byref(s))
        st[i] = s.value
    return u, st


def master_draws(seed, n):
    """The first n uniforms of the master generator of `seed S`."""
    r = RanMars(seed)
    return [r.uniform() for _ in range(n)]


def site_seeds(rseed, n):
    """Strict per-site seeds of a first run whose seed draw is rseed."""
    out = np.empty(n, np.int32)
    lib().orc_site_seeds(float(rseed), int(n), out)
    return out


def script_init(desc, master_seed, lo, hi):
    n = desc.nx * desc.ny * desc.nz
    spin = np.empty(n, np.int32)
    seeds = np.empty(n, np.int32)
    rc = lib().orc_script_init(C.byref(desc), int(master_seed), int(lo), int(hi), spin, seeds)
    if rc:
        raise ValueError("orc_script_init failed: %d" % rc)
    return spin, seeds


def sweeps(desc, spin, seeds, nsweeps, nthreads=0, energies=True):
    """Run nsweeps in place; returns (cumulative naccept[nsweeps], energy[nsweeps]).
    energies=False skips the per-sweep energy pass (energy comes back zero)."""
    nacc = np.zeros(max(nsweeps, 1), np.int64)
    eng = np.zeros(max(nsweeps, 1), np.float64)
    rc = lib().orc_sweeps(C.byref(desc), spin, seeds, int(nsweeps),
                          nacc.ctypes.data, eng.ctypes.data if energies else None,
                          int(nthreads))
    if rc:
        raise ValueError("orc_sweeps failed: %d" % rc)
    return nacc[:nsweeps], eng[:nsweeps]


def energy(desc, spin):
    return lib().orc_energy(C.byref(desc), spin)


def colors(desc):
    n = desc.nx * desc.ny * desc.nz
    L = lib()
    return np.array([L.orc_color(C.byref(desc), g + 1) for g in range(n)], np.int32)


def neighbors(desc):
    n = desc.nx * desc.ny * desc.nz
    mn = lib().orc_maxneigh(desc.lattice)
    out = np.empty(n * mn, np.int32)
    lib().orc_neighbors(C.byref(desc), out)
    return out.reshape(n, mn)


def run_nsweeps(t0, runtime, dt_sweep):
    tend = C.c_double()
    n = lib().orc_run_nsweeps(float(t0), float(runtime), float(dt_sweep), C.byref(tend))
    return n, tend.value
