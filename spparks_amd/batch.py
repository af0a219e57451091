"""ctypes binding of the batched-replica API of libspk_hip.so
(spk_hip_batch_*, include/spk_hip.h).  No fallback: every call goes to the
library and raises on error."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .hip import APPS, LATTICES, Desc, HipError, _check, load_library


class ReplicaDesc(C.Structure):
    _fields_ = [("lattice", Desc), ("seed", C.c_int), ("lo", C.c_int), ("hi", C.c_int),
                ("nsweeps", C.c_int)]


class Request(C.Structure):
    _fields_ = [("replica", ReplicaDesc), ("arrive", C.c_double)]


class Result(C.Structure):
    _fields_ = [("admit", C.c_double), ("finish", C.c_double), ("naccept", C.c_int64),
                ("energy", C.c_double), ("checksum", C.c_uint64), ("status", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("wall", C.c_double), ("site_updates", C.c_int64), ("blocks", C.c_int),
                ("max_live", C.c_int), ("max_pool_used", C.c_int64), ("busy", C.c_double)]


_bound = False


def _bind():
    global _bound
    L = load_library()
    if _bound:
        return L
    vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
    L.spk_hip_batch_create.argtypes = [i64, i32, C.POINTER(vp)]
    L.spk_hip_batch_destroy.argtypes = [vp]
    L.spk_hip_batch_destroy.restype = None
    L.spk_hip_replica_bytes.argtypes = [C.POINTER(Desc), C.POINTER(i64)]
    L.spk_hip_batch_add.argtypes = [vp, C.POINTER(ReplicaDesc), C.POINTER(i32)]
    L.spk_hip_batch_remove.argtypes = [vp, i32]
    L.spk_hip_batch_upload.argtypes = [vp, i32, vp, vp]
    L.spk_hip_batch_download.argtypes = [vp, i32, vp, vp]
    L.spk_hip_batch_run.argtypes = [vp, i32, i32]
    L.spk_hip_batch_status.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64),
                                       C.POINTER(dbl)]
    L.spk_hip_batch_trace.argtypes = [vp, i32, i32, vp, vp]
    L.spk_hip_batch_set_launch.argtypes = [vp, i32, i32]
    L.spk_hip_batch_last_timing.argtypes = [vp, C.POINTER(dbl), C.POINTER(i64), C.POINTER(i64),
                                            C.POINTER(i32), C.POINTER(i32)]
    L.spk_hip_batch_pool.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]
    L.spk_hip_batch_checksum.argtypes = [vp, i32, C.POINTER(C.c_uint64)]
    L.spk_hip_serve.argtypes = [C.POINTER(Request), i32, i64, i32, i32, i32, i32, C.POINTER(Result),
                                C.POINTER(ServeStats)]
    _bound = True
    return L


def replica_desc(app, lattice, nx, ny, nz=1, nspins=2, temperature=0.0, seed=1, lo=1, hi=None,
                 nsweeps=0):
    q = int(nspins) if app.startswith("potts") else 2
    d = Desc(APPS[app], LATTICES[lattice], int(nx), int(ny), int(nz), q, float(temperature))
    return ReplicaDesc(d, int(seed), int(lo), int(q if hi is None else hi), int(nsweeps))


def replica_bytes(app, lattice, nx, ny, nz=1, nspins=2):
    L = _bind()
    d = Desc(APPS[app], LATTICES[lattice], int(nx), int(ny), int(nz), int(nspins), 1.0)
    out = C.c_int64()
    _check(L.spk_hip_replica_bytes(C.byref(d), C.byref(out)))
    return out.value


class Batch:
    """A device pool of replicas advanced together by the batched sweep."""

    def __init__(self, pool_bytes, max_replicas):
        self.L = _bind()
        h = C.c_void_p()
        _check(self.L.spk_hip_batch_create(int(pool_bytes), int(max_replicas), C.byref(h)))
        self.h = h
        self.sizes = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.spk_hip_batch_destroy(self.h)
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

    def add(self, desc: ReplicaDesc):
        """Admit a replica; returns its id, or None when the pool is full."""
        rid = C.c_int()
        rc = self.L.spk_hip_batch_add(self.h, C.byref(desc), C.byref(rid))
        if rc == -3:   # SPK_HIP_ENOMEM: admission refused, batch unchanged
            return None
        _check(rc)
        d = desc.lattice
        self.sizes[rid.value] = d.nx * d.ny * d.nz
        return rid.value

    def remove(self, rid):
        _check(self.L.spk_hip_batch_remove(self.h, int(rid)))
        self.sizes.pop(rid, None)

    def upload(self, rid, spin=None, seed=None):
        sp = None if spin is None else np.ascontiguousarray(spin, np.int32)
        se = None if seed is None else np.ascontiguousarray(seed, np.int32)
        _check(self.L.spk_hip_batch_upload(self.h, int(rid), None if sp is None else sp.ctypes.data,
                                           None if se is None else se.ctypes.data))

    def download(self, rid):
        n = self.sizes[rid]
        spin = np.empty(n, np.int32)
        seed = np.empty(n, np.int32)
        _check(self.L.spk_hip_batch_download(self.h, int(rid), spin.ctypes.data, seed.ctypes.data))
        return spin, seed

    def run(self, max_sweeps, graph=True):
        _check(self.L.spk_hip_batch_run(self.h, int(max_sweeps), 1 if graph else 0))

    def status(self, rid):
        done, left = C.c_int(), C.c_int()
        nacc, eng = C.c_int64(), C.c_double()
        _check(self.L.spk_hip_batch_status(self.h, int(rid), C.byref(done), C.byref(left), C.byref(nacc),
                                           C.byref(eng)))
        return done.value, left.value, nacc.value, eng.value

    def trace(self, rid, n):
        nacc = np.zeros(max(n, 1), np.int64)
        eng = np.zeros(max(n, 1), np.float64)
        _check(self.L.spk_hip_batch_trace(self.h, int(rid), int(n), nacc.ctypes.data, eng.ctypes.data))
        return nacc[:n], eng[:n]

    def set_launch(self, block=0, grid=0):
        _check(self.L.spk_hip_batch_set_launch(self.h, int(block), int(grid)))

    def last_timing(self):
        ms, upd, nl = C.c_double(), C.c_int64(), C.c_int64()
        bs, gr = C.c_int(), C.c_int()
        _check(self.L.spk_hip_batch_last_timing(self.h, C.byref(ms), C.byref(upd), C.byref(nl), C.byref(bs),
                                                C.byref(gr)))
        return {"ms": ms.value, "updates": upd.value, "launches": nl.value, "block": bs.value,
                "grid": gr.value}

    def checksum(self, rid):
        v = C.c_uint64()
        _check(self.L.spk_hip_batch_checksum(self.h, int(rid), C.byref(v)))
        return v.value

    def pool(self):
        used, largest, live = C.c_int64(), C.c_int64(), C.c_int()
        _check(self.L.spk_hip_batch_pool(self.h, C.byref(used), C.byref(largest), C.byref(live)))
        return used.value, largest.value, live.value


def serve(requests, pool_bytes, max_live, block_sweeps=10, block=0, grid=0):
    """Continuous batching of `requests` (dicts with app, lattice, shape,
    nspins, kT, seed, sweeps, arrive, and optionally lo, hi for `set site
    range lo hi`, default 1 Q; sorted by arrive) through spk_hip_serve.
    Returns (results as dicts, stats dict)."""
    L = _bind()
    n = len(requests)
    arr = (Request * max(n, 1))()
    for i, r in enumerate(requests):
        arr[i].replica = replica_desc(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"],
                                      temperature=r["kT"], seed=r["seed"], lo=r.get("lo", 1), hi=r.get("hi"),
                                      nsweeps=r["sweeps"])
        arr[i].arrive = float(r["arrive"])
    out = (Result * max(n, 1))()
    st = ServeStats()
    _check(L.spk_hip_serve(arr, n, int(pool_bytes), int(max_live), int(block_sweeps), int(block), int(grid),
                           out, C.byref(st)))
    res = [{"admit": o.admit, "finish": o.finish, "naccept": o.naccept, "energy": o.energy,
            "checksum": o.checksum, "status": o.status} for o in out[:n]]
    stats = {f: getattr(st, f) for f, _ in ServeStats._fields_}
    return res, stats
