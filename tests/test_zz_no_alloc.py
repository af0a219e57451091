"""spk_hip_sweeps_noalloc (no allocation; accumulator rows in the library's
device array) against spk_hip_sweeps on identically initialised lattices,
through the C ABI: per-sweep Naccept and energy, and the final spins and
seeds, bit for bit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceRows:
    """A caller-owned device buffer of uint64, allocated by the test through
    the HIP runtime libspk_hip.so runs on (torch brings its own runtime, so
    it cannot share pointers with the library in one process)."""

    def __init__(self, n):
        import spparks_amd as S
        S.load_library()                      # loads libamdhip64.so.7 first
        self.rt = C.CDLL("libamdhip64.so.7")  # the same, already loaded, runtime
        self.n = n
        self.ptr = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0
        assert self.rt.hipMemset(self.ptr, C.c_int(0xff), C.c_size_t(8 * n)) == 0
        assert self.rt.hipDeviceSynchronize() == 0

    def host(self):
        out = np.empty(self.n, np.int64)
        assert self.rt.hipMemcpy(C.c_void_p(out.ctypes.data), self.ptr, C.c_size_t(8 * self.n), 2) == 0   # D2H
        return out

    def free(self):
        if self.ptr:
            self.rt.hipFree(self.ptr)
            self.ptr = None


def say(*a):
    print(*a, flush=True)


def fresh(S, app, lat, shape, q, kt, seed):
    L = S.Lattice(app, lat, *shape, nspins=q, temperature=kt)
    s0, rseed = S.strict_init_params(seed)
    L.set_range(s0, 1, q)
    L.init_seeds(rseed)
    return L


def noalloc_run(L, n):
    """n sweeps through the no-allocation entry point; returns cumulative
    Naccept and the energy after each sweep."""
    e0 = L.energy()
    rows = DeviceRows(2 * n)
    try:
        L.sweeps_noalloc(n, rows.ptr.value)
        L.sync()
        r = rows.host()
    finally:
        rows.free()
    nacc = np.cumsum(r[0::2])
    eng = e0 + 2.0 * np.cumsum(r[1::2])
    return nacc, eng


CASES = [
    # app, lattice, shape, Q, kT, seed, sweeps: the first call of each pair
    # is the larger input, the second the smaller
    (("ising", "sq/4n", (512, 384, 1), 2, 2.269, 5, 60), ("potts", "sq/8n", (16, 12, 1), 100, 0.5, 6, 9)),
    (("potts", "sc/26n", (40, 32, 24), 70000, 0.7, 7, 30), ("potts", "sq/4n", (18, 10, 1), 300, 1.1, 8, 4)),
    (("potts/neigh", "sc/26n", (24, 16, 12), 1000, 0.5, 9, 12), ("ising/single", "sq/4n", (66, 10, 1), 2, 1.5, 10, 5)),
]


@pytest.mark.parametrize("pair", CASES, ids=["ising-planes-then-resident", "potts-i32-then-u16", "variants"])
def test_noalloc_sweeps_equal_allocating_sweeps(pair):
    """Two calls in a row, the second on a smaller lattice with fewer
    sweeps (the shared rows keep the first call's contents, so the second
    must clear its own); each against spk_hip_sweeps on a fresh lattice of
    the same script, then 3 more sweeps through spk_hip_sweeps on both."""
    import spparks_amd as S
    for case in pair:
        say("case", case)
        app, lat, shape, q, kt, seed, n = case
        with fresh(S, app, lat, shape, q, kt, seed) as A, fresh(S, app, lat, shape, q, kt, seed) as B:
            _, na, ea = A.sweeps(n)
            nb, eb = noalloc_run(B, n)
            assert np.array_equal(na, nb), case
            assert np.array_equal(ea, eb), case
            sa, ka = A.download()
            sb, kb = B.download()
            assert np.array_equal(sa, sb) and np.array_equal(ka, kb), case
            _, na2, ea2 = A.sweeps(3)
            _, nb2, eb2 = B.sweeps(3)
            assert np.array_equal(na2, nb2) and np.array_equal(ea2, eb2), case
            assert A.checksum() == B.checksum()


def test_noalloc_rejects_bad_arguments():
    """More sweeps than the rows hold, a negative count and a null output
    are refused with a message; zero sweeps does nothing."""
    import spparks_amd as S
    with fresh(S, "potts", "sq/8n", (16, 12, 1), 100, 0.5, 11) as L:
        rows = DeviceRows(2 * 300)
        try:
            with pytest.raises(S.HipError, match=r"outside \[0, 256\]"):
                L.sweeps_noalloc(257, rows.ptr.value)
            with pytest.raises(S.HipError, match=r"outside \[0, 256\]"):
                L.sweeps_noalloc(-1, rows.ptr.value)
            with pytest.raises(S.HipError, match="null argument"):
                L.sweeps_noalloc(1, 0)
            before = L.checksum()
            L.sweeps_noalloc(0, rows.ptr.value)
            L.sync()
            assert L.checksum() == before
            assert (rows.host() == -1).all()   # nothing written for zero sweeps
        finally:
            rows.free()
