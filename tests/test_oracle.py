"""The CPU oracle pinned: against the reference's own RandomPark (oracle/_ref),
the reference run's known answers, and closed forms / invariants."""
import ctypes as C
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------- Park-Miller

def test_park_miller_check_value(orc):
    # Park & Miller (1988): from seed 1 the 10000th state is 1043618065
    s = C.c_int(1)
    for _ in range(10000):
        orc.lib().orc_pm_uniform(C.byref(s))
    assert s.value == 1043618065


def test_park_miller_closed_form(orc):
    # each step is s -> 16807 s mod (2^31-1)
    rng = np.random.default_rng(1)
    for s0 in rng.integers(1, 2**31 - 1, 200):
        u, st = orc.pm_uniform_seq(int(s0), 3)
        x = int(s0)
        for k in range(3):
            x = (16807 * x) % (2**31 - 1)
            assert st[k] == x
            assert u[k] == (1.0 / 2147483647) * x


def test_park_miller_matches_reference(orc):
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(2)
    seeds = list(rng.integers(1, 2**31 - 1, 50)) + [1, 2, 127773, 127772, 2**31 - 2]
    for s0 in seeds:
        u = np.empty(64, np.float64)
        st = np.empty(64, np.int32)
        R.ref_uniform_seq(int(s0), 64, u, st)
        uo, sto = orc.pm_uniform_seq(int(s0), 64)
        assert np.array_equal(st, sto)
        assert np.array_equal(u.view(np.int64), uo.view(np.int64))


def test_irandom_and_from_double_match_reference(orc):
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    L = orc.lib()
    for s0, rng_n in [(1191027839, 100), (5, 2), (987654321, 10000), (123, 1000000)]:
        out = np.empty(500, np.int32)
        R.ref_irandom_seq(s0, rng_n, 500, out)
        s = C.c_int(s0)
        mine = [L.orc_pm_irandom(C.byref(s), rng_n) for _ in range(500)]
        assert list(out) == mine
    for r in [0.0, 1e-12, 0.5546155571937561, 0.9999999999, 0.25]:
        assert R.ref_from_double(r) == L.orc_pm_from_double(r)


def test_tagreset_matches_reference(orc):
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    L = orc.lib()
    offs = np.array([1, 2, 3, 100, 16384, 16777216, 2**31 - 2], np.int64)
    for rseed in [0.9647188782691956, 0.0, 0.999999999, 0.123456789, 1e-9]:
        out = np.empty(len(offs), np.int32)
        R.ref_tagreset(rseed, offs, len(offs), 100, out)
        mine = [L.orc_pm_tagreset(rseed, int(o), 100) for o in offs]
        assert list(out) == mine


def test_tagreset_closed_form(orc):
    # 100 warm-up steps == one multiplication by 16807^100 mod m
    L = orc.lib()
    a100 = pow(16807, 100, 2**31 - 1)
    for rseed, off in [(0.3, 7), (0.77, 123456), (0.0, 5)]:
        s0 = L.orc_pm_tagreset(rseed, off, 0)
        assert L.orc_pm_tagreset(rseed, off, 100) == (s0 * a100) % (2**31 - 1)


# ------------------------------------------------------------- RanMars

def test_ranmars_published_check_values(orc):
    # Marsaglia-Zaman / James (1990) RANMAR test: seeds ij=1802, kl=9373,
    # after 20000 draws the next six are 6533892, 14220222, 7275067,
    # 6172232, 8354498, 10633180 (times 2^-24).  SPPARKS packs (ij, kl)
    # into one seed as ij*30082 + kl + 1.
    m = orc.RanMars(1802 * 30082 + 9373 + 1)
    # init() already consumed one draw
    for _ in range(20000 - 1):
        m.uniform()
    got = [round(m.uniform() * 2**24) for _ in range(6)]
    assert got == [6533892, 14220222, 7275067, 6172232, 8354498, 10633180]


# ------------------------------------------------------------- geometry

@pytest.mark.parametrize("lat,n", [("sq/4n", 4), ("sq/8n", 8), ("sc/6n", 6), ("sc/26n", 26)])
def test_stencil_counts(orc, lat, n):
    offs = np.zeros(81, np.int32)
    assert orc.lib().orc_stencil(orc.LATTICE[lat], offs) == n
    o = offs[:3 * n].reshape(n, 3)
    assert len({tuple(r) for r in o}) == n
    assert not any((r == 0).all() for r in o)
    # symmetric stencil
    assert {tuple(r) for r in o} == {tuple(-r) for r in o}


@pytest.mark.parametrize("lat,shape,K", [("sq/4n", (6, 4, 1), 2), ("sq/8n", (4, 6, 1), 4),
                                         ("sc/6n", (4, 2, 6), 2), ("sc/26n", (4, 6, 2), 8),
                                         ("sc/26n", (2, 2, 2), 8), ("sq/8n", (2, 2, 1), 4)])
def test_coloring_invariants(orc, lat, shape, K):
    d = orc.make_desc("potts", lat, *shape, 5, 1.0)
    assert orc.lib().orc_ncolors(C.byref(d)) == K
    col = orc.colors(d)
    nb = orc.neighbors(d)
    # no site shares a colour with any of its neighbours
    assert not (col[nb] == col[:, None]).any()
    assert sorted(set(col)) == list(range(1, K + 1))
    assert all((col == c).sum() == col.size // K for c in range(1, K + 1))


def test_odd_edge_rejected(orc):
    d = orc.make_desc("potts", "sq/8n", 6, 5, 1, 5, 1.0)
    assert orc.lib().orc_ncolors(C.byref(d)) == 0


# ------------------------------------------------------------- known answers

def test_potts_sq8n_128_known_answers(orc):
    g = json.load(open(os.path.join(GOLD, "potts_sq8n_128_known_answers.json")))
    d = orc.make_desc(g["app"], g["lattice"], g["nx"], g["ny"], g["nz"], g["nspins"], g["temperature"])
    spin, seeds = orc.script_init(d, g["seed"], g["lo"], g["hi"])
    assert orc.energy(d, spin) == g["initial_energy"]
    nacc, eng = orc.sweeps(d, spin, seeds, g["nsweeps"])
    for sweep, na, e in g["rows"]:
        assert nacc[sweep - 1] == na
        assert eng[sweep - 1] == e


def test_run_sweep_counts(orc):
    # time accumulation decides the count (run 0.02 with Q=1e4 is 201 sweeps)
    assert orc.run_nsweeps(0.0, 1.0, 1 / 100)[0] == 100
    assert orc.run_nsweeps(0.0, 0.02, 1 / 10000)[0] == 201
    assert orc.run_nsweeps(0.0, 0.01995, 1 / 10000)[0] == 200
    assert orc.run_nsweeps(0.0, 500.0, 0.5)[0] == 1000
    assert orc.run_nsweeps(0.0, 0.3, 0.5)[0] == 1


# ------------------------------------------------------------- invariants

def _bonds_energy(orc, d, spin):
    nb = orc.neighbors(d)
    return float((spin[nb] != spin[:, None]).sum())


@pytest.mark.parametrize("app,lat,shape,q,kt", [
    ("potts", "sq/8n", (8, 6, 1), 5, 0.7), ("ising", "sq/4n", (10, 8, 1), 2, 2.269),
    ("potts", "sc/26n", (4, 6, 4), 50, 0.5), ("ising", "sc/6n", (4, 4, 6), 2, 1.5),
    ("potts", "sq/4n", (2, 2, 1), 3, 1.0), ("potts", "sc/26n", (2, 2, 2), 4, 0.3)])
def test_energy_is_bond_count(orc, app, lat, shape, q, kt):
    d = orc.make_desc(app, lat, *shape, q, kt)
    spin, seeds = orc.script_init(d, 4242, 1, q)
    assert orc.energy(d, spin) == _bonds_energy(orc, d, spin)
    nacc, eng = orc.sweeps(d, spin, seeds, 3)
    assert eng[-1] == _bonds_energy(orc, d, spin)
    assert (np.diff(nacc) >= 0).all()


def test_zero_temperature_never_uphill(orc):
    d = orc.make_desc("potts", "sq/8n", 16, 16, 1, 10, 0.0)
    spin, seeds = orc.script_init(d, 99, 1, 10)
    e = orc.energy(d, spin)
    for _ in range(5):
        _, eng = orc.sweeps(d, spin, seeds, 1)
        assert eng[0] <= e
        e = eng[0]


def test_single_state_potts_never_changes(orc):
    d = orc.make_desc("potts", "sc/26n", 4, 4, 4, 1, 0.8)
    spin, seeds = orc.script_init(d, 3, 1, 1)
    s0 = seeds.copy()
    nacc, _ = orc.sweeps(d, spin, seeds, 2)
    assert nacc[-1] == 0 and (spin == 1).all()
    # every site still advanced its own generator (level moves draw once)
    a = pow(16807, 2, 2**31 - 1)
    assert np.array_equal(seeds.astype(np.int64), (s0.astype(np.int64) * a) % (2**31 - 1))


def test_site_event_by_hand(orc):
    """One Potts event recomputed from the rule's definition with Python ints
    and math.exp: a dropped term, wrong sign or wrong draw count fails it."""
    import math
    d = orc.make_desc("potts", "sq/4n", 4, 4, 1, 7, 0.9)
    rng = np.random.default_rng(5)
    nb = orc.neighbors(d)
    for trial in range(300):
        spin = rng.integers(1, 8, 16).astype(np.int32)
        g = int(rng.integers(0, 16))
        s0 = int(rng.integers(1, 2**31 - 1))
        # expectation
        old = int(spin[g])
        ei = sum(int(spin[j] != old) for j in nb[g])
        s1 = (16807 * s0) % (2**31 - 1)
        prop = min(int(7 * ((1.0 / 2147483647) * s1)) + 1, 7)
        ef = sum(int(spin[j] != prop) for j in nb[g])
        s_end, new = s1, prop
        if ef > ei:
            s2 = (16807 * s1) % (2**31 - 1)
            s_end = s2
            if (1.0 / 2147483647) * s2 > math.exp((ei - ef) * (1.0 / 0.9)):
                new = old
        seed = C.c_int(s0)
        changed = orc.lib().orc_site_event(C.byref(d), spin, g, C.byref(seed))
        assert seed.value == s_end
        assert spin[g] == new
        assert changed == int(new != old)


@pytest.mark.parametrize("lat,shape", [("sq/4n", (12, 8, 1)), ("sc/26n", (4, 4, 6)),
                                       ("sq/8n", (6, 4, 1)), ("sc/6n", (4, 6, 2))])
def test_ising_is_potts_q2(orc, lat, shape):
    """(int)(2*u)+1 is 1 exactly when u < 0.5 (2*u is exact), so Ising's
    proposal equals Potts Q=2's; the two apps then share site_energy and the
    Metropolis test.  This carries the Potts known answers over to Ising."""
    di = orc.make_desc("ising", lat, *shape, 2, 1.3)
    dp = orc.make_desc("potts", lat, *shape, 2, 1.3)
    si, ki = orc.script_init(di, 777, 1, 2)
    sp, kp = si.copy(), ki.copy()
    ni, ei = orc.sweeps(di, si, ki, 6)
    np_, ep = orc.sweeps(dp, sp, kp, 6)
    assert np.array_equal(si, sp) and np.array_equal(ki, kp)
    assert np.array_equal(ni, np_) and np.array_equal(ei, ep)


@pytest.mark.parametrize("lat", ["sq/4n", "sq/8n", "sc/6n", "sc/26n"])
def test_stencil_order_is_reference_loop_order(orc, lat):
    """The neighbour list order is what offsets_2d / offsets_3d produce
    (src/create_sites.cpp:1303-1371): dx outermost, then dy, then dz, each
    -1..1, kept when the distance is within the lattice's shell.  Order
    matters for potts/neigh and potts/neighonly (first-appearance lists)."""
    import math
    is3 = lat.startswith("sc")
    cut = {"sq/4n": 1.0, "sc/6n": 1.0, "sq/8n": math.sqrt(2.0), "sc/26n": math.sqrt(3.0)}[lat]
    want = []
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in ((-1, 0, 1) if is3 else (0,)):
                r = math.sqrt(i * i + j * j + k * k)
                if 1.0 - 1e-6 < r < cut + 1e-6:
                    want.append((i, j, k))
    nx, ny, nz = (6, 6, 6) if is3 else (6, 6, 1)
    d = orc.make_desc("potts", lat, nx, ny, nz, 3, 1.0)
    nb = orc.neighbors(d)
    g = 2 + nx * (3 + ny * (2 if is3 else 0))   # an interior site: no wrap
    i0, j0, k0 = 2, 3, (2 if is3 else 0)
    got = [((n % nx) - i0, ((n // nx) % ny) - j0, (n // (nx * ny)) - k0) for n in nb[g]]
    assert got == want


def _unique_unlike(vals, me):
    out = []
    for v in vals:
        if v != me and v not in out:
            out.append(v)
    return out


@pytest.mark.parametrize("app", ["potts/neigh", "potts/neighonly", "ising/single"])
@pytest.mark.parametrize("lat", ["sq/4n", "sq/8n", "sc/26n"])
@pytest.mark.parametrize("kt", [0.9, 0.0])
def test_variant_site_event_by_hand(orc, app, lat, kt):
    """One event of each proposal variant recomputed from its definition with
    Python ints and math.exp (src/app_potts_neigh.cpp:71-116,
    src/app_potts_neighonly.cpp:71-118, src/app_ising_single.cpp:39-77):
    which value is proposed, how many draws are taken (none for an
    ising/single downhill flip or a potts/neighonly site with no unlike
    neighbour), and the Metropolis test."""
    import math
    AM = 1.0 / 2147483647
    step = lambda s: (16807 * s) % (2**31 - 1)
    is3 = lat.startswith("sc")
    shape = (4, 4, 4) if is3 else (4, 4, 1)
    q = 2 if app == "ising/single" else 5
    d = orc.make_desc(app, lat, *shape, q, kt)
    nb = orc.neighbors(d)
    ns = nb.shape[1]
    n = nb.shape[0]
    rng = np.random.default_rng(11)
    for trial in range(400):
        # low-entropy neighbourhoods too, so nevent = 0 and 1 both occur
        hi = q if trial % 3 else 2
        spin = rng.integers(1, hi + 1, n).astype(np.int32)
        g = int(rng.integers(0, n))
        s0 = int(rng.integers(1, 2**31 - 1))
        old = int(spin[g])
        vals = [int(spin[j]) for j in nb[g]]
        ei = sum(v != old for v in vals)
        s = s0
        prop = None
        if app == "ising/single":
            prop = 2 if old == 1 else 1
        else:
            uniq = _unique_unlike(vals, old)
            if app == "potts/neigh":
                s = step(s)
                iran = int(ns * (AM * s))
                if iran < len(uniq):
                    prop = uniq[iran]
            elif uniq:
                s = step(s)
                iran = min(int(len(uniq) * (AM * s)), len(uniq) - 1)
                prop = uniq[iran]
        new = old
        if prop is not None:
            ef = sum(v != prop for v in vals)
            new = prop
            if ef > ei:
                if kt == 0.0:
                    new = old
                else:
                    s = step(s)
                    if AM * s > math.exp((ei - ef) * (1.0 / kt)):
                        new = old
        seed = C.c_int(s0)
        changed = orc.lib().orc_site_event(C.byref(d), spin, g, C.byref(seed))
        assert seed.value == s, (trial, app)
        assert spin[g] == new, (trial, app)
        assert changed == int(new != old)


def test_variants_energy_and_draw_invariants(orc):
    """potts/neighonly never draws at a site with no unlike neighbour, so a
    uniform lattice keeps every seed; potts/neigh draws exactly once there;
    energies stay the bond count for all three variants."""
    for app, q in (("potts/neighonly", 4), ("potts/neigh", 4)):
        d = orc.make_desc(app, "sc/26n", 4, 4, 4, q, 0.7)
        spin = np.full(64, 3, np.int32)
        seeds = np.arange(1, 65, dtype=np.int32) * 1000003
        s0 = seeds.copy()
        nacc, _ = orc.sweeps(d, spin, seeds, 2)
        assert nacc[-1] == 0 and (spin == 3).all()
        k = 0 if app == "potts/neighonly" else 2
        a = pow(16807, k, 2**31 - 1)
        assert np.array_equal(seeds.astype(np.int64), (s0.astype(np.int64) * a) % (2**31 - 1))
    for app, q in (("potts/neigh", 30), ("potts/neighonly", 30), ("ising/single", 2)):
        d = orc.make_desc(app, "sq/8n", 12, 10, 1, q, 0.8)
        spin, seeds = orc.script_init(d, 31337, 1, q)
        nacc, eng = orc.sweeps(d, spin, seeds, 4)
        assert eng[-1] == _bonds_energy(orc, d, spin)
        assert (np.diff(nacc) >= 0).all() and nacc[-1] > 0
