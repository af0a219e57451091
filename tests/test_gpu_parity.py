"""GPU parity of the single-lattice path (libspk_hip.so) against the oracle:
bit-exact spins, seeds, cumulative Naccept and energy after every sweep."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_init(S, app, lat, shape, q, kt, seed, lo=1, hi=None):
    hi = q if hi is None else hi
    L = S.Lattice(app, lat, *shape, nspins=q, temperature=kt)
    s0, rseed = S.strict_init_params(seed)
    L.set_range(s0, lo, hi)
    L.init_seeds(rseed)
    return L


def oracle_run(orc, app, lat, shape, q, kt, seed, nsweeps, lo=1, hi=None):
    hi = q if hi is None else hi
    d = orc.make_desc(app, lat, *shape, q, kt)
    spin, seeds = orc.script_init(d, seed, lo, hi)
    spin0, seeds0 = spin.copy(), seeds.copy()
    nacc, eng = orc.sweeps(d, spin, seeds, nsweeps)
    return spin0, seeds0, spin, seeds, nacc, eng


CASES = [
    # app, lattice, shape, Q, kT, seed, sweeps
    ("potts", "sq/8n", (16, 12, 1), 100, 0.5, 56789, 7),
    ("potts", "sq/8n", (2, 2, 1), 3, 0.8, 11, 5),
    ("potts", "sq/4n", (18, 10, 1), 5, 1.1, 5, 6),
    ("ising", "sq/4n", (64, 48, 1), 2, 2.269, 777, 40),
    ("ising", "sq/4n", (2, 2, 1), 2, 2.0, 3, 5),
    ("potts", "sc/26n", (10, 8, 6), 10000, 0.5, 4242, 6),
    ("potts", "sc/26n", (2, 2, 2), 4, 0.6, 8, 5),
    ("potts", "sc/26n", (6, 4, 2), 7, 0.0, 91, 4),
    ("ising", "sc/26n", (8, 6, 4), 2, 3.0, 17, 5),
    ("potts", "sc/6n", (8, 6, 4), 9, 0.9, 23, 5),
    ("ising", "sc/6n", (4, 4, 2), 2, 1.0, 29, 5),
    ("potts", "sq/8n", (34, 66, 1), 1, 0.5, 31, 3),
    ("ising", "sq/8n", (12, 12, 1), 2, 0.0, 37, 4),
    ("potts", "sq/4n", (130, 6, 1), 1000000, 0.15, 41, 3),
    # storage / vector-width variants: u16 + 4 sites per thread, i32 + 4, u8 + 4
    ("potts", "sq/8n", (16, 8, 1), 300, 0.7, 43, 4),
    ("potts", "sq/4n", (16, 6, 1), 65535, 0.9, 47, 3),
    ("potts", "sc/26n", (16, 8, 8), 70000, 0.4, 53, 3),
    ("ising", "sc/6n", (16, 4, 4), 2, 1.7, 59, 3),
    ("potts", "sc/26n", (32, 16, 8), 255, 0.6, 61, 3),
    ("potts", "sc/26n", (32, 16, 8), 256, 0.6, 61, 3),
]


@pytest.mark.parametrize("block,resident", [(512, 0), (64, 0), (0, 1)], ids=["b512", "b64", "resident"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s" % (c[0], c[1], "x".join(map(str, c[2]))))
def test_small_lattice_bit_exact(orc, case, block, resident):
    """Every lattice x app x storage type: per-colour kernels with workgroup
    tiles of two shapes (partial tiles at the sub-grid's edges), and the
    one-workgroup kernel that keeps the whole state in LDS."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed, n = case
    say("case", case, "block", block, "resident", resident)
    spin0, seeds0, spin1, seeds1, nacc, eng = oracle_run(orc, app, lat, shape, q, kt, seed, n)
    with gpu_init(S, app, lat, shape, q, kt, seed) as L:
        L.set_resident(resident)
        if block:
            L.set_launch(block)
        gs0, gk0 = L.download()
        assert np.array_equal(gs0, spin0), "set site range on device"
        assert np.array_equal(gk0, seeds0), "tagreset seeds on device"
        assert L.energy() == orc.energy(orc.make_desc(app, lat, *shape, q, kt), spin0)
        tot, gn, ge = L.sweeps(n)
        gs1, gk1 = L.download()
    assert np.array_equal(gs1, spin1)
    assert np.array_equal(gk1, seeds1)
    assert np.array_equal(gn, nacc)
    assert np.array_equal(ge, eng)
    assert tot == nacc[-1]


VARIANT_CASES = [
    # proposal variants on the ordered neighbour path: every lattice, every
    # storage type, edge-2 boxes (duplicate neighbours), T = 0
    ("potts/neigh", "sq/8n", (16, 12, 1), 100, 0.5, 71, 6),
    ("potts/neigh", "sq/4n", (18, 10, 1), 5, 1.1, 72, 6),
    ("potts/neigh", "sc/26n", (10, 8, 6), 10000, 0.5, 73, 4),
    ("potts/neigh", "sc/6n", (8, 6, 4), 9, 0.0, 74, 4),
    ("potts/neigh", "sq/8n", (2, 2, 1), 3, 0.8, 75, 5),
    ("potts/neighonly", "sq/8n", (34, 66, 1), 70000, 0.5, 76, 4),
    ("potts/neighonly", "sc/26n", (2, 2, 2), 4, 0.6, 77, 5),
    ("potts/neighonly", "sq/4n", (64, 48, 1), 7, 0.7, 78, 8),
    ("potts/neighonly", "sc/6n", (4, 4, 2), 2, 0.0, 79, 4),
    ("ising/single", "sq/4n", (64, 48, 1), 2, 2.269, 80, 20),
    ("ising/single", "sc/26n", (8, 6, 4), 2, 3.0, 81, 5),
    ("ising/single", "sq/8n", (12, 12, 1), 2, 0.0, 82, 4),
    ("ising/single", "sc/6n", (16, 4, 4), 2, 1.7, 83, 5),
]


@pytest.mark.parametrize("case", VARIANT_CASES, ids=lambda c: "%s-%s-%s" % (c[0], c[1], "x".join(map(str, c[2]))))
def test_variant_bit_exact(orc, case):
    """potts/neigh, potts/neighonly and ising/single against the oracle:
    spins, seeds, cumulative Naccept and energy after every sweep."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed, n = case
    say("variant case", case)
    spin0, seeds0, spin1, seeds1, nacc, eng = oracle_run(orc, app, lat, shape, q, kt, seed, n)
    with gpu_init(S, app, lat, shape, q, kt, seed) as L:
        gs0, gk0 = L.download()
        assert np.array_equal(gs0, spin0) and np.array_equal(gk0, seeds0)
        tot, gn, ge = L.sweeps(n)
        gs1, gk1 = L.download()
    assert np.array_equal(gs1, spin1)
    assert np.array_equal(gk1, seeds1)
    assert np.array_equal(gn, nacc)
    assert np.array_equal(ge, eng)


def test_potts_sq8n_128_known_answers_on_gpu():
    import spparks_amd as S
    g = json.load(open(os.path.join(GOLD, "potts_sq8n_128_known_answers.json")))
    with gpu_init(S, g["app"], g["lattice"], (g["nx"], g["ny"], g["nz"]), g["nspins"],
                  g["temperature"], g["seed"]) as L:
        assert L.energy() == g["initial_energy"]
        tot, nacc, eng = L.sweeps(g["nsweeps"])
    for sweep, na, e in g["rows"]:
        assert nacc[sweep - 1] == na
        assert eng[sweep - 1] == e


def test_upload_download_roundtrip_and_sweeps_from_upload(orc):
    import spparks_amd as S
    d = orc.make_desc("potts", "sc/26n", 12, 10, 8, 300, 0.4)
    spin, seeds = orc.script_init(d, 1234, 1, 300)
    with S.Lattice("potts", "sc/26n", 12, 10, 8, nspins=300, temperature=0.4) as L:
        L.upload(spin, seeds)
        s, k = L.download()
        assert np.array_equal(s, spin) and np.array_equal(k, seeds)
        _, gn, ge = L.sweeps(4)
        s, k = L.download()
    nacc, eng = orc.sweeps(d, spin, seeds, 4)
    assert np.array_equal(s, spin) and np.array_equal(k, seeds)
    assert np.array_equal(gn, nacc) and np.array_equal(ge, eng)


def test_invalid_values_rejected():
    import spparks_amd as S
    with S.Lattice("potts", "sq/8n", 8, 8, nspins=5, temperature=1.0) as L:
        bad = np.full(64, 3, np.int32)
        bad[17] = 6
        with pytest.raises(S.HipError, match="invalid values"):
            L.upload(bad, None)
        bad[17] = 0
        with pytest.raises(S.HipError, match="invalid values"):
            L.upload(bad, None)
        with pytest.raises(S.HipError):
            L.upload(None, np.zeros(64, np.int32))


@pytest.mark.parametrize("block", [64, 128, 256, 1024])
def test_launch_shape_independence(block):
    """colour/strict is decomposition independent: any workgroup size gives
    the same trajectory."""
    import spparks_amd as S
    ref = None
    with gpu_init(S, "potts", "sc/26n", (16, 12, 10), 50, 0.7, 99) as L:
        L.set_resident(False)
        L.set_launch(block)
        _, nacc, eng = L.sweeps(5)
        spin, seeds = L.download()
    with gpu_init(S, "potts", "sc/26n", (16, 12, 10), 50, 0.7, 99) as L:
        _, nacc2, eng2 = L.sweeps(5)
        spin2, seeds2 = L.download()
    assert np.array_equal(spin, spin2) and np.array_equal(seeds, seeds2)
    assert np.array_equal(nacc, nacc2) and np.array_equal(eng, eng2)


@pytest.mark.parametrize("case", [("ising", "sq/4n", (256, 192, 1), 2, 2.269, 5),
                                  ("potts", "sq/8n", (130, 66, 1), 300, 0.6, 6),
                                  ("potts", "sc/26n", (12, 10, 8), 70000, 1.1, 7)],
                         ids=["ising-256x192-planes", "potts-sq8n-130x66", "potts-sc26n-i32"])
def test_graph_pieces_of_every_size_equal_oracle(orc, case):
    """Calls of 1..7, 9, 16, 31 and 33 sweeps (graph replays of 32 and of
    each remainder, some prepared ahead with prepare(), the bit planes kept
    across calls) against one oracle run of the same total, sweep by
    sweep."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed = case
    calls = [1, 2, 3, 4, 5, 6, 7, 9, 16, 31, 33]
    _, _, spin1, seeds1, nacc, eng = oracle_run(orc, app, lat, shape, q, kt, seed, sum(calls))
    with gpu_init(S, app, lat, shape, q, kt, seed) as L:
        L.set_resident(False)
        gn, ge, base = [], [], 0
        for k, n in enumerate(calls):
            if k % 3 == 0:
                L.prepare(n)
            tot, a, e = L.sweeps(n)
            gn += [base + int(x) for x in a]
            ge += list(e)
            base += tot
        gs, gk = L.download()
    assert np.array_equal(np.array(gn), nacc) and np.array_equal(np.array(ge), eng)
    assert np.array_equal(gs, spin1) and np.array_equal(gk, seeds1)


@pytest.mark.parametrize("case", [("ising", "sq/4n", (128, 96, 1), 2, 2.269, 21),
                                  ("potts", "sc/26n", (16, 12, 10), 10000, 0.5, 22),
                                  ("potts", "sq/8n", (66, 34, 1), 70000, 0.7, 23)],
                         ids=["ising-planes", "potts-sc26n", "potts-sq8n-i32"])
def test_graph_replays_equal_eager_launches(case):
    """Chunks of sweeps replayed from captured graphs, several replays per
    call and several calls with eager work and synchronisation in between
    (energy, download, temperature change), against eager launches."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed = case
    runs = []
    for chunk in (32, 0):
        with gpu_init(S, app, lat, shape, q, kt, seed) as L:
            L.set_resident(False)
            L.set_graph(chunk)
            out = []
            for n in (70, 33, 64):
                _, gn, ge = L.sweeps(n)
                out.append((gn, ge, L.energy()) + L.download())
            L.set_temperature(kt * 0.5)
            _, gn, ge = L.sweeps(40)
            out.append((gn, ge, L.energy()) + L.download())
            runs.append(out)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_split_calls_equal_one_call():
    """Consecutive runs continue exactly (seeds persist between runs)."""
    import spparks_amd as S
    with gpu_init(S, "ising", "sq/4n", (32, 32, 1), 2, 2.269, 5) as L:
        _, n1, e1 = L.sweeps(3)
        _, n2, e2 = L.sweeps(4)
        a = L.download()
    with gpu_init(S, "ising", "sq/4n", (32, 32, 1), 2, 2.269, 5) as L:
        _, n, e = L.sweeps(7)
        b = L.download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(np.concatenate([n1, n2 + n1[-1]]), n)
    assert np.array_equal(np.concatenate([e1, e2]), e)


def test_temperature_change_between_runs(orc):
    import spparks_amd as S
    d = orc.make_desc("potts", "sq/8n", 20, 20, 1, 8, 1.5)
    spin, seeds = orc.script_init(d, 6, 1, 8)
    with S.Lattice("potts", "sq/8n", 20, 20, nspins=8, temperature=1.5) as L:
        L.upload(spin, seeds)
        L.sweeps(2)
        L.set_temperature(0.0)
        L.sweeps(2)
        L.set_temperature(0.25)
        _, gn, ge = L.sweeps(2)
        s, k = L.download()
    orc.sweeps(d, spin, seeds, 2)
    for t in (0.0, 0.25):
        d = orc.make_desc("potts", "sq/8n", 20, 20, 1, 8, t)
        nacc, eng = orc.sweeps(d, spin, seeds, 2)
    assert np.array_equal(s, spin) and np.array_equal(k, seeds)
    assert np.array_equal(ge, eng)


# ------------------------------------------------------- full-size shapes

FULL = [
    ("ising", "sq/4n", (4096, 4096, 1), 2, 2.269, 12345, 5),
    ("potts", "sc/26n", (256, 256, 256), 10000, 0.5, 24680, 2),
    # the proposal variants at the sizes they are measured at
    ("potts/neigh", "sc/26n", (256, 256, 256), 10000, 0.5, 24680, 2),
    ("potts/neighonly", "sc/26n", (256, 256, 256), 10000, 0.5, 24680, 2),
    ("ising/single", "sq/4n", (4096, 4096, 1), 2, 2.269, 12345, 5),
]


@pytest.mark.parametrize("case", FULL, ids=["ising-sq4n-4096", "potts-sc26n-256", "potts-neigh-sc26n-256",
                                            "potts-neighonly-sc26n-256", "ising-single-sq4n-4096"])
def test_full_size_first_sweeps_bit_exact(orc, case):
    import spparks_amd as S
    app, lat, shape, q, kt, seed, n = case
    say("full-size case", case)
    spin0, seeds0, spin1, seeds1, nacc, eng = oracle_run(orc, app, lat, shape, q, kt, seed, n)
    with gpu_init(S, app, lat, shape, q, kt, seed) as L:
        gs0, gk0 = L.download()
        assert np.array_equal(gs0, spin0)
        assert np.array_equal(gk0, seeds0)
        _, gn, ge = L.sweeps(n)
        gs1, gk1 = L.download()
        assert L.energy() == eng[-1]
    assert np.array_equal(gn, nacc) and np.array_equal(ge, eng)
    assert np.array_equal(gs1, spin1) and np.array_equal(gk1, seeds1)


with open(os.path.join(GOLD, "full_runs.json")) as _f:
    FULL_RUNS = {r["name"]: r for r in json.load(_f)["runs"]}
with open(os.path.join(GOLD, "max_size_runs.json")) as _f:
    MAX_RUNS = {r["name"]: r for r in json.load(_f)["runs"]}


def run_against_fixture(r, mode):
    """Runs fixture entry r on the GPU (mode: bits, bytes, bits-short,
    bits32 or default) and compares Naccept and energy after every sweep and the state
    checksum every r["every"] sweeps.  bits-short splits each stretch into
    calls of 20 and fewer sweeps (planes kept across them)."""
    import spparks_amd as S
    say("fixture run", r["name"], mode)
    with gpu_init(S, r["app"], r["lattice"], tuple(r["shape"]), r["nspins"], r["kT"], r["seed"]) as L:
        if mode == "bytes":
            L.set_bitplanes(False)
        elif mode == "bits32":   # the per-colour bit-plane kernel, not two sweeps per launch
            L.set_bitplanes(32)
        assert str(L.checksum()) == r["checksum"][0]
        base, nacc, eng = 0, [], []
        for i in range(r["sweeps"] // r["every"]):
            left = r["every"]
            while left:
                n = min(left, 20) if mode == "bits-short" else left
                tot, gn, ge = L.sweeps(n)
                nacc += [int(base + x) for x in gn]
                eng += [int(e) for e in ge]
                base += tot
                left -= n
            assert str(L.checksum()) == r["checksum"][i + 1], "state after %d sweeps" % ((i + 1) * r["every"])
    assert nacc == r["naccept"]
    assert eng == r["energy"]


@pytest.mark.parametrize("name,mode", [("ising-sq4n-4096", "bits"), ("ising-sq4n-4096", "bytes"),
                                       ("ising-sq4n-4096", "bits-short"), ("ising-sq4n-4096", "bits32"),
                                       ("potts-sc26n-256", "default")],
                         ids=["ising-sq4n-4096-bits", "ising-sq4n-4096-bytes",
                              "ising-sq4n-4096-bits-short", "ising-sq4n-4096-bits32", "potts-sc26n-256"])
def test_full_length_runs_match_oracle_fixture(name, mode):
    """The whole 1000-sweep (4096^2 Ising) and 200-sweep (256^3 Potts) runs:
    Naccept and energy after every sweep and the state checksum every
    `every` sweeps equal the oracle's, stored in tests/golden/full_runs.json."""
    run_against_fixture(FULL_RUNS[name], mode)


@pytest.mark.parametrize("name,mode", [("ising-sq4n-32768", "bits"), ("ising-sq4n-32768", "bytes"),
                                       ("potts-sc26n-1024", "default")],
                         ids=["ising-sq4n-32768-bits", "ising-sq4n-32768-bytes", "potts-sc26n-1024"])
def test_maximum_size_lattices_match_oracle_fixture(name, mode):
    """Lattices of 2^30 sites (tests/golden/max_size_runs.json): the seed
    array passes 2^32 bytes, so the per-colour kernels run with 64-bit
    offsets, and the bit planes hold 2^29-site colours."""
    r = MAX_RUNS[name]
    assert r["shape"][0] * r["shape"][1] * r["shape"][2] * 4 >= 1 << 32
    run_against_fixture(r, mode)


BIT_CASES = [
    # two-state sq/4n lattices, which sweep on bit planes by default:
    # sub-grid rows of 1, 9, 32, 33, 63 and 65 sites, Potts Q=2, T = 0
    ("ising", "sq/4n", (2, 2, 1), 2, 2.0, 3, 6),
    ("ising", "sq/4n", (18, 10, 1), 2, 2.269, 5, 7),
    ("ising", "sq/4n", (64, 48, 1), 2, 2.269, 777, 9),
    ("ising", "sq/4n", (66, 10, 1), 2, 1.5, 9, 6),
    ("potts", "sq/4n", (126, 6, 1), 2, 0.9, 10, 5),
    ("ising", "sq/4n", (130, 4, 1), 2, 3.5, 11, 5),
    ("ising", "sq/4n", (40, 34, 1), 2, 0.0, 12, 5),
    ("potts", "sq/4n", (64, 2, 1), 2, 0.0, 13, 4),
    ("ising", "sq/4n", (256, 130, 1), 2, 2.269, 14, 6),
    # ising/single: flip proposals, one draw only for an uphill flip; with
    # planes off it runs the one-site-per-thread variant kernel
    ("ising/single", "sq/4n", (18, 10, 1), 2, 2.269, 14, 7),
    ("ising/single", "sq/4n", (66, 34, 1), 2, 1.2, 15, 6),
    ("ising/single", "sq/4n", (130, 4, 1), 2, 3.5, 16, 5),
    ("ising/single", "sq/4n", (40, 8, 1), 2, 0.0, 17, 5),
    ("ising/single", "sq/4n", (256, 130, 1), 2, 2.269, 18, 6),
]


@pytest.mark.parametrize("planes", [8, 16, 32, 0], ids=["bits8", "bits16", "bits32", "bytes"])
@pytest.mark.parametrize("case", BIT_CASES, ids=lambda c: "%s-%s-q%d-T%g" % (c[0], "x".join(map(str, c[2])), c[3], c[4]))
def test_two_state_bit_planes_bit_exact(orc, case, planes):
    """Two-state sq/4n lattices on bit planes (and with them switched off)
    against the oracle: spins, seeds, per-sweep Naccept and energy."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed, n = case
    say("case", case, "bit planes", planes)
    _, _, spin1, seeds1, nacc, eng = oracle_run(orc, app, lat, shape, q, kt, seed, n)
    with gpu_init(S, app, lat, shape, q, kt, seed) as L:
        L.set_resident(False)
        L.set_bitplanes(planes)
        tot, gn, ge = L.sweeps(n)
        gs1, gk1 = L.download()
    assert np.array_equal(gs1, spin1)
    assert np.array_equal(gk1, seeds1)
    assert np.array_equal(gn, nacc)
    assert np.array_equal(ge, eng)


@pytest.mark.parametrize("shape", [(1026, 512, 1), (2574, 530, 1)], ids=["1026x512", "2574x530-pairs"])
def test_bit_planes_equal_bytes_on_a_large_lattice(shape):
    """1026 x 512 Ising (sub-grid rows of 513 sites: a one-bit last word)
    and 2574 x 530 (41-word rows, which sweep two sweeps per launch by
    default, and a band that does not divide the rows): calls of 13, 32, 0,
    5 and 20 sweeps with the planes kept across calls, and download,
    energy, checksum, set_temperature, a seed upload and a change of sites
    per thread between them: bit planes and the byte kernel give the same
    spins, seeds, Naccept, energies and checksums."""
    import spparks_amd as S
    out = []
    for planes in (1, 8, 32, 0):
        say("bit planes", planes, shape)
        with gpu_init(S, "ising", "sq/4n", shape, 2, 2.269, 4321) as L:
            L.set_bitplanes(planes)
            rows = []
            for n in (13, 32, 0):
                rows.append(L.sweeps(n) if n else None)
            rows.append(L.sweeps(5))
            rows.append(L.download())         # the planes go back to bytes here
            rows.append(L.sweeps(20))
            rows.append(L.energy())
            rows.append(L.sweeps(7))
            rows.append(L.checksum())
            L.set_temperature(1.5)            # graphs dropped, planes kept
            rows.append(L.sweeps(9))
            spin, seeds = L.download()
            L.upload(seed=seeds[::-1].copy())  # seeds rewritten while bytes are current
            rows.append(L.sweeps(4))
            if planes:
                L.set_bitplanes(16 if planes != 16 else 32)   # a new plane layout mid-run
            rows.append(L.sweeps(6))
            spin, seeds = L.download()
            out.append((spin, seeds, rows, L.energy()))
    s2, k2, r2, e2 = out[-1]   # the byte kernel
    for s1, k1, r1, e1 in out[:-1]:
        assert np.array_equal(s1, s2) and np.array_equal(k1, k2)
        assert e1 == e2
        for a, b in zip(r1, r2):
            if a is None:
                continue
            if isinstance(a, tuple):
                for x, y in zip(a, b):
                    assert np.array_equal(np.asarray(x), np.asarray(y))
            else:
                assert a == b


PAIR_CASES = [
    # two sweeps per launch: rows of 64 whole plane words (nx 4096: the
    # kernel's full-row form), of 64 words with a partly padded last one
    # (nx 4094) and of 34 and 41 words, bands that do not divide the rows
    # (ny 66, 70, 80, 530), odd sweeps after the pairs, T = 0, Potts Q=2,
    # ising/single
    ("ising", "sq/4n", (4096, 64, 1), 2, 2.269, 31, 9),
    ("potts", "sq/4n", (4096, 66, 1), 2, 0.9, 38, 5),
    ("ising", "sq/4n", (4096, 64, 1), 2, 0.0, 39, 4),
    ("ising", "sq/4n", (2574, 70, 1), 2, 1.5, 32, 6),
    ("potts", "sq/4n", (2114, 66, 1), 2, 0.9, 33, 5),
    ("ising", "sq/4n", (4094, 64, 1), 2, 0.0, 34, 4),
    ("ising", "sq/4n", (4096, 530, 1), 2, 2.269, 37, 12),
    ("ising/single", "sq/4n", (4096, 80, 1), 2, 2.269, 35, 6),
    ("ising/single", "sq/4n", (2178, 64, 1), 2, 0.0, 36, 4),
    # 16 bands and 37 sweeps: a graph of 32 sweeps (16 launches), then one
    # of 4 sweeps and a per-colour sweep
    ("ising", "sq/4n", (4096, 256, 1), 2, 2.269, 40, 37),
]


@pytest.mark.parametrize("graph", [32, 0], ids=["graphs", "eager"])
@pytest.mark.parametrize("case", PAIR_CASES, ids=lambda c: "%s-%s-q%d-T%g" % (c[0], "x".join(map(str, c[2])), c[3], c[4]))
def test_two_sweeps_per_launch_bit_exact(orc, case, graph):
    """Lattices with 33..64-word plane rows sweep two sweeps per launch by
    default: spins, seeds, per-sweep Naccept and energy against the oracle,
    through graph replay and eager launches."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed, n = case
    say("case", case, "graph", graph)
    _, _, spin1, seeds1, nacc, eng = oracle_run(orc, app, lat, shape, q, kt, seed, n)
    with gpu_init(S, app, lat, shape, q, kt, seed) as L:
        L.set_graph(graph)
        tot, gn, ge = L.sweeps(n)
        gs1, gk1 = L.download()
    assert np.array_equal(gn, nacc)
    assert np.array_equal(ge, eng)
    assert np.array_equal(gs1, spin1)
    assert np.array_equal(gk1, seeds1)


@pytest.mark.parametrize("case", [("potts", "sq/8n", (16, 12, 1), 100, 0.5, 56789),
                                  ("ising", "sq/4n", (66, 10, 1), 2, 2.269, 9),
                                  ("potts", "sc/26n", (10, 8, 6), 70000, 0.6, 3)],
                         ids=["potts-sq8n-u8", "ising-sq4n-ragged", "potts-sc26n-i32"])
def test_resident_kernel_equals_per_colour_kernels(case):
    """1100 sweeps in one call (more than one launch's 1024 accumulator
    rows), then 7 more: the one-workgroup LDS kernel and the per-colour
    kernels give the same spins, seeds, Naccept and energies."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed = case
    out = []
    for resident in (1, 0):
        say("case", case, "resident", resident)
        with gpu_init(S, app, lat, shape, q, kt, seed) as L:
            L.set_resident(resident)
            a = L.sweeps(1100)
            b = L.sweeps(7)
            out.append((a, b) + L.download())
    (a1, b1, s1, k1), (a2, b2, s2, k2) = out
    assert np.array_equal(s1, s2) and np.array_equal(k1, k2)
    for x, y in ((a1, a2), (b1, b2)):
        assert x[0] == y[0] and np.array_equal(x[1], y[1]) and np.array_equal(x[2], y[2])


def _random_cases(n=16, seed=20261202):
    """Seeded random lattices: style, app, Q (all three storage types), even
    edges, kT (including 0), sweep counts."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        lat = ["sq/4n", "sq/8n", "sc/6n", "sc/26n"][k % 4]
        app = "ising" if rng.random() < 0.3 else "potts"
        q = 2 if app == "ising" else int(rng.choice([2, 3, 17, 255, 300, 70000]))
        if lat.startswith("sq"):
            shape = (2 * int(rng.integers(1, 40)), 2 * int(rng.integers(1, 24)), 1)
        else:
            shape = (2 * int(rng.integers(1, 9)), 2 * int(rng.integers(1, 7)), 2 * int(rng.integers(1, 6)))
        kt = 0.0 if rng.random() < 0.15 else round(float(rng.uniform(0.2, 3.5)), 3)
        out.append((app, lat, shape, q, kt, int(rng.integers(1, 10 ** 6)), int(rng.integers(2, 6))))
    return out


@pytest.mark.parametrize("case", _random_cases(), ids=lambda c: "%s-%s-%s-q%d-T%g" % (c[0], c[1], "x".join(map(str, c[2])), c[3], c[4]))
def test_random_lattices_every_path_bit_exact(orc, case):
    """Seeded random cases through every path that applies to them (the
    one-workgroup LDS kernel, the per-colour kernels, and for two-state sq/4n
    the bit planes) against the oracle."""
    import spparks_amd as S
    app, lat, shape, q, kt, seed, n = case
    _, _, spin1, seeds1, nacc, eng = oracle_run(orc, app, lat, shape, q, kt, seed, n)
    paths = [(1, 0), (0, 0)]
    if lat == "sq/4n" and q == 2:
        paths.append((0, 1))
    for resident, planes in paths:
        say("case", case, "resident", resident, "bit planes", planes)
        with gpu_init(S, app, lat, shape, q, kt, seed) as L:
            L.set_resident(resident)
            L.set_bitplanes(planes)
            _, gn, ge = L.sweeps(n)
            gs, gk = L.download()
        assert np.array_equal(gs, spin1) and np.array_equal(gk, seeds1), (resident, planes)
        assert np.array_equal(gn, nacc) and np.array_equal(ge, eng), (resident, planes)


def test_two_handles_from_two_threads():
    """Distinct handles driven from distinct threads at once (the library
    calls release the GIL) give the same states as the same runs one after
    the other; each thread's error text is its own."""
    import threading
    import spparks_amd as S
    cases = [("ising", "sq/4n", (512, 384, 1), 2, 2.269, 21, 60),
             ("potts", "sc/26n", (32, 24, 16), 5000, 0.6, 22, 25)]

    def run(case, out, k):
        app, lat, shape, q, kt, seed, n = case
        with gpu_init(S, app, lat, shape, q, kt, seed) as L:
            for _ in range(3):
                L.sweeps(n // 3, per_sweep=False)
            try:
                L.set_launch(100)
            except S.HipError as e:
                out[k] = (L.checksum(), str(e))

    alone = [None, None]
    for k, c in enumerate(cases):
        run(c, alone, k)
    both = [None, None]
    say("two threads")
    ts = [threading.Thread(target=run, args=(c, both, k)) for k, c in enumerate(cases)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in ts)
    assert both == alone
    assert all("power of two" in e for _, e in both)


def test_argument_edges(orc):
    """sweeps(0) is a no-op; a negative sweep count, workgroup sizes that are
    not powers of two in 64..1024 and graph chunks outside 0..1024 are
    rejected; the lattice then runs on exactly as the oracle does."""
    import spparks_amd as S
    case = ("potts", "sq/8n", (16, 12, 1), 100, 0.5, 56789)
    spin0, seeds0, spin1, seeds1, nacc, eng = oracle_run(orc, *case, 3)
    with gpu_init(S, *case) as L:
        assert L.sweeps(0) == (0, None, None)
        gs, gk = L.download()
        assert np.array_equal(gs, spin0) and np.array_equal(gk, seeds0)
        with pytest.raises(S.HipError, match="nsweeps"):
            L.sweeps(-1)
        for b in (0, 32, 100, 2048):
            with pytest.raises(S.HipError, match="power of two"):
                L.set_launch(b)
        for c in (-8, -1, 1025):
            with pytest.raises(S.HipError, match="chunk_sweeps"):
                L.set_graph(c)
        L.set_graph(7)
        _, gn, ge = L.sweeps(3)
        gs, gk = L.download()
    assert np.array_equal(gs, spin1) and np.array_equal(gk, seeds1)
    assert np.array_equal(gn, nacc) and np.array_equal(ge, eng)


def test_bitplane_setting_rejects_bad_values():
    import spparks_amd as S
    with gpu_init(S, "ising", "sq/4n", (64, 48, 1), 2, 2.269, 1) as L:
        for bad in (3, 64, -1):
            with pytest.raises(S.HipError, match="bit planes"):
                L.set_bitplanes(bad)
        L.set_bitplanes(16)
        L.set_resident(False)
        L.sweeps(2)
