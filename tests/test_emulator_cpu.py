"""The kernels' own index math and site rule (spk_device.h), run on the host by
tests/emul/spk_emul under AddressSanitizer, against the oracle.  Catches
layout, wrap and neighbour-set mistakes before anything reaches a GPU."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMUL = os.path.join(HERE, "emul", "spk_emul")

CASES = [
    # sub-grid rows of every length mod 4 (pads of 3, 2, 1, 0 slots)
    ("ising", "sq/4n", (2, 6, 1), 2, 1.8, 71, 4),
    ("ising", "sq/4n", (4, 6, 1), 2, 1.8, 72, 4),
    ("ising", "sq/4n", (6, 4, 1), 2, 1.8, 73, 4),
    ("ising", "sq/4n", (10, 6, 1), 2, 1.8, 74, 4),
    ("potts", "sc/26n", (6, 4, 2), 30, 0.7, 75, 3),
    ("potts", "sc/26n", (10, 6, 4), 3000, 0.5, 76, 3),
    ("potts", "sc/26n", (14, 2, 6), 70000, 0.3, 77, 3),
    ("potts", "sq/8n", (22, 6, 1), 9, 0.9, 78, 3),
    ("potts", "sc/6n", (6, 4, 4), 9, 0.9, 79, 3),
    ("potts", "sq/8n", (16, 8, 1), 300, 0.7, 1, 4),
    ("potts", "sc/26n", (8, 8, 4), 70000, 0.4, 2, 3),
    ("potts", "sq/4n", (16, 6, 1), 65535, 0.9, 3, 3),
    ("ising", "sc/6n", (16, 4, 4), 2, 1.7, 4, 3),
    ("potts", "sq/8n", (16, 12, 1), 100, 0.5, 56789, 5),
    ("potts", "sq/8n", (2, 2, 1), 3, 0.8, 11, 4),
    ("potts", "sq/8n", (2, 6, 1), 3, 0.8, 12, 4),
    ("potts", "sq/4n", (18, 10, 1), 5, 1.1, 5, 4),
    ("potts", "sq/4n", (2, 4, 1), 5, 1.1, 6, 4),
    ("ising", "sq/4n", (64, 48, 1), 2, 2.269, 777, 10),
    ("ising", "sq/4n", (2, 2, 1), 2, 2.0, 3, 4),
    ("potts", "sc/26n", (10, 8, 6), 10000, 0.5, 4242, 4),
    ("potts", "sc/26n", (2, 2, 2), 4, 0.6, 8, 4),
    ("potts", "sc/26n", (2, 4, 6), 4, 0.6, 9, 4),
    ("potts", "sc/26n", (6, 4, 2), 7, 0.0, 91, 4),
    ("ising", "sc/26n", (8, 6, 4), 2, 3.0, 17, 4),
    ("potts", "sc/6n", (8, 6, 4), 9, 0.9, 23, 4),
    ("potts", "sc/6n", (2, 2, 2), 9, 0.9, 24, 4),
    ("ising", "sc/6n", (4, 4, 2), 2, 1.0, 29, 4),
    ("ising", "sq/8n", (12, 12, 1), 2, 0.0, 37, 4),
    # Potts with two states takes the two-state shortcut too
    ("potts", "sq/4n", (10, 6, 1), 2, 0.8, 41, 4),
    ("potts", "sc/26n", (6, 4, 2), 2, 1.2, 42, 4),
    # proposal variants (ordered neighbour path), every lattice, edge-2 boxes
    # (duplicate neighbours), T = 0
    ("potts/neigh", "sq/4n", (10, 6, 1), 5, 0.8, 51, 4),
    ("potts/neigh", "sq/8n", (2, 6, 1), 4, 0.6, 52, 4),
    ("potts/neigh", "sc/26n", (6, 4, 4), 300, 0.5, 53, 3),
    ("potts/neigh", "sc/6n", (4, 4, 2), 7, 0.0, 54, 3),
    ("potts/neighonly", "sq/8n", (16, 12, 1), 100, 0.5, 55, 4),
    ("potts/neighonly", "sc/26n", (2, 2, 2), 3, 0.7, 56, 4),
    ("potts/neighonly", "sq/4n", (18, 10, 1), 70000, 1.1, 57, 3),
    ("potts/neighonly", "sc/6n", (8, 6, 4), 9, 0.0, 58, 3),
    ("ising/single", "sq/4n", (64, 48, 1), 2, 2.269, 59, 5),
    ("ising/single", "sc/26n", (8, 6, 4), 2, 3.0, 60, 4),
    ("ising/single", "sq/8n", (12, 12, 1), 2, 0.0, 61, 4),
    ("ising/single", "sc/6n", (2, 4, 2), 2, 1.5, 62, 4),
    # byte lanes holding values >= 0x80 (u8 storage up to Q = 255)
    ("potts", "sc/26n", (6, 4, 4), 255, 0.6, 43, 4),
    ("potts", "sq/8n", (10, 6, 1), 200, 0.7, 44, 4),
]


def storage_variants(app, q, nx):
    """(stype, vec) pairs the library may pick for this case, plus wider
    storage types, which must give the same trajectory."""
    qmax = q if app.startswith("potts") else 2
    narrow = 0 if qmax <= 255 else 1 if qmax <= 65535 else 2
    return [(t, 4) for t in range(narrow, 3)]


def run_emul(orc, tmp_path, case, env_extra=None, variants=None):
    app, lat, shape, q, kt, seed, n = case
    d = orc.make_desc(app, lat, *shape, q, kt)
    spin, seeds = orc.script_init(d, seed, 1, q)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(spin.tobytes())
        f.write(seeds.tobytes())
    nacc, eng = orc.sweeps(d, spin, seeds, n)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", **(env_extra or {}))
    for stype, vec in variants or storage_variants(app, q, shape[0]):
        r = subprocess.run([EMUL, str(orc.APP[app]), str(orc.LATTICE[lat]), *map(str, shape), str(q),
                            repr(kt), str(n), str(stype), str(vec), str(fin), str(fout)],
                           env=env, capture_output=True, text=True)
        assert r.returncode == 0, (stype, vec, r.stderr[-2000:])
        N = spin.size
        raw = open(fout, "rb").read()
        assert np.array_equal(np.frombuffer(raw, np.int32, N, 0), spin), (stype, vec)
        assert np.array_equal(np.frombuffer(raw, np.int32, N, 4 * N), seeds), (stype, vec)
        assert np.array_equal(np.frombuffer(raw, np.int64, n, 8 * N), nacc), (stype, vec)
        assert np.array_equal(np.frombuffer(raw, np.int64, n, 8 * N + 8 * n).astype(np.float64), eng), (stype, vec)


BIT_CASES = [
    # two-state sq/4n on bit planes: sub-grid rows of 1, 9, 32 (one full
    # word), 33 (a 1-bit last word), 63 and 65 sites; Potts Q=2; T = 0
    ("ising", "sq/4n", (2, 2, 1), 2, 2.0, 3, 6),
    ("ising", "sq/4n", (18, 10, 1), 2, 2.269, 5, 7),
    ("ising", "sq/4n", (64, 48, 1), 2, 2.269, 777, 9),
    ("ising", "sq/4n", (66, 10, 1), 2, 1.5, 9, 6),
    ("potts", "sq/4n", (126, 6, 1), 2, 0.9, 10, 5),
    ("ising", "sq/4n", (130, 4, 1), 2, 3.5, 11, 5),
    ("ising", "sq/4n", (40, 34, 1), 2, 0.0, 12, 5),
    ("potts", "sq/4n", (64, 2, 1), 2, 0.0, 13, 4),
    # ising/single (flip proposals, one draw only when uphill)
    ("ising/single", "sq/4n", (18, 10, 1), 2, 2.269, 14, 7),
    ("ising/single", "sq/4n", (66, 34, 1), 2, 1.2, 15, 6),
    ("ising/single", "sq/4n", (130, 4, 1), 2, 3.5, 16, 5),
    ("ising/single", "sq/4n", (40, 8, 1), 2, 0.0, 17, 5),
]


@pytest.mark.parametrize("spt", [8, 16, 32])
@pytest.mark.parametrize("case", BIT_CASES, ids=lambda c: "%s-%s-q%d-T%g" % (c[0], "x".join(map(str, c[2])), c[3], c[4]))
def test_emulated_bit_planes_match_oracle(orc, tmp_path, case, spt):
    """Two-state sq/4n lattices on bit planes (32 sites per word, shadow
    seeds interleaved by wavefront): row ends inside a word, single-word
    rows, one-bit last words, the seed shadow's one-to-one map, and the
    incremental energy after every sweep."""
    if not os.path.exists(EMUL):
        pytest.skip("tests/emul/spk_emul not built")
    run_emul(orc, tmp_path, case, {"SPK_EMUL_BITS": str(spt)}, variants=[(0, 4)])


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s-q%d" % (c[0], c[1], "x".join(map(str, c[2])), c[3]))
def test_emulated_kernel_matches_oracle(orc, tmp_path, case):
    if not os.path.exists(EMUL):
        pytest.skip("tests/emul/spk_emul not built (run __graft_entry__.build())")
    app, lat, shape, q, kt, seed, n = case
    d = orc.make_desc(app, lat, *shape, q, kt)
    spin, seeds = orc.script_init(d, seed, 1, q)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(spin.tobytes())
        f.write(seeds.tobytes())
    nacc, eng = orc.sweeps(d, spin, seeds, n)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    for stype, vec in storage_variants(app, q, shape[0]):
        r = subprocess.run([EMUL, str(orc.APP[app]), str(orc.LATTICE[lat]), *map(str, shape), str(q),
                            repr(kt), str(n), str(stype), str(vec), str(fin), str(fout)],
                           env=env, capture_output=True, text=True)
        assert r.returncode == 0, (stype, vec, r.stderr[-2000:])
        N = spin.size
        raw = open(fout, "rb").read()
        es = np.frombuffer(raw, np.int32, N, 0)
        ek = np.frombuffer(raw, np.int32, N, 4 * N)
        en = np.frombuffer(raw, np.int64, n, 8 * N)
        ee = np.frombuffer(raw, np.int64, n, 8 * N + 8 * n)
        assert np.array_equal(es, spin), (stype, vec)
        assert np.array_equal(ek, seeds), (stype, vec)
        assert np.array_equal(en, nacc), (stype, vec)
        assert np.array_equal(ee.astype(np.float64), eng), (stype, vec)


PAIR_CASES = [
    # two sweeps per launch (k_sweep_bits_pair): rows of 64 whole plane
    # words (nx 4096: the kernel's full-row form), of 64 words with the last
    # one partly padding (nx 4094) and of 34 and 41 words, bands that do not
    # divide the rows (ny 66, 70, 80), the halo wrapping round the lattice,
    # an odd sweep after the pairs, T = 0, Potts Q=2
    ("ising", "sq/4n", (4096, 64, 1), 2, 2.269, 31, 5),
    ("potts", "sq/4n", (4096, 66, 1), 2, 0.9, 38, 4),
    ("ising", "sq/4n", (4096, 64, 1), 2, 0.0, 39, 3),
    ("ising", "sq/4n", (2574, 70, 1), 2, 1.5, 32, 4),
    ("potts", "sq/4n", (2114, 66, 1), 2, 0.9, 33, 3),
    ("ising", "sq/4n", (4094, 64, 1), 2, 0.0, 34, 4),
    ("ising/single", "sq/4n", (4096, 80, 1), 2, 2.269, 35, 6),
    ("ising/single", "sq/4n", (2178, 64, 1), 2, 0.0, 36, 2),
]


@pytest.mark.parametrize("case", PAIR_CASES, ids=lambda c: "%s-%s-q%d-T%g" % (c[0], "x".join(map(str, c[2])), c[3], c[4]))
def test_emulated_two_sweeps_per_launch_match_oracle(orc, tmp_path, case):
    """The two-sweep band kernel's steps (pair_thread, the LDS planes and
    seed ring, the band and halo rows) run stage by stage on the host
    against the oracle: spins, seeds, Naccept and energy after every sweep
    (sweep 1's counts from stages 0-1, sweep 2's from stages 2-3)."""
    if not os.path.exists(EMUL):
        pytest.skip("tests/emul/spk_emul not built")
    run_emul(orc, tmp_path, case, {"SPK_EMUL_BITS": "pair"}, variants=[(0, 4)])
