"""Batched replicas on the GPU: every replica's trajectory in a mixed batch
equals the oracle's for that replica alone, for eager, graph and autotuned
launches; pool admission, retirement and slot reuse."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# app, lattice, shape, Q, kT, seed, sweeps -- every lattice style and spin
# storage type, rows of every length mod 4, different sweep counts
MIX = [
    ("ising", "sq/4n", (10, 6, 1), 2, 2.0, 101, 7),
    ("ising", "sq/4n", (64, 48, 1), 2, 2.269, 102, 9),
    ("potts", "sq/8n", (22, 6, 1), 9, 0.9, 103, 5),
    ("potts", "sq/8n", (16, 12, 1), 300, 0.5, 104, 6),
    ("potts", "sc/26n", (10, 6, 4), 3000, 0.5, 105, 4),
    ("potts", "sc/26n", (14, 2, 6), 70000, 0.3, 106, 5),
    ("potts", "sc/26n", (8, 8, 8), 100, 0.7, 107, 3),
    ("potts", "sc/6n", (6, 4, 4), 9, 0.9, 108, 6),
    ("ising", "sc/26n", (6, 6, 6), 2, 3.0, 109, 4),
    ("potts", "sq/4n", (130, 6, 1), 1000000, 0.15, 110, 2),
    ("potts", "sc/26n", (2, 2, 2), 4, 0.6, 111, 5),
    ("ising", "sq/8n", (12, 12, 1), 2, 0.0, 112, 4),
]


def oracle_traj(orc, case):
    app, lat, shape, q, kt, seed, n = case
    d = orc.make_desc(app, lat, *shape, q, kt)
    spin, seeds = orc.script_init(d, seed, 1, q)
    e0 = orc.energy(d, spin)
    nacc, eng = orc.sweeps(d, spin, seeds, n)
    return e0, spin, seeds, nacc, eng


def add_all(B, bmod, cases):
    ids = []
    for (app, lat, shape, q, kt, seed, n) in cases:
        rid = B.add(bmod.replica_desc(app, lat, *shape, nspins=q, temperature=kt, seed=seed, nsweeps=n))
        assert rid is not None
        ids.append(rid)
    return ids


@pytest.mark.parametrize("mode", ["eager", "graph", "autotune"])
def test_mixed_batch_matches_oracle(orc, mode):
    from spparks_amd import batch as bmod
    say("batch mode", mode)
    ref = [oracle_traj(orc, c) for c in MIX]
    with bmod.Batch(64 << 20, 32) as B:
        ids = add_all(B, bmod, MIX)
        for rid, r in zip(ids, ref):
            assert B.status(rid)[3] == r[0]          # initial energy
        if mode != "autotune":
            B.set_launch(256, 64)
        # blocks of 3 sweeps: replicas finish in different blocks
        for _ in range(4):
            B.run(3, graph=(mode == "graph"))
        for rid, case, (e0, spin, seeds, nacc, eng) in zip(ids, MIX, ref):
            done, left, na, e = B.status(rid)
            assert done == case[6] and left == 0
            gn, ge = B.trace(rid, case[6])
            assert np.array_equal(gn, nacc), case
            assert np.array_equal(ge, eng), case
            gs, gk = B.download(rid)
            assert np.array_equal(gs, spin), case
            assert np.array_equal(gk, seeds), case


def test_pool_admission_and_slot_reuse(orc):
    from spparks_amd import batch as bmod
    case = ("potts", "sc/26n", (8, 8, 8), 100, 0.7, 107, 3)
    nb = bmod.replica_bytes("potts", "sc/26n", 8, 8, 8, 100)
    with bmod.Batch(3 * nb, 8) as B:
        a = B.add(bmod.replica_desc("potts", "sc/26n", 8, 8, 8, nspins=100, temperature=0.7, seed=1, nsweeps=2))
        b = B.add(bmod.replica_desc("potts", "sc/26n", 8, 8, 8, nspins=100, temperature=0.7, seed=2, nsweeps=2))
        c = B.add(bmod.replica_desc("potts", "sc/26n", 8, 8, 8, nspins=100, temperature=0.7, seed=3, nsweeps=2))
        assert None not in (a, b, c)
        # full: admission refused, batch unchanged
        assert B.add(bmod.replica_desc("potts", "sc/26n", 8, 8, 8, nspins=100, temperature=0.7, seed=4,
                                       nsweeps=2)) is None
        assert B.pool()[2] == 3
        B.set_launch(128, 16)
        B.run(2)
        B.remove(b)
        assert B.pool()[2] == 2
        # the freed extent and slot take a new replica, which then matches the oracle
        app, lat, shape, q, kt, seed, n = case
        d = B.add(bmod.replica_desc(app, lat, *shape, nspins=q, temperature=kt, seed=seed, nsweeps=n))
        assert d is not None
        B.run(5)
        e0, spin, seeds, nacc, eng = oracle_traj(orc, case)
        gs, gk = B.download(d)
        assert np.array_equal(gs, spin) and np.array_equal(gk, seeds)
        assert np.array_equal(B.trace(d, n)[1], eng)
        # the survivors were not disturbed by the churn
        for rid, sd in ((a, 1), (c, 3)):
            o = oracle_traj(orc, ("potts", "sc/26n", (8, 8, 8), 100, 0.7, sd, 2))
            assert np.array_equal(B.download(rid)[0], o[1])


def test_batch_upload_then_run(orc):
    from spparks_amd import batch as bmod
    app, lat, shape, q, kt = "potts", "sq/8n", (14, 10, 1), 50, 0.6
    d = orc.make_desc(app, lat, *shape, q, kt)
    spin, seeds = orc.script_init(d, 77, 1, q)
    with bmod.Batch(8 << 20, 4) as B:
        rid = B.add(bmod.replica_desc(app, lat, *shape, nspins=q, temperature=kt, seed=5, nsweeps=4))
        B.upload(rid, spin, seeds)
        B.set_launch(64, 8)
        B.run(4, graph=True)
        nacc, eng = orc.sweeps(d, spin, seeds, 4)
        gs, gk = B.download(rid)
    assert np.array_equal(gs, spin) and np.array_equal(gk, seeds)


def test_batch_refuses_proposal_variants():
    from spparks_amd import batch as bmod
    import spparks_amd as S
    with bmod.Batch(8 << 20, 4) as B:
        with pytest.raises(S.HipError, match="potts and ising"):
            B.add(bmod.replica_desc("potts/neigh", "sq/8n", 8, 8, 1, nspins=5, temperature=0.5, seed=3, nsweeps=2))
