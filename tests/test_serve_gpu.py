"""Serving on the GPU: continuous batching of a request stream.  Every
request's result is the same as solving it alone; admission respects the
HBM budget; device checksums equal the host restatement."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_stream(n, rate, seed=3):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        L = int(2 * rng.integers(4, 13))
        q = int(rng.choice([50, 3000, 100000]))
        reqs.append({"id": i, "arrive": float(t[i] - t[0]), "app": "potts", "lattice": "sc/26n",
                     "shape": [L, L, L], "nspins": q, "kT": round(float(rng.uniform(0.1, 1.0)), 4),
                     "seed": 10**6 + i, "sweeps": int(rng.integers(3, 25))})
    return reqs


def solo(S, r):
    with S.Lattice(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"], temperature=r["kT"]) as L:
        s0, rseed = S.strict_init_params(r["seed"])
        L.set_range(s0, 1, r["nspins"])
        L.init_seeds(rseed)
        nacc, _, _ = L.sweeps(r["sweeps"], per_sweep=False)
        return L.checksum(), nacc, L.energy(), L.download()


def test_checksum_matches_host_restatement():
    import spparks_amd as S
    from spparks_amd import batch as bmod
    with S.Lattice("potts", "sc/26n", 10, 6, 4, nspins=70000, temperature=0.4) as L:
        s0, rseed = S.strict_init_params(9)
        L.set_range(s0, 1, 70000)
        L.init_seeds(rseed)
        L.sweeps(2)
        sp, sd = L.download()
        assert L.checksum() == S.checksum(sp, sd)
    with bmod.Batch(16 << 20, 4) as B:
        rid = B.add(bmod.replica_desc("ising", "sq/4n", 22, 10, nspins=2, temperature=2.0, seed=4, nsweeps=3))
        B.set_launch(64, 8)
        B.run(3)
        sp, sd = B.download(rid)
        assert B.checksum(rid) == S.checksum(sp, sd)


def test_stream_results_equal_solo_runs(orc):
    import spparks_amd as S
    from spparks_amd import batch as bmod
    reqs = small_stream(24, rate=200.0)
    say("serving", len(reqs), "requests")
    res, st = bmod.serve(reqs, pool_bytes=256 << 20, max_live=64, block_sweeps=4, block=128, grid=64)
    assert st["site_updates"] == sum(np.prod(r["shape"]) * r["sweeps"] for r in reqs)
    for r, o in zip(reqs, res):
        assert o["status"] == 0
        assert o["admit"] >= r["arrive"] - 1e-9 and o["finish"] >= o["admit"]
        cs, nacc, e, _ = solo(S, r)
        assert (o["checksum"], o["naccept"], o["energy"]) == (cs, nacc, e), r
    # and two of them against the oracle directly
    for r in reqs[:2]:
        d = orc.make_desc("potts", "sc/26n", *r["shape"], r["nspins"], r["kT"])
        spin, seeds = orc.script_init(d, r["seed"], 1, r["nspins"])
        orc.sweeps(d, spin, seeds, r["sweeps"])
        assert res[r["id"]]["checksum"] == S.checksum(spin, seeds)


def test_invalid_request_is_refused_alone():
    """Requests that are not valid solves (an odd edge: SPK_HIP_EINVAL; a
    set range beyond Q: SPK_HIP_EVALUE) are refused alone; the requests
    around them are served as usual."""
    import spparks_amd as S
    from spparks_amd import batch as bmod
    reqs = small_stream(5, rate=1000.0, seed=8)
    odd = dict(reqs[1], shape=[9, 8, 8])
    beyond_q = dict(reqs[3], hi=reqs[3]["nspins"] + 1)   # set site range 1 Q+1
    stream = [reqs[0], odd, reqs[2], beyond_q, reqs[4]]
    bad = (odd, beyond_q)
    say("serving with two invalid requests")
    res, st = bmod.serve(stream, pool_bytes=64 << 20, max_live=8, block_sweeps=4, block=128, grid=64)
    assert [o["status"] for o in res] == [0, -1, 0, -5, 0]
    assert st["site_updates"] == sum(np.prod(r["shape"]) * r["sweeps"] for r in stream if r not in bad)
    for r, o in zip(stream, res):
        if r not in bad:
            cs, nacc, e, _ = solo(S, r)
            assert (o["checksum"], o["naccept"], o["energy"]) == (cs, nacc, e)


def test_admission_under_a_tight_budget():
    import spparks_amd as S
    from spparks_amd import batch as bmod
    reqs = small_stream(12, rate=1000.0, seed=5)
    biggest = max(bmod.replica_bytes("potts", "sc/26n", *r["shape"], nspins=r["nspins"]) for r in reqs)
    # room for about three requests at a time, plus one request that can never fit
    huge = dict(reqs[-1], id=len(reqs), shape=[64, 64, 64], arrive=reqs[-1]["arrive"] + 1e-3)
    res, st = bmod.serve(reqs + [huge], pool_bytes=3 * biggest, max_live=8, block_sweeps=5, block=128, grid=64)
    assert res[-1]["status"] == -3
    assert st["max_pool_used"] <= 3 * biggest and st["max_live"] >= 1
    for r, o in zip(reqs, res):
        assert o["status"] == 0
        cs, nacc, e, _ = solo(S, r)
        assert (o["checksum"], o["naccept"], o["energy"]) == (cs, nacc, e)
