"""The batched-ensemble and serving workloads on the GPU at their configured
sizes: every one of the 512 ensemble replicas, advanced together in
autotuned and graph-replayed blocks, and every request of a 200 req/s
serving stream admitted into a 200 GB pool, ends exactly as the same
script run alone through the single-lattice sweep; a few are also checked
against the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def alone(S, r, n):
    """The replica's script through the plain single-lattice sweep: per-sweep
    Naccept and energy, checksum, final energy."""
    with S.Lattice(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"], temperature=r["kT"]) as L:
        s0, rseed = S.strict_init_params(r["seed"])
        L.set_range(s0, 1, r["nspins"])
        L.init_seeds(rseed)
        _, nacc, eng = L.sweeps(n)
        return nacc, eng, L.checksum(), L.energy()


def oracle_state(orc, r, n):
    d = orc.make_desc(r["app"], r["lattice"], *r["shape"], r["nspins"], r["kT"])
    spin, seeds = orc.script_init(d, r["seed"], 1, r["nspins"])
    nacc, eng = orc.sweeps(d, spin, seeds, n)
    return spin, seeds, nacc, eng


def test_ensemble_512_replicas_equal_their_runs_alone(orc):
    """The 512-replica ensemble (256 Ising sq/4n L 64-1024, 256 Potts sc/26n
    L 32-128, Q 100-1e5, seeds 1000+r), 25 sweeps each: a 10-sweep block
    under the autotuner, then graph-replayed 10-sweep blocks.  Every
    replica's per-sweep Naccept and energy, checksum and energy equal its
    run alone; the smallest Ising, the smallest Potts, the smallest int32
    Potts (Q > 65535) and the largest Ising also equal the oracle."""
    import spparks_amd as S
    from spparks_amd import batch as bmod
    from spparks_amd.workloads import ensemble_replicas
    n = 25
    reps = ensemble_replicas(sweeps=n)
    assert len(reps) == 512
    descs = [bmod.replica_desc(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"], temperature=r["kT"],
                               seed=r["seed"], nsweeps=n) for r in reps]
    pool = sum(bmod.replica_bytes(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"]) for r in reps)
    size = lambda r: int(np.prod(r["shape"]))
    ising = [r for r in reps if r["app"] == "ising"]
    potts = [r for r in reps if r["app"] == "potts"]
    picks = [min(ising, key=size), min(potts, key=size),
             min((r for r in potts if r["nspins"] > 65535), key=size), max(ising, key=size)]
    say("ensemble: %d replicas, %d sites; oracle picks %s" % (len(reps), sum(map(size, reps)),
                                                              [(p["id"], p["shape"][0]) for p in picks]))
    with bmod.Batch(pool, len(reps)) as B:
        ids = [B.add(d) for d in descs]
        assert None not in ids
        B.set_launch(0, 0)
        B.run(10, graph=False)      # the autotuner times launch shapes on this block's steps
        tuned = B.last_timing()
        assert tuned["block"] > 0 and tuned["grid"] > 0
        B.run(10, graph=True)
        B.run(10, graph=True)       # 5 sweeps left: a second graph shape
        got = {}
        for rid, r in zip(ids, reps):
            done, left, na, e = B.status(rid)
            assert done == n and left == 0
            gn, ge = B.trace(rid, n)
            got[r["id"]] = (gn, ge, B.checksum(rid), e)
        picked = {p["id"]: B.download(ids[p["id"]]) for p in picks}
    say("batched done; alone runs")
    bad = []
    for r in reps:
        nacc, eng, ck, e = alone(S, r, n)
        gn, ge, gck, ge_end = got[r["id"]]
        if not (np.array_equal(gn, nacc) and np.array_equal(ge, eng) and gck == ck and ge_end == e):
            bad.append(r["id"])
    assert bad == [], "replicas differing from their runs alone: %s" % bad[:20]
    for p in picks:
        say("oracle", p["id"], p["app"], p["shape"], p["nspins"])
        spin, seeds, nacc, eng = oracle_state(orc, p, n)
        gs, gk = picked[p["id"]]
        assert np.array_equal(gs, spin) and np.array_equal(gk, seeds), p
        assert np.array_equal(got[p["id"]][0], nacc) and np.array_equal(got[p["id"]][1], eng), p


def test_serving_stream_results_equal_solo_runs(orc):
    """A 200 req/s Poisson stream of 40 sc/26n Potts grain-growth requests
    (edges 32-192, Q 100-1e6, kT 0.1-1.0, 50-500 sweeps, seeds 10^6+i),
    continuously batched into a 200 GB pool: requests run together
    (several live at once) and each result equals the request solved
    alone; the two smallest also equal the oracle."""
    import spparks_amd as S
    from spparks_amd import batch as bmod
    from spparks_amd.workloads import serving_requests
    reqs = serving_requests(200.0, 40)
    say("serving: %d requests, %.2f G site updates" % (len(reqs), sum(np.prod(r["shape"]) * r["sweeps"]
                                                                    for r in reqs) / 1e9))
    res, st = bmod.serve(reqs, pool_bytes=200 * 10**9, max_live=4096, block_sweeps=10, block=256, grid=1024)
    assert all(o["status"] == 0 for o in res)
    assert st["max_live"] > 1, st
    say("served in %.2f s, %d blocks, up to %d live; solo runs" % (st["wall"], st["blocks"], st["max_live"]))
    bad = []
    for r, o in zip(reqs, res):
        nacc, eng, ck, e = alone(S, r, r["sweeps"])
        if (ck, int(nacc[-1]), e) != (o["checksum"], o["naccept"], o["energy"]):
            bad.append(r["id"])
    assert bad == [], "requests differing from their solo runs: %s" % bad
    work = lambda r: int(np.prod(r["shape"])) * r["sweeps"]
    for r in sorted(reqs, key=work)[:2]:
        say("oracle", r["id"], r["shape"], r["nspins"], r["sweeps"])
        spin, seeds, nacc, eng = oracle_state(orc, r, r["sweeps"])
        o = res[r["id"]]
        assert o["checksum"] == S.checksum(spin, seeds)
        assert o["naccept"] == int(nacc[-1]) and o["energy"] == eng[-1]
