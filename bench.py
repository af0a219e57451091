#!/usr/bin/env python3
"""bench.py -- strict colour sweep throughput on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload ising-sq4n-4096]

A step is one strict sweep (every colour once) over the whole lattice, i.e.
nx*ny*nz site updates.  The headline workload is the 4096x4096 sq/4n Ising
lattice at kT=2.269 (K timed sweeps; its own run is 1000 sweeps).  Inputs are
generated on the device by the `set site range` and strict-seed
initialisation kernels before the timed region, so the timed region contains
only the sweeps.  Prints one JSON line.

The default run also measures every other configuration and nests the
results under "configs": the 128x128 Potts script (one-workgroup kernel,
with a latency figure instead of an HBM fraction), the whole 1000-sweep
4096^2 Ising run, the 256^3 sc/26n Potts run, the 512-replica ensemble
(batched against one at a time, every replica compared), serving streams
at 10, 50 and 200 requests/s, and the three proposal variants.
`--workload` runs one workload alone; `--no-configs` keeps the headline only.
"""
from __future__ import annotations

import argparse
import copy
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# name -> (app, lattice, shape, Q, kT, seed, sweeps, description)
WORKLOADS = {
    "potts-sq8n-128": ("potts", "sq/8n", (128, 128, 1), 100, 0.5, 56789, 100,
             "Potts Q=100 sq/8n 128x128 periodic, kT=0.5, seed 56789"),
    "ising-sq4n-4096": ("ising", "sq/4n", (4096, 4096, 1), 2, 2.269, 12345, 1000,
             "Ising sq/4n 4096x4096 periodic, kT=2.269, seed 12345"),
    "ising-sq4n-8192": ("ising", "sq/4n", (8192, 8192, 1), 2, 2.269, 12345, 250,
                "Ising sq/4n 8192x8192 periodic, kT=2.269, seed 12345 (four times the 4096x4096 lattice)"),
    "potts-sc26n-256": ("potts", "sc/26n", (256, 256, 256), 10000, 0.5, 24680, 200,
             "Potts Q=10000 sc/26n 256^3 periodic, kT=0.5, mask yes, seed 24680"),
    # the strict-capable proposal variants on the same lattices
    "potts-neigh-sc26n-256": ("potts/neigh", "sc/26n", (256, 256, 256), 10000, 0.5, 24680, 50,
             "potts/neigh Q=10000 sc/26n 256^3 periodic, kT=0.5, seed 24680"),
    "potts-neighonly-sc26n-256": ("potts/neighonly", "sc/26n", (256, 256, 256), 10000, 0.5, 24680, 50,
             "potts/neighonly Q=10000 sc/26n 256^3 periodic, kT=0.5, seed 24680"),
    "ising-single-sq4n-4096": ("ising/single", "sq/4n", (4096, 4096, 1), 2, 2.269, 12345, 200,
             "ising/single sq/4n 4096x4096 periodic, kT=2.269, seed 12345"),
}
VARIANTS = ("potts/neigh", "potts/neighonly", "ising/single")

NCOLORS = {"sq/4n": 2, "sq/8n": 4, "sc/6n": 2, "sc/26n": 8}


def spin_bytes(app, q):
    """Device storage of one spin: the narrowest of uint8/uint16/int32
    holding 1..Q."""
    qmax = q if app.startswith("potts") else 2
    return 1 if qmax <= 255 else 2 if qmax <= 65535 else 4


def bytes_per_update(lattice, sbytes):
    """Algorithmic HBM bytes per site update of the per-colour design: each
    colour pass reads the whole spin array (its sites' neighbourhoods cover
    it), reads and writes the int32 seed and writes the spin of its own N/K
    sites: K*s + 8 + s bytes for s-byte spins (4K + 12 for int32)."""
    return NCOLORS[lattice] * sbytes + 8 + sbytes


def load_traffic(workload):
    """rocprofv3-measured HBM traffic of the workload's dominant kernel:
    (bytes per launch or None, the measurement's details or None)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f).get(workload)
    except (OSError, ValueError):
        t = None
    return (t.get("bytes_per_launch") if t else None), t


def host_cores():
    """CPUs this process may run on: its affinity set, capped by a cgroup v2
    CPU quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, math.ceil(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_baseline(app, lattice, shape, q, kt, seed, budget_s, threads):
    """The oracle (C restatement, OpenMP within a colour) on this host's cores:
    whole-lattice sweeps of the same workload until budget_s has elapsed."""
    from oracle import oracle as O
    d = O.make_desc(app, lattice, *shape, q, kt)
    t0 = time.perf_counter()
    spin, seeds = O.script_init(d, seed, 1, q)
    setup_s = time.perf_counter() - t0
    n = shape[0] * shape[1] * shape[2]
    done, t0 = 0, time.perf_counter()
    while True:
        O.sweeps(d, spin, seeds, 1, nthreads=threads, energies=False)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= 1000:
            break
    return {"value": round(n * done / el / 1e9, 6), "unit": "G site-updates/s",
            "cores": threads, "host_cpus": os.cpu_count(), "kind": "port",
            "sample": "%d whole-lattice strict sweeps of the same workload from the same initial "
                      "state, oracle/spk_oracle.c (OpenMP over the sites of one colour) in %.1f s"
                      % (done, el),
            "setup_s": round(setup_s, 4)}


def ensemble_variant(r):
    """The step-kernel variant that runs a replica: ising, or potts-u8 /
    potts-u16 / potts-i32 by spin storage."""
    if r["app"] == "ising":
        return "ising"
    return "potts-" + {1: "u8", 2: "u16", 4: "i32"}[spin_bytes(r["app"], r["nspins"])]


def _hash(a, b):
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    h.update(a.tobytes())
    h.update(b.tobytes())
    return h.hexdigest()


def run_ensemble(args):
    """Ensemble of 512 independent replicas (256 Ising sq/4n, 256 Potts sc/26n),
    100 strict sweeps each.  Batched: one device pool, one persistent kernel
    per (variant, colour step), 25-sweep blocks replayed as HIP graphs, launch
    shape from the autotuner.  Plain: the same replicas one at a time through
    the single-lattice sweep.  Every replica's final spins and seeds must be
    identical between the two."""
    import spparks_amd as S
    from spparks_amd import batch as bmod
    from spparks_amd.workloads import ensemble_replicas

    reps = ensemble_replicas(sweeps=args.steps or 100)
    if args.ensemble_variant:   # the replicas of one step-kernel variant only
        reps = [r for r in reps if ensemble_variant(r) == args.ensemble_variant]
    nsw = reps[0]["sweeps"]
    block = args.block_sweeps
    descs = [bmod.replica_desc(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"], temperature=r["kT"],
                               seed=r["seed"], nsweeps=nsw) for r in reps]
    pool = sum(bmod.replica_bytes(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"]) for r in reps)
    sites = [r["shape"][0] * r["shape"][1] * r["shape"][2] for r in reps]
    updates = sum(sites) * nsw

    def fill(B):
        return [B.add(d) for d in descs]

    # autotune (or, with --block, warm up) on a throwaway copy of the
    # ensemble: results do not depend on the shape, but its sweeps would
    # count against the timed trajectory, and the first launches of each
    # kernel load its code
    with bmod.Batch(pool, len(reps)) as B:
        fill(B)
        B.set_launch(args.block, args.grid or 2048) if args.block else B.set_launch(0, 0)
        B.run(24, graph=False)   # about 20 colour steps per candidate shape
        t = B.last_timing()
        shape = (t["block"], t["grid"])

    with bmod.Batch(pool, len(reps)) as B:
        ids = fill(B)
        B.set_launch(*shape)
        t0 = time.perf_counter()
        dev_ms = 0.0
        nl = 0
        done = 0
        while done < nsw:
            B.run(block, graph=True)
            tm = B.last_timing()
            dev_ms += tm["ms"]
            nl += tm["launches"]
            done += block
        wall_b = time.perf_counter() - t0
        batched = {}
        for rid, r in zip(ids, reps):
            sp, sd = B.download(rid)
            st = B.status(rid)
            batched[r["id"]] = (_hash(sp, sd), st[2], st[3])

    # the same replicas one at a time through the plain single-lattice sweep
    wall_p = 0.0
    mismatch = 0
    for r in ([] if args.skip_plain else reps):
        with S.Lattice(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"], temperature=r["kT"]) as L:
            s0, rseed = S.strict_init_params(r["seed"])
            L.set_range(s0, 1, r["nspins"])
            L.init_seeds(rseed)
            L.prepare(nsw)   # its graphs captured outside the timed call
            L.sync()
            t0 = time.perf_counter()
            nacc, _, _ = L.sweeps(nsw, per_sweep=False)
            wall_p += time.perf_counter() - t0
            sp, sd = L.download()
            e = L.energy()
        h, na, eb = batched[r["id"]]
        if h != _hash(sp, sd) or na != nacc or eb != e:
            mismatch += 1

    value = updates / wall_b / 1e9
    plain = updates / wall_p / 1e9 if wall_p > 0 else None
    bytes_alg = sum(n * nsw * bytes_per_update(r["lattice"], spin_bytes(r["app"], r["nspins"]))
                    for n, r in zip(sites, reps))
    achieved = bytes_alg / (dev_ms / 1e3) / 1e9
    return {
        "metric": "G site-updates/s & % HBM peak; batched vs plain speedup; p99 solve latency",
        "value": round(value, 4), "unit": "G site-updates/s", "n_gpus": 1, "steps": nsw, "warmup": 0,
        "ms_per_step": round(wall_b * 1e3 / nsw, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic: 512 replicas from "
        "spparks_amd.workloads.ensemble_replicas(), each initialised from its own seed",
        "config": {"workload": "ensemble-512: batched ensemble of 512 replicas (256 Ising sq/4n L 64-1024, 256 Potts "
                               "sc/26n L 32-128, Q 100-1e5), %d sweeps each in %d-sweep graph blocks%s"
                               % (nsw, block, "; only the %d %s replicas" % (len(reps), args.ensemble_variant)
                                  if args.ensemble_variant else ""),
                   "replicas": len(reps), "sites": sum(sites), "site_updates": updates,
                   "launch_shape": {"block": shape[0], "grid": shape[1]}},
        "batched_vs_plain_speedup": round(value / plain, 3) if plain else None,
        "plain_one_at_a_time": ({"value": round(plain, 4), "unit": "G site-updates/s", "seconds": round(wall_p, 4)}
                                if plain else None),
        "parity": {"replicas_compared": 0 if args.skip_plain else len(reps), "mismatches": mismatch,
                   "compared": "final spins+seeds hash, Naccept, energy: batched vs plain"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic("ensemble-512")[0],
                     "kernel": "k_batch_step (all variants of the block)",
                     "device_ms": round(dev_ms, 3), "step_launches": nl},
        "cpu_baseline": None,
    }


def run_serving(args):
    """Serving: Poisson streams of 3d Potts grain-growth solves (sc/26n, edges
    32-192, Q 100-1e6, kT 0.1-1.0, 50-500 sweeps) at offered loads of 10, 50
    and 200 requests/s, continuously batched through spk_hip_serve against a
    200 GB HBM budget.  Reports p50/p99 completion latency and aggregate
    site-updates/s per load, and checks sampled requests against solo runs."""
    import numpy as np

    import spparks_amd as S
    from spparks_amd import batch as bmod
    from spparks_amd.workloads import serving_requests

    loads = [(10.0, args.serve_n[0]), (50.0, args.serve_n[1]), (200.0, args.serve_n[2])]
    levels = []
    sample_bad = 0
    sample_n = 0
    for rate, n in loads:
        reqs = serving_requests(rate, n)
        res, st = bmod.serve(reqs, pool_bytes=int(args.serve_budget_gb * 1e9), max_live=4096,
                             block_sweeps=args.serve_block, block=args.block or 256, grid=args.grid or 1024)
        lat = np.array([o["finish"] - r["arrive"] for r, o in zip(reqs, res) if o["status"] == 0])
        offered = rate * np.mean([np.prod(r["shape"]) * r["sweeps"] for r in reqs]) / 1e9
        levels.append({"offered_req_per_s": rate, "requests": n, "completed": int(len(lat)),
                       "refused": int(sum(o["status"] != 0 for o in res)),
                       "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                       "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                       "offered_G_updates_per_s": round(float(offered), 3),
                       "G_site_updates_per_s": round(st["site_updates"] / st["wall"] / 1e9, 3),
                       "gpu_busy_frac": round(st["busy"] / st["wall"], 3),
                       "wall_s": round(st["wall"], 3), "blocks": st["blocks"], "max_live": st["max_live"],
                       "max_pool_GB": round(st["max_pool_used"] / 1e9, 3)})
        # a sample of requests solved alone must give the same state
        rng = np.random.default_rng(int(rate))
        for i in rng.choice(n, size=min(args.serve_check, n), replace=False):
            r, o = reqs[int(i)], res[int(i)]
            with S.Lattice(r["app"], r["lattice"], *r["shape"], nspins=r["nspins"], temperature=r["kT"]) as L:
                s0, rseed = S.strict_init_params(r["seed"])
                L.set_range(s0, 1, r["nspins"])
                L.init_seeds(rseed)
                nacc, _, _ = L.sweeps(r["sweeps"], per_sweep=False)
                solo = (L.checksum(), nacc, L.energy())
            sample_n += 1
            sample_bad += solo != (o["checksum"], o["naccept"], o["energy"])
        print(json.dumps({"serving_level": levels[-1]}), file=sys.stderr, flush=True)
    top = levels[-1]
    return {
        "metric": "G site-updates/s & % HBM peak; batched vs plain speedup; p99 solve latency",
        "value": top["G_site_updates_per_s"], "unit": "G site-updates/s", "n_gpus": 1,
        "steps": sum(l["requests"] for l in levels), "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: Poisson request streams from spparks_amd.workloads.serving_requests()",
        "config": {"workload": "serving-sc26n: serving 3d Potts sc/26n solves at 10/50/200 req/s, %g GB HBM budget, "
                               "%d-sweep blocks" % (args.serve_budget_gb, args.serve_block)},
        "levels": levels,
        "p99_latency_s_at_200": top["p99_latency_s"],
        "parity": {"sampled_requests": sample_n, "mismatches": int(sample_bad),
                   "compared": "checksum, Naccept, energy: in the mix vs solved alone"},
        "roofline": None, "cpu_baseline": None,
    }


def copy_calibration(mib):
    """A plain elementwise kernel of known traffic (torch.neg on int32: reads
    and writes mib MiB each), for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * 2**20 // 4
    a = torch.arange(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    return {"copy_calibration": {"kernel": "torch.neg int32", "launches": 3,
                                 "bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4}}


def run(args):
    import spparks_amd as S

    app, lattice, shape, q, kt, seed, nsw, desc = WORKLOADS[args.workload]
    n = shape[0] * shape[1] * shape[2]
    steps = args.steps if args.steps is not None else nsw
    K = NCOLORS[lattice]
    with S.Lattice(app, lattice, *shape, nspins=q, temperature=kt) as L:
        s0, rseed = S.strict_init_params(seed)
        L.sync()
        t_setup = time.perf_counter()
        L.set_range(s0, 1, q)
        L.init_seeds(rseed)
        L.sync()
        t_setup = time.perf_counter() - t_setup
        if args.block:
            L.set_launch(args.block)
        if args.no_resident:
            L.set_resident(False)
        if args.no_bitplanes:
            L.set_bitplanes(0)
        elif args.bit_sites:
            L.set_bitplanes(args.bit_sites)
        L.set_graph(args.chunk)
        if args.warmup:
            L.sweeps(args.warmup, per_sweep=False)
        # set-up of the timed call: its graphs captured, planes packed
        L.prepare(steps)
        L.sync()
        t0 = time.perf_counter()
        nacc, _, _ = L.sweeps(steps, per_sweep=False)
        L.sync()
        wall = time.perf_counter() - t0
        dev_ms, nlaunch = L.last_timing()
        energy = L.energy()

    updates = n * steps
    value = updates / wall / 1e9
    per_launch_s = dev_ms / 1e3 / nlaunch
    sb = spin_bytes(app, q)
    B = bytes_per_update(lattice, sb)
    tname = {1: "u8", 2: "u16", 4: "i32"}[sb]
    # the whole state (spins and int32 seeds) in one workgroup's 160 KiB of
    # LDS: one launch per call (the library's rule; proposal variants never)
    resident = (not args.no_resident and app not in VARIANTS and n * (sb + 4) <= 160 * 1024
                and nlaunch < steps)
    planes = (lattice == "sq/4n" and (app in ("ising", "ising/single") or q == 2)
              and not args.no_bitplanes and not resident)
    variant = app in VARIANTS and not planes
    launches = None   # (two-sweep launches, per-colour launches) of a call on paired planes
    if variant:
        # one site per thread, neighbours gathered in the reference's order;
        # the same bytes as the per-colour design
        kernel = "k_sweep_variant<%s, %s, %s>" % (lattice, tname, app)
        bytes_per_launch = B * n // K
        B_kernel = B
    elif resident:
        # all sweeps of the call in one single-workgroup launch: the state
        # crosses HBM once each way; the kernel is bound by the latency of
        # its colour passes in LDS, not by HBM
        kernel = "k_sweep_resident<%s, %s> (one workgroup, state in LDS)" % (lattice, tname)
        bytes_per_launch = 2 * n * (sb + 4)
        B_kernel = bytes_per_launch / (n * steps / nlaunch)
    elif planes and nlaunch < K * steps:
        # two sweeps per launch (k_sweep_bits_pair): each site's int32 seed
        # is read and written once per launch and its plane bit read and
        # written once; the halo rows a band reads again are not counted.
        # A call of n sweeps makes 2*(n//4) such launches and 2*(n%4)
        # per-colour launches (8 + 3/8 bytes per update each)
        pairs = (K * steps - nlaunch) // 3
        colour = nlaunch - pairs
        kernel = "k_sweep_bits_pair (two sweeps per launch, bands of 16 rows)"
        bytes_per_launch = int((8 + 2 / 8) * n)
        B_kernel = bytes_per_launch / (2 * n)
        # a per-colour launch counted as its share of a pair launch's bytes
        colour_share = (8 + 3 / 8) * n / K / bytes_per_launch
        per_launch_s = (dev_ms / 1e3) / (pairs + colour * colour_share)
        launches = {"two_sweep": pairs, "per_colour": colour}
    elif planes:
        # two-state sq/4n on bit planes: a colour pass reads and writes the
        # int32 seed of each of its sites and moves one bit per site of its
        # own colour (read, write) and of the other (read)
        kernel = "k_sweep_bits (two-state sq/4n, 32 sites per word)"
        B = 8 + 3 / 8
        bytes_per_launch = int(B * n // K)
        B_kernel = B
    else:
        kernel = "k_sweep_color<%s, %s, 4>" % (lattice, tname)
        bytes_per_launch = B * n // K
        B_kernel = B
    achieved = bytes_per_launch / per_launch_s / 1e9
    # the measured traffic of the kernel that ran: two sweeps per launch is
    # the default where it fits; the per-colour plane kernel and the byte
    # kernel have entries of their own
    traffic, traffic_detail = load_traffic(args.workload + (
        "-bytes" if lattice == "sq/4n" and not planes and not resident and q <= 2 else
        "-percolour" if planes and launches is None else ""))
    if resident:
        traffic, traffic_detail = load_traffic(args.workload + "-resident")
    out = {
        "metric": "G site-updates/s & % HBM peak; batched vs plain speedup; p99 solve latency",
        "value": round(value, 4),
        "unit": "G site-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: set site range 1 Q from the master seed, strict seeds by tagreset",
        "config": {"workload": "%s: %s; a step is one strict sweep of the whole lattice, %d timed"
                               % (args.workload, desc, steps),
                   "app": app, "lattice": lattice, "shape": list(shape), "nspins": q, "kT": kt,
                   "seed": seed, "sites": n, "colors": K, "sweep": "color/strict",
                   "naccept": int(nacc), "final_energy": energy},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_measurement": traffic_detail,
                     "kernel": kernel,
                     "bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": round(per_launch_s * 1e6, 3),
                     "bytes_per_update": round(B_kernel, 3),
                     "launches": launches,
                     # the int32-spin byte models of the per-colour design, for
                     # comparison only: frac above uses what this kernel moves
                     "int32_model_bytes_per_update": {"per_colour_4K_plus_12": 4 * K + 12,
                                                      "compulsory_spin_and_seed_16": 16}},
        "pct_hbm_peak": round(100 * value * 1e9 * B_kernel / (HBM_PEAK_GBS * 1e9), 2),
        "spin_bytes": spin_bytes(app, q),
    }
    # device-side set-up (`set site range` by jump-ahead, strict seeds by
    # tagreset) against the oracle's sequential restatement of the same
    if resident:
        # one workgroup on one CU: the colour passes of a launch follow each
        # other behind a barrier and the state crosses HBM once per launch,
        # so the time per pass, not an HBM fraction, says how close it is
        # to what bounds it (the issue time of a pass's instructions, from
        # the committed SQ counters, when present)
        passes = steps * K
        per_pass_us = dev_ms * 1e3 / passes
        roof = {"bound": "latency", "kernel": kernel, "per_pass_us": round(per_pass_us, 3),
                "passes_per_launch": passes // nlaunch, "avg_launch_us": round(per_launch_s * 1e6, 3),
                "bytes_per_launch": bytes_per_launch, "traffic": traffic, "traffic_measurement": traffic_detail}
        lat = load_traffic(args.workload + "-latency")[1]
        if lat:
            roof["issue_floor_us_per_pass"] = lat["issue_floor_us_per_pass"]
            roof["frac"] = round(lat["issue_floor_us_per_pass"] / per_pass_us, 4)
            roof["issue_floor_measurement"] = lat
        out["roofline"] = roof
        out["pct_hbm_peak"] = None
    if launches is not None:
        # two sweeps per launch move half the bytes of two per-colour sweeps
        # and are bound by VALU issue instead: the issue time of a launch's
        # instructions (committed SQ counters) against its measured time
        v = load_traffic(args.workload + "-valu")[1]
        if v:
            out["roofline"]["valu_issue"] = {"floor_us_per_launch": v["issue_floor_us_per_launch"],
                                             "frac": round(v["issue_floor_us_per_launch"] / (per_launch_s * 1e6), 4),
                                             "measurement": v}
    out["setup"] = {"device_s": round(t_setup, 6),
                    "what": "set site range 1 Q and tagreset seeds for every site, on the device"}
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cores()
        out["cpu_baseline"] = cpu_baseline(app, lattice, shape, q, kt, seed, args.cpu_seconds, threads)
        out["setup"].update(cpu_port_s=out["cpu_baseline"].pop("setup_s"), cpu_threads=threads)
    else:
        out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed sweeps (default: the workload's count)")
    ap.add_argument("--warmup", type=int, default=64,
                    help="untimed sweeps first (captures the sweep graphs)")
    ap.add_argument("--workload", default="ising-sq4n-4096", choices=sorted(WORKLOADS) + ["ensemble-512", "serving-sc26n"])
    ap.add_argument("--serve-n", type=int, nargs=3, default=[150, 500, 1000], metavar=("N10", "N50", "N200"),
                    help="serving: requests per offered load")
    ap.add_argument("--serve-block", type=int, default=10, help="serving: sweeps per batch block")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-check", type=int, default=8, help="serving: requests per load re-solved alone")
    ap.add_argument("--skip-plain", action="store_true",
                    help="ensemble: time the batched run only (no one-at-a-time comparison)")
    ap.add_argument("--block-sweeps", type=int, default=25, help="ensemble: sweeps per graph block")
    ap.add_argument("--ensemble-variant", default="", choices=["", "ising", "potts-u8", "potts-u16", "potts-i32"],
                    help="ensemble: only the replicas one step-kernel variant runs")
    ap.add_argument("--grid", type=int, default=0, help="ensemble: persistent workgroups (with --block)")
    ap.add_argument("--block", type=int, default=0, help="sweep-kernel workgroup size (0 = default)")
    ap.add_argument("--no-bitplanes", action="store_true",
                    help="two-state sq/4n lattices on the byte-per-site kernel instead of bit planes")
    ap.add_argument("--no-resident", action="store_true",
                    help="small lattices through the per-colour kernels instead of one LDS-resident workgroup")
    ap.add_argument("--bit-sites", type=int, default=0, choices=[0, 8, 16, 32],
                    help="bit planes: sites per thread (0 = default)")
    ap.add_argument("--chunk", type=int, default=32, help="sweeps per replayed graph (0 = eager launches)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="headline workload only (no nested measurement of the other configurations)")
    ap.add_argument("--serve-n-nested", type=int, nargs=3, default=[40, 120, 300], metavar=("N10", "N50", "N200"),
                    help="serving requests per offered load in the nested configs")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run only a plain copy-like kernel of MIB MiB (rocprof counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args.copy_calibration)), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(a single lattice would shard by slabs with a per-colour halo; replicas by "
              "replica).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble-512":
        print(json.dumps(run_ensemble(args)), flush=True)
    elif args.workload == "serving-sc26n":
        print(json.dumps(run_serving(args)), flush=True)
    else:
        out = run(args)
        if args.workload == "ising-sq4n-4096" and not args.no_configs:
            out["configs"] = run_configs(args)
        print(json.dumps(out), flush=True)


# keys every single-workload line repeats; a nested config keeps the rest
_SHARED = ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data")


def _nested(r, seconds):
    out = {k: v for k, v in r.items() if k not in _SHARED}
    out["bench_seconds"] = round(seconds, 2)
    return out


def run_configs(args):
    """Every configuration besides the headline, each at its own workload:
    cfg1 potts-sq8n-128 (100 sweeps), cfg2 ising-sq4n-4096 (1000 sweeps),
    cfg3 potts-sc26n-256 (200 sweeps), cfg4 ensemble-512 (100 sweeps, every
    replica against its one-at-a-time run), cfg5 serving-sc26n (shorter
    streams than --workload serving-sc26n), and the proposal variants.  The
    single lattices get a short CPU sample of the oracle each."""
    cfgs = {}
    label = {"potts-sq8n-128": "cfg1", "ising-sq4n-4096": "cfg2", "potts-sc26n-256": "cfg3"}
    for wl in ("potts-sq8n-128", "ising-sq4n-4096", "potts-sc26n-256", "potts-neigh-sc26n-256",
               "potts-neighonly-sc26n-256", "ising-single-sq4n-4096"):
        a = copy.copy(args)
        # warm-up sweeps also bring the clocks up after the CPU baseline (the
        # one-workgroup kernel of cfg1 runs on one CU and feels them most)
        a.workload, a.steps, a.warmup = wl, None, 64
        a.cpu_seconds = 3.0
        a.no_cpu_baseline = args.no_cpu_baseline or wl not in label
        t0 = time.perf_counter()
        r = run(a)
        cfgs[label.get(wl, wl) + ":" + wl] = _nested(r, time.perf_counter() - t0)
        print(json.dumps({"config_done": wl, "value": r["value"]}), file=sys.stderr, flush=True)
    a = copy.copy(args)
    a.steps, a.block, a.grid = None, 0, 0
    t0 = time.perf_counter()
    cfgs["cfg4:ensemble-512"] = _nested(run_ensemble(a), time.perf_counter() - t0)
    print(json.dumps({"config_done": "ensemble-512"}), file=sys.stderr, flush=True)
    a = copy.copy(args)
    a.serve_n = args.serve_n_nested
    a.serve_check = 4
    t0 = time.perf_counter()
    cfgs["cfg5:serving-sc26n"] = _nested(run_serving(a), time.perf_counter() - t0)
    return cfgs


if __name__ == "__main__":
    main()
