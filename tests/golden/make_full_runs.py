"""Writes full_runs.json: the oracle's trajectories of the two full-size
single-lattice workloads, whole length, for the GPU tests to compare
against without re-running the oracle there.

* ising sq/4n 4096x4096, kT 2.269, seed 12345: 1000 sweeps (run 500.0 at
  stats 50.0 in script time);
* potts Q=10000 sc/26n 256x256x256, kT 0.5, seed 24680: 200 sweeps.

With --max it writes max_size_runs.json instead: two sweeps of lattices of
2^30 sites, past the 2^32-byte arrays where the GPU kernels switch to
64-bit offsets (ising sq/4n 32768x32768; potts Q=10000 sc/26n 1024^3).
They need about 10 GB of host memory.

Each entry holds the cumulative Naccept and the energy after every sweep
(src/app_lattice.cpp:917-934, src/diag_energy.cpp:53-66) and the state
checksum (spparks_amd.hip.checksum: spins and strict seeds in site-ID order)
after every `every` sweeps.  The oracle is the plain-C restatement in
oracle/, run with OpenMP inside a colour.

Run: python tests/golden/make_full_runs.py [--max]   (about 5 minutes each on 8 cores)
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from spparks_amd.hip import checksum  # noqa: E402

RUNS = [
    {"name": "ising-sq4n-4096", "app": "ising", "lattice": "sq/4n", "shape": [4096, 4096, 1],
     "nspins": 2, "kT": 2.269, "seed": 12345, "sweeps": 1000, "every": 100},
    {"name": "potts-sc26n-256", "app": "potts", "lattice": "sc/26n", "shape": [256, 256, 256],
     "nspins": 10000, "kT": 0.5, "seed": 24680, "sweeps": 200, "every": 20},
]

MAX_RUNS = [
    {"name": "ising-sq4n-32768", "app": "ising", "lattice": "sq/4n", "shape": [32768, 32768, 1],
     "nspins": 2, "kT": 2.269, "seed": 12345, "sweeps": 2, "every": 1},
    {"name": "potts-sc26n-1024", "app": "potts", "lattice": "sc/26n", "shape": [1024, 1024, 1024],
     "nspins": 10000, "kT": 0.5, "seed": 24680, "sweeps": 2, "every": 1},
]


def run(r):
    d = O.make_desc(r["app"], r["lattice"], *r["shape"], r["nspins"], r["kT"])
    spin, seeds = O.script_init(d, r["seed"], 1, r["nspins"])
    out = dict(r)
    out["checksum"] = [str(checksum(spin, seeds))]
    out["naccept"], out["energy"] = [], []
    base = 0
    for _ in range(r["sweeps"] // r["every"]):
        nacc, eng = O.sweeps(d, spin, seeds, r["every"])
        out["naccept"] += [int(base + x) for x in nacc]
        out["energy"] += [int(e) for e in eng]
        base += int(nacc[-1])
        out["checksum"].append(str(checksum(spin, seeds)))
    return out


def main():
    big = "--max" in sys.argv[1:]
    res = []
    for r in MAX_RUNS if big else RUNS:
        t = time.time()
        res.append(run(r))
        print("%s: %d sweeps in %.1f s" % (r["name"], r["sweeps"], time.time() - t), flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "max_size_runs.json" if big else "full_runs.json")
    with open(path, "w") as f:
        json.dump({"runs": res}, f)
        f.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
