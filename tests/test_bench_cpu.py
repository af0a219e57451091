"""bench.py's host logic: the committed traffic measurements it reads, and
its refusal of more than one GPU.  No GPU is touched."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_every_traffic_lookup_reads_or_is_absent():
    """Every name bench.py looks up in profiles/traffic.json returns a byte
    count or None, and a details record or None, whatever keys the record
    holds (a latency record has no byte count)."""
    b = _bench()
    names = []
    for wl in list(b.WORKLOADS) + ["ensemble-512"]:
        names += [wl, wl + "-bytes", wl + "-resident", wl + "-latency", wl + "-percolour", wl + "-valu"]
    for name in names:
        nbytes, detail = b.load_traffic(name)
        assert nbytes is None or (isinstance(nbytes, int) and nbytes > 0), name
        assert detail is None or isinstance(detail, dict), name
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        table = json.load(f)
    for name, rec in table.items():
        if name.endswith("-latency"):
            assert rec["issue_floor_us_per_pass"] > 0, name
        elif name.endswith("-valu"):
            assert rec["issue_floor_us_per_launch"] > 0, name
        else:
            assert rec["bytes_per_launch"] > 0, name


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
