"""A plain C99 program (tests/abi/consumer.c, built by build()) drives both
libraries through include/spk_hip.h and include/spparks_hip.h alone."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "abi", "consumer")


def consumer(mode, timeout):
    if not os.path.exists(EXE):   # built by build(); a plain gcc line when it is missing
        from spparks_amd import build
        build.build_abi_consumer()
    print("consumer", mode, flush=True)
    p = subprocess.run([EXE, mode], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stdout + p.stderr
    return [line.split() for line in p.stdout.splitlines()]


def test_host_entry_points_from_c():
    import spparks_amd as S
    out = consumer("host", 60)
    assert out[0] == ["ncolors", "8"]
    assert out[1][:2] == ["odd_edge", "-1"] and "incommensurate" in " ".join(out[1])
    with open(os.path.join(ROOT, "tests", "golden", "metropolis_table.json")) as f:
        want = json.load(f)["tables"]["2.269"]
    thr = [l for l in out if l[0] == "thr"]
    assert [int(l[2]) for l in thr] == want["thr"][:5]
    assert [float.fromhex(l[3]) for l in thr[1:]] == [float.fromhex(h) for h in want["exp_bits"][1:5]]
    draws = [float.fromhex(l[2]) for l in out if l[0] == "draw"]
    assert draws == list(S.master_draws(12345, 5))


@pytest.mark.gpu
def test_batch_and_library_from_c():
    """Two replicas in a batch equal the same scripts run alone (checksum,
    Naccept, energy); the reference's library calls run the Potts Q=100
    128x128 script to its recorded energy after 100 sweeps."""
    out = consumer("gpu", 120)
    reps = [l for l in out if l[0] == "replica"]
    assert len(reps) == 2
    for l in reps:
        assert l[3:6] == l[7:10], l
    lib = [l for l in out if l[0] == "library"][0]
    assert lib[1:] == ["nglobal", "16384", "energy", "97818"]


@pytest.mark.gpu
def test_reference_prototypes_and_live_site_array(tmp_path):
    """A C caller that declares the entry points exactly as the reference's
    src/library.h (spparks_open with an MPI_Comm) runs the Potts Q=100
    128x128 script to its recorded energy; "xyz" gives create_sites'
    coordinates; values written through "site" reach the device, and the
    array shows the device's spins after a further run."""
    exe = os.path.join(ROOT, "tests", "abi", "refdecl")
    if not os.path.exists(exe):
        from spparks_amd import build
        build.build_refdecl()
    script = os.path.join(ROOT, "tests", "scripts", "in.potts_sq8n_128")
    print("refdecl", script, flush=True)
    p = subprocess.run([exe, script], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert p.returncode == 0, p.stdout + p.stderr
    out = {l.split()[0]: l.split()[1:] for l in p.stdout.splitlines()}
    assert out["nglobal"] == ["16384", "energy", "97818"]
    assert out["xyz"] == ["16383", "127", "127", "0"]
    assert p.stdout.splitlines()[1].split() == ["xyz", "129", "1", "1", "0"]
    assert out["iarray1_is_site"] == ["1", "iarray2_null", "1", "darray1_null", "1"]
    assert out["striped"] == ["energy", "98304"]
    e, host = out["after_run"][1], out["after_run"][3]
    assert e == host and e != "98304"


def test_dump_writer_formats_like_fprintf(tmp_path):
    """The dump writer's buffered formatting (to_chars for integers, %g for
    coordinates and energies) writes the same bytes as one fprintf per value,
    for every field, 2d and 3d, buffers that fill several times, and two
    snapshots queued through the writer thread into one file."""
    exe = os.path.join(ROOT, "tests", "abi", "dump_format")
    if not os.path.exists(exe):
        from spparks_amd import build
        build.build_dump_format_check()
    p = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and p.stdout.strip() == "ok", p.stdout + p.stderr
