"""The drop-in boundary end to end on the GPU: unchanged SPPARKS input scripts
through the spk_hip program and through libspparks_hip.so (the reference's
library API), checked against the reference run's known answers and the
oracle."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "spparks_amd", "lib", "spk_hip")
SCRIPTS = os.path.join(ROOT, "tests", "scripts")
GOLD = os.path.join(ROOT, "tests", "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_exe(tmp_path, script, *args):
    say("spk_hip", script, args)
    r = subprocess.run([EXE, "-in", script, "-log", "none", *args], capture_output=True, text=True,
                       cwd=tmp_path, timeout=110)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


def stats_rows(out):
    """(time, naccept, nreject, nsweeps, energy) rows of the stats output."""
    rows, on = [], False
    for line in out.splitlines():
        w = line.split()
        if w[:4] == ["Time", "Naccept", "Nreject", "Nsweeps"]:
            on = True
            continue
        if on and len(w) >= 6:
            try:
                rows.append((float(w[0]), int(w[1]), int(w[2]), int(w[3]), float(w[5])))
            except ValueError:
                on = False
        elif on:
            on = False
    return rows


def test_potts_sq8n_128_script_known_answers(tmp_path):
    """in.potts_sq8n_128 unchanged: the stats lines carry the reference run's numbers."""
    g = json.load(open(os.path.join(GOLD, "potts_sq8n_128_known_answers.json")))
    out = run_exe(tmp_path, os.path.join(SCRIPTS, "in.potts_sq8n_128"), "-sf", "hip")
    rows = stats_rows(out)
    assert rows[0] == (0.0, 0, 0, 0, float(g["initial_energy"]))
    assert rows[-1][3] == 100
    by_sweep = {r[3]: r for r in rows}
    for sweep, na, e in g["rows"]:
        t, nacc, nrej, nsw, eng = by_sweep[sweep]
        assert nacc == na and eng == e
        assert nacc + nrej == sweep * 128 * 128
    assert "Loop time of" in out


def test_library_api_matches_oracle(orc, tmp_path, monkeypatch):
    from spparks_amd.spparks import spparks
    monkeypatch.chdir(tmp_path)
    spk = spparks(cmdargs=["-screen", "none", "-log", "none"])
    lines = open(os.path.join(SCRIPTS, "in.potts_sq8n_128")).read().replace("run             1.0", "run 0.1")
    for line in lines.splitlines():
        spk.command(line)
    # time accumulates 0.01 per sweep in floating point: `run 0.1` is 11 sweeps
    n1, t1 = orc.run_nsweeps(0.0, 0.1, 0.01)
    n2, t2 = orc.run_nsweeps(t1, 0.05, 0.01)
    d = orc.make_desc("potts", "sq/8n", 128, 128, 1, 100, 0.5)
    spin, seeds = orc.script_init(d, 56789, 1, 100)
    orc.sweeps(d, spin, seeds, n1)
    n = spk.extract("nglobal", 0)
    assert n == 128 * 128
    site = np.ctypeslib.as_array(spk.extract("site", 1), shape=(n,)).copy()
    seed = np.ctypeslib.as_array(spk.extract("seeds", 1), shape=(n,)).copy()
    assert np.array_equal(site, spin) and np.array_equal(seed, seeds)
    assert spk.energy() == orc.energy(d, spin)
    # a second run at another temperature continues exactly
    spk.command("temperature 0.3")
    spk.command("run 0.05")
    d2 = orc.make_desc("potts", "sq/8n", 128, 128, 1, 100, 0.3)
    orc.sweeps(d2, spin, seeds, n2)
    site = np.ctypeslib.as_array(spk.extract("site", 1), shape=(n,)).copy()
    assert np.array_equal(site, spin)
    assert spk.extract("time", 3) == t2
    assert spk.extract("dimension", 0) == 2
    ids = np.ctypeslib.as_array(spk.extract("id", 1), shape=(n,))
    assert ids[0] == 1 and ids[-1] == n
    spk.close()


def test_sweep_count_from_time_accumulation(tmp_path):
    """run 0.02 with Q=1e4 makes 201 sweeps, run 0.01995 makes 200."""
    base = ("seed 7\napp_style potts 10000\ndimension 3\nlattice sc/26n 1.0\n"
            "region box block 0 4 0 4 0 4\ncreate_box box\ncreate_sites box\n"
            "set site range 1 10000\nsweep color/strict\ntemperature 0.5\ndiag_style energy\n")
    for runtime, n in (("0.02", 201), ("0.01995", 200)):
        p = tmp_path / ("in.t" + runtime)
        p.write_text(base + "run %s\n" % runtime)
        rows = stats_rows(run_exe(tmp_path, str(p)))
        assert rows[-1][3] == n
        assert rows[-1][1] + rows[-1][2] == n * 64


def test_full_size_scripts_first_sweeps(orc, tmp_path):
    """in.ising_sq4n_4096 / in.potts_sc26n_256 unchanged except the run length (-var T): the first
    two sweeps of the 4096^2 Ising and 256^3 Potts runs against the oracle."""
    cases = [("in.ising_sq4n_4096", "ising", "sq/4n", (4096, 4096, 1), 2, 2.269, 12345, "1.0", "0.5"),
             ("in.potts_sc26n_256", "potts", "sc/26n", (256, 256, 256), 10000, 0.5, 24680, "0.00015", "1.0e-4")]
    for script, app, lat, shape, q, kt, seed, T, DT in cases:
        out = run_exe(tmp_path, os.path.join(SCRIPTS, script), "-var", "T", T, "-var", "DT", DT)
        rows = stats_rows(out)
        d = orc.make_desc(app, lat, *shape, q, kt)
        spin, seeds = orc.script_init(d, seed, 1, q)
        e0 = orc.energy(d, spin)
        nacc, eng = orc.sweeps(d, spin, seeds, 2)
        assert rows[0][4] == float("%10g" % e0)
        assert [r[1] for r in rows[1:]] == list(nacc)
        assert [r[4] for r in rows[1:]] == [float("%10g" % e) for e in eng]


def test_run_options_log_undump_and_stats_delay(orc, tmp_path):
    """sweep color/strict/hip, stats with a delay, undump, log, print,
    run ... upto and run ... pre no post no, and a zero-length run that only
    writes a dump: the stats schedule, the files each dump and log get, and
    the final spins against the oracle after the sweeps the stats report."""
    p = tmp_path / "in.opts"
    p.write_text("seed 777\napp_style potts 50\ndimension 2\nlattice sq/8n 1.0\n"
                 "region box block 0 32 0 24 -0.5 0.5\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 50\nsweep color/strict/hip mask no\ntemperature 0.6\n"
                 "diag_style energy\nstats 0.02 delay 0.06\ndump 1 text 0.1 d1.txt id site\nrun 0.2\n"
                 "undump 1\nlog log.second\nprint \"second part\"\nrun 0.4 upto\n"
                 "run 0.1 pre no post no\ndump 3 text 1.0 end.txt id site\nrun 0.0\n")
    out = run_exe(tmp_path, str(p), "-echo", "none")
    rows = stats_rows(out)
    # dt per sweep is 1/Q = 0.02: the first run prints at setup, then from
    # t = 0.06 (the delay) after every sweep
    first = [r[3] for r in rows if r[3] <= 10]
    assert first[:9] == [0, 3, 4, 5, 6, 7, 8, 9, 10]
    snaps = open(tmp_path / "d1.txt").read().split("ITEM: TIMESTEP\n")[1:]
    assert len(snaps) == 3   # t = 0, 0.1, 0.2; undumped before the later runs
    second = open(tmp_path / "log.second").read()
    assert "second part" in second and "Loop time" in second
    nsw = rows[-1][3]
    assert 19 <= nsw <= 21 + 6   # 0.2 + up to 0.4 + 0.1 in steps of 0.02
    d = orc.make_desc("potts", "sq/8n", 32, 24, 1, 50, 0.6)
    spin, seeds = orc.script_init(d, 777, 1, 50)
    nacc, eng = orc.sweeps(d, spin, seeds, nsw)
    assert rows[-1][1] == int(nacc[-1]) and rows[-1][4] == float("%10g" % eng[-1])
    end = open(tmp_path / "end.txt").read().split("ITEM: TIMESTEP\n")[1:]
    assert len(end) == 1
    vals = np.array([[int(x) for x in ln.split()] for ln in end[0].splitlines()[8:]])
    assert np.array_equal(vals[:, 1], spin)


def test_clear_between_two_simulations(orc, tmp_path):
    """clear (src/input.cpp:524-529) destroys the first simulation, lattice
    and dumps included, and keeps the variables: the second simulation in the
    same script runs as it would alone, master seed and all."""
    p = tmp_path / "in.two"
    p.write_text("variable Q index 32\n"
                 "seed 5\napp_style ising\ndimension 2\nlattice sq/4n 1.0\n"
                 "region box block 0 16 0 12 -0.5 0.5\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 2\nsweep color/strict\ntemperature 1.5\ndump 1 text 1.0 first.txt id site\n"
                 "run 2.0\nclear\n"
                 "seed 9\napp_style potts ${Q}\ndimension 3\nlattice sc/26n 1.0\n"
                 "region box block 0 8 0 6 0 4\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 ${Q}\nsweep color/strict\ntemperature 0.7\ndiag_style energy\n"
                 "stats 0.125\nrun 0.375\n")
    out = run_exe(tmp_path, str(p))
    rows = stats_rows(out)
    # Q = 32: a sweep is 1/32 of a time unit, exact in binary
    d = orc.make_desc("potts", "sc/26n", 8, 6, 4, 32, 0.7)
    spin, seeds = orc.script_init(d, 9, 1, 32)
    nacc, eng = orc.sweeps(d, spin, seeds, 12)
    assert [r[3] for r in rows] == [0, 4, 8, 12]
    assert [r[1] for r in rows[1:]] == [int(nacc[k - 1]) for k in (4, 8, 12)]
    assert [r[4] for r in rows[1:]] == [float("%10g" % eng[k - 1]) for k in (4, 8, 12)]
    assert len(open(tmp_path / "first.txt").read().split("ITEM: TIMESTEP\n")) == 4   # t = 0, 1, 2


def test_dump_modify_schedules(orc, tmp_path):
    """dump_modify delay / first / pad and logfreq (src/dump.cpp:644-790)
    with the reference's schedule (Output::setup / compute, next_time,
    src/output.cpp:82-202, :414-437) on an Ising lattice, where a sweep is
    0.5 time units and every time is exact in binary:
      delta 1.5, delay 4: no snapshot before t = 4, then t = 4, 4.5, 6, 7.5
        (next_time rounds up to the delta grid); first 5, pad 3 name them
        d.005 .. d.008;
      delta 0.5, logfreq 2 4.0: t = 0, then 0.5, 1, 2, 4, 8 (two per
        factor of 4)."""
    p = tmp_path / "in.dm"
    p.write_text("seed 41\napp_style ising\ndimension 2\nlattice sq/4n 1.0\n"
                 "region box block 0 20 0 12 -0.5 0.5\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 2\nsweep color/strict\ntemperature 2.1\n"
                 "dump 1 text 1.5 d.* id site\ndump_modify 1 delay 4.0 first 5 pad 3 flush yes sort id\n"
                 "dump 2 text 0.5 log.txt id site\ndump_modify 2 logfreq 2 4.0\nrun 8.0\n")
    run_exe(tmp_path, str(p))
    names = sorted(f.name for f in tmp_path.iterdir() if f.name.startswith("d."))
    assert names == ["d.005", "d.006", "d.007", "d.008"]
    times = [float(open(tmp_path / n).read().split("\n")[1].split()[1]) for n in names]
    assert times == [4.0, 4.5, 6.0, 7.5]
    logsnaps = open(tmp_path / "log.txt").read().split("ITEM: TIMESTEP\n")[1:]
    assert [float(s.split("\n")[0].split()[1]) for s in logsnaps] == [0.0, 0.5, 1.0, 2.0, 4.0, 8.0]
    # the t = 4 snapshot is the state after 8 sweeps
    d = orc.make_desc("ising", "sq/4n", 20, 12, 1, 2, 2.1)
    spin, seeds = orc.script_init(d, 41, 1, 2)
    orc.sweeps(d, spin, seeds, 8)
    vals = np.array([[int(x) for x in ln.split()] for ln in open(tmp_path / "d.005").read().splitlines()[9:]])
    assert np.array_equal(vals[:, 1], spin)


def test_include_dump_one_and_reset_time(orc, tmp_path):
    """include runs another file's commands in place; dump_one writes one
    snapshot of a dump at the current time; reset_time moves the clock that
    stats print without touching the state or the sweep count."""
    (tmp_path / "in.setup").write_text("app_style ising\ndimension 2\nlattice sq/4n 1.0\n"
                                       "region box block 0 24 0 16 -0.5 0.5\ncreate_box box\ncreate_sites box\n")
    p = tmp_path / "in.main"
    p.write_text("seed 31\ninclude in.setup\nset site range 1 2\nsweep color/strict\ntemperature 2.0\n"
                 "diag_style energy\nstats 1.0\ndump 7 text 100.0 once.txt id site\nrun 3.0\n"
                 "dump_one 7\nreset_time 10.0\nrun 2.0\n")
    out = run_exe(tmp_path, str(p))
    rows = stats_rows(out)
    # ising: dt per sweep is 1/2, so run 3.0 is 6 sweeps and run 2.0 four more
    assert [r[3] for r in rows] == [0, 2, 4, 6, 6, 8, 10]
    assert [r[0] for r in rows] == [0.0, 1.0, 2.0, 3.0, 10.0, 11.0, 12.0]
    d = orc.make_desc("ising", "sq/4n", 24, 16, 1, 2, 2.0)
    spin, seeds = orc.script_init(d, 31, 1, 2)
    spin6 = spin.copy()
    sd6 = seeds.copy()
    nacc, eng = orc.sweeps(d, spin6, sd6, 6)
    snaps = open(tmp_path / "once.txt").read().split("ITEM: TIMESTEP\n")[1:]
    assert len(snaps) == 2   # the run's first snapshot at t = 0, then dump_one at t = 3
    assert snaps[1].splitlines()[0].split()[0] == "1"
    vals = np.array([[int(x) for x in ln.split()] for ln in snaps[1].splitlines()[8:]])
    assert np.array_equal(vals[:, 1], spin6)
    nacc10, eng10 = orc.sweeps(d, spin, seeds, 10)
    assert rows[-1][1] == int(nacc10[-1]) and rows[-1][4] == float("%10g" % eng10[-1])


def test_zero_temperature_3d_script_with_default_dump(orc, tmp_path):
    """A 3d grain-growth script at T = 0 with a text dump that names no
    fields (the reference's default: id site x y z, the header calling site
    "type"): every snapshot and every stats line against the oracle, and the
    energy never rises at T = 0."""
    p = tmp_path / "in.gg"
    p.write_text("seed 56789\napp_style potts 100\ndimension 3\nlattice sc/26n 1.0\n"
                 "region box block 0 20 0 20 0 20\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 100\nsweep color/strict\ndiag_style energy\ntemperature 0.0\n"
                 "stats 0.2\ndump 1 text 0.5 dump.gg\nrun 1.0\n")
    out = run_exe(tmp_path, str(p))
    d = orc.make_desc("potts", "sc/26n", 20, 20, 20, 100, 0.0)
    spin, seeds = orc.script_init(d, 56789, 1, 100)
    nacc, eng = orc.sweeps(d, spin.copy(), seeds.copy(), 100)
    rows = stats_rows(out)
    assert [r[3] for r in rows] == [0, 20, 40, 60, 80, 100]
    assert [r[1] for r in rows[1:]] == [int(nacc[k - 1]) for k in (20, 40, 60, 80, 100)]
    assert [r[4] for r in rows[1:]] == [float("%10g" % eng[k - 1]) for k in (20, 40, 60, 80, 100)]
    assert all(b <= a for a, b in zip(eng, eng[1:]))
    snaps = open(tmp_path / "dump.gg").read().split("ITEM: TIMESTEP\n")[1:]
    assert len(snaps) == 3   # t = 0, 0.5, 1.0
    sp, sd = spin.copy(), seeds.copy()
    for snap, nsw in zip(snaps, (0, 50, 50)):
        if nsw:
            orc.sweeps(d, sp, sd, nsw)
        lines = snap.splitlines()
        assert lines[7] == "ITEM: ATOMS id type x y z"
        vals = np.array([[float(x) for x in ln.split()] for ln in lines[8:]])
        assert np.array_equal(vals[:, 0], np.arange(1, 8001))
        assert np.array_equal(vals[:, 1], sp)
        g = np.arange(8000)
        assert np.array_equal(vals[:, 2:], np.stack([g % 20, (g // 20) % 20, g // 400], 1))


def test_dump_text_snapshots(orc, tmp_path):
    p = tmp_path / "in.dump"
    p.write_text("seed 99\napp_style ising\ndimension 2\nlattice sq/4n 1.0\n"
                 "region box block 0 16 0 12 -0.5 0.5\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 2\nsweep color/strict\ntemperature 1.5\n"
                 "dump 1 text 1.0 dump.ising id site x y energy\nrun 2.0\n")
    run_exe(tmp_path, str(p))
    text = open(tmp_path / "dump.ising").read().split("ITEM: TIMESTEP\n")[1:]
    assert len(text) == 3   # t = 0, 1, 2
    d = orc.make_desc("ising", "sq/4n", 16, 12, 1, 2, 1.5)
    spin, seeds = orc.script_init(d, 99, 1, 2)
    nb = orc.neighbors(d)
    for snap, nsw in zip(text, (0, 2, 2)):
        if nsw:
            orc.sweeps(d, spin, seeds, nsw)
        lines = snap.splitlines()
        assert lines[1:3] == ["ITEM: NUMBER OF ATOMS", "192"]
        assert lines[7] == "ITEM: ATOMS id type x y energy"
        vals = np.array([[float(x) for x in ln.split()] for ln in lines[8:]])
        assert np.array_equal(vals[:, 0], np.arange(1, 193))
        assert np.array_equal(vals[:, 1], spin)
        assert np.array_equal(vals[:, 2], np.arange(192) % 16)
        assert np.array_equal(vals[:, 4], (spin[nb] != spin[:, None]).sum(1))


def test_dump_write_error_stops_the_run(tmp_path):
    """Snapshots are written on a background thread; a failed write (a
    full device) still ends the run with an error once they are drained."""
    if not os.path.exists("/dev/full"):
        pytest.skip("no /dev/full")
    p = tmp_path / "in.full"
    p.write_text("seed 7\napp_style ising\ndimension 2\nlattice sq/4n 1.0\n"
                 "region box block 0 64 0 64 -0.5 0.5\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 2\nsweep color/strict\ntemperature 2.0\n"
                 "dump 1 text 1.0 /dev/full id site\nrun 3.0\nprint after\n")
    r = subprocess.run([EXE, "-in", str(p), "-log", "none"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=60)
    assert r.returncode != 0 and "Error writing dump file" in r.stdout + r.stderr
    assert "after" not in r.stdout


def test_invalid_values_error(tmp_path):
    for body, msg in (("", "invalid values"), ("set site value 11\n", "invalid values")):
        p = tmp_path / "in.bad"
        p.write_text("seed 5\napp_style potts 10\ndimension 2\nlattice sq/8n 1.0\n"
                     "region box block 0 8 0 8 -0.5 0.5\ncreate_box box\ncreate_sites box\n" + body +
                     "sweep color/strict\ntemperature 1.0\nrun 1.0\n")
        r = subprocess.run([EXE, "-in", str(p), "-log", "none"], capture_output=True, text=True, cwd=tmp_path,
                           timeout=60)
        assert r.returncode != 0 and msg in r.stdout + r.stderr


@pytest.mark.parametrize("app,q,lat,dims,run,nsw", [
    # potts/neigh[only]: dt_sweep = 1/maxneigh, so run 1.0 on sq/8n is 8 sweeps
    ("potts/neigh", 50, "sq/8n", (16, 12, 1), "1.0", 8),
    ("potts/neighonly", 500, "sc/26n", (8, 6, 4), "0.2", 6),
    # ising/single: dt_sweep = 1
    ("ising/single", 0, "sq/4n", (20, 14, 1), "7.0", 7),
])
def test_variant_scripts_match_oracle(orc, tmp_path, app, q, lat, dims, run, nsw):
    """Unchanged scripts of the proposal variants: the sweep count follows the
    variant's dt_sweep, and every stats row (one per sweep) matches the
    oracle's Naccept and energy."""
    nx, ny, nz = dims
    three = lat.startswith("sc")
    style = "%s %d" % (app, q) if app.startswith("potts") else app
    qq = q if app.startswith("potts") else 2
    region = "region box block 0 %d 0 %d %s\n" % (nx, ny, ("0 %d" % nz) if three else "-0.5 0.5")
    dt = {"potts/neigh": 1.0 / (26 if lat == "sc/26n" else 8), "potts/neighonly": 1.0 / (26 if lat == "sc/26n" else 8),
          "ising/single": 1.0}[app]
    text = ("seed 4321\napp_style %s\ndimension %d\nlattice %s 1.0\n%s"
            "create_box box\ncreate_sites box\nset site range 1 %d\nsweep color/strict\n"
            "temperature 0.6\ndiag_style energy\nstats %r\nrun %s\n"
            % (style, 3 if three else 2, lat, region, qq, dt, run))
    p = tmp_path / "in.variant"
    p.write_text(text)
    rows = stats_rows(run_exe(tmp_path, str(p)))
    assert rows[-1][3] == nsw
    d = orc.make_desc(app, lat, nx, ny, nz, qq, 0.6)
    spin, seeds = orc.script_init(d, 4321, 1, qq)
    e0 = orc.energy(d, spin)
    nacc, eng = orc.sweeps(d, spin, seeds, nsw)
    n = nx * ny * nz
    assert rows[0][1] == 0 and rows[0][4] == e0
    assert len(rows) >= 3
    for r in rows[1:]:
        k = r[3]   # the row's sweep count
        assert r[1] == nacc[k - 1] and r[1] + r[2] == k * n
        assert r[4] == eng[k - 1]


CK_BASE = ("app_style potts 50\ndimension 2\nlattice sq/8n 1.0\nregion box block 0 16 0 12 -0.5 0.5\n"
           "create_box box\ncreate_sites box\n")


def site_file_values(path):
    """(header lines, values array) of a site file written by dump sites."""
    lines = open(path).read().splitlines()
    k = lines.index("Values")
    vals = np.array([[int(w) for w in l.split()] for l in lines[k + 2:] if l.strip()], np.int64)
    return lines[:k], vals


def test_dump_sites_format_and_checkpoint_resume(orc, tmp_path):
    """dump sites writes one site file per snapshot with the reference's
    header (src/dump_sites.cpp:52-66); with this build's seed column such a
    file is a checkpoint: a fresh process (another master seed, no set
    command) that read_sites it continues the trajectory bit for bit."""
    a = tmp_path / "in.a"
    a.write_text("seed 901\n" + CK_BASE + "set site range 1 50\nsweep color/strict\ntemperature 0.6\n"
                 "diag_style energy\nstats 0.02\ndump ck sites 0.1 ck.* id site seed\nrun 0.2\n")
    rows_a = stats_rows(run_exe(tmp_path, str(a)))
    head, v1 = site_file_values(tmp_path / "ck.1")
    assert head[0] == "Site file written by dump sites ck command at time: 1 0.1"
    assert head[1:] == ["", "2 dimension", "192 sites", "id site seed values", "0 16 xlo xhi", "0 12 ylo yhi",
                        "-0.5 0.5 zlo zhi", ""]
    _, v2 = site_file_values(tmp_path / "ck.2")
    assert np.array_equal(v1[:, 0], np.arange(1, 193))
    # the oracle after 5 sweeps (t = 0.1, dt_sweep = 1/50) and at the end of
    # run 0.2: ten steps of 0.02 sum to just below 0.2, so the reference's
    # loop (dt_rkmc fixed at setup) makes an 11th sweep, ending at t = 0.22
    n_end = rows_a[-1][3]
    assert n_end == orc.run_nsweeps(0.0, 0.2, 1.0 / 50)[0] == 11
    d = orc.make_desc("potts", "sq/8n", 16, 12, 1, 50, 0.6)
    spin, seeds = orc.script_init(d, 901, 1, 50)
    orc.sweeps(d, spin, seeds, 5)
    assert np.array_equal(v1[:, 1], spin) and np.array_equal(v1[:, 2], seeds)
    s10, k10 = spin.copy(), seeds.copy()
    nacc, eng = orc.sweeps(d, s10, k10, 5)
    s_end, k_end = spin.copy(), seeds.copy()
    orc.sweeps(d, s_end, k_end, n_end - 5)
    assert np.array_equal(v2[:, 1], s_end) and np.array_equal(v2[:, 2], k_end)

    b = tmp_path / "in.b"
    b.write_text("seed 7777\n" + CK_BASE + "read_sites ck.1\nsweep color/strict\ntemperature 0.6\n"
                 "diag_style energy\nstats 0.02\ndump fin sites 0.1 fin.* id site seed\nrun 0.1\n")
    rows = stats_rows(run_exe(tmp_path, str(b)))
    _, vb = site_file_values(tmp_path / "fin.1")
    # run 0.1 from t = 0 is 5 sweeps: the continuation of ck.1 to sweep 10
    assert rows[-1][3] == 5 and rows[-1][4] == eng[-1]
    assert np.array_equal(vb[:, 1], s10) and np.array_equal(vb[:, 2], k10)


def test_read_sites_values_only_keeps_tagreset_seeds(orc, tmp_path):
    """A site file without the seed column sets spins only; the first run
    then seeds by tagreset from this script's own master draws (no set
    command: the seed draw is the 4th)."""
    rng = np.random.default_rng(3)
    spins = rng.integers(1, 51, 192).astype(np.int32)
    f = tmp_path / "v.sites"
    f.write_text("title\n\n2 dimension\n192 sites\nid site values\n0 16 xlo xhi\n0 12 ylo yhi\n"
                 "-0.5 0.5 zlo zhi\n\nValues\n\n" + "".join("%d %d\n" % (i + 1, s) for i, s in enumerate(spins)))
    b = tmp_path / "in.v"
    b.write_text("seed 4242\n" + CK_BASE + "read_sites v.sites\nsweep color/strict\ntemperature 0.6\n"
                 "diag_style energy\nstats 0.02\ndump fin sites 0.04 out.* id site seed\nrun 0.04\n")
    rows = stats_rows(run_exe(tmp_path, str(b)))
    _, v0 = site_file_values(tmp_path / "out.0")
    assert np.array_equal(v0[:, 1], spins)
    seeds = orc.site_seeds(orc.master_draws(4242, 4)[3], 192)
    assert np.array_equal(v0[:, 2], seeds)
    d = orc.make_desc("potts", "sq/8n", 16, 12, 1, 50, 0.6)
    sp = spins.copy()
    nacc, eng = orc.sweeps(d, sp, seeds, 2)
    _, v1 = site_file_values(tmp_path / "out.1")
    assert np.array_equal(v1[:, 1], sp) and np.array_equal(v1[:, 2], seeds)
    assert rows[-1][1] == nacc[-1] and rows[-1][4] == eng[-1]


@pytest.mark.parametrize("values,msg", [
    ("1 3 4\n", "Incorrect value format"),
    ("1 3\n2 3\n", "Unexpected end of data file"),
    ("".join("%d 51\n" % (i + 1) for i in range(192)), "One or more sites have invalid values"),
])
def test_read_sites_value_errors(tmp_path, values, msg):
    f = tmp_path / "e.sites"
    f.write_text("t\n\n2 dimension\n192 sites\n\nValues\n\n" + values)
    p = tmp_path / "in.e"
    p.write_text("seed 1\n" + CK_BASE + "read_sites e.sites\n")
    r = subprocess.run([EXE, "-in", str(p), "-log", "none"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=110)
    assert r.returncode != 0
    assert msg in r.stdout + r.stderr


@pytest.mark.parametrize("stats,want", [
    ("3.0", [0.0, 1.0, 2.5, 4.0, 5.5]),
    ("0.5", [0.0, 0.5, 1.0, 2.0, 2.5, 3.5, 4.0, 5.0, 5.5])])
def test_dump_tolerance_follows_the_output_schedule(orc, tmp_path, stats, want):
    """dump_modify tol 0.6 on a dump every 1.5 time units (a sweep is 0.5):
    Output::compute runs when the time reaches nextoutput, the earliest of
    the stats time and each dump's next_time, which already has the
    tolerance taken off once; a dump then writes when time >= next_time -
    tol (src/output.cpp:145-202, src/app_lattice.cpp:705-708).  With stats
    every 3.0 the snapshots fall at 0, 1, 2.5, 4 and 5.5; with stats every
    sweep compute runs after each sweep and they fall at 0, 0.5, 1, 2, 2.5,
    3.5, 4, 5 and 5.5.  Each snapshot holds the oracle's state at its time."""
    p = tmp_path / "in.tol"
    p.write_text("seed 43\napp_style ising\ndimension 2\nlattice sq/4n 1.0\n"
                 "region box block 0 20 0 12 -0.5 0.5\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 2\nsweep color/strict\ntemperature 2.1\nstats %s\n"
                 "dump 1 text 1.5 snaps.txt id site\ndump_modify 1 tol 0.6\nrun 6.0\n" % stats)
    run_exe(tmp_path, str(p))
    snaps = open(tmp_path / "snaps.txt").read().split("ITEM: TIMESTEP\n")[1:]
    times = [float(s.split("\n")[0].split()[1]) for s in snaps]
    assert times == want
    d = orc.make_desc("ising", "sq/4n", 20, 12, 1, 2, 2.1)
    spin, seeds = orc.script_init(d, 43, 1, 2)
    done = 0
    for t, s in zip(times, snaps):
        n = int(round(t / 0.5))
        if n > done:
            orc.sweeps(d, spin, seeds, n - done)
            done = n
        vals = np.array([[int(x) for x in ln.split()] for ln in s.splitlines()[8:]])
        assert np.array_equal(vals[:, 1], spin), t


def test_run_upto_before_the_current_time_runs_no_sweep(tmp_path):
    """run X upto with X below the current time is legal (only X < 0 is an
    error, src/app.cpp:126-129): it sets up, prints the stats line and runs
    no sweep."""
    p = tmp_path / "in.upto"
    p.write_text("seed 44\napp_style ising\ndimension 2\nlattice sq/4n 1.0\n"
                 "region box block 0 8 0 8 -0.5 0.5\ncreate_box box\ncreate_sites box\n"
                 "set site range 1 2\nsweep color/strict\ntemperature 2.1\ndiag_style energy\nstats 1.0\n"
                 "run 2.0\nrun 1.0 upto\n")
    rows = stats_rows(run_exe(tmp_path, str(p)))
    assert [r[3] for r in rows] == [0, 2, 4, 4]
