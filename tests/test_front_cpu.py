"""spk_hip front end on the CPU: command parsing and the reference's error
messages for every error that arises before lattice data exist."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "spparks_amd", "lib", "spk_hip")

PRE = """seed 5
app_style potts 10
dimension 2
lattice sq/8n 1.0
region box block 0 8 0 8 -0.5 0.5
create_box box
"""


def run_script(tmp_path, text, *args):
    if not os.path.exists(EXE):
        pytest.skip("spk_hip not built")
    p = tmp_path / "in.test"
    p.write_text(text)
    r = subprocess.run([EXE, "-in", str(p), "-log", "none", *args], capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("text,msg", [
    ("seed 0\n", "Illegal seed command"),
    ("next\n", "Illegal next command"),
    ("next zz\n", "Invalid variable in next command"),
    ("variable a index 1 2\nvariable b loop 2\nnext a b\n", "All variables in next command must be same style"),
    ("variable e equal 1\nnext e\n", "Invalid variable style with next command"),
    ("label\n", "Illegal label command"),
    ("jump\n", "Illegal jump command"),
    ("jump no_such_file.in\n", "Cannot open input script no_such_file.in"),
    ("variable n loop 0\nprint ${n}\n", "Substitution for illegal variable n"),
    ("clear 1\n", "Illegal clear command"),
    ("app_style potts 10\nclear\nrun 1.0\n", "Run command before app_style set"),
    (PRE + "dump 1 text 1.0 f.txt\ndump_modify 2 delay 1\n", "Could not find dump ID in dump_modify command"),
    (PRE + "dump 1 text 1.0 f.txt\ndump_modify 1\n", "Illegal dump_modify command"),
    (PRE + "dump 1 text 1.0 f.txt\ndump_modify 1 delta 0\n", "Illegal dump_modify command"),
    (PRE + "dump 1 text 1.0 f.txt\ndump_modify 1 logfreq 3\n", "Illegal dump_modify command"),
    (PRE + "dump 1 text 1.0 f.txt\ndump_modify 1 bogus 1\n", "Illegal dump_modify command"),
    (PRE + "dump 1 text 1.0 f.txt\ndump_modify 1 sort 2\n", "outside this build"),
    (PRE + "dump 1 text 1.0 f.txt\ndump_modify 1 region box\n", "outside this build"),
    ("include no_such_file.in\n", "Cannot open input script no_such_file.in"),
    ("include\n", "Illegal include command"),
    ("reset_time 1.0\n", "Reset_time command before app_style set"),
    ("app_style potts 10\nreset_time -1.0\n", "Illegal reset_time command"),
    ("app_style potts 10\nreset_time stitch f.db first\n", "outside this build"),
    ("dump_one 1\n", "Dump_one command before app_style set"),
    ("app_style potts 10\ndump_one 1\n", "Could not find dump ID in dump_one command"),
    ("bogus_command 1\n", "Unknown command"),
    ("run 1.0\n", "Run command before app_style set"),
    ("app_style potts 10\ndimension 3\nlattice sq/8n 1.0\n", "Lattice style does not match dimension"),
    ("app_style potts\n", "Illegal app_style command"),
    ("app_style diffusion\n", "Unrecognized app style"),
    ("app_style potts 10\nsweep random\n", "outside this build"),
    ("app_style potts 10\nsweep color/strict mask maybe\n", "Illegal sweep command"),
    ("app_style potts 10\nsector yes\n", "outside this build"),
    ("sweep color/strict\n", "before app_style set"),
    (PRE + "create_sites box\nsweep color/strict\nrun 1.0\n", "ROCm|hip|device"),
    ("app_style ising\ndimension 2\nlattice sq/4n 1.0\nregion box block 0 7 0 8 -0.5 0.5\n"
     "create_box box\ncreate_sites box\nseed 3\nset site range 1 2\n", "incommensurate"),
    ("app_style potts 10\ndimension 2\nlattice sq/8n 1.0\nregion box block 0 8.5 0 8 -0.5 0.5\n"
     "create_box box\ncreate_sites box\n", "not a multiple of lattice spacing"),
    ("variable x index 3\nprint ${y}\n", "illegal variable"),
    ("stats -1\n", "Illegal stats command"),
    # dump sites / read_sites (src/dump_sites.cpp:25-48, src/read_sites.cpp:69-80)
    (PRE + "dump 1 sites 1.0 out.sites id site\n", "Dump sites must write one file per snapshot"),
    (PRE + "dump 1 sites 1.0 out.* site id\n", "Dump sites requires ID as first field"),
    (PRE + "dump 1 sites 1.0 out.* id\n", "Dump sites requires two or more fields"),
    (PRE + "dump 1 text 1.0 out.txt id propensity\n", "Invalid attribute propensity"),
    ("read_sites file.sites\n", "Read_sites command before app_style set"),
    ("app_style potts 10\nread_sites a b\n", "Illegal read_sites command"),
    ("app_style potts 10\nread_sites no_such_file\n", "Cannot open file"),
    ("app_style ising/single 3\n", "Illegal app_style command"),
    ("app_style potts/neigh\n", "Illegal app_style command"),
])
def test_error_messages(tmp_path, text, msg):
    import re
    rc, out = run_script(tmp_path, text)
    assert rc != 0
    assert re.search(msg, out), out


def test_variables_and_print(tmp_path):
    rc, out = run_script(tmp_path, "variable a index 7\nvariable b string hello\nprint \"$a ${b} x\"\n",
                         "-var", "a", "9")
    assert rc == 0, out
    assert "9 hello x" in out


def test_continuation_and_comments(tmp_path):
    rc, out = run_script(tmp_path, "print one & \n two # comment\n# only a comment\n\nprint \"a # b\"\n")
    assert rc == 0, out
    assert "one two" in out and "a # b" in out


def test_only_hip_suffix(tmp_path):
    rc, out = run_script(tmp_path, "print ok\n", "-sf", "opt")
    assert rc != 0 and "hip" in out
    rc, out = run_script(tmp_path, "print ok\n", "-sf", "hip")
    assert rc == 0


def _site_file(tmp_path, body_header, values):
    p = tmp_path / "in.sites"
    p.write_text("title line\n\n" + body_header + values)
    return str(p)


@pytest.mark.parametrize("header,values,msg", [
    ("2 dimension\n64 sites\nid site values\n-0.5 8.5 xlo xhi\n", "\nValues\n\n1 1\n",
     "Data file simulation box different"),
    ("3 dimension\n64 sites\n", "\nValues\n\n", "dimension does not match"),
    ("2 dimension\n64 sites\n", "\nSites\n\n", "Sites section is outside this build"),
    ("2 dimension\n64 sites\n# a comment\n", "\nVelocities\n\n", "Unknown identifier in data file: Velocities"),
    ("2 dimension\n64 sites\n", "\nValues\n\n1 3\n", "Cannot read Values before sites exist"),
    ("2 dimension\n", "", "Site file has no Sites, Neighbors, or Values"),
])
def test_read_sites_header_errors(tmp_path, header, values, msg):
    """read_sites' header checks against the existing box, in the reference's
    order (src/read_sites.cpp:199-314, :122-168); the box exists, the sites
    do not (creating them needs the device)."""
    f = _site_file(tmp_path, header, values)
    rc, out = run_script(tmp_path, PRE + "read_sites %s\n" % f)
    assert rc != 0
    assert msg in out, out


def test_equal_style_variables_substitute_ten_digits(tmp_path):
    """An equal-style variable substitutes "%.10g" of its value
    (src/variable.cpp:349-352), so a temperature written with more digits
    runs at the ten-digit value, as in the reference; index and string
    variables substitute their text."""
    rc, out = run_script(tmp_path, "variable T equal 2.269185314213\nvariable D equal 1.0e-4\n"
                                   "variable N equal 250000\nvariable S index 0.12345678901234\n"
                                   "print ${T} ${D} ${N} ${S}\n")
    assert rc == 0, out
    assert "2.269185314 0.0001 250000 0.12345678901234" in out


def test_loops_with_jump_label_next(tmp_path):
    """An index and a loop variable stepped by next, the script re-read from
    its label by jump until the variable runs out, which skips the final
    jump (src/input.cpp:621-662, src/variable.cpp:242-280); then a jump on
    to another file."""
    (tmp_path / "in.b").write_text("print \"in b\"\n")
    rc, out = run_script(tmp_path, "variable a index 3 5 7\nvariable i loop 2\nlabel top\n"
                                   "print \"a = ${a} i = ${i}\"\nnext a\nnext i\njump in.test top\n"
                                   "print \"after the loop\"\njump in.b\nprint \"never\"\n")
    assert rc == 0, out
    lines = [l for l in out.splitlines() if l.startswith(("a =", "after", "in b", "never"))]
    # i runs out after two passes, which skips the jump; a is still at 7
    assert lines == ["a = 3 i = 1", "a = 5 i = 2", "after the loop", "in b"]


def test_log_command_continues_in_a_new_file(tmp_path):
    rc, out = run_script(tmp_path, "print one\nlog second.log\nprint two\n")
    assert rc == 0, out
    text = (tmp_path / "second.log").read_text()
    assert text.startswith("SPPARKS") and "continued logfile from same input script" in text
    assert "two" in text and "one" not in text
    rc, out = run_script(tmp_path, "log no_such_dir/x.log\n")
    assert rc != 0 and "Cannot open logfile no_such_dir/x.log" in out


def test_clear_inside_a_loop_keeps_the_loop(tmp_path):
    """clear keeps the input state (variables, the file being read, a
    pending label), so a loop around clear runs every pass."""
    rc, out = run_script(tmp_path, "variable i loop 3\nlabel top\nclear\nprint \"pass ${i}\"\n"
                                   "next i\njump in.test top\nprint end\n")
    assert rc == 0, out
    assert [l for l in out.splitlines() if l.startswith(("pass", "end"))] == ["pass 1", "pass 2", "pass 3", "end"]


def test_jump_from_the_library_is_refused():
    """A jump needs an input file to read on from; through spparks_command
    it is an error, and the instance goes on."""
    from spparks_amd.spparks import spparks, SpparksError
    spk = spparks(cmdargs=["-screen", "none", "-log", "none"])
    with pytest.raises(SpparksError, match="Jump command outside an input script"):
        spk.command("jump in.anything")
    spk.command("variable x index 4")
    spk.close()


def test_jump_to_a_missing_label_is_an_error(tmp_path):
    """A jump whose label never comes ends with the reference's error at the
    end of the file (src/input.cpp:155-157), and the instance is not left
    skipping commands; lines passed over while looking for a label are not
    substituted (src/input.cpp:314-317), so an undefined variable there is
    no error."""
    rc, out = run_script(tmp_path, "print start\njump in.test nolabel\nprint never\n")
    assert rc != 0 and "Label wasn't found in input script" in out, out
    assert "never" not in out and out.count("start") >= 1
    rc, out = run_script(tmp_path, "jump in.test skip\nprint ${undefined}\nlabel skip\nprint done\n")
    assert rc == 0, out
    assert "done" in out and "Substitution" not in out
    from spparks_amd.spparks import spparks, SpparksError
    p = tmp_path / "in.nolabel"
    p.write_text("jump %s nolabel\n" % p)
    spk = spparks(cmdargs=["-screen", "none", "-log", "none"])
    with pytest.raises(SpparksError, match="Label wasn't found in input script"):
        spk.file(str(p))
    with pytest.raises(SpparksError, match="Substitution for illegal variable"):
        spk.command("print ${nothing}")   # commands run again: not skipped
    spk.close()
