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
