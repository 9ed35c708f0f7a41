"""The MEX binding shown in INTEGRATION.md (the reference-side drop-in a
maintainer would add in place of common/TimeEvolutionMex.C) compiles as C and
as C++ against a stub mex.h, and, run on a MATLAB-like struct against a
stand-in of qd_time_evolution, passes every field through, reads time.steps
as int32 and writes it back in place, and turns a library error into
mexErrMsgTxt -- the contract of TimeEvolutionMex.C:18-78 and
MatlabStructures.C:55."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "cpu_harness", "mex")


def binding_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", text, re.S)
    src = [b for b in blocks if '#include "mex.h"' in b and "mexFunction" in b]
    assert len(src) == 1
    return src[0]


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_mex_binding_compiles_and_round_trips(tmp_path, lang):
    ext = "c" if lang == "c" else "cpp"
    b = tmp_path / ("TimeEvolutionMexAMD." + ext)
    b.write_text(binding_source())
    h = tmp_path / ("harness." + ext)
    h.write_text(open(os.path.join(HARNESS, "mex_harness.c")).read())
    cc = ["gcc", "-std=c11"] if lang == "c" else ["g++", "-std=c++17"]
    exe = str(tmp_path / "mex_test")
    inc = ["-I", HARNESS, "-I", os.path.join(ROOT, "include")]
    r = subprocess.run(cc + ["-Wall", "-Wextra", "-Werror", "-Wno-missing-field-initializers"] + inc
                       + [str(b), str(h), "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "mex binding ok" in r.stdout
