"""Host checks of the FP64 logarithm used by the fused sweeps.

pf::log_tab (csrc/pf_math.h) replaces the library log in the sweeps'
free-energy term (source/phase_body.inc:167-170) and in the solute's
chemical potential (source/conc_body.inc:125-144).  Its table is regenerated
from the committed generator and compared with the committed header, and
the function is compiled for the host (the same expression as on the
device: frexp and fma are exact IEEE operations on both) and checked
against the long-double logarithm over 8M arguments, including the
neighbourhood of 1 and both edges of every table interval.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "phase-field-project-3d_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_log_table_matches_generator(tmp_path):
    out = tmp_path / "pf_log_table.h"
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_log_table.py"), str(out)],
                   check=True, capture_output=True)
    with open(os.path.join(CSRC, "pf_log_table.h")) as f:
        assert out.read_text() == f.read()


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="no hipcc")
def test_log_accuracy(tmp_path):
    exe = tmp_path / "log_accuracy"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-I", CSRC,
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "log_accuracy.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe), "2000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr  # special values follow log()
    max_ulp, mean_ulp, n = r.stdout.split()[:3]
    assert int(n) >= 4000000
    assert float(max_ulp) <= 1.5, r.stdout
    assert float(mean_ulp) <= 0.35, r.stdout
