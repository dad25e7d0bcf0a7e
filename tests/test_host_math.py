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


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="no hipcc")
def test_ratio_log_accuracy(tmp_path):
    """pf::ln_ratio_near1, the free-energy logarithm ln|A/B| of the fused
    sweeps (source/phase_body.inc:167-174): within 4 ulp of the result where
    the 2 atanh series is used (|A-B|/(A+B) <= 1/16; the logarithm of a
    rounded quotient would be off by up to an ulp of 1 there, many ulps of a
    result near 0), within 5 ulp of 1 elsewhere, and log|A/B| of the special
    operands (0, +-inf, NaN, A = -B)."""
    exe = tmp_path / "ratio_log_accuracy"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-I", CSRC,
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "ratio_log_accuracy.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe), "1000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    near_ulp, far_ulp1, n = r.stdout.split()[:3]
    assert int(n) >= 1300000
    assert float(near_ulp) <= 4.0, r.stdout
    assert float(far_ulp1) <= 5.0, r.stdout


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="no hipcc")
def test_shell_enumeration(tmp_path):
    """pf::shell_flat, the boundary-shell enumeration of the decoration
    kernels: 32-bit and 64-bit index arithmetic agree, and the cells are
    exactly the shell, once each, on 6^3 ragged grids x 2 widths x 2
    x-face options."""
    exe = tmp_path / "shell_enum"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-I", CSRC,
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "shell_enum.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    grids, cells = (int(v) for v in r.stdout.split()[:2])
    assert grids == 6 ** 3 * 4 and cells > 0
