"""The Fortran driver (bin/pf3d): the reference's deck in, the reference's
output events out.  The GPU tests run it on the decks of the unmodified
principal.f90 fixtures and compare, line by line, the console transcript
and performance_log.txt with what the reference itself printed, and the
Tecplot snapshots (8 significant digits) with its files.

Masked in the transcripts: the wall-clock fields of the timing block
(principal.f90:390-398) and its device lines (which device, how many, how
many OpenMP threads), the device banner (principal.f90:46-50), and the lines
of this build's own options, which start with "pf3d:".  The diagnostics line
(L2, max, min, sum c, free energy; e14.6) is compared to 2e-6 relative: the
reductions sum in another order.  The ~5,500 W0_E0 try lines are compared by
count and on the fixture's sample of them."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import gpu_available
from helpers import deck_from_json, load_golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "phase-field-project-3d_amd", "bin", "pf3d")

TRY = "E0-EPSLON_Try==>"
MASKED_KEYS = {"Execution Device", "GPU Device ID", "Available Devices", "OpenMP Threads",
               "Elapsed Time", "Time Since Last Output", "Average Time/Step",
               "Average Time/Output", "Estimated Remaining", "Execution"}
BANNERS = ("### GPU offload ENABLED", "### No OpenMP target device", "### gfx950 HIP device")


def _run(tmp_path, deck, nml=""):
    deck.write(str(tmp_path / "entrada.dat"))
    (tmp_path / "output").mkdir(exist_ok=True)
    if nml:
        (tmp_path / "pf3d.nml").write_text(nml)
    return subprocess.run([EXE], cwd=str(tmp_path), stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=300)


@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
def test_driver_reports_missing_device(tmp_path):
    if gpu_available():
        pytest.skip("a GPU is present")
    z = load_golden("principal_planar.npz")
    r = _run(tmp_path, deck_from_json(z["deck"]))
    assert r.returncode != 0
    assert "pf_init failed" in r.stdout and "no HIP device" in r.stdout


def _parse_fortran_e(t):
    t = t.strip()
    if "E" not in t.upper() and "NAN" not in t.upper():
        for q in range(len(t) - 1, 0, -1):
            if t[q] in "+-":
                t = t[:q] + "E" + t[q:]
                break
    return float(t)


def _masked(lines):
    """The transcript lines that are compared (see the module docstring)."""
    out = []
    for ln in lines:
        if ln.lstrip().startswith(TRY) or ln.startswith("pf3d:") or ln.startswith(BANNERS):
            continue
        if ":" in ln and ln.split(":")[0].strip() in MASKED_KEYS:
            continue
        out.append(ln)
    return out


def _same_diag(a, b):
    ta, tb = a.replace("==>", " ").split(), b.replace("==>", " ").split()
    if ta[0::2] != tb[0::2]:
        return False
    for x, y in zip(ta[1::2], tb[1::2]):
        u, v = _parse_fortran_e(x), _parse_fortran_e(y)
        if not (math.isnan(u) and math.isnan(v)) and not abs(u - v) <= 2e-6 * abs(v):
            return False
    return True


def assert_transcript(got_lines, ref_lines, what):
    got, ref = _masked(got_lines), _masked([str(x) for x in ref_lines])
    for q, (a, b) in enumerate(zip(got, ref)):
        if b.startswith("L2(phi)==>"):
            assert _same_diag(a, b), f"{what} line {q}:\n got {a!r}\n ref {b!r}"
        else:
            assert a == b, f"{what} line {q}:\n got {a!r}\n ref {b!r}"
    assert len(got) == len(ref), (what, len(got), len(ref), got[len(ref):], ref[len(got):])


def assert_try_lines(stdout_lines, z):
    tries = [ln for ln in stdout_lines if ln.lstrip().startswith(TRY)]
    assert len(tries) == int(z["w0e0_tries"])
    for q, ln in zip(z["w0e0_sample_index"], z["w0e0_sample_lines"]):
        assert tries[int(q)] == str(ln)


def assert_snapshots(tmp_path, z):
    files = sorted(os.listdir(tmp_path / "output"))
    assert [f.strip() for f in files] == [str(n).strip() for n in z["filenames"]]
    for q, fn in enumerate(files):
        lines = (tmp_path / "output" / fn).read_text().splitlines()
        assert [ln.strip() for ln in lines[:4]] == [str(h).strip() for h in z[f"e{q}_header"]]
        data = np.array([[float(x) for x in ln.split()] for ln in lines[4:] if ln.strip()])
        ref = z[f"e{q}_xyz_phi_tp_c"]
        assert data.shape == ref.shape
        np.testing.assert_allclose(data, ref, rtol=2e-7, atol=1e-300)


def check_against_reference(tmp_path, r, z):
    assert r.returncode == 0, r.stdout[-3000:]
    assert_snapshots(tmp_path, z)
    out = r.stdout.splitlines()
    assert_try_lines(out, z)
    assert_transcript(out, z["stdout_lines"], "console")
    log = (tmp_path / "performance_log.txt").read_text().splitlines()
    assert_transcript(log, z["perf_log_lines"], "performance_log.txt")


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a gfx950 GPU")
@pytest.mark.parametrize("fused", [".false.", ".true."])
@pytest.mark.parametrize("name", ["planar", "front_j", "stop"])
def test_driver_matches_reference_output(tmp_path, fused, name):
    z = load_golden(f"principal_{name}.npz")
    deck = deck_from_json(z["deck"])
    print(f"[driver] {name} fused={fused}", flush=True)
    r = _run(tmp_path, deck, f"&pf3d\n opt_fused={fused}\n/\n")
    check_against_reference(tmp_path, r, z)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a gfx950 GPU")
def test_driver_restart_matches_reference(tmp_path):
    """tipo_inicial = 4 (entrada.f90:272-321): continue from restart.dat, the
    reference's own output event, with the bottom 2 layers dropped (moving
    frame); the reference's output events after the restart are the target."""
    z = load_golden("principal_restart.npz")
    deck = deck_from_json(z["deck"])
    print("[driver] restart", flush=True)
    (tmp_path / "restart.dat").write_text(str(z["restart_dat"]))
    r = _run(tmp_path, deck)
    assert "deslocamento = 2 camadas" in r.stdout
    check_against_reference(tmp_path, r, z)

