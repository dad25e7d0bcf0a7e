"""The Fortran driver (bin/pf3d): the reference's deck in, the reference's
output events out.  The GPU test runs it on the deck of the unmodified
principal.f90 fixture and compares the Tecplot snapshots (8 significant
digits) and diagnostics lines with what the reference itself wrote."""
import os
import subprocess

import numpy as np
import pytest

from conftest import gpu_available
from helpers import deck_from_json, load_golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "phase-field-project-3d_amd", "bin", "pf3d")


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


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a gfx950 GPU")
@pytest.mark.parametrize("fused", [".false.", ".true."])
def test_driver_matches_reference_output(tmp_path, fused):
    z = load_golden("principal_planar.npz")
    deck = deck_from_json(z["deck"])
    print(f"[driver] fused={fused}", flush=True)
    r = _run(tmp_path, deck, f"&pf3d\n opt_fused={fused}\n/\n")
    assert r.returncode == 0, r.stdout[-3000:]
    files = sorted(os.listdir(tmp_path / "output"))
    ref_names = [str(n).strip() for n in z["filenames"]]
    assert [f.strip() for f in files] == ref_names
    for q, fn in enumerate(files):
        lines = (tmp_path / "output" / fn).read_text().splitlines()
        assert [ln.strip() for ln in lines[:4]] == [str(h).strip() for h in z[f"e{q}_header"]]
        data = np.array([[float(x) for x in ln.split()] for ln in lines[4:] if ln.strip()])
        ref = z[f"e{q}_xyz_phi_tp_c"]
        assert data.shape == ref.shape
        np.testing.assert_allclose(data, ref, rtol=2e-7, atol=1e-300)
    diag = [ln for ln in r.stdout.splitlines() if ln.startswith("L2(phi)")]
    got = np.array([[_parse_fortran_e(t) for t in ln.replace("==>", " ").split()[1::2]]
                    for ln in diag])
    np.testing.assert_allclose(got, z["diag"], rtol=2e-6)
    probes = [ln for ln in r.stdout.splitlines()
              if len(ln.split()) == 3 and "E" in ln and "==>" not in ln]
    assert probes[:len(z["probe_lines"])] == [str(p) for p in z["probe_lines"]]
    assert os.path.exists(tmp_path / "performance_log.txt")


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
    assert r.returncode == 0, r.stdout[-3000:]
    assert "shift = 2 layers" in r.stdout
    files = sorted(os.listdir(tmp_path / "output"))
    assert [f.strip() for f in files] == [str(n).strip() for n in z["filenames"]]
    for q, fn in enumerate(files):
        lines = (tmp_path / "output" / fn).read_text().splitlines()
        data = np.array([[float(x) for x in ln.split()] for ln in lines[4:] if ln.strip()])
        np.testing.assert_allclose(data, z[f"e{q}_xyz_phi_tp_c"], rtol=2e-7, atol=1e-300)
    diag = [ln for ln in r.stdout.splitlines() if ln.startswith("L2(phi)")]
    got = np.array([[_parse_fortran_e(t) for t in ln.replace("==>", " ").split()[1::2]]
                    for ln in diag])
    np.testing.assert_allclose(got, z["diag"], rtol=2e-6)
