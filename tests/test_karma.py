"""The quantitative Karma model (karma/karma3d.f90).

CPU: the oracle (oracle/karma_oracle.c) against the reference program
compiled from karma/karma3d.f90 (tests/golden/karma_n16.npz: its Tecplot
output at steps 20 and 40 of a 16^3 run with noise, 8 significant digits),
the mirror ghosts against the literal x, y, z passes, the noise against the
main model's hash, the namelist reader and tip rule, and the library's
exports.  GPU: the HIP step against the oracle, and the program bin/karma3d
against the reference program's console transcript and files (the regular
run and the tip_stop exit)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import oracle as O
from pfamd import capi
from pfamd.karma import EXPORTS, Karma, KarmaDeck, tips

GOLD = os.path.join(ROOT, "tests", "golden", "karma_n16.npz")
EXE = os.path.join(ROOT, "phase-field-project-3d_amd", "bin", "karma3d")


def _nml():
    z = np.load(GOLD)
    return {k: v for k, v in z["nml"]}


def _oracle_fields(n, steps, Fnoise, R0, record=()):
    cl0, om = O.karma_omega()
    phi, U = O.karma_init(n, R0=R0, Omega=om)
    par = O.karma_par(Fnoise=Fnoise)
    out = {}
    for it in range(1, steps + 1):
        O.karma_step(n, par, it, phi, U)
        if it in record:
            out[it] = (O.karma_interior(n, phi).copy(), O.karma_interior(n, U).copy())
    return out, cl0


def test_oracle_matches_reference_program():
    """Printed with ES15.7 (8 significant digits, relative rounding <= 5e-8):
    the oracle agrees to that rounding at both output steps."""
    nml = _nml()
    n = int(nml["N"])
    z = np.load(GOLD)
    out, cl0 = _oracle_fields(n, 40, float(nml["Fnoise"]), float(nml["R0"]), record=(20, 40))
    k = 0.14
    for it in (20, 40):
        p, u = out[it]
        ph = 0.5 * (1.0 + p)
        cc = cl0 * 0.5 * (1.0 + k - (1.0 - k) * p) * (1.0 + (1.0 - k) * u)
        gp, gc = z[f"phase_{it}"], z[f"conc_{it}"]
        assert np.all(np.abs(ph - gp) <= 6e-8 * np.abs(gp) + 1e-300), it
        assert np.all(np.abs(cc - gc) <= 6e-8 * np.abs(gc)), it
        # the run moved: the interface advanced and U changed from -Omega
        assert gp.max() > 0.99 and gp.min() < 1e-8
    p20, p40 = out[20][0], out[40][0]
    assert np.abs(p40 - p20).max() > 1e-3


def test_oracle_mirror_is_the_three_passes():
    n = 5
    rng = np.random.default_rng(3)
    f = rng.standard_normal((n + 2) ** 3)
    g = f.copy().reshape(n + 2, n + 2, n + 2)  # [k, j, i]
    g[:, :, 0] = g[:, :, 1]; g[:, :, n + 1] = g[:, :, n]      # x pass
    g[:, 0, :] = g[:, 1, :]; g[:, n + 1, :] = g[:, n, :]      # y pass
    g[0, :, :] = g[1, :, :]; g[n + 1, :, :] = g[n, :, :]      # z pass
    O.karma_mirror(n, f)
    assert np.array_equal(f.reshape(n + 2, n + 2, n + 2), g)


def test_oracle_noise_is_the_unsalted_hash():
    """ruido = 2 u - 1 with u the main model's pseudo_random at salt 0."""
    lib = O.load()
    for (i, j, k, it) in [(1, 1, 1, 1), (5, 17, 200, 12345), (200, 1, 3, 45000)]:
        u = lib.pfo_pseudo_random(i, j, k, it, 0, 0)
        assert O.karma_noise(i, j, k, it) == 2.0 * u - 1.0


def test_oracle_flat_states_are_fixed_points():
    """Uniform solid (phi = 1) and uniform liquid at U = 0 do not move:
    phi - phi^3 and the fluxes vanish."""
    n = 6
    par = O.karma_par()
    for pv in (1.0, -1.0):
        phi = np.full((n + 2) ** 3, pv)
        U = np.zeros((n + 2) ** 3)
        O.karma_step(n, par, 1, phi, U)
        assert np.all(phi == pv) and np.all(U == 0.0)


def test_library_exports_karma_symbols():
    hdr = open(os.path.join(ROOT, "include", "pf_karma.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(EXPORTS), declared ^ set(EXPORTS)
    lib = capi.load()
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.gpu
@pytest.mark.parametrize("zmarch", [True, False])
@pytest.mark.parametrize("n,steps,Fnoise", [(16, 40, 0.05), (23, 60, 0.0), (40, 30, 0.02),
                                            (37, 20, 0.01)])
def test_karma_gpu_matches_oracle(n, steps, Fnoise, zmarch):
    """Same operation order with FP contraction off on both sides, correctly
    rounded division and square root on the device, the seed set on the host:
    the GPU fields equal the oracle's bit for bit, in both kernel forms (the
    z-march tiles are 32 x 8 x 8: n = 23, 37, 40 leave partial tiles on every
    axis)."""
    print(f"[karma] n={n} steps={steps} Fnoise={Fnoise} zmarch={zmarch}", flush=True)
    R0 = min(8.0, 0.45 * n * 0.8)
    out, _ = _oracle_fields(n, steps, Fnoise, R0, record=(steps,))
    m = Karma(KarmaDeck(N=n, Fnoise=Fnoise, R0=R0))
    m.set_zmarch(zmarch)
    try:
        m.run(steps // 2)
        m.run(steps - steps // 2)
        phi, U, it = m.fields()
    finally:
        m.close()
    p, u = out[steps]
    dp, du = np.abs(phi - p).max(), np.abs(U - u).max()
    print(f"  max|dphi|={dp:.3e} max|dU|={du:.3e}", flush=True)
    assert it == steps
    assert np.array_equal(phi, p) and np.array_equal(U, u)


@pytest.mark.gpu
def test_karma_gpu_errors():
    with pytest.raises(capi.PfError):
        Karma(KarmaDeck(N=1))


def test_namelist_reader():
    """KarmaDeck.from_nml reads the group as the Fortran namelist does:
    entries in any order and case, comments, commas, d exponents; absent
    entries keep the program's defaults; unknown names are errors."""
    d = KarmaDeck.from_nml("! c\n&KARMA  n = 16, LAMBDA=3.5 ! x\n dl = 2.5d-9\n nout=7 /\n")
    assert (d.N, d.lambda_, d.Dl, d.nout) == (16, 3.5, 2.5e-9, 7)
    assert d.eps4 == KarmaDeck().eps4 and d.Fnoise == 0.0
    assert KarmaDeck.from_nml(KarmaDeck.shipped().to_nml()) == KarmaDeck.shipped()
    z = np.load(GOLD)
    g = KarmaDeck.from_nml("&karma\n" + "\n".join(f"{k} = {v}" for k, v in z["nml"]) + "\n/")
    assert (g.N, g.Fnoise, g.tip_stop, g.R0, g.nsteps, g.nout) == (16, 0.05, 2.0, 6.0, 40, 20)
    with pytest.raises(ValueError):
        KarmaDeck.from_nml("&karma\n bogus = 1\n/")
    # the scales the program prints for the golden run
    sc = g.scales()
    assert f"{sc['D']:7.3f}".strip() == "2.507"
    assert f"{sc['tau0']:10.3E}".strip() == "3.876E-07"


def test_tip_rule():
    n = 9
    phi = -np.ones((n, n, n))
    assert tips(phi) == (0, 0)
    phi[0, 0, :5] = 0.5           # x axis: cells i = 1..5 solid
    phi[0, 0, 7] = 1e-3           # an isolated positive cell further out counts
    for q in range(3):
        phi[q, q, q] = 1.0        # diagonal: i = 1..3
    assert tips(phi) == (8, 3)


@pytest.mark.skipif(not os.path.exists(EXE), reason="program not built")
def test_program_reports_missing_device(tmp_path):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    (tmp_path / "karma.nml").write_text("&karma\n N = 8\n nsteps = 2\n/\n")
    r = subprocess.run([EXE], cwd=str(tmp_path), stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode != 0 and "pf_karma_create failed" in r.stdout


def _program(tmp_path, nml_items):
    (tmp_path / "karma.nml").write_text(
        "&karma\n" + "".join(f" {k} = {v}\n" for k, v in nml_items) + "/\n")
    (tmp_path / "output").mkdir()
    return subprocess.run([EXE], cwd=str(tmp_path), stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=300)


_BANNERS = ("### CPU OpenMP", "### GPU offload ENABLED", "### gfx950 HIP device")


def _console(lines):
    """Transcript lines compared: the device banner dropped and the wall
    clock field of the progress line cut off."""
    return [str(ln).split("  wall=")[0] for ln in lines if not str(ln).startswith(_BANNERS)]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="program not built")
def test_program_matches_reference_program(tmp_path):
    """bin/karma3d on the golden's namelist: the console transcript line by
    line (wall clock and device banner masked), the file names and header
    records exactly, and the fields to the files' 8 significant digits."""
    z = np.load(GOLD)
    print("[karma program] 16^3, 40 steps", flush=True)
    r = _program(tmp_path, z["nml"])
    assert r.returncode == 0, r.stdout[-3000:]
    assert _console(r.stdout.splitlines()) == _console(z["stdout_lines"])
    names = sorted(os.listdir(tmp_path / "output"))
    assert names == [str(x) for x in z["filenames"]]
    n = 16
    for q, fn in enumerate(names):
        it = int(fn[len("karma_"):-len(".dat")])
        lines = (tmp_path / "output" / fn).read_text().splitlines()
        assert lines[:4] == [str(x) for x in z["headers"][q]]
        rows = np.array([[float(x) for x in ln.split()] for ln in lines[4:]])
        assert rows.shape == (n ** 3, 6)
        np.testing.assert_allclose(rows[:, 3].reshape(n, n, n), z[f"phase_{it}"], rtol=6e-8,
                                   atol=1e-300)
        np.testing.assert_allclose(rows[:, 5].reshape(n, n, n), z[f"conc_{it}"], rtol=6e-8)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="program not built")
def test_program_tip_stop_exit(tmp_path):
    """tip_stop = 0.3: the x-axis tip (i = 8 > 0.3 * 16) ends the run at the
    first output step, with the reference's message and one file."""
    z = np.load(GOLD)
    print("[karma program] tip_stop exit", flush=True)
    r = _program(tmp_path, z["stop_nml"])
    assert r.returncode == 0, r.stdout[-3000:]
    assert _console(r.stdout.splitlines()) == _console(z["stop_stdout_lines"])
    assert sorted(os.listdir(tmp_path / "output")) == [str(x) for x in z["stop_filenames"]]
