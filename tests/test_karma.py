"""The quantitative Karma model (karma/karma3d.f90).

CPU: the oracle (oracle/karma_oracle.c) against the reference program
compiled from karma/karma3d.f90 (tests/golden/karma_n16.npz: its Tecplot
output at steps 20 and 40 of a 16^3 run with noise, 8 significant digits),
the mirror ghosts against the literal x, y, z passes, the noise against the
main model's hash, and the library's exports.  GPU: the HIP step against the
oracle."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import oracle as O
from pfamd import capi
from pfamd.karma import EXPORTS, Karma, KarmaDeck

GOLD = os.path.join(ROOT, "tests", "golden", "karma_n16.npz")


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
@pytest.mark.parametrize("n,steps,Fnoise", [(16, 40, 0.05), (23, 60, 0.0), (40, 30, 0.02)])
def test_karma_gpu_matches_oracle(n, steps, Fnoise):
    """Same operation order with FP contraction off on both sides, correctly
    rounded division and square root on the device, the seed set on the host:
    the GPU fields equal the oracle's bit for bit."""
    print(f"[karma] n={n} steps={steps} Fnoise={Fnoise}", flush=True)
    R0 = min(8.0, 0.45 * n * 0.8)
    out, _ = _oracle_fields(n, steps, Fnoise, R0, record=(steps,))
    m = Karma(KarmaDeck(N=n, Fnoise=Fnoise, R0=R0))
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
