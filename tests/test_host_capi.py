"""CPU-side checks of libpfamd.so: the exported C-ABI, the host scalar
physics (bit-exact with the reference fixtures), the initial fields, the deck
parser and the closed form of the boundary decorations.  No GPU needed."""
import os
import re

import numpy as np
import pytest

from helpers import deck_from_json, load_golden
from pfamd import capi
from pfamd.deck import Deck, read_deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def test_library_exports_every_declared_symbol(lib):
    hdr = open(os.path.join(ROOT, "include", "pf_capi.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    assert declared == set(capi.EXPORTS), declared ^ set(capi.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_noise_bit_exact_with_reference(lib):
    z = load_golden("random_kat.npz")
    for t, pr, rs in zip(z["tuples"], z["pseudo_random"], z["ruido_suave"]):
        i, j, k, it, salt = (int(x) for x in t)
        assert lib.pf_pseudo_random(i, j, k, it, salt, 0) == pr
        assert lib.pf_ruido_suave(i, j, k, it, 0) == rs


def test_scalars_bit_exact_with_reference(lib):
    z = load_golden("scalars_kat.npz")
    cols = list(z["columns"])
    variants = [deck_from_json(s) for s in z["variants"]]
    for row in z["rows"]:
        r = dict(zip(cols, row))
        d = variants[int(r["variant"])].resolved()
        dt = lib.pf_time_step(d.dx, d.D_liq)
        assert dt == r["dt"]
        assert lib.pf_noise_scale(d.AAA, dt, d.dx, d.D_liq, d.ruido_tipo) == r["AAA"]
        W0, E0, idx = capi.w0_e0(d.Thin_Interface * d.dx, d.ajuste, d.Sigma, lib)
        assert (W0, E0) == (r["W0"], r["E0"])
        cs, cl = capi.equilibrium(r["TP"], d.Tm, d.Ke, d.Me, lib)
        assert (cs, cl) == (r["Cs_eq"], r["Cl_eq"])
        M0 = lib.pf_auxiliar_m0(r["TP"], cl, cs, d.R, d.Vm, E0, W0, d.Sigma, d.Beta, d.Ke, d.Me,
                                d.D_liq, d.ajuste)
        assert M0 == r["M0"]


@pytest.mark.parametrize("name", ["principal_planar.npz", "principal_front_j.npz"])
def test_w0_e0_try_trace_matches_reference_transcript(lib, name):
    """W0_E0.f90:19 prints |E0 - EPSLON_Try| for every try: the unmodified
    principal.f90's count of such lines and a sample of them (list-directed,
    shortest round-trip digits, so the parse is exact) against
    pf_w0_e0_tries."""
    z = load_golden(name)
    d = deck_from_json(z["deck"]).resolved()
    diff = capi.w0_e0_tries(d.Thin_Interface * d.dx, d.ajuste, d.Sigma, lib)
    assert len(diff) == int(z["w0e0_tries"])
    for q, ln in zip(z["w0e0_sample_index"], z["w0e0_sample_lines"]):
        label, val = str(ln).split()
        assert label == "E0-EPSLON_Try==>"
        assert diff[int(q)] == float(val), (q, ln)
    assert np.isinf(diff[0]) and diff[-1] < 1e-5 <= diff[-2]


@pytest.mark.parametrize("tipo,R", [(0, 0.0), (1, 3.0), (2, 4.0), (3, 0.0), (5, 4.0), (6, 3.0)])
def test_initial_fields_match_oracle(lib, oracle_lib, tipo, R):
    nx, ny, nz = 17, 15, 12
    from pfamd.run import initial_fields

    d = Deck(Nx=nx, Ny=ny, Nz=nz, tipo_inicial=tipo, R_semente=R)
    mine = initial_fields(d, n_seeds=5, seed_rng=20261203, lib=lib)
    ref = oracle_lib.initial_fields(nx, ny, nz, tipo, R, d.resolved().T0, d.C0, d.Ke, 5, 20261203)
    for a, b in zip(mine, ref.as_list()):
        assert np.array_equal(a, b)
    if tipo in (1, 2, 5, 6):
        assert mine[0].sum() > 0  # a seed exists


def test_error_paths(lib):
    p = capi.PfParams()
    p.nx, p.ny, p.nz = 4, 8, 8
    h = capi.C.c_void_p()
    rc = lib.pf_init(capi.C.byref(p), None, None, None, None, None, None, None, capi.C.byref(h))
    assert rc == capi.PF_E_ARG and b"must be >= 5" in lib.pf_last_error()
    assert lib.pf_temperatura(None, capi.C.byref(p)) == capi.PF_E_ARG
    if capi.device_count(lib) == 0:
        p.nx = 8
        rc = lib.pf_init(capi.C.byref(p), None, None, None, None, None, None, None,
                         capi.C.byref(h))
        assert rc == capi.PF_E_NODEV


def test_deck_round_trip_and_reference_deck(tmp_path):
    d = Deck(Nx=33, Ny=34, Nz=35, TP=911.5, tipo_inicial=1, R_semente=4.0, taxa_resfr=0.0,
             n_varreduras=2, h_quintico=1, ruido_tipo=2)
    f = tmp_path / "entrada.dat"
    d.write(str(f))
    assert read_deck(str(f)) == d
    # the reference's own run6 deck (a data fixture copied from entrada.dat)
    r = read_deck(os.path.join(ROOT, "tests", "golden", "entrada_run6.dat"))
    assert (r.Nx, r.Ny, r.Nz) == (340, 340, 700)
    assert r.tipo_inicial == 3 and r.taxa_resfr == 19.84 and r.n_varreduras == 2
    assert (r.h_quintico, r.ruido_tipo) == (0, 0)
    assert r.Ke == float(np.float32(0.14)) and r.WriteFileD_timeCounter == 7500
    # optional records missing -> reference defaults (entrada.f90:61-89)
    lines = open(str(f)).read().splitlines()
    (tmp_path / "short.dat").write_text("\n".join(lines[:39]) + "\n")
    s = read_deck(str(tmp_path / "short.dat")).resolved()
    assert (s.TP, s.tipo_inicial, s.taxa_resfr, s.n_varreduras, s.h_quintico, s.ruido_tipo) == \
        (919.0, 0, 0.0, 1, 1, 2)


def _fortran_phase_decoration(F0):
    """equacao_campo_fase.f90:151-167 literally: sequential array sections."""
    f = F0.copy()  # index [i,j,k], 0-based
    nx, ny, nz = f.shape
    f[:, 0, :] = f[:, 1, :]; f[:, 1, :] = f[:, 2, :]
    f[:, ny - 1, :] = f[:, ny - 2, :]; f[:, ny - 2, :] = f[:, ny - 3, :]
    f[0, :, :] = f[1, :, :]; f[1, :, :] = f[2, :, :]
    f[nx - 1, :, :] = f[nx - 2, :, :]; f[nx - 2, :, :] = f[nx - 3, :, :]
    f[:, :, 0] = f[:, :, 1]; f[:, :, 1] = f[:, :, 2]
    f[:, :, nz - 1] = f[:, :, nz - 2]; f[:, :, nz - 2] = f[:, :, nz - 3]
    return f


def _phase_map(i, n):  # 1-based, as pf_common.h phase_map
    return i + 1 if i <= 2 else (i - 1 if i >= n - 1 else i)


@pytest.mark.parametrize("shape", [(5, 5, 5), (6, 7, 5), (9, 8, 10), (12, 5, 7)])
def test_phase_decoration_closed_form(shape):
    rng = np.random.default_rng(sum(shape))
    F0 = rng.standard_normal(shape)
    ref = _fortran_phase_decoration(F0)
    nx, ny, nz = shape
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            for k in range(1, nz + 1):
                src = (_phase_map(i, nx) - 1, _phase_map(j, ny) - 1, _phase_map(k, nz) - 1)
                assert ref[i - 1, j - 1, k - 1] == F0[src]


def _fortran_solute_decoration(S):
    """concentracao_soluto.f90:116-124 literally (snapshot S = c_old)."""
    c = S.copy()
    nx, ny, nz = S.shape
    c[:, 0, :] = S[:, 1, :]; c[:, ny - 1, :] = S[:, ny - 2, :]
    c[0, :, :] = S[1, :, :]; c[nx - 1, :, :] = S[nx - 2, :, :]
    c[:, :, 0] = S[:, :, 1]; c[:, :, nz - 1] = S[:, :, nz - 2]
    return c


@pytest.mark.parametrize("shape", [(5, 5, 5), (6, 7, 5), (9, 8, 10)])
def test_solute_decoration_closed_form(shape):
    rng = np.random.default_rng(7 + sum(shape))
    S = rng.standard_normal(shape)
    ref = _fortran_solute_decoration(S)
    nx, ny, nz = shape
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            for k in range(1, nz + 1):
                si, sj, sk = i, j, k
                if k in (1, nz):
                    sk = 2 if k == 1 else nz - 1
                elif i in (1, nx):
                    si = 2 if i == 1 else nx - 1
                elif j in (1, ny):
                    sj = 2 if j == 1 else ny - 1
                assert ref[i - 1, j - 1, k - 1] == S[si - 1, sj - 1, sk - 1]
