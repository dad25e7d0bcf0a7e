"""Pins the C oracle against the reference compiled from source.

Every fixture in tests/golden/ was written by tests/golden/make_golden.py from
the reference's own Fortran (flang -O2) -- its kernels driven by our harness,
and its unmodified principal.f90.  The oracle must reproduce them bit for bit
(NaN positions included); diagnostics sums, whose order the reference leaves
to OpenMP, to 1e-14 relative.
"""
import glob
import json
import os

import numpy as np
import pytest

from helpers import (FIELDS, SCALARS, deck_from_json, fortran_e, initial_from_deck,
                     load_golden, params_from_deck, same_bits)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_pseudo_random_kat(oracle_lib):
    O = oracle_lib
    z = load_golden("random_kat.npz")
    for t, pr, rs, r4 in zip(z["tuples"], z["pseudo_random"], z["ruido_suave"], z["ran_real4"]):
        i, j, k, it, salt = (int(x) for x in t)
        assert O.pseudo_random(i, j, k, it, salt) == pr
        assert O.ruido_suave(i, j, k, it) == rs
        assert float(np.float32(O.pseudo_random(i, j, k, it, salt))) == r4


def test_pseudo_random_properties(oracle_lib):
    """Properties fixed by the algorithm, independent of the fixtures: range
    [0,1], the seed term vanishes at seed 0 and changes the stream otherwise,
    and ruido_suave reduces to the lattice value at lattice points."""
    O = oracle_lib
    vals = [O.pseudo_random(i, 3, 5, 7, 1) for i in range(1, 2000)]
    assert min(vals) >= 0.0 and max(vals) <= 1.0
    assert abs(np.mean(vals) - 0.5) < 0.03
    assert O.pseudo_random(4, 5, 6, 7, 1, 0) != O.pseudo_random(4, 5, 6, 7, 1, 1)
    for (i, j, k) in [(6, 12, 18), (0, 0, 0), (36, 6, 60)]:
        assert O.ruido_suave(i, j, k, 45) == O.pseudo_random(i // 6, j // 6, k // 6, 45 // 20, 3)


def test_scalars_kat(oracle_lib):
    O = oracle_lib
    z = load_golden("scalars_kat.npz")
    cols = list(z["columns"])
    variants = [deck_from_json(s) for s in z["variants"]]
    for row in z["rows"]:
        r = dict(zip(cols, row))
        d = variants[int(r["variant"])].resolved()
        dt = O.load().pfo_delta_t(d.dx, d.D_liq)
        assert dt == r["dt"]
        assert O.load().pfo_noise_scale(d.AAA, dt, d.dx, d.D_liq, d.ruido_tipo) == r["AAA"]
        W0, E0, idx = O.w0_e0(d.Thin_Interface * d.dx, d.ajuste, d.Sigma)
        assert (W0, E0) == (r["W0"], r["E0"]) and idx >= 0
        TP = r["TP"]
        Ke = float(np.float32(d.Ke))
        cs, cl = ((d.Tm - TP) * Ke) / d.Me, (d.Tm - TP) / d.Me
        assert (cs, cl) == (r["Cs_eq"], r["Cl_eq"])
        M0 = O.auxiliar(TP, cl, cs, d.R, d.Vm, E0, W0, d.Sigma, d.Beta, d.Ke, d.Me, d.D_liq,
                        d.ajuste)
        assert M0 == r["M0"], (M0, r["M0"])


def test_w0_e0_closed_form(oracle_lib):
    """W0_E0 stops at the first grid point within 1e-5 of the fixed point
    E0 = ajuste*Sigma*3*sqrt(2)/sqrt(W0(E0)); check against the closed form
    E0* = sqrt(ajuste*Sigma*3*sqrt(2)*Lambda/1.5556349186104)."""
    O = oracle_lib
    lam = 4 * 1.5e-7
    W0, E0, idx = O.w0_e0(lam, 2.0, float(np.float32(0.093)))
    star = np.sqrt(float(np.float32(2.0 * np.float32(0.093))) * 4.24264068711929 * lam
                   / 1.5556349186104)
    assert abs(E0 - star) / star < 0.02
    t = idx * float(np.float32(1e-7))
    assert abs(E0 - t) < 1e-5
    assert np.isclose(W0, (1.5556349186104 / lam * E0) ** 2, rtol=0.05)


UNIT = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "unit_*.npz")))


@pytest.mark.parametrize("name", UNIT)
def test_unit_call_bitwise(oracle_lib, name):
    O = oracle_lib
    z = load_golden(name)
    deck = deck_from_json(z["deck"])
    nml = json.loads(str(z["nml"]))
    p, L = params_from_deck(deck, O)
    if "it_set" in nml:
        p.it = int(round(nml["it_set"]))
    if "tp_set" in nml:
        p.TP = nml["tp_set"]
    F = O.Fields(None, [z[f"in_{nm}"] for nm in FIELDS])
    call = str(z["call"])
    if call == "temperatura":
        O.temperatura(p, F)
        assert same_bits(F.T_new, z["out_T_new"])
    elif call == "fase":
        probe = O.equacao_campo_fase(p, F, want_probe=True)
        assert same_bits(F.phi_new, z["out_phi_new"])
        assert same_bits(F.dpp, z["out_dpp"])
        # the printed probe line (format 9(1X,e14.4), equacao_campo_fase.f90:174)
        assert [fortran_e(v, 4) for v in probe] == [fortran_e(v, 4) for v in z["probe_print"]]
    elif call == "soluto":
        O.concentracao_soluto(p, F)
        assert same_bits(F.c_new, z["out_c_new"])
        assert same_bits(F.c_old, z["out_c_old"])
    elif call == "diag":
        out = O.diagnosticos(p, F.phi_new, F.c_new)
        ref = z["scalars"][SCALARS.index("L2"):SCALARS.index("L2") + 5]
        np.testing.assert_allclose(out, ref, rtol=1e-14)
    else:
        raise AssertionError(call)


RUNS = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "run_*.npz")))


@pytest.mark.parametrize("name", RUNS)
def test_time_loop_bitwise(oracle_lib, name):
    O = oracle_lib
    z = load_golden(name)
    deck = deck_from_json(z["deck"])
    p, L = params_from_deck(deck, O)
    F = initial_from_deck(deck, O)
    done = 0
    for st in z["dumps"]:
        O.run_steps(p, L, int(st) - done, F)
        done = int(st)
        # state after the step == what the harness dumped before the end-of-step copy
        for nm in ("phi_new", "c_new", "T_new", "dpp"):
            assert same_bits(getattr(F, nm), z[f"s{st}_{nm}"]), (name, st, nm)
        s = dict(zip(SCALARS, z[f"s{st}_scalars"]))
        assert p.it == s["it"]
        assert float(L.time) == s["time"]
        assert (p.TP, p.M0, p.Cs_eq, p.Cl_eq) == (s["TP"], s["M0"], s["Cs_eq"], s["Cl_eq"])
        diag = O.diagnosticos(p, F.phi_new, F.c_new)
        ref = np.array([s["L2"], s["phi_max"], s["phi_min"], s["sum_c"], s["free_energy"]])
        np.testing.assert_allclose(diag, ref, rtol=1e-14, equal_nan=True)


@pytest.mark.parametrize("name", ["principal_sphere.npz", "principal_planar.npz",
                                  "principal_front_j.npz", "principal_stop.npz"])
def test_unmodified_principal_output_event(oracle_lib, name):
    """The unmodified principal.f90: its output events (it = N+1, 2N+1, or
    the stoptime step) write phi, TP and c with 8 significant digits; the
    oracle's loop must agree to that precision (NaN included) -- this pins
    the loop order (TP lag, deck M0 at step 1, REAL(4) time) against the real
    driver, not only our harness.  principal_front_j starts from the
    original front (tipo_inicial = 0), so it pins that
    initial condition too (entrada.f90:237-260)."""
    O = oracle_lib
    z = load_golden(name)
    deck = deck_from_json(z["deck"])
    p, L = params_from_deck(deck, O)
    F = initial_from_deck(deck, O)
    done = 0
    for q, it_ev in enumerate(int(x) for x in z["event_its"]):
        O.run_steps(p, L, it_ev - done, F)
        done = it_ev
        data = z[f"e{q}_xyz_phi_tp_c"]
        np.testing.assert_allclose(F.phi_new, data[:, 3], rtol=1e-7, atol=1e-300, equal_nan=True)
        np.testing.assert_allclose(F.c_new, data[:, 5], rtol=1e-7, atol=1e-300, equal_nan=True)
        np.testing.assert_allclose(data[:, 4], p.TP, rtol=1e-7)
        # filename: saida.f90:20-28 pads the path to 52 characters and writes
        # IT with I6 and time with f14.7 (blanks kept by trim/adjustl)
        fname = str(z["filenames"][q])
        assert fname.strip().startswith("saida_")
        assert fname.split("_time_")[0].split("saida_")[1].strip() == str(it_ev)
        assert abs(float(fname.split("_time_")[1].strip()[:-5]) - float(L.time)) < 1e-7


def _thermal_case(O, nx=6, ny=5, nz=7, D=0.37):
    deck = deck_from_json(load_golden("principal_planar.npz")["deck"])
    deck.Nx, deck.Ny, deck.Nz = nx, ny, nz
    p, _ = params_from_deck(deck, O)
    p.dx, p.dy, p.dz = 1.0e-7, 1.7e-7, 2.9e-7  # distinct, so a swapped axis shows
    p.D = D
    F = O.initial_fields(nx, ny, nz, 0, 0.0, 900.0, 0.02, 0.14)
    F.dpp[:] = 0.0  # no latent-heat source: only the D term and TP remain
    return p, F


def _temperatura(O, p, F, T):
    G = F.copy()
    G.T_old = np.ascontiguousarray(T.reshape(-1, order="F"))
    O.temperatura(p, G)
    return G.T_new.reshape((p.nx, p.ny, p.nz), order="F")


def test_thermal_term_spike_closed_form(oracle_lib):
    """temp_body.inc with D != 0 (the reference never sets D, so no fixture
    holds this branch).  With T_old = TP everywhere but one cell m, where it
    is TP + s, and no source: each face neighbour of m along x, y, z gets
    TP + D dt s / dx^2, / dy^2, / dz^2; m itself gets TP (the stencil's centre
    is TP, not T_old(m)); every other interior cell gets TP; the boundary
    keeps T_old."""
    O = oracle_lib
    p, F = _thermal_case(O)
    TP, s = p.TP, 3.0
    T = np.full((p.nx, p.ny, p.nz), TP)
    m = (2, 2, 3)
    T[m] += s
    got = _temperatura(O, p, F, T)
    want = np.full_like(T, TP)
    for ax, h in ((0, p.dx), (1, p.dy), (2, p.dz)):
        for sg in (-1, 1):
            q = list(m)
            q[ax] += sg
            want[tuple(q)] = TP + p.D * p.dt * s / h ** 2
    inner = (slice(1, -1),) * 3
    edge = np.ones_like(T, dtype=bool)
    edge[inner] = False
    np.testing.assert_allclose(got[inner], want[inner], rtol=1e-15, atol=0)
    assert np.array_equal(got[edge], T[edge])


def test_thermal_term_brute_force_superposition(oracle_lib):
    """The D term is linear in T_old - TP: on a tiny grid, a random field's
    update equals TP plus the sum, over every cell, of that cell's spike
    response (one oracle call per cell) scaled by its deviation."""
    O = oracle_lib
    p, F = _thermal_case(O, nx=5, ny=4, nz=6)
    TP = p.TP
    rng = np.random.default_rng(5)
    dev = rng.normal(0.0, 2.0, (p.nx, p.ny, p.nz))
    got = _temperatura(O, p, F, TP + dev)
    acc = np.zeros_like(dev)
    for m in np.ndindex(dev.shape):
        T = np.full_like(dev, TP)
        T[m] += 1.0
        acc += (_temperatura(O, p, F, T) - TP) * dev[m]
    inner = (slice(1, -1),) * 3
    # interior: superposition of the spike responses (boundary spikes act
    # only through their interior neighbours)
    resp = np.zeros_like(dev)
    resp[inner] = acc[inner]
    np.testing.assert_allclose(got[inner] - TP, resp[inner], rtol=1e-9, atol=1e-9 * TP)


def test_growth_deck_amplifies_one_ulp(oracle_lib):
    """The reference scheme amplifies rounding on the growth deck (64x64x128,
    one seed r=5 at the bottom centre, held at TP = 905 K, noise off): moving
    phi by +-1 ulp (a seeded random sign per cell) after step 3 and running to
    step 20 departs from the unperturbed run by >= 1e-12 in phi, four orders
    of magnitude above the perturbation -- the fast-growing tail of the
    front, where h'(phi) > 0 and the driving force make phi grow by a factor
    of two per step, carries the difference.  A form that rounds differently
    from the reference (the fused sweeps) therefore differs from the oracle
    by the same spread whatever its accuracy; the GPU tests of that deck
    bound it by a multiple of the spread (test_gpu_configs.py)."""
    from pfamd.deck import Deck

    O = oracle_lib
    O.set_threads(max(1, min(8, os.cpu_count() or 1)))
    deck = Deck(Nx=64, Ny=64, Nz=128, tipo_inicial=5, R_semente=5.0, h_quintico=0, ruido_tipo=0,
                AAA=0.0, n_varreduras=2, stoptime=0.0, taxa_resfr=0.0, TP=905.0)
    d = deck.resolved()

    def run(seed):
        p, L = params_from_deck(deck, O)
        F = O.initial_fields(d.Nx, d.Ny, d.Nz, d.tipo_inicial, d.R_semente, d.T0, d.C0, d.Ke)
        O.run_steps(p, L, 3, F)
        if seed:
            s = np.random.default_rng(seed).integers(-1, 2, size=np.asarray(F.phi_new).shape)
            for arr in (F.phi_new, F.phi_old):
                a = np.asarray(arr)
                up, dn = np.nextafter(a, np.inf), np.nextafter(a, -np.inf)
                a[:] = np.maximum(np.where(s > 0, up, np.where(s < 0, dn, a)), 0.0)
        O.run_steps(p, L, 17, F)
        return F.phi_new.copy()

    ref = run(0)
    spread = max(float(np.max(np.abs(run(seed) - ref))) for seed in (1, 2))
    assert 1e-12 <= spread <= 1e-9, spread
