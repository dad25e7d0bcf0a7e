"""Shared helpers for the tests: deck -> pf_params / pf_loop, fixtures."""
from __future__ import annotations

import json
import os

import numpy as np

from pfamd.capi import PfLoop, PfParams
from pfamd.deck import Deck

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("phi_old", "phi_new", "c_old", "c_new", "T_old", "T_new", "dpp")
SCALARS = ("it", "time", "TP", "M0", "E0", "W0", "Cs_eq", "Cl_eq", "AAA", "DELTA_T",
           "L2", "phi_max", "phi_min", "sum_c", "free_energy", "n_varreduras")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def deck_from_json(s) -> Deck:
    d = json.loads(str(s))
    return Deck(**d)


def params_from_deck(deck: Deck, O, precision=0, seed=0):
    """What entrada + W0_E0 leave in the module for the kernels, computed with
    the oracle's scalar helpers (the oracle is the checker in tests)."""
    d = deck.resolved()
    p = PfParams()
    p.nx, p.ny, p.nz = d.Nx, d.Ny, d.Nz
    p.it = 0
    p.n_varreduras = d.n_varreduras
    p.h_quintico = d.h_quintico
    p.ruido_tipo = d.ruido_tipo
    p.precision = precision
    p.seed = seed
    p.dx, p.dy, p.dz = d.dx, d.dy, d.dz
    p.dt = O.load().pfo_delta_t(d.dx, d.D_liq)
    p.D = 0.0
    p.delta_H, p.cp_s = d.delta_H, d.Cp_s
    p.TP = d.TP
    p.Tm, p.Me, p.Vm, p.R = d.Tm, d.Me, d.Vm, d.R
    lam = d.Thin_Interface * d.dx
    W0, E0, _ = O.w0_e0(lam, d.ajuste, d.Sigma)
    p.E0, p.W0, p.M0 = E0, W0, d.M0
    p.delta_eps = d.DELTA_EPSLON
    p.AAA = O.load().pfo_noise_scale(d.AAA, p.dt, d.dx, d.D_liq, d.ruido_tipo)
    p.D_liq, p.D_sol = d.D_liq, d.D_sol
    p.Ke = d.Ke
    p.Cs_eq = ((d.Tm - p.TP) * float(np.float32(d.Ke))) / d.Me
    p.Cl_eq = (d.Tm - p.TP) / d.Me
    L = PfLoop()
    L.T0, L.taxa_resfr = d.T0, d.taxa_resfr
    L.Sigma, L.Beta, L.ajuste, L.time = d.Sigma, d.Beta, d.ajuste, 0.0
    return p, L


def initial_from_deck(deck: Deck, O):
    d = deck.resolved()
    return O.initial_fields(d.Nx, d.Ny, d.Nz, d.tipo_inicial, d.R_semente, d.T0, d.C0, d.Ke)


def same_bits(a, b):
    """Bitwise equality that treats NaN == NaN (same payload not required)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a.view(np.int64) == b.view(np.int64)) | both_nan))


def max_abs(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b))
    return float(np.nanmax(d)) if d.size else 0.0


def fortran_e(x: float, d: int) -> str:
    """Fortran Ew.d edit descriptor mantissa/exponent text: 0.dddd E+ee
    (returned as 'd.dddd' digits + exponent, blanks stripped)."""
    if x == 0.0:
        return "0." + "0" * d + "E+00"
    if x != x:
        return "NaN"
    from decimal import Decimal, ROUND_HALF_EVEN
    sign = "-" if x < 0 else ""
    v = Decimal(abs(x))
    e = v.adjusted() + 1
    m = (v.scaleb(-e)).quantize(Decimal(1).scaleb(-d), rounding=ROUND_HALF_EVEN)
    if m >= 1:
        e += 1
        m = (v.scaleb(-e)).quantize(Decimal(1).scaleb(-d), rounding=ROUND_HALF_EVEN)
    es = f"{'+' if e >= 0 else '-'}{abs(e):02d}"
    return f"{sign}{m}E{es}"
