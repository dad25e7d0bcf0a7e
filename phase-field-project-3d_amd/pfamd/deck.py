"""entrada.dat reader/writer.

Mirrors the reference's positional list-directed reads
(source/entrada.f90:17-89): three header records are skipped, then each
`read (1,*)` takes the first value(s) of the next record that holds a value
(a blank record is skipped, as list-directed input does), and the rest of the
record is ignored.  Records 40-43 are optional; when absent the reference's
defaults apply (entrada.f90:61-89).  Fortran `d` exponents are accepted.

REAL(4) quantities of the reference (Ke, Sigma, Beta, ajuste; variaveis.f90:28)
are returned rounded to float32 so that every consumer sees the values the
reference computes with.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

# (name, kind) in read order; kind: 'f8' double, 'f4' REAL(4), 'i' integer.
_REQUIRED = [
    ("Tm", "f8"), ("T0", "f8"), ("delta_H", "f8"), ("E0_deck", "f8"), ("W0_deck", "f8"),
    ("M0", "f8"), ("D_s", "f8"), ("D_l", "f8"), ("Cp_s", "f8"), ("Cp_l", "f8"),
    ("Nx", "i"), ("Ny", "i"), ("Nz", "i"), ("dy", "f8"), ("dx", "f8"), ("dz", "f8"),
    ("small", "f8"), ("ANI", "f8"), ("it_max", "i"), ("AAA", "f8"), ("Teta0", "f8"),
    ("DELTA_EPSLON", "f8"), ("DELTA_T_deck", "f8"), ("D_sol", "f8"), ("D_liq", "f8"),
    ("Ke", "f4"), ("Me", "f8"), ("Vm", "f8"), ("R", "f8"), ("C0", "f8"),
    ("Thin_Interface", "f8"), ("Sigma", "f4"), ("Beta", "f4"), ("ajuste", "f4"),
    ("stoptime", "f8"), ("WriteFileD_timeCounter", "i"),
]


def _f4(x: float) -> float:
    return float(np.float32(x))


def _num(tok: str, kind: str):
    t = tok.strip().rstrip(",")
    if kind == "i":
        return int(t)
    v = float(t.replace("d", "e").replace("D", "e"))
    return _f4(v) if kind == "f4" else v


@dataclasses.dataclass
class Deck:
    """The values `entrada` reads, before its dt/AAA rescale."""

    Tm: float = 933.3
    T0: float = 922.0
    delta_H: float = 0.891002e9
    E0_deck: float = 1.055e-3
    W0_deck: float = 673.2e3
    M0: float = 1.9e-3
    D_s: float = 6.59477338e-5
    D_l: float = 3.23646776e-5
    Cp_s: float = 2.926560e6
    Cp_l: float = 2.626320e6
    Nx: int = 64
    Ny: int = 64
    Nz: int = 128
    dy: float = 1.5e-7
    dx: float = 1.5e-7
    dz: float = 1.5e-7
    small: float = 1e-23
    ANI: float = 4.0
    it_max: int = 1000000000
    AAA: float = 0.03
    Teta0: float = 0.0
    DELTA_EPSLON: float = 0.03
    DELTA_T_deck: float = 1.5e-7
    D_sol: float = 3.0e-13
    D_liq: float = 3.0e-9
    Ke: float = _f4(0.14)
    Me: float = 640.0
    Vm: float = 1.13e-5
    R: float = 8.31451
    C0: float = 0.0196
    Thin_Interface: float = 4.0
    Sigma: float = _f4(0.093)
    Beta: float = 1.0
    ajuste: float = 2.0
    stoptime: float = 0.085
    WriteFileD_timeCounter: int = 7500
    # optional records 40-43 (None = record absent -> reference default)
    TP: Optional[float] = 905.0
    tipo_inicial: Optional[int] = 3
    R_semente: Optional[float] = 0.0
    taxa_resfr: Optional[float] = 19.84
    n_varreduras: Optional[int] = 2
    h_quintico: Optional[int] = 0
    ruido_tipo: Optional[int] = 0

    # -- what entrada derives (entrada.f90:61-89 defaults, :159-185) ---------
    def resolved(self) -> "Deck":
        """Copy with the optional-record defaults applied (entrada.f90:61-89)."""
        d = dataclasses.replace(self)
        if d.TP is None:
            d.TP = 919.0
        if d.tipo_inicial is None or d.R_semente is None:
            d.tipo_inicial, d.R_semente = 0, 0.0
        if d.taxa_resfr is None or d.n_varreduras is None:
            d.taxa_resfr, d.n_varreduras = 0.0, 1
        d.n_varreduras = max(1, int(d.n_varreduras))
        if d.taxa_resfr > 0.0:
            d.TP = d.T0
        if d.h_quintico is None or d.ruido_tipo is None:
            d.h_quintico, d.ruido_tipo = 1, 2
        return d

    def write(self, path: str) -> None:
        """Write a deck the reference's `entrada` parses back to these values."""

        def f(x: float) -> str:
            return repr(float(x)).replace("e", "d") if "e" in repr(float(x)) else repr(float(x)) + "d0"

        lines = ["+++ phase-field deck (pfamd) +++", "", ""]
        for name, kind in _REQUIRED:
            v = getattr(self, name)
            lines.append(f"{int(v)}   | {name}" if kind == "i" else f"{f(v)}   | {name}")
        opt = []
        if self.TP is not None:
            opt.append(f"{f(self.TP)}   | TP")
            if self.tipo_inicial is not None:
                opt.append(f"{int(self.tipo_inicial)}  {f(self.R_semente)}   | tipo_inicial, R_semente")
                if self.taxa_resfr is not None:
                    opt.append(f"{f(self.taxa_resfr)}  {int(self.n_varreduras)}   | taxa_resfr, n_varreduras")
                    if self.h_quintico is not None:
                        opt.append(f"{int(self.h_quintico)}  {int(self.ruido_tipo)}   | h_quintico, ruido_tipo")
        lines += opt
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")


def _value_records(lines: List[str]):
    """Yield the token lists of records that hold at least one value."""
    for ln in lines:
        body = ln.split("|", 1)[0]
        toks = body.replace(",", " ").split()
        if toks:
            yield toks


def read_deck(path: str) -> Deck:
    """Parse an entrada.dat the way entrada.f90:19-89 does."""
    with open(path, "r", encoding="latin-1") as fh:
        lines = fh.read().splitlines()
    recs = _value_records(lines[3:])
    vals = {}
    for name, kind in _REQUIRED:
        try:
            toks = next(recs)
        except StopIteration:
            raise ValueError(f"entrada.dat ends before '{name}'") from None
        vals[name] = _num(toks[0], kind)
    opt = [("TP",), ("tipo_inicial", "R_semente"), ("taxa_resfr", "n_varreduras"),
           ("h_quintico", "ruido_tipo")]
    kinds = {"TP": "f8", "tipo_inicial": "i", "R_semente": "f8", "taxa_resfr": "f8",
             "n_varreduras": "i", "h_quintico": "i", "ruido_tipo": "i"}
    for names in opt:
        try:
            toks = next(recs)
            if len(toks) < len(names):
                raise ValueError
            for nm, tk in zip(names, toks):
                vals[nm] = _num(tk, kinds[nm])
        except (StopIteration, ValueError):
            for nm in names:
                vals[nm] = None
    return Deck(**vals)


@dataclasses.dataclass
class Derived:
    """Scalars `entrada` and `W0_E0` derive from the deck."""

    dt: float
    AAA: float
    Lambda: float
    W0: float
    E0: float
    w0e0_index: int


def derive(deck: Deck, lib=None) -> Derived:
    """dt, AAA rescale, Lambda (entrada.f90:161-185) and W0/E0 (W0_E0.f90)
    computed through the library's host helpers (bit-exact with the
    reference)."""
    from . import capi

    lib = lib or capi.load()
    d = deck.resolved()
    dt = lib.pf_time_step(d.dx, d.D_liq)
    aaa = lib.pf_noise_scale(d.AAA, dt, d.dx, d.D_liq, int(d.ruido_tipo))
    lam = d.Thin_Interface * d.dx
    W0, E0, idx = capi.w0_e0(lam, d.ajuste, d.Sigma, lib)
    return Derived(dt=dt, AAA=aaa, Lambda=lam, W0=W0, E0=E0, w0e0_index=idx)


__all__ = ["Deck", "read_deck", "derive", "Derived"]
