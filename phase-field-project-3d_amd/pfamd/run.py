"""Host-side set-up and time loop around the C-ABI.

This mirrors what the reference's `entrada`, `W0_E0` and `principal` do on
the host (source/entrada.f90:159-262, source/W0_E0.f90, source/principal.f90:
93-259) using the library's own host helpers, and drives the device either
through the drop-in calls (one per reference subroutine) or through the fused
pf_run_steps path.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from . import capi
from .capi import PfLoop, PfParams
from .deck import Deck

FIELD_NAMES = ("phi_old", "phi_new", "c_old", "c_new", "T_old", "T_new", "dpp")


@dataclasses.dataclass
class Setup:
    p: PfParams
    L: PfLoop
    deck: Deck
    w0e0_index: int


def setup_from_deck(deck: Deck, precision: int = capi.PF_FP64, seed: int = 0,
                    D: float = 0.0, lib=None) -> Setup:
    """pf_params / pf_loop as `entrada` + `W0_E0` leave the module."""
    lib = lib or capi.load()
    d = deck.resolved()
    p = PfParams()
    p.nx, p.ny, p.nz = d.Nx, d.Ny, d.Nz
    p.it = 0
    p.n_varreduras = int(d.n_varreduras)
    p.h_quintico = int(d.h_quintico)
    p.ruido_tipo = int(d.ruido_tipo)
    p.precision = int(precision)
    p.seed = int(seed)
    p.dx, p.dy, p.dz = d.dx, d.dy, d.dz
    p.dt = lib.pf_time_step(d.dx, d.D_liq)
    p.D = float(D)
    p.delta_H, p.cp_s = d.delta_H, d.Cp_s
    p.TP = d.TP
    p.Tm, p.Me, p.Vm, p.R = d.Tm, d.Me, d.Vm, d.R
    W0, E0, idx = capi.w0_e0(d.Thin_Interface * d.dx, d.ajuste, d.Sigma, lib)
    p.E0, p.W0, p.M0 = E0, W0, d.M0
    p.delta_eps = d.DELTA_EPSLON
    p.AAA = lib.pf_noise_scale(d.AAA, p.dt, d.dx, d.D_liq, int(d.ruido_tipo))
    p.D_liq, p.D_sol = d.D_liq, d.D_sol
    p.Ke = d.Ke
    p.Cs_eq, p.Cl_eq = capi.equilibrium(p.TP, p.Tm, p.Ke, p.Me, lib)
    L = PfLoop()
    L.T0, L.taxa_resfr = d.T0, d.taxa_resfr
    L.Sigma, L.Beta, L.ajuste, L.time = d.Sigma, d.Beta, d.ajuste, 0.0
    return Setup(p=p, L=L, deck=d, w0e0_index=idx)


def initial_fields(deck: Deck, n_seeds: int = 0, seed_rng: int = 0, lib=None):
    """The seven initial fields (entrada.f90:226-262 and the seeded modes)."""
    lib = lib or capi.load()
    d = deck.resolved()
    n = d.Nx * d.Ny * d.Nz
    arrs = [np.empty(n) for _ in range(7)]
    capi.check(lib.pf_initial_fields(d.Nx, d.Ny, d.Nz, int(d.tipo_inicial), float(d.R_semente),
                                     d.T0, d.C0, d.Ke, int(n_seeds), int(seed_rng), *arrs), lib)
    return arrs


class Simulation:
    """One device-resident instance: the reference's module state + fields."""

    def __init__(self, deck: Deck, precision: int = capi.PF_FP64, seed: int = 0,
                 n_seeds: int = 0, seed_rng: int = 0, fields=None, D: float = 0.0, lib=None):
        self.lib = lib or capi.load()
        s = setup_from_deck(deck, precision, seed, D, self.lib)
        self.p, self.L, self.deck = s.p, s.L, s.deck
        if fields is None:
            fields = initial_fields(deck, n_seeds, seed_rng, self.lib)
        self.ctx = capi.Context(self.p, fields, self.lib)

    # principal.f90:93-226, drop-in calls in the reference's order
    def step_dropin(self, nsteps: int = 1, probes: Optional[list] = None):
        p, L, ctx = self.p, self.L, self.ctx
        for _ in range(nsteps):
            p.it += 1
            L.time = float(np.float32(p.dt + float(L.time)))
            ctx.temperatura(p)
            if L.taxa_resfr > 0.0:
                p.TP = L.T0 - L.taxa_resfr * p.dt * float(p.it)
            p.Cs_eq, p.Cl_eq = capi.equilibrium(p.TP, p.Tm, p.Ke, p.Me, self.lib)
            for sw in range(max(1, p.n_varreduras)):
                if sw > 0:
                    ctx.phase_new_to_old()
                pr = ctx.equacao_campo_fase(p, want_probe=probes is not None)
                if probes is not None:
                    probes.append(pr)
            ctx.concentracao_soluto(p)
            p.M0 = self.lib.pf_auxiliar_m0(p.TP, p.Cl_eq, p.Cs_eq, p.R, p.Vm, p.E0, p.W0,
                                           L.Sigma, L.Beta, p.Ke, p.Me, p.D_liq, L.ajuste)
            ctx.end_step()

    def run(self, nsteps: int, probes: bool = False):
        return self.ctx.run_steps(self.p, self.L, nsteps, probes)

    def diagnostics(self):
        return self.ctx.diagnosticos(self.p)

    def fields(self):
        return dict(zip(FIELD_NAMES, self.ctx.download_all()))

    def save(self, path: str):
        """Exact checkpoint (fields in their stored precision, scalars,
        tiling, solute-mass ledger)."""
        self.ctx.checkpoint_save(self.p, self.L, path)

    @classmethod
    def resume(cls, path: str, deck: Optional[Deck] = None, lib=None) -> "Simulation":
        """A Simulation continuing from a checkpoint written by save()."""
        self = cls.__new__(cls)
        self.lib = lib or capi.load()
        self.ctx, self.p, self.L = capi.Context.from_checkpoint(path, self.lib)
        self.deck = deck
        return self

    def close(self):
        self.ctx.close()
