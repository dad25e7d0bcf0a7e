"""Ensembles of independent solidification problems, solved batched
(pf_batch_*: one launch per sweep for all of them) or one after another
(pf_run_steps per problem).

The ensemble configuration: 64 instances with grid edges drawn
log-uniformly from 32-160 per axis (z >= x, y), cooling rates uniform in
5-50 K/s, 1-8 seeds of radius 4 on the bottom plane, effective noise
amplitude uniform in 0-0.02, FP32-mixed, 5,000 steps each, everything drawn
from seed 7.  The deck's other parameters are the run6 Al-Cu deck's.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import capi
from .deck import Deck
from .run import Simulation, initial_fields, setup_from_deck

# effective/deck noise amplitude for ruido_tipo >= 1: sqrt(dt / (dx^2/(5 D)))
# with dt = dx^2/(7.5 D) (entrada.f90:173-185)
NOISE_SCALE = math.sqrt(5.0 / 7.5)


@dataclass
class Problem:
    """One solidification run: deck, precision, noise seed, seeded initial
    condition and step count."""
    deck: Deck
    precision: int = capi.PF_FP32_MIXED
    noise_seed: int = 1
    n_seeds: int = 1
    seed_rng: int = 0
    steps: int = 5000
    tag: str = ""

    @property
    def cells(self) -> int:
        return self.deck.Nx * self.deck.Ny * self.deck.Nz

    def setup(self, lib=None):
        s = setup_from_deck(self.deck, self.precision, self.noise_seed, lib=lib)
        f = initial_fields(self.deck, self.n_seeds, self.seed_rng, lib)
        return s, f


def draw_problem(rng: np.random.Generator, edge_lo: int, edge_hi: int, max_seeds: int,
                 precision: int, steps: int, tag: str = "") -> Problem:
    e = np.exp(rng.uniform(math.log(edge_lo), math.log(edge_hi), size=3))
    nx, ny, nz = (int(round(v)) for v in e)
    nz = max(nz, nx, ny)
    taxa = float(rng.uniform(5.0, 50.0))
    n_seeds = int(rng.integers(1, max_seeds + 1))
    a_eff = float(rng.uniform(0.0, 0.02))
    deck = Deck(Nx=nx, Ny=ny, Nz=nz, tipo_inicial=6, R_semente=4.0, h_quintico=0, ruido_tipo=1,
                AAA=a_eff / NOISE_SCALE, taxa_resfr=taxa, n_varreduras=2, stoptime=0.0)
    return Problem(deck=deck, precision=precision, noise_seed=int(rng.integers(1, 2**31)),
                   n_seeds=n_seeds, seed_rng=int(rng.integers(1, 2**62)), steps=steps, tag=tag)


def ensemble_problems(n: int = 64, seed: int = 7, steps: int = 5000,
                      precision: int = capi.PF_FP32_MIXED) -> List[Problem]:
    rng = np.random.default_rng(seed)
    return [draw_problem(rng, 32, 160, 8, precision, steps, tag=f"e{q}") for q in range(n)]


@dataclass
class RunResult:
    wall_s: float
    cell_steps: int
    results: list = field(default_factory=list)  # (phi, c, T, diag) per problem if kept

    @property
    def solves_per_s(self) -> float:
        return len(self.results) / self.wall_s if self.results else float("nan")

    @property
    def mlups(self) -> float:
        return self.cell_steps / self.wall_s / 1e6


def run_batched(problems: List[Problem], graph_steps: int = 16, autotune: bool = True,
                keep: bool = False, lib=None) -> RunResult:
    """All problems in one batch, timed from the first admission to the last
    completion (initial fields are prepared on the host beforehand)."""
    lib = lib or capi.load()
    prepared = [p.setup(lib) for p in problems]
    b = capi.Batch(lib=lib)
    try:
        b.set_graph_steps(graph_steps)
        t0 = time.perf_counter()
        slots = [b.add(s.p, s.L, pr.steps, f[0], f[2], f[4]) for pr, (s, f) in zip(problems, prepared)]
        if autotune:
            b.autotune(reps=1)
        b.run(max(pr.steps for pr in problems))
        b.synchronize()
        wall = time.perf_counter() - t0
        out = []
        for pr, sl in zip(problems, slots):
            if keep:
                phi, c, T = b.download(sl, pr.cells)
                out.append((phi, c, T, b.diagnosticos(sl)))
            else:
                out.append(None)
        return RunResult(wall, sum(pr.cells * pr.steps for pr in problems), out)
    finally:
        b.close()


def run_sequential(problems: List[Problem], keep: bool = False, lib=None) -> RunResult:
    """Each problem alone through the single-instance fused loop, one after
    another, timed from its upload to its completion."""
    lib = lib or capi.load()
    prepared = [p.setup(lib) for p in problems]
    wall, out = 0.0, []
    for pr, (s, f) in zip(problems, prepared):
        t0 = time.perf_counter()
        sim = Simulation(pr.deck, precision=pr.precision, seed=pr.noise_seed, fields=f, lib=lib)
        sim.run(pr.steps)
        sim.ctx.synchronize()
        wall += time.perf_counter() - t0
        if keep:
            g = sim.fields()
            out.append((g["phi_new"], g["c_new"], g["T_new"], sim.diagnostics()))
        else:
            out.append(None)
        sim.close()
    return RunResult(wall, sum(pr.cells * pr.steps for pr in problems), out)
