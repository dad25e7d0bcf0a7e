"""Benchmark of the coupled phase-field time step on one MI355X.

Workload (the columnar-growth configuration): 256 x 256 x 528 cells, Al-Cu
(run6 deck parameters: c0 = 0.0196 = 4.5 wt% Cu, T0 = 922 K, cooling 19.84
K/s), 16 seeds of radius 4 on the z = 1 plane at SplitMix64(20261203)
positions, zero-mean cell noise with effective amplitude 0.01 (deck
0.0122474) from the counter-based stream with seed 1, two phase sweeps per
step, cubic h(phi), diagnostics every 1000 steps (inside the timed region).

One "step" is one reference time step (two phase sweeps, solute,
temperature, mobility update) over the whole grid with the fields resident
in HBM.  Prints one JSON line: MLUPS of the FP64 drop-in form (value), the
FP32-mixed form beside it (and its accuracy against FP64 on a small grid),
the roofline of the dominant kernel from HIP events, the batched ensemble
(solves/s batched vs one after another), the serving loop (p50/p99 and
solves/s under a Poisson stream) and the reference's own CPU path timed on
this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "phase-field-project-3d_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

NX, NY, NZ = 256, 256, 528
N_SEEDS, SEED_RNG, NOISE_SEED = 16, 20261203, 1
DIAG_EVERY = 1000
HBM_PEAK_GBS = 8000.0
# Compulsory HBM bytes per cell of each fused kernel (FP64; half for FP32):
#   sweep A reads phi, c and writes phi                       -> 3 fields
#   sweep B reads phi, c and writes phi, c, T (T is not read:
#   the temperature update's Laplacian has D = 0)              -> 5 fields
# fields moved per cell by each sweep kernel (compulsory: every field read or
# written once); sweep B reads phi, c and writes phi, c, T when fused and
# phi, T when split, with the solute kernel reading phi, c and writing c
FIELDS_PER_CELL = {"sweep_A": 3, "sweep_B": 5, "solute": 3}
FIELDS_PER_CELL_SPLIT = {"sweep_A": 3, "sweep_B": 4, "solute": 3}


def workload_deck():
    from pfamd.deck import Deck

    return Deck(Nx=NX, Ny=NY, Nz=NZ, tipo_inicial=6, R_semente=4.0, h_quintico=0, ruido_tipo=1,
                AAA=0.0122474, taxa_resfr=19.84, n_varreduras=2, stoptime=0.0,
                WriteFileD_timeCounter=DIAG_EVERY)


_TORCH = None


def _torch_init():
    """Imports torch (for its device synchronize) before any GPU work: an
    import between the tiling and the timed steps would leave the GPU idle
    long enough for its clocks to drop, and a window of a few dozen steps is
    too short for them to come back (measured: 20 steps after 0.2 s idle
    1.10 ms/step, without the idle 1.00 ms/step)."""
    global _TORCH
    try:
        import torch

        _TORCH = torch if torch.cuda.is_available() else None
    except Exception:
        _TORCH = None


def _sync(sim):
    sim.ctx.synchronize()
    if _TORCH is not None:
        _TORCH.cuda.synchronize()


def parse_tiling(spec):
    """'tA:gA,tB:gB,tS:gS,split' (tile shape : work-group target of sweep A,
    sweep B and the split form's solute kernel, then 0/1 for the split form)
    -> ((tA, gA), (tB, gB), (tS, gS), split): a fixed tiling in place of the
    autotuner, so that a profiling run launches the same kernels as a timed
    one (the bench line's tiling.spec gives the autotuner's pick in this
    form)."""
    try:
        parts = spec.split(",")
        kinds = tuple(tuple(int(x) for x in p.split(":")) for p in parts[:3])
        split = int(parts[3])
        assert len(parts) == 4 and all(len(k) == 2 for k in kinds)
    except (ValueError, AssertionError, IndexError):
        raise SystemExit("--tiling wants tA:gA,tB:gB,tS:gS,split")
    return kinds + (split,)


def tiling_spec(sim):
    k = [sim.ctx.sweep_tiling(q) for q in range(3)]
    return ",".join(f"{t}:{g}" for t, g, _ in k) + f",{int(sim.ctx.sweep_split()[0])}"


def run_form(precision, steps, warmup, timing=True, autotune=True, tiling=None):
    """Warm-up, then exactly `steps` timed steps; returns timings."""
    from pfamd.run import Simulation

    deck = workload_deck()
    t0 = time.time()
    sim = Simulation(deck, precision=precision, seed=NOISE_SEED, n_seeds=N_SEEDS,
                     seed_rng=SEED_RNG)
    t_init = time.time() - t0
    # pick the sweep tiling on this grid (times every candidate on the
    # initial fields, writing into scratch buffers: the state is untouched;
    # results do not depend on the tiling), then the warm-up steps and the
    # timed steps back to back, with no host work in between that would let
    # the GPU idle
    cands = []
    if tiling is not None:
        for q in range(3):
            sim.ctx.set_sweep_tiling(q, *tiling[q])
        sim.ctx.set_sweep_split(bool(tiling[3]), tiling[2][0])
    elif autotune:
        cands = sim.ctx.autotune(sim.p, reps=5)  # the runners-up are within 1-2%
    split = sim.ctx.sweep_split()
    tiling = sim.ctx.tiling()
    spec = tiling_spec(sim)
    # the warm-up steps give every kernel's duration (HIP events around each
    # launch); the timed steps record events only around the launches of the
    # kernel that takes the most time -- each record costs the stream a few
    # microseconds, ~4% of an FP64 step and ~7% of an FP32 step when every
    # launch is recorded
    # (the first warm-up step is not timed: its launches carry one-time costs,
    # the first launch of each kernel several times its usual duration)
    warm = {}
    if warmup:
        sim.run(1)
        if warmup > 1:
            sim.ctx.timing(timing)
            sim.run(warmup - 1)
            if timing:
                warm = sim.ctx.timing_read()
                sim.ctx.timing(False)
    _sync(sim)
    dom = None
    steps_kinds = [k for k in ("sweep_A", "sweep_B", "solute") if warm.get(k, (0, 0))[1] > 0]
    if steps_kinds:
        dom = max(steps_kinds, key=lambda k: warm[k][0])
        sim.ctx.timing_kinds([dom])
    sim.ctx.timing(timing)
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        chunk = min(DIAG_EVERY - (sim.p.it % DIAG_EVERY), steps - done)
        sim.run(chunk)
        done += chunk
        if sim.p.it % DIAG_EVERY == 0:
            sim.diagnostics()  # output-event reduction (synchronous)
    _sync(sim)
    wall = time.perf_counter() - t0
    kt = sim.ctx.timing_read()
    sim.ctx.timing(False)
    sim.ctx.timing_kinds(None)
    diag = sim.diagnostics()
    nbytes = sim.ctx.device_bytes()
    sim.close()
    warm_avg = {k: round(v[0] / v[1], 4) for k, v in warm.items() if v[1] > 0}
    return {"wall_s": wall, "kernels": kt, "diag": diag.tolist(), "init_s": t_init,
            "device_bytes": nbytes, "tiling": tiling, "split": split, "spec": spec,
            "autotune": cands, "dominant": dom, "warmup_avg_ms": warm_avg}


PMC_SUMMARY = os.path.join(ROOT, "profiles", "fused_sweeps_256x256x528_pmc.md")


def traffic_key(res, kernel, prec):
    """The summary's key of a timed kernel: its kind (the split form's
    phase + T kernel is `sweep_PT`), precision and tile shape."""
    from pfamd.capi import TILE_SHAPES

    ta, tb = res["tiling"][0], res["tiling"][1]
    split, ts = res["split"][0], res["split"][1]
    kind, tile = {"sweep_A": ("sweep_A", ta), "solute": ("solute", ts),
                  "sweep_B": ("sweep_PT" if split else "sweep_B", tb)}[kernel]
    return f"{kind}_{prec}_" + "x".join(map(str, TILE_SHAPES[tile]))


def pmc_traffic(kernel):
    """HBM bytes per launch of a kernel (a traffic_key) from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated on a copy kernel), or
    None when the summary does not cover that kernel and tile shape."""
    import re

    try:
        txt = open(PMC_SUMMARY).read()
    except OSError:
        return None
    if kernel is None:
        return None
    m = re.search(rf"^traffic_bytes_per_launch: {kernel} (\d+)", txt, re.M)
    return int(m.group(1)) if m else None


def roofline(res, precision_bytes):
    ncell = NX * NY * NZ
    best = None
    fields = FIELDS_PER_CELL_SPLIT if res["split"][0] else FIELDS_PER_CELL
    for nm, nf in fields.items():
        ms, cnt = res["kernels"].get(nm, (0.0, 0))
        if cnt == 0:
            continue
        avg_s = ms / cnt / 1e3
        alg = nf * precision_bytes * ncell
        entry = {"kernel": nm, "avg_ms": avg_s * 1e3, "alg_bytes": alg,
                 "achieved_gbs": alg / avg_s / 1e9, "share": ms}
        if best is None or ms > best["share"]:
            best = entry
    return best


def precision_decks():
    """The precision study's configurations on the 64x64x128 grid, one seed
    of radius 5 at the bottom centre, noise off: the cooling deck (T0 = 922 K
    is above the liquidus, so the seed melts back) and an active-growth
    variant held at TP = 905 K."""
    from pfamd.deck import Deck

    base = dict(Nx=64, Ny=64, Nz=128, tipo_inicial=5, R_semente=5.0, h_quintico=0, ruido_tipo=0,
                AAA=0.0, n_varreduras=2, stoptime=0.0)
    return {"cooling 19.84 K/s (melt-back)": Deck(taxa_resfr=19.84, **base),
            "held at TP=905 K (growth)": Deck(taxa_resfr=0.0, TP=905.0, **base)}


def perturbed_fields(deck, rel=2.0 ** -25, seed=5):
    """The deck's initial fields with every interior c multiplied by
    1 + u, u uniform in [-rel, rel] (half an FP32 ulp: the size of the
    rounding FP32 storage applies once), then rescaled so that the interior
    FP64 mass equals the unperturbed one."""
    import numpy as np

    from pfamd.run import initial_fields

    d = deck.resolved()
    f = initial_fields(deck)
    c = f[2].reshape(d.Nz, d.Ny, d.Nx)
    inner = (slice(1, -1),) * 3
    m0 = float(np.sum(c[inner], dtype=np.float64))
    u = np.random.default_rng(seed).uniform(-rel, rel, size=c[inner].shape)
    c[inner] = c[inner] * (1.0 + u)
    c[inner] = c[inner] * (m0 / float(np.sum(c[inner], dtype=np.float64)))
    return f


def precision_study(checkpoints=(5000, 10000, 20000)):
    """FP32-mixed against FP64: per-field max-abs, RMS and relative L2
    differences and the interior solute mass at each checkpoint (both runs
    on the GPU, in lockstep).  Beside them, a second FP64 run started from
    the same fields with c perturbed by half an FP32 ulp per cell (same
    interior mass): how far the FP64 scheme's own mass trajectory moves
    under a perturbation of the size FP32 storage applies once."""
    import numpy as np

    from pfamd import capi
    from pfamd.run import Simulation

    out = {}
    for name, deck in precision_decks().items():
        d = deck.resolved()
        shape = (d.Nz, d.Ny, d.Nx)
        sims = {"fp64": Simulation(deck), "fp32": Simulation(deck, precision=capi.PF_FP32_MIXED),
                "fp64_perturbed": Simulation(deck, fields=perturbed_fields(deck))}

        def interior_mass(c):
            return float(np.sum(c.reshape(shape)[1:-1, 1:-1, 1:-1], dtype=np.float64))

        m0 = interior_mass(sims["fp64"].fields()["c_new"])
        rows, done = [], 0
        for cp in checkpoints:
            for sim in sims.values():
                sim.run(cp - done)
            done = cp
            f64, f32 = sims["fp64"].fields(), sims["fp32"].fields()
            row = {"step": cp}
            for nm, key in (("phi", "phi_new"), ("c", "c_new"), ("T", "T_new")):
                a, b = f32[key], f64[key]
                nb = float(np.linalg.norm(b))
                row[f"{nm}_maxabs"] = float(np.max(np.abs(a - b)))
                row[f"{nm}_rms"] = float(np.sqrt(np.mean((a - b) ** 2)))
                row[f"{nm}_rel_l2"] = float(np.linalg.norm(a - b) / nb) if nb > 1e-6 else None
            M64, M32 = interior_mass(f64["c_new"]), interior_mass(f32["c_new"])
            Mp = interior_mass(sims["fp64_perturbed"].ctx.download()[1])
            row["mass_fp64_perturbed_vs_fp64"] = (Mp - M64) / M64
            led = sims["fp32"].ctx.mass_ledger()
            row["phi_max_fp64"] = float(np.max(f64["phi_new"]))
            row["mass_fp64_change"] = (M64 - m0) / m0
            row["mass_fp32_vs_fp64"] = (M32 - M64) / M64
            row["mass_fp32_vs_ledger"] = (M32 - led[0]) / led[0]
            rows.append(row)
            log(f"[bench] precision study {name} step {cp}: {row}")
        for sim in sims.values():
            sim.close()
        out[name] = rows
    return {"config": "64x64x128, seed r=5 at the bottom centre, noise off, 2 sweeps/step; "
                      "FP32-mixed (mass correction every 100 steps) vs FP64 on the GPU",
            "checkpoints": out}


def batched_study(steps=5000, graph_steps=16):
    """The 64-instance ensemble batched vs one after another (the same
    problems, single-instance fused loop).  graph_steps 0 launches the
    batched steps one by one instead of replaying a captured graph (for a
    kernel trace: under rocprofv3 the batched form's graph capture ended in
    a segmentation fault inside the HIP runtime)."""
    from pfamd import ensemble

    probs = ensemble.ensemble_problems(64, 7, steps)
    rb = ensemble.run_batched(probs, graph_steps=graph_steps)
    log(f"[bench] batched ensemble: {rb.wall_s:.2f} s, {rb.mlups:.0f} MLUPS")
    rs = ensemble.run_sequential(probs)
    log(f"[bench] sequential ensemble: {rs.wall_s:.2f} s, {rs.mlups:.0f} MLUPS")
    return {"config": "64 instances, edges log-uniform 32-160 (z >= x,y), cooling U[5,50] K/s, "
                      "1-8 bottom seeds r=4, effective noise U[0,0.02], FP32-mixed, "
                      f"{steps} steps each, seed 7; one hipGraph-replayed launch per sweep",
            "cells_total": sum(p.cells for p in probs), "graph_steps": graph_steps,
            "solves_per_s": round(len(probs) / rb.wall_s, 3),
            "sequential_solves_per_s": round(len(probs) / rs.wall_s, 3),
            "speedup_vs_sequential": round(rs.wall_s / rb.wall_s, 3),
            "mlups_batched": round(rb.mlups, 1), "mlups_sequential": round(rs.mlups, 1),
            "wall_s": round(rb.wall_s, 3), "sequential_wall_s": round(rs.wall_s, 3)}


def serving_study(n=24, rate=4.0, peak_mlups=None, graph_steps=16):
    """Poisson request stream served with continuous batching and served
    one request at a time (FIFO), same stream.  Beside the ratio, the bound
    that the stream itself sets: its cell-steps run at the headline grid's
    rates (peak_mlups, per precision, measured in the same bench run) give
    the shortest makespan any server could reach, hence the highest solves/s
    and ratio to the sequential server."""
    from pfamd import capi, serve

    reqs = serve.request_stream(n=n, rate=rate, seed=11)
    rb = serve.serve_batched(reqs, graph_steps=graph_steps)
    sb = {k: round(v, 3) for k, v in rb.summary().items()}
    log(f"[bench] serving (batched): {sb}")
    rs = serve.serve_sequential(reqs)
    ss = {k: round(v, 3) for k, v in rs.summary().items()}
    log(f"[bench] serving (sequential): {ss}")
    out = {"config": f"{n} requests, Poisson {rate}/s (seed 11), edges log-uniform 32-256 "
                     "(z >= x,y), cooling U[5,50] K/s, 1-16 seeds, 1k-20k steps, FP64 or "
                     "FP32-mixed 50/50; admission by free HBM, 200-step ticks ending at the "
                     "first completion, bounded overtaking",
           "cell_steps_total": sum(r.problem.cells * r.problem.steps for r in reqs),
           "batched": sb, "sequential": ss,
           "solves_per_s_ratio": round(sb["solves_per_s"] / ss["solves_per_s"], 3),
           "gpu_busy_frac_batched": round(rb.gpu_wait_s / rb.makespan, 3)}
    if peak_mlups:
        t = sum(r.problem.cells * r.problem.steps / (peak_mlups[r.problem.precision] * 1e6)
                for r in reqs)
        out["bound"] = {"note": "every cell-step at the headline grid's rate of its precision",
                        "peak_mlups": {"fp64": round(peak_mlups[capi.PF_FP64], 1),
                                       "fp32_mixed": round(peak_mlups[capi.PF_FP32_MIXED], 1)},
                        "makespan_s": round(t, 3), "solves_per_s": round(n / t, 3),
                        "ratio_vs_sequential": round(n / t / ss["solves_per_s"], 3)}
    return out


def karma_study(n=200, warm=20, steps=500, cpu_steps=40, cpu=True):
    """The quantitative Karma model (karma/karma3d.f90) on the karma.nml the
    reference ships (200^3, lambda 4, eps4 0.02, dx 0.8, dt 0.008, R0 8,
    Fnoise 0.01): GPU steps timed after a warm-up, and the reference program
    compiled from its source (oracle/_ref) on the same 200^3 grid for a
    bounded number of steps on this host's physical cores, as it times
    itself (its loop wall clock, printed before its output file, whose
    writing is not waited for: the process is ended once the line is out).

    Its step is FP64-VALU bound (correctly rounded divisions and square
    roots per face, every expression in the reference's order): the roofline
    reported is the FP64 issue rate, from the instruction count per cell of
    the two kernels."""
    import re
    import subprocess
    import tempfile
    import time

    import numpy as np

    from pfamd.karma import Karma, KarmaDeck, tips

    deck = KarmaDeck.shipped()
    deck.N = n
    m = Karma(deck)
    try:
        m.run(warm)
        m.synchronize()
        t0 = time.perf_counter()
        m.run(steps)
        m.synchronize()
        wall = time.perf_counter() - t0
        phi, _, it = m.fields()
    finally:
        m.close()
    mlups = n ** 3 * steps / wall / 1e6
    # FP64 VALU instructions per cell-step of the two z-march kernels (their
    # loops' static counts; rcp/rsq are quarter-rate: 3 extra issue slots)
    valu = {"phase": 481 + 3 * 21, "solute": 362 + 3 * (7 + 6)}
    per_cell = sum(valu.values())
    valu_peak = 256 * 4 * 16 * 2.4e9  # lane-instructions/s (FP64 at 16 lanes/SIMD/clk)
    out = {"config": f"{n}^3, karma.nml as shipped (lambda 4, eps4 0.02, dx 0.8, dt 0.008, R0 8, "
                     f"Fnoise 0.01), {warm} warm-up + {steps} timed steps, FP64, z-march kernels",
           "mlups": round(mlups, 1), "ms_per_step": round(wall / steps * 1e3, 4),
           "roofline": {"bound": "valu_fp64",
                        "fp64_issue_slots_per_cell_step": per_cell,
                        "achieved_tinstr_s": round(mlups * 1e6 * per_cell / 1e12, 2),
                        "peak_tinstr_s": round(valu_peak / 1e12, 2),
                        "frac": round(mlups * 1e6 * per_cell / valu_peak, 3),
                        "hbm_alg_bytes_per_cell_step": 64,
                        "hbm_frac": round(mlups * 1e6 * 64 / 8.0e12, 4)},
           "tip_cells": dict(zip(("x_axis", "diagonal"), tips(phi))),
           "reference_published_mlups": 200.0,
           "reference_published_note": "200^3 for 45k steps in ~30 min on the reference's GPU"}
    log(f"[bench] karma {n}^3: {out['mlups']} MLUPS, {out['ms_per_step']} ms/step")
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_karma3d")
    if cpu and os.path.exists(exe):
        threads = host_cpu()["physical_cores"]
        with tempfile.TemporaryDirectory() as d:
            # one progress line, at the last step
            cpu_deck = KarmaDeck.shipped()
            cpu_deck.N, cpu_deck.nsteps, cpu_deck.nout, cpu_deck.tip_stop = (n, cpu_steps,
                                                                             cpu_steps, 2.0)
            with open(os.path.join(d, "karma.nml"), "w") as f:
                f.write(cpu_deck.to_nml())
            os.makedirs(os.path.join(d, "output"))
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                       OMP_PLACES="cores", OMP_TARGET_OFFLOAD="DISABLED")
            pr = subprocess.Popen([exe], cwd=d, env=env, stdout=subprocess.PIPE, text=True)
            wline = None
            t_end = time.time() + 300
            for line in pr.stdout:
                if "wall=" in line:
                    wline = line
                    break
                if time.time() > t_end:
                    break
            pr.kill()
            pr.wait()
        w = re.findall(r"wall=\s*([0-9.]+)\s*s", wline or "")
        if w and float(w[-1]) > 0:
            cpu = n ** 3 * cpu_steps / float(w[-1]) / 1e6
            out["cpu_reference"] = {"value": round(cpu, 2), "unit": "MLUPS", "cores": threads,
                                    "kind": "reference",
                                    "sample": f"{n}^3 (the same grid and namelist), {cpu_steps} "
                                              f"steps, its own loop wall clock {w[-1]} s "
                                              "(0.1 s resolution)"}
            out["gpu_vs_cpu"] = round(mlups / cpu, 1)
            log(f"[bench] karma CPU reference: {cpu:.1f} MLUPS on {threads} threads")
    return out


def host_cpu():
    """The host's CPU as the baseline runs on it: model, the logical CPUs this
    process may use, the physical cores among them (SMT siblings counted
    once) and the NUMA nodes (from /proc/cpuinfo and sysfs)."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else \
        list(range(os.cpu_count() or 1))
    model, family, mnum = "unknown", None, None
    try:
        for ln in open("/proc/cpuinfo"):
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "cpu family" and family is None:
                family = v
            elif k == "model" and mnum is None:
                mnum = v
    except OSError:
        pass
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology"
        try:
            pkg = open(f"{base}/physical_package_id").read().strip()
            core = open(f"{base}/core_id").read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    try:
        nodes = sorted(d for d in os.listdir("/sys/devices/system/node") if d.startswith("node"))
    except OSError:
        nodes = []
    return {"model": model, "cpu_family": family, "cpu_model": mnum, "logical_cpus": len(cpus),
            "physical_cores": len(cores), "numa_nodes": max(1, len(nodes))}


def cpu_baseline():
    """The reference's own OpenMP CPU path (oracle/_ref/ref_harness: its
    kernels compiled from source with flang -O2 -fopenmp), timed on every
    physical core this process may use (OMP_PLACES=cores,
    OMP_PROC_BIND=close, as the reference's set_OpenMP_environment.bat) on
    20 steps of the same grid after 2 warm-up steps (about 10 s of CPU work
    on 16 cores).  Falls back to the C oracle (a faithful restatement) when
    the reference build is absent."""
    cpu = host_cpu()
    threads = cpu["physical_cores"]
    nwarm, nsteps = 2, 20
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    if os.path.exists(exe):
        wd = tempfile.mkdtemp(prefix="pfcpu_")
        try:
            d = workload_deck()
            d.tipo_inicial = 3  # the reference has no seeded mode; cost is data-independent
            d.write(os.path.join(wd, "entrada.dat"))
            with open(os.path.join(wd, "harness.nml"), "w") as fh:
                fh.write(f"&harness\n mode='time', nsteps={nsteps}, nwarm={nwarm}\n/\n")
            subprocess.run([exe], cwd=wd, env=env, stdout=subprocess.DEVNULL, check=True,
                           timeout=600)
            t = float(open(os.path.join(wd, "time.txt")).read().split()[0])
        finally:
            shutil.rmtree(wd, ignore_errors=True)
        kind = "reference"
    else:
        import numpy as np

        import oracle as O
        from pfamd.run import setup_from_deck

        d = workload_deck()
        s = setup_from_deck(d)
        O.set_threads(threads)
        F = O.initial_fields(NX, NY, NZ, 3, 0.0, 922.0, d.C0, d.Ke)
        O.run_steps(s.p, s.L, nwarm, F)
        t0 = time.perf_counter()
        O.run_steps(s.p, s.L, nsteps, F)
        t = time.perf_counter() - t0
        kind = "port"
        _ = np
    mlups = NX * NY * NZ * nsteps / t / 1e6
    return {"value": mlups, "unit": "MLUPS", "cores": threads, "kind": kind,
            "sample": f"{nsteps} steps of the {NX}x{NY}x{NZ} FP64 step after {nwarm} warm-up, "
                      f"OMP_NUM_THREADS={threads} (every physical core), OMP_PLACES=cores, "
                      f"OMP_PROC_BIND=close, {t:.2f} s",
            "host": cpu}


def log(msg):
    """Progress lines go to stderr; stdout carries only the JSON result line."""
    print(msg, file=sys.stderr, flush=True)


def tiling_summary(res):
    from pfamd.capi import TILE_SHAPES

    out = {}
    for nm, t, seg in (("sweep_A", res["tiling"][0], res["tiling"][2]),
                       ("sweep_B", res["tiling"][1], res["tiling"][3])):
        out[nm] = {"tile": "x".join(map(str, TILE_SHAPES[t])), "z_planes_per_group": seg}
    out["spec"] = res["spec"]
    on, ts, segs = res["split"]
    out["sweep_B"]["form"] = "split: phase + T kernel, then the solute kernel" if on else "fused"
    if on:
        out["solute"] = {"tile": "x".join(map(str, TILE_SHAPES[ts])), "z_planes_per_group": segs}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-precision-study", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-karma", action="store_true")
    ap.add_argument("--graph-steps", type=int, default=16,
                    help="steps per captured graph of the batched and serving legs (0: no graph)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tiling")
    ap.add_argument("--tiling", type=parse_tiling, default=None,
                    help="FP64 tiling tA:gA,tB:gB,tS:gS,split instead of the autotuner "
                         "(for profiling runs)")
    ap.add_argument("--tiling32", type=parse_tiling, default=None,
                    help="the same for the FP32-mixed form")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (instances of this "
              "solver shard across GPUs with no communication, see DESIGN.md)", file=sys.stderr)
        return 2
    from pfamd import capi

    if capi.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    _torch_init()
    log(f"[bench] FP64 {NX}x{NY}x{NZ}: {a.warmup} warm-up + {a.steps} timed steps")
    r64 = run_form(capi.PF_FP64, a.steps, a.warmup, autotune=not a.no_autotune, tiling=a.tiling)
    ncell = NX * NY * NZ
    mlups = ncell * a.steps / r64["wall_s"] / 1e6
    log(f"[bench] FP64 autotune {r64['autotune']}")
    log(f"[bench] FP64 {mlups:.1f} MLUPS, tiling {r64['spec']}, kernels {r64['kernels']}")
    rl = roofline(r64, 8)
    out = {
        "metric": "MLUPS", "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": r64["wall_s"] / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "columnar growth 256x256x528, 16 seeds r=4, noise 0.01, "
                               "2 sweeps/step, diagnostics every 1000 steps",
                   "grid": [NX, NY, NZ], "cells": ncell, "form": "FP64 drop-in (fused sweeps)"},
        "roofline": {"bound": "hbm", "achieved": round(rl["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(rl["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(traffic_key(r64, rl["kernel"], "fp64")),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x "
                                     "1.0, factors from a copy of known size)",
                     "traffic_kernel": traffic_key(r64, rl["kernel"], "fp64"),
                     "limiter": "VALU issue and latency: 208-226 (sweep A) and 219-239 "
                                "(phase + T) VALU (waves without / with the noise hash) and "
                                "63-66 scalar instructions per cell, VALU issuing 56-61% of "
                                "the time, 3 waves/SIMD at 167-168 VGPRs (FP64 "
                                "register demand; a fourth wave spills ~40 VGPRs)",
                     "kernel": rl["kernel"], "avg_ms": round(rl["avg_ms"], 4),
                     "alg_bytes_per_launch": rl["alg_bytes"]},
        "step_bytes_per_cell_alg": 64,
        "whole_step_roofline": {"bytes_per_cell_step": 72,
                                "note": "the survey's two-sweep minimum with T read (72 B FP64); "
                                        "with D = 0 this schedule reads no T (64 B)",
                                "achieved": round(mlups * 1e6 * 72 / 1e9, 1),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(mlups * 1e6 * 72 / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels_ms_total": {k: round(v[0], 3) for k, v in r64["kernels"].items() if v[1]},
        "kernels_launches": {k: v[1] for k, v in r64["kernels"].items() if v[1]},
        "kernels_timed_note": "HIP events around the dominant kernel's launches only in the "
                              "timed steps; every kernel's average from the warm-up steps",
        "kernels_avg_ms_warmup": r64["warmup_avg_ms"],
        "whole_step_alg_gbs": round(mlups * 1e6 * 64 / 1e9, 1),
        "diag_final": r64["diag"],
        "tiling": tiling_summary(r64),
    }
    gbs_copy = capi.stream_copy_gbs(2 << 30, 20)
    out["roofline"]["measured_copy_gbs"] = round(gbs_copy, 1)
    out["roofline"]["frac_of_measured_copy"] = round(rl["achieved_gbs"] / gbs_copy, 4)
    log(f"[bench] STREAM copy {gbs_copy:.0f} GB/s")
    if not a.no_fp32:
        log("[bench] FP32-mixed")
        r32 = run_form(capi.PF_FP32_MIXED, a.steps, a.warmup, autotune=not a.no_autotune,
                       tiling=a.tiling32)
        m32 = ncell * a.steps / r32["wall_s"] / 1e6
        log(f"[bench] FP32 autotune {r32['autotune']}")
        log(f"[bench] FP32 {m32:.1f} MLUPS, tiling {r32['spec']}, kernels {r32['kernels']}")
        rl32 = roofline(r32, 4)
        out["fp32_mixed"] = {"value": round(m32, 2), "unit": "MLUPS",
                             "ms_per_step": r32["wall_s"] / a.steps * 1e3,
                             "speedup_vs_fp64": round(m32 / mlups, 3),
                             "roofline_kernel": rl32["kernel"],
                             "roofline_achieved_gbs": round(rl32["achieved_gbs"], 1),
                             "roofline_avg_ms": round(rl32["avg_ms"], 4),
                             "roofline_traffic": pmc_traffic(traffic_key(r32, rl32["kernel"],
                                                                         "fp32")),
                             "kernels_avg_ms_warmup": r32["warmup_avg_ms"],
                             "tiling": tiling_summary(r32),
                             "diag_final": r32["diag"]}
    if not a.no_fp32 and not a.no_precision_study:
        log("[bench] precision study (FP32-mixed vs FP64, 64x64x128)")
        out["fp32_mixed"]["precision_study"] = precision_study()
    if not a.no_batched:
        log("[bench] batched ensemble (64 instances)")
        out["batched"] = batched_study(graph_steps=a.graph_steps)
    if not a.no_serving:
        log("[bench] serving (Poisson stream)")
        peaks = {capi.PF_FP64: mlups}
        if "fp32_mixed" in out:
            peaks[capi.PF_FP32_MIXED] = out["fp32_mixed"]["value"]
        out["serving"] = serving_study(peak_mlups=peaks if len(peaks) == 2 else None,
                                       graph_steps=a.graph_steps)
    if not a.no_karma:
        log("[bench] Karma model")
        out["karma"] = karma_study(cpu=not a.no_cpu_baseline)
    if not a.no_cpu_baseline:
        log("[bench] CPU baseline (reference OpenMP path)")
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
        out["gpu_vs_cpu"] = round(mlups / cb["value"], 1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
