"""GPU tests of the batched form (pf_batch_*): every instance of a batch
gives the same bits as the single-instance fused loop (pf_run_steps) on the
same problem, alone or in any mix, whatever the tiling and graph length."""
import numpy as np
import pytest

from conftest import gpu_available
from pfamd import capi
from pfamd.deck import Deck
from pfamd.run import Simulation, initial_fields, setup_from_deck

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a gfx950 GPU")]


def _deck(n, **kw):
    base = dict(Nx=n[0], Ny=n[1], Nz=n[2], tipo_inicial=5, R_semente=5.0, h_quintico=0,
                ruido_tipo=1, AAA=0.0122474, taxa_resfr=0.0, TP=905.0, n_varreduras=2,
                stoptime=0.0)
    base.update(kw)
    return Deck(**base)


PROBLEMS = {
    "fp64_seed_noise": (_deck((40, 36, 50)), capi.PF_FP64, 3),
    "fp32_planar_cooling": (_deck((34, 30, 44), tipo_inicial=3, taxa_resfr=19.84, ruido_tipo=0,
                                  AAA=0.03), capi.PF_FP32_MIXED, 0),
    "fp64_quintic_smooth": (_deck((48, 40, 64), h_quintico=1, ruido_tipo=2), capi.PF_FP64, 0),
    "fp32_seed_small": (_deck((24, 26, 30), R_semente=4.0), capi.PF_FP32_MIXED, 5),
}


def _add(batch, name, nsteps):
    deck, prec, seed = PROBLEMS[name]
    s = setup_from_deck(deck, prec, seed)
    f = initial_fields(deck)
    return batch.add(s.p, s.L, nsteps, f[0], f[2], f[4])


def _single(name, nsteps):
    deck, prec, seed = PROBLEMS[name]
    sim = Simulation(deck, precision=prec, seed=seed)
    sim.run(nsteps)
    f = sim.fields()
    out = (f["phi_new"], f["c_new"], f["T_new"], sim.p.it, sim.p.TP, sim.p.M0, float(sim.L.time),
           sim.diagnostics())
    sim.close()
    return out


def _from_batch(batch, slot, name):
    deck = PROBLEMS[name][0]
    n = deck.Nx * deck.Ny * deck.Nz
    phi, c, T = batch.download(slot, n)
    p, L = batch.scalars(slot)
    return phi, c, T, p.it, p.TP, p.M0, float(L.time), batch.diagnosticos(slot)


def _same(a, b, label):
    for q, nm in enumerate(("phi", "c", "T")):
        assert np.array_equal(a[q], b[q]), f"{label}: {nm} differs (max {np.max(np.abs(a[q] - b[q]))})"
    assert a[3:7] == b[3:7], f"{label}: scalars {a[3:7]} vs {b[3:7]}"
    assert np.array_equal(a[7], b[7]), f"{label}: diagnostics"


def test_batch_equals_single_instance():
    """Four uneven instances, FP64 and FP32, 30 steps as 3 graph replays of 8
    plus 6 plain steps: each equals pf_run_steps on its own, bit for bit."""
    print("[batch] 4 instances vs single-instance runs", flush=True)
    with capi.Batch() as b:
        b.set_graph_steps(8)
        slots = {nm: _add(b, nm, 30) for nm in PROBLEMS}
        b.run(30)
        for nm, sl in slots.items():
            assert b.status(sl) == (30, 0)
            _same(_from_batch(b, sl, nm), _single(nm, 30), nm)


def test_batch_alone_equals_in_mix():
    """An instance run alone and run in a changing mix (others admitted,
    finished and removed while it runs, tiling autotuned in between) gives
    the same bits; an instance stops at its own step count."""
    print("[batch] alone vs mix", flush=True)
    with capi.Batch() as alone:
        s = _add(alone, "fp64_seed_noise", 40)
        alone.run(40)
        ref = _from_batch(alone, s, "fp64_seed_noise")
    with capi.Batch() as mix:
        mix.set_graph_steps(5)
        x = _add(mix, "fp64_seed_noise", 40)
        short = _add(mix, "fp32_seed_small", 12)
        mix.run(15)  # `short` stops at its own 12 steps
        assert mix.status(short) == (12, 0)
        _same(_from_batch(mix, short, "fp32_seed_small"), _single("fp32_seed_small", 12), "short")
        mix.remove(short)
        y = _add(mix, "fp64_quintic_smooth", 25)
        mix.autotune(reps=1)
        mix.run(25)
        assert mix.status(x) == (40, 0)
        _same(_from_batch(mix, x, "fp64_seed_noise"), ref, "mix")
        _same(_from_batch(mix, y, "fp64_quintic_smooth"), _single("fp64_quintic_smooth", 25), "y")


def test_batch_tiling_and_budget():
    print("[batch] tiling and budget", flush=True)
    deck, prec, seed = PROBLEMS["fp64_seed_noise"]
    n = deck.Nx * deck.Ny * deck.Nz
    with capi.Batch(hbm_budget=2 * 48 * n + 100) as b:
        cls = b.size_class(prec, deck.Nx, deck.Ny)
        outs = []
        for shape, target in ((0, 64), (3, 4096), (capi.batch_tiles(1, 2, 3), 512)):
            b.set_tiling(cls, shape)
            b.set_group_target(target)
            assert b.tiling(cls) == shape
            if shape >= 0x1000:
                assert b.tiling_kinds(cls) == (1, 2, 3)
            sl = _add(b, "fp64_seed_noise", 10)
            b.run(10)
            outs.append(_from_batch(b, sl, "fp64_seed_noise"))
            b.remove(sl)
        _same(outs[0], outs[1], "tilings")
        _same(outs[0], outs[2], "tilings per kernel kind")
        _add(b, "fp64_seed_noise", 1)
        _add(b, "fp64_seed_noise", 1)
        with pytest.raises(capi.PfError):
            _add(b, "fp64_seed_noise", 1)  # over the budget
        used, budget = b.bytes()
        assert used == 2 * 48 * n and budget == 2 * 48 * n + 100


def test_batch_fp32_mass_correction_equals_single():
    """FP32 instances keep the solute-mass ledger in the batch as
    pf_run_steps does (a correction every 100 steps by default): 230 steps
    across two corrections, in a mix with an FP64 instance and with a graph
    length that does not divide the period, equal the single-instance run
    split over calls (100 + 30 + 100), bit for bit, ledger included.  With
    the correction off the two forms agree as well, and differ from the
    corrected run."""
    print("[batch] FP32 solute-mass correction", flush=True)
    names = ("fp32_planar_cooling", "fp32_seed_small")
    with capi.Batch() as b:
        b.set_graph_steps(7)
        slots = {nm: _add(b, nm, 230) for nm in names}
        mix = _add(b, "fp64_seed_noise", 150)
        off = _add(b, "fp32_seed_small", 230)
        b.set_mass_correction(off, 0)
        b.run(230)
        got = {nm: (_from_batch(b, sl, nm), b.mass_ledger(sl)) for nm, sl in slots.items()}
        got_off = _from_batch(b, off, "fp32_seed_small")
        assert b.mass_ledger(off)[1:] == (0.0, 0.0)  # never corrected
        assert b.status(mix) == (150, 0)
        with pytest.raises(capi.PfError):
            b.set_mass_correction(slots[names[0]], 50)  # only before the first step
    for nm in names:
        deck, prec, seed = PROBLEMS[nm]
        sim = Simulation(deck, precision=prec, seed=seed)
        for n in (100, 30, 100):
            sim.run(n)
        f = sim.fields()
        ref = (f["phi_new"], f["c_new"], f["T_new"], sim.p.it, sim.p.TP, sim.p.M0,
               float(sim.L.time), sim.diagnostics())
        led = sim.ctx.mass_ledger()
        sim.close()
        _same(got[nm][0], ref, nm)
        assert got[nm][1] == led, (nm, got[nm][1], led)
        assert led[2] != 0.0  # corrected at least once
    deck, prec, seed = PROBLEMS["fp32_seed_small"]
    sim = Simulation(deck, precision=prec, seed=seed)
    sim.ctx.set_mass_correction(0)
    sim.run(230)
    f = sim.fields()
    ref_off = (f["phi_new"], f["c_new"], f["T_new"], sim.p.it, sim.p.TP, sim.p.M0,
               float(sim.L.time), sim.diagnostics())
    sim.close()
    _same(got_off, ref_off, "correction off")
    assert not np.array_equal(got_off[1], got["fp32_seed_small"][0][1])


def test_batch_split_and_fused_sweep_b_agree():
    """Sweep B of the batched steps fused or split into the phase + T and the
    solute kernels: the same bits, FP64 and FP32 (with a mass correction)."""
    print("[batch] split vs fused sweep B", flush=True)
    outs = []
    for split in (True, False):
        with capi.Batch() as b:
            b.set_sweep_split(split)
            b.set_graph_steps(6)
            slots = {nm: _add(b, nm, 105) for nm in ("fp64_seed_noise", "fp32_seed_small")}
            b.run(105)
            outs.append({nm: (_from_batch(b, sl, nm), b.mass_ledger(sl)) for nm, sl in slots.items()})
    for nm in outs[0]:
        _same(outs[0][nm][0], outs[1][nm][0], nm)
        assert outs[0][nm][1] == outs[1][nm][1], nm
