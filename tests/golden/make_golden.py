"""Writes the golden fixtures in this folder from the reference's own CPU path.

The reference (ileaof/phase-field-project-3D) is compiled from its sources
where they lie by oracle/ref/Makefile (AMD flang, -O2 -fopenmp) into
oracle/_ref/ref_harness -- its kernels plus a harness of ours in place of
principal.f90 -- and oracle/_ref/ref_principal (its unmodified program).  This
script drives both on small decks and stores inputs and outputs as .npz:

  random_kat.npz     pseudo_random / ruido_suave / REAL(4) ran for tuples
  scalars_kat.npz    dt, AAA rescale, W0/E0 search, M0(TP), Cs_eq/Cl_eq
  unit_*.npz         one call of temperatura / equacao_campo_fase /
                     concentracao_soluto / diagnosticos on random fields
  run_*.npz          the time loop (principal.f90:93-226) for a few steps
  principal_*.npz    the unmodified principal.f90's Tecplot output events,
                     console transcript and performance_log.txt
                     (principal_restart: continued from a restart.dat;
                     principal_stop: ended by stoptime)

Run here (needs /root/reference and flang):  python tests/golden/make_golden.py
The fixtures are data only; no reference source is stored.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "phase-field-project-3d_amd"))
from pfamd.deck import Deck  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")
HARNESS = os.path.join(REF_DIR, "ref_harness")
PRINCIPAL = os.path.join(REF_DIR, "ref_principal")
NSCAL = 16
SCALARS = ("it", "time", "TP", "M0", "E0", "W0", "Cs_eq", "Cl_eq", "AAA", "DELTA_T",
           "L2", "phi_max", "phi_min", "sum_c", "free_energy", "n_varreduras")
FIELDS = ("phi_old", "phi_new", "c_old", "c_new", "T_old", "T_new", "dpp")


def build_ref():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)


def _nml(d: dict) -> str:
    parts = []
    for k, v in d.items():
        if isinstance(v, str):
            parts.append(f"{k}='{v}'")
        elif isinstance(v, (list, tuple)):
            parts.append(f"{k}=" + ",".join(str(x) for x in v))
        else:
            parts.append(f"{k}=" + (repr(v).replace("e", "d") if isinstance(v, float) else str(v)))
    return "&harness\n " + ",\n ".join(parts) + "\n/\n"


def run_harness(deck: Deck, nml: dict, init=None, tuples=None):
    wd = tempfile.mkdtemp(prefix="pfref_")
    try:
        deck.write(os.path.join(wd, "entrada.dat"))
        if init is not None:
            np.concatenate([np.asarray(a, dtype=np.float64).reshape(-1) for a in init]).tofile(
                os.path.join(wd, "init.bin"))
            nml = dict(nml, init_file="init.bin")
        if tuples is not None:
            with open(os.path.join(wd, "tuples.txt"), "w") as fh:
                fh.write(f"{len(tuples)}\n")
                for t in tuples:
                    fh.write(" ".join(str(int(x)) for x in t) + "\n")
        with open(os.path.join(wd, "harness.nml"), "w") as fh:
            fh.write(_nml(nml))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([HARNESS], cwd=wd, env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"ref_harness failed: {r.stderr[-2000:]}")
        out = {"stdout_tail": r.stdout[-4000:]}
        for fn in os.listdir(wd):
            if fn.endswith(".bin") and fn not in ("init.bin",):
                out[fn] = np.fromfile(os.path.join(wd, fn), dtype=np.float64)
        return out
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def fortran_float(t: str) -> float:
    """Parse Fortran E-edited output; an exponent of 3 digits drops the 'E'
    (e.g. '-0.179769+309')."""
    t = t.strip()
    if "E" not in t.upper() and "NAN" not in t.upper() and "INF" not in t.upper():
        for q in range(len(t) - 1, 0, -1):
            if t[q] in "+-":
                t = t[:q] + "E" + t[q:]
                break
    return float(t)


def split_dump(raw: np.ndarray, n: int):
    f = {nm: raw[i * n:(i + 1) * n].copy() for i, nm in enumerate(FIELDS)}
    s = dict(zip(SCALARS, raw[7 * n:7 * n + NSCAL].tolist()))
    return f, s


def tiny_deck(**kw) -> Deck:
    d = Deck(Nx=6, Ny=6, Nz=6, stoptime=0.0, WriteFileD_timeCounter=1000000)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {name}: {os.path.getsize(path)} bytes")


# ---------------------------------------------------------------------------
def make_random_kat():
    rng = np.random.default_rng(1234)
    tuples = []
    for i in (1, 2, 3, 5, 6, 7, 64, 127, 256, 528, 1000):
        for it in (0, 1, 2, 19, 20, 21, 1000, 20000):
            tuples.append((i, (i * 7) % 300 + 1, (i * 13) % 500 + 1, it, 1))
    for _ in range(200):
        tuples.append((int(rng.integers(0, 4000)), int(rng.integers(0, 4000)),
                       int(rng.integers(0, 4000)), int(rng.integers(0, 10**6)),
                       int(rng.integers(1, 4))))
    tuples.append((2**31 - 1, 2**31 - 1, 2**31 - 1, 2**31 - 1, 3))
    out = run_harness(tiny_deck(), {"mode": "random"}, tuples=tuples)
    v = out["random.bin"].reshape(-1, 3)
    save("random_kat.npz", tuples=np.array(tuples, dtype=np.int64), pseudo_random=v[:, 0],
         ruido_suave=v[:, 1], ran_real4=v[:, 2])


def make_scalars_kat():
    rows, meta = [], []
    variants = [
        dict(),
        dict(Thin_Interface=2.0),
        dict(dx=2.5e-7, dy=2.5e-7, dz=2.5e-7),
        dict(ajuste=1.0, Beta=2.5),
        dict(D_liq=2.0e-9, ruido_tipo=1),
        dict(Sigma=float(np.float32(0.12)), ruido_tipo=2, h_quintico=1),
    ]
    for vi, kw in enumerate(variants):
        for TP in (905.0, 910.0, 919.0, 921.96, 922.0):
            d = tiny_deck(**kw)
            out = run_harness(d, {"mode": "unit", "unit_call": "auxiliar", "tp_set": TP})
            n = d.Nx * d.Ny * d.Nz
            _, s = split_dump(out["dump_0.bin"], n)
            rows.append([vi, TP, s["DELTA_T"], s["AAA"], s["E0"], s["W0"], s["M0"], s["Cs_eq"],
                         s["Cl_eq"]])
        meta.append(dict(Deck().__dict__, **{k: v for k, v in kw.items()}))
    save("scalars_kat.npz", rows=np.array(rows),
         columns=np.array(["variant", "TP", "dt", "AAA", "E0", "W0", "M0", "Cs_eq", "Cl_eq"]),
         variants=np.array([json.dumps(dict(tiny_deck(**kw).__dict__)) for kw in variants]))


def random_fields(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    phi = rng.uniform(0, 1, n)
    u = rng.uniform(0, 1, n)
    phi[u < 0.3] = 0.0
    phi[(u >= 0.3) & (u < 0.45)] = 1.0
    phi[(u >= 0.45) & (u < 0.55)] = rng.uniform(0.93, 0.97, int(((u >= 0.45) & (u < 0.55)).sum()))
    c = 0.0196 * rng.uniform(0.3, 1.7, n)
    T = 922.0 + rng.uniform(-2, 2, n)
    dpp = rng.uniform(-0.05, 0.05, n)
    phi_new = rng.uniform(0, 1, n)
    c_new = 0.0196 * rng.uniform(0.3, 1.7, n)
    T_new = 922.0 + rng.uniform(-2, 2, n)
    return [phi, phi_new, c, c_new, T, T_new, dpp]


def make_unit():
    nx, ny, nz = 12, 11, 13
    n = nx * ny * nz
    cases = []
    for hq in (0, 1):
        cases.append(("temperatura", dict(h_quintico=hq), {}))
        cases.append(("soluto", dict(h_quintico=hq), {}))
        for rt in (0, 1, 2):
            cases.append(("fase", dict(h_quintico=hq, ruido_tipo=rt, AAA=0.05),
                          {"it_set": 37.0, "tp_set": 910.0}))
    cases.append(("diag", dict(), {}))
    for ci, (call, kw, extra) in enumerate(cases):
        d = Deck(Nx=nx, Ny=ny, Nz=nz, stoptime=0.0, taxa_resfr=0.0, TP=912.0, **kw)
        init = random_fields(nx, ny, nz, 100 + ci)
        nml = {"mode": "unit", "unit_call": call}
        nml.update(extra)
        out = run_harness(d, nml, init=init)
        f, s = split_dump(out["dump_0.bin"], n)
        arrays = {f"in_{nm}": a for nm, a in zip(FIELDS, init)}
        arrays.update({f"out_{nm}": f[nm] for nm in FIELDS})
        arrays["scalars"] = np.array([s[k] for k in SCALARS])
        arrays["deck"] = np.array(json.dumps(d.__dict__))
        arrays["call"] = np.array(call)
        arrays["nml"] = np.array(json.dumps(nml))
        if call == "fase":
            # probe print, equacao_campo_fase.f90:174 (format 9(1X,e14.4))
            toks = out["stdout_tail"].strip().splitlines()[-1].split()
            arrays["probe_print"] = np.array([float(t) for t in toks])
        save(f"unit_{ci:02d}_{call}.npz", **arrays)


RUNS = [
    ("sphere_cubic_pos", dict(Nx=14, Ny=14, Nz=20, tipo_inicial=1, R_semente=4.0, h_quintico=0,
                              ruido_tipo=0, AAA=0.03, taxa_resfr=19.84, n_varreduras=2), 12, [1, 5, 12]),
    ("corner_quintic_smooth", dict(Nx=12, Ny=12, Nz=16, tipo_inicial=2, R_semente=5.0,
                                   h_quintico=1, ruido_tipo=2, AAA=0.03, taxa_resfr=0.0,
                                   n_varreduras=1, TP=905.0), 10, [10]),
    ("planar_cubic_zeromean", dict(Nx=12, Ny=12, Nz=16, tipo_inicial=3, R_semente=0.0,
                                   h_quintico=0, ruido_tipo=1, AAA=0.03, taxa_resfr=19.84,
                                   n_varreduras=2), 10, [10]),
    ("growth_noiseoff", dict(Nx=16, Ny=16, Nz=24, tipo_inicial=1, R_semente=5.0, h_quintico=0,
                             ruido_tipo=0, AAA=0.0, taxa_resfr=0.0, TP=905.0, n_varreduras=2),
     30, [30]),
]


def make_runs():
    for name, kw, nsteps, dumps in RUNS:
        d = Deck(stoptime=0.0, WriteFileD_timeCounter=1000000, **kw)
        n = d.Nx * d.Ny * d.Nz
        out = run_harness(d, {"mode": "run", "nsteps": nsteps, "dump_steps": dumps})
        arrays = {"deck": np.array(json.dumps(d.__dict__)), "nsteps": np.array(nsteps),
                  "dumps": np.array(dumps)}
        for st in dumps:
            f, s = split_dump(out[f"dump_{st}.bin"], n)
            for nm in ("phi_new", "c_new", "T_new", "dpp"):
                arrays[f"s{st}_{nm}"] = f[nm]
            arrays[f"s{st}_scalars"] = np.array([s[k] for k in SCALARS])
        save(f"run_{name}.npz", **arrays)


PRINCIPAL_DECKS = {
    # a centred sphere in a tight box: the reference itself goes NaN here
    "sphere": dict(Nx=12, Ny=12, Nz=16, tipo_inicial=1, R_semente=4.0, h_quintico=0,
                   ruido_tipo=0, AAA=0.03, taxa_resfr=19.84, n_varreduras=2),
    # the run6 deck's own settings (planar base front, cubic h, positive noise)
    "planar": dict(Nx=14, Ny=12, Nz=18, tipo_inicial=3, R_semente=0.0, h_quintico=0,
                   ruido_tipo=0, AAA=0.03, taxa_resfr=19.84, n_varreduras=2, TP=905.0),
    # the original front (tipo_inicial = 0: phi = 1 for j <= 4), quintic h,
    # smooth zero-mean noise, fixed TP, one sweep
    "front_j": dict(Nx=12, Ny=14, Nz=10, tipo_inicial=0, R_semente=0.0, h_quintico=1,
                    ruido_tipo=2, AAA=0.03, taxa_resfr=0.0, n_varreduras=1, TP=905.0),
}

# E0-EPSLON_Try lines (W0_E0.f90:19, one per try, ~5,500 of them) kept in the
# fixtures: the first and last five and every 100th.
W0E0_TRY = "E0-EPSLON_Try==>"


def _try_sample(n):
    return sorted(set(range(min(5, n))) | set(range(0, n, 100)) | set(range(max(0, n - 5), n)))


def _run_principal(d, files=None):
    """Runs the unmodified principal.f90 on deck d (extra input files in
    `files`: name -> text) and records its output events (it = N+1, 2N+1,
    principal.f90:161-187), diagnostics lines and probe prints."""
    wd = tempfile.mkdtemp(prefix="pfprin_")
    try:
        d.write(os.path.join(wd, "entrada.dat"))
        for fn, text in (files or {}).items():
            with open(os.path.join(wd, fn), "w") as fh:
                fh.write(text)
        os.makedirs(os.path.join(wd, "output"))
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([PRINCIPAL], cwd=wd, env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:])
        files = sorted(os.listdir(os.path.join(wd, "output")))
        assert len(files) == 2, files
        arrays = {"deck": np.array(json.dumps(d.__dict__)), "filenames": np.array(files)}
        texts = []
        for q, fn in enumerate(files):
            with open(os.path.join(wd, "output", fn)) as fh:
                text = fh.read()
            texts.append(text)
            lines = text.splitlines()
            arrays[f"e{q}_header"] = np.array(lines[:4])
            arrays[f"e{q}_xyz_phi_tp_c"] = np.array(
                [[float(x) for x in ln.split()] for ln in lines[4:] if ln.strip()])
        # the console transcript and performance_log.txt, for the line-by-line
        # comparison (the try lines apart: their count and a sample)
        out = r.stdout.splitlines()
        tries = [ln for ln in out if ln.lstrip().startswith(W0E0_TRY)]
        idx = _try_sample(len(tries))
        arrays["stdout_lines"] = np.array([ln for ln in out
                                           if not ln.lstrip().startswith(W0E0_TRY)])
        arrays["w0e0_tries"] = np.array(len(tries))
        arrays["w0e0_sample_index"] = np.array(idx, dtype=np.int64)
        arrays["w0e0_sample_lines"] = np.array([tries[q] for q in idx])
        with open(os.path.join(wd, "performance_log.txt")) as fh:
            arrays["perf_log_lines"] = np.array(fh.read().splitlines())
        diag = [ln for ln in r.stdout.splitlines() if ln.startswith("L2(phi)")]
        arrays["diag"] = np.array([[fortran_float(t) for t in ln.replace("==>", " ").split()[1::2]]
                                   for ln in diag])
        arrays["probe_lines"] = np.array([ln for ln in r.stdout.splitlines()
                                          if len(ln.split()) == 3 and "E" in ln
                                          and "==>" not in ln][:36])
        return arrays, texts
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def make_principal(name="sphere"):
    """The unmodified principal.f90 on a small deck: its output events
    (it = 5, 9) write Tecplot POINT files with 8 significant digits."""
    d = Deck(stoptime=0.0, it_max=9, WriteFileD_timeCounter=4, **PRINCIPAL_DECKS[name])
    arrays, _ = _run_principal(d)
    arrays["event_its"] = np.array([5, 9])
    save(f"principal_{name}.npz", **arrays)


def make_principal_stop():
    """The planar deck stopped by stoptime (principal.f90:231-255): 6.5 time
    steps, so the run ends at it = 7 with the termination banner, a final
    snapshot (E_R < N) and the final timing block; one regular event at 5."""
    base = dict(PRINCIPAL_DECKS["planar"])
    d = Deck(stoptime=0.0, it_max=50, WriteFileD_timeCounter=4, **base)
    dt = d.dx * d.dx / (7.5 * d.D_liq)
    d.stoptime = 6.5 * dt
    arrays, _ = _run_principal(d)
    arrays["event_its"] = np.array([5, 7])
    save("principal_stop.npz", **arrays)


def make_restart():
    """principal.f90 continued from restart.dat (tipo_inicial = 4,
    entrada.f90:272-321): the planar deck's second output event is the
    restart file, shifted down by k_desl = R_semente = 2 layers."""
    base = dict(PRINCIPAL_DECKS["planar"])
    d0 = Deck(stoptime=0.0, it_max=9, WriteFileD_timeCounter=4, **base)
    _, texts = _run_principal(d0)
    base.update(tipo_inicial=4, R_semente=2.0)
    d1 = Deck(stoptime=0.0, it_max=9, WriteFileD_timeCounter=4, **base)
    arrays, _ = _run_principal(d1, {"restart.dat": texts[1]})
    arrays["restart_dat"] = np.array(texts[1])
    arrays["event_its"] = np.array([5, 9])
    save("principal_restart.npz", **arrays)


def main(only=None):
    """All fixtures, or only the principal_* ones (`make_golden.py principal`)."""
    build_ref()
    if only != "principal":
        make_random_kat()
        make_scalars_kat()
        make_unit()
        make_runs()
    for name in PRINCIPAL_DECKS:
        make_principal(name)
    make_principal_stop()
    make_restart()


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
