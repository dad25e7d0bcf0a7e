"""ctypes binding of the C oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / the CPU baseline.  Nothing in
the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    from pfamd.capi import PfParams, PfLoop  # shared POD layouts

    i32, u64, f64, f32 = C.c_int32, C.c_uint64, C.c_double, C.c_float
    lib.pfo_pseudo_random.argtypes = [i32, i32, i32, i32, i32, u64]
    lib.pfo_pseudo_random.restype = f64
    lib.pfo_ruido_suave.argtypes = [i32, i32, i32, i32, u64]
    lib.pfo_ruido_suave.restype = f64
    lib.pfo_w0_e0.argtypes = [f64, f32, f32, C.POINTER(f64), C.POINTER(f64)]
    lib.pfo_w0_e0.restype = i32
    lib.pfo_auxiliar.argtypes = [f64, f64, f64, f64, f64, f64, f64, f32, f32, f32, f64, f64, f32]
    lib.pfo_auxiliar.restype = f64
    lib.pfo_delta_t.argtypes = [f64, f64]
    lib.pfo_delta_t.restype = f64
    lib.pfo_noise_scale.argtypes = [f64, f64, f64, f64, i32]
    lib.pfo_noise_scale.restype = f64
    lib.pfo_initial_fields.argtypes = [i32, i32, i32, i32, f64, f64, f64, f32, i32, u64] + [_dp] * 7
    lib.pfo_initial_fields.restype = C.c_int
    P = C.POINTER(PfParams)
    lib.pfo_temperatura.argtypes = [P, _dp, _dp, _dp, _dp]
    lib.pfo_temperatura.restype = None
    lib.pfo_equacao_campo_fase.argtypes = [P, _dp, _dp, _dp, _dp, C.c_void_p]
    lib.pfo_equacao_campo_fase.restype = None
    lib.pfo_concentracao_soluto.argtypes = [P, _dp, _dp, _dp]
    lib.pfo_concentracao_soluto.restype = None
    lib.pfo_diagnosticos.argtypes = [P, _dp, _dp, _dp]
    lib.pfo_diagnosticos.restype = None
    lib.pfo_run_steps.argtypes = [P, C.POINTER(PfLoop), i32] + [_dp] * 7 + [C.c_void_p]
    lib.pfo_run_steps.restype = C.c_int
    lib.pfo_set_threads.argtypes = [i32]
    lib.pfo_set_threads.restype = None
    lib.pfo_get_threads.argtypes = []
    lib.pfo_get_threads.restype = i32
    _lib = lib
    return lib


def set_threads(n: int) -> None:
    load().pfo_set_threads(int(n))


def pseudo_random(i, j, k, it, salt, seed=0):
    return load().pfo_pseudo_random(i, j, k, it, salt, seed)


def ruido_suave(i, j, k, it, seed=0):
    return load().pfo_ruido_suave(i, j, k, it, seed)


def w0_e0(lam, ajuste, sigma):
    W0, E0 = C.c_double(), C.c_double()
    idx = load().pfo_w0_e0(lam, ajuste, sigma, C.byref(W0), C.byref(E0))
    return W0.value, E0.value, idx


def auxiliar(TP, Cl_eq, Cs_eq, R, Vm, E0, W0, Sigma, Beta, Ke, Me, D_liq, ajuste):
    return load().pfo_auxiliar(TP, Cl_eq, Cs_eq, R, Vm, E0, W0, Sigma, Beta, Ke, Me, D_liq, ajuste)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


class Fields:
    """The seven reference fields, flat column-major FP64 arrays."""

    NAMES = ("phi_old", "phi_new", "c_old", "c_new", "T_old", "T_new", "dpp")

    def __init__(self, n, arrays=None):
        if arrays is None:
            arrays = [np.zeros(n) for _ in self.NAMES]
        for nm, a in zip(self.NAMES, arrays):
            setattr(self, nm, _f(a).copy())

    def copy(self):
        return Fields(None, [getattr(self, nm) for nm in self.NAMES])

    def as_list(self):
        return [getattr(self, nm) for nm in self.NAMES]


def initial_fields(nx, ny, nz, tipo, R, T0, C0, Ke, n_seeds=0, seed_rng=0):
    n = nx * ny * nz
    F = Fields(n)
    load().pfo_initial_fields(nx, ny, nz, tipo, R, T0, C0, Ke, n_seeds, seed_rng, *F.as_list())
    return F


def temperatura(p, F):
    load().pfo_temperatura(C.byref(p), F.T_old, F.phi_old, F.dpp, F.T_new)


def equacao_campo_fase(p, F, want_probe=False):
    pr = (C.c_double * 3)()
    load().pfo_equacao_campo_fase(C.byref(p), F.phi_old, F.c_old, F.phi_new, F.dpp,
                                  C.cast(pr, C.c_void_p) if want_probe else None)
    return list(pr) if want_probe else None


def concentracao_soluto(p, F):
    load().pfo_concentracao_soluto(C.byref(p), F.phi_old, F.c_old, F.c_new)


def diagnosticos(p, phi_new, c_new):
    out = np.zeros(5)
    load().pfo_diagnosticos(C.byref(p), _f(phi_new), _f(c_new), out)
    return out


def run_steps(p, L, nsteps, F, probes=False):
    pr = None
    if probes:
        pr = np.zeros(3 * max(1, p.n_varreduras) * nsteps)
    load().pfo_run_steps(C.byref(p), C.byref(L), nsteps, *F.as_list(),
                         pr.ctypes.data_as(C.c_void_p) if pr is not None else None)
    return pr


# ---- Karma model (karma/karma3d.f90), oracle/karma_oracle.c ----------------

class KarmaPar(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("eps4", C.c_double), ("dx", C.c_double),
                ("dt", C.c_double), ("Fnoise", C.c_double), ("k", C.c_double),
                ("a2", C.c_double)]


def karma_par(lambda_=4.0, eps4=0.02, dx=0.8, dt=0.008, Fnoise=0.0, k=0.14):
    """The step parameters of karma.nml (karma3d.f90:71-86); a2 = 0.6267."""
    return KarmaPar(lambda_, eps4, dx, dt, Fnoise, k, 0.6267)


def karma_omega(k=0.14, Tm=933.3, mliq=640.0, C0=0.0196, TP=909.5):
    """(c_l0, Omega) of karma3d.f90:112-114."""
    cl0 = (Tm - TP) / mliq
    return cl0, (cl0 - C0) / ((1.0 - k) * cl0)


def _karma_lib():
    lib = load()
    if not getattr(lib, "_karma_set", False):
        i32, f64 = C.c_int32, C.c_double
        lib.pfo_karma_init.argtypes = [i32, f64, f64, f64, _dp, _dp]
        lib.pfo_karma_init.restype = None
        lib.pfo_karma_mirror.argtypes = [i32, _dp]
        lib.pfo_karma_mirror.restype = None
        lib.pfo_karma_noise.argtypes = [i32, i32, i32, i32]
        lib.pfo_karma_noise.restype = f64
        lib.pfo_karma_step.argtypes = [i32, C.POINTER(KarmaPar), i32, _dp, _dp]
        lib.pfo_karma_step.restype = None
        lib._karma_set = True
    return lib


def karma_init(n, dx=0.8, R0=8.0, Omega=None):
    """Padded (n+2)^3 phi and U (index i + (n+2)(j + (n+2)k)) of the corner
    seed."""
    if Omega is None:
        Omega = karma_omega()[1]
    s = (n + 2) ** 3
    phi, U = np.zeros(s), np.zeros(s)
    _karma_lib().pfo_karma_init(n, dx, R0, Omega, phi, U)
    return phi, U


def karma_step(n, par, it, phi, U):
    _karma_lib().pfo_karma_step(n, C.byref(par), it, phi, U)


def karma_mirror(n, f):
    _karma_lib().pfo_karma_mirror(n, f)


def karma_noise(i, j, k, it):
    return _karma_lib().pfo_karma_noise(i, j, k, it)


def karma_interior(n, f):
    """The cells 1..n of a padded field as an (n, n, n) array indexed [k, j, i]."""
    return np.asarray(f).reshape(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]
