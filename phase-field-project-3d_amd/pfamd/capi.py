"""ctypes binding of the C-ABI (include/pf_capi.h) exported by libpfamd.so.

This is the Python mirror of the Fortran ISO_C_BINDING module
(fortran/pf_capi.f90): the same entry points, the same POD structs.  The
library path is the in-tree build (lib/libpfamd.so); loading fails loudly if
it is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libpfamd.so")

PF_OK, PF_E_ARG, PF_E_HIP, PF_E_NOMEM, PF_E_STATE, PF_E_NODEV = 0, -1, -2, -3, -4, -5
PF_FP64, PF_FP32_MIXED = 0, 1


def batch_tiles(a: int, b: int, s: int) -> int:
    """PF_BATCH_TILES: a batch size class's tile shapes for sweep A, sweep B
    (the split form's phase + T kernel) and the split form's solute kernel."""
    return 0x1000 | a | (b << 4) | (s << 8)


class PfParams(C.Structure):
    """struct pf_params (include/pf_capi.h)."""

    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("it", C.c_int32),
        ("n_varreduras", C.c_int32), ("h_quintico", C.c_int32), ("ruido_tipo", C.c_int32),
        ("precision", C.c_int32), ("seed", C.c_uint64),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("D", C.c_double), ("delta_H", C.c_double), ("cp_s", C.c_double), ("TP", C.c_double),
        ("Tm", C.c_double), ("Me", C.c_double), ("Vm", C.c_double), ("R", C.c_double),
        ("E0", C.c_double), ("W0", C.c_double), ("M0", C.c_double), ("delta_eps", C.c_double),
        ("AAA", C.c_double), ("D_liq", C.c_double), ("D_sol", C.c_double),
        ("Cs_eq", C.c_double), ("Cl_eq", C.c_double), ("Ke", C.c_float), ("reserved", C.c_int32),
    ]

    def copy(self) -> "PfParams":
        q = PfParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(PfParams))
        return q

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PfLoop(C.Structure):
    """struct pf_loop (include/pf_capi.h)."""

    _fields_ = [("T0", C.c_double), ("taxa_resfr", C.c_double), ("Sigma", C.c_float),
                ("Beta", C.c_float), ("ajuste", C.c_float), ("time", C.c_float)]

    def copy(self) -> "PfLoop":
        q = PfLoop()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(PfLoop))
        return q


class PfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pf error {code}: {msg}")
        self.code = code


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None

# Every symbol declared in include/pf_capi.h (tests check the library exports them).
EXPORTS = (
    "pf_version", "pf_last_error", "pf_device_count", "pf_ctx_device", "pf_pseudo_random", "pf_ruido_suave",
    "pf_w0_e0", "pf_w0_e0_tries", "pf_auxiliar_m0", "pf_equilibrium", "pf_time_step", "pf_noise_scale",
    "pf_initial_fields", "pf_init", "pf_finalize", "pf_temperatura", "pf_equacao_campo_fase",
    "pf_phase_new_to_old", "pf_concentracao_soluto", "pf_end_step", "pf_diagnosticos",
    "pf_download", "pf_download_all", "pf_run_steps", "pf_synchronize", "pf_device_bytes",
    "pf_timing_enable", "pf_timing_kinds", "pf_timing_read", "pf_set_fused", "pf_stream_copy", "pf_set_tiling",
    "pf_get_tiling", "pf_autotune", "pf_set_mass_correction", "pf_mass_ledger",
    "pf_checkpoint_save", "pf_checkpoint_load", "pf_set_sweep_split",
    "pf_get_sweep_split", "pf_set_sweep_tiling", "pf_get_sweep_tiling",
    "pf_batch_create", "pf_batch_destroy", "pf_batch_add", "pf_batch_remove", "pf_batch_run",
    "pf_batch_set_graph_steps", "pf_batch_synchronize", "pf_batch_status", "pf_batch_scalars",
    "pf_batch_download", "pf_batch_diagnosticos", "pf_batch_bytes", "pf_batch_size_class",
    "pf_batch_set_tiling", "pf_batch_get_tiling", "pf_batch_set_group_target",
    "pf_batch_autotune", "pf_batch_set_mass_correction", "pf_batch_mass_ledger",
    "pf_diagnosticos_noalloc", "pf_batch_set_sweep_split",
)
TILE_SHAPES = ((64, 8), (32, 8), (32, 16), (16, 16))
PF_KT_NAMES = ("sweep_A", "sweep_B", "boundary", "temperatura", "dropin", "mass", "solute")


def load(path: str = LIB_PATH):
    """Load libpfamd.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(path)
    i32, u64, f64, f32, vp = C.c_int32, C.c_uint64, C.c_double, C.c_float, C.c_void_p
    P, L = C.POINTER(PfParams), C.POINTER(PfLoop)
    sig = {
        "pf_version": ([], C.c_char_p),
        "pf_last_error": ([], C.c_char_p),
        "pf_device_count": ([C.POINTER(i32)], C.c_int),
        "pf_ctx_device": ([C.c_void_p, C.POINTER(i32)], C.c_int),
        "pf_pseudo_random": ([i32, i32, i32, i32, i32, u64], f64),
        "pf_ruido_suave": ([i32, i32, i32, i32, u64], f64),
        "pf_w0_e0": ([f64, f32, f32, C.POINTER(f64), C.POINTER(f64)], i32),
        "pf_w0_e0_tries": ([f64, f32, f32, C.POINTER(f64), i32], i32),
        "pf_auxiliar_m0": ([f64, f64, f64, f64, f64, f64, f64, f32, f32, f32, f64, f64, f32], f64),
        "pf_equilibrium": ([f64, f64, f32, f64, C.POINTER(f64), C.POINTER(f64)], None),
        "pf_time_step": ([f64, f64], f64),
        "pf_noise_scale": ([f64, f64, f64, f64, i32], f64),
        "pf_initial_fields": ([i32, i32, i32, i32, f64, f64, f64, f32, i32, u64] + [_dp] * 7, C.c_int),
        "pf_init": ([P] + [vp] * 7 + [C.POINTER(vp)], C.c_int),
        "pf_finalize": ([vp], C.c_int),
        "pf_temperatura": ([vp, P], C.c_int),
        "pf_equacao_campo_fase": ([vp, P, vp], C.c_int),
        "pf_phase_new_to_old": ([vp], C.c_int),
        "pf_concentracao_soluto": ([vp, P], C.c_int),
        "pf_end_step": ([vp], C.c_int),
        "pf_diagnosticos": ([vp, P, vp], C.c_int),
        "pf_download": ([vp, vp, vp, vp], C.c_int),
        "pf_download_all": ([vp] + [vp] * 7, C.c_int),
        "pf_set_tiling": ([vp, i32, i32, i32], C.c_int),
        "pf_get_tiling": ([vp, vp], C.c_int),
        "pf_autotune": ([vp, P, i32, vp, vp], C.c_int),
        "pf_set_mass_correction": ([vp, i32], C.c_int),
        "pf_set_sweep_split": ([vp, i32, i32], C.c_int),
        "pf_get_sweep_split": ([vp, vp], C.c_int),
        "pf_set_sweep_tiling": ([vp, i32, i32, i32], C.c_int),
        "pf_get_sweep_tiling": ([vp, i32, vp], C.c_int),
        "pf_checkpoint_save": ([vp, P, L, C.c_char_p], C.c_int),
        "pf_checkpoint_load": ([C.c_char_p, P, L, C.POINTER(vp)], C.c_int),
        "pf_batch_create": ([C.c_size_t, C.POINTER(vp)], C.c_int),
        "pf_batch_destroy": ([vp], C.c_int),
        "pf_batch_add": ([vp, P, L, i32, vp, vp, vp, C.POINTER(i32)], C.c_int),
        "pf_batch_remove": ([vp, i32], C.c_int),
        "pf_batch_run": ([vp, i32], C.c_int),
        "pf_batch_set_graph_steps": ([vp, i32], C.c_int),
        "pf_batch_synchronize": ([vp], C.c_int),
        "pf_batch_status": ([vp, i32, C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "pf_batch_scalars": ([vp, i32, P, L], C.c_int),
        "pf_batch_download": ([vp, i32, vp, vp, vp], C.c_int),
        "pf_batch_diagnosticos": ([vp, i32, vp], C.c_int),
        "pf_batch_bytes": ([vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
        "pf_batch_size_class": ([i32, i32, i32], i32),
        "pf_batch_set_tiling": ([vp, i32, i32], C.c_int),
        "pf_batch_get_tiling": ([vp, i32, C.POINTER(i32)], C.c_int),
        "pf_batch_set_group_target": ([vp, i32], C.c_int),
        "pf_batch_autotune": ([vp, i32, i32], C.c_int),
        "pf_batch_set_mass_correction": ([vp, i32, i32], C.c_int),
        "pf_batch_set_sweep_split": ([vp, i32], C.c_int),
        "pf_batch_mass_ledger": ([vp, i32, vp], C.c_int),
        "pf_diagnosticos_noalloc": ([P, vp, vp, vp, vp], C.c_int),
        "pf_mass_ledger": ([vp, vp], C.c_int),
        "pf_run_steps": ([vp, P, L, i32, vp], C.c_int),
        "pf_synchronize": ([vp], C.c_int),
        "pf_device_bytes": ([vp, C.POINTER(C.c_size_t)], C.c_int),
        "pf_timing_enable": ([vp, i32], C.c_int),
        "pf_timing_kinds": ([vp, C.c_uint32], C.c_int),
        "pf_timing_read": ([vp, vp, vp], C.c_int),
        "pf_set_fused": ([vp, i32], C.c_int),
        "pf_stream_copy": ([C.c_size_t, i32, C.POINTER(f64)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int, lib=None):
    if rc != PF_OK:
        lib = lib or load()
        raise PfError(rc, (lib.pf_last_error() or b"").decode())
    return rc


def device_count(lib=None) -> int:
    lib = lib or load()
    n = C.c_int32(0)
    check(lib.pf_device_count(C.byref(n)), lib)
    return n.value


def stream_copy_gbs(nbytes: int, iters: int = 20, lib=None) -> float:
    lib = lib or load()
    g = C.c_double(0.0)
    check(lib.pf_stream_copy(int(nbytes), int(iters), C.byref(g)), lib)
    return g.value


def w0_e0(lam, ajuste, sigma, lib=None):
    lib = lib or load()
    W0, E0 = C.c_double(), C.c_double()
    idx = lib.pf_w0_e0(lam, ajuste, sigma, C.byref(W0), C.byref(E0))
    return W0.value, E0.value, idx


def w0_e0_tries(lam, ajuste, sigma, lib=None):
    """|E0 - try| of every try of the W0/E0 search, up to the accepted one."""
    lib = lib or load()
    _, _, idx = w0_e0(lam, ajuste, sigma, lib)
    n = idx + 1 if idx >= 0 else 10000001
    out = np.zeros(n)
    got = lib.pf_w0_e0_tries(lam, ajuste, sigma, out.ctypes.data_as(C.POINTER(C.c_double)), n)
    return out[:got] if got > 0 else out


def equilibrium(TP, Tm, Ke, Me, lib=None):
    lib = lib or load()
    cs, cl = C.c_double(), C.c_double()
    lib.pf_equilibrium(TP, Tm, Ke, Me, C.byref(cs), C.byref(cl))
    return cs.value, cl.value


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """Owns one pf_ctx: the seven device-resident fields of one instance."""

    def __init__(self, p: PfParams, fields, lib=None):
        self.lib = lib or load()
        self.n = p.nx * p.ny * p.nz
        arrs = [None if f is None else np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
                for f in fields]
        for a in arrs:
            if a is not None and a.size != self.n:
                raise ValueError("field size does not match the grid")
        h = C.c_void_p()
        check(self.lib.pf_init(C.byref(p), *[_ptr(a) for a in arrs], C.byref(h)), self.lib)
        self.h = h

    @classmethod
    def from_checkpoint(cls, path: str, lib=None):
        """(Context, PfParams, PfLoop) from a pf_checkpoint_save file."""
        self = cls.__new__(cls)
        self.lib = lib or load()
        p, L, h = PfParams(), PfLoop(), C.c_void_p()
        check(self.lib.pf_checkpoint_load(path.encode(), C.byref(p), C.byref(L), C.byref(h)),
              self.lib)
        self.h = h
        self.n = p.nx * p.ny * p.nz
        return self, p, L

    # -- drop-in calls -----------------------------------------------------
    def temperatura(self, p):
        check(self.lib.pf_temperatura(self.h, C.byref(p)), self.lib)

    def equacao_campo_fase(self, p, want_probe=False):
        pr = (C.c_double * 3)()
        check(self.lib.pf_equacao_campo_fase(self.h, C.byref(p),
                                             C.cast(pr, C.c_void_p) if want_probe else None), self.lib)
        return list(pr) if want_probe else None

    def phase_new_to_old(self):
        check(self.lib.pf_phase_new_to_old(self.h), self.lib)

    def concentracao_soluto(self, p):
        check(self.lib.pf_concentracao_soluto(self.h, C.byref(p)), self.lib)

    def end_step(self):
        check(self.lib.pf_end_step(self.h), self.lib)

    def diagnosticos(self, p):
        out = np.zeros(5)
        check(self.lib.pf_diagnosticos(self.h, C.byref(p), _ptr(out)), self.lib)
        return out

    def download(self):
        phi, c, T = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        check(self.lib.pf_download(self.h, _ptr(phi), _ptr(c), _ptr(T)), self.lib)
        return phi, c, T

    def download_all(self):
        arrs = [np.empty(self.n) for _ in range(7)]
        check(self.lib.pf_download_all(self.h, *[_ptr(a) for a in arrs]), self.lib)
        return arrs

    # -- fused path ----------------------------------------------------------
    def run_steps(self, p, L, nsteps, probes=False):
        pr = np.zeros(3 * max(1, p.n_varreduras) * nsteps) if probes else None
        check(self.lib.pf_run_steps(self.h, C.byref(p), C.byref(L), int(nsteps), _ptr(pr)), self.lib)
        return pr

    def synchronize(self):
        check(self.lib.pf_synchronize(self.h), self.lib)

    def device_bytes(self):
        b = C.c_size_t(0)
        check(self.lib.pf_device_bytes(self.h, C.byref(b)), self.lib)
        return b.value

    def set_fused(self, on: bool):
        check(self.lib.pf_set_fused(self.h, 1 if on else 0), self.lib)

    def set_tiling(self, tile_a: int, tile_b: int, target_groups: int = 1024):
        check(self.lib.pf_set_tiling(self.h, tile_a, tile_b, target_groups), self.lib)

    def set_sweep_split(self, on: bool, tile_s: int = 0):
        check(self.lib.pf_set_sweep_split(self.h, 1 if on else 0, int(tile_s)), self.lib)

    def sweep_split(self):
        """(split on, solute kernel's tile shape, its z-planes per segment)."""
        out = np.zeros(3, dtype=np.int32)
        check(self.lib.pf_get_sweep_split(self.h, out.ctypes.data_as(C.c_void_p)), self.lib)
        return (bool(out[0]), int(out[1]), int(out[2]))

    def set_sweep_tiling(self, kind: int, tile: int, target: int):
        """Tile shape and work-group target of kernel kind 0 (sweep A), 1
        (sweep B / its phase + T kernel) or 2 (the split form's solute kernel)."""
        check(self.lib.pf_set_sweep_tiling(self.h, int(kind), int(tile), int(target)), self.lib)

    def sweep_tiling(self, kind: int):
        """(tile, target, z-planes per segment) of kernel kind 0, 1 or 2."""
        out = np.zeros(3, dtype=np.int32)
        check(self.lib.pf_get_sweep_tiling(self.h, int(kind), out.ctypes.data_as(C.c_void_p)),
              self.lib)
        return tuple(int(v) for v in out)

    def tiling(self):
        """(tile_a, tile_b, z-planes per segment A, B)."""
        out = np.zeros(4, dtype=np.int32)
        check(self.lib.pf_get_tiling(self.h, out.ctypes.data_as(C.c_void_p)), self.lib)
        return tuple(int(v) for v in out)

    def autotune(self, p, reps: int = 3):
        """Times every tiling candidate; keeps the fastest.  Returns a list of
        (kind, tile, target_groups, ms) with kind A (sweep A), B (fused sweep
        B), P (phase + T of the split form), S (its solute kernel)."""
        n = C.c_int32(0)
        ms = np.zeros(4 * 256)  # pf_autotune writes at most 256 candidates
        check(self.lib.pf_autotune(self.h, C.byref(p), reps, C.byref(n), _ptr(ms)), self.lib)
        return [("ABPS"[int(ms[4 * q])], int(ms[4 * q + 1]), int(ms[4 * q + 2]),
                 float(ms[4 * q + 3])) for q in range(n.value)]

    def checkpoint_save(self, p, L, path: str):
        check(self.lib.pf_checkpoint_save(self.h, C.byref(p), C.byref(L), path.encode()),
              self.lib)

    def set_mass_correction(self, every: int):
        check(self.lib.pf_set_mass_correction(self.h, int(every)), self.lib)

    def mass_ledger(self):
        """(ledger, interior sum at the last correction, last factor)."""
        out = np.zeros(3)
        check(self.lib.pf_mass_ledger(self.h, _ptr(out)), self.lib)
        return tuple(float(v) for v in out)

    def timing(self, on: bool):
        check(self.lib.pf_timing_enable(self.h, 1 if on else 0), self.lib)

    def timing_kinds(self, names=None):
        """Record only launches of these kinds (names of PF_KT_NAMES; None:
        every kind)."""
        mask = 0xFFFFFFFF if names is None else sum(1 << PF_KT_NAMES.index(n) for n in names)
        check(self.lib.pf_timing_kinds(self.h, mask), self.lib)

    def timing_read(self):
        """{kind: (total_ms, launches)} since timing(True)."""
        ms = np.zeros(len(PF_KT_NAMES))
        cnt = np.zeros(len(PF_KT_NAMES), dtype=np.int64)
        check(self.lib.pf_timing_read(self.h, _ptr(ms), cnt.ctypes.data_as(C.c_void_p)), self.lib)
        return {nm: (float(m), int(c)) for nm, m, c in zip(PF_KT_NAMES, ms, cnt)}

    def close(self):
        if self.h:
            check(self.lib.pf_finalize(self.h), self.lib)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.pf_finalize(self.h)
                self.h = None
        except Exception:
            pass


class Batch:
    """Many independent instances advanced together (pf_batch_*)."""

    def __init__(self, hbm_budget: int = 0, lib=None):
        self.lib = lib or load()
        h = C.c_void_p()
        check(self.lib.pf_batch_create(C.c_size_t(hbm_budget), C.byref(h)), self.lib)
        self.h = h

    def add(self, p, L, nsteps: int, phi, c, T) -> int:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi, c, T)]
        slot = C.c_int32(-1)
        check(self.lib.pf_batch_add(self.h, C.byref(p), C.byref(L), int(nsteps),
                                    *[_ptr(a) for a in arrs], C.byref(slot)), self.lib)
        return slot.value

    def remove(self, slot: int):
        check(self.lib.pf_batch_remove(self.h, int(slot)), self.lib)

    def run(self, nsteps: int):
        check(self.lib.pf_batch_run(self.h, int(nsteps)), self.lib)

    def set_graph_steps(self, k: int):
        check(self.lib.pf_batch_set_graph_steps(self.h, int(k)), self.lib)

    def synchronize(self):
        check(self.lib.pf_batch_synchronize(self.h), self.lib)

    def status(self, slot: int):
        done, left = C.c_int32(0), C.c_int32(0)
        check(self.lib.pf_batch_status(self.h, int(slot), C.byref(done), C.byref(left)), self.lib)
        return done.value, left.value

    def scalars(self, slot: int):
        p, L = PfParams(), PfLoop()
        check(self.lib.pf_batch_scalars(self.h, int(slot), C.byref(p), C.byref(L)), self.lib)
        return p, L

    def download(self, slot: int, n: int):
        phi, c, T = np.empty(n), np.empty(n), np.empty(n)
        check(self.lib.pf_batch_download(self.h, int(slot), _ptr(phi), _ptr(c), _ptr(T)),
              self.lib)
        return phi, c, T

    def diagnosticos(self, slot: int):
        out = np.zeros(5)
        check(self.lib.pf_batch_diagnosticos(self.h, int(slot), _ptr(out)), self.lib)
        return out

    def bytes(self):
        u, b = C.c_size_t(0), C.c_size_t(0)
        check(self.lib.pf_batch_bytes(self.h, C.byref(u), C.byref(b)), self.lib)
        return u.value, b.value

    def size_class(self, precision: int, nx: int, ny: int) -> int:
        return int(self.lib.pf_batch_size_class(int(precision), int(nx), int(ny)))

    def set_tiling(self, size_class: int, shape: int):
        check(self.lib.pf_batch_set_tiling(self.h, int(size_class), int(shape)), self.lib)

    def tiling(self, size_class: int) -> int:
        """One shape when every kernel kind of the class has it, else the
        batch_tiles(a, b, s) value."""
        sh = C.c_int32(0)
        check(self.lib.pf_batch_get_tiling(self.h, int(size_class), C.byref(sh)), self.lib)
        return sh.value

    def tiling_kinds(self, size_class: int):
        """(sweep A, sweep B, solute) tile shapes of the class."""
        v = self.tiling(size_class)
        return (v, v, v) if v < 0x1000 else (v & 0xf, (v >> 4) & 0xf, (v >> 8) & 0xf)

    def set_group_target(self, target_groups: int):
        check(self.lib.pf_batch_set_group_target(self.h, int(target_groups)), self.lib)

    def autotune(self, reps: int = 2, size_class: int = -1):
        check(self.lib.pf_batch_autotune(self.h, int(size_class), int(reps)), self.lib)

    def set_mass_correction(self, slot: int, every: int):
        check(self.lib.pf_batch_set_mass_correction(self.h, int(slot), int(every)), self.lib)

    def set_sweep_split(self, on: bool):
        check(self.lib.pf_batch_set_sweep_split(self.h, 1 if on else 0), self.lib)

    def mass_ledger(self, slot: int):
        out = np.zeros(3)
        check(self.lib.pf_batch_mass_ledger(self.h, int(slot), _ptr(out)), self.lib)
        return tuple(float(v) for v in out)

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.pf_batch_destroy(self.h), self.lib)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.pf_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass
