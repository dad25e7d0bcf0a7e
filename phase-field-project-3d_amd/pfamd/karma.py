"""The quantitative Karma model (karma/karma3d.f90) through the C-ABI of
include/pf_karma.h: a 1/8 seed at the corner of an n^3 box with mirror
faces, FP64 fields resident on the GPU."""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import capi

EXPORTS = ("pf_karma_create", "pf_karma_run", "pf_karma_download", "pf_karma_synchronize",
           "pf_karma_destroy")


class KarmaParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad_", C.c_int32), ("lambda_", C.c_double),
                ("eps4", C.c_double), ("dx", C.c_double), ("dt", C.c_double), ("R0", C.c_double),
                ("Fnoise", C.c_double), ("k", C.c_double), ("Tm", C.c_double),
                ("mliq", C.c_double), ("C0", C.c_double), ("TP", C.c_double)]


@dataclasses.dataclass
class KarmaDeck:
    """karma.nml with the program's defaults (karma/karma3d.f90:71-84)."""
    N: int = 200
    lambda_: float = 4.0
    eps4: float = 0.02
    dx: float = 0.8
    dt: float = 0.008
    R0: float = 8.0
    Fnoise: float = 0.0
    k: float = 0.14
    Tm: float = 933.3
    mliq: float = 640.0
    C0: float = 0.0196
    TP: float = 909.5

    def params(self) -> KarmaParams:
        return KarmaParams(self.N, 0, self.lambda_, self.eps4, self.dx, self.dt, self.R0,
                           self.Fnoise, self.k, self.Tm, self.mliq, self.C0, self.TP)


def _lib():
    lib = capi.load()
    if not getattr(lib, "_karma_sig", False):
        i32, vp = C.c_int32, C.c_void_p
        lib.pf_karma_create.argtypes = [C.POINTER(KarmaParams), C.POINTER(vp)]
        lib.pf_karma_run.argtypes = [vp, i32]
        lib.pf_karma_download.argtypes = [vp, vp, vp, C.POINTER(i32)]
        lib.pf_karma_synchronize.argtypes = [vp]
        lib.pf_karma_destroy.argtypes = [vp]
        for nm in EXPORTS:
            getattr(lib, nm).restype = C.c_int
        lib._karma_sig = True
    return lib


class Karma:
    def __init__(self, deck: KarmaDeck):
        self.deck = deck
        self.lib = _lib()
        self.h = C.c_void_p()
        p = deck.params()
        capi.check(self.lib.pf_karma_create(C.byref(p), C.byref(self.h)), self.lib)

    def run(self, nsteps: int):
        capi.check(self.lib.pf_karma_run(self.h, int(nsteps)), self.lib)

    def synchronize(self):
        capi.check(self.lib.pf_karma_synchronize(self.h), self.lib)

    def fields(self):
        """(phi, U, it): the cells 1..N as (N, N, N) arrays indexed [k, j, i]."""
        n = self.deck.N
        phi, U = np.empty(n ** 3), np.empty(n ** 3)
        it = C.c_int32(0)
        capi.check(self.lib.pf_karma_download(self.h, phi.ctypes.data, U.ctypes.data,
                                              C.byref(it)), self.lib)
        return phi.reshape(n, n, n), U.reshape(n, n, n), it.value

    def close(self):
        if self.h:
            rc = self.lib.pf_karma_destroy(self.h)
            self.h = C.c_void_p()
            capi.check(rc, self.lib)
