"""The quantitative Karma model (karma/karma3d.f90) through the C-ABI of
include/pf_karma.h: a 1/8 seed at the corner of an n^3 box with mirror
faces, FP64 fields resident on the GPU.  The program around the step (the
namelist, the tip and the output files) is bin/karma3d
(fortran/karma3d.f90); KarmaDeck and tips() give the same namelist and tip
rule to Python callers."""
from __future__ import annotations

import ctypes as C
import dataclasses
import re

import numpy as np

from . import capi

EXPORTS = ("pf_karma_create", "pf_karma_run", "pf_karma_download", "pf_karma_synchronize",
           "pf_karma_destroy", "pf_karma_set_zmarch")


class KarmaParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad_", C.c_int32), ("lambda_", C.c_double),
                ("eps4", C.c_double), ("dx", C.c_double), ("dt", C.c_double), ("R0", C.c_double),
                ("Fnoise", C.c_double), ("k", C.c_double), ("Tm", C.c_double),
                ("mliq", C.c_double), ("C0", C.c_double), ("TP", C.c_double)]


# thin-interface constants of karma/karma3d.f90:91
A1, A2 = 0.8839, 0.6267


@dataclasses.dataclass
class KarmaDeck:
    """The namelist /karma/ with the program's compiled-in defaults
    (karma/karma3d.f90:75-89); `shipped()` is the karma.nml the reference
    ships next to it, and `from_nml()` reads one."""
    N: int = 200
    lambda_: float = 4.0
    eps4: float = 0.02
    dx: float = 0.8
    dt: float = 0.008
    R0: float = 8.0
    Fnoise: float = 0.0
    nsteps: int = 100000
    nout: int = 2000
    tip_stop: float = 0.9
    k: float = 0.14
    Tm: float = 933.3
    mliq: float = 640.0
    C0: float = 0.0196
    TP: float = 909.5
    sigma: float = 0.093
    dH: float = 0.891002e9
    Dl: float = 3.0e-9

    @classmethod
    def shipped(cls) -> "KarmaDeck":
        """karma/karma.nml:5-18 (noise on, 400k steps, a file every 5000)."""
        return cls(Fnoise=0.01, nsteps=400000, nout=5000)

    @classmethod
    def from_nml(cls, text: str) -> "KarmaDeck":
        """Entries of the group &karma ... / (names case-insensitive, '!'
        comments, ',' or newlines between entries, Fortran d/D exponents);
        absent entries keep the defaults.  Unknown names raise, as the
        Fortran read fails on them."""
        body = []
        for ln in text.splitlines():
            body.append(ln.split("!", 1)[0])
        src = "\n".join(body)
        low = src.lower()
        a = low.find("&karma")
        if a < 0:
            raise ValueError("no &karma group")
        b = low.find("/", a)
        if b < 0:
            raise ValueError("&karma group not closed by '/'")
        names = {f.name.rstrip("_").lower(): f for f in dataclasses.fields(cls)}
        d = cls()
        for item in re.split(r"[,\n]", src[a + len("&karma"):b]):
            if not item.strip():
                continue
            if "=" not in item:
                raise ValueError(f"bad namelist entry {item.strip()!r}")
            key, val = (t.strip() for t in item.split("=", 1))
            f = names.get(key.lower())
            if f is None:
                raise ValueError(f"unknown namelist entry {key!r}")
            v = float(val.lower().replace("d", "e"))
            setattr(d, f.name, int(v) if f.type in (int, "int") else v)
        return d

    def to_nml(self) -> str:
        out = []
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            out.append(f" {f.name.rstrip('_')} = {v!r}")
        return "&karma\n" + "\n".join(out) + "\n/\n"

    def scales(self) -> dict:
        """karma/karma3d.f90:110-116: D, c_l0, Omega, Gamma, d0, W, tau0."""
        D = A2 * self.lambda_
        cl0 = (self.Tm - self.TP) / self.mliq
        Omega = (cl0 - self.C0) / ((1.0 - self.k) * cl0)
        Gam = self.sigma * self.Tm / self.dH
        d0 = Gam / (self.mliq * (1.0 - self.k) * cl0)
        W = self.lambda_ / A1 * d0
        tau0 = A2 * self.lambda_ * W * W / self.Dl
        return dict(D=D, cl0=cl0, Omega=Omega, Gamma=Gam, d0=d0, W=W, tau0=tau0)

    def params(self) -> KarmaParams:
        return KarmaParams(self.N, 0, self.lambda_, self.eps4, self.dx, self.dt, self.R0,
                           self.Fnoise, self.k, self.Tm, self.mliq, self.C0, self.TP)


def tips(phi) -> tuple:
    """(tip_x, tip_d) of karma/karma3d.f90:294-298 for phi indexed [k, j, i]:
    the largest 1-based i with phi(i,1,1) > 0 and with phi(i,i,i) > 0 (0 when
    there is none)."""
    n = phi.shape[0]
    ax = np.nonzero(phi[0, 0, :] > 0.0)[0]
    dg = np.nonzero(phi[np.arange(n), np.arange(n), np.arange(n)] > 0.0)[0]
    return (int(ax[-1]) + 1 if ax.size else 0, int(dg[-1]) + 1 if dg.size else 0)


def _lib():
    lib = capi.load()
    if not getattr(lib, "_karma_sig", False):
        i32, vp = C.c_int32, C.c_void_p
        lib.pf_karma_create.argtypes = [C.POINTER(KarmaParams), C.POINTER(vp)]
        lib.pf_karma_run.argtypes = [vp, i32]
        lib.pf_karma_download.argtypes = [vp, vp, vp, C.POINTER(i32)]
        lib.pf_karma_synchronize.argtypes = [vp]
        lib.pf_karma_destroy.argtypes = [vp]
        lib.pf_karma_set_zmarch.argtypes = [vp, i32]
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

    def set_zmarch(self, on: bool):
        capi.check(self.lib.pf_karma_set_zmarch(self.h, int(bool(on))), self.lib)

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
