"""Builds libpfamd.so (gfx950 HIP kernels + the C-ABI) and the Fortran driver.

  python phase-field-project-3d_amd/build.py [--verbose] [--no-fortran]

Every source is compiled on each call (hipcc cross-compiles for gfx950
without a GPU).  Outputs stay in-tree: lib/libpfamd.so and bin/pf3d.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
OBJ_DIR = os.path.join(PKG, "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["pf_kernels_dropin.hip", "pf_fused.hip", "pf_batch.hip", "pf_runtime.hip", "pf_karma.hip",
               "pf_host.cpp"]
FORTRAN_SOURCES = ["pf_capi.f90", "pf_state.f90", "pf_kernels.f90", "pf_io.f90", "principal.f90"]
KARMA_SOURCES = ["pf_capi.f90", "karma_capi.f90", "karma3d.f90"]

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # keep hipcc from forming packed FP32 ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b + c becomes an FMA only within one source expression (decided by
    # the front end), never across statements by the code generator: how it
    # fuses across statements depends on the surrounding code, which differs
    # between kernels that inline the same arithmetic (own vs halo node,
    # tile shapes, the single-instance and batched kernels), and results
    # must not depend on which kernel computed a value
    "-ffp-contract=on",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(ln for ln in r.stdout.splitlines()
                    if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)
    return out


def build_lib(verbose=False, extra=()):
    os.makedirs(OBJ_DIR, exist_ok=True)

    def one(src):
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-x", "hip", *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd, verbose)
        return obj

    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    out = os.path.join(LIB_DIR, "libpfamd.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    return out


def build_fortran(verbose=False):
    """The Fortran programs -- bin/pf3d (principal + the drop-in
    subroutines) and bin/karma3d (the Karma model's program) -- linked
    against libpfamd.so with an rpath to the in-tree lib/."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(BIN_DIR, exist_ok=True)
    exes = []
    for name, files in (("pf3d", FORTRAN_SOURCES), ("karma3d", KARMA_SOURCES)):
        srcs = [os.path.join(FSRC, f) for f in files]
        if not all(os.path.exists(s) for s in srcs):
            continue
        mod = os.path.join(OBJ_DIR, "mod_" + name)
        os.makedirs(mod, exist_ok=True)
        exe = os.path.join(BIN_DIR, name)
        _run([FLANG, "-O2", "-module-dir", mod, "-I", mod, *srcs, "-o", exe + ".tmp",
              "-L", LIB_DIR, "-lpfamd", f"-Wl,-rpath,{LIB_DIR}", "-Wl,-rpath,$ORIGIN/../lib"],
             verbose)
        os.replace(exe + ".tmp", exe)
        exes.append(exe)
    return exes[0] if exes else None


def build(verbose=False, fortran=True):
    lib = build_lib(verbose)
    exe = build_fortran(verbose) if fortran else None
    return lib, exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-fortran", action="store_true")
    ap.add_argument("--resource-usage", action="store_true",
                    help="print per-kernel VGPR/LDS/occupancy from the compiler")
    a = ap.parse_args(argv)
    if a.resource_usage:
        build_lib(True, extra=["-Rpass-analysis=kernel-resource-usage"])
    lib, exe = build(a.verbose, not a.no_fortran)
    print(lib)
    if exe:
        print(exe)


if __name__ == "__main__":
    sys.exit(main())
_ = shutil
