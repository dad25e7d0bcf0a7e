"""GPU test of pf_diagnosticos_noalloc, the diagnostics entry point that
allocates nothing: called through the C-ABI on device arrays the test
allocates itself (HIP runtime through ctypes), its output equals
pf_diagnosticos of a context holding the same fields -- on three calls in a
row, each on a smaller grid than the one before (the fixed scratch array
keeps the previous call's partial sums).  FP64: bit for bit.  FP32: the
caller's arrays hold full c values, while a context stores c - c_off and
sums (c - c_off) + c_off in FP64; the four phi-based outputs agree bit for
bit and the solute total to 1e-8 relative (the FP32 rounding of each c,
unbiased, summed over the grid)."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available
from pfamd import capi
from pfamd.deck import Deck
from pfamd.run import Simulation

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a gfx950 GPU")]

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    capi.load()  # libpfamd.so pulls in the HIP runtime
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _dev(hip, host):
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), host.nbytes) == 0
    assert hip.hipMemcpy(ptr, host.ctypes.data_as(C.c_void_p), host.nbytes, H2D) == 0
    return ptr


def _deck(n):
    return Deck(Nx=n[0], Ny=n[1], Nz=n[2], tipo_inicial=5, R_semente=5.0, h_quintico=0,
                ruido_tipo=1, AAA=0.0122474, taxa_resfr=0.0, TP=905.0, n_varreduras=2,
                stoptime=0.0)


def test_noalloc_diagnostics_equal_context():
    hip = _hip()
    lib = capi.load()
    for n, prec in (((70, 41, 60), capi.PF_FP64), ((40, 36, 52), capi.PF_FP64),
                    ((24, 22, 30), capi.PF_FP32_MIXED)):
        print(f"[noalloc] {n} precision {prec}", flush=True)
        sim = Simulation(_deck(n), precision=prec, seed=3)
        sim.run(12)
        f = sim.fields()
        ref = sim.diagnostics()
        p = sim.p.copy()
        sim.close()
        dt = np.float64 if prec == capi.PF_FP64 else np.float32
        phi = np.ascontiguousarray(f["phi_new"].astype(dt))  # exact: the stored values
        c = np.ascontiguousarray(f["c_new"].astype(dt))
        dphi, dc = _dev(hip, phi), _dev(hip, c)
        dout = _dev(hip, np.full(5, np.nan))
        try:
            capi.check(lib.pf_diagnosticos_noalloc(C.byref(p), dphi, dc, dout, None), lib)
            assert hip.hipDeviceSynchronize() == 0
            out = np.zeros(5)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dout, out.nbytes, D2H) == 0
        finally:
            for q in (dphi, dc, dout):
                hip.hipFree(q)
        print(f"  {out} vs {ref}", flush=True)
        if prec == capi.PF_FP64:
            assert np.array_equal(out, ref), (n, prec)
        else:
            assert np.array_equal(out[[0, 1, 2, 4]], ref[[0, 1, 2, 4]]), (n, prec)
            assert abs(out[3] - ref[3]) <= 1e-8 * abs(ref[3]), (n, prec)
    # rejected before any device work: a grid below the minimum, null fields
    tiny = _dev(hip, np.zeros(8))
    try:
        bad = p.copy()
        bad.nx = 2
        with pytest.raises(capi.PfError):
            capi.check(lib.pf_diagnosticos_noalloc(C.byref(bad), tiny, tiny, tiny, None), lib)
        with pytest.raises(capi.PfError):
            capi.check(lib.pf_diagnosticos_noalloc(C.byref(p), None, tiny, tiny, None), lib)
    finally:
        hip.hipFree(tiny)
