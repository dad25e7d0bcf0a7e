"""pfamd -- Python host side of the MI355X coupled phase-field time step.

The product is libpfamd.so (HIP kernels for gfx950 behind the C-ABI in
include/pf_capi.h) and the Fortran driver in fortran/.  This package is the
Python mirror of the driver's host logic (deck reading, scalar set-up, the
time loop around the C-ABI) used by the tests, the bench and the batched /
serving layers.
"""
from . import capi, deck  # noqa: F401
from .capi import PfLoop, PfParams, load  # noqa: F401

__all__ = ["capi", "deck", "PfParams", "PfLoop", "load"]
