"""MI355X-native walker propagation for DMCBS (diffusion Monte Carlo, BCS pairing).

The device library libdmc_gpu.so (HIP kernels for gfx950 behind the C-ABI of
include/dmc_gpu.h) does the work; this package is a thin host-side binding whose classes
mirror the reference's Fortran module interfaces (step, stack, wavefunction).  The modules
``observables`` (g(r), density profile, S(q), n(k): gofr, sofq, pdistribution) and ``deck``
(the positional input deck and the walker file fort.9) complete the host side.
"""
from .lib import (DMC_FP64, DMC_MIXED, DMC_ALLPARTICLE, DMC_SINGLEPARTICLE, DMC_VMC,
                  DMC_METROPOLIS, DMCError, load)
from .api import Params, Engine, Batch

__all__ = ["Params", "Engine", "Batch", "DMCError", "load", "DMC_FP64", "DMC_MIXED",
           "DMC_ALLPARTICLE", "DMC_SINGLEPARTICLE", "DMC_VMC", "DMC_METROPOLIS"]
