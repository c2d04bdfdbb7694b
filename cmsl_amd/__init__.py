"""cmsl_amd -- MI355X-native HZO phase-field time step (TDGL + Jacobi-PCG Poisson).

The compute path is libhzo.so (hand-written HIP for gfx950 behind the C ABI
in include/hzo.h); this package is its host-side mirror of the reference's
setup and step loop.
"""
from . import _lib
from ._lib import HzoError, device_count, lib
from .film import (Batch, Film, FilmInputs, landau_raw, make_film_inputs, nondim, read_landau_table,
                   triangle_wave)

__all__ = ["Batch", "Film", "FilmInputs", "HzoError", "device_count", "landau_raw", "lib", "make_film_inputs",
           "nondim", "read_landau_table", "triangle_wave", "_lib"]
