"""MI355X-native replica-exchange Monte Carlo docking core (GeauxDock REMC).

The product is the C ABI of libgdremc.so (include/gd_remc.h): HIP kernels
for gfx950 behind a drop-in for the reference's Run().  `remc` marshals
numpy arrays into that ABI; `synth` builds the seeded benchmark complexes.
"""
from . import remc  # noqa: F401

__all__ = ["remc"]
