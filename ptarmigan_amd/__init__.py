"""MI355X-native implementation of Ptarmigan's per-particle Monte Carlo loop.

The transport runs in libptg.so (HIP kernels for gfx950 behind the C ABI in
include/ptg.h); this package is the host-side mirror of the reference's
collide()/Field interface plus the build script. There is no CPU fallback:
importing the engine without a built library raises LibraryMissing.
"""
from ._abi import LibraryMissing  # noqa: F401
from .engine import (Beam, CollideOptions, Engine, Envelope, Fn, Laser, LaserKind, laser_energy,  # noqa: F401
                     Polarization, PtgError, Species, empty_particles)
