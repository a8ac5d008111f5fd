"""MI355X (gfx950) implementation of ohhelppic3d's per-timestep loop.

The compute path is libohpic.so (hand-written HIP kernels behind the C ABI in
include/ohpic.h); this package holds its ctypes bindings and a host-side
mirror of the reference's mover / interpolator / scatter / boundary /
field-solver interfaces.  A Fortran iso_c_binding mirror lives in fortran/.
"""
from ._lib import (FIELD_EB, FIELD_EX, FIELD_EY, FIELD_EZ, FIELD_PHI, FIELD_RHO, FP64, MIXED,
                   WEIGHTS_REFERENCE, WEIGHTS_STANDARD, OhpicError, available)
from .pic import (Group, LinearInterpolator, LinearScatter, ParticleBoundaries, ParticleMoverBoris,
                  PoissonFieldSolver, SimConfig, Simulation, charge_per_macro_particle,
                  autotune, create_particle_mover, estimate_footprint, fft3d_r2r, mainstep)

__all__ = [
    "FIELD_EB", "FIELD_EX", "FIELD_EY", "FIELD_EZ", "FIELD_PHI", "FIELD_RHO", "FP64", "MIXED",
    "WEIGHTS_REFERENCE", "WEIGHTS_STANDARD", "OhpicError", "available", "Group", "autotune",
    "LinearInterpolator",
    "LinearScatter", "ParticleBoundaries", "ParticleMoverBoris", "PoissonFieldSolver",
    "SimConfig", "Simulation", "charge_per_macro_particle", "create_particle_mover",
    "estimate_footprint", "fft3d_r2r", "mainstep",
]
