"""sph2s_final_amd -- MI355X-native SPH2S solver step.

HIP kernels for gfx950 behind the C ABI of include/sph2s.h (libsph2s.so,
built in-tree by ``sph2s_final_amd.build``), with a host-side mirror of the
reference's Solver / ParticleData / SolverConfiguration / Grid interface.
There is no CPU fallback: without the built library every call raises.
"""
from ._capi import (LibraryNotBuilt, SPHError, lib, FIELD_POSITION, FIELD_VELOCITY,  # noqa: F401
                    FIELD_DENSITY, FIELD_ID, FIELD_SORTED_HASH, FIELD_SORTED_INDEX,
                    FIELD_CELL_START, FIELD_CELL_END, FIELD_ACCEL, FIELD_TAG, FIELD_BLEND,
                    FIELD_STATE, FIELD_COLOR, SPH_DEBUG_ACCEL, SPH_DEBUG_COLOR, SPH_FLAG_XSPH,
                    SPH_EMPTY_CELL, SPH_FLAG_KAHAN,
                    SPH_FLAG_WALL_CLAMP, SPH2_COUPLING_FIXED, SPH2_COUPLING_BLENDED,
                    SPH2_SPLIT_LATTICE, SPH2_SPLIT_REFERENCE, SPH_RENDER_COLORS,
                    SPH_DEVERR_SORT_WAIT, SPH_DEVERR_CAPACITY, SPH_DEVERR_NONFINITE)
from . import autotune  # noqa: F401
from .solver import (BatchSolver, Grid, ParticleData, Solver, SolverConfiguration,  # noqa: F401
                     TwoScaleSolver)

__all__ = ["BatchSolver", "Grid", "ParticleData", "Solver", "SolverConfiguration", "TwoScaleSolver",
           "SPHError",
           "LibraryNotBuilt", "lib"]


def device_count() -> int:
    return int(lib().sph_device_count())
