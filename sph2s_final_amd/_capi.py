"""ctypes binding of include/sph2s.h (libsph2s.so).

The shared library is loaded from this package directory only; there is no
CPU fallback: if it is missing, every call raises ``LibraryNotBuilt``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_float, c_int, c_int32, c_size_t,
                    c_uint32, c_uint64, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
# SPH2S_LIBRARY may name another build of the same library (e.g. a tuning
# variant compiled with different -D options); default is the in-tree build.
LIB_PATH = os.environ.get("SPH2S_LIBRARY") or os.path.join(HERE, "libsph2s.so")

SPH_OK = 0
SPH_ERR_INVALID_ARGUMENT = 1
SPH_ERR_HIP = 2
SPH_ERR_CAPACITY = 3
SPH_ERR_NONFINITE = 4
SPH_ERR_UNSUPPORTED = 5
SPH_ERR_NO_DEVICE = 6
SPH_ERR_DEVICE = 7

SPH_EMPTY_CELL = 0xFFFFFFFF
SPH_FLAG_WALL_CLAMP = 1
SPH_FLAG_KAHAN = 2
SPH_DEBUG_ACCEL = 1
SPH_DEBUG_COLOR = 2
SPH_FLAG_XSPH = 4

FIELD_POSITION = 0
FIELD_VELOCITY = 1
FIELD_DENSITY = 2
FIELD_ID = 3
FIELD_SORTED_HASH = 4
FIELD_SORTED_INDEX = 5
FIELD_CELL_START = 6
FIELD_CELL_END = 7
FIELD_ACCEL = 8
FIELD_TAG = 9
FIELD_BLEND = 10
FIELD_STATE = 11
FIELD_COLOR = 12

SPH2_COUPLING_FIXED = 0
SPH2_COUPLING_BLENDED = 1
SPH2_SPLIT_LATTICE = 0
SPH2_SPLIT_REFERENCE = 1

SPH_DEVERR_SORT_WAIT = 1
SPH_DEVERR_CAPACITY = 2
SPH_DEVERR_NONFINITE = 4

SPH_RENDER_COLORS = 1

PASS_NAMES = ["hash_reset", "radix_sort", "reorder_cells", "density", "force_integrate"]


class LibraryNotBuilt(RuntimeError):
    """libsph2s.so is absent: run ``__graft_entry__.build()`` first."""


class SPHError(RuntimeError):
    """A non-zero status from the C ABI (message from sph_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[sph status {code}] {msg}")
        self.code = code


class SphGrid(Structure):
    _fields_ = [("origin", c_float * 3), ("dims", c_int32 * 3), ("spacing", c_float)]


class SphConfig(Structure):
    _fields_ = [
        ("h", c_float), ("mass", c_float), ("rest_density", c_float),
        ("stiffness", c_float), ("viscosity", c_float), ("tension", c_float),
        ("gravity", c_float * 3), ("restitution", c_float),
        ("domain_lo", c_float * 3), ("domain_hi", c_float * 3),
        ("grid", SphGrid), ("flags", c_uint32),
        ("speed_sound", c_float), ("boundary_mass", c_float), ("xsph", c_float),
    ]


class SphViews(Structure):
    _fields_ = [
        ("pos4", c_void_p), ("vel4", c_void_p), ("density", c_void_p),
        ("sorted_hash", c_void_p), ("sorted_index", c_void_p),
        ("cell_start", c_void_p), ("cell_end", c_void_p), ("accel4", c_void_p),
        ("n", c_uint32), ("num_cells", c_uint32), ("stream", c_void_p),
    ]


class SphRenderBuffers(Structure):
    _fields_ = [("positions", c_void_p), ("color_values", c_void_p), ("states", c_void_p),
                ("n", c_uint32), ("max_n", c_uint32)]


class Sph2Options(Structure):
    _fields_ = [("coupling", c_uint32), ("split_offsets", c_uint32), ("merge", c_uint32),
                ("blend_increment", c_float), ("velocity_cap", c_float)]


# (name, restype, argtypes) of every exported entry point of include/sph2s.h.
SIGNATURES = [
    ("sph_make_grid", c_int, [POINTER(c_float), POINTER(c_float), c_float, POINTER(SphGrid)]),
    ("sph_grid_num_cells", c_uint32, [POINTER(SphGrid)]),
    ("sph_config_init", c_int, [POINTER(SphConfig), c_float, c_float, c_float, c_float, c_float,
                                POINTER(c_float), POINTER(c_float)]),
    ("sph_make_configuration", c_int, [POINTER(c_float), POINTER(c_float), c_float, c_uint32,
                                       c_uint32, c_float, c_float, c_float, c_float, c_float,
                                       POINTER(SphConfig)]),
    ("sph_scene_lattice", c_int, [POINTER(c_int32), c_float, POINTER(c_float), c_float, c_uint64,
                                  c_void_p]),
    ("sph_scene_gaussian_velocity", c_int, [c_uint32, c_float, c_uint64, c_void_p]),
    ("sph_create", c_int, [POINTER(SphConfig), c_void_p, c_void_p, c_uint32, c_uint32,
                           POINTER(c_void_p)]),
    ("sph_destroy", c_int, [c_void_p]),
    ("sph_advance", c_int, [c_void_p, c_float, c_uint32]),
    ("sph_advance_eager", c_int, [c_void_p, c_float, c_uint32]),
    ("sph_synchronize", c_int, [c_void_p]),
    ("sph_set_state", c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
    ("sph_set_boundary", c_int, [c_void_p, c_void_p, c_uint32]),
    ("sph_num_particles", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph_get_config", c_int, [c_void_p, POINTER(SphConfig)]),
    ("sph_get_views", c_int, [c_void_p, POINTER(SphViews)]),
    ("sph_download", c_int, [c_void_p, c_int, c_void_p, c_size_t]),
    ("sph_export_positions_aos", c_int, [c_void_p, c_void_p, c_int]),
    ("sph_set_debug", c_int, [c_void_p, c_uint32]),
    ("sph_device_errors", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph_count_pairs", c_int, [c_void_p, POINTER(c_uint64)]),
    ("sph_profile_passes", c_int, [c_void_p, c_float, c_uint32, POINTER(c_float)]),
    ("sph_autotune", c_int, [c_void_p, c_float, c_uint32, POINTER(c_uint32), POINTER(c_float)]),
    ("sph_set_tune", c_int, [c_void_p, c_uint32]),
    ("sph_get_tune", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph2_create", c_int, [POINTER(SphConfig), POINTER(c_float), POINTER(c_float), c_void_p,
                            c_void_p, c_uint32, c_uint32, c_uint32, POINTER(c_void_p)]),
    ("sph2_destroy", c_int, [c_void_p]),
    ("sph2_set_state", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_uint32]),
    ("sph2_advance", c_int, [c_void_p, c_float, c_uint32]),
    ("sph2_advance_eager", c_int, [c_void_p, c_float, c_uint32]),
    ("sph2_apply_split_merge", c_int, [c_void_p]),
    ("sph2_synchronize", c_int, [c_void_p]),
    ("sph2_set_debug", c_int, [c_void_p, c_uint32]),
    ("sph2_counts", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph2_device_errors", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph2_download", c_int, [c_void_p, c_int, c_int, c_void_p, c_size_t]),
    ("sph2_set_options", c_int, [c_void_p, POINTER(Sph2Options)]),
    ("sph2_set_boundary", c_int, [c_void_p, c_void_p, c_uint32]),
    ("sph2_get_options", c_int, [c_void_p, POINTER(Sph2Options)]),
    ("sph2_set_tune", c_int, [c_void_p, c_int, c_uint32]),
    ("sph2_get_tune", c_int, [c_void_p, c_int, POINTER(c_uint32)]),
    ("sph2_autotune", c_int, [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_float)]),
    ("sph2_set_blend", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_uint32]),
    ("sph2_mark_split", c_int, [c_void_p, c_void_p, c_uint32]),
    ("sph2_export_render", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("sph2_mark_surface_map", c_int, [c_void_p, c_void_p, c_uint32, c_int]),
    ("sph2_detect_surface", c_int, [c_void_p, c_float, c_int]),
    ("sph_export_render", c_int, [c_void_p, c_void_p, c_void_p]),
    ("sph_enable_render_buffers", c_int, [c_void_p, c_uint32]),
    ("sph_get_render_buffers", c_int, [c_void_p, POINTER(SphRenderBuffers)]),
    ("sph2_enable_render_buffers", c_int, [c_void_p, c_uint32]),
    ("sph2_get_render_buffers", c_int, [c_void_p, c_int, POINTER(SphRenderBuffers)]),
    ("sph_config_set_xsph", c_int, [POINTER(SphConfig), c_float]),
    ("sph_batch_create", c_int, [POINTER(SphConfig), c_uint32, c_void_p, c_void_p, c_void_p,
                                 POINTER(c_void_p)]),
    ("sph_batch_destroy", c_int, [c_void_p]),
    ("sph_batch_advance", c_int, [c_void_p, c_float, c_uint32]),
    ("sph_batch_advance_eager", c_int, [c_void_p, c_float, c_uint32]),
    ("sph_batch_synchronize", c_int, [c_void_p]),
    ("sph_batch_set_debug", c_int, [c_void_p, c_uint32]),
    ("sph_batch_info", c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                               c_void_p, c_void_p]),
    ("sph_batch_device_errors", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph_batch_profile_passes", c_int, [c_void_p, c_float, c_uint32, POINTER(c_float)]),
    ("sph_batch_download", c_int, [c_void_p, c_int, c_void_p, c_size_t]),
    ("sph_batch_autotune", c_int, [c_void_p, c_float, c_uint32, POINTER(c_uint32),
                                   POINTER(c_float)]),
    ("sph_batch_set_tune", c_int, [c_void_p, c_uint32]),
    ("sph_batch_get_tune", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph_rollout_create", c_int, [POINTER(SphConfig), c_uint32, c_void_p, c_float, c_uint32,
                                   c_uint32, POINTER(c_void_p)]),
    ("sph_rollout_destroy", c_int, [c_void_p]),
    ("sph_rollout_info", c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32),
                                 POINTER(c_size_t)]),
    ("sph_rollout_forward", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p]),
    ("sph_rollout_backward", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    ("sph_rollout_synchronize", c_int, [c_void_p]),
    ("sph_rollout_device_errors", c_int, [c_void_p, POINTER(c_uint32)]),
    ("sph_radix_sort_temp_bytes", c_size_t, [c_uint32]),
    ("sph_radix_sort_pairs", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32,
                                     c_void_p, c_size_t, c_void_p]),
    ("sph_radix_sort_pairs_noalloc", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32,
                                             c_uint32, c_void_p]),
    ("sph_last_error", c_char_p, []),
    ("sph_device_count", c_int, []),
    ("sph_version", c_char_p, []),
]

_lib = None


def lib() -> ctypes.CDLL:
    """Load libsph2s.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (there is no CPU fallback)")
        # torch ships its own libamdhip64 with the same soname as the system
        # one.  Loading torch first makes libsph2s.so bind to that already
        # loaded runtime, so the process has a single HIP runtime and device
        # pointers and streams can be shared with torch tensors.  Loading
        # libsph2s.so first would leave torch with a second runtime that sees
        # no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SPH_OK:
        msg = lib().sph_last_error()
        raise SPHError(rc, msg.decode() if msg else "")


def f3(v) -> ctypes.Array:
    return (c_float * 3)(*[float(x) for x in v])
