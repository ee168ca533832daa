"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the host oracle (liboracle.so).

The oracle is a CPU restatement of the reference SPH2S step (FP32 hashing with
the reference's IEEE operation order, FP64 physics; see sph_oracle.h).  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use it, and
only as the checker / the timed CPU baseline -- never as the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, Structure, c_double, c_float, c_int, c_int32, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OracleGrid(Structure):
    _fields_ = [("origin", c_float * 3), ("dims", c_int32 * 3), ("spacing", c_float)]


class OracleParams(Structure):
    _fields_ = [
        ("h", c_double), ("mass", c_double), ("rest_density", c_double),
        ("stiffness", c_double), ("viscosity", c_double), ("tension", c_double),
        ("gravity", c_double * 3), ("restitution", c_double),
        ("lo", c_double * 3), ("hi", c_double * 3),
        ("wall_clamp", c_int32), ("h_f32", c_float), ("grid", OracleGrid),
        ("xsph", c_double), ("use_xsph", c_int32),
    ]


class OracleWalls(Structure):
    _fields_ = [("pos", c_void_p), ("n", c_uint32), ("mass", c_double),
                ("speed_sound", c_double)]


def walls(pos, mass: float, speed_sound: float):
    """(OracleWalls, keep-alive array) for boundary particle positions (N, 3)."""
    arr = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    return OracleWalls(arr.ctypes.data_as(c_void_p), arr.shape[0], float(mass),
                       float(speed_sound)), arr


class OracleTsOptions(Structure):
    _fields_ = [("coupling", c_uint32), ("split_offsets", c_uint32), ("merge", c_uint32),
                ("blend_increment", c_float), ("velocity_cap", c_float)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = POINTER(OracleParams)
        vp = c_void_p
        L.oracle_hash.argtypes = [POINTER(OracleGrid), vp, c_uint32, vp]
        L.oracle_stable_sort.argtypes = [vp, c_uint32, vp, vp]
        L.oracle_cell_ranges.argtypes = [vp, c_uint32, c_uint32, vp, vp]
        L.oracle_step.argtypes = [P, vp, vp, c_uint32, c_double, vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_rollout.argtypes = [P, vp, vp, c_uint32, c_double, c_uint32, vp, vp, vp]
        W = POINTER(OracleWalls)
        L.oracle_step_walls.argtypes = [P, W, vp, vp, c_uint32, c_double, vp, vp, vp, vp, vp, vp,
                                        vp, vp]
        L.oracle_rollout_walls.argtypes = [P, W, vp, vp, c_uint32, c_double, c_uint32, vp, vp,
                                           vp]
        L.oracle_rollout_rho.argtypes = [P, W, vp, vp, c_uint32, c_double, c_uint32, vp, vp, vp,
                                         vp]
        L.oracle_energy_colors.argtypes = [P, vp, vp, c_uint32, vp, vp]
        L.oracle_surface_indicator.argtypes = [P, vp, vp, c_uint32, vp, vp, c_uint32, vp]
        L.oracle_boundary_weight.argtypes = [c_double, c_double, c_double]
        L.oracle_boundary_weight.restype = c_double
        L.oracle_rollout_loss_grad.argtypes = [P, vp, vp, c_uint32, c_double, c_uint32, vp, vp,
                                               vp, vp, vp, vp]
        L.oracle_rollout_loss_grad.restype = c_double
        L.oracle_density_bruteforce.argtypes = [P, vp, c_uint32, vp]
        L.oracle_accel_bruteforce.argtypes = [P, vp, vp, vp, c_uint32, vp]
        for f in ("oracle_poly6", "oracle_spiky_grad_mag", "oracle_visc_laplacian"):
            getattr(L, f).argtypes = [c_double, c_double]
            getattr(L, f).restype = c_double
        L.cpu_solver_create.argtypes = [P, vp, vp, c_uint32]
        L.cpu_solver_create.restype = vp
        L.cpu_solver_step.argtypes = [vp, c_float, c_uint32]
        L.cpu_solver_get.argtypes = [vp, vp, vp, vp, vp]
        L.cpu_solver_threads.restype = c_int
        L.cpu_solver_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


def params_from_config(cfg) -> OracleParams:
    """Oracle parameters from a library configuration (sph_config)."""
    p = OracleParams()
    p.h = float(cfg.h)
    p.mass = float(cfg.mass)
    p.rest_density = float(cfg.rest_density)
    p.stiffness = float(cfg.stiffness)
    p.viscosity = float(cfg.viscosity)
    p.tension = float(cfg.tension)
    for d in range(3):
        p.gravity[d] = float(cfg.gravity[d])
        p.lo[d] = float(cfg.domain_lo[d])
        p.hi[d] = float(cfg.domain_hi[d])
        p.grid.origin[d] = cfg.grid.origin[d]
        p.grid.dims[d] = cfg.grid.dims[d]
    p.grid.spacing = cfg.grid.spacing
    p.restitution = float(cfg.restitution)
    p.wall_clamp = 1 if (cfg.flags & 1) else 0
    p.h_f32 = cfg.h
    p.xsph = float(getattr(cfg, "xsph", 0.0))
    p.use_xsph = 1 if (cfg.flags & 4) else 0
    return p


def num_cells(p: OracleParams) -> int:
    return int(p.grid.dims[0]) * int(p.grid.dims[1]) * int(p.grid.dims[2])


def hash_positions(p: OracleParams, pos: np.ndarray) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.empty(pos.shape[0], np.uint32)
    lib().oracle_hash(ctypes.byref(p.grid), _p(pos), pos.shape[0], _p(out))
    return out


def stable_sort(keys: np.ndarray):
    keys = np.ascontiguousarray(keys, np.uint32)
    ks = np.empty_like(keys)
    perm = np.empty_like(keys)
    lib().oracle_stable_sort(_p(keys), keys.size, _p(ks), _p(perm))
    return ks, perm


def cell_ranges(sorted_keys: np.ndarray, cells: int):
    sk = np.ascontiguousarray(sorted_keys, np.uint32)
    st = np.empty(cells, np.uint32)
    en = np.empty(cells, np.uint32)
    lib().oracle_cell_ranges(_p(sk), sk.size, cells, _p(st), _p(en))
    return st, en


def step(p: OracleParams, pos: np.ndarray, vel: np.ndarray, dt: float, wall=None) -> dict:
    """One FP64 step; ``wall`` = walls(...)[0] adds boundary particles."""
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    n = pos.shape[0]
    C = num_cells(p)
    out = dict(hash=np.empty(n, np.uint32), perm=np.empty(n, np.uint32),
               cell_start=np.empty(C, np.uint32), cell_end=np.empty(C, np.uint32),
               rho=np.empty(n), acc=np.empty((n, 3)), pos=np.empty((n, 3)), vel=np.empty((n, 3)))
    lib().oracle_step_walls(ctypes.byref(p), None if wall is None else ctypes.byref(wall),
                            _p(pos), _p(vel), n, float(dt), _p(out["hash"]), _p(out["perm"]),
                            _p(out["cell_start"]), _p(out["cell_end"]), _p(out["rho"]),
                            _p(out["acc"]), _p(out["pos"]), _p(out["vel"]))
    return out


def rollout(p: OracleParams, pos: np.ndarray, vel: np.ndarray, dt: float, steps: int,
            wall=None, with_density: bool = False):
    """(x, v, ids) after ``steps`` FP64 steps, in the last step's sorted order;
    with_density adds the last step's densities in the same order."""
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    n = pos.shape[0]
    x = np.empty((n, 3))
    v = np.empty((n, 3))
    ids = np.empty(n, np.uint32)
    rho = np.empty(n)
    lib().oracle_rollout_rho(ctypes.byref(p), None if wall is None else ctypes.byref(wall),
                             _p(pos), _p(vel), n, float(dt), int(steps), _p(x), _p(v),
                             _p(ids), _p(rho))
    return (x, v, ids, rho) if with_density else (x, v, ids)


def energy_colors(p: OracleParams, pos, vel):
    """(colour, |x_i - x_c|) per input particle of the Mexican-hat energy."""
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    n = pos.shape[0]
    col = np.empty(n)
    dist = np.empty(n)
    lib().oracle_energy_colors(ctypes.byref(p), _p(pos), _p(vel), n, _p(col), _p(dist))
    return col, dist


def surface_indicator(p_low: OracleParams, pos_l, rho_l, pos_h, rho_h) -> np.ndarray:
    pos_l = np.ascontiguousarray(pos_l, np.float32)
    pos_h = np.ascontiguousarray(pos_h, np.float32).reshape(-1, 3)
    rho_l = np.ascontiguousarray(rho_l, np.float64)
    rho_h = np.ascontiguousarray(rho_h, np.float64)
    out = np.empty(pos_l.shape[0])
    lib().oracle_surface_indicator(ctypes.byref(p_low), _p(pos_l), _p(rho_l), pos_l.shape[0],
                                   _p(pos_h) if len(pos_h) else None,
                                   _p(rho_h) if len(rho_h) else None, pos_h.shape[0], _p(out))
    return out


def boundary_weight(r: float, h: float, speed_sound: float) -> float:
    return lib().oracle_boundary_weight(float(r), float(h), float(speed_sound))


def rollout_loss_grad(p: OracleParams, pos0, vel0, dt: float, steps: int, target,
                      grads: bool = True) -> dict:
    """FP64 rollout and the adjoint of L = mean((x_T - target)^2) over the 3n
    components (particles matched by input index): returns loss, x_final and,
    with ``grads``, dL/dpos0, dL/dvel0, dL/dstiffness, dL/dviscosity."""
    pos0 = np.ascontiguousarray(pos0, np.float64)
    vel0 = np.ascontiguousarray(vel0, np.float64)
    target = np.ascontiguousarray(target, np.float64)
    n = pos0.shape[0]
    out = dict(x_final=np.empty((n, 3)))
    gk, gm = ctypes.c_double(), ctypes.c_double()
    if grads:
        out["g_pos0"] = np.empty((n, 3))
        out["g_vel0"] = np.empty((n, 3))
    loss = lib().oracle_rollout_loss_grad(
        ctypes.byref(p), _p(pos0), _p(vel0), n, float(dt), int(steps), _p(target),
        _p(out["x_final"]), _p(out["g_pos0"]) if grads else None,
        _p(out["g_vel0"]) if grads else None, ctypes.byref(gk) if grads else None,
        ctypes.byref(gm) if grads else None)
    out["loss"] = float(loss)
    if grads:
        out["g_stiffness"] = gk.value
        out["g_viscosity"] = gm.value
    return out


def density_bruteforce(p: OracleParams, pos: np.ndarray) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    out = np.empty(pos.shape[0])
    lib().oracle_density_bruteforce(ctypes.byref(p), _p(pos), pos.shape[0], _p(out))
    return out


def accel_bruteforce(p: OracleParams, pos, vel, rho) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    out = np.empty((pos.shape[0], 3))
    lib().oracle_accel_bruteforce(ctypes.byref(p), _p(pos), _p(vel), _p(rho), pos.shape[0],
                                  _p(out))
    return out


def poly6(r: float, h: float) -> float:
    return lib().oracle_poly6(r, h)


def spiky_grad_mag(r: float, h: float) -> float:
    return lib().oracle_spiky_grad_mag(r, h)


def visc_laplacian(r: float, h: float) -> float:
    return lib().oracle_visc_laplacian(r, h)


class CpuSolver:
    """FP32 OpenMP CPU step (the bench's CPU baseline)."""

    def __init__(self, p: OracleParams, pos: np.ndarray, vel: np.ndarray):
        self._pos = np.ascontiguousarray(pos, np.float32)
        self._vel = np.ascontiguousarray(vel, np.float32)
        self.n = self._pos.shape[0]
        self._h = lib().cpu_solver_create(ctypes.byref(p), _p(self._pos), _p(self._vel), self.n)

    def step(self, dt: float, steps: int = 1) -> None:
        lib().cpu_solver_step(self._h, float(dt), int(steps))

    def state(self):
        x = np.empty((self.n, 3), np.float32)
        v = np.empty((self.n, 3), np.float32)
        r = np.empty(self.n, np.float32)
        ids = np.empty(self.n, np.uint32)
        lib().cpu_solver_get(self._h, _p(x), _p(v), _p(r), _p(ids))
        return x, v, r, ids

    @staticmethod
    def threads() -> int:
        return int(lib().cpu_solver_threads())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cpu_solver_destroy(self._h)
            self._h = None


class OracleTwoScale:
    """FP64 two-scale step (test infrastructure): same coupling and split/merge
    rules as the HIP two-scale solver."""

    FIELDS = {"pos": (0, np.float64, 3), "vel": (1, np.float64, 3), "id": (2, np.uint32, 1),
              "tag": (3, np.uint32, 1), "rho": (4, np.float64, 1), "acc": (5, np.float64, 3),
              "hash": (6, np.uint32, 1), "perm": (7, np.uint32, 1),
              "blend": (10, np.float32, 1), "state": (11, np.uint32, 1)}
    CURRENT = ("pos", "vel", "id", "tag", "blend", "state")

    def __init__(self, p_low: OracleParams, grid_high, region_lo, region_hi, max_low: int):
        L = lib()
        if not hasattr(L, "_ts_bound"):
            P = POINTER(OracleParams)
            vp = c_void_p
            L.oracle_ts_create.argtypes = [P, POINTER(OracleGrid), vp, vp, c_uint32]
            L.oracle_ts_create.restype = vp
            L.oracle_ts_destroy.argtypes = [vp]
            L.oracle_ts_set_state.argtypes = [vp, vp, vp, vp, c_uint32, vp, vp, vp, vp, c_uint32]
            L.oracle_ts_step.argtypes = [vp, c_double]
            L.oracle_ts_split_merge.argtypes = [vp]
            L.oracle_ts_counts.argtypes = [vp, vp]
            L.oracle_ts_get.argtypes = [vp, c_int, c_int, vp]
            L.oracle_ts_set_options.argtypes = [vp, POINTER(OracleTsOptions)]
            L.oracle_ts_set_blend.argtypes = [vp, c_int, vp, vp]
            L.oracle_ts_mark_split.argtypes = [vp, vp, c_uint32]
            L.oracle_ts_set_walls.argtypes = [vp, POINTER(OracleWalls)]
            L.oracle_ts_plant_mistake.argtypes = [c_int]
            L._ts_bound = True
        g = OracleGrid()
        for d in range(3):
            g.origin[d] = grid_high.origin[d]
            g.dims[d] = grid_high.dims[d]
        g.spacing = grid_high.spacing
        self._grid_high = g
        self._lo = np.ascontiguousarray(region_lo, np.float32)
        self._hi = np.ascontiguousarray(region_hi, np.float32)
        self._h = L.oracle_ts_create(ctypes.byref(p_low), ctypes.byref(g), _p(self._lo),
                                     _p(self._hi), int(max_low))
        self.cells = [num_cells(p_low), int(g.dims[0]) * int(g.dims[1]) * int(g.dims[2])]

    def set_state(self, pos_l, vel_l, ids_l=None, pos_h=None, vel_h=None, ids_h=None,
                  tags_h=None):
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return _p(a)
        nh = 0 if pos_h is None else len(pos_h)
        lib().oracle_ts_set_state(self._h, arr(pos_l, np.float32), arr(vel_l, np.float32),
                                  arr(ids_l, np.uint32), len(pos_l), arr(pos_h, np.float32),
                                  arr(vel_h, np.float32), arr(ids_h, np.uint32),
                                  arr(tags_h, np.uint32), nh)

    def set_options(self, coupling=0, split_offsets=0, merge=1, blend_increment=0.0125,
                    velocity_cap=0.0):
        o = OracleTsOptions(int(coupling), int(split_offsets), int(merge), float(blend_increment),
                            float(velocity_cap))
        lib().oracle_ts_set_options(self._h, ctypes.byref(o))

    def set_blend(self, res: int, lam=None, state=None):
        la = None if lam is None else np.ascontiguousarray(lam, np.float32)
        st = None if state is None else np.ascontiguousarray(state, np.uint32)
        lib().oracle_ts_set_blend(self._h, int(res), None if la is None else _p(la),
                                  None if st is None else _p(st))

    def set_walls(self, pos, mass: float, speed_sound: float):
        w, keep = walls(pos, mass, speed_sound)
        lib().oracle_ts_set_walls(self._h, ctypes.byref(w) if len(keep) else None)

    def mark_split(self, ids):
        ids = np.ascontiguousarray(ids, np.uint32)
        lib().oracle_ts_mark_split(self._h, _p(ids) if len(ids) else None, len(ids))

    @staticmethod
    def plant_mistake(which: int) -> None:
        """Plant one deliberate mistake in the fixed coupling's cross sums
        (oracle_ts_plant_mistake; 0 removes it).  Tests use it to show that
        their checks catch such a mistake."""
        lib().oracle_ts_plant_mistake.argtypes = [c_int]
        lib().oracle_ts_plant_mistake(int(which))

    def step(self, dt: float) -> None:
        lib().oracle_ts_step(self._h, float(dt))

    def split_merge(self) -> None:
        lib().oracle_ts_split_merge(self._h)

    def counts(self):
        c = np.empty(4, np.uint32)
        lib().oracle_ts_counts(self._h, _p(c))
        return tuple(int(v) for v in c)

    def get(self, res: int, name: str) -> np.ndarray:
        c = self.counts()
        if name in ("start", "end"):
            out = np.empty(self.cells[res], np.uint32)
            lib().oracle_ts_get(self._h, res, 8 if name == "start" else 9, _p(out))
            return out
        fid, dt, k = self.FIELDS[name]
        n = c[res] if name in self.CURRENT else c[2 + res]
        out = np.empty((n, k) if k > 1 else n, dt)
        lib().oracle_ts_get(self._h, res, fid, _p(out))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_ts_destroy(self._h)
            self._h = None
