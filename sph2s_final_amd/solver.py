"""Host-side mirror of the reference's simulation API over the C ABI.

Names follow the reference (SPH2S_FINAL_SINGLE_SCALE/Solver.h, ParticleData.h,
Grid.h): ``Grid.MakeGrid``, ``SolverConfiguration.MakeConfiguration``,
``ParticleData`` with ``NumParticles``/``MaxParticles``/``dPositions`` and
``Solver(fluid, boundary, configuration).Advance(dt)``.  Errors raise
``SPHError`` the way ``CUDA::SafeCall`` throws std::runtime_error
(SPH2S_FINAL_TWO_SCALE/cuda.h:115-129).
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_float, c_uint32, c_void_p

import numpy as np

from . import _capi as C


def _f32(a, shape_last=3) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if arr.ndim == 1:
        arr = arr.reshape(-1, shape_last)
    return arr


class Grid:
    """struct Grid (SPH2S_FINAL_SINGLE_SCALE/Grid.h:9-59)."""

    @staticmethod
    def MakeGrid(origin, end, spacing: float) -> C.SphGrid:
        g = C.SphGrid()
        C.check(C.lib().sph_make_grid(C.f3(origin), C.f3(end), float(spacing), byref(g)))
        return g

    @staticmethod
    def ComputeNumGridCells(grid: C.SphGrid) -> int:
        return int(C.lib().sph_grid_num_cells(byref(grid)))

    @staticmethod
    def ComputeVolume(grid: C.SphGrid) -> float:
        """Grid.h:48-54: the int cell count times spacing^3, in FP32."""
        n = np.float32(int(grid.dims[0]) * int(grid.dims[1]) * int(grid.dims[2]))
        s = np.float32(grid.spacing)
        return float(np.float32(np.float32(np.float32(n * s) * s) * s))


class SolverConfiguration:
    """SolverConfiguration (SPH2S_FINAL_SINGLE_SCALE/Solver.h:13-61)."""

    @staticmethod
    def MakeConfiguration(origin, end, volume, numParticles, avgParticles, restDensity,
                          bulkModulus, viscosity, speedSound, tensionCoefficient,
                          epsilon: float = 0.0) -> C.SphConfig:
        """Solver.h:15-48: h from the average neighbour count, m = rho0 V / N."""
        del epsilon  # XSPH coefficient: the XSPH integrator is disabled in the reference
        cfg = C.SphConfig()
        C.check(C.lib().sph_make_configuration(
            C.f3(origin), C.f3(end), float(volume), int(numParticles), int(avgParticles),
            float(restDensity), float(bulkModulus), float(viscosity), float(speedSound),
            float(tensionCoefficient), byref(cfg)))
        return cfg

    @staticmethod
    def FromConstants(h, mass, rest_density, stiffness, viscosity, domain_lo, domain_hi, *,
                      tension=0.0, gravity=(0.0, -9.81, 0.0), restitution=0.5,
                      wall_clamp=True, kahan=True) -> C.SphConfig:
        """Explicit constants (the benchmark scenes give h and m directly)."""
        cfg = C.SphConfig()
        C.check(C.lib().sph_config_init(byref(cfg), float(h), float(mass), float(rest_density),
                                        float(stiffness), float(viscosity), C.f3(domain_lo),
                                        C.f3(domain_hi)))
        cfg.tension = float(tension)
        for d in range(3):
            cfg.gravity[d] = float(gravity[d])
        cfg.restitution = float(restitution)
        cfg.flags = (C.SPH_FLAG_WALL_CLAMP if wall_clamp else 0) | (C.SPH_FLAG_KAHAN if kahan else 0)
        return cfg

    @staticmethod
    def SetXSPH(cfg: C.SphConfig, epsilon: float) -> C.SphConfig:
        """XSPHCoeff epsilon; > 0 switches the single-scale step to the XSPH
        integrator (integrateXSPHD, Solver.cu:757-853, per-axis correction)."""
        C.check(C.lib().sph_config_set_xsph(byref(cfg), float(epsilon)))
        return cfg


class ParticleData:
    """ParticleData (SPH2S_FINAL_SINGLE_SCALE/ParticleData.h:9-53).

    Holds host arrays until a Solver takes it; from then on ``dPositions``,
    ``dColorValues`` and ``dStates`` are the device addresses of the renderer's
    buffers that the solver rewrites every step (3 floats, 1 float and 1 int
    per particle; sph_enable_render_buffers) and ``dDensities`` that of the
    last step's densities.
    """

    def __init__(self, maxParticles: int, positions=None, velocities=None):
        self.MaxParticles = int(maxParticles)
        self.positions = None if positions is None else _f32(positions)
        self.velocities = None if velocities is None else _f32(velocities)
        self.NumParticles = 0 if self.positions is None else int(self.positions.shape[0])
        if self.NumParticles > self.MaxParticles:
            raise ValueError("NumParticles exceeds MaxParticles")
        self._solver: "Solver | None" = None

    @staticmethod
    def CreateParticleBox(grid: C.SphGrid) -> "ParticleData":
        """A lattice at origin + (i, j, k) * spacing, x fastest (ParticleData.cpp:141-176)."""
        nx, ny, nz = (int(grid.dims[d]) for d in range(3))
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        o = np.array([grid.origin[d] for d in range(3)], dtype=np.float32)
        sp = np.float32(grid.spacing)
        pos = np.stack([o[0] + i.ravel().astype(np.float32) * sp,
                        o[1] + j.ravel().astype(np.float32) * sp,
                        o[2] + k.ravel().astype(np.float32) * sp], axis=1).astype(np.float32)
        return ParticleData(pos.shape[0], pos)

    @staticmethod
    def CreateParticleBoxCanvas(grid: C.SphGrid, border: int) -> "ParticleData":
        """The shell of `border` lattice layers around a grid's box: every point
        origin + (i, j, k) * spacing with -border <= i < dims + border (and
        likewise j, k) that lies outside the box itself, x fastest
        (ParticleData.cpp:178-209).  The reference builds its walls this way."""
        b = int(border)
        nx, ny, nz = (int(grid.dims[d]) for d in range(3))
        k, j, i = np.meshgrid(np.arange(-b, nz + b), np.arange(-b, ny + b), np.arange(-b, nx + b),
                              indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        shell = ~((i >= 0) & (i < nx) & (j >= 0) & (j < ny) & (k >= 0) & (k < nz))
        o = np.array([grid.origin[d] for d in range(3)], dtype=np.float32)
        sp = np.float32(grid.spacing)
        pos = np.stack([o[0] + i[shell].astype(np.float32) * sp,
                        o[1] + j[shell].astype(np.float32) * sp,
                        o[2] + k[shell].astype(np.float32) * sp], axis=1).astype(np.float32)
        return ParticleData(pos.shape[0], pos)

    @staticmethod
    def Union(a: "ParticleData", b: "ParticleData") -> "ParticleData":
        """Concatenation of two particle sets (ParticleData.cpp:68-139)."""
        if a is None or b is None:
            return None
        pos = np.concatenate([a.host_positions(), b.host_positions()])
        vel = np.concatenate([a.host_velocities(), b.host_velocities()])
        return ParticleData(a.MaxParticles + b.MaxParticles, pos, vel)

    def host_positions(self) -> np.ndarray:
        if self._solver is not None:
            return self._solver.positions()
        return np.zeros((0, 3), np.float32) if self.positions is None else self.positions

    def host_velocities(self) -> np.ndarray:
        if self._solver is not None:
            return self._solver.velocities()
        if self.velocities is None:
            return np.zeros((self.NumParticles, 3), np.float32)
        return self.velocities

    def _render(self) -> C.SphRenderBuffers:
        if self._solver is None:
            raise RuntimeError("ParticleData is not bound to a Solver")
        return self._solver.render_buffers()

    @property
    def dPositions(self) -> int:
        return self._render().positions

    @property
    def dColorValues(self) -> int:
        return self._render().color_values

    @property
    def dStates(self) -> int:
        return self._render().states

    @property
    def dDensities(self) -> int:
        if self._solver is None:
            raise RuntimeError("ParticleData is not bound to a Solver")
        return self._solver.views().density


class Solver:
    """Solver (SPH2S_FINAL_SINGLE_SCALE/Solver.h:109-132) on one MI355X.

    ``boundaryData`` holds the static wall particles (boundary-particle walls,
    sph_set_boundary); it may be None (walls are then only the configuration's
    clamp box, if its flag is set).
    """

    def __init__(self, fluidData: ParticleData, boundaryData=None,
                 configuration: C.SphConfig | None = None):
        if configuration is None:
            raise ValueError("configuration is required")
        self._h = c_void_p()
        pos = fluidData.positions if fluidData.positions is not None else np.zeros((0, 3), np.float32)
        vel = fluidData.velocities
        n = int(pos.shape[0])
        C.check(C.lib().sph_create(byref(configuration),
                                   pos.ctypes.data_as(c_void_p) if n else None,
                                   vel.ctypes.data_as(c_void_p) if vel is not None else None,
                                   n, int(fluidData.MaxParticles), byref(self._h)))
        self.configuration = configuration
        self.fluid = fluidData
        fluidData._solver = self
        if boundaryData is not None and getattr(boundaryData, "NumParticles", 0):
            self.set_boundary(boundaryData.host_positions())
        self._render_on = False

    # -- reference API --
    def Bind(self) -> None:
        """No-op: the configuration travels as a kernel argument, not a global."""

    def Advance(self, dt: float, steps: int = 1) -> None:
        C.check(C.lib().sph_advance(self._h, float(dt), int(steps)))

    # -- extras --
    def set_boundary(self, positions) -> None:
        """Static boundary particles (N, 3), hashed and sorted once on the fluid
        grid; every step adds their boundary acceleration.  An empty array
        removes them."""
        pos = _f32(positions) if positions is not None and len(positions) else None
        n = 0 if pos is None else int(pos.shape[0])
        C.check(C.lib().sph_set_boundary(self._h, None if pos is None else
                                         pos.ctypes.data_as(c_void_p), n))

    def advance_eager(self, dt: float, steps: int = 1) -> None:
        C.check(C.lib().sph_advance_eager(self._h, float(dt), int(steps)))

    def synchronize(self) -> None:
        C.check(C.lib().sph_synchronize(self._h))

    def set_debug(self, flags: int) -> None:
        C.check(C.lib().sph_set_debug(self._h, int(flags)))

    def render_tensors(self, color: bool = False):
        """The renderer's buffers as torch CUDA tensors (DLPack-exchangeable):
        positions (N, 3) and, with ``color`` (needs SPH_DEBUG_COLOR), the
        colour values (N,), in the state's order."""
        import torch
        n = self.NumParticles
        xyz = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        col = torch.empty((n,), dtype=torch.float32, device="cuda") if color else None
        C.check(C.lib().sph_export_render(self._h, c_void_p(xyz.data_ptr()) if n else None,
                                          c_void_p(col.data_ptr()) if (color and n) else None))
        self.synchronize()
        return (xyz, col) if color else xyz

    def enable_render_buffers(self, colors: bool = False) -> None:
        """Solver-owned renderer buffers rewritten by every step
        (sph_enable_render_buffers); ``colors`` runs the colour pass too."""
        C.check(C.lib().sph_enable_render_buffers(self._h, C.SPH_RENDER_COLORS if colors else 0))
        self._render_on = True

    def render_buffers(self) -> C.SphRenderBuffers:
        """Device addresses of the renderer buffers (enabled on first use)."""
        if not self._render_on:
            self.enable_render_buffers()
        r = C.SphRenderBuffers()
        C.check(C.lib().sph_get_render_buffers(self._h, byref(r)))
        return r

    def device_errors(self) -> int:
        """Device-side error bits since the last query (SPH_DEVERR_*); synchronises."""
        bits = c_uint32()
        C.check(C.lib().sph_device_errors(self._h, byref(bits)))
        return int(bits.value)

    def count_pairs(self) -> tuple[int, int]:
        """(candidates, pairs with 0 < r < h) that one neighbour pass of the most
        recent step walked; (0, 0) before the first step.  Synchronises."""
        out = (ctypes.c_uint64 * 2)()
        C.check(C.lib().sph_count_pairs(self._h, out))
        return int(out[0]), int(out[1])

    def profile_passes(self, dt: float, steps: int) -> np.ndarray:
        """Runs ``steps`` eager steps with hipEvents around each pass; returns
        the summed milliseconds per pass (PASS_NAMES order)."""
        ms = (ctypes.c_float * len(C.PASS_NAMES))()
        C.check(C.lib().sph_profile_passes(self._h, float(dt), int(steps), ms))
        return np.array(ms, np.float64)

    def autotune(self, dt: float, reps: int = 5):
        """Times the neighbour-pass launch variants on the current state and
        keeps the fastest (results are bit-identical for every variant).
        Returns (tune word, [density default ms, density tuned ms, force
        default ms, force tuned ms])."""
        t = c_uint32()
        ms = (ctypes.c_float * 4)()
        C.check(C.lib().sph_autotune(self._h, float(dt), int(reps), byref(t), ms))
        return int(t.value), [float(v) for v in ms]

    def set_tune(self, tune: int) -> None:
        C.check(C.lib().sph_set_tune(self._h, int(tune)))

    def get_tune(self) -> int:
        t = c_uint32()
        C.check(C.lib().sph_get_tune(self._h, byref(t)))
        return int(t.value)

    def set_state(self, positions, velocities=None) -> None:
        pos = _f32(positions)
        vel = None if velocities is None else _f32(velocities)
        C.check(C.lib().sph_set_state(self._h, pos.ctypes.data_as(c_void_p),
                                      None if vel is None else vel.ctypes.data_as(c_void_p),
                                      int(pos.shape[0])))
        self.fluid.NumParticles = int(pos.shape[0])

    @property
    def NumParticles(self) -> int:
        n = c_uint32()
        C.check(C.lib().sph_num_particles(self._h, byref(n)))
        return int(n.value)

    def views(self) -> C.SphViews:
        v = C.SphViews()
        C.check(C.lib().sph_get_views(self._h, byref(v)))
        return v

    def download(self, field: int) -> np.ndarray:
        n = self.NumParticles
        cells = int(self.views().num_cells)
        if field in (C.FIELD_POSITION, C.FIELD_VELOCITY, C.FIELD_ACCEL):
            out = np.empty((n, 3), np.float32)
        elif field in (C.FIELD_DENSITY, C.FIELD_COLOR):
            out = np.empty(n, np.float32)
        elif field in (C.FIELD_ID, C.FIELD_SORTED_HASH, C.FIELD_SORTED_INDEX):
            out = np.empty(n, np.uint32)
        elif field in (C.FIELD_CELL_START, C.FIELD_CELL_END):
            out = np.empty(cells, np.uint32)
        else:
            raise ValueError(f"unknown field {field}")
        C.check(C.lib().sph_download(self._h, int(field), out.ctypes.data_as(c_void_p), out.nbytes))
        return out

    def positions(self) -> np.ndarray:
        return self.download(C.FIELD_POSITION)

    def velocities(self) -> np.ndarray:
        return self.download(C.FIELD_VELOCITY)

    def densities(self) -> np.ndarray:
        return self.download(C.FIELD_DENSITY)

    def ids(self) -> np.ndarray:
        return self.download(C.FIELD_ID)

    def close(self) -> None:
        if self._h:
            C.lib().sph_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class TwoScaleSolver:
    """Two-scale Solver (SPH2S_FINAL_TWO_SCALE/Solver.h:109-144): a
    low-resolution set (h0, m0) and a high-resolution set (h0/2, m0/8) coupled
    through the complementary averaged-kernel sums, with 1->8 split inside
    ``region`` and 8->1 merge of sibling groups that leave it.  Resolution 0 is
    the low set, 1 the high set."""

    def __init__(self, configuration: C.SphConfig, region_lo, region_hi, positions,
                 velocities=None, max_low: int | None = None, max_high: int | None = None):
        pos = _f32(positions)
        vel = None if velocities is None else _f32(velocities)
        n = int(pos.shape[0])
        self.max_low = int(max_low if max_low is not None else n)
        self.max_high = int(max_high if max_high is not None else 8 * self.max_low)
        self._h = c_void_p()
        C.check(C.lib().sph2_create(byref(configuration), C.f3(region_lo), C.f3(region_hi),
                                    pos.ctypes.data_as(c_void_p) if n else None,
                                    None if vel is None else vel.ctypes.data_as(c_void_p),
                                    n, self.max_low, self.max_high, byref(self._h)))
        self.configuration = configuration

    def Advance(self, dt: float, steps: int = 1) -> None:
        C.check(C.lib().sph2_advance(self._h, float(dt), int(steps)))

    def advance_eager(self, dt: float, steps: int = 1) -> None:
        C.check(C.lib().sph2_advance_eager(self._h, float(dt), int(steps)))

    def apply_split_merge(self) -> None:
        C.check(C.lib().sph2_apply_split_merge(self._h))

    def set_boundary(self, positions) -> None:
        """Static boundary particles (N, 3), sorted once on the low-resolution
        grid; both sets get their boundary acceleration (h0, m0)."""
        pos = _f32(positions) if positions is not None and len(positions) else None
        n = 0 if pos is None else int(pos.shape[0])
        C.check(C.lib().sph2_set_boundary(self._h, None if pos is None else
                                          pos.ctypes.data_as(c_void_p), n))

    def synchronize(self) -> None:
        C.check(C.lib().sph2_synchronize(self._h))

    def set_debug(self, flags: int) -> None:
        C.check(C.lib().sph2_set_debug(self._h, int(flags)))

    def set_state(self, pos_l, vel_l=None, ids_l=None, pos_h=None, vel_h=None, ids_h=None,
                  tags_h=None) -> None:
        pl = _f32(pos_l)
        vl = None if vel_l is None else _f32(vel_l)
        il = None if ids_l is None else np.ascontiguousarray(ids_l, np.uint32)
        nh = 0 if pos_h is None else len(pos_h)
        ph = None if pos_h is None else _f32(pos_h)
        vh = None if vel_h is None else _f32(vel_h)
        ih = None if ids_h is None else np.ascontiguousarray(ids_h, np.uint32)
        th = None if tags_h is None else np.ascontiguousarray(tags_h, np.uint32)
        ptr = lambda a: None if a is None else a.ctypes.data_as(c_void_p)  # noqa: E731
        C.check(C.lib().sph2_set_state(self._h, ptr(pl), ptr(vl), ptr(il), len(pl), ptr(ph),
                                       ptr(vh), ptr(ih), ptr(th), nh))

    def set_options(self, coupling: int = C.SPH2_COUPLING_FIXED,
                    split_offsets: int = C.SPH2_SPLIT_LATTICE, merge: bool = True,
                    blend_increment: float = 0.0125, velocity_cap: float = 0.0) -> None:
        """Coupling options (sph2_set_options): fixed or blended coupling
        (the reference's lambda blending and fade timeline), lattice or
        reference split offsets, merging on/off, the blend increment per step
        (SolverConfiguration::BlendIncrement) and the velocity cap factor c of
        |v| <= c h / dt (0 = off)."""
        o = C.Sph2Options(int(coupling), int(split_offsets), int(bool(merge)),
                          float(blend_increment), float(velocity_cap))
        C.check(C.lib().sph2_set_options(self._h, byref(o)))

    def get_options(self) -> dict:
        o = C.Sph2Options()
        C.check(C.lib().sph2_get_options(self._h, byref(o)))
        return {"coupling": int(o.coupling), "split_offsets": int(o.split_offsets),
                "merge": bool(o.merge), "blend_increment": float(o.blend_increment),
                "velocity_cap": float(o.velocity_cap)}

    def set_tune(self, resolution: int, word: int) -> None:
        """Tune word of one set's own-set passes (sph2_set_tune)."""
        C.check(C.lib().sph2_set_tune(self._h, int(resolution), int(word)))

    def get_tune(self, resolution: int) -> int:
        w = c_uint32()
        C.check(C.lib().sph2_get_tune(self._h, int(resolution), byref(w)))
        return int(w.value)

    def autotune(self, reps: int = 5):
        """Times both sets' own-set pass variants on the current state and
        applies the fastest (sph2_autotune).  Returns ([low word, high word],
        [per set: density default, tuned, force default, tuned ms])."""
        words = (c_uint32 * 2)()
        ms = (c_float * 8)()
        C.check(C.lib().sph2_autotune(self._h, int(reps), words, ms))
        return [int(words[0]), int(words[1])], [float(x) for x in ms]

    def set_blend(self, resolution: int, lam=None, state=None) -> None:
        """Replace one set's blend weights and/or state words (blended coupling)."""
        n = self.counts()[resolution]
        la = None if lam is None else np.ascontiguousarray(lam, np.float32)
        st = None if state is None else np.ascontiguousarray(state, np.uint32)
        for a in (la, st):
            if a is not None and a.shape != (n,):
                raise ValueError(f"expected {n} entries")
        ptr = lambda a: None if a is None else a.ctypes.data_as(c_void_p)  # noqa: E731
        C.check(C.lib().sph2_set_blend(self._h, int(resolution), ptr(la), ptr(st), n))

    def render_tensors(self, resolution: int, color: bool = False):
        """(positions (N, 3), colour (N,) or None, state words (N,) int32) of one
        set as torch CUDA tensors in the current order (sph2_export_render)."""
        import torch
        n = self.counts()[resolution]
        xyz = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        col = torch.empty((n,), dtype=torch.float32, device="cuda") if color else None
        st = torch.empty((n,), dtype=torch.int32, device="cuda")
        if n:
            C.check(C.lib().sph2_export_render(self._h, int(resolution), c_void_p(xyz.data_ptr()),
                                               c_void_p(col.data_ptr()) if color else None,
                                               c_void_p(st.data_ptr())))
        self.synchronize()
        return xyz, col, st

    def enable_render_buffers(self, colors: bool = False) -> None:
        """Solver-owned renderer buffers of both sets, rewritten by every step
        (sph2_enable_render_buffers)."""
        C.check(C.lib().sph2_enable_render_buffers(self._h,
                                                   C.SPH_RENDER_COLORS if colors else 0))

    def render_buffers(self, resolution: int) -> C.SphRenderBuffers:
        r = C.SphRenderBuffers()
        C.check(C.lib().sph2_get_render_buffers(self._h, int(resolution), byref(r)))
        return r

    def mark_surface_map(self, index_map, split: bool = True) -> None:
        """getSurfaceIDs: ``index_map`` is a CUDA uint32/int32 tensor of
        (low slot + 1) entries, 0 = background (blended coupling)."""
        m = index_map.contiguous()
        C.check(C.lib().sph2_mark_surface_map(self._h, c_void_p(m.data_ptr()), int(m.numel()),
                                              int(bool(split))))

    def detect_surface(self, threshold: float, split: bool = True) -> None:
        """Colour-field surface detector on the last step's state (blended coupling)."""
        C.check(C.lib().sph2_detect_surface(self._h, float(threshold), int(bool(split))))

    def mark_split(self, ids) -> None:
        """Request the split of the low-resolution particles with these ids."""
        ids = np.ascontiguousarray(ids, np.uint32)
        C.check(C.lib().sph2_mark_split(self._h, ids.ctypes.data_as(c_void_p) if len(ids) else None,
                                        len(ids)))

    def counts(self) -> tuple[int, int, int, int]:
        """(low, high) now and (low, high) the last step ran with."""
        c = (c_uint32 * 4)()
        C.check(C.lib().sph2_counts(self._h, c))
        return tuple(int(v) for v in c)

    def device_errors(self) -> int:
        bits = c_uint32()
        C.check(C.lib().sph2_device_errors(self._h, byref(bits)))
        return int(bits.value)

    def download(self, resolution: int, field: int) -> np.ndarray:
        cnt = self.counts()
        n_now, n_step = cnt[resolution], cnt[2 + resolution]
        if field in (C.FIELD_POSITION, C.FIELD_VELOCITY):
            out = np.empty((n_now, 3), np.float32)
        elif field == C.FIELD_ACCEL:
            out = np.empty((n_step, 3), np.float32)
        elif field in (C.FIELD_ID, C.FIELD_TAG, C.FIELD_STATE):
            out = np.empty(n_now, np.uint32)
        elif field in (C.FIELD_BLEND, C.FIELD_COLOR):
            out = np.empty(n_now, np.float32)
        elif field == C.FIELD_DENSITY:
            out = np.empty(n_step, np.float32)
        elif field in (C.FIELD_SORTED_HASH, C.FIELD_SORTED_INDEX):
            out = np.empty(n_step, np.uint32)
        elif field in (C.FIELD_CELL_START, C.FIELD_CELL_END):
            g = self.configuration.grid
            dims = [int(g.dims[d]) for d in range(3)]
            if resolution == 1:
                ext = [self.configuration.domain_hi[d] - g.origin[d] for d in range(3)]
                hh = np.float32(0.5) * np.float32(self.configuration.h)
                dims = [int(np.ceil(np.float32(e) / hh)) for e in ext]
            out = np.empty(dims[0] * dims[1] * dims[2], np.uint32)
        else:
            raise ValueError(f"unknown field {field}")
        C.check(C.lib().sph2_download(self._h, int(resolution), int(field),
                                      out.ctypes.data_as(c_void_p), out.nbytes))
        return out

    def close(self) -> None:
        if self._h:
            C.lib().sph2_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class BatchSolver:
    """Independent single-scale scenes advanced together: one segmented
    hash/sort and one launch per pass for every scene, replayed as a hipGraph.
    ``configurations[s]`` holds scene s's own grid, stiffness, viscosity and
    domain; ``positions[s]`` / ``velocities[s]`` its (N_s, 3) state.  Each
    scene's results equal a ``Solver`` running that scene alone."""

    def __init__(self, configurations, positions, velocities=None):
        cfgs = list(configurations)
        if not cfgs or len(positions) != len(cfgs):
            raise ValueError("one configuration and one position array per scene")
        pos = [_f32(p) for p in positions]
        self.counts = np.array([p.shape[0] for p in pos], np.uint32)
        allpos = np.ascontiguousarray(np.concatenate(pos, axis=0))
        allvel = None
        if velocities is not None:
            allvel = np.ascontiguousarray(np.concatenate([_f32(v) for v in velocities], axis=0))
            if allvel.shape != allpos.shape:
                raise ValueError("velocities must match positions")
        arr = (C.SphConfig * len(cfgs))(*cfgs)
        self._h = c_void_p()
        C.check(C.lib().sph_batch_create(arr, len(cfgs), self.counts.ctypes.data_as(c_void_p),
                                         allpos.ctypes.data_as(c_void_p),
                                         None if allvel is None else allvel.ctypes.data_as(c_void_p),
                                         byref(self._h)))
        self.configurations = cfgs
        ns, n, nc = c_uint32(), c_uint32(), c_uint32()
        self.part_offsets = np.zeros(len(cfgs) + 1, np.uint32)
        self.cell_offsets = np.zeros(len(cfgs) + 1, np.uint32)
        C.check(C.lib().sph_batch_info(self._h, byref(ns), byref(n), byref(nc),
                                       self.part_offsets.ctypes.data_as(c_void_p),
                                       self.cell_offsets.ctypes.data_as(c_void_p)))
        self.n_scenes, self.n, self.num_cells = int(ns.value), int(n.value), int(nc.value)

    def Advance(self, dt: float, steps: int = 1) -> None:
        C.check(C.lib().sph_batch_advance(self._h, float(dt), int(steps)))

    def advance_eager(self, dt: float, steps: int = 1) -> None:
        C.check(C.lib().sph_batch_advance_eager(self._h, float(dt), int(steps)))

    def synchronize(self) -> None:
        C.check(C.lib().sph_batch_synchronize(self._h))

    def set_debug(self, flags: int) -> None:
        C.check(C.lib().sph_batch_set_debug(self._h, int(flags)))

    def device_errors(self) -> int:
        bits = c_uint32()
        C.check(C.lib().sph_batch_device_errors(self._h, byref(bits)))
        return int(bits.value)

    def autotune(self, dt: float, reps: int = 5):
        """As Solver.autotune, over the whole batch."""
        t = c_uint32()
        ms = (ctypes.c_float * 4)()
        C.check(C.lib().sph_batch_autotune(self._h, float(dt), int(reps), byref(t), ms))
        return int(t.value), [float(v) for v in ms]

    def set_tune(self, tune: int) -> None:
        C.check(C.lib().sph_batch_set_tune(self._h, int(tune)))

    def get_tune(self) -> int:
        t = c_uint32()
        C.check(C.lib().sph_batch_get_tune(self._h, byref(t)))
        return int(t.value)

    def profile_passes(self, dt: float, steps: int) -> np.ndarray:
        ms = (ctypes.c_float * len(C.PASS_NAMES))()
        C.check(C.lib().sph_batch_profile_passes(self._h, float(dt), int(steps), ms))
        return np.array(ms, np.float64)

    def download(self, field: int) -> np.ndarray:
        """One field over the whole batch, scene-major (``part_offsets`` /
        ``cell_offsets`` give each scene's slice)."""
        if field in (C.FIELD_POSITION, C.FIELD_VELOCITY, C.FIELD_ACCEL):
            out = np.empty((self.n, 3), np.float32)
        elif field == C.FIELD_DENSITY:
            out = np.empty(self.n, np.float32)
        elif field in (C.FIELD_ID, C.FIELD_TAG, C.FIELD_SORTED_HASH, C.FIELD_SORTED_INDEX):
            out = np.empty(self.n, np.uint32)
        elif field in (C.FIELD_CELL_START, C.FIELD_CELL_END):
            out = np.empty(self.num_cells, np.uint32)
        else:
            raise ValueError(f"unknown field {field}")
        C.check(C.lib().sph_batch_download(self._h, int(field), out.ctypes.data_as(c_void_p),
                                           out.nbytes))
        return out

    def scene_slice(self, s: int) -> slice:
        return slice(int(self.part_offsets[s]), int(self.part_offsets[s + 1]))

    def cell_slice(self, s: int) -> slice:
        return slice(int(self.cell_offsets[s]), int(self.cell_offsets[s + 1]))

    def close(self) -> None:
        if self._h:
            C.lib().sph_batch_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
