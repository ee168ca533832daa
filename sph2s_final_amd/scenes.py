"""Benchmark scenes: seeded dam-break lattices with fixed physical constants.

Shared constants: h = 0.0457, lattice spacing s = h/2, rho0 = 1000,
m = rho0 s^3, stiffness k = 3, viscosity mu = 3.5, tension 0,
g = (0, -9.81, 0), dt = 1e-3, wall clamp with restitution 0.5, grid cell = h.
A lattice of n = (nx, ny, nz) particles sits at (i + 1/2) s from the domain
minimum with SplitMix64 jitter U[-0.01 s, 0.01 s) per axis; the domain is
[0, n_d / 32] on each axis, so the 32^3 lattice fills [0, 0.731]^3 of the unit
box and larger lattices keep that block-to-domain ratio.

Named workloads:
  dam_break_32k  32^3 = 32,768 particles, domain [0, 1]^3, seed 1
  dam_break_1m   100^3 = 1,000,000 particles, domain [0, 3.125]^3, seed 2
  two_scale_250k 63^3 = 250,047 low-resolution particles, domain [0, 1.969]^3,
                 seed 3, high-resolution region x in [0.6 L, L], y in [0, 0.5 L],
                 z in [0, L] (20% of the domain volume, the far-wall impact zone)
  walled_dam_break_215k  the reference's own single-scale scene
                 (SPH2S_FINAL_SINGLE_SCALE/main.cpp:161-250): a 42 x 95 x 54
                 fluid block in a tank of boundary particles with a pillar,
                 h from 35 neighbours, k = 30, mu = 5, sigma = 0.8, c = 88.1472,
                 no wall clamp
"""
from __future__ import annotations

from dataclasses import dataclass
from ctypes import c_int32, c_void_p

import numpy as np

from . import _capi as C
from .solver import SolverConfiguration

H = 0.0457
SPACING = H / 2.0
REST_DENSITY = 1000.0
STIFFNESS = 3.0
VISCOSITY = 3.5
DT = 1e-3
JITTER = 0.01
REFERENCE_LATTICE = 32  # lattice edge that fills the unit domain


@dataclass
class Scene:
    name: str
    config: C.SphConfig
    positions: np.ndarray   # (N, 3) float32
    velocities: np.ndarray  # (N, 3) float32
    dt: float = DT

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def particle_mass(spacing: float = SPACING, rest_density: float = REST_DENSITY) -> float:
    return float(np.float32(rest_density * spacing ** 3))


def lattice(n, spacing=SPACING, lo=(0.0, 0.0, 0.0), jitter=JITTER, seed=1) -> np.ndarray:
    n = [int(v) for v in n]
    out = np.empty((n[0] * n[1] * n[2], 3), np.float32)
    C.check(C.lib().sph_scene_lattice((c_int32 * 3)(*n), float(spacing), C.f3(lo), float(jitter),
                                      int(seed), out.ctypes.data_as(c_void_p)))
    return out


def gaussian_velocities(count: int, sigma: float, seed: int) -> np.ndarray:
    out = np.empty((int(count), 3), np.float32)
    C.check(C.lib().sph_scene_gaussian_velocity(int(count), float(sigma), int(seed),
                                                out.ctypes.data_as(c_void_p)))
    return out


def dam_break(n=(32, 32, 32), seed=1, *, stiffness=STIFFNESS, viscosity=VISCOSITY,
              velocity_sigma=0.0, velocity_seed=None, name=None, kahan=True) -> Scene:
    n = tuple(int(v) for v in n)
    hi = tuple(v / REFERENCE_LATTICE for v in n)
    cfg = SolverConfiguration.FromConstants(H, particle_mass(), REST_DENSITY, stiffness,
                                            viscosity, (0.0, 0.0, 0.0), hi, kahan=kahan)
    pos = lattice(n, seed=seed)
    if velocity_sigma > 0.0:
        vseed = (seed * 0x9E3779B1 + 7) if velocity_seed is None else velocity_seed
        vel = gaussian_velocities(pos.shape[0], velocity_sigma, vseed)
    else:
        vel = np.zeros_like(pos)
    return Scene(name or f"dam_break_{n[0]}x{n[1]}x{n[2]}", cfg, pos, vel)


def dam_break_32k() -> Scene:
    return dam_break((32, 32, 32), seed=1, name="dam_break_32k")


def dam_break_1m() -> Scene:
    return dam_break((100, 100, 100), seed=2, name="dam_break_1m")


@dataclass
class WallScene:
    """A fluid block and static boundary particles (boundary-particle walls)."""
    name: str
    config: C.SphConfig
    positions: np.ndarray   # (N, 3) float32 fluid
    velocities: np.ndarray
    walls: np.ndarray       # (M, 3) float32 boundary particles
    dt: float = DT

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def walled_dam_break(fluid_lo=(0.04, 0.04, 0.0), fluid_hi=(0.35, 0.75, 0.4),
                     tank_lo=(0.0, 0.0, -0.1), tank_hi=(1.6, 1.5, 0.5),
                     pillar=((0.7, 0.0, 0.22), (0.75, 1.5, 0.25)), spacing=0.0075, border=5,
                     domain_lo=(-0.1, -0.1, -0.2), domain_hi=(1.7, 1.6, 0.6), tension=0.8,
                     name=None) -> WallScene:
    """initSim of SPH2S_FINAL_SINGLE_SCALE/main.cpp:161-250: the fluid is
    CreateParticleBox over [fluid_lo, fluid_hi) at `spacing`; the walls are the
    Union of CreateParticleBoxCanvas shells (`border` layers) around the tank
    and around the pillar; the configuration is MakeConfiguration over the
    domain with the fluid's grid volume, 35 neighbours, rho0 = 1000, k = 30,
    mu = 5, c = 88.1472 and the tension coefficient.  The defaults are the
    reference's values; smaller boxes give test-sized scenes."""
    from .solver import Grid, ParticleData
    fgrid = Grid.MakeGrid(fluid_lo, fluid_hi, spacing)
    fluid = ParticleData.CreateParticleBox(fgrid)
    shells = [ParticleData.CreateParticleBoxCanvas(Grid.MakeGrid(tank_lo, tank_hi, spacing), border)]
    if pillar is not None:
        shells.append(ParticleData.CreateParticleBoxCanvas(Grid.MakeGrid(pillar[0], pillar[1],
                                                                         spacing), border))
    walls = shells[0]
    for extra in shells[1:]:
        walls = ParticleData.Union(walls, extra)
    volume = Grid.ComputeVolume(fgrid)
    cfg = SolverConfiguration.MakeConfiguration(domain_lo, domain_hi, volume, fluid.NumParticles,
                                                35, 1000.0, 30.0, 5.0, 88.1472, tension)
    pos = fluid.positions
    return WallScene(name or f"walled_dam_break_{pos.shape[0]}", cfg, pos, np.zeros_like(pos),
                     walls.positions)


def walled_dam_break_215k() -> WallScene:
    return walled_dam_break(name="walled_dam_break_215k")


def walled_tank_small() -> WallScene:
    """A test-sized walled scene: 8 x 16 x 8 fluid particles in a tank with a
    pillar, same spacing and constants as the reference scene."""
    return walled_dam_break(fluid_lo=(0.015, 0.015, 0.015), fluid_hi=(0.075, 0.135, 0.075),
                            tank_lo=(0.0, 0.0, 0.0), tank_hi=(0.2, 0.2, 0.09),
                            pillar=((0.12, 0.0, 0.03), (0.135, 0.2, 0.045)), border=3,
                            domain_lo=(-0.05, -0.05, -0.05), domain_hi=(0.25, 0.25, 0.14),
                            name="walled_tank_small")


@dataclass
class TwoScaleScene:
    name: str
    config: C.SphConfig
    positions: np.ndarray
    velocities: np.ndarray
    region_lo: tuple
    region_hi: tuple
    dt: float = DT

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def impact_region(extent):
    """Axis-aligned box over the far-wall impact zone: 20% of the domain volume."""
    L = [float(v) for v in extent]
    return (0.6 * L[0], 0.0, 0.0), (L[0], 0.5 * L[1], L[2])


def two_scale_dam_break(n=63, seed=3, name=None) -> TwoScaleScene:
    sc = dam_break((n, n, n), seed=seed)
    hi = [sc.config.domain_hi[d] for d in range(3)]
    lo_r, hi_r = impact_region(hi)
    return TwoScaleScene(name or f"two_scale_{n}", sc.config, sc.positions, sc.velocities, lo_r,
                         hi_r)


def two_scale_250k() -> TwoScaleScene:
    return two_scale_dam_break(63, seed=3, name="two_scale_250k")


@dataclass
class BatchScenes:
    name: str
    scenes: list          # list[Scene]
    stiffness: np.ndarray
    viscosity: np.ndarray
    dt: float = DT

    @property
    def counts(self) -> np.ndarray:
        return np.array([s.n for s in self.scenes], np.int64)

    @property
    def n(self) -> int:
        return int(self.counts.sum())


def truncated_dam_break(count: int, seed: int, *, stiffness=STIFFNESS, viscosity=VISCOSITY,
                        velocity_sigma=0.0, name=None) -> Scene:
    """A dam break of exactly ``count`` particles: the first ``count`` points of
    an a x a x ceil(count / a^2) lattice, a = ceil(count^(1/3)) (the top z layer
    is partial), in the cubic domain [0, a / 32]^3."""
    count = int(count)
    a = int(np.ceil(round(count ** (1.0 / 3.0), 9)))
    while a ** 3 < count:
        a += 1
    nz = -(-count // (a * a))
    sc = dam_break((a, a, nz), seed=seed, stiffness=stiffness, viscosity=viscosity,
                   velocity_sigma=velocity_sigma, name=name)
    hi = (a / REFERENCE_LATTICE,) * 3
    cfg = SolverConfiguration.FromConstants(H, particle_mass(), REST_DENSITY, stiffness, viscosity,
                                            (0.0, 0.0, 0.0), hi, kahan=True)
    return Scene(sc.name, cfg, np.ascontiguousarray(sc.positions[:count]),
                 np.ascontiguousarray(sc.velocities[:count]))


def batched_dam_breaks(n_scenes=64, seed=4, n_min=4096, n_max=262144, k_range=(1.0, 5.0),
                       mu_range=(1.0, 8.0), name=None) -> BatchScenes:
    """Independent dam breaks of uneven size: N_s = round(exp U[ln n_min,
    ln n_max]), stiffness U[k_range], viscosity U[mu_range], drawn in that order
    per scene from numpy's PCG64 seeded with ``seed``; scene s's lattice jitter
    uses seed 1000 * seed + s."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes, ks, mus = [], [], []
    for s in range(int(n_scenes)):
        cnt = int(round(float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))))
        k = float(np.float32(rng.uniform(*k_range)))
        mu = float(np.float32(rng.uniform(*mu_range)))
        scenes.append(truncated_dam_break(cnt, 1000 * seed + s, stiffness=k, viscosity=mu,
                                          name=f"scene_{s}"))
        ks.append(k)
        mus.append(mu)
    return BatchScenes(name or f"batched_{n_scenes}", scenes, np.array(ks, np.float32),
                       np.array(mus, np.float32))


def batched_64() -> BatchScenes:
    """64 scenes, N_s log-uniform in [4096, 262144], seed 4."""
    return batched_dam_breaks(64, seed=4, name="batched_64")


def rollout_batch(n_scenes=16, shape=(64, 32, 32), seed=5, sigma=0.05, k_range=(1.0, 5.0),
                  mu_range=(1.0, 8.0), name=None) -> BatchScenes:
    """Scenes of the differentiable-rollout task: ``n_scenes`` dam breaks of
    one lattice shape (domain shape / 32), initial velocities N(0, sigma^2) per
    component, and hidden per-scene stiffness and viscosity drawn from numpy's
    PCG64(seed) (stiffness then viscosity per scene).  The configurations carry
    the shared constants' stiffness and viscosity (the starting guess); the
    hidden values are in ``stiffness`` / ``viscosity``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scs, ks, mus = [], [], []
    for s in range(int(n_scenes)):
        ks.append(float(np.float32(rng.uniform(*k_range))))
        mus.append(float(np.float32(rng.uniform(*mu_range))))
        sc = dam_break(shape, seed=1000 * seed + s, name=f"scene_{s}")
        sc.velocities = gaussian_velocities(sc.n, sigma, 7000 + 1000 * seed + s)
        scs.append(sc)
    return BatchScenes(name or f"rollout_{n_scenes}x{shape[0] * shape[1] * shape[2]}", scs,
                       np.array(ks, np.float32), np.array(mus, np.float32))


def rollout_16x64k() -> BatchScenes:
    """16 scenes x 65,536 particles (64 x 32 x 32), seed 5."""
    return rollout_batch(16, (64, 32, 32), seed=5, name="rollout_16x64k")
