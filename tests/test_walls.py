"""Boundary-particle walls (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:169-196,
311-344, 426-476, 687-710, 1003-1059; ParticleData.cpp:68-209): static wall
particles sorted once on the fluid grid, and a Monaghan-style boundary
acceleration added to every fluid particle.

CPU tests pin the oracle's boundary term:

* the weight B(q) r / (0.02 c^2) is the piecewise polynomial 2/3, 2q - 3/2 q^2,
  1/2 (2 - q)^2, which is C1 at q = 2/3 and q = 1 and vanishes at q = 2
  (r = h); a dropped or altered coefficient breaks the continuity;
* one fluid particle next to one wall particle accelerates by g plus B(r) times
  the mass ratio, away from the wall;
* the grid walk over the sorted walls equals an all-pairs sum in numpy;
* a particle at the centre of a symmetric tank feels no net boundary force;
* CreateParticleBoxCanvas builds exactly the shell around the box.

GPU tests compare the HIP step with walls with the oracle: sorted hashes,
permutation and cell ranges bit-exact, density within 1e-5 relative,
acceleration (walls included) within 1e-5 of max |a|, positions within 1e-6;
graph replay equals eager launches bit for bit; the two-scale solver with walls
matches the oracle under both couplings.
"""
import numpy as np
import pytest

import oracle as O
from sph2s_final_amd import (FIELD_ACCEL, FIELD_CELL_END, FIELD_CELL_START, FIELD_DENSITY,
                             FIELD_ID, FIELD_POSITION, FIELD_SORTED_HASH, FIELD_SORTED_INDEX,
                             FIELD_VELOCITY, SPH2_COUPLING_BLENDED, SPH_DEBUG_ACCEL, Grid,
                             ParticleData, Solver, TwoScaleSolver, scenes)


def _bw_numpy(r, h, c):
    q = 2.0 * r / h
    base = np.where(q < 2.0 / 3.0, 2.0 / 3.0,
                    np.where(q < 1.0, 2.0 * q - 1.5 * q * q,
                             np.where(q < 2.0, 0.5 * (2.0 - q) ** 2, 0.0)))
    return 0.02 * c * c / r * base


# ---------------------------------------------------------------- CPU: oracle
def test_boundary_weight_is_c1_spline():
    h, c = 0.0152, 88.1472
    k = 0.02 * c * c

    def shape(q):  # B r / (0.02 c^2) as a function of q = 2 r / h
        r = q * h / 2.0
        return O.boundary_weight(r, h, c) * r / k
    eps = 1e-7
    for q0 in (2.0 / 3.0, 1.0):
        left, right = shape(q0 - eps), shape(q0 + eps)
        assert abs(left - right) < 1e-6
        dl = (shape(q0 - eps) - shape(q0 - 2 * eps)) / eps
        dr = (shape(q0 + 2 * eps) - shape(q0 + eps)) / eps
        assert abs(dl - dr) < 1e-4
    assert shape(2.0 - 1e-9) < 1e-15 and shape(2.0) == 0.0 and shape(2.5) == 0.0
    assert abs(shape(0.1) - 2.0 / 3.0) < 1e-15
    qs = np.linspace(0.05, 1.99, 97)
    ref = _bw_numpy(qs * h / 2, h, c)
    got = np.array([O.boundary_weight(q * h / 2, h, c) for q in qs])
    assert np.allclose(got, ref, rtol=1e-14, atol=0)


def _cfg_small():
    return scenes.walled_tank_small().config


def test_single_particle_next_to_one_wall_particle():
    cfg = _cfg_small()
    p = O.params_from_config(cfg)
    h = float(np.float32(cfg.h))
    for d in (0.2 * h, 0.4 * h, 0.7 * h, 0.95 * h):
        fluid = np.array([[0.1, 0.1, 0.05]], np.float32)
        wall = fluid.astype(np.float64) - np.array([d, 0.0, 0.0])
        w, keep = O.walls(wall.astype(np.float32), cfg.boundary_mass, cfg.speed_sound)
        out = O.step(p, fluid, np.zeros_like(fluid), 1e-3, wall=w)
        r = float(fluid[0, 0]) - float(keep[0, 0])
        ratio = float(cfg.boundary_mass) / (float(cfg.mass) + float(cfg.boundary_mass))
        b = O.boundary_weight(r, float(cfg.h), float(cfg.speed_sound)) * ratio
        assert b > 0
        assert np.allclose(out["acc"][0], [b, p.gravity[1], 0.0], rtol=1e-12, atol=1e-9)


def test_wall_grid_walk_equals_all_pairs():
    sc = scenes.walled_tank_small()
    p = O.params_from_config(sc.config)
    rng = np.random.default_rng(3)
    # fluid spread over the tank interior, many within h of a wall
    fluid = (rng.random((600, 3)) * np.array([0.19, 0.19, 0.085])).astype(np.float32)
    vel = np.zeros_like(fluid)
    w, keep = O.walls(sc.walls, sc.config.boundary_mass, sc.config.speed_sound)
    with_w = O.step(p, fluid, vel, 1e-3, wall=w)
    without = O.step(p, fluid, vel, 1e-3)
    b_grid = with_w["acc"] - without["acc"]
    x = fluid[with_w["perm"]].astype(np.float64)
    wx = keep.astype(np.float64)
    d = x[:, None, :] - wx[None, :, :]
    r = np.linalg.norm(d, axis=2)
    h = float(sc.config.h)
    mask = (r < h) & (r > 0)
    rs = np.where(mask, r, 1.0)
    ratio = float(sc.config.boundary_mass) / (float(sc.config.mass) + float(sc.config.boundary_mass))
    wgt = np.where(mask, _bw_numpy(rs, h, float(sc.config.speed_sound)) * ratio / rs, 0.0)
    b_ref = (wgt[:, :, None] * d).sum(axis=1)
    assert np.count_nonzero(np.any(mask, axis=1)) > 100
    scale = np.max(np.abs(b_ref))
    assert np.max(np.abs(b_grid - b_ref)) < 1e-9 * scale


def test_symmetric_tank_centre_feels_no_wall_force():
    s = 0.0075
    g = Grid.MakeGrid((0.0, 0.0, 0.0), (3 * s, 3 * s, 3 * s), s)   # a few cells a side
    walls = ParticleData.CreateParticleBoxCanvas(g, 3).positions
    cfg = _cfg_small()
    p = O.params_from_config(cfg)
    dims = np.array([int(g.dims[d]) for d in range(3)])
    centre = (0.5 * (dims - 1) * np.float32(s)).astype(np.float32)[None, :]
    w, _ = O.walls(walls, cfg.boundary_mass, cfg.speed_sound)
    out = O.step(p, centre, np.zeros_like(centre), 1e-3, wall=w)
    one_sided = O.boundary_weight(s, float(cfg.h), float(cfg.speed_sound))
    assert np.max(np.abs(out["acc"][0] - [0.0, p.gravity[1], 0.0])) < 1e-9 * one_sided


def test_canvas_builder_is_the_shell():
    g = Grid.MakeGrid((0.1, 0.2, 0.3), (0.1 + 5 * 0.01, 0.2 + 4 * 0.01, 0.3 + 3 * 0.01), 0.01)
    dims = [int(g.dims[d]) for d in range(3)]
    b = 2
    pd = ParticleData.CreateParticleBoxCanvas(g, b)
    pos = pd.positions
    total = np.prod([n + 2 * b for n in dims]) - np.prod(dims)
    assert pd.NumParticles == pos.shape[0] == total
    idx = np.rint((pos - np.array([0.1, 0.2, 0.3], np.float32)) / np.float32(0.01)).astype(int)
    inside = np.all((idx >= 0) & (idx < np.array(dims)), axis=1)
    assert not inside.any()
    assert len({tuple(v) for v in idx}) == total
    assert idx.min() == -b and np.all(idx.max(axis=0) == np.array(dims) + b - 1)
    # x fastest
    assert np.all(np.diff(idx[:3, 0]) == 1)


def test_walls_hold_the_oracle_fluid():
    sc = scenes.walled_tank_small()
    p = O.params_from_config(sc.config)
    w, _ = O.walls(sc.walls, sc.config.boundary_mass, sc.config.speed_sound)
    x, v, _ = O.rollout(p, sc.positions, sc.velocities, sc.dt, 60, wall=w)
    assert np.all(np.isfinite(x))
    s = 0.0075
    assert np.all(x > -s) and np.all(x < np.array([0.2, 0.2, 0.09]) + s)


# ---------------------------------------------------------------- GPU parity
def _compare_single(s, out, n):
    assert np.array_equal(s.download(FIELD_SORTED_HASH), out["hash"])
    assert np.array_equal(s.download(FIELD_SORTED_INDEX), out["perm"])
    assert np.array_equal(s.download(FIELD_CELL_START), out["cell_start"])
    assert np.array_equal(s.download(FIELD_CELL_END), out["cell_end"])
    rho = s.download(FIELD_DENSITY)
    acc = s.download(FIELD_ACCEL)
    rho_err = np.max(np.abs(rho - out["rho"]) / out["rho"])
    acc_err = np.max(np.abs(acc - out["acc"])) / np.max(np.abs(out["acc"]))
    dx = np.max(np.abs(s.download(FIELD_POSITION) - out["pos"]))
    print(f"  rho {rho_err:.2e} acc {acc_err:.2e} pos {dx:.2e}", flush=True)
    assert rho_err <= 1e-5 and acc_err <= 1e-5 and dx <= 1e-6


@pytest.mark.gpu
def test_walls_step_parity():
    sc = scenes.walled_tank_small()
    print(f"case walls_step n={sc.n} walls={len(sc.walls)}", flush=True)
    p = O.params_from_config(sc.config)
    w, _ = O.walls(sc.walls, sc.config.boundary_mass, sc.config.speed_sound)
    vel = scenes.gaussian_velocities(sc.n, 0.5, 4)
    fluid = ParticleData(sc.n, sc.positions, vel)
    walls = ParticleData(len(sc.walls), sc.walls)
    s = Solver(fluid, walls, sc.config)
    try:
        s.set_debug(SPH_DEBUG_ACCEL)
        out = O.step(p, sc.positions, vel, sc.dt, wall=w)
        s.Advance(sc.dt)
        s.synchronize()
        assert s.device_errors() == 0
        _compare_single(s, out, sc.n)
        # a state pressed against the walls: the wall term dominates
        x = s.download(FIELD_POSITION)
        xw = x.copy()
        xw[:, 0] = np.float32(0.002) + (x[:, 0] - x[:, 0].min()) * np.float32(0.2)
        s.set_state(xw, np.zeros_like(xw))
        s.set_boundary(sc.walls)
        out = O.step(p, xw, np.zeros_like(xw), sc.dt, wall=w)
        s.Advance(sc.dt)
        s.synchronize()
        _compare_single(s, out, sc.n)
        assert np.max(np.abs(out["acc"][:, 0])) > 100.0
    finally:
        s.close()


@pytest.mark.gpu
def test_walls_graph_matches_eager_and_contains():
    sc = scenes.walled_tank_small()
    print(f"case walls_rollout n={sc.n}", flush=True)
    runs = []
    for eager in (False, True):
        s = Solver(ParticleData(sc.n, sc.positions), ParticleData(len(sc.walls), sc.walls),
                   sc.config)
        (s.advance_eager if eager else s.Advance)(sc.dt, 150)
        s.synchronize()
        assert s.device_errors() == 0
        runs.append((s.download(FIELD_POSITION), s.download(FIELD_VELOCITY), s.download(FIELD_ID)))
        s.close()
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)
    x = runs[0][0]
    s_ = 0.0075
    assert np.all(np.isfinite(x))
    assert np.all(x > -s_) and np.all(x < np.array([0.2, 0.2, 0.09]) + s_)


@pytest.mark.gpu
@pytest.mark.parametrize("coupling", [0, SPH2_COUPLING_BLENDED])
def test_two_scale_walls_parity(coupling):
    sc = scenes.walled_tank_small()
    print(f"case two_scale_walls coupling={coupling} n={sc.n}", flush=True)
    lo, hi = (0.0, 0.0, 0.0), (0.05, 0.2, 0.09)        # refine next to the x = 0 wall
    gh = Grid.MakeGrid([sc.config.grid.origin[d] for d in range(3)],
                       [sc.config.domain_hi[d] for d in range(3)], 0.5 * sc.config.h)
    o = O.OracleTwoScale(O.params_from_config(sc.config), gh, lo, hi, sc.n)
    o.set_state(sc.positions, sc.velocities)
    o.set_walls(sc.walls, sc.config.boundary_mass, sc.config.speed_sound)
    t = TwoScaleSolver(sc.config, lo, hi, sc.positions, sc.velocities, max_low=sc.n,
                       max_high=8 * sc.n)
    try:
        t.set_debug(SPH_DEBUG_ACCEL)
        t.set_boundary(sc.walls)
        if coupling:
            t.set_options(coupling=coupling, blend_increment=0.25, velocity_cap=0.8)
            o.set_options(coupling=1, blend_increment=0.25, velocity_cap=0.8)
        o.split_merge()
        t.apply_split_merge()
        for step in range(3):
            print(f"  step {step}", flush=True)
            o.step(sc.dt)
            t.Advance(sc.dt)
            t.synchronize()
            assert t.device_errors() == 0
            assert t.counts() == o.counts()
            for r in (0, 1):
                if o.counts()[2 + r] == 0:
                    continue
                rho, rho_ref = t.download(r, FIELD_DENSITY), o.get(r, "rho")
                acc, acc_ref = t.download(r, FIELD_ACCEL), o.get(r, "acc")
                rho_err = np.max(np.abs(rho - rho_ref) / rho_ref)
                acc_err = np.max(np.abs(acc - acc_ref)) / np.max(np.abs(acc_ref))
                dx = np.max(np.abs(t.download(r, FIELD_POSITION) - o.get(r, "pos")), initial=0.0)
                print(f"  res {r}: rho {rho_err:.2e} acc {acc_err:.2e} pos {dx:.2e}", flush=True)
                assert rho_err <= 1e-5 and acc_err <= 1e-5 and dx <= 1e-6
                assert np.array_equal(t.download(r, FIELD_ID), o.get(r, "id"))
            # re-seed the oracle from the GPU state so FP32/FP64 drift cannot flip a split
            o.set_state(t.download(0, FIELD_POSITION), t.download(0, FIELD_VELOCITY),
                        t.download(0, FIELD_ID), t.download(1, FIELD_POSITION),
                        t.download(1, FIELD_VELOCITY), t.download(1, FIELD_ID),
                        t.download(1, 9))
            if coupling:
                from sph2s_final_amd import FIELD_BLEND, FIELD_STATE
                for r in (0, 1):
                    o.set_blend(r, t.download(r, FIELD_BLEND), t.download(r, FIELD_STATE))
    finally:
        t.close()
