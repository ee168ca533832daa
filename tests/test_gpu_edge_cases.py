"""Edge cases of the single-scale and batched steps on the GPU (through the C ABI).

Empty, tiny, ragged, extremely dense and face-aligned inputs.  Each is checked
against the FP64 oracle with the bars of test_gpu_step (hashes, permutation and
cell ranges bit-exact; density 1e-5 relative; acceleration 1e-5 of max |a*|),
and the smallest ones also against closed forms.  Cells at the x-bin tables'
limit (254 particles binned, 255 not) are checked against the oracle.  A cell
of 66,000 particles would take the oracle minutes, so that case is checked
against a brute-force density of sampled
particles (to 1e-4: its rows are single FP32 runs of 66,000 terms) and the
momentum balance of the pressure pairs.
"""
import numpy as np
import pytest

import oracle as O
from sph2s_final_amd import (FIELD_ACCEL, FIELD_CELL_END, FIELD_CELL_START, FIELD_DENSITY,
                             FIELD_POSITION, FIELD_SORTED_HASH, FIELD_SORTED_INDEX,
                             FIELD_VELOCITY, SPH_DEBUG_ACCEL, SPH_EMPTY_CELL, BatchSolver,
                             ParticleData, Solver, scenes)

pytestmark = pytest.mark.gpu

RHO_TOL = 1e-5
ACC_TOL = 1e-5


def _config():
    """The 8^3 dam break's configuration: domain [0, 0.25]^3, grid cell h."""
    return scenes.dam_break((8, 8, 8), seed=1).config


def _gpu_step(cfg, pos, vel, dt=scenes.DT, max_n=None):
    n = len(pos)
    with Solver(ParticleData(max_n or max(n, 1), pos if n else None, vel if n else None), None,
                cfg) as s:
        s.set_debug(SPH_DEBUG_ACCEL)
        s.Advance(dt)
        s.synchronize()
        assert s.device_errors() == 0
        return dict(hash=s.download(FIELD_SORTED_HASH), perm=s.download(FIELD_SORTED_INDEX),
                    cell_start=s.download(FIELD_CELL_START), cell_end=s.download(FIELD_CELL_END),
                    rho=s.densities(), acc=s.download(FIELD_ACCEL), pos=s.positions(),
                    vel=s.velocities())


def _check(name, got, ref, rho_tol=RHO_TOL, acc_tol=ACC_TOL, pos_tol=1e-6):
    assert np.array_equal(got["hash"], ref["hash"]), name
    assert np.array_equal(got["perm"], ref["perm"]), name
    assert np.array_equal(got["cell_start"], ref["cell_start"]), name
    assert np.array_equal(got["cell_end"], ref["cell_end"]), name
    rho_err = np.max(np.abs(got["rho"] - ref["rho"]) / ref["rho"])
    acc_err = np.max(np.abs(got["acc"] - ref["acc"])) / np.max(np.abs(ref["acc"]))
    pos_err = np.max(np.abs(got["pos"] - ref["pos"]))
    print(f"case {name}: n={len(got['rho'])} rho {rho_err:.3e} acc {acc_err:.3e} "
          f"pos {pos_err:.3e}", flush=True)
    assert rho_err <= rho_tol, name
    assert acc_err <= acc_tol, name
    assert pos_err <= pos_tol, name


def _oracle_step(cfg, pos, vel, dt=scenes.DT):
    return O.step(O.params_from_config(cfg), pos, vel, dt)


def _self_density(cfg):
    """m W_poly6(0) = m 315 / (64 pi h^3), the density of a particle alone."""
    return float(np.float32(cfg.mass)) * 315.0 / (64.0 * np.pi * float(cfg.h) ** 3)


def test_empty_solver_steps_and_stays_empty():
    cfg = _config()
    print("case empty solver", flush=True)
    with Solver(ParticleData(64), None, cfg) as s:
        s.Advance(scenes.DT, 3)
        s.synchronize()
        assert s.device_errors() == 0
        assert s.positions().shape == (0, 3)
        assert s.densities().shape == (0,)
        cs = s.download(FIELD_CELL_START)
        assert np.all(cs == SPH_EMPTY_CELL)


def test_lone_particle_closed_form():
    # No neighbour: the density is the self term, the acceleration is g, and
    # one symplectic Euler step gives v = v0 + dt g, x = x0 + dt v.
    cfg = _config()
    x0 = np.array([[0.11, 0.13, 0.12]], np.float32)
    v0 = np.array([[0.2, -0.1, 0.05]], np.float32)
    got = _gpu_step(cfg, x0, v0)
    ref = _oracle_step(cfg, x0, v0)
    _check("lone particle", got, ref)
    assert abs(got["rho"][0] / _self_density(cfg) - 1.0) < 1e-6
    g = np.array([cfg.gravity[d] for d in range(3)], np.float64)
    assert np.allclose(got["acc"][0, :3], g, rtol=0, atol=1e-6)
    v1 = v0[0].astype(np.float64) + scenes.DT * g
    assert np.allclose(got["vel"][0], v1, rtol=0, atol=1e-6)
    assert np.allclose(got["pos"][0], x0[0] + scenes.DT * v1, rtol=0, atol=1e-7)


def test_pair_is_symmetric_and_conserves_momentum():
    # Two particles at rest 0.4 h apart: equal densities (self term plus one
    # pair term), and the pressure accelerations cancel: a1 + a2 = 2 g.
    cfg = _config()
    h = float(cfg.h)
    x = np.array([[0.10, 0.12, 0.12], [0.10 + 0.4 * h, 0.12, 0.12]], np.float32)
    v = np.zeros_like(x)
    got = _gpu_step(cfg, x, v)
    ref = _oracle_step(cfg, x, v)
    _check("pair", got, ref)
    assert got["rho"][0] == got["rho"][1]
    r = float(x[1, 0] - x[0, 0])
    w = 315.0 / (64.0 * np.pi * h ** 9) * (h * h - r * r) ** 3
    assert abs(got["rho"][0] / (_self_density(cfg) + float(np.float32(cfg.mass)) * w) - 1) < 1e-6
    g = np.array([cfg.gravity[d] for d in range(3)], np.float64)
    a = got["acc"][:, :3].astype(np.float64)
    assert np.max(np.abs(a[0] + a[1] - 2 * g)) <= 1e-6 * np.max(np.abs(a))
    assert abs(a[0, 0]) > 1.0  # the pair does push


@pytest.mark.parametrize("count", [2, 63, 65, 255, 257, 1000, 4097])
def test_ragged_counts_match_oracle(count):
    # Counts that leave partial waves and blocks, from the first `count`
    # points of a jittered lattice (a partial top layer).
    sc = scenes.truncated_dam_break(count, seed=31 + count, velocity_sigma=0.2)
    got = _gpu_step(sc.config, sc.positions, sc.velocities)
    _check(f"ragged {count}", got, _oracle_step(sc.config, sc.positions, sc.velocities))


def test_dense_cluster_overflowing_the_lds_tiles():
    # 3,000 particles inside one cell's interior: every stencil row's run is
    # longer than the density and force LDS tiles, so both passes read those
    # rows from global memory, and every particle queues ~3,000 neighbours.
    cfg = _config()
    h = float(cfg.h)
    rng = np.random.default_rng(5)
    c = np.array([2.5, 2.5, 2.5]) * h  # the centre of cell (2, 2, 2)
    pos = (c + rng.uniform(-0.3 * h, 0.3 * h, size=(3000, 3))).astype(np.float32)
    vel = rng.normal(0.0, 0.1, size=(3000, 3)).astype(np.float32)
    got = _gpu_step(cfg, pos, vel)
    ref = _oracle_step(cfg, pos, vel)
    _check("dense cluster", got, ref)
    cs, ce = got["cell_start"], got["cell_end"]
    assert int(np.max(ce[cs != SPH_EMPTY_CELL] - cs[cs != SPH_EMPTY_CELL])) == 3000


def test_face_aligned_lattice_matches_oracle():
    # Particles exactly on cell faces (x = i h/2 from the origin, no jitter):
    # the hash of each must round like the reference's FP32 division.
    cfg = _config()
    s = float(np.float32(cfg.h)) / 2.0
    i = np.arange(9)
    gx, gy, gz = np.meshgrid(i, i, i, indexing="ij")
    pos = (np.stack([gx, gy, gz], axis=-1).reshape(-1, 3) * s + 0.0).astype(np.float32)
    vel = np.zeros_like(pos)
    got = _gpu_step(cfg, pos, vel)
    _check("face-aligned lattice", got, _oracle_step(cfg, pos, vel))


@pytest.mark.parametrize("count", [254, 255])
def test_cells_at_the_x_bin_limit(count):
    # One cell holding `count` particles spread through it, inside a lattice
    # of spacing h/2 around it: 254 is the most the 8-bit x-bin counts take
    # (the cell is binned and runs through it are trimmed at its bins), 255
    # the fewest that leave it unbinned (walked whole).  Both against the FP64
    # oracle: hashes, permutation and ranges bit-exact; a miscounted or
    # wrapped bin would drop or double neighbours, an O(1) error.  Density and
    # acceleration are held to the usual 1e-5 (measured 7-9e-7).
    cfg = _config()
    h = float(cfg.h)
    rng = np.random.default_rng(100 + count)
    lo = np.array([3.0, 3.0, 3.0]) * h
    full = lo + rng.uniform(0.02 * h, 0.98 * h, size=(count, 3))
    i = np.arange(8)
    gx, gy, gz = np.meshgrid(i, i, i, indexing="ij")
    lat = (np.stack([gx, gy, gz], axis=-1).reshape(-1, 3) + 0.5) * (h / 2) + 1.5 * h
    inside = np.all((lat >= lo) & (lat < lo + h), axis=1)
    pos = np.concatenate([full, lat[~inside]]).astype(np.float32)
    vel = np.zeros_like(pos)
    print(f"case x-bin limit cell count={count} n={len(pos)}", flush=True)
    got = _gpu_step(cfg, pos, vel)
    ref = _oracle_step(cfg, pos, vel)
    cs, ce = got["cell_start"], got["cell_end"]
    occupied = cs != SPH_EMPTY_CELL
    assert int(np.max(ce[occupied] - cs[occupied])) == count
    _check(f"x-bin limit {count}", got, ref)


def test_cell_too_full_for_the_x_bins():
    # 66,000 particles in one cell: more than the x-bin counts hold (254 per
    # cell), so the cell is marked unbinned and its rows are walked whole.  Checked
    # against a brute-force FP64 density of 256 sampled particles (every pair
    # is within h), and, at rest, the pressure pairs' momentum balance:
    # sum_i a_i = N g (the pair terms (A_i + A_j) F(r) x_ij are antisymmetric;
    # the viscosity terms vanish with v = 0).
    cfg = _config()
    h = float(cfg.h)
    rng = np.random.default_rng(9)
    n = 66000
    c = np.array([3.5, 3.5, 3.5]) * h
    pos = (c + rng.uniform(-0.2 * h, 0.2 * h, size=(n, 3))).astype(np.float32)
    vel = np.zeros_like(pos)
    print(f"case unbinned cell n={n}", flush=True)
    with Solver(ParticleData(n, pos, vel), None, cfg) as s:
        s.set_debug(SPH_DEBUG_ACCEL)
        s.Advance(scenes.DT)
        s.synchronize()
        assert s.device_errors() == 0
        ids = s.ids()
        rho = s.download(FIELD_DENSITY)
        acc = s.download(FIELD_ACCEL)[:, :3].astype(np.float64)
    assert np.all(np.isfinite(acc))
    x = pos.astype(np.float64)
    m = float(np.float32(cfg.mass))
    c6 = 315.0 / (64.0 * np.pi * h ** 9)
    pick = rng.choice(n, 256, replace=False)
    # Each row run is summed in plain FP32 in index order (run sums are then
    # Kahan-combined), so a run's relative error can reach its length times
    # the FP32 unit (66,000 x 6e-8 = 4e-3).  Runs at fluid densities hold tens
    # to hundreds of terms and meet 1e-5; this one run of 66,000 is held to 1e-4.
    err = 0.0
    for k in pick:
        i = int(ids[k])
        r2 = np.sum((x - x[i]) ** 2, axis=1)
        ref = m * c6 * np.sum(np.where(r2 < h * h, (h * h - r2) ** 3, 0.0))
        err = max(err, abs(rho[k] / ref - 1.0))
    print(f"  density vs brute force: max rel err {err:.3e}", flush=True)
    assert err <= 1e-4
    g = np.array([cfg.gravity[d] for d in range(3)], np.float64)
    imbalance = np.abs(acc.sum(axis=0) - n * g) / np.abs(acc).sum(axis=0).max()
    print(f"  momentum imbalance {imbalance.max():.3e}", flush=True)
    assert imbalance.max() <= 1e-4


def test_ragged_batch_with_empty_and_single_particle_scenes():
    # Scenes of 0, 1, 65 and 700 particles in one batch: each non-empty scene
    # steps bit for bit like its own solver, and the empty one stays empty.
    base = [scenes.truncated_dam_break(k, seed=40 + k, velocity_sigma=0.2) for k in (1, 65, 700)]
    empty = scenes.Scene("empty", _config(), np.zeros((0, 3), np.float32),
                         np.zeros((0, 3), np.float32))
    order = [base[1], empty, base[0], base[2]]
    print(f"case ragged batch counts={[s.n for s in order]}", flush=True)
    with BatchSolver([s.config for s in order], [s.positions for s in order],
                     [s.velocities for s in order]) as b:
        b.Advance(scenes.DT, 5)
        b.synchronize()
        assert b.device_errors() == 0
        pos, vel, rho = (b.download(f) for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_DENSITY))
        for k, sc in enumerate(order):
            sl = b.scene_slice(k)
            if sc.n == 0:
                assert sl.stop - sl.start == 0
                continue
            with Solver(ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config) as one:
                one.Advance(scenes.DT, 5)
                one.synchronize()
                assert np.array_equal(pos[sl], one.positions()), k
                assert np.array_equal(vel[sl], one.velocities()), k
                assert np.array_equal(rho[sl], one.densities()), k
