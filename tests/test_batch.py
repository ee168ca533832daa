"""Batched independent scenes: one segmented hash/sort and one launch per pass
for every scene.

Bars:
  * each scene of a batch equals a single-scene Solver run of the same scene
    bit for bit (sorted hashes and permutation after removing the scene's cell
    and particle offsets, cell ranges, density, acceleration, positions and
    velocities), after one step and after a graph-replayed rollout: per-scene
    reduction order is the same as the single-scene path's;
  * each scene's first step matches the FP64 oracle to the single-scale bars
    (integers bit-exact, density 1e-5 relative, acceleration 1e-5 of max |a|).
CPU tests pin the scene generator (counts, parameter ranges, determinism).
"""
import numpy as np
import pytest

import oracle as O
from sph2s_final_amd import (FIELD_ACCEL, FIELD_CELL_END, FIELD_CELL_START, FIELD_DENSITY,
                             FIELD_POSITION, FIELD_SORTED_HASH, FIELD_SORTED_INDEX, FIELD_TAG,
                             FIELD_VELOCITY, SPH_DEBUG_ACCEL, SPH_EMPTY_CELL, BatchSolver,
                             ParticleData, Solver, scenes)

EMPTY = SPH_EMPTY_CELL


# ---------------------------------------------------------------- CPU: scenes
def test_batched_scene_generator_ranges_and_determinism():
    a = scenes.batched_dam_breaks(16, seed=11, n_min=512, n_max=8192)
    b = scenes.batched_dam_breaks(16, seed=11, n_min=512, n_max=8192)
    assert np.array_equal(a.counts, b.counts)
    assert np.array_equal(a.stiffness, b.stiffness) and np.array_equal(a.viscosity, b.viscosity)
    assert a.counts.min() >= 512 and a.counts.max() <= 8192
    assert np.all((a.stiffness >= 1) & (a.stiffness <= 5))
    assert np.all((a.viscosity >= 1) & (a.viscosity <= 8))
    for s, sc in enumerate(a.scenes):
        assert np.array_equal(sc.positions, b.scenes[s].positions)
        assert sc.config.stiffness == a.stiffness[s] and sc.config.viscosity == a.viscosity[s]
        hi = np.array([sc.config.domain_hi[d] for d in range(3)], np.float32)
        assert np.all(sc.positions > 0) and np.all(sc.positions < hi)


def test_truncated_dam_break_exact_count():
    for cnt in (4096, 4097, 5000, 262144):
        sc = scenes.truncated_dam_break(cnt, seed=1)
        assert sc.n == cnt
        a = int(round(sc.config.domain_hi[0] * scenes.REFERENCE_LATTICE))
        assert a ** 3 >= cnt > (a - 1) ** 3


def test_batched_64_matches_spec():
    b = scenes.batched_64()
    assert len(b.scenes) == 64
    assert b.counts.min() >= 4096 and b.counts.max() <= 262144
    assert 3.0e6 < b.n < 5.0e6  # E[sum N_s] = 64 (262144 - 4096) / ln 64 = 3.97e6


# ---------------------------------------------------------------- GPU
def _small_batch(n_scenes=6, seed=7, sigma=0.3):
    bs = scenes.batched_dam_breaks(n_scenes, seed=seed, n_min=512, n_max=20000)
    for s, sc in enumerate(bs.scenes):
        sc.velocities = scenes.gaussian_velocities(sc.n, sigma, 100 + s)
    return bs


def _batch(bs, debug=True):
    b = BatchSolver([sc.config for sc in bs.scenes], [sc.positions for sc in bs.scenes],
                    [sc.velocities for sc in bs.scenes])
    if debug:
        b.set_debug(SPH_DEBUG_ACCEL)
    return b


def _single(sc, debug=True):
    s = Solver(ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config)
    if debug:
        s.set_debug(SPH_DEBUG_ACCEL)
    return s


def _scene_fields(b, s, fields):
    sl, cl = b.scene_slice(s), b.cell_slice(s)
    p0, c0 = np.uint32(b.part_offsets[s]), np.uint32(b.cell_offsets[s])
    out = {}
    for name, f in fields.items():
        v = b.download(f)
        if f in (FIELD_CELL_START, FIELD_CELL_END):
            v = v[cl].copy()
            m = v != EMPTY
            v[m] -= p0
        else:
            v = v[sl].copy()
            if f == FIELD_SORTED_HASH:
                v -= c0
            elif f == FIELD_SORTED_INDEX:
                v -= p0
        out[name] = v
    return out


STEP_FIELDS = dict(hash=FIELD_SORTED_HASH, perm=FIELD_SORTED_INDEX, cell_start=FIELD_CELL_START,
                   cell_end=FIELD_CELL_END, rho=FIELD_DENSITY, acc=FIELD_ACCEL,
                   pos=FIELD_POSITION, vel=FIELD_VELOCITY)


@pytest.mark.gpu
def test_batch_equals_single_scene_runs():
    bs = _small_batch()
    print(f"case batch_equals_single counts={bs.counts.tolist()}", flush=True)
    with _batch(bs) as b:
        b.Advance(bs.dt)
        b.synchronize()
        assert b.device_errors() == 0
        tags = b.download(FIELD_TAG)
        for s, sc in enumerate(bs.scenes):
            assert np.all(tags[b.scene_slice(s)] == s)
            got = _scene_fields(b, s, STEP_FIELDS)
            with _single(sc) as one:
                one.Advance(sc.dt)
                one.synchronize()
                for name, f in STEP_FIELDS.items():
                    ref = one.download(f)
                    assert np.array_equal(got[name], ref), (s, name)
        # Graph-replayed rollout: still bit-identical per scene.
        b.Advance(bs.dt, 19)
        b.synchronize()
        assert b.device_errors() == 0
        for s, sc in enumerate(bs.scenes):
            got = _scene_fields(b, s, dict(pos=FIELD_POSITION, vel=FIELD_VELOCITY,
                                           rho=FIELD_DENSITY))
            with _single(sc, debug=False) as one:
                one.Advance(sc.dt, 20)
                one.synchronize()
                assert np.array_equal(got["pos"], one.positions()), s
                assert np.array_equal(got["vel"], one.velocities()), s
                assert np.array_equal(got["rho"], one.densities()), s


@pytest.mark.gpu
def test_batch_first_step_matches_oracle():
    bs = _small_batch(n_scenes=4, seed=9)
    print(f"case batch_oracle counts={bs.counts.tolist()}", flush=True)
    with _batch(bs) as b:
        b.advance_eager(bs.dt)
        b.synchronize()
        for s, sc in enumerate(bs.scenes):
            got = _scene_fields(b, s, STEP_FIELDS)
            ref = O.step(O.params_from_config(sc.config), sc.positions, sc.velocities, sc.dt)
            for k in ("hash", "perm", "cell_start", "cell_end"):
                assert np.array_equal(got[k], ref[k]), (s, k)
            rho_err = np.max(np.abs(got["rho"] - ref["rho"]) / ref["rho"])
            acc_err = np.max(np.abs(got["acc"] - ref["acc"])) / np.max(np.abs(ref["acc"]))
            print(f"  scene {s} n={sc.n}: rho {rho_err:.2e} acc {acc_err:.2e}", flush=True)
            assert rho_err <= 1e-5 and acc_err <= 1e-5


@pytest.mark.gpu
def test_batch_64_scenes_graph_equals_eager_and_single_runs():
    bs = scenes.batched_64()
    print(f"case batched_64 n={bs.n}", flush=True)
    with _batch(bs, debug=False) as a, _batch(bs, debug=False) as e:
        a.Advance(bs.dt, 5)
        e.advance_eager(bs.dt, 5)
        a.synchronize()
        e.synchronize()
        assert a.device_errors() == 0 and e.device_errors() == 0
        pa, pe = a.download(FIELD_POSITION), e.download(FIELD_POSITION)
        assert np.all(np.isfinite(pa))
        assert np.array_equal(pa, pe)
        order = np.argsort(bs.counts)
        for s in (int(order[0]), int(order[-1])):
            sc = bs.scenes[s]
            with _single(sc, debug=False) as one:
                one.Advance(sc.dt, 5)
                one.synchronize()
                assert np.array_equal(pa[a.scene_slice(s)], one.positions()), s


@pytest.mark.gpu
def test_batch_particles_outside_their_grids_match_oracle():
    # Each scene gets a few particles outside its own grid (below it, and a
    # close pair beyond its top corner); they are hashed into the scene's
    # boundary cells and must keep their neighbours and self term, as in the
    # oracle's unpruned stencil.
    bs = _small_batch(n_scenes=3, seed=11)
    for sc in bs.scenes:
        cfg = sc.config
        o = np.array([cfg.grid.origin[d] for d in range(3)], np.float64)
        ext = np.array([cfg.grid.dims[d] for d in range(3)], np.float64) * cfg.grid.spacing
        h = cfg.h
        pos = sc.positions.astype(np.float64)
        base = pos[np.argmin(pos[:, 1])]
        pos[:4] = [base + [0.0, -(base[1] - o[1]) - 0.4 * h, 0.0],
                   base + [0.1 * h, -(base[1] - o[1]) - 2.2 * h, 0.0],
                   o + ext + [0.3 * h, 0.2 * h, 0.1 * h],
                   o + ext + [0.6 * h, 0.3 * h, 0.3 * h]]
        sc.positions = pos.astype(np.float32)
    print(f"case batch_outside counts={bs.counts.tolist()}", flush=True)
    with _batch(bs) as b:
        b.advance_eager(bs.dt)
        b.synchronize()
        assert b.device_errors() == 0
        for s, sc in enumerate(bs.scenes):
            got = _scene_fields(b, s, STEP_FIELDS)
            ref = O.step(O.params_from_config(sc.config), sc.positions, sc.velocities, sc.dt)
            for k in ("hash", "perm", "cell_start", "cell_end"):
                assert np.array_equal(got[k], ref[k]), (s, k)
            rho_err = np.max(np.abs(got["rho"] - ref["rho"]) / ref["rho"])
            acc_err = np.max(np.abs(got["acc"] - ref["acc"])) / np.max(np.abs(ref["acc"]))
            print(f"  scene {s} n={sc.n}: rho {rho_err:.2e} acc {acc_err:.2e}", flush=True)
            assert rho_err <= 1e-5 and acc_err <= 1e-5
