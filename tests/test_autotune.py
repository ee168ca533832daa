"""Launch variants and the autotuner.

Every tune word must give bit-identical results (the variants only change
which runs are staged in LDS, when queued pairs are evaluated and which block
handles which particles), so a tuned solver steps exactly like an untuned one.
CPU tests cover the tune-word description and the cache front end.
"""
import json

import numpy as np
import pytest

from sph2s_final_amd import (FIELD_DENSITY, FIELD_POSITION, FIELD_VELOCITY, BatchSolver,
                             ParticleData, SPHError, Solver, autotune, scenes)

WORDS = [0x0000, 0x0001, 0x0002, 0x0003, 0x0004, 0x0005, 0x0010, 0x0020, 0x0030, 0x0040,
         0x0050, 0x0060, 0x0070, 0x0080, 0x0090, 0x1100, 0x2200, 0x3300, 0x4400, 0x4452,
         0x1231, 0x2393, 0x4485, 0x10000, 0x20000, 0x40000, 0x80000, 0x50000, 0xA0000,
         0x91100, 0x64400, 0x100000, 0x400000, 0x500000, 0x501100, 0x424400]


# ---------------------------------------------------------------- CPU
def test_describe_tune_words():
    d = autotune.describe(0)
    assert d == {"density_tile": 1280, "force_tile": 640, "force_list_cap": 16,
                 "force_list_bits": 16, "density_blocks_per_cu": None,
                 "force_blocks_per_cu": None, "density_block_threads": 256,
                 "force_block_threads": 256, "density_particles_per_lane": 1,
                 "force_particles_per_lane": 1}
    d = autotune.describe(0x4432)
    assert d["density_tile"] == 1536
    assert (d["force_tile"], d["force_list_cap"], d["force_list_bits"]) == (768, 8, 32)
    assert d["density_blocks_per_cu"] == 8 and d["force_blocks_per_cu"] == 8
    d = autotune.describe(0x91100)
    assert (d["density_block_threads"], d["force_block_threads"]) == (128, 512)
    assert d["density_blocks_per_cu"] == 1 and d["force_blocks_per_cu"] == 1
    d = autotune.describe(0x501100)
    assert (d["density_particles_per_lane"], d["force_particles_per_lane"]) == (2, 2)
    assert d["density_block_threads"] == 256 and d["density_blocks_per_cu"] == 1


class _FakeSolver:
    def __init__(self):
        self.tuned, self.set = 0, []

    def autotune(self, dt, reps):
        self.tuned += 1
        return 0x1231, [1.0, 0.9, 2.0, 1.5]

    def set_tune(self, w):
        self.set.append(w)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    s = _FakeSolver()
    r1 = autotune.tune(s, "single", 1000, 50, 1e-3, cache_path=path)
    assert not r1["cached"] and r1["tune"] == 0x1231 and s.tuned == 1
    r2 = autotune.tune(s, "single", 1000, 50, 1e-3, cache_path=path)
    assert r2["cached"] and r2["tune"] == 0x1231 and s.tuned == 1 and s.set == [0x1231]
    r3 = autotune.tune(s, "single", 2000, 50, 1e-3, cache_path=path)  # another workload
    assert not r3["cached"] and s.tuned == 2
    with open(path) as f:
        assert len(json.load(f)) == 2


# ---------------------------------------------------------------- GPU
def _run_single(sc, word, steps):
    with Solver(ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config) as s:
        s.set_tune(word)
        assert s.get_tune() == word
        s.Advance(sc.dt, steps)
        s.synchronize()
        assert s.device_errors() == 0
        return s.positions(), s.velocities(), s.densities()


@pytest.mark.gpu
def test_every_variant_is_bit_identical_single():
    sc = scenes.dam_break((40, 30, 36), seed=3, velocity_sigma=0.5)
    print(f"case variants single n={sc.n}", flush=True)
    ref = _run_single(sc, 0, 8)
    for w in WORDS[1:]:
        got = _run_single(sc, w, 8)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), hex(w)


@pytest.mark.gpu
def test_every_variant_is_bit_identical_batched():
    bs = scenes.batched_dam_breaks(5, seed=13, n_min=2000, n_max=40000)
    print(f"case variants batched counts={bs.counts.tolist()}", flush=True)
    outs = []
    for w in (0, 0x0032, 0x3341, 0x4450, 0x0073, 0x2295, 0x50000, 0xA3300):
        with BatchSolver([s.config for s in bs.scenes], [s.positions for s in bs.scenes]) as b:
            b.set_tune(w)
            b.Advance(bs.dt, 6)
            b.synchronize()
            outs.append([b.download(f) for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_DENSITY)])
    for o in outs[1:]:
        for a, c in zip(outs[0], o):
            assert np.array_equal(a, c)


@pytest.mark.gpu
def test_autotune_keeps_state_and_results():
    sc = scenes.dam_break((48, 48, 48), seed=4)
    print(f"case autotune n={sc.n}", flush=True)
    with Solver(ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config) as s, \
            Solver(ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config) as ref:
        s.Advance(sc.dt, 3)
        ref.Advance(sc.dt, 3)
        s.synchronize()
        before = s.positions()
        word, ms = s.autotune(sc.dt, reps=3)
        print(f"  tune {word:#06x} {autotune.describe(word)} ms {ms}", flush=True)
        assert s.get_tune() == word
        assert np.array_equal(s.positions(), before)
        assert ms[1] <= ms[0] and ms[3] <= ms[2]
        s.Advance(sc.dt, 5)
        ref.Advance(sc.dt, 5)
        s.synchronize()
        ref.synchronize()
        assert np.array_equal(s.positions(), ref.positions())
        assert np.array_equal(s.velocities(), ref.velocities())


@pytest.mark.gpu
def test_batch_autotune_and_invalid_words():
    bs = scenes.batched_dam_breaks(4, seed=14, n_min=4096, n_max=30000)
    with BatchSolver([s.config for s in bs.scenes], [s.positions for s in bs.scenes]) as b:
        word, ms = b.autotune(bs.dt, reps=2)
        assert b.get_tune() == word and len(ms) == 4
        for bad in (0x6, 0xA0, 0x500, 0x5000, 0x30000, 0xC0000, 0x10001, 0x40010,
                    0x100000, 0x400000, 0x1000000):
            with pytest.raises(SPHError):
                b.set_tune(bad)
        assert b.get_tune() == word


@pytest.mark.gpu
def test_single_solver_refuses_invalid_words():
    sc = scenes.dam_break((12, 12, 12), seed=6)
    with Solver(ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config) as s:
        s.set_tune(0x500000)  # two particles per lane in both passes
        for bad in (0x200000, 0x800000, 0x100001, 0x110000, 0x400010, 0x440000, 0x1000000):
            with pytest.raises(SPHError):
                s.set_tune(bad)
        assert s.get_tune() == 0x500000
