"""Two-scale step: split (1 -> 8) / merge (8 -> 1) bookkeeping and the coupled
two-resolution step.

CPU tests pin the two-scale oracle with invariants: a split conserves mass,
momentum and the centroid and places the 8 children at +/- h1/4 per axis;
a merge conserves mass and momentum and puts the merged particle at the
children's centroid; counts move by exactly -1/+8 per split and +1/-8 per
merge; ids and group tags follow the documented scheme.

GPU tests compare the HIP two-scale solver with the oracle on the same FP32
state: flags, counts, ids, tags and the compacted order bit-exact; positions
and velocities produced by split/merge within 1e-6 (FP32 vs FP64 offsets and
means); one coupled step: per resolution the sorted hashes, permutation and
cell ranges bit-exact, density within 1e-5 relative and acceleration within
1e-5 of max |a|.
"""
import numpy as np
import pytest

import oracle as O
from sph2s_final_amd import (FIELD_ACCEL, FIELD_CELL_END, FIELD_CELL_START, FIELD_DENSITY,
                             FIELD_ID, FIELD_POSITION, FIELD_SORTED_HASH, FIELD_SORTED_INDEX,
                             FIELD_TAG, FIELD_VELOCITY, SPH_DEBUG_ACCEL, Grid, TwoScaleSolver,
                             scenes)

MASK = (1 << 29) - 1


def _oracle(sc, max_low=None):
    p = O.params_from_config(sc.config)
    gh = Grid.MakeGrid((0.0, 0.0, 0.0), [sc.config.domain_hi[d] for d in range(3)],
                       0.5 * sc.config.h)
    return O.OracleTwoScale(p, gh, sc.region_lo, sc.region_hi, max_low or sc.n)


def _split_state(sc):
    """The initial lattice after one split/merge pass (both sets populated)."""
    o = _oracle(sc)
    o.set_state(sc.positions, sc.velocities)
    o.split_merge()
    return o


# ---------------------------------------------------------------- CPU: oracle
def test_oracle_split_conserves_mass_momentum_centroid():
    sc = scenes.two_scale_dam_break(12, seed=3)
    vel = scenes.gaussian_velocities(sc.n, 0.3, 5)
    o = _oracle(sc)
    o.set_state(sc.positions, vel)
    o.split_merge()
    nl, nh, _, _ = o.counts()
    lo, hi = np.array(sc.region_lo, np.float32), np.array(sc.region_hi, np.float32)
    inside = np.all((sc.positions >= lo) & (sc.positions <= hi), axis=1)
    assert nl == sc.n - inside.sum() and nh == 8 * inside.sum()
    xh, vh, ih, th = o.get(1, "pos"), o.get(1, "vel"), o.get(1, "id"), o.get(1, "tag")
    parents = np.flatnonzero(inside)
    d = 0.25 * np.float32(0.5 * np.float32(sc.config.h))
    for q, pid in enumerate(parents):
        kids = slice(8 * q, 8 * q + 8)
        assert np.array_equal(th[kids] & MASK, np.full(8, pid))
        assert np.array_equal(th[kids] >> 29, np.arange(8))
        assert np.array_equal(ih[kids], sc.n + 8 * pid + np.arange(8))
        assert np.allclose(xh[kids].mean(axis=0), sc.positions[pid], atol=1e-12)
        assert np.allclose(vh[kids].mean(axis=0), vel[pid], atol=1e-12)
        offs = xh[kids] - sc.positions[pid].astype(np.float64)
        assert np.allclose(np.abs(offs), d, rtol=1e-6)
        k = np.arange(8)
        assert np.array_equal(offs > 0, np.stack([(k >> a) & 1 for a in range(3)], 1) == 1)


def test_oracle_merge_group_outside_region():
    sc = scenes.two_scale_dam_break(12, seed=3)
    o = _oracle(sc)
    o.set_state(sc.positions, sc.velocities)
    o.split_merge()
    nl0, nh0, _, _ = o.counts()
    xl, vl, il = o.get(0, "pos"), o.get(0, "vel"), o.get(0, "id")
    xh, vh, ih, th = o.get(1, "pos"), o.get(1, "vel"), o.get(1, "id"), o.get(1, "tag")
    # Move the first two groups far outside the region (x = 0.05), give them velocities.
    rng = np.random.default_rng(0)
    for q in range(2):
        xh[8 * q:8 * q + 8, 0] = 0.05 + 0.001 * rng.random(8)
        vh[8 * q:8 * q + 8] = rng.normal(0, 0.2, (8, 3))
    o.set_state(xl.astype(np.float32), vl.astype(np.float32), il, xh.astype(np.float32),
                vh.astype(np.float32), ih, th)
    xs, vs = xh.astype(np.float32).astype(np.float64), vh.astype(np.float32).astype(np.float64)
    o.split_merge()
    nl, nh, _, _ = o.counts()
    assert (nl, nh) == (nl0 + 2, nh0 - 16)
    xm, vm, im = o.get(0, "pos")[-2:], o.get(0, "vel")[-2:], o.get(0, "id")[-2:]
    for q in range(2):
        g = th[8 * q] & MASK
        assert im[q] == g
        assert np.allclose(xm[q], xs[8 * q:8 * q + 8].mean(axis=0), atol=1e-12)
        assert np.allclose(vm[q], vs[8 * q:8 * q + 8].mean(axis=0), atol=1e-12)


def test_oracle_two_scale_steps_conserve_mass():
    sc = scenes.two_scale_dam_break(10, seed=4)
    o = _oracle(sc)
    o.set_state(sc.positions, sc.velocities)
    for _ in range(5):
        o.step(sc.dt)
        nl, nh, _, _ = o.counts()
        assert 8 * nl + nh == 8 * sc.n
    assert np.all(np.isfinite(o.get(0, "pos"))) and np.all(np.isfinite(o.get(1, "pos")))


# ---------------------------------------------------------------- GPU parity
def _gpu(sc, debug=True):
    s = TwoScaleSolver(sc.config, sc.region_lo, sc.region_hi, sc.positions, sc.velocities,
                       max_low=sc.n, max_high=8 * sc.n)
    if debug:
        s.set_debug(SPH_DEBUG_ACCEL)
    return s


def _set_both(solver, o):
    solver.set_state(o.get(0, "pos").astype(np.float32), o.get(0, "vel").astype(np.float32),
                     o.get(0, "id"), o.get(1, "pos").astype(np.float32),
                     o.get(1, "vel").astype(np.float32), o.get(1, "id"), o.get(1, "tag"))


def _compare_state(s, o, atol_x=1e-6, atol_v=1e-6):
    assert s.counts()[:2] == o.counts()[:2]
    for r in (0, 1):
        assert np.array_equal(s.download(r, FIELD_ID), o.get(r, "id")), r
        if r == 1:
            assert np.array_equal(s.download(r, FIELD_TAG), o.get(r, "tag"))
        dx = np.max(np.abs(s.download(r, FIELD_POSITION) - o.get(r, "pos")), initial=0.0)
        dv = np.max(np.abs(s.download(r, FIELD_VELOCITY) - o.get(r, "vel")), initial=0.0)
        print(f"  res {r}: pos {dx:.2e} vel {dv:.2e}", flush=True)
        assert dx <= atol_x and dv <= atol_v


@pytest.mark.gpu
def test_split_merge_stage_parity():
    sc = scenes.two_scale_dam_break(14, seed=5)
    print(f"case split_merge_stage n={sc.n}", flush=True)
    vel = scenes.gaussian_velocities(sc.n, 0.2, 9)
    o = _oracle(sc)
    o.set_state(sc.positions, vel)
    with _gpu(sc) as s:
        s.set_state(sc.positions, vel)
        o.split_merge()
        s.apply_split_merge()
        _compare_state(s, o)
        # Second pass on the split state after moving some groups out: merges.
        xh, vh = o.get(1, "pos"), o.get(1, "vel")
        xh[:24, 0] = 0.02
        o.set_state(o.get(0, "pos").astype(np.float32), o.get(0, "vel").astype(np.float32),
                    o.get(0, "id"), xh.astype(np.float32), vh.astype(np.float32),
                    o.get(1, "id"), o.get(1, "tag"))
        _set_both(s, o)
        o.split_merge()
        s.apply_split_merge()
        _compare_state(s, o)
        assert s.device_errors() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(12, 3), (16, 7)])
def test_two_scale_step_parity(n, seed):
    sc = scenes.two_scale_dam_break(n, seed=seed)
    print(f"case two_scale_step n={sc.n}", flush=True)
    o = _split_state(sc)
    with _gpu(sc) as s:
        _set_both(s, o)
        o.step(sc.dt)
        s.Advance(sc.dt)
        s.synchronize()
        assert s.device_errors() == 0
        assert s.counts() == o.counts()
        for r in (0, 1):
            assert np.array_equal(s.download(r, FIELD_SORTED_HASH), o.get(r, "hash"))
            assert np.array_equal(s.download(r, FIELD_SORTED_INDEX), o.get(r, "perm"))
            assert np.array_equal(s.download(r, FIELD_CELL_START), o.get(r, "start"))
            assert np.array_equal(s.download(r, FIELD_CELL_END), o.get(r, "end"))
            rho, rho_ref = s.download(r, FIELD_DENSITY), o.get(r, "rho")
            acc, acc_ref = s.download(r, FIELD_ACCEL), o.get(r, "acc")
            rho_err = np.max(np.abs(rho - rho_ref) / rho_ref)
            acc_err = np.max(np.abs(acc - acc_ref)) / np.max(np.abs(acc_ref))
            print(f"  res {r}: rho {rho_err:.2e} acc {acc_err:.2e}", flush=True)
            assert rho_err <= 1e-5 and acc_err <= 1e-5
        _compare_state(s, o, atol_x=1e-6, atol_v=1e-5)


@pytest.mark.gpu
def test_two_scale_rollout_conserves_mass_and_matches_eager():
    sc = scenes.two_scale_dam_break(16, seed=2)
    print(f"case two_scale_rollout n={sc.n}", flush=True)
    with _gpu(sc, debug=False) as a, _gpu(sc, debug=False) as b:
        a.Advance(sc.dt, 60)
        b.advance_eager(sc.dt, 60)
        a.synchronize()
        b.synchronize()
        assert a.device_errors() == 0 and b.device_errors() == 0
        ca, cb = a.counts(), b.counts()
        assert ca == cb
        assert 8 * ca[0] + ca[1] == 8 * sc.n
        assert ca[1] > 0
        for r in (0, 1):
            for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_ID):
                assert np.array_equal(a.download(r, f), b.download(r, f))
            x = a.download(r, FIELD_POSITION)
            assert np.all(np.isfinite(x))


@pytest.mark.gpu
def test_two_scale_250k_full_size():
    """The two-scale workload at its own size (63^3 = 250,047 low particles,
    about 460k active after the first split): 100 graph-replayed steps equal
    100 eager steps bit for bit, particle count is conserved (8 n_low + n_high
    = 8 N), the state stays finite with no device error; then one step from
    that GPU state matches the oracle step from the same FP32 state (per set:
    sorted hashes, permutation and cell ranges bit-exact, density within 1e-5
    relative, acceleration within 1e-5 of max |a|, and the split/merge
    outcome: counts, ids and tags)."""
    sc = scenes.two_scale_250k()
    steps = 100
    print(f"case two_scale_250k n={sc.n} steps={steps}", flush=True)
    with _gpu(sc, debug=False) as a, _gpu(sc, debug=False) as b:
        a.Advance(sc.dt, steps)
        b.advance_eager(sc.dt, steps)
        a.synchronize()
        b.synchronize()
        assert a.device_errors() == 0 and b.device_errors() == 0
        ca, cb = a.counts(), b.counts()
        print(f"  counts after {steps} steps: {ca[:2]}", flush=True)
        assert ca == cb
        assert 8 * ca[0] + ca[1] == 8 * sc.n and ca[1] > 100_000
        for r in (0, 1):
            for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_ID):
                assert np.array_equal(a.download(r, f), b.download(r, f)), (r, f)
            assert np.array_equal(a.download(1, FIELD_TAG), b.download(1, FIELD_TAG))
            assert np.all(np.isfinite(a.download(r, FIELD_POSITION)))
            assert np.all(np.isfinite(a.download(r, FIELD_VELOCITY)))
        # One more step from the GPU state, on the GPU and on the oracle.
        o = _oracle(sc)
        o.set_state(a.download(0, FIELD_POSITION), a.download(0, FIELD_VELOCITY),
                    a.download(0, FIELD_ID), a.download(1, FIELD_POSITION),
                    a.download(1, FIELD_VELOCITY), a.download(1, FIELD_ID),
                    a.download(1, FIELD_TAG))
        a.set_debug(SPH_DEBUG_ACCEL)
        a.Advance(sc.dt)
        a.synchronize()
        o.step(sc.dt)
        assert a.device_errors() == 0
        assert a.counts() == o.counts()
        for r in (0, 1):
            assert np.array_equal(a.download(r, FIELD_SORTED_HASH), o.get(r, "hash"))
            assert np.array_equal(a.download(r, FIELD_SORTED_INDEX), o.get(r, "perm"))
            assert np.array_equal(a.download(r, FIELD_CELL_START), o.get(r, "start"))
            assert np.array_equal(a.download(r, FIELD_CELL_END), o.get(r, "end"))
            rho, rho_ref = a.download(r, FIELD_DENSITY), o.get(r, "rho")
            acc, acc_ref = a.download(r, FIELD_ACCEL), o.get(r, "acc")
            rho_err = np.max(np.abs(rho - rho_ref) / rho_ref)
            acc_err = np.max(np.abs(acc - acc_ref)) / np.max(np.abs(acc_ref))
            print(f"  res {r}: rho {rho_err:.2e} acc {acc_err:.2e} "
                  f"(max |a*| {np.max(np.abs(acc_ref)):.1f})", flush=True)
            assert rho_err <= 1e-5 and acc_err <= 1e-5
        # Velocities take dt * (the acceleration error); positions one more dt.
        amax = max(np.max(np.abs(o.get(r, "acc"))) for r in (0, 1))
        _compare_state(a, o, atol_x=1e-6 + sc.dt ** 2 * 1e-5 * amax,
                       atol_v=1e-6 + sc.dt * 1e-5 * amax)


# ---------------------------------------------------------------- GPU: tune words
TS_WORDS = [(0x0000, 0x0000), (0x0032, 0x0041), (0x10000, 0x80000), (0x100000, 0x400000),
            (0x501100, 0x0000), (0x0000, 0x424400), (0x4485, 0x2393)]


def _ts_run(sc, words, steps):
    with _gpu(sc, debug=False) as s:
        s.set_tune(0, words[0])
        s.set_tune(1, words[1])
        assert (s.get_tune(0), s.get_tune(1)) == tuple(words)
        s.Advance(sc.dt, steps)
        s.synchronize()
        assert s.device_errors() == 0
        return s.counts(), [s.download(r, f) for r in (0, 1)
                            for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_ID)]


@pytest.mark.gpu
def test_two_scale_tune_words_are_bit_identical():
    # The own-set passes' launch variants only change which runs are staged in
    # LDS, when queued pairs are evaluated and which block handles which
    # particles, so every word steps exactly like the default.
    sc = scenes.two_scale_dam_break(16, seed=5)
    print(f"case two_scale tune words n={sc.n}", flush=True)
    ref_counts, ref = _ts_run(sc, TS_WORDS[0], 12)
    for words in TS_WORDS[1:]:
        counts, got = _ts_run(sc, words, 12)
        assert counts == ref_counts, words
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), words


@pytest.mark.gpu
def test_two_scale_autotune_keeps_state_and_results():
    sc = scenes.two_scale_dam_break(16, seed=6)
    print(f"case two_scale autotune n={sc.n}", flush=True)
    with _gpu(sc, debug=False) as s, _gpu(sc, debug=False) as ref:
        s.Advance(sc.dt, 4)
        ref.Advance(sc.dt, 4)
        s.synchronize()
        before = [s.download(r, FIELD_POSITION) for r in (0, 1)]
        words, ms = s.autotune(reps=2)
        print(f"  words {[hex(w) for w in words]} ms {ms}", flush=True)
        assert [s.get_tune(0), s.get_tune(1)] == words and len(ms) == 8
        for r in (0, 1):
            assert np.array_equal(s.download(r, FIELD_POSITION), before[r])
        assert ms[1] <= ms[0] and ms[5] <= ms[4]
        s.Advance(sc.dt, 6)
        ref.Advance(sc.dt, 6)
        s.synchronize()
        ref.synchronize()
        assert s.counts() == ref.counts()
        for r in (0, 1):
            for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_ID):
                assert np.array_equal(s.download(r, f), ref.download(r, f))
        for bad_res, bad_word in ((2, 0), (-1, 0), (0, 0x6), (1, 0x200000), (0, 0x1000000)):
            with pytest.raises(Exception):
                s.set_tune(bad_res, bad_word)
        assert [s.get_tune(0), s.get_tune(1)] == words
