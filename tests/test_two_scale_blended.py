"""Two-scale step under the reference's blended coupling
(SPH2S_FINAL_TWO_SCALE/Solver.cu:215-499, 732-1083, 1163-1289): per-particle
blend weights lambda on every neighbour term, integrateD's fade-in/fade-out
state machine and velocity cap, and the r/sqrt(3) split offsets
(insertHighResParticles, Solver.cu:541-586).

CPU tests pin the FP64 oracle's blended step without retyping it:

* lambda = 1 everywhere reduces it to the fixed coupling (densities and
  accelerations equal);
* lambda_j = 0 for a subset equals removing that subset from the state
  (densities and accelerations of the others equal), which fails if a term
  misses its lambda_j or takes lambda_i instead;
* a uniform lambda = c scales every density by c;
* the velocity cap leaves |v| <= c h / dt along the uncapped direction and
  leaves slower particles untouched;
* a reference-offset split puts the 8 children on the corners of a cube
  centred on the parent whose half-diagonal r satisfies 4/3 pi (2 r)^3 rho = m0;
* the fade timeline: lambda_parent + lambda_child = 1 at every step, the
  parent leaves and the children settle on the same step, and the effective
  mass m0 sum lambda_low + m1 sum lambda_high is conserved through splits and
  merges; ids stay unique; a split waits for its group's children and a merge
  for the group's low particle.

GPU tests compare the HIP blended step with the oracle re-seeded from the GPU
state every step: counts, ids, tags, lambda and state words bit-exact; sorted
hashes, permutations and cell ranges bit-exact; densities within 1e-5
relative, accelerations within 1e-5 of max |a|, positions within 1e-6 and
velocities within 1e-5 (FP32 vs FP64).  Graph replay equals eager launches
bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from sph2s_final_amd import (FIELD_ACCEL, FIELD_BLEND, FIELD_CELL_END, FIELD_CELL_START,
                             FIELD_DENSITY, FIELD_ID, FIELD_POSITION, FIELD_SORTED_HASH,
                             FIELD_SORTED_INDEX, FIELD_STATE, FIELD_TAG, FIELD_VELOCITY,
                             SPH2_COUPLING_BLENDED, SPH2_SPLIT_LATTICE, SPH2_SPLIT_REFERENCE,
                             SPH_DEBUG_ACCEL, Grid, TwoScaleSolver, scenes)

MASK = (1 << 29) - 1
EMPTY_REGION = ((-1.0, -1.0, -1.0), (-0.5, -0.5, -0.5))


def _oracle(sc, region=None, params=None):
    p = params or O.params_from_config(sc.config)
    gh = Grid.MakeGrid((0.0, 0.0, 0.0), [sc.config.domain_hi[d] for d in range(3)],
                       0.5 * sc.config.h)
    lo, hi = region or (sc.region_lo, sc.region_hi)
    return O.OracleTwoScale(p, gh, lo, hi, sc.n)


def _split_state(sc, vel=None):
    """Both sets populated: the initial lattice after one fixed split pass."""
    o = _oracle(sc)
    o.set_state(sc.positions, sc.velocities if vel is None else vel)
    o.split_merge()
    return o


def _copy_state(dst, src):
    """dst (oracle) <- src (oracle) state, FP32-rounded as a GPU state would be."""
    dst.set_state(src.get(0, "pos").astype(np.float32), src.get(0, "vel").astype(np.float32),
                  src.get(0, "id"), src.get(1, "pos").astype(np.float32),
                  src.get(1, "vel").astype(np.float32), src.get(1, "id"), src.get(1, "tag"))


def _blended(o, **kw):
    args = dict(coupling=1, split_offsets=0, merge=1, blend_increment=0.0125, velocity_cap=0.0)
    args.update(kw)
    o.set_options(**args)


def _by_id(o, res, name):
    ids = o.get(res, "id")
    order = np.argsort(ids)
    return ids[order], o.get(res, name)[order]


def _eff_mass(o, m0):
    return m0 * o.get(0, "blend").astype(np.float64).sum() + \
        (m0 / 8.0) * o.get(1, "blend").astype(np.float64).sum()


# ---------------------------------------------------------------- CPU: oracle
def test_blended_lambda_one_equals_fixed():
    sc = scenes.two_scale_dam_break(10, seed=4)
    base = _split_state(sc, scenes.gaussian_velocities(sc.n, 0.3, 2))
    fixed, blend = _oracle(sc, EMPTY_REGION), _oracle(sc, EMPTY_REGION)
    _copy_state(fixed, base)
    _copy_state(blend, base)
    fixed.set_options(coupling=0, merge=0)
    _blended(blend, merge=0)
    fixed.step(sc.dt)
    blend.step(sc.dt)
    for r in (0, 1):
        rf, rb = fixed.get(r, "rho"), blend.get(r, "rho")
        af, ab = fixed.get(r, "acc"), blend.get(r, "acc")
        assert len(rf) > 0
        assert np.max(np.abs(rb - rf) / rf) < 1e-12
        assert np.max(np.abs(ab - af)) < 1e-10 * np.max(np.abs(af))
        assert np.array_equal(fixed.get(r, "id"), blend.get(r, "id"))
        assert np.max(np.abs(fixed.get(r, "pos") - blend.get(r, "pos"))) < 1e-12


def _pre_state(o):
    return {r: {k: o.get(r, k).copy() for k in ("pos", "vel", "id", "tag")} for r in (0, 1)}


def test_blended_zero_weight_equals_removal():
    sc = scenes.two_scale_dam_break(10, seed=4)
    base = _split_state(sc, scenes.gaussian_velocities(sc.n, 0.3, 3))
    nl, nh = base.counts()[:2]
    rng = np.random.default_rng(0)
    drop = [rng.random(nl) < 0.2, rng.random(nh) < 0.3]
    st = _pre_state(base)
    # blended, dropped particles at lambda = 0 (settled state, so lambda stays 0)
    blend = _oracle(sc, EMPTY_REGION)
    _copy_state(blend, base)
    _blended(blend, merge=0)
    for r in (0, 1):
        blend.set_blend(r, np.where(drop[r], 0.0, 1.0).astype(np.float32))
    blend.step(sc.dt)
    # fixed, dropped particles removed
    fixed = _oracle(sc, EMPTY_REGION)
    k = [~drop[0], ~drop[1]]
    fixed.set_state(st[0]["pos"][k[0]].astype(np.float32), st[0]["vel"][k[0]].astype(np.float32),
                    st[0]["id"][k[0]], st[1]["pos"][k[1]].astype(np.float32),
                    st[1]["vel"][k[1]].astype(np.float32), st[1]["id"][k[1]], st[1]["tag"][k[1]])
    fixed.set_options(coupling=0, merge=0)
    fixed.step(sc.dt)
    for r in (0, 1):
        ids_b = st[r]["id"][blend.get(r, "perm")]
        ids_f = st[r]["id"][k[r]][fixed.get(r, "perm")]
        keep = np.isin(ids_b, ids_f)
        rb, ab = blend.get(r, "rho")[keep], blend.get(r, "acc")[keep]
        ob = np.argsort(ids_b[keep])
        of = np.argsort(ids_f)
        rf, af = fixed.get(r, "rho"), fixed.get(r, "acc")
        assert np.array_equal(ids_b[keep][ob], ids_f[of])
        assert np.max(np.abs(rb[ob] - rf[of]) / rf[of]) < 1e-12
        assert np.max(np.abs(ab[ob] - af[of])) < 1e-10 * np.max(np.abs(af))


def test_blended_uniform_weight_scales_density():
    sc = scenes.two_scale_dam_break(10, seed=4)
    base = _split_state(sc)
    one, half = _oracle(sc, EMPTY_REGION), _oracle(sc, EMPTY_REGION)
    for o, lam in ((one, 1.0), (half, 0.375)):
        _copy_state(o, base)
        _blended(o, merge=0)
        for r in (0, 1):
            o.set_blend(r, np.full(base.counts()[r], lam, np.float32))
        o.step(sc.dt)
    for r in (0, 1):
        assert np.max(np.abs(half.get(r, "rho") - 0.375 * one.get(r, "rho")) /
                      one.get(r, "rho")) < 1e-12


@pytest.mark.parametrize("coupling", [0, 1])
def test_velocity_cap(coupling):
    sc = scenes.two_scale_dam_break(10, seed=4)
    p = O.params_from_config(sc.config)
    p.wall_clamp = 0
    vel = scenes.gaussian_velocities(sc.n, 14.0, 11)
    base = _split_state(sc, vel)
    dt, cap = sc.dt, 0.8
    runs = []
    for c in (0.0, cap):
        o = _oracle(sc, EMPTY_REGION, params=p)
        _copy_state(o, base)
        o.set_options(coupling=coupling, merge=0, velocity_cap=c)
        o.step(dt)
        runs.append(o)
    capped = 0
    for r in (0, 1):
        h = np.float32(sc.config.h) * (np.float32(0.5) if r else np.float32(1.0))
        vmax = float(np.float32(cap) * h / np.float32(dt))
        _, v0 = _by_id(runs[0], r, "vel")
        _, v1 = _by_id(runs[1], r, "vel")
        n0 = np.linalg.norm(v0, axis=1)
        fast = n0 > vmax
        capped += int(fast.sum())
        assert np.array_equal(v1[~fast], v0[~fast])
        assert np.allclose(v1[fast], v0[fast] * (vmax / n0[fast])[:, None], rtol=1e-12, atol=0)
        assert np.all(np.linalg.norm(v1, axis=1) <= vmax * (1 + 1e-12))
    assert capped > 0


def test_reference_split_offsets_and_states():
    sc = scenes.two_scale_dam_break(10, seed=4)
    o = _oracle(sc, EMPTY_REGION)
    o.set_state(sc.positions, sc.velocities)
    _blended(o, split_offsets=1)
    marked = np.array([5, 17, 123, 400], np.uint32)
    o.mark_split(marked)
    state = o.get(0, "state")
    assert np.array_equal(np.flatnonzero(state == 3), marked)
    pre_ids = o.get(0, "id").copy()
    o.step(sc.dt)
    m0 = float(np.float32(sc.config.mass))
    rho_sorted = o.get(0, "rho")
    ids_sorted = pre_ids[o.get(0, "perm")]
    il, xl = _by_id(o, 0, "pos")
    _, sl = _by_id(o, 0, "state")
    _, ll = _by_id(o, 0, "blend")
    ih, xh = _by_id(o, 1, "pos")
    _, sh = _by_id(o, 1, "state")
    _, lh = _by_id(o, 1, "blend")
    assert len(ih) == 8 * len(marked) and len(il) == sc.n
    for p in marked:
        q = int(np.flatnonzero(il == p)[0])
        assert sl[q] == 1 and ll[q] == 1.0            # the parent fades out
        kids = np.flatnonzero((ih - sc.n) // 8 == p)
        assert len(kids) == 8
        assert np.all(sh[kids] == 6) and np.all(lh[kids] == 0.0)
        rho = rho_sorted[np.flatnonzero(ids_sorted == p)[0]]
        off = xh[kids] - xl[q]
        assert np.allclose(off.mean(axis=0), 0.0, atol=1e-12)
        r = np.linalg.norm(off, axis=1)
        assert np.allclose(r, r[0], rtol=1e-12)
        assert np.allclose(np.abs(off), r[0] / np.sqrt(3.0), rtol=1e-12)
        # the child cube's half-diagonal r: a sphere of radius 2 r holds m0 at rho
        assert abs(4.0 / 3.0 * np.pi * (2.0 * r[0]) ** 3 * rho / m0 - 1.0) < 1e-12
        k = (ih[kids] - sc.n) % 8
        signs = np.stack([(k >> a) & 1 for a in range(3)], 1) == 1
        assert np.array_equal(off > 0, signs)


@pytest.mark.parametrize("inc", [0.25, 0.3])
def test_fade_timeline(inc):
    sc = scenes.two_scale_dam_break(10, seed=4)
    o = _oracle(sc, EMPTY_REGION)
    o.set_state(sc.positions, sc.velocities)
    _blended(o, blend_increment=inc, merge=0)
    marked = np.array([3, 250, 777], np.uint32)
    o.mark_split(marked)
    o.step(sc.dt)
    m0 = float(np.float32(sc.config.mass))
    e0 = _eff_mass(o, m0)
    lam_p, lam_c = np.float32(1.0), np.float32(0.0)
    for k in range(1, 10):
        o.step(sc.dt)
        lam_p = np.float32(lam_p - np.float32(inc))
        lam_c = np.float32(lam_c + np.float32(inc))
        il, ll = _by_id(o, 0, "blend")
        ih, lh = _by_id(o, 1, "blend")
        _, sh = _by_id(o, 1, "state")
        parents = np.isin(il, marked)
        if lam_p <= 0:
            assert not parents.any() and len(il) == sc.n - len(marked)
            assert lam_c >= 1.0 and np.all(lh == 1.0) and np.all(sh == 4)
            break
        assert parents.sum() == len(marked)
        assert np.all(ll[parents] == lam_p) and np.all(lh == lam_c) and np.all(sh == 6)
        assert abs(_eff_mass(o, m0) - e0) < 1e-6 * e0
    else:
        pytest.fail("the parents never left")
    assert abs(_eff_mass(o, m0) - e0) < 1e-6 * e0


def _translate_groups(xh, groups, x_target):
    """Move whole sibling groups (8 consecutive children) rigidly to x ~ x_target."""
    for q in groups:
        blk = slice(8 * q, 8 * q + 8)
        xh[blk, 0] += x_target - xh[blk, 0].mean()
    return xh


def test_faded_merge_and_merge_waits_for_parent():
    sc = scenes.two_scale_dam_break(10, seed=4)
    base = _split_state(sc)
    th = base.get(1, "tag")
    g0, g1 = int(th[0] & MASK), int(th[8] & MASK)
    xh = _translate_groups(base.get(1, "pos"), (0, 1), 0.08)
    # a low particle with id g1 still alive (fading out, lambda 1): g1 must wait
    xl = np.vstack([base.get(0, "pos"), xh[8:16].mean(axis=0)]).astype(np.float32)
    vl = np.vstack([base.get(0, "vel"), np.zeros((1, 3))]).astype(np.float32)
    il = np.append(base.get(0, "id"), np.uint32(g1))
    o = _oracle(sc)
    o.set_state(xl, vl, il, xh.astype(np.float32), base.get(1, "vel").astype(np.float32),
                base.get(1, "id"), th)
    _blended(o, blend_increment=0.25)
    st = o.get(0, "state")
    st[-1] = 1
    o.set_blend(0, None, st)
    m0 = float(np.float32(sc.config.mass))

    def eff():
        # the effective mass without the added (unpaired) fading-out particle
        ids, lam, st = o.get(0, "id"), o.get(0, "blend"), o.get(0, "state")
        extra = (ids == g1) & (st == 1)
        return _eff_mass(o, m0) - m0 * float(lam[extra].sum())
    e0 = eff()
    o.step(sc.dt)
    il, sl = _by_id(o, 0, "state")
    _, ll = _by_id(o, 0, "blend")
    assert np.sum(il == g0) == 1 and sl[il == g0][0] == 2 and ll[il == g0][0] == 0.0
    assert np.sum(il == g1) == 1 and sl[il == g1][0] == 1 and ll[il == g1][0] == 0.75
    ih, sh = _by_id(o, 1, "state")
    grp = (ih - sc.n) // 8
    assert np.all(sh[grp == g0] == 5) and np.all(sh[grp == g1] == 4)
    assert abs(eff() - e0) < 1e-6 * e0
    for _ in range(4):
        o.step(sc.dt)
        assert abs(eff() - e0) < 1e-6 * e0
        for r in (0, 1):
            ids = o.get(r, "id")
            assert len(np.unique(ids)) == len(ids)
    il, sl = _by_id(o, 0, "state")
    ih, _ = _by_id(o, 1, "state")
    grp = (ih - sc.n) // 8
    assert not np.any(grp == g0)                            # faded out and removed
    assert np.sum(il == g0) == 1 and sl[il == g0][0] == 0   # settled
    # the old g1 parent left on the 4th step; the group merged then
    assert np.sum(il == g1) == 1 and sl[il == g1][0] == 2
    assert np.sum(grp == g1) == 8


def test_split_waits_for_live_children():
    sc = scenes.two_scale_dam_break(10, seed=4)
    base = _split_state(sc)
    th = base.get(1, "tag")
    g = int(th[0] & MASK)
    xl = np.vstack([base.get(0, "pos"), base.get(1, "pos")[0:8].mean(axis=0)]).astype(np.float32)
    vl = np.vstack([base.get(0, "vel"), np.zeros((1, 3))]).astype(np.float32)
    il = np.append(base.get(0, "id"), np.uint32(g))
    o = _oracle(sc, EMPTY_REGION)
    o.set_state(xl, vl, il, base.get(1, "pos").astype(np.float32),
                base.get(1, "vel").astype(np.float32), base.get(1, "id"), th)
    _blended(o, merge=0)
    o.mark_split([g])
    nh = o.counts()[1]
    o.step(sc.dt)
    il2, sl = _by_id(o, 0, "state")
    assert o.counts()[1] == nh and sl[il2 == g][0] == 3     # still waiting
    # without the children the request goes through
    keep = (base.get(1, "tag") & MASK) != g
    xl2, vl2 = o.get(0, "pos").astype(np.float32), o.get(0, "vel").astype(np.float32)
    il3, st3 = o.get(0, "id"), o.get(0, "state")
    o.set_state(xl2, vl2, il3, o.get(1, "pos")[keep].astype(np.float32),
                o.get(1, "vel")[keep].astype(np.float32), o.get(1, "id")[keep],
                o.get(1, "tag")[keep])
    o.set_blend(0, None, st3)
    o.step(sc.dt)
    ih, sh = _by_id(o, 1, "state")
    kids = (ih - sc.n) // 8 == g
    assert kids.sum() == 8 and np.all(sh[kids] == 6)
    il4, sl4 = _by_id(o, 0, "state")
    assert sl4[il4 == g][0] == 1


def test_blended_rollout_invariants():
    sc = scenes.two_scale_dam_break(10, seed=6)
    o = _oracle(sc)
    o.set_state(sc.positions, scenes.gaussian_velocities(sc.n, 0.5, 8))
    _blended(o, blend_increment=0.25, split_offsets=1, velocity_cap=0.8)
    m0 = float(np.float32(sc.config.mass))
    e0 = _eff_mass(o, m0)
    for _ in range(12):
        o.step(sc.dt)
        assert abs(_eff_mass(o, m0) - e0) < 1e-5 * e0
        for r in (0, 1):
            ids = o.get(r, "id")
            assert len(np.unique(ids)) == len(ids)
            lam, st = o.get(r, "blend"), o.get(r, "state")
            assert np.all((lam >= -1e-6) & (lam <= 1.0))
            assert np.all((st & 4) == (4 if r else 0))
    assert o.counts()[1] > 0


# ---------------------------------------------------------------- GPU parity
def _gpu(sc, region=None):
    lo, hi = region or (sc.region_lo, sc.region_hi)
    s = TwoScaleSolver(sc.config, lo, hi, sc.positions, sc.velocities, max_low=sc.n,
                       max_high=8 * sc.n)
    s.set_debug(SPH_DEBUG_ACCEL)
    return s


def _seed_oracle_from_gpu(o, s):
    o.set_state(s.download(0, FIELD_POSITION), s.download(0, FIELD_VELOCITY),
                s.download(0, FIELD_ID), s.download(1, FIELD_POSITION),
                s.download(1, FIELD_VELOCITY), s.download(1, FIELD_ID), s.download(1, FIELD_TAG))
    for r in (0, 1):
        o.set_blend(r, s.download(r, FIELD_BLEND), s.download(r, FIELD_STATE))


def _compare_step(s, o, atol_x=1e-6, atol_v=1e-5):
    assert s.device_errors() == 0
    assert s.counts() == o.counts()
    for r in (0, 1):
        assert np.array_equal(s.download(r, FIELD_SORTED_HASH), o.get(r, "hash"))
        assert np.array_equal(s.download(r, FIELD_SORTED_INDEX), o.get(r, "perm"))
        assert np.array_equal(s.download(r, FIELD_CELL_START), o.get(r, "start"))
        assert np.array_equal(s.download(r, FIELD_CELL_END), o.get(r, "end"))
        rho, rho_ref = s.download(r, FIELD_DENSITY), o.get(r, "rho")
        acc, acc_ref = s.download(r, FIELD_ACCEL), o.get(r, "acc")
        if len(rho_ref):
            rho_err = np.max(np.abs(rho - rho_ref) / rho_ref)
            acc_err = np.max(np.abs(acc - acc_ref)) / np.max(np.abs(acc_ref))
            print(f"  res {r}: rho {rho_err:.2e} acc {acc_err:.2e}", flush=True)
            assert rho_err <= 1e-5 and acc_err <= 1e-5
        assert np.array_equal(s.download(r, FIELD_ID), o.get(r, "id"))
        if r == 1:
            assert np.array_equal(s.download(r, FIELD_TAG), o.get(r, "tag"))
        assert np.array_equal(s.download(r, FIELD_BLEND), o.get(r, "blend"))
        assert np.array_equal(s.download(r, FIELD_STATE), o.get(r, "state"))
        dx = np.max(np.abs(s.download(r, FIELD_POSITION) - o.get(r, "pos")), initial=0.0)
        dv = np.max(np.abs(s.download(r, FIELD_VELOCITY) - o.get(r, "vel")), initial=0.0)
        print(f"  res {r}: pos {dx:.2e} vel {dv:.2e}", flush=True)
        assert dx <= atol_x and dv <= atol_v


@pytest.mark.gpu
@pytest.mark.parametrize("offsets", [SPH2_SPLIT_LATTICE, SPH2_SPLIT_REFERENCE])
def test_blended_step_parity(offsets):
    sc = scenes.two_scale_dam_break(14, seed=5)
    print(f"case blended_step offsets={offsets} n={sc.n}", flush=True)
    base = _split_state(sc, scenes.gaussian_velocities(sc.n, 0.3, 4))
    o = _oracle(sc)
    with _gpu(sc) as s:
        s.set_state(base.get(0, "pos").astype(np.float32), base.get(0, "vel").astype(np.float32),
                    base.get(0, "id"), base.get(1, "pos").astype(np.float32),
                    base.get(1, "vel").astype(np.float32), base.get(1, "id"), base.get(1, "tag"))
        s.set_options(coupling=SPH2_COUPLING_BLENDED, split_offsets=offsets,
                      blend_increment=0.25, velocity_cap=0.8)
        rng = np.random.default_rng(1)
        for r in (0, 1):
            n = s.counts()[r]
            s.set_blend(r, rng.random(n).astype(np.float32))
        s.mark_split(np.arange(0, sc.n, 97, dtype=np.uint32))
        _blended(o, split_offsets=offsets, blend_increment=0.25, velocity_cap=0.8)
        for step in range(4):
            print(f"  step {step}", flush=True)
            _seed_oracle_from_gpu(o, s)
            o.step(sc.dt)
            s.Advance(sc.dt)
            s.synchronize()
            _compare_step(s, o)


@pytest.mark.gpu
def test_blended_graph_matches_eager_and_conserves():
    sc = scenes.two_scale_dam_break(16, seed=2)
    print(f"case blended_rollout n={sc.n}", flush=True)
    with _gpu(sc) as a, _gpu(sc) as b:
        for s in (a, b):
            s.set_debug(0)
            s.set_options(coupling=SPH2_COUPLING_BLENDED, split_offsets=SPH2_SPLIT_REFERENCE,
                          blend_increment=0.1, velocity_cap=0.8)
        a.Advance(sc.dt, 40)
        b.advance_eager(sc.dt, 40)
        a.synchronize()
        b.synchronize()
        assert a.device_errors() == 0 and b.device_errors() == 0
        assert a.counts() == b.counts()
        m0 = float(np.float32(sc.config.mass))
        eff = m0 * a.download(0, FIELD_BLEND).astype(np.float64).sum() + \
            m0 / 8 * a.download(1, FIELD_BLEND).astype(np.float64).sum()
        assert abs(eff - m0 * sc.n) < 1e-5 * m0 * sc.n
        assert a.counts()[1] > 0
        for r in (0, 1):
            for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_ID, FIELD_BLEND, FIELD_STATE):
                assert np.array_equal(a.download(r, f), b.download(r, f))
            ids = a.download(r, FIELD_ID)
            assert len(np.unique(ids)) == len(ids)
            assert np.all(np.isfinite(a.download(r, FIELD_POSITION)))
