"""Independent pins of the oracle's two-scale cross-resolution coupling.

The reference couples the two resolutions through complementary sums
(TWO_SCALE/Solver.cu:252-290 density, 394-499 force, 960-1000 the stencil
over the other set's grid): a particle of set A sees every particle j of the
other set B within the LOW-resolution radius h0, with B's mass m_B and the
averaged kernels 1/2 (W_h0 + W_h1), 1/2 (grad W_h0 + grad W_h1) and
1/2 (lapl W_h0 + lapl W_h1), where the h1 kernel is 0 beyond h1.

These CPU tests check the oracle's fixed coupling without reusing its code:

* closed forms for one low and one high particle at r < h1, h1 <= r < h0 and
  r >= h0 (densities, pressure + viscosity + tension accelerations), written
  from Mueller, Charypar and Gross (SCA 2003) kernels;
* momentum: with no walls in reach, sum_low m0 a + sum_high m1 a = M g, since
  every cross pressure and viscosity pair term is antisymmetric;
* an O(N^2) numpy restatement of both sets' densities and accelerations on
  about 2,000 particles, which pins the stencil's reach (a low particle spans
  up to 5 cells per axis of the h/2 grid);
* planted mistakes: with each of the own set's mass in place of the other's,
  cutoff h1 in place of h0, a dropped 1/2, or a stencil reaching only h1,
  planted in the oracle, at least the checks listed for it fail.
"""
import numpy as np
import pytest

import oracle as O
from sph2s_final_amd import Grid, scenes

PI = np.pi


# Mueller et al. 2003 kernels (r >= 0 arrays), zero at and beyond h.
def w_poly6(r, h):
    return np.where(r < h, 315.0 / (64.0 * PI * h ** 9) * np.clip(h * h - r * r, 0, None) ** 3,
                    0.0)


def spiky_mag(r, h):
    """|grad W_spiky| = 45/(pi h^6) (h - r)^2, pointing from j to i (repulsive)."""
    return np.where(r < h, 45.0 / (PI * h ** 6) * np.clip(h - r, 0, None) ** 2, 0.0)


def visc_lapl(r, h):
    return np.where(r < h, 45.0 / (PI * h ** 6) * np.clip(h - r, 0, None), 0.0)


def _scene(tension=0.0):
    sc = scenes.two_scale_dam_break(10, seed=3)
    sc.config.tension = tension
    return sc


def _oracle(sc):
    p = O.params_from_config(sc.config)
    gh = Grid.MakeGrid((0.0, 0.0, 0.0), [sc.config.domain_hi[d] for d in range(3)],
                       0.5 * sc.config.h)
    return O.OracleTwoScale(p, gh, sc.region_lo, sc.region_hi, 1 << 20)


def _constants(sc):
    """h0, h1, m0, m1, rho0, k, mu, sigma, g exactly as the two-scale step rounds them."""
    h0f = np.float32(sc.config.h)
    h1f = np.float32(0.5) * h0f
    m0f = np.float32(sc.config.mass)
    m1f = m0f / np.float32(8.0)
    g = np.array([sc.config.gravity[d] for d in range(3)], np.float64)
    return (float(h0f), float(h1f), float(m0f), float(m1f), float(sc.config.rest_density),
            float(sc.config.stiffness), float(sc.config.viscosity), float(sc.config.tension), g)


def _step(sc, xl, vl, xh, vh, tags_h=None):
    """One oracle step; returns (sorted x, v, rho, acc) per set, pre-integration."""
    o = _oracle(sc)
    nh = len(xh)
    ih = np.arange(nh, dtype=np.uint32) + (1 << 20)
    th = np.arange(nh, dtype=np.uint32) if tags_h is None else tags_h
    o.set_state(xl, vl, None, xh, vh, ih, th)
    o.step(sc.dt)
    out = []
    for r, (x, v) in enumerate(((xl, vl), (xh, vh))):
        perm = o.get(r, "perm")
        out.append((x.astype(np.float64)[perm], v.astype(np.float64)[perm], o.get(r, "rho"),
                    o.get(r, "acc")))
    return out


# ---------------------------------------------------------------- closed forms
def _pair_case(sc, r_over_h0):
    h0 = _constants(sc)[0]
    c = np.array([0.15, 0.16, 0.14])
    u = np.array([1.0, 0.3, -0.5])
    u /= np.linalg.norm(u)
    xl = c.astype(np.float32)[None, :]
    xh = (c + r_over_h0 * h0 * u).astype(np.float32)[None, :]
    vl = np.array([[0.1, -0.2, 0.05]], np.float32)
    vh = np.array([[-0.3, 0.1, 0.2]], np.float32)
    return xl, vl, xh, vh


def _pair_closed_form(sc, xl, vl, xh, vh):
    h0, h1, m0, m1, rho0, k, mu, sigma, g = _constants(sc)
    x_l, x_h = xl[0].astype(np.float64), xh[0].astype(np.float64)
    v_l, v_h = vl[0].astype(np.float64), vh[0].astype(np.float64)
    d = x_l - x_h                               # x_lh
    r = float(np.sqrt(d @ d))
    inr = r < h0
    w_avg = 0.5 * (w_poly6(r, h0) + w_poly6(r, h1)) if inr else 0.0
    rho_l = m0 * w_poly6(0.0, h0) + m1 * w_avg
    rho_h = m1 * w_poly6(0.0, h1) + m0 * w_avg
    A_l, A_h = k * (rho_l - rho0) / rho_l ** 2, k * (rho_h - rho0) / rho_h ** 2
    s_avg = 0.5 * (spiky_mag(r, h0) + spiky_mag(r, h1)) if inr else 0.0
    l_avg = 0.5 * (visc_lapl(r, h0) + visc_lapl(r, h1)) if inr else 0.0
    # a_i = sum_j [ m_j (A_i + A_j) |grad W| x_ij / r + mu m_j / (rho_i rho_j) lapl (v_j - v_i)
    #               - sigma m_j / rho_i W x_ij ] + g
    a_l = (m1 * (A_l + A_h) * s_avg * d / r + mu * m1 / (rho_l * rho_h) * l_avg * (v_h - v_l)
           - sigma * m1 / rho_l * w_avg * d + g)
    a_h = (m0 * (A_h + A_l) * s_avg * (-d) / r + mu * m0 / (rho_l * rho_h) * l_avg * (v_l - v_h)
           - sigma * m0 / rho_h * w_avg * (-d) + g)
    return float(rho_l), float(rho_h), a_l, a_h


def _pair_error(sc, r_over_h0):
    xl, vl, xh, vh = _pair_case(sc, r_over_h0)
    (_, _, rho_l, acc_l), (_, _, rho_h, acc_h) = _step(sc, xl, vl, xh, vh)
    cl, ch, al, ah = _pair_closed_form(sc, xl, vl, xh, vh)
    e_rho = max(abs(rho_l[0] - cl) / cl, abs(rho_h[0] - ch) / ch)
    scale = max(np.abs(al - _constants(sc)[8]).max(), np.abs(ah - _constants(sc)[8]).max(), 1.0)
    e_acc = max(np.abs(acc_l[0] - al).max(), np.abs(acc_h[0] - ah).max()) / scale
    return e_rho, e_acc


# r/h0: 0.3 and 0.45 (< h1 = h0/2), 0.6 and 0.95 (h1 <= r < h0), 1.2 (beyond h0)
PAIR_RADII = [0.3, 0.45, 0.6, 0.95, 1.2]


@pytest.mark.parametrize("tension", [0.0, 0.8])
@pytest.mark.parametrize("r_over_h0", PAIR_RADII)
def test_cross_pair_closed_form(r_over_h0, tension):
    sc = _scene(tension)
    e_rho, e_acc = _pair_error(sc, r_over_h0)
    assert e_rho < 1e-12 and e_acc < 1e-12, (e_rho, e_acc)


def test_cross_pair_beyond_h0_is_uncoupled():
    sc = _scene()
    xl, vl, xh, vh = _pair_case(sc, 1.2)
    (_, _, rho_l, acc_l), (_, _, rho_h, acc_h) = _step(sc, xl, vl, xh, vh)
    h0, h1, m0, m1 = _constants(sc)[:4]
    g = _constants(sc)[8]
    assert rho_l[0] == pytest.approx(m0 * 315.0 / (64.0 * PI * h0 ** 3), rel=1e-13)
    assert rho_h[0] == pytest.approx(m1 * 315.0 / (64.0 * PI * h1 ** 3), rel=1e-13)
    assert np.allclose(acc_l[0], g, rtol=0, atol=1e-13)
    assert np.allclose(acc_h[0], g, rtol=0, atol=1e-13)


# ---------------------------------------------------------------- momentum
def _cluster(seed=11, n_low=250, n_high=700):
    """Random low and high particles interleaved in a box well inside the domain."""
    rng = np.random.default_rng(seed)
    lo, hi = 0.10, 0.21
    xl = rng.uniform(lo, hi, (n_low, 3)).astype(np.float32)
    xh = rng.uniform(lo, hi, (n_high, 3)).astype(np.float32)
    vl = rng.normal(0.0, 0.3, (n_low, 3)).astype(np.float32)
    vh = rng.normal(0.0, 0.3, (n_high, 3)).astype(np.float32)
    return xl, vl, xh, vh


def _momentum_residual(sc, case):
    h0, h1, m0, m1 = _constants(sc)[:4]
    g = _constants(sc)[8]
    (_, _, _, acc_l), (_, _, _, acc_h) = _step(sc, *case)
    total = m0 * acc_l.sum(axis=0) + m1 * acc_h.sum(axis=0)
    mass = m0 * len(acc_l) + m1 * len(acc_h)
    scale = m0 * np.abs(acc_l).sum() + m1 * np.abs(acc_h).sum()
    return float(np.abs(total - mass * g).max() / scale)


def test_cross_momentum_invariant():
    sc = _scene()
    res = _momentum_residual(sc, _cluster())
    assert res < 1e-12, res


# ---------------------------------------------------------------- all pairs
def _all_pairs(sc, xs, vs):
    """Both sets' densities and accelerations by an O(N^2) sum over every pair.
    xs, vs: [low, high] sorted FP64 arrays."""
    h0, h1, m0, m1, rho0, k, mu, sigma, g = _constants(sc)
    hh, mm = (h0, h1), (m0, m1)
    rho = []
    for a in (0, 1):
        b = 1 - a
        ra = np.linalg.norm(xs[a][:, None, :] - xs[a][None, :, :], axis=2)
        rb = np.linalg.norm(xs[a][:, None, :] - xs[b][None, :, :], axis=2)
        own = mm[a] * w_poly6(ra, hh[a]).sum(axis=1)
        cross = mm[b] * (0.5 * (w_poly6(rb, h0) + w_poly6(rb, h1)) * (rb < h0)).sum(axis=1)
        rho.append(own + cross)
    press = [k * (r - rho0) / r ** 2 for r in rho]
    acc = []
    for a in (0, 1):
        b = 1 - a
        ai = np.zeros_like(xs[a])
        for other, avg in ((a, False), (b, True)):
            d = xs[a][:, None, :] - xs[other][None, :, :]        # x_ij
            r = np.linalg.norm(d, axis=2)
            m = mm[other]
            if avg:
                live = (r > 0) & (r < h0)
                s = 0.5 * (spiky_mag(r, h0) + spiky_mag(r, h1))
                lap = 0.5 * (visc_lapl(r, h0) + visc_lapl(r, h1))
                w = 0.5 * (w_poly6(r, h0) + w_poly6(r, h1))
            else:
                live = (r > 0) & (r < hh[a])
                s, lap, w = spiky_mag(r, hh[a]), visc_lapl(r, hh[a]), w_poly6(r, hh[a])
            rs = np.where(live, r, 1.0)
            pres = m * (press[a][:, None] + press[other][None, :]) * s / rs * live
            vis = mu * m / (rho[a][:, None] * rho[other][None, :]) * lap * live
            ten = sigma * m / rho[a][:, None] * w * live
            ai += ((pres - ten)[:, :, None] * d).sum(axis=1)
            ai += (vis[:, :, None] * (vs[other][None, :, :] - vs[a][:, None, :])).sum(axis=1)
        acc.append(ai + g)
    return rho, acc


def _split_case(seed=3):
    sc = _scene()
    o = _oracle(sc)
    o.set_state(sc.positions, scenes.gaussian_velocities(sc.n, 0.3, seed))
    o.split_merge()
    xl, vl = o.get(0, "pos").astype(np.float32), o.get(0, "vel").astype(np.float32)
    xh, vh = o.get(1, "pos").astype(np.float32), o.get(1, "vel").astype(np.float32)
    return sc, (xl, vl, xh, vh, o.get(1, "tag"))


def _all_pairs_error(sc, case):
    res = _step(sc, *case)
    xs, vs = [res[0][0], res[1][0]], [res[0][1], res[1][1]]
    rho, acc = _all_pairs(sc, xs, vs)
    e_rho = max(np.max(np.abs(res[r][2] - rho[r]) / rho[r]) for r in (0, 1))
    scale = max(np.abs(acc[r]).max() for r in (0, 1))
    e_acc = max(np.max(np.abs(res[r][3] - acc[r])) for r in (0, 1)) / scale
    return float(e_rho), float(e_acc)


def test_cross_all_pairs_split_interface():
    """The split dam break: the interface between the sets runs along the
    region's faces, so low particles reach high particles up to h0 away."""
    sc, case = _split_case()
    nl, nh = len(case[0]), len(case[2])
    assert nl > 500 and nh > 500
    e_rho, e_acc = _all_pairs_error(sc, case)
    assert e_rho < 1e-12 and e_acc < 1e-12, (e_rho, e_acc)


def test_cross_all_pairs_interleaved_cluster():
    sc = _scene(tension=0.8)
    e_rho, e_acc = _all_pairs_error(sc, _cluster(seed=5))
    assert e_rho < 1e-12 and e_acc < 1e-12, (e_rho, e_acc)


# ---------------------------------------------------------------- planted mistakes
MISTAKES = {1: "own mass for the other set", 2: "cutoff h1 for h0", 3: "no 1/2 in the averages",
            4: "stencil reach h1"}


@pytest.fixture
def planted():
    yield O.OracleTwoScale.plant_mistake
    O.OracleTwoScale.plant_mistake(0)


@pytest.mark.parametrize("which", sorted(MISTAKES))
def test_pins_catch_planted_mistakes(planted, which):
    """Each planted mistake is caught: by the closed forms (some radius), by
    the all-pairs sums, and -- for the mass swap, the only one that breaks the
    pairwise antisymmetry -- by the momentum invariant."""
    sc = _scene(tension=0.8)
    sc0 = _scene()
    planted(which)
    closed = max(max(_pair_error(sc, r)) for r in PAIR_RADII)
    _, split = _split_case()
    pairs = max(max(_all_pairs_error(sc0, split)), max(_all_pairs_error(sc, _cluster(seed=5))))
    momentum = _momentum_residual(sc0, _cluster())
    planted(0)
    print(f"planted {MISTAKES[which]}: closed {closed:.2e} all-pairs {pairs:.2e} "
          f"momentum {momentum:.2e}")
    assert pairs > 1e-6, "all-pairs sums missed the mistake"
    if which != 4:
        assert closed > 1e-6, "closed forms missed the mistake"
    if which == 1:
        assert momentum > 1e-6, "momentum invariant missed the mass swap"
