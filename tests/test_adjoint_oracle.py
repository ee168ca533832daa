"""Pins of the FP64 oracle adjoint (CPU only).

The oracle's reverse-mode gradients of L = mean((x_T - target)^2) with respect
to stiffness, viscosity and the initial velocities and positions are checked
two independent ways:
  * central finite differences of the oracle's own FP64 forward rollout
    (16^3 = 4,096 particles, 20 steps, N(0, 0.05^2) initial velocities,
    targets from hidden stiffness/viscosity): relative error <= 1e-5;
  * an independent restatement of the step in torch float64 over all pairs
    (no grid, no sort; pair masks held constant), differentiated by torch
    autograd, on 8^3 particles with and without wall contacts: 1e-9.
A planted error in any adjoint term (a sign, a dropped pair term, a wrong
power of rho) moves the gradients far beyond either bound.
"""
import numpy as np
import pytest
import torch

import oracle as O
from sph2s_final_amd import scenes


def _setup(n, steps, seed=1, sigma=0.05, k_hidden=2.2, mu_hidden=5.0):
    sc = scenes.dam_break((n, n, n), seed=seed)
    vel = scenes.gaussian_velocities(sc.n, sigma, 77 + seed).astype(np.float64)
    p = O.params_from_config(sc.config)
    ph = O.params_from_config(sc.config)
    ph.stiffness, ph.viscosity = k_hidden, mu_hidden
    pos = sc.positions.astype(np.float64)
    target = O.rollout_loss_grad(ph, pos, vel, sc.dt, steps, np.zeros_like(pos),
                                 grads=False)["x_final"]
    return sc, p, pos, vel, target


def _loss(p, pos, vel, dt, steps, target):
    return O.rollout_loss_grad(p, pos, vel, dt, steps, target, grads=False)["loss"]


def test_oracle_adjoint_matches_central_differences():
    steps = 20
    sc, p, pos, vel, target = _setup(16, steps)
    g = O.rollout_loss_grad(p, pos, vel, sc.dt, steps, target)
    assert g["loss"] > 0
    # stiffness and viscosity
    for name, attr in (("g_stiffness", "stiffness"), ("g_viscosity", "viscosity")):
        base = getattr(p, attr)
        eps = 1e-3 * base
        setattr(p, attr, base + eps)
        lp = _loss(p, pos, vel, sc.dt, steps, target)
        setattr(p, attr, base - eps)
        lm = _loss(p, pos, vel, sc.dt, steps, target)
        setattr(p, attr, base)
        fd = (lp - lm) / (2 * eps)
        rel = abs(g[name] - fd) / abs(fd)
        print(f"{name}: adjoint {g[name]:.9e} fd {fd:.9e} rel {rel:.2e}")
        assert rel <= 1e-5
    # a few initial velocity and position components with the largest gradients
    rng = np.random.default_rng(0)
    for key, arr in (("g_vel0", vel), ("g_pos0", pos)):
        flat = np.abs(g[key]).ravel()
        picks = list(np.argsort(flat)[-3:]) + list(rng.integers(0, flat.size, 2))
        scale = np.max(flat)
        for idx in picks:
            i, d = divmod(int(idx), 3)
            eps = 1e-6
            a = arr.copy()
            a[i, d] += eps
            lp = _loss(p, a if key == "g_pos0" else pos, a if key == "g_vel0" else vel, sc.dt,
                       steps, target)
            a[i, d] -= 2 * eps
            lm = _loss(p, a if key == "g_pos0" else pos, a if key == "g_vel0" else vel, sc.dt,
                       steps, target)
            fd = (lp - lm) / (2 * eps)
            err = abs(g[key][i, d] - fd) / scale
            print(f"{key}[{i},{d}]: adjoint {g[key][i, d]:.6e} fd {fd:.6e} err/max {err:.2e}")
            assert err <= 1e-5


def _torch_rollout(p, pos, vel, k, mu, dt, steps):
    """All-pairs FP64 restatement of the step (no grid); masks are constants."""
    h, m = p.h, p.mass
    c6 = 315.0 / (64.0 * np.pi * h ** 9)
    c45 = 45.0 / (np.pi * h ** 6)
    g = torch.tensor([p.gravity[d] for d in range(3)], dtype=torch.float64)
    lo = torch.tensor([p.lo[d] for d in range(3)], dtype=torch.float64)
    hi = torch.tensor([p.hi[d] for d in range(3)], dtype=torch.float64)
    x, v = pos, vel
    for _ in range(steps):
        d = x[:, None, :] - x[None, :, :]
        r2 = (d * d).sum(-1)
        with torch.no_grad():
            near = r2 < h * h
            pair = near & (r2 > 0)
        rho = m * c6 * torch.where(near, (h * h - r2) ** 3, torch.zeros_like(r2)).sum(1)
        A = k * (rho - p.rest_density) / rho ** 2
        B = 1.0 / rho
        r = torch.sqrt(torch.where(pair, r2, torch.ones_like(r2)))
        hr = torch.where(pair, h - r, torch.zeros_like(r2))
        F = hr * hr / r
        P = (((A[:, None] + A[None, :]) * F)[..., None] * d).sum(1)
        V = ((hr * B[None, :])[..., None] * (v[None, :, :] - v[:, None, :])).sum(1)
        a = m * c45 * P + mu * m * c45 * B[:, None] * V + g
        v = v + dt * a
        x = x + dt * v
        with torch.no_grad():
            below, above = x < lo, x > hi
            hit = below | above
        x = torch.where(below, lo.expand_as(x), torch.where(above, hi.expand_as(x), x))
        v = torch.where(hit, -p.restitution * v, v)
    return x


@pytest.mark.parametrize("sigma,steps", [(0.3, 3), (5.0, 4)])
def test_oracle_adjoint_matches_torch_autograd(sigma, steps):
    sc = scenes.dam_break((8, 8, 8), seed=3)
    p = O.params_from_config(sc.config)
    vel = scenes.gaussian_velocities(sc.n, sigma, 5).astype(np.float64)
    pos = sc.positions.astype(np.float64)
    target = pos + 0.01 * np.random.default_rng(1).standard_normal(pos.shape)
    g = O.rollout_loss_grad(p, pos, vel, sc.dt, steps, target)
    x0 = torch.tensor(pos, requires_grad=True)
    v0 = torch.tensor(vel, requires_grad=True)
    k = torch.tensor(float(p.stiffness), dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(float(p.viscosity), dtype=torch.float64, requires_grad=True)
    xT = _torch_rollout(p, x0, v0, k, mu, sc.dt, steps)
    assert np.allclose(xT.detach().numpy(), g["x_final"], rtol=0, atol=1e-12)
    if sigma > 1.0:  # the wall clamp is exercised
        lo = np.array([p.lo[d] for d in range(3)])
        assert np.any(g["x_final"] == lo)
    loss = ((xT - torch.tensor(target)) ** 2).mean()
    assert abs(loss.item() - g["loss"]) <= 1e-12 * g["loss"]
    loss.backward()
    for name, t in (("g_pos0", x0), ("g_vel0", v0)):
        ref = t.grad.numpy()
        err = np.max(np.abs(g[name] - ref)) / np.max(np.abs(ref))
        print(f"{name}: err/max {err:.2e}")
        assert err <= 1e-9
    for name, t in (("g_stiffness", k), ("g_viscosity", mu)):
        rel = abs(g[name] - t.grad.item()) / abs(t.grad.item())
        print(f"{name}: {g[name]:.9e} vs {t.grad.item():.9e} rel {rel:.2e}")
        assert rel <= 1e-9
