"""Results must not depend on what earlier kernels left in LDS.

LDS is not cleared between kernels, so a slot that a kernel reads without
having written it holds whatever the previous workgroup on that CU left
there, NaN bit patterns included.  The staged scans load whole groups of
candidates, padded past the end of a staged range, and must keep those loads
out of every sum (masked, or finite copies multiplied by a zero weight).
Each test runs a pass twice from the same state, once after kernels that leave
NaN patterns in the LDS of every CU (torch's sort and reductions of an all-NaN
tensor stage their data there), and requires bit-identical results.  This is
heuristic (a kernel's workgroups need not land where the NaNs were left), but
it exercises the case that once turned every adjoint gradient NaN.
"""
import numpy as np
import pytest
import torch

from sph2s_final_amd import (FIELD_DENSITY, FIELD_POSITION, FIELD_VELOCITY, ParticleData,
                             Solver, TwoScaleSolver, scenes)
from sph2s_final_amd.autograd import Rollout

pytestmark = pytest.mark.gpu


def _poison_lds():
    x = torch.full((1 << 22,), float("nan"), device="cuda")
    torch.sort(x)
    y = torch.full((2048, 4096), float("nan"), device="cuda")
    y.sum(dim=0)
    y.amax(dim=1)
    torch.cuda.synchronize()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")


def test_single_step_ignores_stale_lds():
    sc = scenes.dam_break((20, 20, 20), seed=11, velocity_sigma=0.3)
    print(f"case stale lds single n={sc.n}", flush=True)
    outs = []
    for poison in (False, True):
        with Solver(ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config) as s:
            s.Advance(sc.dt, 3)
            s.synchronize()
            if poison:
                _poison_lds()
            s.advance_eager(sc.dt, 2)
            s.synchronize()
            assert s.device_errors() == 0
            outs.append([s.positions(), s.velocities(), s.densities()])
    for a, b in zip(*outs):
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b)


def test_two_scale_step_ignores_stale_lds():
    sc = scenes.two_scale_dam_break(16, seed=12)
    print(f"case stale lds two-scale n={sc.n}", flush=True)
    outs = []
    for poison in (False, True):
        with TwoScaleSolver(sc.config, sc.region_lo, sc.region_hi, sc.positions, sc.velocities,
                            max_low=sc.n, max_high=8 * sc.n) as t:
            t.Advance(sc.dt, 3)
            t.synchronize()
            if poison:
                _poison_lds()
            t.advance_eager(sc.dt, 2)
            t.synchronize()
            assert t.device_errors() == 0
            outs.append([t.download(r, f) for r in (0, 1)
                         for f in (FIELD_POSITION, FIELD_VELOCITY, FIELD_DENSITY)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_adjoint_gradients_ignore_stale_lds():
    bs = scenes.batched_dam_breaks(3, seed=21, n_min=2000, n_max=9000)
    for s in bs.scenes:
        s.velocities = scenes.gaussian_velocities(s.n, 0.1, 300 + s.n)
    pos = _dev(np.concatenate([s.positions for s in bs.scenes]))
    vel = _dev(np.concatenate([s.velocities for s in bs.scenes]))
    print(f"case stale lds adjoint counts={bs.counts.tolist()}", flush=True)
    grads = []
    with Rollout([s.config for s in bs.scenes], bs.counts, bs.dt, 12, 4) as r:
        seed = None
        for poison in (False, True):
            x = r.forward(pos, vel)
            if seed is None:
                seed = torch.full_like(x, 1e-4)
            if poison:
                _poison_lds()
            gp, gv, gk, gm = r.backward(seed)
            torch.cuda.synchronize()
            grads.append([gp.cpu().numpy(), gv.cpu().numpy(), np.asarray(gk), np.asarray(gm)])
        assert r.device_errors() == 0
    for a, b in zip(*grads):
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b)
