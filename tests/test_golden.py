"""The golden fixtures of tests/golden (written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture (integers exactly, floating
point to 1e-12 relative).  GPU: the HIP path on the fixture inputs matches the
fixture outputs to the parity bars (integers bit-exact; density and
acceleration 1e-5; split/merge positions 1e-6; gradients 1e-3 of max).
"""
import os

import numpy as np
import pytest

import oracle as O
from sph2s_final_amd import (FIELD_ACCEL, FIELD_CELL_END, FIELD_CELL_START, FIELD_ID,
                             FIELD_POSITION, FIELD_SORTED_HASH, FIELD_SORTED_INDEX, FIELD_TAG,
                             SPH_DEBUG_ACCEL, Grid, ParticleData, Solver, TwoScaleSolver, scenes)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _close(a, b, rtol=1e-12):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(np.asarray(a, np.float64) - b)) <= rtol * scale


def _step_scene():
    return scenes.dam_break((6, 6, 6), seed=21, velocity_sigma=0.4, velocity_seed=22)


def _two_scale_scene():
    return scenes.two_scale_dam_break(8, seed=23)


# ---------------------------------------------------------------- CPU
def test_scene_generator_reproduces_fixture_inputs():
    f = _load("step_6x6x6.npz")
    sc = _step_scene()
    assert np.array_equal(sc.positions, f["pos"]) and np.array_equal(sc.velocities, f["vel"])
    g = _load("two_scale_8.npz")
    assert np.array_equal(_two_scale_scene().positions, g["pos"])


def test_oracle_reproduces_step_fixture():
    f = _load("step_6x6x6.npz")
    sc = _step_scene()
    r = O.step(O.params_from_config(sc.config), f["pos"], f["vel"], float(f["dt"]))
    for k in ("hash", "perm", "cell_start", "cell_end"):
        assert np.array_equal(r[k], f[k]), k
    for k, fk in (("rho", "rho"), ("acc", "acc"), ("pos", "pos_out"), ("vel", "vel_out")):
        assert _close(r[k], f[fk]), k


def test_oracle_reproduces_two_scale_fixture():
    f = _load("two_scale_8.npz")
    sc = _two_scale_scene()
    p = O.params_from_config(sc.config)
    gh = Grid.MakeGrid((0.0, 0.0, 0.0), [sc.config.domain_hi[d] for d in range(3)],
                       0.5 * sc.config.h)
    o = O.OracleTwoScale(p, gh, tuple(f["region_lo"]), tuple(f["region_hi"]), sc.n)
    o.set_state(f["pos"], f["vel"])
    o.split_merge()
    assert np.array_equal(np.array(o.counts(), np.uint32), f["split_counts"])
    for r in (0, 1):
        assert np.array_equal(o.get(r, "id"), f[f"split_{r}_id"])
        assert _close(o.get(r, "pos"), f[f"split_{r}_pos"])
    assert np.array_equal(o.get(1, "tag"), f["split_1_tag"])
    o.step(float(f["dt"]))
    assert np.array_equal(np.array(o.counts(), np.uint32), f["step_counts"])
    for r in (0, 1):
        for k in ("id", "hash", "perm"):
            assert np.array_equal(o.get(r, k), f[f"step_{r}_{k}"]), (r, k)
        for k in ("pos", "vel", "rho"):
            assert _close(o.get(r, k), f[f"step_{r}_{k}"]), (r, k)


def test_oracle_reproduces_adjoint_fixture():
    f = _load("adjoint_5x5x5.npz")
    sc = scenes.dam_break((5, 5, 5), seed=24)
    g = O.rollout_loss_grad(O.params_from_config(sc.config), f["pos"].astype(np.float64),
                            f["vel"].astype(np.float64), float(f["dt"]), int(f["steps"]),
                            f["target"])
    assert abs(g["loss"] - float(f["loss"])) <= 1e-12 * float(f["loss"])
    for k in ("x_final", "g_pos0", "g_vel0"):
        assert _close(g[k], f[k], 1e-10), k
    for k in ("g_stiffness", "g_viscosity"):
        assert abs(g[k] - float(f[k])) <= 1e-10 * abs(float(f[k])), k


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_step_matches_fixture():
    f = _load("step_6x6x6.npz")
    sc = _step_scene()
    print("case golden step_6x6x6", flush=True)
    with Solver(ParticleData(sc.n, f["pos"], f["vel"]), None, sc.config) as s:
        s.set_debug(SPH_DEBUG_ACCEL)
        s.Advance(float(f["dt"]))
        s.synchronize()
        assert np.array_equal(s.download(FIELD_SORTED_HASH), f["hash"])
        assert np.array_equal(s.download(FIELD_SORTED_INDEX), f["perm"])
        assert np.array_equal(s.download(FIELD_CELL_START), f["cell_start"])
        assert np.array_equal(s.download(FIELD_CELL_END), f["cell_end"])
        rho_err = np.max(np.abs(s.densities() - f["rho"]) / f["rho"])
        acc_err = np.max(np.abs(s.download(FIELD_ACCEL) - f["acc"])) / np.max(np.abs(f["acc"]))
        print(f"  rho {rho_err:.2e} acc {acc_err:.2e}", flush=True)
        assert rho_err <= 1e-5 and acc_err <= 1e-5
        assert np.max(np.abs(s.positions() - f["pos_out"])) <= 1e-6


@pytest.mark.gpu
def test_gpu_two_scale_split_matches_fixture():
    f = _load("two_scale_8.npz")
    sc = _two_scale_scene()
    print("case golden two_scale_8", flush=True)
    with TwoScaleSolver(sc.config, tuple(f["region_lo"]), tuple(f["region_hi"]), f["pos"],
                        f["vel"], max_low=sc.n, max_high=8 * sc.n) as t:
        t.apply_split_merge()
        assert np.array_equal(np.array(t.counts()[:2], np.uint32), f["split_counts"][:2])
        for r in (0, 1):
            assert np.array_equal(t.download(r, FIELD_ID), f[f"split_{r}_id"])
            assert np.max(np.abs(t.download(r, FIELD_POSITION) - f[f"split_{r}_pos"]),
                          initial=0.0) <= 1e-6
        assert np.array_equal(t.download(1, FIELD_TAG), f["split_1_tag"])


@pytest.mark.gpu
def test_gpu_gradients_match_fixture():
    import torch
    from sph2s_final_amd.autograd import Rollout
    f = _load("adjoint_5x5x5.npz")
    sc = scenes.dam_break((5, 5, 5), seed=24)
    print("case golden adjoint_5x5x5", flush=True)
    seed = 2.0 * (f["x_final"] - f["target"]) / f["target"].size
    with Rollout([sc.config], [sc.n], float(f["dt"]), int(f["steps"]), 2) as r:
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()  # noqa: E731
        r.forward(dev(f["pos"]), dev(f["vel"]))
        gp, gv, gk, gm = r.backward(dev(seed))
        torch.cuda.synchronize()
    for name, got in (("g_pos0", gp.cpu().numpy()), ("g_vel0", gv.cpu().numpy())):
        err = np.max(np.abs(got - f[name])) / np.max(np.abs(f[name]))
        print(f"  {name} {err:.2e}", flush=True)
        assert err <= 1e-3
    assert abs(gk[0] - float(f["g_stiffness"])) <= 1e-3 * abs(float(f["g_stiffness"]))
    assert abs(gm[0] - float(f["g_viscosity"])) <= 1e-3 * abs(float(f["g_viscosity"]))
