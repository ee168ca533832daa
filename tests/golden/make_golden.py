"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no tests, golden vectors or fixtures, so these vectors are
outputs of the oracle (oracle/), whose functions are pinned independently by
closed forms, brute force, invariants, finite differences and a torch float64
restatement (tests/test_oracle.py, tests/test_adjoint_oracle.py).  They freeze
the oracle's results on small inputs so that a later change to the oracle or
to the GPU path is caught against fixed data:

  step_6x6x6.npz       one single-scale step of a 6^3 dam break with random
                       velocities: inputs, sorted hashes, permutation, cell
                       ranges, density, acceleration, integrated state
  two_scale_8.npz      one split/merge pass and one coupled two-scale step of
                       an 8^3 dam break: counts, ids, tags, positions
  adjoint_5x5x5.npz    a 4-step rollout's loss and its gradients with respect
                       to initial positions, velocities, stiffness, viscosity

Run from the repository root:  python tests/golden/make_golden.py
(numpy .npz without pickled objects; load with numpy.load(allow_pickle=False)).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from sph2s_final_amd import Grid, scenes  # noqa: E402


def step_fixture():
    sc = scenes.dam_break((6, 6, 6), seed=21, velocity_sigma=0.4, velocity_seed=22)
    p = O.params_from_config(sc.config)
    r = O.step(p, sc.positions, sc.velocities, sc.dt)
    np.savez(os.path.join(HERE, "step_6x6x6.npz"), pos=sc.positions, vel=sc.velocities,
             dt=np.float64(sc.dt), hash=r["hash"], perm=r["perm"], cell_start=r["cell_start"],
             cell_end=r["cell_end"], rho=r["rho"], acc=r["acc"], pos_out=r["pos"],
             vel_out=r["vel"])


def two_scale_fixture():
    sc = scenes.two_scale_dam_break(8, seed=23)
    p = O.params_from_config(sc.config)
    gh = Grid.MakeGrid((0.0, 0.0, 0.0), [sc.config.domain_hi[d] for d in range(3)],
                       0.5 * sc.config.h)
    o = O.OracleTwoScale(p, gh, sc.region_lo, sc.region_hi, sc.n)
    o.set_state(sc.positions, sc.velocities)
    o.split_merge()
    split = {f"split_{r}_{k}": o.get(r, k) for r in (0, 1) for k in ("pos", "vel", "id")}
    split["split_1_tag"] = o.get(1, "tag")
    split_counts = np.array(o.counts(), np.uint32)
    o.step(sc.dt)
    stepped = {f"step_{r}_{k}": o.get(r, k) for r in (0, 1)
               for k in ("pos", "vel", "id", "rho", "hash", "perm")}
    np.savez(os.path.join(HERE, "two_scale_8.npz"), pos=sc.positions, vel=sc.velocities,
             region_lo=np.array(sc.region_lo), region_hi=np.array(sc.region_hi),
             dt=np.float64(sc.dt), split_counts=split_counts,
             step_counts=np.array(o.counts(), np.uint32), **split, **stepped)


def adjoint_fixture():
    sc = scenes.dam_break((5, 5, 5), seed=24)
    vel = scenes.gaussian_velocities(sc.n, 0.2, 25).astype(np.float64)
    p = O.params_from_config(sc.config)
    pos = sc.positions.astype(np.float64)
    target = pos + 0.003 * np.random.default_rng(26).standard_normal(pos.shape)
    g = O.rollout_loss_grad(p, pos, vel, sc.dt, 4, target)
    np.savez(os.path.join(HERE, "adjoint_5x5x5.npz"), pos=sc.positions, vel=vel.astype(np.float32),
             target=target, dt=np.float64(sc.dt), steps=np.int64(4), loss=np.float64(g["loss"]),
             x_final=g["x_final"], g_pos0=g["g_pos0"], g_vel0=g["g_vel0"],
             g_stiffness=np.float64(g["g_stiffness"]), g_viscosity=np.float64(g["g_viscosity"]))


if __name__ == "__main__":
    step_fixture()
    two_scale_fixture()
    adjoint_fixture()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
