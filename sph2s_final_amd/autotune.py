"""Autotuner front end: picks the launch variant of the density and force
passes for a workload and remembers it.

The library's tuner (sph_autotune / sph_batch_autotune) times every variant
of the two neighbour passes -- LDS tile size, neighbour-list capacity and grid
shape (one block per 256 particles, or a persistent grid of 1..8 blocks per
CU) -- on the solver's current state and keeps the fastest; every variant
gives bit-identical results, so tuning never changes the simulation.  This
module caches the chosen tune word in a JSON file keyed by the device and the
workload's shape (kind, particle count, cell count, scene count), so a later
run of the same workload reuses it without timing.

Tune word fields (include/sph2s.h): bits 0-3 density tile, 4-7 force
tile/list, 8-11 density grid, 12-15 force grid, 16-17 density block size,
18-19 force block size, 20-21 density and 22-23 force particles per lane
(particles per wavefront = 64 x particles per lane; single scenes only).
"""
from __future__ import annotations

import json
import os

DENSITY_TILES = (1280, 512, 1536, 2048, 768, 1024)
# (LDS tile, list capacity, list entry bits)
FORCE_VARIANTS = ((640, 16, 16), (768, 16, 16), (768, 24, 16), (768, 8, 32), (512, 16, 16),
                  (1024, 16, 16), (512, 8, 32), (640, 24, 16), (1024, 8, 32), (512, 24, 16))
GRID_BLOCKS_PER_CU = (None, 1, 2, 4, 8)  # None: one block per block-size particles
BLOCK_THREADS = (256, 128, 512)


def describe(tune: int) -> dict:
    """The launch shape a tune word selects."""
    d, f = tune & 0xF, (tune >> 4) & 0xF
    dg, fg = (tune >> 8) & 0xF, (tune >> 12) & 0xF
    db, fb = (tune >> 16) & 0x3, (tune >> 18) & 0x3
    dp, fp = (tune >> 20) & 0x3, (tune >> 22) & 0x3
    return {"density_tile": DENSITY_TILES[d], "force_tile": FORCE_VARIANTS[f][0],
            "force_list_cap": FORCE_VARIANTS[f][1], "force_list_bits": FORCE_VARIANTS[f][2],
            "density_blocks_per_cu": GRID_BLOCKS_PER_CU[dg],
            "force_blocks_per_cu": GRID_BLOCKS_PER_CU[fg],
            "density_block_threads": BLOCK_THREADS[db], "force_block_threads": BLOCK_THREADS[fb],
            "density_particles_per_lane": 1 + dp, "force_particles_per_lane": 1 + fp}


def workload_key(kind: str, particles: int, cells: int, scenes: int = 1, device: str = "") -> str:
    return f"{device}|{kind}|n={int(particles)}|cells={int(cells)}|scenes={int(scenes)}"


def _device_name() -> str:
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_name(0)
    except Exception:
        pass
    return "unknown"


def load_cache(path: str) -> dict:
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_cache(path: str, cache: dict) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(solver, kind: str, particles: int, cells: int, dt: float, scenes: int = 1,
         cache_path: str | None = None, reps: int = 5, retune: bool = False) -> dict:
    """Applies a cached tune word for this workload, or runs the tuner on the
    solver's current state and caches the result.  ``solver`` is a Solver or
    BatchSolver.  Returns {"tune", "cached", "launch", "ms"}."""
    key = workload_key(kind, particles, cells, scenes, _device_name())
    cache = load_cache(cache_path) if cache_path else {}
    if not retune and key in cache:
        word = int(cache[key]["tune"])
        solver.set_tune(word)
        return {"tune": word, "cached": True, "launch": describe(word),
                "ms": cache[key].get("ms")}
    word, ms = solver.autotune(dt, reps)
    if cache_path:
        cache[key] = {"tune": word, "ms": ms, "launch": describe(word)}
        save_cache(cache_path, cache)
    return {"tune": word, "cached": False, "launch": describe(word), "ms": ms}
