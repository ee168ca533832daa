"""CPU tests of the C ABI surface: the library loads, exports every symbol
include/sph2s.h declares, and its host-only helpers (grid, configuration,
scene generator, argument validation) behave; no GPU compute is called."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

import sph2s_final_amd as S
from sph2s_final_amd import _capi as C
from sph2s_final_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    declared = _declared_symbols()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sph_[a-z0-9_]+)", out))
    assert declared <= exported, declared - exported
    bound = {name for name, _, _ in C.SIGNATURES}
    assert declared <= bound, declared - bound


def test_version_and_device_count_without_gpu():
    assert b"gfx950" in C.lib().sph_version()
    assert S.device_count() >= 0


def test_make_grid_matches_reference_rule():
    g = S.Grid.MakeGrid((0.0, 0.0, 0.0), (3.125, 1.0, 0.5), 0.0457)
    want = [math.ceil(np.float32(v) / np.float32(0.0457)) for v in (3.125, 1.0, 0.5)]
    assert [g.dims[d] for d in range(3)] == want == [69, 22, 11]
    assert S.Grid.ComputeNumGridCells(g) == 69 * 22 * 11


def test_make_configuration_formula():
    cfg = S.SolverConfiguration.MakeConfiguration((-0.1, -0.1, -0.2), (1.7, 1.6, 0.6), 0.05,
                                                  20000, 35, 1000.0, 30.0, 5.0, 88.1472, 0.8)
    h = (0.75 * 0.05 * 35 / (math.pi * 20000)) ** (1 / 3)
    assert cfg.h == pytest.approx(h, rel=1e-6)
    assert cfg.mass == pytest.approx(1000.0 * 0.05 / 20000, rel=1e-6)
    assert cfg.grid.spacing == pytest.approx(cfg.h)
    assert cfg.tension == pytest.approx(0.8)


def test_lattice_generator_is_seeded_and_jitter_bounded():
    a = scenes.lattice((5, 4, 3), seed=7)
    b = scenes.lattice((5, 4, 3), seed=7)
    c = scenes.lattice((5, 4, 3), seed=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    s = scenes.SPACING
    i = np.arange(60)
    base = np.stack([(i % 5 + 0.5) * s, ((i // 5) % 4 + 0.5) * s, (i // 20 + 0.5) * s], 1)
    assert np.max(np.abs(a - base)) <= 0.01 * s * (1 + 1e-5)
    assert np.max(np.abs(a - base)) > 0.005 * s


def test_lattice_generator_matches_splitmix64_definition():
    # SplitMix64 stream, three draws per particle (x, y, z), u = (z >> 11) * 2^-53.
    def splitmix(state):
        state = (state + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return state, z ^ (z >> 31)

    st = 3
    s = float(np.float32(scenes.SPACING))  # the C ABI takes float spacing / jitter
    jf = float(np.float32(0.01))
    got = scenes.lattice((2, 2, 1), seed=3)
    for p in range(4):
        idx = (p % 2, p // 2, 0)
        for d in range(3):
            st, z = splitmix(st)
            u = (z >> 11) * 2.0 ** -53
            want = np.float32((idx[d] + 0.5) * s + (2 * u - 1) * jf * s)
            assert got[p, d] == want


def test_dam_break_scenes_match_benchmark_definition():
    sc = scenes.dam_break_1m()
    assert sc.n == 1_000_000
    assert [sc.config.grid.dims[d] for d in range(3)] == [69, 69, 69]
    assert [sc.config.domain_hi[d] for d in range(3)] == [3.125] * 3
    assert sc.config.mass == pytest.approx(1000.0 * (0.0457 / 2) ** 3, rel=1e-6)
    assert (sc.config.stiffness, sc.config.viscosity) == (3.0, 3.5)
    small = scenes.dam_break_32k()
    assert small.n == 32768 and [small.config.grid.dims[d] for d in range(3)] == [22] * 3


def test_gaussian_velocity_statistics():
    v = scenes.gaussian_velocities(200_000, 0.05, 11)
    assert abs(v.mean()) < 1e-3
    assert v.std() == pytest.approx(0.05, rel=0.01)


def test_invalid_configuration_is_rejected_before_any_device_work():
    sc = scenes.dam_break((4, 4, 4))
    cfg = sc.config
    cfg.h = -1.0
    h = ctypes.c_void_p()
    rc = C.lib().sph_create(ctypes.byref(cfg), sc.positions.ctypes.data_as(ctypes.c_void_p), None,
                            sc.n, sc.n, ctypes.byref(h))
    assert rc == C.SPH_ERR_INVALID_ARGUMENT
    assert b"h must be > 0" in C.lib().sph_last_error()


def test_solver_without_device_raises(has_gpu):
    if has_gpu:
        pytest.skip("a device is visible")
    sc = scenes.dam_break((4, 4, 4))
    with pytest.raises(S.SPHError) as e:
        S.Solver(S.ParticleData(sc.n, sc.positions), None, sc.config)
    assert e.value.code == C.SPH_ERR_NO_DEVICE


def test_particle_box_and_union_mirror_reference():
    g = S.Grid.MakeGrid((0.04, 0.04, 0.0), (0.1, 0.1, 0.05), 0.0075)
    box = S.ParticleData.CreateParticleBox(g)
    assert box.NumParticles == S.Grid.ComputeNumGridCells(g)
    assert np.allclose(box.positions[1] - box.positions[0], [0.0075, 0, 0], atol=1e-7)
    u = S.ParticleData.Union(box, box)
    assert u.NumParticles == 2 * box.NumParticles and u.MaxParticles == 2 * box.MaxParticles


def test_make_configuration_keeps_wall_constants():
    cfg = S.SolverConfiguration.MakeConfiguration((-0.1, -0.1, -0.2), (1.7, 1.6, 0.6), 0.05,
                                                  20000, 35, 1000.0, 30.0, 5.0, 88.1472, 0.8)
    assert cfg.speed_sound == pytest.approx(88.1472)
    assert cfg.boundary_mass == cfg.mass          # Solver.h:38
    d = scenes.dam_break_32k().config
    assert d.speed_sound == pytest.approx(88.1472) and d.boundary_mass == d.mass


def test_cpp_facade_compiles_and_links(tmp_path):
    """The header-only C++ facade builds a reference-style program (fluid box,
    canvas walls, Solver, Advance) and links against libsph2s.so."""
    src = tmp_path / "facade.cpp"
    src.write_text(r'''
#include "sph2s/Solver.hpp"
using namespace sph2s;
int main() {
    Grid fluidGrid = Grid::MakeGrid(make_float3(0.04f, 0.04f, 0.0f), make_float3(0.1f, 0.1f, 0.05f), 0.0075f);
    ParticleData* fluid = ParticleData::CreateParticleBox(fluidGrid);
    ParticleData* a = ParticleData::CreateParticleBoxCanvas(
        Grid::MakeGrid(make_float3(0.0f, 0.0f, 0.0f), make_float3(0.2f, 0.2f, 0.1f), 0.0075f), 3);
    ParticleData* b = ParticleData::CreateParticleBoxCanvas(
        Grid::MakeGrid(make_float3(0.12f, 0.0f, 0.03f), make_float3(0.135f, 0.2f, 0.045f), 0.0075f), 3);
    ParticleData* walls = ParticleData::Union(a, b);
    SolverConfiguration c = SolverConfiguration::MakeConfiguration(
        make_float3(-0.05f, -0.05f, -0.05f), make_float3(0.25f, 0.25f, 0.14f),
        Grid::ComputeVolume(fluidGrid), fluid->MaxParticles, 35, 1000.0f, 30.0f, 5.0f, 88.1472f, 0.8f);
    if (sph_device_count() == 0) return 0;
    Solver s(fluid, walls, &c);
    s.Advance(0.001f);
    s.Synchronize();
    return fluid->NumParticles > 0 ? 0 : 1;
}
''')
    exe = tmp_path / "facade"
    res = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}",
                          str(src), "-o", str(exe), f"-L{os.path.dirname(C.LIB_PATH)}", "-lsph2s",
                          f"-Wl,-rpath,{os.path.dirname(C.LIB_PATH)}"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_cpp_two_scale_facade_compiles_and_links(tmp_path):
    """The two-scale reference's program shape compiles against the facade's
    sph2s::two_scale names: MakeConfiguration with blendIncrement as its 11th
    argument and the reference's configuration members, Solver(fluid,
    fluidHigh, boundary, &config), and ParticleData's renderer pointers."""
    src = tmp_path / "facade2.cpp"
    src.write_text(r'''
#include "sph2s/Solver.hpp"
using namespace sph2s::two_scale;
int main() {
    Grid particleGrid = Grid::MakeGrid(make_float3(0.02f, 0.04f, -0.05f),
                                       make_float3(0.35f, 0.75f, 0.45f), 0.0075f);
    ParticleData* fluid = ParticleData::CreateParticleBox(particleGrid);
    ParticleData* fluidHigh = new ParticleData(fluid->MaxParticles * 8);
    fluidHigh->NumParticles = 0;
    SolverConfiguration config = SolverConfiguration::MakeConfiguration(
        make_float3(-0.1f, -0.1f, -0.2f), make_float3(1.7f, 1.6f, 0.6f),
        Grid::ComputeVolume(particleGrid), fluid->MaxParticles, 35, 1000.0f, 30.0f, 5.0f,
        88.1472f, 0.8f, 0.0125f);
    if (!(config.EffectiveRadius[1] == 0.5f * config.EffectiveRadius[0])) return 2;
    if (!(config.FluidParticleMass[1] == 1.0f / 8.0f * config.FluidParticleMass[0])) return 3;
    if (config.BlendIncrement != 0.0125f || config.Grid[1].spacing != config.EffectiveRadius[1])
        return 4;
    if (sph_device_count() == 0) return 0;
    Solver s(fluid, fluidHigh, nullptr, &config);
    s.Advance(0.001f);
    float* xyz = fluid->dPositions;
    float* col = fluid->dColorValues;
    int* st = fluidHigh->dStates;
    return (xyz && col && st) ? 0 : 1;
}
''')
    exe = tmp_path / "facade2"
    res = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}",
                          str(src), "-o", str(exe), f"-L{os.path.dirname(C.LIB_PATH)}", "-lsph2s",
                          f"-Wl,-rpath,{os.path.dirname(C.LIB_PATH)}"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, (run.returncode, run.stderr)


FACADE_DEMO = os.path.join(ROOT, "tests", "_bin", "facade_demo")


def _read_sets(path, n_sets):
    """facade_demo's output: NumParticles of each set, then per set the
    renderer buffers read through ParticleData's device pointers: dPositions
    (3n float32), dColorValues (n float32), dStates (n int32)."""
    raw = path.read_bytes()
    counts = [int(v) for v in np.frombuffer(raw[:4 * n_sets], np.uint32)]
    off, sets = 4 * n_sets, []
    for n in counts:
        xyz = np.frombuffer(raw[off:off + 12 * n], np.float32).reshape(n, 3)
        off += 12 * n
        col = np.frombuffer(raw[off:off + 4 * n], np.float32)
        off += 4 * n
        st = np.frombuffer(raw[off:off + 4 * n], np.int32)
        off += 4 * n
        sets.append((xyz, col, st))
    assert off == len(raw)
    return counts, sets


def _run_demo(*args):
    assert os.path.exists(FACADE_DEMO), "tests/_bin/facade_demo is not built (build())"
    return subprocess.run([FACADE_DEMO, *[str(a) for a in args]], capture_output=True, text=True,
                          timeout=120)


@pytest.mark.gpu
def test_cpp_facade_program_steps_like_the_python_path(tmp_path):
    """A reference-style host program written against the C++ facade only
    (tests/facade_demo.cpp: fluid box, canvas tank and pillar walls,
    MakeConfiguration, Solver::Advance once per step) runs on the GPU.  The
    positions it copies from fluid->dPositions (3 floats per particle, the
    renderer's layout) equal, bit for bit, those of the same scene
    (walled_tank_small, built with the Python mirror of the same classes)
    stepped through the Python path; dColorValues and dStates are 0, as the
    single-scale reference never writes them."""
    steps = 6
    out = tmp_path / "xyz.bin"
    print(f"case facade program steps={steps}", flush=True)
    res = _run_demo(steps, out)
    assert res.returncode == 0, res.stderr
    (n,), [(xyz, col, st)] = _read_sets(out, 1)

    sc = scenes.walled_tank_small()  # the same scene through the Python mirror classes
    assert n == sc.n
    with S.Solver(S.ParticleData(sc.n, sc.positions), S.ParticleData(len(sc.walls), sc.walls),
                  sc.config) as s:
        for _ in range(steps):
            s.Advance(0.001)
        s.synchronize()
        ref = s.positions()
        assert s.device_errors() == 0
    assert np.all(np.isfinite(ref))
    assert not np.array_equal(ref, sc.positions)  # the fluid moved
    assert np.array_equal(xyz, ref)
    assert not col.any() and not st.any()


@pytest.mark.gpu
def test_cpp_facade_two_scale_program_steps_like_the_python_path(tmp_path):
    """The facade's TwoScaleSolver (the reference's two-scale Solver(fluid,
    fluidHigh, boundary, cfg) plus a split region) in the facade program:
    after 6 steps (splits in the region box) its (low, high) counts and both
    sets' dPositions equal, bit for bit, the same scene stepped through the
    Python TwoScaleSolver; dStates hold the fixed coupling's 0 (low) and 4
    (high)."""
    steps = 6
    out = tmp_path / "two.bin"
    print(f"case facade two-scale program steps={steps}", flush=True)
    res = _run_demo("two", steps, out)
    assert res.returncode == 0, res.stderr
    (n_low, n_high), [(low, _, st_l), (high, _, st_h)] = _read_sets(out, 2)

    grid = S.Grid.MakeGrid((0.02, 0.02, 0.02), (0.2, 0.2, 0.2), 0.01)
    fluid = S.ParticleData.CreateParticleBox(grid)
    cfg = S.SolverConfiguration.MakeConfiguration(
        (0.0, 0.0, 0.0), (0.4, 0.4, 0.4), S.Grid.ComputeVolume(grid), fluid.MaxParticles, 35,
        1000.0, 30.0, 5.0, 88.1472, 0.0)
    t = S.TwoScaleSolver(cfg, (0.02, 0.02, 0.02), (0.1, 0.1, 0.1), fluid.positions, None,
                         max_low=fluid.MaxParticles, max_high=8 * fluid.MaxParticles)
    try:
        for _ in range(steps):
            t.Advance(0.001)
        t.synchronize()
        c = t.counts()
        assert t.device_errors() == 0
        assert (n_low, n_high) == c[:2]
        assert n_high > 0 and n_low < fluid.MaxParticles  # the region split
        assert np.array_equal(low, t.download(0, S.FIELD_POSITION))
        assert np.array_equal(high, t.download(1, S.FIELD_POSITION))
        assert not st_l.any() and np.all(st_h == 4)
    finally:
        t.close()


@pytest.mark.gpu
def test_cpp_facade_reference_two_scale_program(tmp_path):
    """The reference's own two-scale program shape (TWO_SCALE/main.cpp:150-254)
    in sph2s::two_scale names: MakeConfiguration with the blend increment as
    its 11th argument, Solver(fluid, fluidHigh, boundary, &config), splits
    requested by id after the first step.  Counts and both sets' renderer
    buffers (dPositions, dColorValues, dStates) equal, bit for bit, the same
    run through the Python TwoScaleSolver with the blended coupling,
    reference split offsets, no merging, the 0.8 h / dt velocity cap and no
    split region."""
    steps = 8
    out = tmp_path / "ref2.bin"
    print(f"case facade reference two-scale program steps={steps}", flush=True)
    res = _run_demo("ref2", steps, out)
    assert res.returncode == 0, res.stderr
    (n_low, n_high), sets = _read_sets(out, 2)

    grid = S.Grid.MakeGrid((0.02, 0.02, 0.02), (0.14, 0.2, 0.14), 0.01)
    fluid = S.ParticleData.CreateParticleBox(grid)
    walls = S.ParticleData.CreateParticleBoxCanvas(
        S.Grid.MakeGrid((0.0, 0.0, 0.0), (0.3, 0.3, 0.3), 0.01), 2)
    cfg = S.SolverConfiguration.MakeConfiguration(
        (-0.05, -0.05, -0.05), (0.35, 0.35, 0.35), S.Grid.ComputeVolume(grid), fluid.MaxParticles,
        35, 1000.0, 30.0, 5.0, 88.1472, 0.8)
    t = S.TwoScaleSolver(cfg, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), fluid.positions, None,
                         max_low=fluid.MaxParticles, max_high=8 * fluid.MaxParticles)
    try:
        t.set_boundary(walls.host_positions())
        t.set_options(coupling=S.SPH2_COUPLING_BLENDED, split_offsets=S.SPH2_SPLIT_REFERENCE,
                      merge=False, blend_increment=0.0125, velocity_cap=0.8)
        t.set_debug(S.SPH_DEBUG_COLOR)
        for k in range(steps):
            t.Advance(0.001)
            if k == 0:
                t.mark_split(np.arange(0, fluid.MaxParticles, 7))
        t.synchronize()
        assert t.device_errors() == 0
        c = t.counts()
        assert (n_low, n_high) == c[:2]
        assert n_high > 8 * 300  # the marked particles split
        for r, (xyz, col, st) in enumerate(sets):
            assert np.array_equal(xyz, t.download(r, S.FIELD_POSITION)), r
            assert np.array_equal(col, t.download(r, S.FIELD_COLOR)), r
            assert np.array_equal(st.view(np.uint32), t.download(r, S.FIELD_STATE)), r
        assert np.any(sets[0][1] != 0)  # colour values are written
        assert np.all((sets[1][2] & 4) == 4)  # high-set state words
    finally:
        t.close()


@pytest.mark.gpu
def test_split_beyond_capacity_is_reported(tmp_path):
    """A split that exceeds max_high drops the excess children (writes stay
    within the capacity) and raises SPH_DEVERR_CAPACITY: synchronize reports
    SPH_ERR_CAPACITY until the bits are read, the count stays at the capacity."""
    sc = scenes.two_scale_dam_break(12, seed=3)
    print(f"case split overflow n={sc.n} max_high=64", flush=True)
    t = S.TwoScaleSolver(sc.config, sc.region_lo, sc.region_hi, sc.positions, sc.velocities,
                         max_low=sc.n, max_high=64)
    try:
        t.Advance(sc.dt)
        with pytest.raises(C.SPHError) as e:
            t.synchronize()
        assert e.value.code == C.SPH_ERR_CAPACITY
        assert t.counts()[1] == 64
        assert t.device_errors() & C.SPH_DEVERR_CAPACITY
        t.synchronize()  # the bits were cleared
    finally:
        t.close()


@pytest.mark.gpu
def test_nonfinite_state_is_reported():
    """A NaN velocity (nothing in a step indexes by velocity) turns the
    integrated state of that particle and its neighbours non-finite: the
    force pass raises SPH_DEVERR_NONFINITE and synchronize reports
    SPH_ERR_NONFINITE.  A finite run reports nothing."""
    sc = scenes.dam_break((10, 10, 10), seed=2)
    vel = sc.velocities.copy()
    vel[17, 0] = np.nan
    print(f"case non-finite velocity n={sc.n}", flush=True)
    with S.Solver(S.ParticleData(sc.n, sc.positions, vel), None, sc.config) as s:
        s.Advance(sc.dt)
        with pytest.raises(C.SPHError) as e:
            s.synchronize()
        assert e.value.code == C.SPH_ERR_NONFINITE
        assert s.device_errors() == C.SPH_DEVERR_NONFINITE
        assert not np.all(np.isfinite(s.positions()))
    with S.Solver(S.ParticleData(sc.n, sc.positions, sc.velocities), None, sc.config) as s:
        s.Advance(sc.dt, 3)
        s.synchronize()
        assert s.device_errors() == 0
