"""Scenes from .sim scripts on the HIP path (fs_world_*) against the CPU
checker driven through the same frame loop as SIMULATION_WORLD
(SIMULATION_WORLD.h:273-330): the level set's ComputeCurvatures, then one
CFL step (dt = CFL / rate with |g_y|) or num_substeps steps of the script's
dt, with the level-set sources applied at frame 0 and in every Sourcing
phase.

Bit-exact: the initial level set (the air/water sphere of radius 1/3 and
the sources' min/max), the first frame's CFL step.  Frames: within 1e-8 when
every solve reaches 1e-11 relative residual."""
import os

import numpy as np
import pytest

from fluid_simulator_backup_amd import _abi
from helpers import rel_err

pytestmark = pytest.mark.gpu

SCRIPTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scripts")
TIGHT = dict(stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-11)


def _checker(oracle_lib, hip_lib, path, **opts):
    """The CPU checker set up as the world sets itself up."""
    info, src = _abi.parse_world(hip_lib, path, **opts)
    g = info.grid
    s = _abi.Solver(oracle_lib, "fso_", info.params, g)
    dx, dy, dz = ((g.x_max - g.x_min) / (g.nx - 1), (g.y_max - g.y_min) / (g.ny - 1),
                  (g.z_max - g.z_min) / (g.nz - 1))
    z, y, x = np.meshgrid(g.z_min + np.arange(g.nz) * dz, g.y_min + np.arange(g.ny) * dy,
                          g.x_min + np.arange(g.nx) * dx, indexing="ij")
    phi = np.sqrt(x * x + y * y + z * z) - 1.0 / 3  # SIMULATION_WORLD.cpp:116-128
    nx, ny, nz = g.nx, g.ny, g.nz
    s.set_fields(phi, np.zeros((nz, ny, nx + 1)), np.zeros((nz, ny + 1, nx)), np.zeros((nz + 1, ny, nx)))
    if src:
        s.set_sources([dict(center=tuple(a.center), radius=a.radius, half_width=a.half_width, mode=a.mode,
                            start_frame=a.start_frame, stop_frame=a.stop_frame) for a in src])
        s.set_frame(0)
        s.apply_sources()
    return s, info, phi


def _checker_frame(s, info, frame):
    s.set_frame(frame)
    s.compute_curvatures()
    if info.cfl == 0 or info.frame_rate == 0:
        for _ in range(info.num_substeps):
            s.step(info.dt)
        return info.dt
    dt = s.cfl_dt(info.cfl, 0)
    s.step(dt)
    return dt


@pytest.mark.parametrize("script,frames", [("air_water_bubble.sim", 3), ("air_water_sources.sim", 3)])
def test_world_frames_match_checker(hip_lib, oracle_lib, script, frames):
    path = os.path.join(SCRIPTS, script)
    world = _abi.World(hip_lib, path, 0, **TIGHT)
    cpu, info, phi0 = _checker(oracle_lib, hip_lib, path, **TIGHT)
    fg, fc = world.solver.get_fields(), cpu.get_fields()
    assert np.array_equal(fg["phi"], fc["phi"])
    if info.n_sources:
        assert np.abs(fc["phi"] - phi0).max() > 0.1  # the sources changed the level set
    else:
        assert np.array_equal(fg["phi"], phi0)
    accu = 0.0
    for f in range(frames):
        dt = _checker_frame(cpu, info, f)
        st = world.advance_frame()
        wi = world.info()
        assert wi.frame == f + 1
        if info.cfl != 0:
            accu += dt
            if f == 0:
                assert wi.accu_dt == dt
            assert abs(wi.accu_dt - accu) <= 1e-9 * accu
        assert st["pcg_iterations"] > 0
    fg, fc = world.solver.get_fields(), cpu.get_fields()
    fg["p"] -= fg["p"].mean()
    fc["p"] -= fc["p"].mean()
    scale = max(np.linalg.norm(fc[c]) for c in ("u", "v", "w"))
    print(script, {c: rel_err(fg[c], fc[c]) for c in fg})
    for c in ("phi", "p"):
        assert rel_err(fg[c], fc[c]) < 1e-8, (c, rel_err(fg[c], fc[c]))
    for c in ("u", "v", "w"):
        assert np.linalg.norm(fg[c] - fc[c]) / scale < 1e-8, c
    world.close()


def test_world_sources_follow_their_frame_window(hip_lib):
    """SOURCE_A adds its sphere in frames 0 and 1: at the cell nearest its
    centre (signed distance -0.104 there) the level set stays well inside
    the air while the window is open."""
    path = os.path.join(SCRIPTS, "air_water_sources.sim")
    world = _abi.World(hip_lib, path, 0, **TIGHT)
    info, src = _abi.parse_world(hip_lib, path)
    g = info.grid
    a = src[0]
    d = 2.0 / 16
    # the cell whose half-offset centre (CellCenter) is nearest SOURCE_A's centre
    ijk = [int(round((a.center[ax] - lo) / d - 0.5)) for ax, lo in enumerate((g.x_min, g.y_min, g.z_min))]
    val = lambda: world.solver.get_fields()["phi"][ijk[2], ijk[1], ijk[0]]
    assert val() < -0.1
    for _ in range(2):
        world.advance_frame()
        assert val() < -0.05
    world.close()


@pytest.mark.parametrize("precision", [_abi.FS_FP32_REFINE, _abi.FS_FP32_MG])
def test_world_fp32_refine_and_fixed_schedule(hip_lib, precision):
    """The script's scene with an FP32 projection inside FP64 refinement
    (Jacobi or multigrid) and the fixed schedule advances and keeps every
    solve at its relative residual."""
    path = os.path.join(SCRIPTS, "air_water_bubble.sim")
    world = _abi.World(hip_lib, path, 0, precision=precision, schedule=_abi.FS_SCHEDULE_FIXED,
                       stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8)
    for _ in range(2):
        st = world.advance_frame()
        assert st["residual_rel"] <= 1e-8
    f = world.solver.get_fields()
    assert all(np.isfinite(f[c]).all() for c in f)
    world.close()


@pytest.mark.parametrize("script,parallel,cfl", [("oil_water_pipe.sim", 1, 0.9),
                                                 ("oil_water_pipe_vertical.sim", 0, 0.8)])
def test_pipe_world_matches_checker(hip_lib, oracle_lib, script, parallel, cfl):
    """An oil_water_simulation script (the parallel or the vertical pipe)
    through fs_world_*: the world's pipe handle against the CPU checker's pipe
    driven through the same loop (the CFL step with the curvature of the
    current level set, then ComputeCurvatures and the step), bit for bit over
    three frames."""
    path = os.path.join(SCRIPTS, script)
    world = _abi.World(hip_lib, path, 0)
    info = world.info()
    assert info.oil_water == 1 and info.cfl == cfl and info.pipe.is_parallel == parallel
    gpu = _abi.Pipe(hip_lib, "fs_", info.params, info.pipe, info.grid, handle=hip_lib.fs_world_solver(world.h))
    cpu = _abi.Pipe(oracle_lib, "fsop_", info.params, info.pipe, info.grid)
    cpu.initial_fields()
    for c, a in gpu.get_fields().items():
        assert np.array_equal(a, cpu.get_fields()[c]), c
    accu = 0.0
    for f in range(3):
        dt = cpu.cfl_dt(info.cfl)
        sc = cpu.advance_frame(dt)
        sg = world.advance_frame()
        accu += dt
        assert world.info().accu_dt == accu
        assert sg["pcg_iterations"] == sc["pcg_iterations"]
    fg, fc = gpu.get_fields(), cpu.get_fields()
    for c in fc:
        assert np.array_equal(fg[c], fc[c]), c
    world.close()
