"""Surface tension on the HIP path vs the CPU checker, through the C-ABI.

Bit-exact: the curvature field and its maximum near the interface (central
differences, the one-sided fallback, the clamp; no libm call but sqrt, no
reduction but a max), the pressure right-hand side with the jump terms, and
the CFL time step.  Solves: within the solve tolerance, as for the other
projections.
"""
import numpy as np
import pytest

from fluid_simulator_backup_amd import _abi, scenes
from helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu

SIGMA = 0.0728  # air/water, 3D_script.sim:191
SCHEDULES = [_abi.FS_SCHEDULE_COMPAT, _abi.FS_SCHEDULE_FIXED]


def _pair_on(hip_lib, oracle_lib, phi, grid, **over):
    kw = dict(surface_tension=SIGMA)
    kw.update(over)
    P = _abi.default_params(**kw)
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    rng = np.random.default_rng(2)
    vel = (0.1 * rng.standard_normal((nz, ny, nx + 1)), 0.1 * rng.standard_normal((nz, ny + 1, nx)),
           0.1 * rng.standard_normal((nz + 1, ny, nx)))
    gpu = _abi.Solver(hip_lib, "fs_", P, grid, 0)
    cpu = _abi.Solver(oracle_lib, "fso_", P, grid)
    for s in (gpu, cpu):
        s.set_fields(phi, *vel)
    return gpu, cpu


@pytest.mark.parametrize("sched", SCHEDULES)
@pytest.mark.parametrize("name,n", [("rising_drop", 24), ("bubbles", 20)])
def test_curvature_bit_exact(hip_lib, oracle_lib, sched, name, n):
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=n, surface_tension=SIGMA, schedule=sched)
    (kg, mg), (kc, mc) = gpu.curvature(), cpu.curvature()
    assert np.array_equal(kg, kc), rel_err(kg, kc)
    assert mg == mc and mg > 0


@pytest.mark.parametrize("sched", SCHEDULES)
def test_curvature_ragged_grid_fallback_and_clamp(hip_lib, oracle_lib, sched):
    """Uneven spacing (dz < dy < dx) so the clamp and the dz quirk act; a
    kink whose central differences vanish (one-sided fallback) and a sphere
    of a cell's size (clamped)."""
    g = _abi.unit_cube_grid(13, 17, 21)
    k, j, i = np.meshgrid(np.arange(21), np.arange(17), np.arange(13), indexing="ij")
    x, y, z = (i + .5) / 13, (j + .5) / 17, (k + .5) / 21
    kink = np.abs(i - 6) / 13 - 0.5 * np.abs(k - 10) / 21 + 0.3
    ball = np.sqrt((x - x[2, 8, 1]) ** 2 + (y - y[2, 8, 1]) ** 2 + (z - z[2, 8, 1]) ** 2) - 0.5 / 21
    phi = np.minimum(kink, ball)
    gpu, cpu = _pair_on(hip_lib, oracle_lib, phi, g, schedule=sched)
    (kg, mg), (kc, mc) = gpu.curvature(), cpu.curvature()
    assert np.array_equal(kg, kc), rel_err(kg, kc)
    assert mg == mc
    assert kc.min() == -21.0  # clamp at -1/dz


@pytest.mark.parametrize("sched", SCHEDULES)
@pytest.mark.parametrize("name,n", [("rising_drop", 24), ("dam_break", 20), ("bubbles", 16)])
def test_pressure_system_with_jump_bit_exact(hip_lib, oracle_lib, sched, name, n):
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=n, surface_tension=SIGMA, schedule=sched)
    ag, ac = gpu.assemble(1e-3), cpu.assemble(1e-3)
    for key in ("diag", "bx", "by", "bz", "rhs"):
        assert np.array_equal(ag[key], ac[key]), (key, rel_err(ag[key], ac[key]))
    # the jump terms are there: the right-hand side differs from sigma = 0's
    _, c0 = make_pair(None, oracle_lib, name, n=n, schedule=sched)
    assert rel_err(ac["rhs"], c0.assemble(1e-3)["rhs"]) > 1e-6


@pytest.mark.parametrize("sched", SCHEDULES)
def test_projection_with_jump(hip_lib, oracle_lib, sched):
    gpu, cpu = make_pair(hip_lib, oracle_lib, "rising_drop", n=24, surface_tension=SIGMA, schedule=sched,
                         stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-12)
    gpu.projection(2e-3)
    cpu.projection(2e-3)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    fg["p"] -= fg["p"].mean()
    fc["p"] -= fc["p"].mean()
    for c in ("p", "u", "v", "w"):
        assert rel_err(fg[c], fc[c]) < 1e-8, (c, rel_err(fg[c], fc[c]))


@pytest.mark.parametrize("sched", SCHEDULES)
@pytest.mark.parametrize("name", ["rising_drop", "dam_break"])
def test_frames_with_surface_tension(hip_lib, oracle_lib, sched, name):
    """Two CFL frames with sigma = 0.0728: the first CFL step agrees bit for
    bit (velocity maxima and curvature maximum are exact on equal fields),
    the second to the solve tolerance; the fields to 1e-8 when each solve
    reaches 4e-11 relative residual."""
    over = dict(surface_tension=SIGMA, schedule=sched, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=4e-11)
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=20, **over)
    for f in range(2):
        dg, dc = gpu.cfl_dt(0.5, 1), cpu.cfl_dt(0.5, 1)
        assert dg == dc if f == 0 else abs(dg - dc) < 1e-8 * dc, (f, dg, dc)
        gpu.advance_frame(dc)
        cpu.advance_frame(dc)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    fg["p"] -= fg["p"].mean()
    fc["p"] -= fc["p"].mean()
    scale = max(np.linalg.norm(fc[c]) for c in ("u", "v", "w"))
    for c in ("phi", "p"):
        assert rel_err(fg[c], fc[c]) < 1e-8, (c, rel_err(fg[c], fc[c]))
    for c in ("u", "v", "w"):
        assert np.linalg.norm(fg[c] - fc[c]) / scale < 1e-8, c


def test_fp32_refined_projection_with_jump(hip_lib, oracle_lib):
    over = dict(surface_tension=SIGMA, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8)
    g32, cpu = make_pair(hip_lib, oracle_lib, "rising_drop", n=32, precision=_abi.FS_FP32_REFINE, **over)
    st = g32.projection(2e-3)
    cpu.projection(2e-3)
    assert st["residual_rel"] <= 1e-8
    fg, fc = g32.get_fields(), cpu.get_fields()
    for c in ("w", "p"):
        a, b = fg[c], fc[c]
        if c == "p":
            a, b = a - a.mean(), b - b.mean()
        assert rel_err(a, b) < 1e-5, (c, rel_err(a, b))


def test_flat_level_set_reports_nonfinite(hip_lib):
    P = _abi.default_params(surface_tension=SIGMA)
    s = _abi.Solver(hip_lib, "fs_", P, _abi.unit_cube_grid(8), 0)
    s.set_fields(np.full((8, 8, 8), 0.05))
    with pytest.raises(_abi.FsError) as e:
        s.curvature()
    assert e.value.code == _abi.FS_ERR_NONFINITE
    with pytest.raises(_abi.FsError) as e:
        s.cfl_dt(0.5, 1)
    assert e.value.code == _abi.FS_ERR_NONFINITE


@pytest.mark.parametrize("sched", SCHEDULES)
def test_frames_with_csf_model(hip_lib, oracle_lib, sched):
    """CSF_model with sigma = 0.0728 in the air/water path (no force, jump
    condition zeroed from the first Sourcing on): two CFL frames against the
    checker to 1e-8 (the CFL keeps its surface-tension term: the first dt bit
    for bit), and the GPU's fields equal its own frames with sigma = 0 and the
    same dt sequence bit for bit."""
    over = dict(schedule=sched, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=4e-11)
    gpu, cpu = make_pair(hip_lib, oracle_lib, "rising_drop", n=20, surface_tension=SIGMA, csf_model=1, **over)
    plain, _ = make_pair(hip_lib, oracle_lib, "rising_drop", n=20, surface_tension=0.0, **over)
    for f in range(2):
        dg, dc = gpu.cfl_dt(0.5, 1), cpu.cfl_dt(0.5, 1)
        assert dg == dc if f == 0 else abs(dg - dc) < 1e-8 * dc, (f, dg, dc)
        gpu.advance_frame(dc)
        cpu.advance_frame(dc)
        plain.advance_frame(dc)
    fg, fc, fp = gpu.get_fields(), cpu.get_fields(), plain.get_fields()
    for c in ("phi", "u", "v", "w", "p"):
        assert np.array_equal(fg[c], fp[c]), c
    fg["p"] -= fg["p"].mean()
    fc["p"] -= fc["p"].mean()
    scale = max(np.linalg.norm(fc[c]) for c in ("u", "v", "w"))
    for c in ("phi", "p"):
        assert rel_err(fg[c], fc[c]) < 1e-8, (c, rel_err(fg[c], fc[c]))
    for c in ("u", "v", "w"):
        assert np.linalg.norm(fg[c] - fc[c]) / scale < 1e-8, c


@pytest.mark.parametrize("method", [_abi.FS_CURV_NORMALS, _abi.FS_CURV_BOTH, _abi.FS_CURV_NONE])
@pytest.mark.parametrize("sched", SCHEDULES)
def test_curvature_methods_bit_exact(hip_lib, oracle_lib, sched, method):
    """curvature_method: by normals, both passes, or none, bit for bit with the
    checker -- before any projection (compatible schedule: zero normals) and
    after one (its ComputeNormals), on a ragged grid that reaches the clamp."""
    g = _abi.unit_cube_grid(23, 19, 21)
    k, j, i = np.meshgrid(np.arange(21), np.arange(19), np.arange(23), indexing="ij")
    x, y, z = (i + .5) / 23, (j + .5) / 19, (k + .5) / 21
    phi = np.minimum(np.sqrt((x - .5) ** 2 + (y - .45) ** 2 + (z - .55) ** 2) - 0.27,
                     np.sqrt((x - .2) ** 2 + (y - .7) ** 2 + (z - .3) ** 2) - 0.6 / 21)
    gpu, cpu = _pair_on(hip_lib, oracle_lib, phi, g, schedule=sched, curvature_method=method,
                        stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-10)
    for when in ("before", "after"):
        (kg, mg), (kc, mc) = gpu.curvature(), cpu.curvature()
        assert np.array_equal(kg, kc), (when, rel_err(kg, kc))
        assert mg == mc, (when, mg, mc)
        if when == "before":
            for s in (gpu, cpu):
                s.projection(1e-3, allow_not_converged=True)
    if method == _abi.FS_CURV_NONE:
        assert not kg.any()


@pytest.mark.parametrize("sched", SCHEDULES)
def test_frames_with_curvature_by_normals(hip_lib, oracle_lib, sched):
    """Two CFL frames with sigma = 0.0728 and the curvature by normals: the
    first CFL step bit for bit, the fields to 1e-8 of the checker when each
    solve reaches 4e-11 relative residual."""
    over = dict(surface_tension=SIGMA, schedule=sched, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=4e-11,
                curvature_method=_abi.FS_CURV_NORMALS)
    gpu, cpu = make_pair(hip_lib, oracle_lib, "rising_drop", n=20, **over)
    for f in range(2):
        dg, dc = gpu.cfl_dt(0.5, 1), cpu.cfl_dt(0.5, 1)
        assert dg == dc if f == 0 else abs(dg - dc) < 1e-8 * dc, (f, dg, dc)
        gpu.advance_frame(dc)
        cpu.advance_frame(dc)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    fg["p"] -= fg["p"].mean()
    fc["p"] -= fc["p"].mean()
    scale = max(np.linalg.norm(fc[c]) for c in ("u", "v", "w"))
    for c in ("phi", "p"):
        assert rel_err(fg[c], fc[c]) < 1e-8, (c, rel_err(fg[c], fc[c]))
    for c in ("u", "v", "w"):
        assert np.linalg.norm(fg[c] - fc[c]) / scale < 1e-8, c
