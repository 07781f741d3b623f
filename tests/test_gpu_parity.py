"""HIP path vs the CPU checker, through the C-ABI, on seeded scenes.

Bit-exact where the reference arithmetic has no libm call and no
reduction: cell classification, divergence, GFM face coefficients, the
diagonal with the Neumann pin, the right-hand side (after the mean removal,
summed in the same plane order), the operator product, WENO5/ENO3 advection
stages and the Godunov reinitialization sweeps.  Tolerances elsewhere:
  - smoothed Heaviside (sin) in viscosity / gradient update: 1e-13 relative;
  - PCG results (dot products summed in a different order): the solve
    tolerance, checked as relative max-norm error <= 1e-7 on p and the
    corrected velocity for an absolute tolerance of 1e-9.
"""
import numpy as np
import pytest

from fluid_simulator_backup_amd import _abi
from helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["rising_drop", "dam_break"])
def test_fields_round_trip(hip_lib, oracle_lib, name):
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=12)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    for c in fg:
        assert np.array_equal(fg[c], fc[c]), c


@pytest.mark.parametrize("name,n", [("rising_drop", 24), ("dam_break", 20), ("bubbles", 16)])
def test_pressure_system_bit_exact(hip_lib, oracle_lib, name, n):
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=n)
    dt = 1e-3
    ag, ac = gpu.assemble(dt), cpu.assemble(dt)
    for k in ("diag", "bx", "by", "bz", "rhs"):
        assert np.array_equal(ag[k], ac[k]), (k, rel_err(ag[k], ac[k]))
    x = np.random.default_rng(3).standard_normal((n, n, n))
    assert np.array_equal(gpu.apply_operator(x), cpu.apply_operator(x))


@pytest.mark.parametrize("sched", [_abi.FS_SCHEDULE_COMPAT, _abi.FS_SCHEDULE_FIXED])
def test_ragged_grid_pressure_system(hip_lib, oracle_lib, sched):
    """Unequal cell counts and spacings: the compatible schedule's matrix
    (1/dx^2 diagonal for every face) and the fixed schedule's (each face over
    its own d^2) are both reproduced bit for bit."""
    from fluid_simulator_backup_amd import scenes
    phi, vel, P, _ = scenes.make("rising_drop", 16, schedule=sched)
    g = _abi.unit_cube_grid(13, 9, 11)
    z, y, x = np.meshgrid((np.arange(11) + .5) / 11, (np.arange(9) + .5) / 9, (np.arange(13) + .5) / 13,
                          indexing="ij")
    phi = np.sqrt((x - .5) ** 2 + (y - .5) ** 2 + (z - .4) ** 2) - .25
    rng = np.random.default_rng(1)
    u, v, w = rng.standard_normal((11, 9, 14)), rng.standard_normal((11, 10, 13)), rng.standard_normal((12, 9, 13))
    gpu = _abi.Solver(hip_lib, "fs_", P, g, 0)
    cpu = _abi.Solver(oracle_lib, "fso_", P, g)
    for s in (gpu, cpu):
        s.set_fields(phi, u, v, w)
    ag, ac = gpu.assemble(0.01), cpu.assemble(0.01)
    for k in ("diag", "bx", "by", "bz", "rhs"):
        assert np.array_equal(ag[k], ac[k]), k
    xx = rng.standard_normal((11, 9, 13))
    assert np.array_equal(gpu.apply_operator(xx), cpu.apply_operator(xx))


def test_phase_sequence(hip_lib, oracle_lib):
    """Phases in the order a frame calls them, compared after each."""
    gpu, cpu = make_pair(hip_lib, oracle_lib, "rising_drop", n=24, tolerance=1e-9)
    dt = 2e-3
    for s in (gpu, cpu):
        s.compute_curvatures()
    # reinitialization: WENO5 + Godunov + clamp, no reduction
    for s in (gpu, cpu):
        s.phase("reinitialization", dt)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    assert np.array_equal(fg["phi"], fc["phi"]), rel_err(fg["phi"], fc["phi"])
    # projection: assembly exact, PCG to 1e-9 absolute
    sg, sc = gpu.projection(dt), cpu.projection(dt)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    assert abs(sg["pcg_iterations"] - sc["pcg_iterations"]) <= 2
    for c in ("p", "w", "u", "v"):
        assert rel_err(fg[c], fc[c]) < 1e-7, (c, rel_err(fg[c], fc[c]))
    assert np.array_equal(fg["u"], fc["u"]) and np.array_equal(fg["v"], fc["v"])  # x, y not corrected
    # level-set advection reads the ghost copies made above: exact
    for s in (gpu, cpu):
        s.phase("levelset_advection", dt)
    fg2, fc2 = gpu.get_fields(), cpu.get_fields()
    assert np.array_equal(fg2["phi"], fc2["phi"])
    # velocity advection: increments from exact ghost copies
    for s in (gpu, cpu):
        s.phase("velocity_advection", dt)
    fg3, fc3 = gpu.get_fields(), cpu.get_fields()
    for c in ("u", "v"):
        assert np.array_equal(fg3[c], fc3[c]), c
    assert rel_err(fg3["w"] - fg2["w"], fc3["w"] - fc2["w"]) < 1e-12
    for s in (gpu, cpu):
        s.phase("apply_gravity", dt)
        s.phase("sourcing", dt)
    fg4, fc4 = gpu.get_fields(), cpu.get_fields()
    for c in ("u", "v", "w"):
        assert rel_err(fg4[c], fc4[c]) < 1e-7, c


@pytest.mark.parametrize("schedule", [_abi.FS_SCHEDULE_COMPAT, _abi.FS_SCHEDULE_FIXED])
def test_fast_weno_phases_within_rounding(hip_lib, oracle_lib, schedule):
    """fast_weno: the WENO5 derivatives in fewer operations (product-form
    weights, fused multiply-adds) give the checker's level-set stage, WENO5
    velocity stages and 25 reinitialization sweeps to within rounding (the
    reference's operation order gives them bit for bit, test_phase_sequence)."""
    gpu, cpu = make_pair(hip_lib, oracle_lib, "rising_drop", n=24, schedule=schedule, use_5th_weno_v=1,
                         fast_weno=1)
    dt = 2e-3
    for s in (gpu, cpu):
        s.compute_curvatures()
        s.phase("reinitialization", dt)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    e = rel_err(fg["phi"], fc["phi"])
    print("reinitialization, 25 sweeps: rel err", e)
    assert 0 < e < 1e-13
    for s in (gpu, cpu):
        s.projection(dt)  # fills the velocity ghost copies the compatible stages read
        s.phase("levelset_advection", dt)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    e = rel_err(fg["phi"], fc["phi"])
    print("level-set stage: rel err", e)
    assert e < 1e-13
    before = gpu.get_fields()
    for s in (gpu, cpu):
        s.set_fields(u=before["u"], v=before["v"], w=before["w"])  # the same start on both sides
        s.phase("velocity_advection", dt)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    for c in ("u", "v", "w"):
        e = rel_err(fg[c] - before[c], fc[c] - before[c])
        print(c, "velocity stage increment: rel err", e)
        assert e < 1e-11, c


@pytest.mark.parametrize("schedule", [_abi.FS_SCHEDULE_COMPAT, _abi.FS_SCHEDULE_FIXED])
def test_fast_reinit_z_march_equals_per_cell(hip_lib, oracle_lib, monkeypatch, schedule):
    """The fast sweep marching z (each column's WENO5 triples sliding with k,
    8 planes per thread) gives the per-cell fast sweep's bits, also in a
    last chunk shorter than 8 planes (21 = 2 x 8 + 5)."""
    out = []
    for per_cell in (False, True):
        if per_cell:
            monkeypatch.setenv("FS_REINIT_PER_CELL", "1")
        else:
            monkeypatch.delenv("FS_REINIT_PER_CELL", raising=False)
        gpu, _ = make_pair(hip_lib, oracle_lib, "rising_drop", n=21, schedule=schedule, fast_weno=1)
        gpu.compute_curvatures()
        gpu.phase("reinitialization", 2e-3)
        out.append(gpu.get_fields()["phi"])
        gpu.close()
    assert np.array_equal(out[0], out[1])


def _frames(hip_lib, oracle_lib, name, n, frames, velocity_scale=False, **over):
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=n, moving=False, **over)
    its = []
    for f in range(frames):
        # the CFL rule before every frame (SIMULATION_WORLD.h:288-312), from
        # the checker's velocity, and the same dt on both sides
        dt = cpu.cfl_dt(0.5, 1)
        sg = gpu.advance_frame(dt)
        sc = cpu.advance_frame(dt)
        its.append((sg["pcg_iterations"], sc["pcg_iterations"]))
    fg, fc = gpu.get_fields(), cpu.get_fields()
    # the pressure's additive constant is held only by the weak Neumann pin:
    # compare it with the mean removed
    fg["p"] = fg["p"] - fg["p"].mean()
    fc["p"] = fc["p"] - fc["p"].mean()
    errs = {c: rel_err(fg[c], fc[c]) for c in ("phi", "u", "v", "w", "p")}
    if velocity_scale:
        # velocity components against the velocity's overall size (a nearly
        # still component would otherwise weigh solver-tolerance noise)
        scale = max(np.linalg.norm(fc[c]) for c in ("u", "v", "w"))
        for c in ("u", "v", "w"):
            errs[c] = np.linalg.norm(fg[c] - fc[c]) / scale
    print(name, n, "iterations gpu/cpu", its, "errors", errs)
    return its, errs


@pytest.mark.parametrize("name,n,frames", [("rising_drop", 24, 3), ("dam_break", 24, 3)])
def test_frames_fp64_tight(hip_lib, oracle_lib, name, n, frames):
    """Solve to 4e-11 relative residual (just above the attainable FP64
    residual of the density ratio 1000 scenes): the frames agree to 1e-8."""
    its, errs = _frames(hip_lib, oracle_lib, name, n, frames, stop_rule=_abi.FS_STOP_REL_RESIDUAL,
                        rel_tolerance=4e-11)
    for c, e in errs.items():
        assert e < 1e-8, (c, e)


@pytest.mark.parametrize("name", ["rising_drop", "dam_break"])
def test_semi_implicit_viscosity_phase(hip_lib, oracle_lib, name):
    """The three semi-implicit viscosity solves on a random field across a
    two-phase interface: rows assembled identically (same entry order), the
    solves to 1e-13 relative residual agree to 1e-10."""
    over = dict(semi_implicit_viscosity=1, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-13,
                water_viscosity=0.1, air_viscosity=0.05)
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=20, **over)
    rng = np.random.default_rng(7)
    f = gpu.get_fields()
    vel = [rng.standard_normal(f[c].shape) for c in ("u", "v", "w")]
    for s in (gpu, cpu):
        s.set_fields(u=vel[0], v=vel[1], w=vel[2])
        s.phase("sourcing", 1e-2)
    fg, fc = gpu.get_fields(), cpu.get_fields()
    for c in ("u", "v", "w"):
        assert rel_err(fg[c], fc[c]) < 1e-10, (c, rel_err(fg[c], fc[c]))
        assert rel_err(fg[c], vel["uvw".index(c)]) > 1e-4  # the solve did change the field


@pytest.mark.parametrize("name", ["dam_break", "rising_drop"])
def test_frames_semi_implicit_rk3(hip_lib, oracle_lib, name):
    """TVD-RK3 frames with semi-implicit viscosity (three viscosity solves
    and three projections per step) agree to 1e-8 after 2 frames when every
    solve reaches 4e-11 relative residual."""
    its, errs = _frames(hip_lib, oracle_lib, name, 20, 2, order_for_time_advancing=3, semi_implicit_viscosity=1,
                        stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=4e-11)
    for c, e in errs.items():
        assert e < 1e-8, (c, e)


@pytest.mark.parametrize("name,order", [("dam_break", 2), ("dam_break", 3), ("rising_drop", 3)])
def test_frames_runge_kutta_orders(hip_lib, oracle_lib, name, order):
    """Orders 2 and 3 of the step schedule (two / three projections per
    step, stage blends, RK reinitialization) agree to 1e-8 after 2 frames
    when each solve reaches 4e-11 relative residual."""
    its, errs = _frames(hip_lib, oracle_lib, name, 20, 2, order_for_time_advancing=order,
                        stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=4e-11)
    for c, e in errs.items():
        assert e < 1e-8, (c, e)


@pytest.mark.parametrize("name,order,over", [
    ("dam_break", 1, dict(use_5th_weno_v=1)),
    ("rising_drop", 1, dict()),
    ("dam_break", 3, dict(use_5th_weno_v=1)),
    ("rising_drop", 2, dict(semi_implicit_viscosity=1)),
])
def test_frames_fixed_schedule(hip_lib, oracle_lib, name, order, over):
    """The fixed schedule (ghost copies refreshed before each advection stage,
    every face corrected with the matrix's coefficients, wall faces zeroed
    after forces, each WENO5 velocity component on its own ghost copy, the
    Euclidean Godunov norm) agrees with the CPU checker's to 1e-8 after 2
    frames when each solve reaches 4e-11 relative residual (velocity
    components measured against the whole velocity's norm)."""
    its, errs = _frames(hip_lib, oracle_lib, name, 20, 2, velocity_scale=True, order_for_time_advancing=order,
                        schedule=_abi.FS_SCHEDULE_FIXED, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=4e-11,
                        **over)
    for c, e in errs.items():
        assert e < 1e-8, (c, e)


@pytest.mark.parametrize("name,n", [("rising_drop", 32), ("dam_break", 24), ("bubbles", 32)])
def test_fp32_refined_projection_matches_fp64(hip_lib, oracle_lib, name, n):
    """FP32 PCG with FP64 refinement reaches 1e-8 relative residual (checked
    in FP64 on the device) and its fields match the FP64 path within what
    two solutions at that residual can differ by."""
    over = dict(stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8)
    g64, cpu = make_pair(hip_lib, oracle_lib, name, n=n, **over)
    g32, _ = make_pair(hip_lib, oracle_lib, name, n=n, precision=_abi.FS_FP32_REFINE, **over)
    dt = 2e-3
    s64, s32, sc = g64.projection(dt), g32.projection(dt), cpu.projection(dt)
    print(name, n, "fp64", s64, "fp32+refine", s32)
    assert s32["residual_rel"] <= 1e-8 and s64["residual_rel"] <= 1e-8
    assert s32["pcg_iterations"] <= 1.5 * s64["pcg_iterations"] + 20
    f64, f32, fc = g64.get_fields(), g32.get_fields(), cpu.get_fields()
    for c in ("w", "p"):
        a, b = f32[c], fc[c]
        if c == "p":
            a, b = a - a.mean(), b - b.mean()
        assert rel_err(a, b) < 1e-5, (c, rel_err(a, b))


def test_fp32_refined_projection_unequal_spacing(hip_lib, oracle_lib):
    """FP32+refine on unequal spacings (40 x 28 x 34 cells on the unit
    cube, density ratio 1000) in the fixed schedule, whose flux-form matrix
    keeps constants in its null space: it reaches 1e-10 relative residual and
    matches the CPU checker's FP64 solve; the compatible schedule refuses the
    combination."""
    nx, ny, nz = 40, 28, 34
    g = _abi.unit_cube_grid(nx, ny, nz)
    k, j, i = np.meshgrid((np.arange(nz) + .5) / nz, (np.arange(ny) + .5) / ny, (np.arange(nx) + .5) / nx,
                          indexing="ij")
    phi = np.sqrt((i - .5) ** 2 + (j - .45) ** 2 + (k - .4) ** 2) - .2
    rng = np.random.default_rng(3)
    vel = (rng.standard_normal((nz, ny, nx + 1)), rng.standard_normal((nz, ny + 1, nx)),
           rng.standard_normal((nz + 1, ny, nx)))
    over = dict(stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-10, max_iteration=20000)
    P32 = _abi.default_params(schedule=_abi.FS_SCHEDULE_FIXED, precision=_abi.FS_FP32_REFINE, **over)
    P64 = _abi.default_params(schedule=_abi.FS_SCHEDULE_FIXED, **over)
    gpu = _abi.Solver(hip_lib, "fs_", P32, g, 0)
    cpu = _abi.Solver(oracle_lib, "fso_", P64, g)
    for s in (gpu, cpu):
        s.set_fields(phi, *vel)
    st = gpu.projection(1e-2)
    cpu.projection(1e-2)
    assert st["residual_rel"] <= 1e-10, st
    fg, fc = gpu.get_fields(), cpu.get_fields()
    for c in ("p", "u", "v", "w"):
        a, b = fg[c], fc[c]
        if c == "p":
            a, b = a - a.mean(), b - b.mean()
        assert rel_err(a, b) < 1e-6, (c, rel_err(a, b))
    with pytest.raises(_abi.FsError) as e:
        _abi.Solver(hip_lib, "fs_", _abi.default_params(precision=_abi.FS_FP32_REFINE, **over), g, 0)
    assert e.value.code == _abi.FS_ERR_UNSUPPORTED


@pytest.mark.parametrize("sched", [_abi.FS_SCHEDULE_COMPAT, _abi.FS_SCHEDULE_FIXED])
def test_spmv_coefficients_from_phi(hip_lib, monkeypatch, sched):
    """From 128^3 on, the fused SpMV forms the face coefficients from the
    level set instead of reading the stored arrays.  FP64: the solve is the
    same bit for bit (same formula, same operands).  FP32+refine: the inner
    operator rounds differently but reaches 1e-8 and matches the FP64 fields
    to the solve's accuracy.  (FS_BETA_ARRAYS selects the stored arrays.)"""
    n = 128
    over = dict(stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8, schedule=sched)
    out = {}
    for prec in (_abi.FS_FP64, _abi.FS_FP32_REFINE):
        for arrays in (True, False):
            if arrays:
                monkeypatch.setenv("FS_BETA_ARRAYS", "1")
            else:
                monkeypatch.delenv("FS_BETA_ARRAYS", raising=False)
            from fluid_simulator_backup_amd import scenes
            from helpers import smooth_velocity
            phi, _, P, g = scenes.make("rising_drop", n, precision=prec, **over)
            gpu = _abi.Solver(hip_lib, "fs_", P, g, 0)
            gpu.set_fields(phi, *smooth_velocity(n))
            st = gpu.projection(2e-3)
            out[prec, arrays] = (st, gpu.get_fields())
            gpu.close()
    (s_a, f_a), (s_p, f_p) = out[_abi.FS_FP64, True], out[_abi.FS_FP64, False]
    assert s_a["pcg_iterations"] == s_p["pcg_iterations"]
    for c in ("p", "u", "v", "w"):
        assert np.array_equal(f_a[c], f_p[c]), c
    s32, f32 = out[_abi.FS_FP32_REFINE, False]
    assert s32["residual_rel"] <= 1e-8, s32
    for c in ("p", "w"):
        a, b = f32[c], f_a[c]
        if c == "p":
            a, b = a - a.mean(), b - b.mean()
        assert rel_err(a, b) < 1e-5, (c, rel_err(a, b))


@pytest.mark.parametrize("name,n", [("rising_drop", 32), ("dam_break", 32), ("bubbles", 40)])
def test_multigrid_refined_projection(hip_lib, oracle_lib, name, n):
    """FS_FP32_MG: CG preconditioned by a multigrid V-cycle in FP32, inside
    the same FP64 refinement.  It reaches 1e-10 relative residual (the FP64
    residual decides), matches the CPU checker's FP64 solve to the solve's
    accuracy, and needs far fewer inner iterations than Jacobi."""
    over = dict(stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-10)
    gmg, cpu = make_pair(hip_lib, oracle_lib, name, n=n, precision=_abi.FS_FP32_MG, **over)
    gj, _ = make_pair(hip_lib, oracle_lib, name, n=n, precision=_abi.FS_FP32_REFINE, **over)
    smg, sj = gmg.projection(2e-3), gj.projection(2e-3)
    cpu.projection(2e-3)
    print(name, n, "mg", smg["pcg_iterations"], smg["refinement_sweeps"], "jacobi", sj["pcg_iterations"])
    assert smg["residual_rel"] <= 1e-10, smg
    assert 3 * smg["pcg_iterations"] < sj["pcg_iterations"], (smg, sj)
    fg, fc = gmg.get_fields(), cpu.get_fields()
    for c in ("p", "u", "v", "w"):
        a, b = fg[c], fc[c]
        if c == "p":
            a, b = a - a.mean(), b - b.mean()
        assert rel_err(a, b) < 1e-7, (c, rel_err(a, b))


def test_multigrid_odd_and_unequal_grids(hip_lib, oracle_lib):
    """Odd cell counts (ceil coarsening, clamped transfer at the walls) and
    unequal spacings (fixed schedule): the multigrid refinement reaches
    1e-10 and matches the FP64 checker."""
    nx, ny, nz = 37, 29, 41
    g = _abi.unit_cube_grid(nx, ny, nz)
    k, j, i = np.meshgrid((np.arange(nz) + .5) / nz, (np.arange(ny) + .5) / ny, (np.arange(nx) + .5) / nx,
                          indexing="ij")
    phi = np.minimum(np.sqrt((i - .4) ** 2 + (j - .55) ** 2 + (k - .45) ** 2) - .2, 0.8 - k)
    rng = np.random.default_rng(8)
    vel = (rng.standard_normal((nz, ny, nx + 1)), rng.standard_normal((nz, ny + 1, nx)),
           rng.standard_normal((nz + 1, ny, nx)))
    over = dict(stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-10, schedule=_abi.FS_SCHEDULE_FIXED)
    gpu = _abi.Solver(hip_lib, "fs_", _abi.default_params(precision=_abi.FS_FP32_MG, **over), g, 0)
    cpu = _abi.Solver(oracle_lib, "fso_", _abi.default_params(**over), g)
    for s in (gpu, cpu):
        s.set_fields(phi, *vel)
    st = gpu.projection(1e-2)
    cpu.projection(1e-2)
    assert st["residual_rel"] <= 1e-10, st
    fg, fc = gpu.get_fields(), cpu.get_fields()
    for c in ("p", "u", "v", "w"):
        a, b = fg[c], fc[c]
        if c == "p":
            a, b = a - a.mean(), b - b.mean()
        assert rel_err(a, b) < 1e-7, (c, rel_err(a, b))


def test_multigrid_frames(hip_lib, oracle_lib):
    """Two frames of the dam break with the multigrid refinement agree with
    the CPU checker's FP64 frames to 1e-8 when both solve to 4e-11."""
    its, errs = _frames(hip_lib, oracle_lib, "dam_break", 24, 2, velocity_scale=True, precision=_abi.FS_FP32_MG,
                        stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=4e-11)
    for c, e in errs.items():
        assert e < 1e-8, (c, e)


def test_multigrid_fused_sweeps_match_half_sweeps(hip_lib, monkeypatch):
    """The fused red-black sweep (both colours in one launch, out of place)
    and the LDS coarsest-level solve give the values of in-place half sweeps
    bit for bit: same arithmetic per cell, and red-black updates within a
    colour are independent.  An odd, unequal grid so tiles and z-chunks end
    inside the box."""
    from fluid_simulator_backup_amd import scenes
    phi, vel, P, g = scenes.make("bubbles", 40, precision=_abi.FS_FP32_MG, stop_rule=_abi.FS_STOP_REL_RESIDUAL,
                                 rel_tolerance=1e-8, schedule=_abi.FS_SCHEDULE_FIXED)
    g = _abi.unit_cube_grid(37, 41, 35)
    z, y, x = np.meshgrid((np.arange(35) + .5) / 35, (np.arange(41) + .5) / 41, (np.arange(37) + .5) / 37,
                          indexing="ij")
    phi = np.minimum(np.sqrt((x - .5) ** 2 + (y - .45) ** 2 + (z - .4) ** 2) - .2, z - .75)
    rng = np.random.default_rng(2)
    vel = (0.1 * rng.standard_normal((35, 41, 38)), 0.1 * rng.standard_normal((35, 42, 37)),
           0.1 * rng.standard_normal((36, 41, 37)))
    out = []
    for split in (True, False):
        if split:
            monkeypatch.setenv("FS_MG_SPLIT_SMOOTH", "1")
        else:
            monkeypatch.delenv("FS_MG_SPLIT_SMOOTH", raising=False)
        s = _abi.Solver(hip_lib, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        st = [s.advance_frame(2e-3) for _ in range(2)]
        out.append((s.get_fields(), [t["pcg_iterations"] for t in st]))
        s.close()
    assert out[0][1] == out[1][1]
    for c in out[0][0]:
        assert np.array_equal(out[0][0][c], out[1][0][c]), c


@pytest.mark.parametrize("name,order,masked", [("rising_drop", 1, None), ("dam_break", 2, None), ("rising_drop", 1, "7"),
                                               ("dam_break", 3, "0"), ("rising_drop", 2, "-1")])
def test_concurrent_reinitialization_matches_sequential_order(hip_lib, monkeypatch, name, order, masked):
    """The step's reinitialization on the second stream (on a copy of the
    advected level set, with the assembly's ghost fill replayed), its later
    sweeps on the all-CU stream from the switch point on (adaptive, or fixed
    by FS_REINIT_MASKED_SWEEPS: 7 within the first call, 0 at once, -1
    never), gives the frames of the sequential order bit for bit
    (FS_SERIAL_REINIT)."""
    from fluid_simulator_backup_amd import scenes
    phi, vel, P, g = scenes.make(name, 24, precision=_abi.FS_FP32_MG, stop_rule=_abi.FS_STOP_REL_RESIDUAL,
                                 rel_tolerance=1e-8, order_for_time_advancing=order)
    out = []
    if masked is None:
        monkeypatch.delenv("FS_REINIT_MASKED_SWEEPS", raising=False)
    else:
        monkeypatch.setenv("FS_REINIT_MASKED_SWEEPS", masked)
    for serial in (True, False):
        if serial:
            monkeypatch.setenv("FS_SERIAL_REINIT", "1")
        else:
            monkeypatch.delenv("FS_SERIAL_REINIT", raising=False)
        s = _abi.Solver(hip_lib, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        dt = s.cfl_dt(0.5, 1)
        for _ in range(3):
            s.advance_frame(dt)
        out.append(s.get_fields())
        s.close()
    for c in out[0]:
        assert np.array_equal(out[0][c], out[1][c]), c


def test_frame_kernel_sampling_leaves_frames_unchanged(hip_lib):
    """The in-loop probes of the reinitialization sweep and the finest-level
    smoother time one launch per call, count every launch, and change no
    result: the sampled frames equal unsampled ones bit for bit."""
    from fluid_simulator_backup_amd import scenes
    phi, vel, P, g = scenes.make("rising_drop", 20, precision=_abi.FS_FP32_MG, stop_rule=_abi.FS_STOP_REL_RESIDUAL,
                                 rel_tolerance=1e-8)
    out = []
    for sampling in (False, True):
        s = _abi.Solver(hip_lib, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        dt = s.cfl_dt(0.5, 1)
        _abi.set_kernel_sampling(s, sampling)
        for _ in range(2):
            s.advance_frame(dt)
        if sampling:
            avg, n, launches = _abi.frame_kernel_samples(s, _abi.FS_KS_REINIT)
            assert avg > 0 and n >= 2 and launches == n * P.iteration_number_for_reinitialization
            avg, n, launches = _abi.frame_kernel_samples(s, _abi.FS_KS_MG_SMOOTH)
            assert avg > 0 and n >= 2 and launches % 4 == 0 and launches >= 4 * n
        out.append(s.get_fields())
        s.close()
    for c in out[0]:
        assert np.array_equal(out[0][c], out[1][c]), c


@pytest.mark.parametrize("name,n,frames", [("rising_drop", 24, 4), ("dam_break", 24, 4)])
def test_frames_fp64_reference_stop_rule(hip_lib, oracle_lib, name, n, frames):
    """The reference's absolute stop rule at its script tolerance (1e-5):
    iteration counts within 2 of each other, fields within 1e-4."""
    its, errs = _frames(hip_lib, oracle_lib, name, n, frames)
    for a, b in its:
        assert abs(a - b) <= 2, its
    for c, e in errs.items():
        assert e < 1e-4, (c, e)


@pytest.mark.parametrize("precision,mode,mixed", [(_abi.FS_FP32_REFINE, "graph", False),
                                                   (_abi.FS_FP32_MG, "graph", False),
                                                   (_abi.FS_FP32_MG, "graph", True),
                                                   (_abi.FS_FP32_MG, "eager", True),
                                                   (_abi.FS_FP32_MG, "per-scene", False),
                                                   (_abi.FS_FP32_MG, "per-scene", True)])
def test_batch_matches_scenes_advanced_alone(hip_lib, monkeypatch, precision, mode, mixed):
    """A batch of uneven scenes (FP32 projection inside FP64 refinement, one
    host thread per scene stream) gives, bit for bit and after every frame,
    what each scene gives advanced alone: every kernel and reduction is
    deterministic and no state is shared.  The batch's reinitialization is
    each scene's own stage on its stream ("per-scene", the default), or the
    opt-in scene-table stage (FS_BATCH_TABLE: one launch per kernel for all
    scenes) replayed from its two captured graphs frame after frame
    ("graph") or the same launches without the graphs ("eager"); `mixed` gives every second scene the fixed schedule and
    the fast WENO5 arithmetic."""
    from fluid_simulator_backup_amd import scenes
    monkeypatch.delenv("FS_BATCH_NO_GRAPH", raising=False)
    monkeypatch.delenv("FS_BATCH_TABLE", raising=False)
    if mode != "per-scene":
        monkeypatch.setenv("FS_BATCH_TABLE", "1")
    if mode == "eager":
        monkeypatch.setenv("FS_BATCH_NO_GRAPH", "1")
    sc = scenes.batch_scenes(6, seed=5, edges=(12, 16, 20, 24), precision=precision,
                             stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8, max_iteration=5000)
    if mixed:
        for q, s in enumerate(sc):
            if q % 2:
                s[3].schedule = _abi.FS_SCHEDULE_FIXED
                s[3].fast_weno = 1
    batch = _abi.Batch(hip_lib, [s[3] for s in sc], [s[4] for s in sc], 0, 4)
    alone = []
    for (name, phi, vel, P, g), bs in zip(sc, batch.scenes):
        bs.set_fields(phi, *vel)
        s = _abi.Solver(hip_lib, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        alone.append(s)
    for f in range(4):
        dts = [s.cfl_dt(0.5, 1) for s in alone]
        st_b = batch.advance_frame(dts)
        st_a = [s.advance_frame(dt) for s, dt in zip(alone, dts)]
        for a, b in zip(st_a, st_b):
            assert a["pcg_iterations"] == b["pcg_iterations"]
        for q, (s, bs) in enumerate(zip(alone, batch.scenes)):
            fa, fb = s.get_fields(), bs.get_fields()
            for c in ("phi", "u", "v", "w", "p"):
                assert np.array_equal(fa[c], fb[c]), (f, q, c)
    batch.close()
