"""Oil/water pipe mode (fs_create_pipe) on the GPU against the CPU checker
(oracle/fs_oracle_pipe.c), phase by phase and frame by frame.

Every kernel keeps the reference's operation order and both sides sum in the
same fixed order, so the comparisons are bit for bit: fields after each phase,
the persistent ghost copies each phase leaves (the reference's stale values
outside the pipe included), the pressure and the solves' iteration counts."""
import sys

import numpy as np
import pytest

from fluid_simulator_backup_amd import _abi

pytestmark = pytest.mark.gpu

PHASES = ["levelset_advection", "velocity_advection", "gravity", "sourcing", "projection", "reinitialization"]
# the persistent arrays each phase writes
PHASE_ARRAYS = {
    "levelset_advection": ["ghost_v", "ghost_w", "scalar_ghost"],
    "velocity_advection": ["ghost_u", "ghost_v", "ghost_w", "adv_u_y", "adv_u_z", "adv_v_y", "adv_v_z", "adv_w_y",
                           "adv_w_z"],
    "gravity": ["scalar_ghost"],
    "sourcing": ["vis_u_y", "vis_u_z", "vis_v_y", "vis_v_z", "vis_w_y", "vis_w_z", "scalar_ghost"],
    "projection": ["divergence", "jump", "pressure", "phi", "ghost_v", "ghost_w", "boundary"],
    "reinitialization": ["scalar_ghost"],
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(hip_lib, oracle_lib, n=20, **kw):
    P, Q = _abi.pipe_params(**kw)
    g = _abi.pipe_grid(n)
    gpu = _abi.Pipe(hip_lib, "fs_", P, Q, g)
    cpu = _abi.Pipe(oracle_lib, "fsop_", P, Q, g)
    for s in (gpu, cpu):
        s.initial_fields()
    return gpu, cpu


def assert_fields_equal(gpu, cpu, what):
    fg, fc = gpu.get_fields(), cpu.get_fields()
    for c in ("phi", "u", "v", "w", "p"):
        d = float(np.abs(fg[c] - fc[c]).max())
        assert np.array_equal(fg[c], fc[c]), f"{what}: {c} differs by {d}"


def test_pipe_constants_and_initial_state(hip_lib, oracle_lib):
    say("[pipe] constants and initial state, 20^3")
    gpu, cpu = pair(hip_lib, oracle_lib)
    assert gpu.constants() == cpu.constants()
    assert_fields_equal(gpu, cpu, "initial")
    assert np.array_equal(gpu.array("boundary"), cpu.array("boundary"))


@pytest.mark.parametrize("warm", [0, 1])
def test_pipe_phases_bit_exact(hip_lib, oracle_lib, warm):
    """Each phase alone from the same state: from the initial state and after
    one full RK2 frame (v, w non-zero, the ghost copies holding the values
    the phases left)."""
    gpu, cpu = pair(hip_lib, oracle_lib)
    dt = cpu.cfl_dt(0.9)
    assert gpu.cfl_dt(0.9) == dt
    for _ in range(warm):
        say(f"[pipe] warm frame, dt={dt}")
        sg, sc = gpu.advance_frame(dt), cpu.advance_frame(dt)
        assert sg["pcg_iterations"] == sc["pcg_iterations"]
        assert_fields_equal(gpu, cpu, "warm frame")
    for name in PHASES:
        say(f"[pipe] phase {name} warm={warm}")
        sg, sc = gpu.phase(name, dt), cpu.phase(name, dt)
        assert_fields_equal(gpu, cpu, name)
        for arr in PHASE_ARRAYS[name]:
            a, b = gpu.array(arr), cpu.array(arr)
            assert np.array_equal(a, b), f"{name}: array {arr} differs by {np.abs(a - b).max()}"
        if name in ("sourcing", "projection"):
            assert sg["viscosity_iterations"] == sc["viscosity_iterations"]
            assert sg["pcg_iterations"] == sc["pcg_iterations"]


@pytest.mark.parametrize("order", [1, 2, 3])
def test_pipe_frames_bit_exact(hip_lib, oracle_lib, order):
    """Three frames with the CFL step before each (CFL 0.9, the script's)."""
    gpu, cpu = pair(hip_lib, oracle_lib, n=24, order_for_time_advancing=order)
    for f in range(3):
        dt = cpu.cfl_dt(0.9)
        assert gpu.cfl_dt(0.9) == dt
        say(f"[pipe] order {order} frame {f} dt={dt}")
        sg, sc = gpu.advance_frame(dt), cpu.advance_frame(dt)
        assert sg["pcg_iterations"] == sc["pcg_iterations"]
        assert sg["viscosity_iterations"] == sc["viscosity_iterations"]
        assert sg["max_velocity"] == sc["max_velocity"]
        assert_fields_equal(gpu, cpu, f"order {order} frame {f}")


def test_pipe_refusals(hip_lib):
    say("[pipe] refusals")
    g = _abi.pipe_grid(20)
    for kw, code in (({"is_parallel": 0}, _abi.FS_ERR_UNSUPPORTED),
                     ({"dimensionless_form": 0}, _abi.FS_ERR_UNSUPPORTED),
                     ({"boundary_order_velocity": 2}, _abi.FS_ERR_UNSUPPORTED),
                     ({"csf_model": 1}, _abi.FS_ERR_UNSUPPORTED),
                     ({"precision": _abi.FS_FP32_MG}, _abi.FS_ERR_UNSUPPORTED),
                     ({"a": 1.5}, _abi.FS_ERR_BAD_ARG)):
        P, Q = _abi.pipe_params(**kw)
        with pytest.raises(_abi.FsError) as e:
            _abi.Pipe(hip_lib, "fs_", P, Q, g)
        assert e.value.code == code, kw
    P, Q = _abi.pipe_params()
    s = _abi.Pipe(hip_lib, "fs_", P, Q, g)
    # the air/water-only entry points refuse a pipe handle
    rc = hip_lib.fs_compute_curvatures(s.h)
    assert rc == _abi.FS_ERR_UNSUPPORTED


def test_pipe_shipped_scene_at_its_size(hip_lib, oracle_lib):
    """The reference script's scene at its own 51^3 grid, eight frames with
    its CFL 0.9: bit for bit every frame, and each projection converged to
    the script's tolerance (1e-6 on sqrt(r^T D^-1 r))."""
    gpu, cpu = pair(hip_lib, oracle_lib, n=51)
    for f in range(8):
        dt = cpu.cfl_dt(0.9)
        assert gpu.cfl_dt(0.9) == dt
        say(f"[pipe] 51^3 frame {f} dt={dt}")
        sg, sc = gpu.advance_frame(dt), cpu.advance_frame(dt)
        assert sg["pcg_iterations"] == sc["pcg_iterations"]
        assert sg["residual_precond"] < 1e-6
        assert_fields_equal(gpu, cpu, f"51^3 frame {f}")
