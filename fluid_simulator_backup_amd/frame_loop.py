"""Frame loop over one solver handle: SIMULATION_WORLD::AdvanceOneFrameThread.

The reference recomputes the time step before every frame,
dt = DetermineTimeStep() * CFL (SIMULATION_WORLD.h:288-312, the rate from
EULERIAN_FLUID_SOLVER_3D::CFLOneTimeStep, :2673-2785), then advances one
step.  A dt held from the first frame is not the reference's loop: the
faces accelerate under gravity and the CFL number of a fixed dt grows
without bound.  The loop below is shared by the bench and the tests, for the
HIP library and for the CPU checker alike (both expose cfl_dt and
advance_frame through ``_abi.Solver``).
"""
import ctypes as C

from . import _abi


class FrameError(RuntimeError):
    """A frame of the loop failed; names the frame, its dt and the residual."""

    def __init__(self, frame, dt, code, message, stats=None):
        self.frame, self.dt, self.code, self.stats = frame, dt, code, stats
        res = f", residual_rel {stats['residual_rel']:.3e}" if stats else ""
        its = f", {stats['pcg_iterations']} PCG iterations" if stats else ""
        super().__init__(f"frame {frame} (dt {dt:.6e}): {_abi.ERROR_NAMES.get(code, code)}: {message}{res}{its}")


def advance_frame_rc(solver, dt):
    """(return code, statistics dict) of one fs_advance_frame call."""
    st = _abi.FsStepStats()
    rc = solver._f("advance_frame")(solver.h, dt, C.byref(st))
    return rc, st.as_dict()


def _message(solver):
    if solver.prefix == "fs_":
        m = solver.lib.fs_last_error(solver.h)
        return m.decode() if m else ""
    return "CPU checker"


def courant(solver, dt, max_velocity):
    """dt * (max|u|/dx + max|v|/dy + max|w|/dz) on the MAC grid (the c_f
    term of CFLOneTimeStep, :2697-2700)."""
    g = solver.grid
    dx = (g.x_max - g.x_min) / (g.nx - 1)
    dy = (g.y_max - g.y_min) / (g.ny - 1)
    dz = (g.z_max - g.z_min) / (g.nz - 1)
    return dt * (max_velocity[0] / dx + max_velocity[1] / dy + max_velocity[2] / dz)


def advance_frames(solver, frames, cfl=0.5, gravity_norm=1, on_frame=None):
    """Advance `frames` frames, each with dt = cfl / rate recomputed from the
    current velocity.  Returns one record per frame: the frame, dt, the
    Courant number of dt at the frame's start velocity (cfl_start), and the
    step statistics.  Raises FrameError on the first frame whose step fails
    (non-convergence included) or whose dt cannot be formed."""
    recs = []
    vmax = None
    for f in range(frames):
        try:
            dt = solver.cfl_dt(cfl, gravity_norm)
        except _abi.FsError as e:
            raise FrameError(f, float("nan"), e.code, f"cfl_dt: {e}") from None
        rc, st = advance_frame_rc(solver, dt)
        if rc != _abi.FS_OK:
            raise FrameError(f, dt, rc, _message(solver), st)
        rec = {"frame": f, "dt": dt, "cfl_start": courant(solver, dt, vmax) if vmax is not None else 0.0}
        rec.update(st)
        recs.append(rec)
        # the projection's max |u| per axis is the velocity the next frame's
        # CFL rule reads (reinitialization changes only the level set)
        vmax = st["max_velocity"]
        if on_frame is not None:
            on_frame(rec)
    return recs
