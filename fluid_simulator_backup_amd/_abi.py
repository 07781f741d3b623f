"""ctypes view of include/fs_abi.h.

The structs and the ``Solver`` wrapper are shared by the HIP library
(``libfs_hip.so``, prefix ``fs_``) and by the CPU checker used in tests
(prefix ``fso_``), so a parity test drives both sides with the same calls.
"""
import ctypes as C
import os

import numpy as np

FS_OK = 0
FS_ERR_BAD_ARG = -1
FS_ERR_OOM = -2
FS_ERR_HIP = -3
FS_ERR_NOT_CONVERGED = -4
FS_ERR_NONFINITE = -5
FS_ERR_UNSUPPORTED = -6

FS_FP64 = 0
FS_FP32_REFINE = 1
FS_FP32_MG = 2
FS_STOP_ABS_PRECOND = 0
FS_STOP_REL_RESIDUAL = 1
FS_SCHEDULE_COMPAT = 0
FS_SCHEDULE_FIXED = 1
FS_CURV_LEVELSET, FS_CURV_NORMALS, FS_CURV_BOTH, FS_CURV_NONE = 0, 1, 2, 3
FS_SOURCE_NONE = 0
FS_SOURCE_ADD = 1
FS_SOURCE_DELETE = 2
FS_KS_REINIT = 0
FS_KS_MG_SMOOTH = 1
FS_LAYOUT_INTERIOR = 0
FS_LAYOUT_GHOSTED = 1
FS_MEM_HOST = 0
FS_MEM_DEVICE = 1

ERROR_NAMES = {
    FS_ERR_BAD_ARG: "FS_ERR_BAD_ARG",
    FS_ERR_OOM: "FS_ERR_OOM",
    FS_ERR_HIP: "FS_ERR_HIP",
    FS_ERR_NOT_CONVERGED: "FS_ERR_NOT_CONVERGED",
    FS_ERR_NONFINITE: "FS_ERR_NONFINITE",
    FS_ERR_UNSUPPORTED: "FS_ERR_UNSUPPORTED",
}

_D = C.c_double
_I = C.c_int


class FsGrid(C.Structure):
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I),
                ("x_min", _D), ("y_min", _D), ("z_min", _D),
                ("x_max", _D), ("y_max", _D), ("z_max", _D)]


class FsParams(C.Structure):
    _fields_ = [
        ("air_bubble_rising", _I),
        ("water_density", _D), ("air_density", _D),
        ("water_viscosity", _D), ("air_viscosity", _D),
        ("gravity", _D * 3),
        ("surface_tension", _D),
        ("use_delta_function_formulation", _I),
        ("order_for_time_advancing", _I),
        ("sussmanstyle_reinitialization", _I),
        ("scaling_number_for_reinitialization", _I),
        ("iteration_number_for_reinitialization", _I),
        ("levelset_epsilon", _D),
        ("velocity_epsilon", _D),
        ("use_5th_weno_v", _I),
        ("tolerance", _D),
        ("max_iteration", _I),
        ("use_previous_pressure", _I),
        ("stop_rule", _I),
        ("rel_tolerance", _D),
        ("precision", _I),
        ("schedule", _I),
        ("is_levelset_advection_active", _I),
        ("is_velocity_advection_active", _I),
        ("is_external_force_active", _I),
        ("is_sourcing_active", _I),
        ("is_projection_active", _I),
        ("semi_implicit_viscosity", _I),
        ("fast_weno", _I),
        ("csf_model", _I),
        ("curvature_method", _I),
    ]


class FsStepStats(C.Structure):
    _fields_ = [("pcg_iterations", _I), ("refinement_sweeps", _I),
                ("residual_precond", _D), ("residual_rel", _D),
                ("max_velocity", _D * 3), ("projection_ms", _D), ("step_ms", _D),
                ("viscosity_iterations", _I)]

    def as_dict(self):
        return {"pcg_iterations": self.pcg_iterations,
                "refinement_sweeps": self.refinement_sweeps,
                "residual_precond": self.residual_precond,
                "residual_rel": self.residual_rel,
                "max_velocity": list(self.max_velocity),
                "projection_ms": self.projection_ms,
                "step_ms": self.step_ms,
                "viscosity_iterations": self.viscosity_iterations}


def default_params(**kw):
    """Air/water defaults of the reference script (3D_script.sim:90-178),
    with zero surface tension, gravity along -z."""
    p = FsParams()
    p.air_bubble_rising = 1
    p.water_density, p.air_density = 1000.0, 1.0
    p.water_viscosity, p.air_viscosity = 1e-3, 1.8e-5
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, 0.0, -9.8
    p.surface_tension = 0.0
    p.use_delta_function_formulation = 1
    p.order_for_time_advancing = 1
    p.sussmanstyle_reinitialization = 1
    p.scaling_number_for_reinitialization = 5
    p.iteration_number_for_reinitialization = 25
    p.levelset_epsilon = 1e-6
    p.velocity_epsilon = 1e-6
    p.use_5th_weno_v = 0
    p.tolerance = 1e-5
    p.max_iteration = 10000
    p.use_previous_pressure = 1
    p.stop_rule = FS_STOP_ABS_PRECOND
    p.rel_tolerance = 1e-8
    p.precision = FS_FP64
    p.schedule = FS_SCHEDULE_COMPAT
    p.is_levelset_advection_active = 1
    p.is_velocity_advection_active = 1
    p.is_external_force_active = 1
    p.is_sourcing_active = 1
    p.is_projection_active = 1
    p.semi_implicit_viscosity = 0
    p.fast_weno = 0
    p.csf_model = 0
    p.curvature_method = FS_CURV_LEVELSET
    for k, v in kw.items():
        if k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def unit_cube_grid(nx, ny=None, nz=None):
    """Cell-centred grid on the unit cube: first/last centres at dx/2 and
    1-dx/2, so that the reference's (max-min)/(res-1) spacing is 1/n."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    g = FsGrid()
    g.nx, g.ny, g.nz = nx, ny, nz
    g.x_min, g.y_min, g.z_min = 0.5 / nx, 0.5 / ny, 0.5 / nz
    g.x_max, g.y_max, g.z_max = 1 - 0.5 / nx, 1 - 0.5 / ny, 1 - 0.5 / nz
    return g


_PD = C.POINTER(C.c_double)


class FsLevelsetSource(C.Structure):
    _fields_ = [("center", _D * 3), ("radius", _D), ("half_width", _D), ("mode", _I),
                ("start_frame", _I), ("stop_frame", _I)]


class FsWorldOptions(C.Structure):
    _fields_ = [("precision", _I), ("stop_rule", _I), ("rel_tolerance", _D), ("schedule", _I)]





def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)


def bind(lib, prefix):
    """Declare the argument types of the ABI functions in ``lib``."""
    f = lambda n: getattr(lib, prefix + n)
    vp = C.c_void_p
    f("create").argtypes = ([C.POINTER(FsParams), C.POINTER(FsGrid)] +
                            ([_I] if prefix == "fs_" else []) + [C.POINTER(vp)])
    f("create").restype = _I
    f("destroy").argtypes = [vp]
    f("destroy").restype = None
    f("set_fields").argtypes = [vp, _PD, _PD, _PD, _PD, _PD]
    f("get_fields").argtypes = [vp, _PD, _PD, _PD, _PD, _PD]
    for n in ("write_levelset", "read_levelset", "write_checkpoint", "read_checkpoint"):
        f(n).argtypes = [vp, C.c_char_p]
    for n in ("advance_frame", "step"):
        f(n).argtypes = [vp, _D, C.POINTER(FsStepStats)]
    f("cfl_dt").argtypes = [vp, _D, _I, C.POINTER(_D)]
    f("compute_curvatures").argtypes = [vp]
    f("set_levelset_sources").argtypes = [vp, _I, C.POINTER(FsLevelsetSource)]
    f("set_frame").argtypes = [vp, _I]
    f("apply_levelset_sources").argtypes = [vp]
    f("get_curvature").argtypes = [vp, _PD, _PD]
    for n in ("levelset_advection", "velocity_advection", "apply_gravity", "sourcing", "reinitialization"):
        f(n).argtypes = [vp, _D]
    f("projection").argtypes = [vp, _D, C.POINTER(FsStepStats)]
    f("assemble_pressure_system").argtypes = [vp, _D, _PD, _PD, _PD, _PD, _PD]
    f("apply_pressure_operator").argtypes = [vp, _PD, _PD]
    if prefix == "fs_":
        lib.fs_time_pcg_kernels.argtypes = [vp, _I, _I, _PD]
        lib.fs_set_kernel_sampling.argtypes = [vp, _I]
        lib.fs_get_kernel_samples.argtypes = [vp, _I, _PD, C.POINTER(_I)]
        lib.fs_get_frame_kernel_samples.argtypes = [vp, _I, _PD, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        lib.fs_set_fields_ex.argtypes = [vp, vp, vp, vp, vp, vp, _I, _I]
        lib.fs_get_fields_ex.argtypes = [vp, vp, vp, vp, vp, vp, _I, _I]
        lib.fs_set_async.argtypes = [vp, _I]
        lib.fs_synchronize.argtypes = [vp]
        lib.fs_get_stream.argtypes = [vp, C.POINTER(vp)]
        lib.fs_last_error.argtypes = [vp]
        lib.fs_last_error.restype = C.c_char_p
        lib.fs_adjoint_projection.argtypes = [vp, _D, _PD, _PD, _PD, C.POINTER(FsStepStats)]
        lib.fs_adjoint_sourcing.argtypes = [vp, _D, _PD, _PD, _PD]
        lib.fs_weno5_transport.argtypes = [vp, _D, _PD, _PD, _PD, _PD, _PD]
        lib.fs_adjoint_weno5_transport.argtypes = [vp, _D] + [_PD] * 9
        lib.fs_adjoint_step.argtypes = [vp, _D, _PD, _PD, _PD, _PD, C.POINTER(FsStepStats)]
        lib.fs_levelset_loss.argtypes = [vp, _PD, C.POINTER(_D)]
        lib.fs_adjoint_rollout.argtypes = ([vp, _I, _D, _I, _PD, C.POINTER(_D)] + [_PD] * 4 +
                                           [C.POINTER(FsStepStats)])
        lib.fs_batch_create.argtypes = [_I, C.POINTER(FsParams), C.POINTER(FsGrid), _I, _I, C.POINTER(vp)]
        lib.fs_batch_destroy.argtypes = [vp]
        lib.fs_batch_destroy.restype = None
        lib.fs_batch_scene.argtypes = [vp, _I]
        lib.fs_batch_scene.restype = vp
        lib.fs_batch_advance_frame.argtypes = [vp, _PD, C.POINTER(FsStepStats), C.POINTER(_I)]
        lib.fs_batch_last_error.argtypes = [vp]
        lib.fs_batch_last_error.restype = C.c_char_p
        cs = C.c_char_p
        lib.fs_script_open.argtypes = [cs, C.POINTER(vp)]
        lib.fs_script_close.argtypes = [vp]
        lib.fs_script_close.restype = None
        lib.fs_script_get_int.argtypes = [vp, cs, cs, _I, C.POINTER(_I)]
        lib.fs_script_get_float.argtypes = [vp, cs, cs, _D, C.POINTER(_D)]
        lib.fs_script_get_bool.argtypes = [vp, cs, cs, _I, C.POINTER(_I)]
        lib.fs_script_get_vector3.argtypes = [vp, cs, cs, _PD, _PD]
        lib.fs_script_get_int3.argtypes = [vp, cs, cs, C.POINTER(_I), C.POINTER(_I)]
        lib.fs_script_get_string.argtypes = [vp, cs, cs, cs, C.c_char_p, _I]
        lib.fs_script_child_count.argtypes = [vp, cs, C.POINTER(_I)]
        lib.fs_script_child_name.argtypes = [vp, cs, _I, C.c_char_p, _I]
        lib.fs_world_create.argtypes = [cs, C.POINTER(FsWorldOptions), _I, C.POINTER(vp)]
        lib.fs_world_destroy.argtypes = [vp]
        lib.fs_world_destroy.restype = None
        lib.fs_world_last_error.argtypes = [vp]
        lib.fs_world_last_error.restype = C.c_char_p
        lib.fs_world_solver.argtypes = [vp]
        lib.fs_world_solver.restype = vp
        lib.fs_world_get_info.argtypes = [vp, C.POINTER(FsWorldInfo)]
        lib.fs_world_advance_frame.argtypes = [vp, C.POINTER(FsStepStats)]
        lib.fs_world_parse.argtypes = [cs, C.POINTER(FsWorldOptions), C.POINTER(FsWorldInfo),
                                       C.POINTER(FsLevelsetSource), _I]
    return lib


class FsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Solver:
    """Thin wrapper over one handle of either library.

    Mirrors EULERIAN_FLUID_SOLVER_3D's call surface: fields in and out,
    one frame / one step, the phase functions and the CFL time step."""

    def __init__(self, lib, prefix, params, grid, device=0, handle=None):
        self.lib, self.prefix = lib, prefix
        self.params, self.grid = params, grid
        self.nx, self.ny, self.nz = grid.nx, grid.ny, grid.nz
        self.owned = handle is None
        if handle is not None:  # a scene of a batch: the batch owns it
            self.h = C.c_void_p(handle)
            return
        h = C.c_void_p()
        args = [C.byref(params), C.byref(grid)] + ([device] if prefix == "fs_" else []) + [C.byref(h)]
        rc = self._f("create")(*args)
        if rc != FS_OK:
            raise FsError(rc, "create failed")
        self.h = h

    def _f(self, n):
        return getattr(self.lib, self.prefix + n)

    def _check(self, rc, what, allow=()):
        if rc != FS_OK and rc not in allow:
            msg = what
            if self.prefix == "fs_":
                m = self.lib.fs_last_error(self.h)
                msg = f"{what}: {m.decode() if m else ''}"
            raise FsError(rc, msg)
        return rc

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # shapes, x fastest: arrays are indexed [k, j, i]
    def shape(self, comp):
        nx, ny, nz = self.nx, self.ny, self.nz
        return {"phi": (nz, ny, nx), "p": (nz, ny, nx), "u": (nz, ny, nx + 1),
                "v": (nz, ny + 1, nx), "w": (nz + 1, ny, nx)}[comp]

    def set_fields(self, phi=None, u=None, v=None, w=None, p=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        phi, u, v, w, p = map(conv, (phi, u, v, w, p))
        self._check(self._f("set_fields")(self.h, _ptr(phi), _ptr(u), _ptr(v), _ptr(w), _ptr(p)), "set_fields")

    def get_fields(self):
        out = {c: np.empty(self.shape(c)) for c in ("phi", "u", "v", "w", "p")}
        self._check(self._f("get_fields")(self.h, *[_ptr(out[c]) for c in ("phi", "u", "v", "w", "p")]),
                    "get_fields")
        return out

    def write_levelset(self, path):
        self._check(self._f("write_levelset")(self.h, os.fsencode(path)), "write_levelset")

    def read_levelset(self, path):
        self._check(self._f("read_levelset")(self.h, os.fsencode(path)), "read_levelset")

    def write_checkpoint(self, path):
        self._check(self._f("write_checkpoint")(self.h, os.fsencode(path)), "write_checkpoint")

    def read_checkpoint(self, path):
        self._check(self._f("read_checkpoint")(self.h, os.fsencode(path)), "read_checkpoint")

    def advance_frame(self, dt, allow_not_converged=False):
        st = FsStepStats()
        self._check(self._f("advance_frame")(self.h, dt, C.byref(st)), "advance_frame",
                    (FS_ERR_NOT_CONVERGED,) if allow_not_converged else ())
        return st.as_dict()

    def step(self, dt, allow_not_converged=False):
        st = FsStepStats()
        self._check(self._f("step")(self.h, dt, C.byref(st)), "step",
                    (FS_ERR_NOT_CONVERGED,) if allow_not_converged else ())
        return st.as_dict()

    def cfl_dt(self, cfl, gravity_norm=1):
        d = C.c_double()
        self._check(self._f("cfl_dt")(self.h, cfl, gravity_norm, C.byref(d)), "cfl_dt")
        return d.value

    def compute_curvatures(self):
        self._check(self._f("compute_curvatures")(self.h), "compute_curvatures")

    def curvature(self):
        """(curvature (nz, ny, nx), abs_max_curvature) of the current level set."""
        k = np.empty((self.nz, self.ny, self.nx))
        m = C.c_double()
        self._check(self._f("get_curvature")(self.h, _ptr(k), C.byref(m)), "get_curvature")
        return k, m.value

    def set_sources(self, sources):
        """Level-set sources: a list of dicts with center, radius, half_width,
        mode and optional start_frame / stop_frame."""
        arr = (FsLevelsetSource * max(1, len(sources)))()
        for a, d in zip(arr, sources):
            for i in range(3):
                a.center[i] = d["center"][i]
            a.radius, a.half_width, a.mode = d["radius"], d["half_width"], d["mode"]
            a.start_frame = d.get("start_frame", -(2 ** 31 - 1))
            a.stop_frame = d.get("stop_frame", 2 ** 31 - 1)
        self._check(self._f("set_levelset_sources")(self.h, len(sources), arr), "set_levelset_sources")

    def set_frame(self, frame):
        self._check(self._f("set_frame")(self.h, frame), "set_frame")

    def apply_sources(self):
        self._check(self._f("apply_levelset_sources")(self.h), "apply_levelset_sources")

    def phase(self, name, dt):
        self._check(self._f(name)(self.h, dt), name)

    def projection(self, dt, allow_not_converged=False):
        st = FsStepStats()
        self._check(self._f("projection")(self.h, dt, C.byref(st)), "projection",
                    (FS_ERR_NOT_CONVERGED,) if allow_not_converged else ())
        return st.as_dict()

    def assemble(self, dt):
        nx, ny, nz = self.nx, self.ny, self.nz
        out = {"diag": np.empty((nz, ny, nx)), "bx": np.empty((nz, ny, nx + 1)),
               "by": np.empty((nz, ny + 1, nx)), "bz": np.empty((nz + 1, ny, nx)),
               "rhs": np.empty((nz, ny, nx))}
        self._check(self._f("assemble_pressure_system")(self.h, dt, *[_ptr(out[c]) for c in
                                                                      ("diag", "bx", "by", "bz", "rhs")]),
                    "assemble")
        return out

    def adjoint_projection(self, dt, lu, lv, lw):
        """lam -> P^T lam for the projection at the current level set."""
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (lu, lv, lw)]
        st = FsStepStats()
        self._check(self.lib.fs_adjoint_projection(self.h, dt, *[_ptr(a) for a in out], C.byref(st)),
                    "adjoint_projection")
        return out, st.as_dict()

    def adjoint_sourcing(self, dt, lu, lv, lw):
        """lam -> V^T lam for the explicit viscous update at the current level set."""
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (lu, lv, lw)]
        self._check(self.lib.fs_adjoint_sourcing(self.h, dt, *[_ptr(a) for a in out]), "adjoint_sourcing")
        return out

    def weno5_transport(self, dt, phi, u, v, w):
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        phi, u, v, w = map(c, (phi, u, v, w))
        out = np.empty_like(phi)
        self._check(self.lib.fs_weno5_transport(self.h, dt, _ptr(phi), _ptr(u), _ptr(v), _ptr(w), _ptr(out)),
                    "weno5_transport")
        return out

    def adjoint_weno5_transport(self, dt, phi, u, v, w, lam):
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        phi, u, v, w, lam = map(c, (phi, u, v, w, lam))
        out = [np.empty_like(phi), np.empty_like(u), np.empty_like(v), np.empty_like(w)]
        self._check(self.lib.fs_adjoint_weno5_transport(self.h, dt, _ptr(phi), _ptr(u), _ptr(v), _ptr(w), _ptr(lam),
                                                        *[_ptr(a) for a in out]), "adjoint_weno5_transport")
        return out

    def adjoint_step(self, dt, lphi, lu, lv, lw, allow_not_converged=False):
        """Adjoint of one fs_step at the current fields: the adjoint of the
        step's output in, of its input out (the fields are left as they were)."""
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (lphi, lu, lv, lw)]
        st = FsStepStats()
        rc = self.lib.fs_adjoint_step(self.h, dt, *[_ptr(a) for a in out], C.byref(st))
        self._check(rc, "adjoint_step", allow=(FS_ERR_NOT_CONVERGED,) if allow_not_converged else ())
        return out, st.as_dict()

    def levelset_loss(self, phi_target):
        t = np.ascontiguousarray(phi_target, dtype=np.float64)
        out = _D()
        self._check(self.lib.fs_levelset_loss(self.h, _ptr(t), C.byref(out)), "levelset_loss")
        return out.value

    def adjoint_rollout(self, steps, dt, checkpoint_every, phi_target, allow_not_converged=False):
        """Loss of the level set after `steps` steps and its gradient with
        respect to the initial (phi, u, v, w); the fields are left at the
        initial state."""
        t = np.ascontiguousarray(phi_target, dtype=np.float64)
        f = self.get_fields()
        grads = [np.empty_like(f[c]) for c in ("phi", "u", "v", "w")]
        loss = _D()
        st = FsStepStats()
        self._check(self.lib.fs_adjoint_rollout(self.h, steps, dt, checkpoint_every, _ptr(t), C.byref(loss),
                                                *[_ptr(g) for g in grads], C.byref(st)), "adjoint_rollout",
                    (FS_ERR_NOT_CONVERGED,) if allow_not_converged else ())
        return loss.value, grads, st.as_dict()

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.empty_like(x)
        self._check(self._f("apply_pressure_operator")(self.h, _ptr(x), _ptr(q)), "apply_operator")
        return q


class Batch:
    """Many independent scenes advanced together (fs_batch_*)."""

    def __init__(self, lib, params, grids, device=0, threads=0):
        self.lib = lib
        n = len(params)
        self.params = (FsParams * n)(*params)
        self.grids = (FsGrid * n)(*grids)
        h = C.c_void_p()
        rc = lib.fs_batch_create(n, self.params, self.grids, device, threads, C.byref(h))
        if rc != FS_OK:
            raise FsError(rc, "fs_batch_create failed")
        self.h = h
        self.scenes = [Solver(lib, "fs_", params[i], grids[i], handle=lib.fs_batch_scene(h, i)) for i in range(n)]

    def advance_frame(self, dts, allow_not_converged=False):
        n = len(self.scenes)
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        st = (FsStepStats * n)()
        rcs = (_I * n)()
        rc = self.lib.fs_batch_advance_frame(self.h, _ptr(dt), st, rcs)
        bad = [r for r in rcs if r != FS_OK and not (allow_not_converged and r == FS_ERR_NOT_CONVERGED)]
        if bad:
            m = self.lib.fs_batch_last_error(self.h)
            raise FsError(bad[0], m.decode() if m else "batch")
        return [x.as_dict() for x in st]

    def close(self):
        if getattr(self, "h", None):
            for sc in self.scenes:
                sc.h = None
            self.lib.fs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def time_pcg_kernels(solver, precision, reps):
    out = np.zeros(3)
    solver._check(solver.lib.fs_time_pcg_kernels(solver.h, precision, reps, _ptr(out)), "time_pcg_kernels")
    return out


def set_kernel_sampling(solver, enable):
    solver._check(solver.lib.fs_set_kernel_sampling(solver.h, int(bool(enable))), "set_kernel_sampling")


def kernel_samples(solver, precision):
    """Mean in-loop duration (ms) and sample count of the three PCG kernels."""
    avg = np.zeros(3)
    cnt = (_I * 3)()
    solver._check(solver.lib.fs_get_kernel_samples(solver.h, precision, _ptr(avg), cnt), "get_kernel_samples")
    return avg, np.array(list(cnt))


def frame_kernel_samples(solver, kernel):
    """Mean in-loop duration (ms), timed launches and all launches of a
    frame kernel (FS_KS_REINIT, FS_KS_MG_SMOOTH) since sampling was enabled."""
    avg = C.c_double()
    n, launches = C.c_longlong(), C.c_longlong()
    solver._check(solver.lib.fs_get_frame_kernel_samples(solver.h, kernel, C.byref(avg), C.byref(n),
                                                         C.byref(launches)), "get_frame_kernel_samples")
    return avg.value, n.value, launches.value


def load(path, prefix):
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    return bind(C.CDLL(path), prefix)


def _opts(precision=FS_FP64, stop_rule=FS_STOP_ABS_PRECOND, rel_tolerance=1e-8, schedule=FS_SCHEDULE_COMPAT):
    o = FsWorldOptions()
    o.precision, o.stop_rule, o.rel_tolerance, o.schedule = precision, stop_rule, rel_tolerance, schedule
    return o


class Script:
    """A parsed .sim script (SCRIPT_READER semantics) of the HIP library."""

    def __init__(self, lib, path):
        self.lib = lib
        self.h = C.c_void_p()
        rc = lib.fs_script_open(os.fsencode(path), C.byref(self.h))
        if rc != FS_OK:
            raise FsError(rc, f"cannot read {path}")

    def close(self):
        if self.h:
            self.lib.fs_script_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def int(self, block, key, default=0):
        v = _I()
        self.lib.fs_script_get_int(self.h, block.encode(), key.encode(), default, C.byref(v))
        return v.value

    def float(self, block, key, default=0.0):
        v = _D()
        self.lib.fs_script_get_float(self.h, block.encode(), key.encode(), default, C.byref(v))
        return v.value

    def bool(self, block, key, default=False):
        v = _I()
        self.lib.fs_script_get_bool(self.h, block.encode(), key.encode(), int(default), C.byref(v))
        return bool(v.value)

    def vector3(self, block, key, default=(0.0, 0.0, 0.0)):
        d, o = (_D * 3)(*default), (_D * 3)()
        self.lib.fs_script_get_vector3(self.h, block.encode(), key.encode(), d, o)
        return tuple(o)

    def int3(self, block, key, default=(0, 0, 0)):
        d, o = (_I * 3)(*default), (_I * 3)()
        self.lib.fs_script_get_int3(self.h, block.encode(), key.encode(), d, o)
        return tuple(o)

    def string(self, block, key, default="Null"):
        buf = C.create_string_buffer(4096)
        self.lib.fs_script_get_string(self.h, block.encode(), key.encode(), default.encode(), buf, 4096)
        return buf.value.decode()

    def children(self, block=""):
        n = _I()
        self.lib.fs_script_child_count(self.h, block.encode(), C.byref(n))
        out = []
        for i in range(n.value):
            buf = C.create_string_buffer(256)
            self.lib.fs_script_child_name(self.h, block.encode(), i, buf, 256)
            out.append(buf.value.decode())
        return out


def parse_world(lib, path, **opts):
    """(FsWorldInfo, [FsLevelsetSource]) a script sets up, without a device."""
    info = FsWorldInfo()
    src = (FsLevelsetSource * 64)()
    o = _opts(**opts)
    rc = lib.fs_world_parse(os.fsencode(path), C.byref(o), C.byref(info), src, 64)
    if rc != FS_OK:
        raise FsError(rc, f"fs_world_parse {path}")
    return info, [src[i] for i in range(min(64, info.n_sources))]


class World:
    """SIMULATION_WORLD from a .sim script on the HIP library."""

    def __init__(self, lib, path, device=0, **opts):
        self.lib = lib
        self.h = C.c_void_p()
        o = _opts(**opts)
        rc = lib.fs_world_create(os.fsencode(path), C.byref(o), device, C.byref(self.h))
        if rc != FS_OK:
            raise FsError(rc, f"fs_world_create {path}")
        info = self.info()
        self.solver = Solver(lib, "fs_", info.params, info.grid, device, handle=lib.fs_world_solver(self.h))

    def info(self):
        i = FsWorldInfo()
        self.lib.fs_world_get_info(self.h, C.byref(i))
        return i

    def advance_frame(self, allow_not_converged=False):
        st = FsStepStats()
        rc = self.lib.fs_world_advance_frame(self.h, C.byref(st))
        if rc != FS_OK and not (allow_not_converged and rc == FS_ERR_NOT_CONVERGED):
            m = self.lib.fs_world_last_error(self.h)
            raise FsError(rc, f"advance_frame: {m.decode() if m else ''}")
        return st.as_dict()

    def close(self):
        if self.h:
            self.lib.fs_world_destroy(self.h)
            self.h = C.c_void_p()


# ---- oil/water pipe mode ---------------------------------------------------

FS_PIPE_LEVELSET_ADVECTION, FS_PIPE_VELOCITY_ADVECTION, FS_PIPE_GRAVITY = 0, 1, 2
FS_PIPE_SOURCING, FS_PIPE_PROJECTION, FS_PIPE_REINITIALIZATION = 3, 4, 5
FS_PIPE_KS_SPMV = 0
FS_PIPE_KS_UPDATE = 1
PIPE_ARRAYS = ["phi", "scalar_ghost", "ghost_u", "ghost_v", "ghost_w",
               "adv_u_y", "adv_u_z", "adv_v_y", "adv_v_z", "adv_w_y", "adv_w_z",
               "vis_u_y", "vis_u_z", "vis_v_y", "vis_v_z", "vis_w_y", "vis_w_z",
               "boundary", "divergence", "jump", "pressure"]
_PIPE_ARRAY_COMP = {"ghost_u": 1, "ghost_v": 2, "ghost_w": 3, "adv_u_y": 1, "adv_u_z": 1, "adv_v_y": 2,
                    "adv_v_z": 2, "adv_w_y": 3, "adv_w_z": 3, "vis_u_y": 1, "vis_u_z": 1, "vis_v_y": 2,
                    "vis_v_z": 2, "vis_w_y": 3, "vis_w_z": 3}


class FsPipeParams(C.Structure):
    _fields_ = [("is_parallel", _I), ("a", _D), ("R2", _D), ("A_0", _D), ("alpha", _D),
                ("m", _D), ("eta", _D), ("K", _D), ("g", _D),
                ("oil_density", _D), ("oil_viscosity", _D),
                ("dimensionless_form", _I), ("cylindrical_boundary", _I), ("boundary_order_velocity", _I)]


class FsWorldInfo(C.Structure):
    _fields_ = [("params", FsParams), ("grid", FsGrid),
                ("dt", _D), ("max_dt", _D), ("cfl", _D), ("frame_rate", _D), ("accu_dt", _D),
                ("num_substeps", _I), ("frame", _I), ("last_frame", _I), ("n_sources", _I),
                ("oil_water", _I), ("pipe", FsPipeParams)]


# the pipe keys of the shipped scripts: (is_parallel, a, A_0, alpha, K)
PIPE_SCRIPTS = {
    "3D_script.sim": (1, 1.28, 0.001, 2.0, -0.4552),          # parallel, 3D_script.sim:35-39, 229
    "3D_script_1.sim": (0, 1.61, 0.1, 2.4, -2.03030303),     # vertical, 3D_script_1.sim:35-39, 229
    "3D_script_3.sim": (0, 1.7, 0.01, 1.975, -0.542709),     # vertical, 3D_script_3.sim:35-38, 229
}


def pipe_params(script="3D_script.sim", **kw):
    """fs_params and fs_pipe_params of one of the reference's oil/water
    scripts (3D_script.sim:32-40, 195-290 and the vertical pipes of
    3D_script_1.sim and 3D_script_3.sim): core-annular flow, dimensionless
    form.  3D_script_2.sim is 3D_script.sim's pipe with the vertical grid."""
    p = default_params(water_density=0.995, water_viscosity=0.01, surface_tension=8.54,
                       gravity=(0.0, -980.0, 0.0), order_for_time_advancing=2,
                       scaling_number_for_reinitialization=2, iteration_number_for_reinitialization=10,
                       tolerance=1e-6)
    q = FsPipeParams()
    par, a, A_0, alpha, K = PIPE_SCRIPTS[script]
    q.is_parallel = par
    q.a, q.R2, q.A_0, q.alpha = a, 0.47625, A_0, alpha
    q.m, q.eta, q.K, q.g = 0.00166, 1.1, K, -980.0
    q.oil_density, q.oil_viscosity = 0.905, 6.01
    q.dimensionless_form, q.cylindrical_boundary, q.boundary_order_velocity = 1, 1, 1
    for k, v in kw.items():
        if hasattr(q, k) and k not in ("water_density",):
            setattr(q, k, v)
        elif k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        else:
            setattr(p, k, v)
    return p, q


def pipe_grid(n=51):
    """GRID_STRUCTURE_3D_PARALLEL of the script (3D_script.sim:86-93) at n^3."""
    g = FsGrid()
    g.nx = g.ny = g.nz = n
    g.x_min, g.y_min, g.z_min = 0.0, -1.570796, -1.570796
    g.x_max, g.y_max, g.z_max = 3.141592, 1.570796, 1.570796
    return g


def vertical_pipe_grid(nxz=71, ny=51, half=1.8326, length=2.618):
    """GRID_STRUCTURE_3D_VERTICAL of 3D_script_1.sim (:78-85): nxz x ny x nxz
    nodes, x and z in [-half, half], y (the pipe's axis) in [0, length];
    3D_script_3.sim has half 2.2270, length 3.1814."""
    g = FsGrid()
    g.nx, g.ny, g.nz = nxz, ny, nxz
    g.x_min, g.y_min, g.z_min = -half, 0.0, -half
    g.x_max, g.y_max, g.z_max = half, length, half
    return g


def bind_pipe(lib, prefix):
    vp = C.c_void_p
    if prefix == "fs_":
        lib.fs_create_pipe.argtypes = [C.POINTER(FsParams), C.POINTER(FsPipeParams), C.POINTER(FsGrid), _I,
                                       C.POINTER(vp)]
        lib.fs_last_error.argtypes = [vp]
        lib.fs_last_error.restype = C.c_char_p
        names = {"init": "fs_pipe_initial_fields", "constants": "fs_pipe_constants", "phase": "fs_pipe_phase",
                 "array": "fs_pipe_get_array", "step": "fs_step", "frame": "fs_advance_frame",
                 "set": "fs_set_fields", "get": "fs_get_fields", "destroy": "fs_destroy"}
        getattr(lib, names["init"]).argtypes = [vp]
        getattr(lib, names["constants"]).argtypes = [vp, _PD]
        lib.fs_cfl_dt.argtypes = [vp, _D, _I, C.POINTER(_D)]
        lib.fs_set_kernel_sampling.argtypes = [vp, _I]
        lib.fs_pipe_kernel_samples.argtypes = [vp, _I, _PD, _PD, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    else:
        lib.fsop_create.argtypes = [C.POINTER(FsParams), C.POINTER(FsPipeParams), C.POINTER(FsGrid), C.POINTER(vp)]
        names = {"init": "fsop_initial_fields", "constants": "fsop_constants", "phase": "fsop_phase",
                 "array": "fsop_get_array", "step": "fsop_step", "frame": "fsop_advance_frame",
                 "set": "fsop_set_fields", "get": "fsop_get_fields", "destroy": "fsop_destroy"}
        lib.fsop_initial_fields.argtypes = [vp]
        lib.fsop_constants.argtypes = [vp, _PD]
        lib.fsop_cfl_dt.argtypes = [vp, _D, C.POINTER(_D)]
        lib.fsop_compute_curvatures.argtypes = [vp]
    getattr(lib, names["phase"]).argtypes = [vp, _I, _D, C.POINTER(FsStepStats)]
    getattr(lib, names["array"]).argtypes = [vp, _I, _PD]
    getattr(lib, names["step"]).argtypes = [vp, _D, C.POINTER(FsStepStats)]
    getattr(lib, names["frame"]).argtypes = [vp, _D, C.POINTER(FsStepStats)]
    getattr(lib, names["set"]).argtypes = [vp, _PD, _PD, _PD, _PD, _PD]
    getattr(lib, names["get"]).argtypes = [vp, _PD, _PD, _PD, _PD, _PD]
    getattr(lib, names["destroy"]).argtypes = [vp]
    getattr(lib, names["destroy"]).restype = None
    return names


class Pipe:
    """One oil/water pipe scene on either library: the HIP library's
    fs_create_pipe handle (prefix "fs_") or the CPU checker's (prefix "fsop_")."""

    PHASES = {"levelset_advection": FS_PIPE_LEVELSET_ADVECTION, "velocity_advection": FS_PIPE_VELOCITY_ADVECTION,
              "gravity": FS_PIPE_GRAVITY, "sourcing": FS_PIPE_SOURCING, "projection": FS_PIPE_PROJECTION,
              "reinitialization": FS_PIPE_REINITIALIZATION}

    def __init__(self, lib, prefix, params, pipe, grid, device=0, handle=None):
        self.lib, self.prefix, self.params, self.pipe, self.grid = lib, prefix, params, pipe, grid
        self.nx, self.ny, self.nz = grid.nx, grid.ny, grid.nz
        self.names = bind_pipe(lib, prefix)
        self.owned = handle is None
        if handle is not None:  # a world's solver: the world owns it
            self.h = C.c_void_p(handle)
            return
        h = C.c_void_p()
        if prefix == "fs_":
            rc = lib.fs_create_pipe(C.byref(params), C.byref(pipe), C.byref(grid), device, C.byref(h))
        else:
            rc = lib.fsop_create(C.byref(params), C.byref(pipe), C.byref(grid), C.byref(h))
        if rc != FS_OK:
            raise FsError(rc, "create_pipe failed")
        self.h = h

    def _call(self, key, *args, allow=()):
        rc = getattr(self.lib, self.names[key])(self.h, *args)
        if rc != FS_OK and rc not in allow:
            msg = key
            if self.prefix == "fs_":
                m = self.lib.fs_last_error(self.h)
                msg = f"{key}: {m.decode() if m else ''}"
            raise FsError(rc, msg)
        return rc

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                getattr(self.lib, self.names["destroy"])(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, comp):
        return Solver.shape(self, comp)

    def initial_fields(self):
        self._call("init")

    def constants(self):
        out = np.empty(7)
        self._call("constants", _ptr(out))
        return dict(zip(("R1", "f", "A", "V0", "J", "Re", "We"), out.tolist()))

    def set_fields(self, phi=None, u=None, v=None, w=None, p=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        arrs = list(map(conv, (phi, u, v, w, p)))
        self._call("set", *[_ptr(a) for a in arrs])

    def get_fields(self):
        out = {c: np.empty(self.shape(c)) for c in ("phi", "u", "v", "w", "p")}
        self._call("get", *[_ptr(out[c]) for c in ("phi", "u", "v", "w", "p")])
        return out

    def phase(self, name, dt):
        st = FsStepStats()
        self._call("phase", self.PHASES[name], dt, C.byref(st))
        return st.as_dict()

    def step(self, dt):
        st = FsStepStats()
        self._call("step", dt, C.byref(st))
        return st.as_dict()

    def advance_frame(self, dt):
        st = FsStepStats()
        self._call("frame", dt, C.byref(st))
        return st.as_dict()

    def cfl_dt(self, cfl):
        d = C.c_double()
        if self.prefix == "fs_":
            rc = self.lib.fs_cfl_dt(self.h, cfl, 0, C.byref(d))
        else:
            rc = self.lib.fsop_cfl_dt(self.h, cfl, C.byref(d))
        if rc != FS_OK:
            raise FsError(rc, "cfl_dt")
        return d.value

    def set_sampling(self, enable):
        """In-loop timing of the PCG loop kernels (HIP library only)."""
        rc = self.lib.fs_set_kernel_sampling(self.h, int(bool(enable)))
        if rc != FS_OK:
            raise FsError(rc, "set_kernel_sampling")

    def kernel_samples(self, kernel):
        """(mean ms, mean algorithmic bytes, timed launches, launches) of
        FS_PIPE_KS_SPMV or FS_PIPE_KS_UPDATE since sampling was enabled."""
        ms, by = C.c_double(), C.c_double()
        n, launches = C.c_longlong(), C.c_longlong()
        rc = self.lib.fs_pipe_kernel_samples(self.h, kernel, C.byref(ms), C.byref(by), C.byref(n), C.byref(launches))
        if rc != FS_OK:
            raise FsError(rc, "pipe_kernel_samples")
        return ms.value, by.value, n.value, launches.value

    def array(self, name):
        """A persistent array over its box grown by 3 ghost layers, [k, j, i]."""
        c = _PIPE_ARRAY_COMP.get(name, 0)
        ni, nj, nk = self.nx + (c == 1), self.ny + (c == 2), self.nz + (c == 3)
        out = np.empty((nk + 6, nj + 6, ni + 6))
        self._call("array", PIPE_ARRAYS.index(name), _ptr(out))
        return out
