"""fs_cfl_dt_noalloc, the handle-free CFL time step that allocates nothing,
against fs_cfl_dt on a handle holding the same fields (bit for bit), on two
calls in a row with the second on a smaller grid, and its refusals."""
import ctypes as C

import numpy as np
import pytest

from fluid_simulator_backup_amd import _abi, scenes
from helpers import smooth_velocity

pytestmark = pytest.mark.gpu


def _bind(lib):
    f = lib.fs_cfl_dt_noalloc
    f.argtypes = [C.POINTER(_abi.FsParams), C.POINTER(_abi.FsGrid), C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                  C.c_int, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _case(hip_lib, f, name, n, seed, gravity_norm):
    import torch
    phi, _, P, g = scenes.make(name, n)
    vel = smooth_velocity(n, seed=seed, amp=0.7)
    s = _abi.Solver(hip_lib, "fs_", P, g, 0)
    s.set_fields(phi, *vel)
    want = s.cfl_dt(0.5, gravity_norm)
    s.close()
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in vel]
    dt = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    rc = f(C.byref(P), C.byref(g), dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), 0.5, gravity_norm,
           dt.data_ptr(), C.c_void_p(stream.cuda_stream))
    assert rc == _abi.FS_OK
    torch.cuda.synchronize()
    got = float(dt.cpu()[0])
    print(name, n, "noalloc", got, "handle", want)
    return got, want


def test_noalloc_cfl_matches_handle(hip_lib):
    f = _bind(hip_lib)
    # two calls in a row, the second on a smaller grid: the partials of the
    # first must not leak into the second
    for name, n, seed, gn in (("rising_drop", 40, 3, 1), ("dam_break", 20, 9, 0)):
        got, want = _case(hip_lib, f, name, n, seed, gn)
        assert got == want


def test_noalloc_cfl_refusals(hip_lib):
    import torch
    f = _bind(hip_lib)
    P = _abi.default_params()
    small = torch.zeros(8, dtype=torch.float64, device="cuda")
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    # a grid beyond the scratch array's 2048 * 2^17 faces per component is
    # refused before any device work (the buffers are never read)
    g = _abi.unit_cube_grid(1000)
    p = small.data_ptr()
    assert f(C.byref(P), C.byref(g), p, p, p, 0.5, 1, dt.data_ptr(), None) == _abi.FS_ERR_OOM
    g = _abi.unit_cube_grid(16)
    assert f(C.byref(P), C.byref(g), None, p, p, 0.5, 1, dt.data_ptr(), None) == _abi.FS_ERR_BAD_ARG
    assert f(C.byref(P), C.byref(g), p, p, p, -1.0, 1, dt.data_ptr(), None) == _abi.FS_ERR_BAD_ARG
    Pt = _abi.default_params(surface_tension=0.07)
    assert f(C.byref(Pt), C.byref(g), p, p, p, 0.5, 1, dt.data_ptr(), None) == _abi.FS_ERR_UNSUPPORTED
    torch.cuda.synchronize()
