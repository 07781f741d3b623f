"""The FP64 reference gradient: the CPU checker's own source compiled in
reverse mode (oracle/fs_oracle_ad.cpp, operator-overloading tape), pinned
before it is trusted:
  - its forward values equal the plain checker's bit for bit (the active
    type changes no operation);
  - its gradient of a rollout's level-set loss matches central differences
    of the plain checker's forward along seeded directions, at density ratio
    1.25 where the fixed schedule's map is smooth over the difference steps;
  - the loss's own derivative alone (zero steps) is the closed form
    2 (H(phi) - H(target)) H'(phi).
The GPU adjoint is compared with it in tests/test_gpu_adjoint_step.py."""
import ctypes as C

import numpy as np
import pytest

from fluid_simulator_backup_amd import _abi, scenes
from helpers import ad_gradient

FIXED = dict(schedule=_abi.FS_SCHEDULE_FIXED, use_5th_weno_v=1, stop_rule=_abi.FS_STOP_REL_RESIDUAL,
             rel_tolerance=1e-13, max_iteration=20000)
PD = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(PD)


def _loss(phi, target, eps):
    def H(f):
        return np.where(f < -eps, 0.0, np.where(f > eps, 1.0, 0.5 + f / (2 * eps) + np.sin(np.pi * f / eps) / (2 * np.pi)))
    d = H(phi) - H(target)
    dH = np.where(np.abs(phi) > eps, 0.0, (1 + np.cos(np.pi * phi / eps)) / (2 * eps))
    return float((d * d).sum()), 2 * d * dH


def _scene(n, ratio, seed):
    phi, _, P, g = scenes.make("slosh", n, air_density=1000.0 / ratio, **FIXED)
    vel = list(scenes.gaussian_random_velocity(n, seed, sigma_cells=max(2.0, n / 16)))
    return phi, vel, P, g


def _checker_rollout(oracle_lib, P, g, x, steps, dt):
    s = _abi.Solver(oracle_lib, "fso_", P, g)
    oracle_lib.fso_set_viscosity_inplace(s.h, 0)
    s.set_fields(x[0], x[1], x[2], x[3], p=np.zeros_like(x[0]))
    for _ in range(steps):
        s.step(dt, allow_not_converged=True)
    f = s.get_fields()
    s.close()
    return f


def test_reverse_build_forward_is_the_checker_bit_for_bit(oracle_lib, oracle_ad):
    n, steps, dt = 12, 2, 4e-3
    phi, vel, P, g = _scene(n, 1000.0, 40)
    ref = _checker_rollout(oracle_lib, P, g, [phi] + vel, steps, dt)
    f = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in [phi] + vel + [np.zeros_like(phi)]]
    rc = oracle_ad.fsa_rollout_values(C.byref(P), C.byref(g), *[_p(a) for a in f], 0, steps, dt)
    assert rc in (_abi.FS_OK, _abi.FS_ERR_NOT_CONVERGED), rc
    for c, a in zip(("phi", "u", "v", "w", "p"), f):
        assert np.array_equal(a, ref[c]), c


def test_loss_gradient_closed_form(oracle_ad):
    n = 12
    phi, vel, P, g = _scene(n, 1000.0, 40)
    target = phi + 0.02 * np.sin(7 * np.arange(phi.size).reshape(phi.shape))
    loss, grads, _ = ad_gradient(oracle_ad, P, g, [phi] + vel, 0, 4e-3, target)
    l2, lphi = _loss(phi, target, 1.5 / n)
    assert abs(loss - l2) <= 1e-13 * l2
    assert np.allclose(grads[0], lphi, rtol=1e-13, atol=1e-13 * np.abs(lphi).max())
    for gv in grads[1:]:
        assert not gv.any()


def test_reverse_gradient_matches_checker_differences(oracle_lib, oracle_ad):
    """Three steps at 12^3, density ratio 1.25, a target from another
    velocity: the reverse gradient along two smooth unit-RMS velocity
    directions and one level-set direction agrees with Richardson-
    extrapolated central differences (steps 1e-6, 5e-7) of the plain
    checker to 1e-6."""
    n, steps, dt = 12, 3, 4e-3
    phi, vel_t, P, g = _scene(n, 1.25, 40)
    target = _checker_rollout(oracle_lib, P, g, [phi] + vel_t, steps, dt)["phi"]
    _, vel, _, _ = _scene(n, 1.25, 41)
    x0 = [phi] + vel
    loss, grads, nodes = ad_gradient(oracle_ad, P, g, x0, steps, dt, target)
    print("loss", loss, "tape nodes", nodes)

    def L(x):
        return _loss(_checker_rollout(oracle_lib, P, g, x, steps, dt)["phi"], target, 1.5 / n)[0]

    assert abs(L(x0) - loss) <= 1e-12 * loss
    from scipy.ndimage import gaussian_filter
    dphi = gaussian_filter(np.random.default_rng(5).standard_normal(phi.shape), 2.0, mode="reflect")
    dirs = [[np.zeros_like(phi)] + list(scenes.gaussian_random_velocity(n, 200 + q, sigma_cells=2.0, std=1.0))
            for q in range(2)]
    dirs.append([dphi / dphi.std()] + [np.zeros_like(a) for a in vel])
    for d in dirs:
        fd = {}
        for h in (1e-6, 5e-7):
            xp = [a + h * b for a, b in zip(x0, d)]
            xm = [a - h * b for a, b in zip(x0, d)]
            fd[h] = (L(xp) - L(xm)) / (2 * h)
        rich = (4 * fd[5e-7] - fd[1e-6]) / 3
        ad = sum(float(np.vdot(a, b)) for a, b in zip(grads, d))
        print("central", fd, "richardson", rich, "reverse", ad, "rel", abs(rich - ad) / abs(ad))
        assert abs(rich - ad) <= 1e-6 * abs(ad)
