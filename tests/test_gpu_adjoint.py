"""Adjoints of the GPU timestep pieces against the CPU checker's forward.

For a map F that is linear at a fixed level set, the adjoint F^T is pinned
by the dot-product identity <F x, lam> = <x, F^T lam> for random x and lam,
with F evaluated by the CPU checker (the oracle's forward projection) and
F^T by the HIP library: a wrong sign, a missing face, a transposed operand or
a wrong solve fails it far beyond the solve tolerance."""
import numpy as np
import pytest

from fluid_simulator_backup_amd import _abi
from helpers import make_pair

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return sum(float(np.vdot(x, y)) for x, y in zip(a, b))


@pytest.mark.parametrize("name,schedule", [("rising_drop", _abi.FS_SCHEDULE_COMPAT), ("dam_break", _abi.FS_SCHEDULE_COMPAT),
                                           ("rising_drop", _abi.FS_SCHEDULE_FIXED)])
def test_adjoint_projection_dot_product(hip_lib, oracle_lib, name, schedule):
    n, dt = 16, 2e-3
    over = dict(stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-13, max_iteration=20000, schedule=schedule)
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=n, **over)
    rng = np.random.default_rng(11)
    f = cpu.get_fields()
    x = [rng.standard_normal(f[c].shape) for c in ("u", "v", "w")]
    lam = [rng.standard_normal(f[c].shape) for c in ("u", "v", "w")]
    cpu.set_fields(u=x[0], v=x[1], w=x[2])
    cpu.projection(dt)
    fo = cpu.get_fields()
    px = [fo["u"], fo["v"], fo["w"]]
    ptl, st = gpu.adjoint_projection(dt, *lam)
    # compare the parts beyond the identity: <(P - I) x, lam> both ways
    base = _dot(x, lam)
    c_fwd, c_adj = _dot(px, lam) - base, _dot(x, ptl) - base
    scale = np.sqrt(_dot(px, px) * _dot(lam, lam))
    print(name, schedule, "correction fwd", c_fwd, "adj", c_adj, "iterations", st["pcg_iterations"])
    assert abs(c_fwd) > 1e-5 * scale  # the projection changes these fields
    assert abs(c_fwd - c_adj) < 1e-8 * abs(c_fwd), (c_fwd, c_adj)


@pytest.mark.parametrize("name", ["rising_drop", "dam_break"])
def test_adjoint_sourcing_dot_product(hip_lib, oracle_lib, name):
    """Explicit viscosity (the checker's out-of-place order, which the HIP
    path runs) against its adjoint; viscosities raised so the update is far
    from the identity."""
    n, dt = 16, 2e-3
    over = dict(water_viscosity=0.5, air_viscosity=0.2)
    gpu, cpu = make_pair(hip_lib, oracle_lib, name, n=n, **over)
    rng = np.random.default_rng(12)
    f = cpu.get_fields()
    x = [rng.standard_normal(f[c].shape) for c in ("u", "v", "w")]
    lam = [rng.standard_normal(f[c].shape) for c in ("u", "v", "w")]
    cpu.set_fields(u=x[0], v=x[1], w=x[2])
    cpu.phase("sourcing", dt)
    fo = cpu.get_fields()
    vx = [fo["u"], fo["v"], fo["w"]]
    vtl = gpu.adjoint_sourcing(dt, *lam)
    base = _dot(x, lam)
    c_fwd, c_adj = _dot(vx, lam) - base, _dot(x, vtl) - base
    scale = np.sqrt(_dot(vx, vx) * _dot(lam, lam))
    print(name, "correction fwd", c_fwd, "adj", c_adj)
    assert abs(c_fwd) > 1e-5 * scale
    assert abs(c_fwd - c_adj) < 1e-10 * abs(c_fwd), (c_fwd, c_adj)
    # the forward state is untouched by the adjoint call: a step still
    # matches the checker bit for bit in its exact phases
    gpu.set_fields(u=x[0], v=x[1], w=x[2])
    gpu.phase("sourcing", dt)
    fg = gpu.get_fields()
    for c in ("u", "v", "w"):
        assert np.abs(fg[c] - fo[c]).max() <= 1e-12 * np.abs(fo[c]).max()


def _transport_inputs(n, seed):
    from fluid_simulator_backup_amd import scenes
    x, y, z = scenes.cell_centres(n)
    rng = np.random.default_rng(seed)
    # a sphere distance (kinked at its centre) plus noise: the WENO5 weights
    # move away from their linear values, so their derivatives matter
    phi = np.sqrt((x - .45) ** 2 + (y - .5) ** 2 + (z - .55) ** 2) - .3 + 0.02 * rng.standard_normal(x.shape)
    h = 1.0 / n
    fx = np.arange(n + 1) * h
    c = (np.arange(n) + .5) * h
    # velocities bounded away from zero: the upwind side never switches
    u = (0.5 + 0.2 * np.sin(2 * np.pi * fx))[None, None, :] + 0 * c[:, None, None] + 0 * c[None, :, None]
    v = (-0.4 + 0.15 * np.cos(2 * np.pi * fx))[None, :, None] + 0 * c[:, None, None] + 0 * c[None, None, :]
    w = (0.3 + 0.1 * np.sin(2 * np.pi * fx))[:, None, None] + 0 * c[None, :, None] + 0 * c[None, None, :]
    return phi, u, v, w


def test_weno5_transport_exact_on_linear_level_set(hip_lib, oracle_lib):
    """WENO5 differentiates linear data exactly: away from the walls (where
    the constant extension bends the data) phi' = phi - dt (u a + v b + w c)."""
    n, dt = 16, 1e-2
    gpu, _ = make_pair(hip_lib, oracle_lib, "rising_drop", n=n)
    from fluid_simulator_backup_amd import scenes
    x, y, z = scenes.cell_centres(n)
    phi = 0.3 * x - 0.7 * y + 0.2 * z + 0.1
    u, v, w = np.full((n, n, n + 1), 0.4), np.full((n, n + 1, n), -0.25), np.full((n + 1, n, n), 0.6)
    out = gpu.weno5_transport(dt, phi, u, v, w)
    expect = phi - dt * (0.4 * 0.3 + (-0.25) * (-0.7) + 0.6 * 0.2)
    inner = (slice(3, -3),) * 3
    assert np.abs(out[inner] - expect[inner]).max() < 1e-13


def test_adjoint_weno5_transport_finite_differences(hip_lib, oracle_lib):
    n, dt = 16, 2e-2
    gpu, _ = make_pair(hip_lib, oracle_lib, "rising_drop", n=n)
    phi, u, v, w = _transport_inputs(n, 3)
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(phi.shape)
    lphi, lu, lv, lw = gpu.adjoint_weno5_transport(dt, phi, u, v, w, lam)
    def fd(d, h):
        fp = gpu.weno5_transport(dt, *[a + h * b for a, b in zip((phi, u, v, w), d)])
        fm = gpu.weno5_transport(dt, *[a - h * b for a, b in zip((phi, u, v, w), d)])
        return float(np.vdot(lam, (fp - fm) / (2 * h)))

    for trial in range(4):
        d = [rng.standard_normal(a.shape) for a in (phi, u, v, w)]
        # central differences err by O(h^2) (the weights vary fast where the
        # clamped wall data is nearly flat): Richardson-extrapolate h, h/2
        f1, f2 = fd(d, 1e-6), fd(d, 5e-7)
        rich = (4 * f2 - f1) / 3
        ad = _dot([lphi, lu, lv, lw], d)
        print("fd", f1, f2, "richardson", rich, "adjoint", ad)
        assert abs(rich - ad) < 1e-7 * max(abs(ad), 1.0), (rich, ad)
    # the weight derivatives are not negligible: with them removed (linear
    # weights would make the map affine in phi) the check above would fail
    # by far more than its tolerance on this kinked, noisy level set
    fp = gpu.weno5_transport(dt, phi + 1e-3 * lam, u, v, w)
    fm = gpu.weno5_transport(dt, phi - 1e-3 * lam, u, v, w)
    f0 = gpu.weno5_transport(dt, phi, u, v, w)
    assert np.abs(fp + fm - 2 * f0).max() > 1e-9  # second differences: the map is not affine in phi


def test_autograd_loss_through_transport(hip_lib, oracle_lib):
    """The adjoint wrapped as a torch.autograd.Function (test harness only):
    gradient of sum (H_eps(phi') - H_eps(phi_target))^2 with respect to the
    three velocity fields, against central finite differences along random
    directions."""
    torch = pytest.importorskip("torch")
    n, dt = 12, 2e-2
    gpu, _ = make_pair(hip_lib, oracle_lib, "rising_drop", n=n)
    phi0, u0, v0, w0 = _transport_inputs(n, 5)
    eps = 1.5 / n

    class Transport(torch.autograd.Function):
        @staticmethod
        def forward(ctx, phi, u, v, w):
            ctx.save_for_backward(phi, u, v, w)
            out = gpu.weno5_transport(dt, *[t.detach().numpy() for t in (phi, u, v, w)])
            return torch.from_numpy(out)

        @staticmethod
        def backward(ctx, g):
            phi, u, v, w = ctx.saved_tensors
            res = gpu.adjoint_weno5_transport(dt, *[t.detach().numpy() for t in (phi, u, v, w)],
                                              g.detach().contiguous().numpy())
            return tuple(torch.from_numpy(r) for r in res)

    def heaviside(p):
        inside = 0.5 * (1 + p / eps + torch.sin(np.pi * p / eps) / np.pi)
        return torch.where(p < -eps, torch.zeros_like(p), torch.where(p > eps, torch.ones_like(p), inside))

    target = heaviside(torch.from_numpy(phi0 + 0.05))

    def loss_of(u, v, w):
        out = Transport.apply(torch.from_numpy(phi0), u, v, w)
        return ((heaviside(out) - target) ** 2).sum()

    uu, vv, ww = [torch.from_numpy(a.copy()).requires_grad_(True) for a in (u0, v0, w0)]
    loss = loss_of(uu, vv, ww)
    gu, gv, gw = torch.autograd.grad(loss, (uu, vv, ww))
    rng = np.random.default_rng(6)
    def fd(d, h):
        with torch.no_grad():
            lp = loss_of(*[torch.from_numpy(a + h * b) for a, b in zip((u0, v0, w0), d)]).item()
            lm = loss_of(*[torch.from_numpy(a - h * b) for a, b in zip((u0, v0, w0), d)]).item()
        return (lp - lm) / (2 * h)

    for _ in range(3):
        d = [rng.standard_normal(a.shape) for a in (u0, v0, w0)]
        rich = (4 * fd(d, 5e-7) - fd(d, 1e-6)) / 3
        ad = sum(float((g.numpy() * b).sum()) for g, b in zip((gu, gv, gw), d))
        print("loss fd", rich, "adjoint", ad)
        assert abs(rich - ad) < 1e-6 * max(abs(ad), 1e-3), (rich, ad)
