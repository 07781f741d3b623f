"""Synthetic seeded scenes on the unit cube (inputs only, numpy on the host).

Every scene returns the initial level set phi (cell centres, x fastest,
indexed [k, j, i]) and zero MAC velocities, plus the fs_params overrides
(densities, viscosities, phase labels).  Random numbers come from SplitMix64
so that every consumer reproduces the same scene bit for bit.

phi < 0 marks the phase the reference calls "air" in air_bubble_rising mode
(PROJECTION_3D.h:605-620); scenes whose light phase sits at phi <= 0 use
air_bubble_rising = 1.
"""
import numpy as np

from . import _abi

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _M64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def cell_centres(n):
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return x, y, z


def zero_velocity(n):
    return (np.zeros((n, n, n + 1)), np.zeros((n, n + 1, n)), np.zeros((n + 1, n, n)))


def dam_break(n, seed=0):
    """Water column x<0.4, z<0.6 with a 4-mode seeded surface ripple.

    Water sits at phi <= 0 (water_drop mode, PROJECTION_3D.h:623-639)."""
    rng = SplitMix64(seed)
    x, y, z = cell_centres(n)
    ks = [(1, 0), (0, 1), (1, 1), (2, 1)]
    h = np.full_like(x, 0.6)
    for kx, ky in ks:
        a = 0.01 * rng.uniform()
        th = 2 * np.pi * rng.uniform()
        h = h + a * np.cos(2 * np.pi * (kx * x + ky * y) + th)
    phi = np.maximum(x - 0.4, z - h)
    params = dict(air_bubble_rising=0, water_density=1000.0, air_density=1.0,
                  water_viscosity=1e-3, air_viscosity=1.8e-5)
    return phi, zero_velocity(n), params


def _real_sh(l, m, ct, ph):
    """Real spherical harmonic (unnormalised Legendre form) for the drop."""
    from scipy.special import lpmv
    return lpmv(m, l, ct) * np.cos(m * ph)


def rising_drop(n, seed=1):
    """Oil drop r=0.15 at (0.5,0.5,0.25) in water, 2% radius perturbation
    over 6 spherical-harmonic modes; oil 800 kg/m^3, 0.05 Pa s inside
    (phi <= 0), water 1000, 1e-3 outside."""
    rng = SplitMix64(seed)
    modes = [(2, 0), (2, 2), (3, 1), (3, 3), (4, 0), (4, 2)]
    c = np.array([rng.uniform() * 2 - 1 for _ in modes])
    c = c / np.sqrt((c * c).sum())
    x, y, z = cell_centres(n)
    dx, dy, dz = x - 0.5, y - 0.5, z - 0.25
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    ct = np.where(r > 0, dz / np.maximum(r, 1e-300), 1.0)
    ph = np.arctan2(dy, dx)
    pert = np.zeros_like(r)
    for (l, m), cc in zip(modes, c):
        pert += cc * _real_sh(l, m, ct, ph)
    phi = r - 0.15 * (1 + 0.02 * pert)
    params = dict(air_bubble_rising=1, water_density=1000.0, air_density=800.0,
                  water_viscosity=1e-3, air_viscosity=0.05)
    return phi, zero_velocity(n), params


def bubbles(n, seed=2, count=64):
    """Water pool z<0.7 with `count` non-overlapping seeded air bubbles,
    radii U(0.02, 0.06); air (phi <= 0) 1 kg/m^3."""
    rng = SplitMix64(seed)
    cs, rs = [], []
    tries = 0
    while len(cs) < count and tries < 100000:
        tries += 1
        r = 0.02 + 0.04 * rng.uniform()
        c = np.array([r + (1 - 2 * r) * rng.uniform(), r + (1 - 2 * r) * rng.uniform(),
                      r + (0.7 - 2 * r) * rng.uniform()])
        if all(np.linalg.norm(c - c2) > r + r2 for c2, r2 in zip(cs, rs)):
            cs.append(c)
            rs.append(r)
    x, y, z = cell_centres(n)
    phi = 0.7 - z  # air above the pool has phi <= 0
    for c, r in zip(cs, rs):
        phi = np.minimum(phi, np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r)
    params = dict(air_bubble_rising=1, water_density=1000.0, air_density=1.0,
                  water_viscosity=1e-3, air_viscosity=1.8e-5)
    return phi, zero_velocity(n), params


def slosh(n, seed=4):
    """Sloshing tank: water below the tilted free surface z = 0.5 + 0.1 (x -
    0.5) (phi <= 0, water_drop mode as the dam break), air above; densities
    and viscosities as the dam break."""
    x, y, z = cell_centres(n)
    phi = z - (0.5 + 0.1 * (x - 0.5))
    params = dict(air_bubble_rising=0, water_density=1000.0, air_density=1.0,
                  water_viscosity=1e-3, air_viscosity=1.8e-5)
    return phi, zero_velocity(n), params


def gaussian_random_velocity(n, seed, sigma_cells=8.0, std=0.05):
    """MAC velocity whose components are white noise (numpy's PCG64 stream
    of `seed`) smoothed by a separable Gaussian of sigma_cells cells and
    rescaled to standard deviation `std`; the normal wall faces are zero."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    out = []
    for shape, axis in (((n, n, n + 1), 2), ((n, n + 1, n), 1), ((n + 1, n, n), 0)):
        f = gaussian_filter(rng.standard_normal(shape), sigma=sigma_cells, mode="reflect")
        f *= std / f.std()
        idx = [slice(None)] * 3
        for wall in (0, -1):
            idx[axis] = wall
            f[tuple(idx)] = 0
        out.append(f)
    return tuple(out)


SCENES = {"dam_break": dam_break, "rising_drop": rising_drop, "bubbles": bubbles, "slosh": slosh}


def make(name, n, seed=None, **param_overrides):
    """(phi, (u, v, w), FsParams, FsGrid) for a named scene at n^3."""
    fn = SCENES[name]
    phi, vel, pkw = fn(n) if seed is None else fn(n, seed)
    pkw.update(param_overrides)
    return phi, vel, _abi.default_params(**pkw), _abi.unit_cube_grid(n)


def batch_scenes(count=48, seed=3, edges=(32, 48, 64, 96, 128), **param_overrides):
    """Uneven scenes for batched serving: per scene an edge from `edges`, a
    dam break or a rising drop by coin flip, a density ratio log-uniform in
    [1.25, 1000] (heavy phase 1000 kg/m^3), scene seed 1000*seed + s.  Returns
    a list of (name, phi, (u, v, w), FsParams, FsGrid)."""
    rng = SplitMix64(seed)
    out = []
    for s in range(count):
        n = edges[min(int(rng.uniform() * len(edges)), len(edges) - 1)]
        name = "dam_break" if rng.uniform() < 0.5 else "rising_drop"
        ratio = float(np.exp(np.log(1.25) + (np.log(1000.0) - np.log(1.25)) * rng.uniform()))
        phi, vel, pkw = SCENES[name](n, 1000 * seed + s)
        pkw["water_density"] = 1000.0
        pkw["air_density"] = 1000.0 / ratio
        pkw.update(param_overrides)
        out.append((name, phi, vel, _abi.default_params(**pkw), _abi.unit_cube_grid(n)))
    return out
