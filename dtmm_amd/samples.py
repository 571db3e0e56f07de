"""Synthetic samples used by the parity tests and the benchmark.

``random_director_data`` is the smooth random nematic texture of the
512x512x128 benchmark workload: tilt and twist are sums of 16 random
Fourier modes (integer wave vectors with |k| <= 8 cycles per box, amplitudes
N(0, 0.3) rad, uniform phases) on top of a planar base tilt and a 90 degree
linear twist across the thickness.  ``droplet_array_data`` places radial
nematic droplets on a regular grid in an isotropic host.
"""
import numpy as np

from .data import director2data, refind2eps


def _smooth_field(rng, shape, nmodes=16, kmax=8, sigma=0.3):
    nz, ny, nx = shape
    out = np.zeros(shape, np.float64)
    z = np.arange(nz) / nz
    y = np.arange(ny) / ny
    x = np.arange(nx) / nx
    for _ in range(nmodes):
        while True:
            k = rng.integers(-kmax, kmax + 1, size=3)
            if 0 < np.sqrt((k ** 2).sum()) <= kmax:
                break
        a = rng.normal(0., sigma)
        psi = rng.uniform(0., 2 * np.pi)
        ez = np.exp(2j * np.pi * k[0] * z)
        ey = np.exp(2j * np.pi * k[1] * y)
        ex = np.exp(2j * np.pi * k[2] * x)
        out += a * (np.exp(1j * psi) * ez[:, None, None] * ey[None, :, None] * ex[None, None, :]).real
    return out


def random_director(shape, seed=0, theta0=np.pi / 2, twist=np.pi / 2):
    """Unit director field (nz, ny, nx, 3) of the smooth random texture."""
    rng = np.random.default_rng(seed)
    theta = theta0 + _smooth_field(rng, shape)
    phi = _smooth_field(rng, shape)
    nz = shape[0]
    phi = phi + twist * (np.arange(nz) / max(nz - 1, 1))[:, None, None]
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], -1)


def random_director_data(shape, seed=0, no=1.5, ne=1.7, theta0=np.pi / 2, twist=np.pi / 2):
    """Optical block (thickness, eps, angles) of the smooth random texture."""
    return director2data(random_director(shape, seed, theta0, twist), no=no, ne=ne)


def droplet_array_data(shape, radius, centres, no=1.5, ne=1.6, nhost=1.5):
    """Radial nematic droplets at `centres` ((x, y) offsets in pixels from the
    box centre, droplets centred in z) in an isotropic host."""
    nz, ny, nx = shape
    az, ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    zz, yy, xx = np.meshgrid(az, ay, ax, indexing="ij")
    director = np.zeros(shape + (3,), np.float64)
    mask = np.zeros(shape, bool)
    for cx, cy in centres:
        dx, dy = xx - cx, yy - cy
        r = np.sqrt(dx ** 2 + dy ** 2 + zz ** 2)
        m = (r <= radius) & (r != 0)
        director[..., 0][m] = dx[m] / r[m]
        director[..., 1][m] = dy[m] / r[m]
        director[..., 2][m] = zz[m] / r[m]
        mask |= r <= radius
    return director2data(director, mask=mask, no=no, ne=ne, nhost=nhost)


__all__ = ["random_director", "random_director_data", "droplet_array_data", "refind2eps"]
