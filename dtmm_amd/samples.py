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
    """Sum of nmodes random Fourier modes a cos(2 pi k.r / N + psi), as one
    complex matrix product of the z factors with the (y, x) factors."""
    nz, ny, nx = shape
    z = np.arange(nz) / nz
    y = np.arange(ny) / ny
    x = np.arange(nx) / nx
    ez = np.empty((nmodes, nz), np.complex128)
    exy = np.empty((nmodes, ny * nx), np.complex128)
    for m in range(nmodes):
        while True:
            k = rng.integers(-kmax, kmax + 1, size=3)
            if 0 < np.sqrt((k ** 2).sum()) <= kmax:
                break
        a = rng.normal(0., sigma)
        psi = rng.uniform(0., 2 * np.pi)
        ez[m] = a * np.exp(1j * psi) * np.exp(2j * np.pi * k[0] * z)
        exy[m] = (np.exp(2j * np.pi * k[1] * y)[:, None] *
                  np.exp(2j * np.pi * k[2] * x)[None, :]).ravel()
    return (ez.T @ exy).real.reshape(shape)


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
    box centre, droplets centred in z) in an isotropic host; voxel
    coordinates are half-pixel centred as in dtmm/data.py:773-777."""
    nz, ny, nx = shape
    az, ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    director = np.zeros(shape + (3,), np.float64)
    mask = np.zeros(shape, bool)
    for cx, cy in centres:
        # bounding box of the droplet (whole grid lines within radius)
        sl = []
        for a, c in ((az, 0.), (ay, cy), (ax, cx)):
            idx = np.nonzero(np.abs(a - c) <= radius)[0]
            sl.append(slice(idx[0], idx[-1] + 1) if len(idx) else slice(0, 0))
        zz, yy, xx = np.meshgrid(az[sl[0]], ay[sl[1]] - cy, ax[sl[2]] - cx, indexing="ij")
        r = np.sqrt(xx ** 2 + yy ** 2 + zz ** 2)
        m = (r <= radius) & (r != 0)
        d = director[sl[0], sl[1], sl[2]]
        d[..., 0][m] = xx[m] / r[m]
        d[..., 1][m] = yy[m] / r[m]
        d[..., 2][m] = zz[m] / r[m]
        mask[sl[0], sl[1], sl[2]] |= r <= radius
    return director2data(director, mask=mask, no=no, ne=ne, nhost=nhost)


def droplet_grid_centres(size, n=3):
    """(x, y) offsets of an n x n grid of droplets spaced size / n apart,
    centred on the box."""
    step = size / n
    offs = [(i - (n - 1) / 2.) * step for i in range(n)]
    return [(x, y) for y in offs for x in offs]


__all__ = ["random_director", "random_director_data", "droplet_array_data",
           "droplet_grid_centres", "refind2eps"]
