"""Optical-data model: blocks of (thickness, eps eigenvalues, Euler angles).

Host-side NumPy input builders with the reference's names and argument
meanings (dtmm/data.py).  They build inputs; the propagation itself runs
only on the GPU (``dtmm_amd.transfer``).
"""
import numpy as np

FDTYPE = np.float64
CDTYPE = np.complex128


def refind2eps(refind):
    """Refractive index -> epsilon, eps = n**2 (dtmm/data.py:1059-1062)."""
    return np.asarray(refind) ** 2


def uniaxial_order(order, eps):
    """Uniaxial (order > 0), isotropic (order == 0) or unchanged (order < 0)
    eigenvalues from a diagonal tensor (dtmm/data.py:1215-1274)."""
    eps = np.asarray(eps)
    order = np.asarray(order, FDTYPE)
    m = (eps[..., 0] + eps[..., 1] + eps[..., 2]) / 3.
    delta = eps[..., 2] - (eps[..., 0] + eps[..., 1]) / 2.
    o = np.broadcast_to(order, m.shape)
    e1 = np.where(o == 0., m, m - 1. / 3. * o * delta)
    e3 = np.where(o == 0., m, m + 2. / 3. * o * delta)
    e1 = np.where(o >= 0, e1, eps[..., 0])
    e2 = np.where(o >= 0, e1, eps[..., 1])
    e3 = np.where(o >= 0, e3, eps[..., 2])
    return np.stack([e1, e2, e3], -1)


def director2order(director):
    """Order parameter = sqrt(|director|) (dtmm/data.py:962-972)."""
    d = np.asarray(director, FDTYPE)
    return np.sqrt(np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2 + d[..., 2] ** 2))


def director2angles(director):
    """Director -> (psi=0, theta, phi) (dtmm/data.py:975-989)."""
    d = np.asarray(director, FDTYPE)
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    out = np.empty(d.shape, FDTYPE)
    out[..., 0] = 0.
    out[..., 1] = np.arctan2(np.sqrt(x ** 2 + y ** 2), z)
    out[..., 2] = np.arctan2(y, x)
    return out


def director2data(director, mask=None, no=1.5, ne=1.6, nhost=None, scale_factor=1.,
                  thickness=None):
    """Optical block from a director field (dtmm/data.py:277-320)."""
    director = np.asarray(director, FDTYPE)
    material = np.empty(director.shape, FDTYPE)
    material[...] = refind2eps([no, no, ne])[None, ...]
    material = uniaxial_order(director2order(director) / scale_factor, material)
    if mask is not None:
        material[np.logical_not(mask), :] = refind2eps([nhost, nhost, nhost])[None, ...]
    if thickness is None:
        thickness = np.ones(shape=(material.shape[0],))
    return thickness, material, director2angles(director)


def _r3(shape):
    """Half-pixel-centred coordinates (dtmm/data.py:773-777)."""
    az, ay, ax = [np.arange(-l / 2. + .5, l / 2. + .5) for l in shape]
    zz, yy, xx = np.meshgrid(az, ay, ax, indexing="ij")
    return xx, yy, zz


def nematic_droplet_director(shape, radius, profile="r", retmask=False):
    """Director of a nematic droplet (dtmm/data.py:803-856)."""
    nz, ny, nx = shape
    out = np.zeros(shape=(nz, ny, nx, 3), dtype=FDTYPE)
    xx, yy, zz = _r3(shape)
    r = (xx ** 2 + yy ** 2 + zz ** 2) ** 0.5
    mask = (r <= radius)
    m = np.logical_and(mask, r != 0.)
    rm = r[m]
    if profile == "r":
        out[..., 0][m] = xx[m] / rm
        out[..., 1][m] = yy[m] / rm
        out[..., 2][m] = zz[m] / rm
    elif isinstance(profile, str):
        index = {"x": 0, "y": 1, "z": 2}
        if profile not in index:
            raise ValueError("Unsupported profile type!")
        out[..., index[profile]][m] = 1.
    else:
        x, y, z = profile
        out[..., 0][m] = x
        out[..., 1][m] = y
        out[..., 2][m] = z
    if retmask:
        return mask, out
    return out


def nematic_droplet_data(shape, radius, profile="r", no=1.5, ne=1.6, nhost=1.5):
    """Optical block of a nematic droplet in a host (dtmm/data.py:897-927)."""
    mask, director = nematic_droplet_director(shape, radius, profile=profile, retmask=True)
    return director2data(director, mask=mask, no=no, ne=ne, nhost=nhost)


def validate_optical_block(data, shape=None):
    """Cast a block and add singleton axes so it broadcasts to (L, H, W, 3)
    (dtmm/data.py:534-618).  Callable (dispersive) materials are not
    supported by the engine and raise."""
    thickness, material, angles = data
    if callable(material):
        raise NotImplementedError("dispersive (callable) materials are not supported")
    thickness = np.asarray(thickness, dtype=FDTYPE)
    material = np.asarray(material)
    if np.issubdtype(material.dtype, np.complexfloating):
        material = np.asarray(material, dtype=CDTYPE)
    else:
        material = np.asarray(material, dtype=FDTYPE)
    angles = np.asarray(angles, dtype=FDTYPE)
    if thickness.ndim == 0:
        thickness = thickness[None]
        material = material[None, ...]
        angles = angles[None, ...]
    elif thickness.ndim != 1:
        raise ValueError("Thickess dimension should be 1.")
    n = len(thickness)
    if material.ndim == 2:
        material = material[:, None, None, :]
    elif material.ndim == 3:
        material = material[:, None, :, :]
    elif material.ndim != 4:
        raise ValueError("Invalid material dimensions.")
    bshape = (n, 1, 1, 3) if shape is None else (n, shape[0], shape[1], 3)
    np.broadcast_shapes(material.shape, bshape)
    np.broadcast_shapes(angles.shape, bshape)
    return thickness, material, angles


def validate_optical_data(data, shape=None):
    """List of validated blocks (dtmm/data.py:620-643)."""
    if isinstance(data, list):
        if shape is None:
            raise ValueError("For heterogeneous data, you must provide the `shape` argument.")
        return [validate_optical_block(d, shape) for d in data]
    return validate_optical_data([data], shape if shape is not None else (1, 1))


def isotropic_effective_eps(block):
    """Volume mean of the isotropic part of eps of a block, the effective
    medium of the diffraction step for eff_data=0 (dtmm/transfer.py:34-41)."""
    d, material, angles = block
    return uniaxial_order(0., material).mean(axis=(0, 1, 2))
