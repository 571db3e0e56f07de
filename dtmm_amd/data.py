"""Optical-data model: blocks of (thickness, eps eigenvalues, Euler angles).

Host-side NumPy input builders with the reference's names, arguments and
results (dtmm/data.py): director and droplet samples, the order-parameter
scaling of eps, block validation and the dispersive (wavelength-dependent)
materials.  They build inputs; the propagation itself runs only on the GPU
(``dtmm_amd.transfer``).
"""
import numpy as np

FDTYPE = np.float64
CDTYPE = np.complex128

# unit director of the fixed droplet profiles (dtmm/data.py:803-856)
_PROFILE_AXES = {"x": (1., 0., 0.), "y": (0., 1., 0.), "z": (0., 0., 1.)}


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
    """Optical block (thickness, eps, angles) of a director field
    (dtmm/data.py:277-320).

    The length of the director sets the nematic order S = sqrt(|n|) /
    scale_factor, which scales the anisotropy of the (no^2, no^2, ne^2)
    tensor about its mean (``uniaxial_order``); voxels where `mask` is False
    hold the isotropic host nhost^2.  Layer thicknesses default to 1.
    """
    n = np.asarray(director, FDTYPE)
    pure = np.broadcast_to(refind2eps(np.asarray([no, no, ne], FDTYPE)), n.shape)
    eps = uniaxial_order(director2order(n) / scale_factor, pure)
    if mask is not None:
        host = refind2eps(np.full(3, nhost, FDTYPE))
        eps = np.where(np.asarray(mask, bool)[..., None], eps, host)
    d = np.ones(n.shape[0]) if thickness is None else thickness
    return d, eps, director2angles(n)


def _voxel_centres(shape):
    """Voxel-centre coordinates (z, y, x) of a box, each relative to the box
    centre and shaped to broadcast against the box; an even edge has no
    voxel at 0 (dtmm/data.py:773-777)."""
    return np.ix_(*[np.arange(n, dtype=FDTYPE) - (n - 1) / 2. for n in shape])


def nematic_droplet_director(shape, radius, profile="r", retmask=False):
    """Director of a nematic droplet of `radius` centred in the box
    (dtmm/data.py:803-856).  profile "r" points along the radius vector; "x",
    "y", "z" or a 3-vector is a uniform director.  Voxels outside the droplet
    (and its centre, for "r") hold a zero director.  retmask: also the
    droplet mask r <= radius."""
    z, y, x = _voxel_centres(shape)
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    inside = r <= radius
    filled = inside & (r != 0.)
    if isinstance(profile, str):
        if profile == "r":
            with np.errstate(invalid="ignore", divide="ignore"):
                vec = np.stack(np.broadcast_arrays(x / r, y / r, z / r), -1)
        elif profile in _PROFILE_AXES:
            vec = np.asarray(_PROFILE_AXES[profile], FDTYPE)
        else:
            raise ValueError("Unsupported profile type!")
    else:
        vec = np.asarray(profile, FDTYPE).reshape(3)
    director = np.where(filled[..., None], vec, FDTYPE(0.))
    return (inside, director) if retmask else director


def sphere_mask(shape, radius, offset=(0, 0, 0)):
    """True inside a sphere of `radius` about the box centre shifted by
    offset (x, y, z) (dtmm/data.py:780-805).  The reference adds the z offset
    where it subtracts the x and y ones (``zz--offset[2]``); kept for
    identical masks."""
    z, y, x = _voxel_centres(shape)
    dx, dy, dz = offset
    return np.sqrt((x - dx) ** 2 + (y - dy) ** 2 + (z + dz) ** 2) <= radius


def cholesteric_director(shape, pitch, hand="left"):
    """In-plane director rotating by 2 pi per `pitch` layers, layer 0 along
    x; "left" turns clockwise seen from +z (dtmm/data.py:864-895)."""
    sign = {"left": -1., "right": 1.}.get(hand)
    if sign is None:
        raise ValueError("Unknown handedness '{}'".format(hand))
    nz, ny, nx = shape
    phi = sign * 2 * np.pi / pitch * np.arange(nz)
    out = np.zeros((nz, ny, nx, 3), FDTYPE)
    out[..., 0] = np.cos(phi)[:, None, None]
    out[..., 1] = np.sin(phi)[:, None, None]
    return out


def cholesteric_droplet_data(shape, radius, pitch, hand="left", no=1.5, ne=1.6, nhost=1.5):
    """Optical block of a cholesteric droplet in a host (dtmm/data.py
    :929-960)."""
    return director2data(cholesteric_director(shape, pitch, hand=hand),
                         mask=sphere_mask(shape, radius), no=no, ne=ne, nhost=nhost)


def nematic_droplet_data(shape, radius, profile="r", no=1.5, ne=1.6, nhost=1.5):
    """Optical block of a nematic droplet in a host (dtmm/data.py:897-927)."""
    mask, director = nematic_droplet_director(shape, radius, profile=profile, retmask=True)
    return director2data(director, mask=mask, no=no, ne=ne, nhost=nhost)


# ---------------------------------------------------------------------------
# dispersive materials (dtmm/data.py:1094-1204)
# ---------------------------------------------------------------------------

class EpsilonDispersive:
    """A wavelength-dependent eps: per-voxel coefficients of shape
    (..., 3, n), one row of n coefficients per eigenvalue.  Calling the
    object with a wavelength in nanometres returns the eps eigenvalues
    (..., 3) there.  Built from `values`, or as `shape` + (3, n) coefficients
    of a non-dispersive eps = 1 (first coefficient 1, the others 0)."""

    def __init__(self, values=None, shape=None, n=None, broadcast=False, copy=False):
        n = 1 if n is None else int(n)
        if n < 0:
            raise ValueError("n must be >= 0")
        full = (tuple(shape) if shape is not None else ()) + (3, n)
        if values is None:
            if shape is None:
                raise ValueError("Coefficients were not set, so you need to provide n and "
                                 "shape parameters.")
            c = np.zeros(full, FDTYPE)
            c[..., 0] = 1.
        else:
            c = np.asarray(values, FDTYPE)
            # at least one row per eigenvalue
            c = np.broadcast_to(c, np.broadcast_shapes(c.shape, (3, c.shape[-1] if c.ndim else n)))
            if broadcast:
                c = np.broadcast_to(c, np.broadcast_shapes(c.shape, full))
            if copy:
                c = c.copy()
        self.coefficients = c

    def __getitem__(self, index):
        return type(self)(self.coefficients[index])

    def __call__(self, wavelength):
        raise NotImplementedError("a dispersive material defines __call__(wavelength)")


class EpsilonCauchy(EpsilonDispersive):
    """Cauchy dispersion n = c0 + c1 / w^2 + c2 / w^4 + ..., eps = n^2, with
    w the wavelength in microns (dtmm/data.py:1110-1129, :1183-1193)."""

    def __call__(self, wavelength):
        c = self.coefficients
        inv_w2 = 1. / (np.asarray(wavelength, FDTYPE) / 1000.) ** 2
        n = c[..., 0]
        for i in range(1, c.shape[-1]):
            n = n + c[..., i] * inv_w2 ** i
        return n ** 2


class EpsilonSellmeier(EpsilonDispersive):
    """Sellmeier dispersion eps = c0 + w^2 c1 / (w^2 - c2) + w^2 c3 / (w^2 - c4)
    + ..., w in microns, an odd number of coefficients
    (dtmm/data.py:1072-1093, :1195-1204)."""

    def __call__(self, wavelength):
        c = self.coefficients
        w2 = (np.asarray(wavelength, FDTYPE) / 1000.) ** 2
        eps = c[..., 0]
        for i in range((c.shape[-1] - 1) // 2):
            eps = eps + w2 * c[..., 1 + 2 * i] / (w2 - c[..., 2 + 2 * i])
        return eps


def is_optical_block_dispersive(block):
    """Whether the block's eps is a callable of the wavelength
    (dtmm/data.py:1131-1134)."""
    return callable(block[1])


def is_optical_data_dispersive(optical_data):
    """Whether any block of the optical data is dispersive
    (dtmm/data.py:1136-1142)."""
    return any(is_optical_block_dispersive(b) for b in optical_data)


# ---------------------------------------------------------------------------
# validation (dtmm/data.py:534-682)
# ---------------------------------------------------------------------------

def _with_lateral_axes(a, rank, what):
    """Insert the missing (height, width) singleton axes after the layer axis
    so that `a` has `rank` axes: per-layer values (n, 3) become (n, 1, 1, 3)
    and row data (n, W, 3) become (n, 1, W, 3)."""
    missing = rank - a.ndim
    if not 0 <= missing <= 2:
        raise ValueError("Invalid %s dimensions." % what)
    return a.reshape(a.shape[:1] + (1,) * missing + a.shape[1:])


def validate_optical_block(data, shape=None, wavelength=None):
    """Checks one block (thickness, eps, angles) and returns it cast to the
    engine's types with the layer axis first (dtmm/data.py:534-618).

    A scalar thickness is a one-layer block.  eps with fewer than four axes
    gains singleton lateral axes; eps and angles must broadcast to
    (layers, height, width, 3) of `shape`.  A dispersive eps is evaluated at
    `wavelength` (nm) when one is given; without one, an
    ``EpsilonDispersive`` keeps its coefficients (with a trailing coefficient
    axis) and any other callable is an error."""
    if isinstance(data, list):
        if shape is None:
            raise ValueError("For heterogeneous data, you must provide the `shape` argument.")
        return [validate_optical_block(b, shape, wavelength) for b in data]
    thickness, material, angles = data
    kind = None
    if callable(material):
        if wavelength is not None:
            material = material(wavelength)
        elif isinstance(material, EpsilonDispersive):
            kind = type(material)
            material = material.coefficients
        else:
            raise ValueError("Epsilon is a callable. You must provide wavelength! Use "
                             "split_wavelength = True.")
    thickness = np.asarray(thickness, FDTYPE)
    material = np.asarray(material)
    material = material.astype(CDTYPE if np.iscomplexobj(material) else FDTYPE, copy=False)
    angles = np.asarray(angles, FDTYPE)
    if thickness.ndim > 1:
        raise ValueError("thickness must be a scalar or a 1D array of layer thicknesses.")
    if thickness.ndim == 0:
        thickness, material, angles = thickness[None], material[None], angles[None]
    n = len(thickness)
    lateral = (1, 1) if shape is None else tuple(shape)
    rank = 4 if kind is None else 5
    material = _with_lateral_axes(material, rank, "material" if kind is None else
                                  "material coefficients")
    np.broadcast_shapes(material.shape, (n,) + lateral + (3,) + (() if kind is None else (1,)))
    np.broadcast_shapes(angles.shape, (n,) + lateral + (3,))
    if kind is not None:
        material = kind(material)
    return thickness, material, angles


def validate_optical_data(data, shape=None, wavelength=None):
    """List of validated blocks (dtmm/data.py:620-643); a single block is
    taken as a one-block list."""
    if not isinstance(data, list):
        return validate_optical_data([data], shape if shape is not None else (1, 1), wavelength)
    if shape is None:
        raise ValueError("For heterogeneous data, you must provide the `shape` argument.")
    return [validate_optical_block(b, shape, wavelength) for b in data]


def evaluate_optical_block(block, wavelength=550):
    """The block with a dispersive eps evaluated at `wavelength` (nm)
    (dtmm/data.py:682-687)."""
    d, epsv, epsa = block
    return d, (epsv(wavelength) if callable(epsv) else epsv), epsa


def isotropic_effective_eps(block):
    """Volume mean of the isotropic part of eps of a block, the effective
    medium of the diffraction step for eff_data=0 (dtmm/transfer.py:34-41)."""
    d, material, angles = block
    return uniaxial_order(0., material).mean(axis=(0, 1, 2))


# ---------------------------------------------------------------------------
# Effective media of the diffraction step (eff_data of transfer_field)
# ---------------------------------------------------------------------------

_SYMMETRY_NAMES = {"isotropic": 0, "uniaxial": 1, "biaxial": 2}


def _symmetry_id(value):
    if isinstance(value, str):
        if value not in _SYMMETRY_NAMES:
            raise ValueError("Invalid symmetry argument.")
        return _SYMMETRY_NAMES[value]
    if value in (0, 1, 2, -1):
        return int(value)
    raise ValueError("Invalid symmetry argument.")


def parse_symmetry(arg):
    """An int symmetry (0 isotropic, 1 uniaxial, 2 biaxial, -1 unchanged) or
    a tuple of per-layer symmetries (dtmm/data.py:1341-1362)."""
    if isinstance(arg, str):
        return _symmetry_id(arg)
    try:
        items = list(arg)
    except TypeError:
        return _symmetry_id(arg)
    return tuple(parse_symmetry(a) for a in items)


def eig_symmetry(order, eig):
    """Eigenvalues made isotropic (order 0), uniaxial (1: the first two
    averaged, the third kept) or left as they are (2 or -1)
    (dtmm/data.py:1276-1306)."""
    o = np.array(order, int)
    return uniaxial_order(np.where(o > 1, -1, o), eig)


def _sym_matrix(t6):
    t = np.asarray(t6)
    xx, yy, zz, xy, xz, yz = (t[..., i] for i in range(6))
    return np.stack([np.stack([xx, xy, xz], -1), np.stack([xy, yy, yz], -1),
                     np.stack([xz, yz, zz], -1)], -2)


def tensor_eig(t6):
    """Eigenvalues w (..., 3) and eigenframe v (..., 3, 3), eigenvectors as
    columns, of real symmetric tensors (xx, yy, zz, xy, xz, yz), ordered as
    the reference orders them (dtmm/linalg.py:364-424): ascending with a
    right-handed frame, or -- when the off-diagonal part, relative to the
    largest element, has a squared norm of at most 1e-12 -- the diagonal in
    its own order with the identity frame; then the eigenvalue farthest from
    the other two is moved to index 2 by a cyclic shift.  The uniaxial
    effective medium keeps that eigenvalue as its extraordinary one."""
    t = np.asarray(t6, FDTYPE)
    w, v = np.linalg.eigh(_sym_matrix(t))
    v[..., :, 2] = np.cross(v[..., :, 0], v[..., :, 1])
    big = np.abs(t).max(axis=-1)
    inv = np.divide(1., big, out=np.zeros_like(big), where=big > 0)
    off = ((t[..., 3:] * inv[..., None]) ** 2).sum(-1)
    flat = off <= 1e-12
    w = np.where(flat[..., None], t[..., :3], w)
    v = np.where(flat[..., None, None], np.eye(3), v)
    # distance of each pair: g01 = |w1 - w0|, g02 = |w2 - w0|, g12 = |w2 - w1|
    g01 = np.abs(w[..., 1] - w[..., 0])
    g02 = np.abs(w[..., 2] - w[..., 0])
    g12 = np.abs(w[..., 2] - w[..., 1])
    shift = (g01 > g02) | (g01 > g12)
    # pair (1, 2) closest: w0 stands out and goes last; else pair (0, 2): w1
    last0 = shift & (g02 > g12)
    last1 = shift & ~(g02 > g12)
    order = np.broadcast_to(np.arange(3), w.shape).copy()
    order[last0] = (1, 2, 0)
    order[last1] = (2, 0, 1)
    w = np.take_along_axis(w, order, -1)
    v = np.take_along_axis(v, order[..., None, :], -1)
    return w, v


def rotation_angles(r):
    """Euler angles (psi, theta, phi) of a rotation r = Rz(phi) Ry(theta)
    Rz(psi) with the reference's rule for theta = 0, where only phi is kept,
    as arccos of r[1, 1] (dtmm/rotation.py:233-260)."""
    r = np.asarray(r, FDTYPE)
    theta = np.arccos(np.clip(r[..., 2, 2], -1., 1.))
    flat = ((r[..., 0, 2] == 0.) & (r[..., 1, 2] == 0.)) | (theta == 0.)
    phi = np.where(flat, np.arccos(np.clip(r[..., 1, 1], -1., 1.)),
                   np.arctan2(r[..., 1, 2], r[..., 0, 2]))
    psi = np.where(flat, 0., np.arctan2(r[..., 2, 1], -r[..., 2, 0]))
    return np.stack([psi, theta, phi], -1)


def rotation_matrix(angles):
    """Rz(phi) Ry(theta) Rz(psi) of angles (..., 3) = (psi, theta, phi)
    (dtmm/rotation.py:199-227)."""
    a = np.asarray(angles, FDTYPE)
    cps, sps = np.cos(a[..., 0]), np.sin(a[..., 0])
    cth, sth = np.cos(a[..., 1]), np.sin(a[..., 1])
    cph, sph = np.cos(a[..., 2]), np.sin(a[..., 2])
    return np.stack([
        np.stack([cph * cth * cps - sph * sps, -cph * cth * sps - sph * cps, cph * sth], -1),
        np.stack([sph * cth * cps + cph * sps, cph * cps - sph * cth * sps, sph * sth], -1),
        np.stack([-sth * cps, sth * sps, cth], -1)], -2)


def epsva2eps(epsv, epsa):
    """Laboratory-frame tensors (xx, yy, zz, xy, xz, yz) of eigenvalues and
    Euler angles (dtmm/data.py:515-532)."""
    r = rotation_matrix(epsa)
    e = np.asarray(epsv)
    comp = lambda i, j: (r[..., i, :] * r[..., j, :] * e).sum(-1)
    return np.stack([comp(0, 0), comp(1, 1), comp(2, 2), comp(0, 1), comp(0, 2), comp(1, 2)], -1)


def eps2epsva(eps):
    """Eigenvalues and Euler angles of tensors (..., 6) (dtmm/data.py:489-513).
    Only real tensors are taken: for complex ones the reference warns that
    its eigenframe is unreliable."""
    t = np.asarray(eps)
    if np.iscomplexobj(t):
        if np.any(t.imag != 0):
            raise NotImplementedError("anisotropic effective media of absorbing (complex) "
                                      "tensors are not supported")
        t = t.real
    w, v = tensor_eig(t)
    return w, rotation_angles(v)


def effective_from_tensors(tensors, symmetry):
    """Effective layers (epsv (n, 3), epsa (n, 3)) of one block from the mean
    laboratory-frame tensor of each of its n layers (equal voxel counts), as
    effective_block forms them (dtmm/data.py:1364-1419): one symmetry averages
    the whole block, a sequence of per-layer symmetries each layer."""
    t = np.asarray(tensors)
    n = len(t)
    order = np.asarray(parse_symmetry(symmetry), int)
    if order.ndim == 0 and order == 0:
        ev = np.broadcast_to(((t[:, 0] + t[:, 1] + t[:, 2]) / 3.).mean(), (3,))
        return np.tile(ev, (n, 1)), np.zeros((n, 3))
    if order.ndim == 0:
        w, a = eps2epsva(t.mean(0))
        w = eig_symmetry(order, w)
        return np.tile(w, (n, 1)), np.tile(a, (n, 1))
    if len(order) != n:
        raise ValueError("Shape of the symmetry argument is incompatible with block size.")
    w, a = eps2epsva(t)
    return eig_symmetry(order, w), a


def effective_block(optical_block, symmetry=0):
    """Effective block (d, epsv, epsa) of an optical block with the material
    averaged over the block, or over each layer for per-layer symmetries
    (dtmm/data.py:1364-1419)."""
    d, epsv, epsa = validate_optical_block(optical_block)
    L = len(d)
    shape = (L,) + tuple(np.broadcast_shapes(np.shape(epsv)[1:], np.shape(epsa)[1:]))
    ev = np.broadcast_to(epsv, shape)
    t = epsva2eps(ev, np.broadcast_to(epsa, shape)).reshape(L, -1, 6).mean(1)
    order = np.asarray(parse_symmetry(symmetry), int)
    if order.ndim == 0 and order == 0:
        e = uniaxial_order(0., ev).reshape(-1, 3).mean(0)
        return d, np.tile(e, (L, 1)), np.zeros((L, 3))
    w, a = effective_from_tensors(t, symmetry)
    return d, w, a


def effective_data(optical_data, symmetry=0):
    """effective_block of every block (dtmm/data.py:1421-1443)."""
    blocks = optical_data if isinstance(optical_data, list) else [optical_data]
    out = [effective_block(b, symmetry) for b in blocks]
    return out if isinstance(optical_data, list) else out[0]

