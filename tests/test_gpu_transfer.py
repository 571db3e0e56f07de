"""GPU parity: dtmm_amd.transfer_field (HIP, gfx950) against the complex128
oracle restatement of the reference, on the same inputs.

Tolerances (max|F - F_ref| / max|F_ref| over the output field):
  c128 GPU path: 1e-9 (round-off only);
  c64 GPU path:  1e-4 (the target stated for complex64 propagation).
"""
import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from helpers import droplet_case, flush, random_case, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}


def _compare(fd, optical, method, precision, beta=0., phi=0., **kw):
    field, wl, ps = fd
    ref = oracle.transfer_field((field, wl, ps), optical, beta=beta, phi=phi, method=method, **kw)[0]
    dt = np.complex64 if precision == "c64" else np.complex128
    flush("  GPU transfer %s %s field %s" % (method, precision, field.shape))
    out = dm.transfer_field((field.astype(dt), wl, ps), optical, beta=beta, phi=phi,
                            method=method, precision=precision, **kw)[0]
    assert out.shape == ref.shape
    err = rel_err(out, ref)
    ierr = rel_err(dm.field2intensity(out), oracle.field2intensity(ref))
    flush("  rel field err %.3e  intensity err %.3e" % (err, ierr))
    assert np.isfinite(err)
    assert err < TOL[precision], err
    return err


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_droplet_small(method, precision):
    flush("case droplet_small %s %s" % (method, precision))
    fd, optical = droplet_case(shape=(12, 32, 32))
    _compare(fd, optical, method, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_random_nonsquare_two_pols(method, precision):
    flush("case random_nonsquare %s %s" % (method, precision))
    fd, optical = random_case(shape=(10, 24, 40), wavelengths=(450., 650.), jones=None)
    _compare(fd, optical, method, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_multi_beam(method, precision):
    flush("case multi_beam %s %s" % (method, precision))
    beta = [0., 0.2, 0.15]
    phi = [0., 1.0, -2.0]
    fd, optical = random_case(shape=(6, 32, 48), wavelengths=(500., 600.), jones=None,
                              beta=beta, phi=phi)
    _compare(fd, optical, method, precision, beta=beta, phi=phi)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_nstep_and_offaxis_single_beam(method):
    flush("case nstep2 offaxis %s" % method)
    fd, optical = random_case(shape=(5, 32, 32), wavelengths=(550.,), jones=(1, 1j),
                              beta=0.1, phi=0.3)
    _compare(fd, optical, method, "c128", beta=0.1, phi=0.3, nstep=2)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_two_blocks_different_effective_media(method):
    """Blocks with different mean eps: the 4x4 path uses the dense k-space
    boundary operator between them."""
    flush("case two_blocks %s" % method)
    fd, b1 = random_case(shape=(4, 32, 32), wavelengths=(550.,), jones=(1, 0))
    _, b2 = droplet_case(shape=(5, 32, 32), radius=12, ne=1.75, nhost=1.55)
    _compare(fd, [b1[0], b2[0]], method, "c128", nin=1.0, nout=1.33)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_absorbing_material(method):
    flush("case absorbing %s" % method)
    fd, optical = random_case(shape=(6, 32, 32), wavelengths=(550.,), jones=(0, 1))
    d, eps, ang = optical[0]
    eps = eps + 0.01j
    _compare(fd, [(d, eps, ang)], method, "c128")


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_config1_droplet_full_size(precision):
    """96x96 pixels, 60 layers, radius-30 radial droplet, 550 nm x-polarised
    Blackman-windowed normal incidence, 4x4, diffraction 1, npass 1."""
    flush("case config1 %s" % precision)
    fd, optical = droplet_case(shape=(60, 96, 96), radius=30, wavelengths=(550.,), jones=(1, 0))
    _compare(fd, optical, "4x4", precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("n", [3, 5])
def test_diffraction_windows(n, precision):
    """diffraction = n > 1: the field is split by the n x n k-space tent
    windows, each propagated with its own mean beam, and summed."""
    flush("case windows n=%d %s" % (n, precision))
    fd, optical = random_case(shape=(6, 32, 48), wavelengths=(450., 650.), jones=None)
    _compare(fd, optical, "4x4", precision, diffraction=n)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [2, 4])
def test_diffraction_windows_multi_beam(n):
    flush("case windows multi_beam n=%d" % n)
    beta = [0., 0.2, 0.15]
    phi = [0., 1.0, -2.0]
    fd, optical = random_case(shape=(4, 32, 32), wavelengths=(500., 600.), jones=None,
                              beta=beta, phi=phi)
    _compare(fd, optical, "4x4", "c128", beta=beta, phi=phi, diffraction=n)


@pytest.mark.gpu
def test_diffraction_windows_nstep_offaxis():
    """Sub-steps run inside each window (propagate_4x4.py:408-413)."""
    flush("case windows nstep2 offaxis")
    fd, optical = random_case(shape=(4, 32, 32), wavelengths=(550.,), jones=(1, 1j),
                              beta=0.1, phi=0.3)
    _compare(fd, optical, "4x4", "c128", beta=0.1, phi=0.3, nstep=2, diffraction=3)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,precision", [((3, 32, 32), "c128"), ((2, 16, 1024), "c64")])
def test_diffraction_windows_nstep_two_pols(shape, precision):
    """Windows with sub-steps and two rays per beam, where the row kernel
    takes one ray per work-group (complex128, and 1024-point rows)."""
    flush("case windows nstep2 two pols %s %s" % (shape, precision))
    fd, optical = random_case(shape=shape, wavelengths=(500., 650.), jones=None,
                              beta=0.1, phi=0.3)
    _compare(fd, optical, "4x4", precision, beta=0.1, phi=0.3, nstep=2, diffraction=3)


@pytest.mark.gpu
def test_diffraction_windows_two_blocks():
    flush("case windows two_blocks")
    fd, b1 = random_case(shape=(3, 32, 32), wavelengths=(550.,), jones=(1, 0))
    _, b2 = droplet_case(shape=(4, 32, 32), radius=12, ne=1.75, nhost=1.55)
    _compare(fd, [b1[0], b2[0]], "4x4", "c128", nin=1.0, nout=1.33, diffraction=3)


def _compare_npass(fd, optical, precision, **kw):
    """npass > 1: output field and the input-side field left in the input
    array (reflections) against the oracle."""
    field, wl, ps = fd
    f_ref = np.array(field, dtype=np.complex128)
    ref = oracle.transfer_field((f_ref, wl, ps), optical, method="4x4", **kw)[0]
    dt = np.complex64 if precision == "c64" else np.complex128
    f_gpu = np.array(field, dtype=dt)
    flush("  GPU npass transfer %s field %s %s" % (precision, field.shape, kw))
    out = dm.transfer_field((f_gpu, wl, ps), optical, method="4x4", precision=precision, **kw)[0]
    err = rel_err(out, ref)
    rerr = rel_err(f_gpu, f_ref)
    flush("  rel field err %.3e  input-side err %.3e" % (err, rerr))
    assert err < TOL[precision], err
    assert rerr < TOL[precision], rerr


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_npass3_reflections(precision):
    flush("case npass3 %s" % precision)
    fd, optical = random_case(shape=(6, 32, 48), wavelengths=(450., 650.), jones=None)
    _compare_npass(fd, optical, precision, beta=0., phi=0., npass=3, nin=1.0, nout=1.33)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("diffraction,npass,norm", [(1, 3, 1), (0, 3, 1), (1, 2, 0), (0, 4, 0)])
def test_reflection2_npass(diffraction, npass, norm, precision):
    """4x4 reflection mode 2 with npass > 1: real-space passes with the
    normalisations between them (dtmm/transfer.py:719-725, :815-872)."""
    flush("case refl2 npass d=%d npass=%d norm=%d %s" % (diffraction, npass, norm, precision))
    fd, optical = random_case(shape=(5, 32, 48), wavelengths=(480., 640.), jones=None,
                              beta=0.1, phi=0.4)
    _compare_npass(fd, optical, precision, beta=0.1, phi=0.4, npass=npass, norm=norm,
                   diffraction=diffraction, reflection=2, nin=1.0, nout=1.4)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("diffraction,npass", [(3, 1), (2, 1), (3, 2)])
def test_reflection2_windows(diffraction, npass, precision):
    """4x4 reflection mode 2 with diffraction > 1: per layer, the real-space
    parts of the tent windows through the sub-steps with their mean beams,
    summed (dtmm/propagate_4x4.py:176-204); two oblique beams, sub-steps."""
    flush("case refl2 windows d=%d npass=%d %s" % (diffraction, npass, precision))
    beta, phi = [0.1, 0.2], [0.4, -1.0]
    fd, optical = random_case(shape=(3, 24, 32), wavelengths=(500., 620.), jones=None,
                              beta=beta, phi=phi)
    if npass == 1:
        _compare(fd, optical, "4x4", precision, beta=beta, phi=phi, reflection=2,
                 diffraction=diffraction, nstep=2, nin=1.1, nout=1.3)
    else:
        _compare_npass(fd, optical, precision, beta=beta, phi=phi, npass=npass, reflection=2,
                       diffraction=diffraction, nin=1.1, nout=1.3)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("diffraction", [1, 0])
def test_2x2_reflection2(diffraction, precision):
    """2x2 reflection mode 2 (real space, per-pixel interface transmission
    into every layer, dtmm/propagate_2x2.py:91-190): two oblique beams,
    sub-steps, two blocks with different effective media, and the bulk."""
    flush("case 2x2 refl2 d=%d %s" % (diffraction, precision))
    beta, phi = [0.05, 0.15], [0.3, -1.2]
    fd, b1 = random_case(shape=(3, 32, 48), wavelengths=(470., 630.), jones=None,
                         beta=beta, phi=phi)
    _, b2 = droplet_case(shape=(2, 32, 48), radius=10, ne=1.72, nhost=1.55)
    kw = dict(beta=beta, phi=phi, reflection=2, diffraction=diffraction, nstep=2, nin=1.1,
              nout=1.35)
    _compare(fd, [b1[0], b2[0]], "2x2", precision, **kw)
    field, wl, ps = fd
    ref = oracle.transfer_field((field, wl, ps), [b1[0], b2[0]], method="2x2", ret_bulk=True,
                                **kw)[0]
    dt = np.complex64 if precision == "c64" else np.complex128
    out = dm.transfer_field((field.astype(dt), wl, ps), [b1[0], b2[0]], method="2x2",
                            precision=precision, ret_bulk=True, **kw)[0]
    assert out.shape == ref.shape
    for j in range(out.shape[0]):
        assert rel_err(out[j], ref[j]) < TOL[precision], j


@pytest.mark.gpu
def test_illumination_rays_multi_beam():
    """Beams from illumination_rays (dtmm/field.py:371-404): the nine rays of
    a 3-pixel aperture (the corner centres lie at r = sqrt(2) / 1.5 < 1) at
    NA 0.1, both polarisations, through a droplet."""
    flush("case illumination rays")
    beta, phi, inten = dm.illumination_rays(0.1, 3, smooth=0.)
    assert len(beta) == 9
    _, optical = droplet_case(shape=(6, 24, 24), radius=8)
    fd = dm.illumination_data((24, 24), [520.], pixelsize=100, beta=beta, phi=phi,
                              intensity=inten, jones=None, window=dm.blackman((24, 24)))
    _compare(fd, optical, "4x4", "c128", beta=beta, phi=phi)


@pytest.mark.gpu
@pytest.mark.parametrize("npass,norm", [(2, 1), (4, 0), (5, 1)])
def test_npass_norms(npass, norm):
    flush("case npass=%d norm=%d" % (npass, norm))
    fd, optical = droplet_case(shape=(8, 32, 32), radius=12, ne=1.75, nhost=1.55,
                               wavelengths=(550.,), jones=(1, 0))
    _compare_npass(fd, optical, "c128", beta=0., phi=0., npass=npass, norm=norm, nin=1.2,
                   nout=1.0)


@pytest.mark.gpu
def test_npass_with_windows_multi_beam():
    flush("case npass3 windows multi_beam")
    beta = [0., 0.15]
    phi = [0., 0.8]
    fd, optical = random_case(shape=(4, 32, 32), wavelengths=(500., 600.), jones=None,
                              beta=beta, phi=phi)
    _compare_npass(fd, optical, "c128", beta=beta, phi=phi, npass=3, diffraction=3, nin=1.0,
                   nout=1.5, nstep=2)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
@pytest.mark.parametrize("shape", [(3, 32, 1024), (3, 1024, 32)])
def test_1024_point_lines(shape, method, precision):
    """1024-point rows (two waves per line) and columns."""
    flush("case 1024 lines %s %s %s" % (shape, method, precision))
    fd, optical = random_case(shape=shape, wavelengths=(500., 600.), jones=None)
    _compare(fd, optical, method, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(2, 32, 1024), (2, 1024, 32)])
def test_1024_point_lines_windows_npass(shape):
    flush("case 1024 lines windows npass %s" % (shape,))
    beta = [0.1, 0.1]
    phi = [0., 2.1]
    fd, optical = random_case(shape=shape, wavelengths=(550.,), jones=None, beta=beta, phi=phi)
    _compare_npass(fd, optical, "c64", beta=beta, phi=phi, npass=3, diffraction=3, nin=1.52,
                   nout=1.52)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_2x2_reflection1(precision):
    """2x2 with reflection mode 1: Fresnel transmission at the input, block
    and output interfaces (propagate_2x2.py:28-30)."""
    flush("case 2x2 reflection1 %s" % precision)
    fd, optical = random_case(shape=(6, 32, 48), wavelengths=(450., 650.), jones=None)
    _compare(fd, optical, "2x2", precision, reflection=1, nin=1.0, nout=1.33)


@pytest.mark.gpu
def test_2x2_reflection1_two_blocks_multi_beam():
    flush("case 2x2 reflection1 two blocks multi beam")
    beta = [0., 0.2]
    phi = [0., 1.0]
    fd, b1 = random_case(shape=(3, 32, 32), wavelengths=(550.,), jones=None, beta=beta, phi=phi)
    _, b2 = droplet_case(shape=(4, 32, 32), radius=12, ne=1.75, nhost=1.55)
    _compare(fd, [b1[0], b2[0]], "2x2", "c128", beta=beta, phi=phi, reflection=1, nin=1.2,
             nout=1.0, nstep=2)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_default_method(precision):
    """transfer_field without method: the configured '"2x2"' default, which
    selects reflection mode 1 (transfer.py:459)."""
    flush("case default method %s" % precision)
    fd, optical = droplet_case(shape=(8, 32, 32), radius=12)
    _compare(fd, optical, None, precision)


@pytest.mark.gpu
def test_transfer_batch_uneven_samples():
    """Samples of different grid, depth, beams and method in one batched
    engine call, each against the oracle."""
    flush("case batch of uneven samples")
    beta7 = [0.] + [0.2] * 6
    phi7 = [0.] + [k * np.pi / 3 for k in range(6)]
    s1 = random_case(shape=(5, 32, 48), wavelengths=(450., 650.), jones=None)
    s2 = droplet_case(shape=(9, 24, 24), radius=9, wavelengths=(550.,), jones=(1, 0))
    s3 = random_case(shape=(3, 16, 32), wavelengths=(500., 600.), jones=None, beta=beta7,
                     phi=phi7, ne=1.8)
    samples = [
        {"field_data": s1[0], "optical_data": s1[1], "beta": 0., "phi": 0.},
        {"field_data": s2[0], "optical_data": s2[1], "beta": 0., "phi": 0., "method": "2x2"},
        {"field_data": s3[0], "optical_data": s3[1], "beta": beta7, "phi": phi7},
    ]
    for precision, dt in (("c128", np.complex128), ("c64", np.complex64)):
        batch = [dict(s, field_data=(s["field_data"][0].astype(dt),) + tuple(s["field_data"][1:]))
                 for s in samples]
        res = dm.transfer_batch(batch, method="4x4", precision=precision)
        assert len(res) == 3
        for smp, (out, wl, ps) in zip(samples, res):
            fd = smp["field_data"]
            ref = oracle.transfer_field(fd, smp["optical_data"], beta=smp["beta"], phi=smp["phi"],
                                        method=smp.get("method", "4x4"))[0]
            err = rel_err(out, ref)
            flush("  %s sample %s err %.3e" % (precision, out.shape, err))
            assert err < TOL[precision], err


@pytest.mark.gpu
def test_transfer_graph_replays_match_eager_batch():
    """The batch captured as one HIP graph: every replay, with eager engine
    work and host synchronisation between replays and new inputs set before
    one of them, gives exactly the eager batch's output (same kernels, same
    inputs), including a sample with npass = 3 that overwrites its input."""
    flush("case transfer graph replays")
    s1 = random_case(shape=(5, 32, 48), wavelengths=(450., 650.), jones=None, seed=3)
    s2 = droplet_case(shape=(9, 24, 24), radius=9, wavelengths=(550.,), jones=(1, 0))
    s3 = random_case(shape=(4, 32, 32), wavelengths=(500.,), jones=(1, 0), seed=5)
    samples = [
        {"field_data": s1[0], "optical_data": s1[1], "beta": 0., "phi": 0.},
        {"field_data": s2[0], "optical_data": s2[1], "beta": 0., "phi": 0., "method": "2x2"},
        {"field_data": s3[0], "optical_data": s3[1], "beta": 0., "phi": 0., "npass": 3,
         "nin": 1.0, "nout": 1.3},
    ]

    def batch(fields):
        b = []
        for s, f in zip(samples, fields):
            b.append(dict(s, field_data=(f.copy(),) + tuple(s["field_data"][1:])))
        return b

    fields0 = [s["field_data"][0].astype(np.complex64) for s in samples]
    rng = np.random.default_rng(11)
    fields1 = [(f * (0.5 + rng.uniform(size=f.shape))).astype(np.complex64) for f in fields0]
    eager0 = [r[0] for r in dm.transfer_batch(batch(fields0), method="4x4", precision="c64")]
    eager1 = [r[0] for r in dm.transfer_batch(batch(fields1), method="4x4", precision="c64")]
    g = dm.TransferGraph(batch(fields0), method="4x4", precision="c64")
    for rep, (fields, eager) in enumerate([(None, eager0), (None, eager0), (fields1, eager1),
                                           (None, eager1), (fields0, eager0)]):
        if fields is not None:
            g.set_inputs(fields)
        g.replay()
        out = [r[0] for r in g.results()]
        # eager engine work between replays
        dm.transfer_field((fields0[1].copy(),) + tuple(samples[1]["field_data"][1:]),
                          samples[1]["optical_data"], beta=0., phi=0., method="2x2",
                          precision="c64")
        for i, (a, b) in enumerate(zip(out, eager)):
            flush("  replay %d sample %d max diff %.3e" % (rep, i, float(np.abs(a - b).max())))
            assert np.array_equal(a, b), (rep, i)


def _oracle_analyser_loss(fd, d, eps, ang, roi, jones=(1, 1), n=1.5, beta=0., phi=0., nstep=1):
    """The analysed-intensity loss on the oracle's output field."""
    field, wl, ps = fd
    out = oracle.transfer_field((field, wl, ps), [(d, eps, ang)], beta=beta, phi=phi, nstep=nstep,
                                method="4x4")[0]
    m = dm.polarizer4x4(jones, n)
    y0, y1, x0, x1 = roi
    sub = out[..., :, y0:y1, x0:x1]
    g = np.stack([sum(m[i, j] * sub[..., j, :, :] for j in range(4)) for i in range(4)], axis=-3)
    return float(((g[..., 0, :, :] * g[..., 1, :, :].conj()).real -
                  (g[..., 2, :, :] * g[..., 3, :, :].conj()).real).sum())


@pytest.mark.gpu
def test_adjoint_gradient_matches_oracle_finite_differences():
    """Gradient of the analysed intensity (45 degree analyser, central
    region) with respect to theta and phi of single voxels: the adjoint
    transfer (c128) against central differences of the oracle loss, and the
    c64 adjoint against the c128 one."""
    flush("case adjoint gradient")
    # 256-point rows: the layer operator of every pixel runs on threads of
    # all four component waves
    fd, optical = random_case(shape=(3, 16, 256), wavelengths=(500., 600.), jones=(1, 0))
    d, eps, ang = optical[0]
    eps = np.ascontiguousarray(np.broadcast_to(eps, ang.shape))
    ang = np.ascontiguousarray(ang)
    roi = (4, 12, 64, 192)
    tr = dm.AdjointTransfer(fd, [(d, eps, ang)], beta=0., phi=0., precision="c128")
    out = tr.forward()
    loss, seed = dm.analyser_intensity_loss(out, (1, 1), 1.5, roi)
    ref_loss = _oracle_analyser_loss(fd, d, eps, ang, roi)
    flush("  loss gpu %.12g oracle %.12g" % (loss, ref_loss))
    assert abs(loss - ref_loss) < 1e-9 * abs(ref_loss)
    grad = tr.backward(seed).cpu().numpy()
    gmax = np.abs(grad).max()
    assert gmax > 0
    rng = np.random.default_rng(5)
    h = 1e-5
    for i in range(8):
        l, y = rng.integers(0, 3), rng.integers(0, 16)
        x = rng.integers(64 * (i % 4), 64 * (i % 4) + 64)
        for a in (1, 2):
            ap = ang.copy()
            am = ang.copy()
            ap[l, y, x, a] += h
            am[l, y, x, a] -= h
            fdv = (_oracle_analyser_loss(fd, d, eps, ap, roi) -
                   _oracle_analyser_loss(fd, d, eps, am, roi)) / (2 * h)
            g = grad[a - 1, l, y, x]
            flush("  voxel (%d,%d,%d) angle %d: adjoint %.9e fd %.9e" % (l, y, x, a, g, fdv))
            assert abs(g - fdv) < 1e-6 * gmax
    tr64 = dm.AdjointTransfer((fd[0].astype(np.complex64),) + tuple(fd[1:]), [(d, eps, ang)],
                              beta=0., phi=0., precision="c64")
    out64 = tr64.forward()
    _, seed64 = dm.analyser_intensity_loss(out64, (1, 1), 1.5, roi)
    g64 = tr64.backward(seed64).cpu().numpy()
    err = np.abs(g64 - grad).max() / gmax
    flush("  c64 vs c128 gradient err %.3e" % err)
    assert err < 1e-3


@pytest.mark.gpu
def test_adjoint_gradient_beams_and_substeps():
    """Two oblique beams of two polarisations each (four rays) and two
    sub-steps per layer: the c128 adjoint gradient against central
    differences of the oracle loss, which sums over every ray."""
    flush("case adjoint gradient, 2 beams x 2 polarisations, nstep 2")
    beta, phi = np.array([0., 0.15]), np.array([0., 0.7])
    fd, optical = random_case(shape=(2, 16, 64), wavelengths=(500., 600.), beta=beta, phi=phi)
    d, eps, ang = optical[0]
    eps = np.ascontiguousarray(np.broadcast_to(eps, ang.shape))
    ang = np.ascontiguousarray(ang)
    roi = (4, 12, 16, 48)
    tr = dm.AdjointTransfer(fd, [(d, eps, ang)], beta=beta, phi=phi, nstep=2, precision="c128")
    out = tr.forward()
    loss, seed = dm.analyser_intensity_loss(out, (1, 1), 1.5, roi)
    ref_loss = _oracle_analyser_loss(fd, d, eps, ang, roi, beta=beta, phi=phi, nstep=2)
    flush("  loss gpu %.12g oracle %.12g" % (loss, ref_loss))
    assert abs(loss - ref_loss) < 1e-9 * abs(ref_loss)
    grad = tr.backward(seed).cpu().numpy()
    gmax = np.abs(grad).max()
    assert gmax > 0
    rng = np.random.default_rng(11)
    h = 1e-5
    for _ in range(4):
        l, y, x = rng.integers(0, 2), rng.integers(4, 12), rng.integers(16, 48)
        for a in (1, 2):
            ap = ang.copy()
            am = ang.copy()
            ap[l, y, x, a] += h
            am[l, y, x, a] -= h
            fdv = (_oracle_analyser_loss(fd, d, eps, ap, roi, beta=beta, phi=phi, nstep=2) -
                   _oracle_analyser_loss(fd, d, eps, am, roi, beta=beta, phi=phi, nstep=2)) / (2 * h)
            g = grad[a - 1, l, y, x]
            flush("  voxel (%d,%d,%d) angle %d: adjoint %.9e fd %.9e" % (l, y, x, a, g, fdv))
            assert abs(g - fdv) < 1e-6 * gmax


@pytest.mark.gpu
def test_gradient_graph_replays_match_eager():
    """The gradient evaluation captured as one HIP graph: every replay, with
    eager work and synchronisations in between, equals the eager result."""
    flush("case gradient graph replays")
    fd, optical = random_case(shape=(3, 32, 64), wavelengths=(500., 600.), jones=(1, 0))
    roi = (8, 24, 16, 48)
    tr = dm.AdjointTransfer(fd, optical, beta=0., phi=0., precision="c128")
    out = tr.forward()
    loss0, seed = dm.analyser_intensity_loss(out, (1, 1), 1.5, roi)
    g0 = tr.backward(seed).cpu().numpy()
    replay = tr.gradient_graph(
        lambda f: dm.analyser_intensity_loss(f, (1, 1), 1.5, roi, as_tensor=True))
    import torch
    for i in range(3):
        loss, grad = replay()
        torch.cuda.synchronize()
        g = grad.cpu().numpy()
        flush("  replay %d loss %.12g max diff %.3e" % (i, float(loss.item()), np.abs(g - g0).max()))
        assert abs(float(loss.item()) - loss0) <= 1e-12 * abs(loss0)
        np.testing.assert_allclose(g, g0, rtol=0, atol=1e-12 * np.abs(g0).max())
        # eager work between replays
        out2 = dm.transfer_field(fd, optical, beta=0., phi=0., method="4x4")[0]
        assert np.isfinite(out2).all()


def _biaxial_case(shape=(6, 24, 40), seed=5, wavelengths=(450., 650.), jones=None, beta=0.,
                  phi=0., uniaxial_every_other=True):
    """Random texture with biaxial principal values (1.5^2, 1.6^2, 1.75^2)
    and a random third Euler angle psi; every other layer uniaxial, so
    layers with and without the numeric eigensolve alternate."""
    fd, optical = random_case(shape=shape, seed=seed, wavelengths=wavelengths, jones=jones,
                              beta=beta, phi=phi)
    d, eps, ang = optical[0]
    rng = np.random.default_rng(seed)
    eps = np.array(np.broadcast_to(eps, ang.shape))
    eps[..., 0], eps[..., 1], eps[..., 2] = 1.5 ** 2, 1.6 ** 2, 1.75 ** 2
    if uniaxial_every_other:
        eps[1::2, ..., 1] = eps[1::2, ..., 0]
    ang = np.array(ang)
    ang[..., 0] = rng.uniform(-np.pi, np.pi, shape)
    return fd, [(d, eps, ang)]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_biaxial_layers(method, precision):
    """Biaxial voxels: per-layer table of the numeric eigensolve (Hessenberg
    + shifted QR, Poynting sort) against the oracle's numpy.linalg.eig."""
    flush("case biaxial %s %s" % (method, precision))
    fd, optical = _biaxial_case()
    _compare(fd, optical, method, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_biaxial_multi_beam_nstep(method):
    flush("case biaxial multi beam nstep %s" % method)
    beta, phi = [0., 0.2], [0., 0.8]
    fd, optical = _biaxial_case(shape=(4, 32, 48), beta=beta, phi=phi)
    _compare(fd, optical, method, "c128", beta=beta, phi=phi, nstep=2)


@pytest.mark.gpu
def test_biaxial_windows_nstep():
    """Biaxial layers with sub-steps inside each window: the window's
    operator table is built by the first sub-step and reused by the next."""
    flush("case biaxial windows nstep2")
    fd, optical = _biaxial_case(shape=(3, 32, 32), uniaxial_every_other=False)
    _compare(fd, optical, "4x4", "c128", diffraction=3, nstep=2)


@pytest.mark.gpu
def test_biaxial_windows_npass():
    """Biaxial layers under diffraction 3 (table per window beam) and npass 2."""
    flush("case biaxial windows npass")
    fd, optical = _biaxial_case(shape=(4, 32, 32), uniaxial_every_other=False)
    _compare(fd, optical, "4x4", "c128", diffraction=3)
    _compare_npass(fd, optical, "c128", beta=0., phi=0., npass=2, nin=1.0, nout=1.5)


@pytest.mark.gpu
def test_biaxial_adjoint_is_refused():
    """The adjoint transfer has no biaxial derivatives: it raises."""
    flush("case biaxial adjoint refused")
    fd, optical = _biaxial_case(shape=(2, 16, 32), jones=(1, 0))
    tr = dm.AdjointTransfer(fd, optical, beta=0., phi=0., precision="c128")
    with pytest.raises(NotImplementedError):
        tr.forward()


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_launch_shapes_do_not_change_results_and_tuner_picks_one(method):
    """Every (column lines, row rays) launch shape and every wavelength split
    of the row kernels gives the same output bit for bit; the autotuner picks
    one of them and the tuned transfer matches the oracle."""
    import torch
    from dtmm_amd import tune
    flush("case launch shapes %s" % method)
    fd, optical = random_case(shape=(6, 64, 96), wavelengths=(450., 550., 650.), jones=None)
    field, wl, ps = fd
    job = dm.transfer._prepare((field.astype(np.complex64), wl, ps), optical, beta=0., phi=0.,
                               method=method, precision="c64")
    cands = tune.candidates(job.plan)
    assert (4 if method == "4x4" else 2) in [c for c, _ in cands]
    assert len(cands) >= 4
    outs = {}
    for c, r in cands:
        tune.set_tiles(job.plan, c, r)
        assert tune.get_tiles(job.plan) == (c, r)
        job.plan.execute(job.stack, job.fin, job.fout, workspace=job.ws)
        torch.cuda.synchronize()
        outs[(c, r)] = job.fout.cpu().numpy()
    # wavelength splits of the row kernels: 2 gives uneven chunks (2 + 1 of
    # 3 wavelengths), 8 is capped at the 3 wavelengths
    for sp in (1, 2, 3, 8):
        tune.set_row_split(job.plan, sp)
        assert tune.get_row_split(job.plan) == sp
        job.plan.execute(job.stack, job.fin, job.fout, workspace=job.ws)
        torch.cuda.synchronize()
        outs[("split", sp)] = job.fout.cpu().numpy()
    tune.set_row_split(job.plan, 0)
    first = next(iter(outs.values()))
    for k, v in outs.items():
        assert np.array_equal(v, first), k
    res = tune.tune_plan(job.plan, job.stack, job.fin, job.fout, job.ws, cache=False)
    flush("  tuned %s split %s: %s %s" % (res["tiles"], res["row_split"],
                                          {k: round(v, 3) for k, v in res["ms"].items()},
                                          {k: round(v, 3) for k, v in res["split_ms"].items()}))
    assert res["tiles"] in cands and set(res["ms"]) == set(cands)
    assert res["row_split"] in [0] + tune.split_candidates(job.plan)
    out = dm.transfer_field((field.astype(np.complex64), wl, ps), optical, beta=0., phi=0.,
                            method=method, precision="c64", tune=True)[0]
    ref = oracle.transfer_field(fd, optical, beta=0., phi=0., method=method)[0]
    assert rel_err(out, ref) < TOL["c64"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_no_diffraction(method, precision):
    """diffraction=0: 4x4 in reflection mode 2 (the reference's choice for
    diffraction 0) and 2x2 in reflection mode 0, per-pixel stacks in real
    space; two beams, sub-steps and different input / output media."""
    flush("case no diffraction %s %s" % (method, precision))
    beta = [0., 0.2]
    phi = [0., 1.0]
    fd, optical = random_case(shape=(6, 32, 48), wavelengths=(450., 650.), jones=None,
                              beta=beta, phi=phi)
    _compare(fd, optical, method, precision, beta=beta, phi=phi, diffraction=0, nstep=2,
             nin=1.0, nout=1.33)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_no_diffraction_biaxial_absorbing(method):
    flush("case no diffraction biaxial absorbing %s" % method)
    fd, optical = random_case(shape=(4, 24, 32), wavelengths=(550.,), jones=(1, 1j), beta=0.15,
                              phi=0.4)
    d, eps, ang = optical[0]
    eps = np.array(eps, complex)
    eps[..., 1] = eps[..., 1] * 1.05 + 0.01j
    ang = np.array(ang)
    ang[..., 0] = 0.3
    _compare(fd, [(d, eps, ang)], method, "c128", beta=0.15, phi=0.4, diffraction=0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_reflection2_diffraction(precision):
    """4x4 reflection mode 2 with diffraction 1: per-pixel half steps around
    the k-space E diffraction, field in real space; two beams, sub-steps,
    different input / output media."""
    flush("case reflection 2 diffraction 1 %s" % precision)
    beta = [0., 0.15]
    phi = [0., -0.7]
    fd, optical = random_case(shape=(5, 32, 48), wavelengths=(450., 650.), jones=None,
                              beta=beta, phi=phi)
    _compare(fd, optical, "4x4", precision, beta=beta, phi=phi, reflection=2, diffraction=1,
             nstep=2, nin=1.0, nout=1.33)


@pytest.mark.gpu
def test_reflection2_diffraction_blocks_biaxial():
    flush("case reflection 2 diffraction 1 blocks biaxial")
    fd, b1 = random_case(shape=(3, 32, 32), wavelengths=(550.,), jones=(1, 0))
    _, b2 = droplet_case(shape=(3, 32, 32), radius=12, ne=1.75, nhost=1.55)
    d, eps, ang = b1[0]
    eps = np.array(eps)
    eps[..., 1] *= 1.03
    _compare(fd, [(d, eps, ang), b2[0]], "4x4", "c128", reflection=2, diffraction=1, nin=1.2,
             nout=1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("method,reflection,diffraction", [("4x4", 1, 1), ("4x4", 2, 1),
                                                           ("4x4", 2, 0), ("2x2", 0, 1),
                                                           ("2x2", 1, 1), ("2x2", 0, 0)])
def test_ret_bulk(method, reflection, diffraction, precision):
    """ret_bulk: the field after every layer (plus input and output) against
    the oracle's bulk, with sub-steps and two beams."""
    flush("case ret_bulk %s reflection %d diffraction %d %s"
          % (method, reflection, diffraction, precision))
    beta = [0., 0.15]
    phi = [0., 0.9]
    fd, optical = random_case(shape=(4, 32, 48), wavelengths=(450., 650.), jones=None,
                              beta=beta, phi=phi)
    field, wl, ps = fd
    kw = dict(beta=beta, phi=phi, method=method, reflection=reflection, diffraction=diffraction,
              nstep=2, nin=1.2, nout=1.0)
    ref = oracle.transfer_field((field, wl, ps), optical, ret_bulk=True, **kw)[0]
    dt = np.complex64 if precision == "c64" else np.complex128
    out = dm.transfer_field((field.astype(dt), wl, ps), optical, precision=precision,
                            ret_bulk=True, **kw)[0]
    assert out.shape == ref.shape == (6,) + field.shape
    for j in range(out.shape[0]):
        err = rel_err(out[j], ref[j])
        flush("  slot %d rel err %.3e" % (j, err))
        assert err < TOL[precision], (j, err)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("diffraction,nstep,biax", [(-1, 1, False), (np.inf, 2, True)])
def test_full_diffraction(diffraction, nstep, biax, precision):
    """diffraction < 0 or inf: exact per-mode propagation of every FFT mode
    with its own beam at every pixel (dtmm/propagate_4x4.py:515-564), both
    polarisations, two wavelengths, sub-steps, and biaxial voxels."""
    flush("case full diffraction %s nstep=%d biaxial=%s %s" % (diffraction, nstep, biax,
                                                                 precision))
    if biax:
        fd, optical = _biaxial_case(shape=(3, 16, 24), wavelengths=(480., 620.), jones=None)
    else:
        fd, optical = random_case(shape=(3, 16, 24), wavelengths=(480., 620.), jones=None)
    _compare(fd, optical, "4x4", precision, diffraction=diffraction, nstep=nstep)
