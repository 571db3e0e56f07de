"""GPU parity at the line lengths and launch shapes the benchmarks time.

The row and column kernels are compiled per line length (128, 192, 256,
384, 512, 1024 points), so each size the benchmark workloads run is checked
against the complex128 oracle here, on the workloads' own sample kinds with
few layers:

* the 512 x 512 smooth random texture (both methods, both precisions, every
  launch shape the autotuner may pick);
* one batched call with 128^2, 192^2 and 384^2 samples and the seven
  hexagonal beams of the uneven batch;
* 256 x 256 with the three wavelengths of the gradient workload;
* 1024 x 1024 lines on both axes;
* the reflection workload's physics on a 96 x 96 crop: a 3 x 3 array of
  radial droplets in n = 1.5 between n = 1.52 glass, three beams at
  beta 0.1, phi 0/120/240 deg, both polarisations, two wavelengths, 4x4,
  diffraction 5 (25 tent windows) with npass 3.

Tolerances as in test_gpu_transfer: 1e-9 (c128) and 1e-4 (c64) of
max|F - F_ref| / max|F_ref|.
"""
import functools

import numpy as np
import pytest

import dtmm_amd as dm
import oracle
from dtmm_amd.samples import droplet_array_data, droplet_grid_centres, random_director_data
from helpers import flush, rel_err

TOL = {"c128": 1e-9, "c64": 1e-4}
DT = {"c128": np.complex128, "c64": np.complex64}
HEX_BEAMS = ([0.] + [0.2] * 6, [0.] + [k * np.pi / 3 for k in range(6)])
C5_BEAMS = ([0.1, 0.1, 0.1], [0., 2 * np.pi / 3, 4 * np.pi / 3])


def _texture(size, layers, wavelengths, jones=None, beta=0., phi=0., seed=0):
    optical = [random_director_data((layers, size, size), seed=seed, no=1.5, ne=1.7)]
    fd = dm.illumination_data((size, size), list(wavelengths), pixelsize=100, beta=beta, phi=phi,
                              jones=jones, window=dm.blackman((size, size)))
    return fd, optical


@functools.lru_cache(maxsize=None)
def _c2_case(method):
    oracle.set_fft_workers(8)
    fd, optical = _texture(512, 3, (450., 650.))
    ref = oracle.transfer_field(fd, optical, beta=0., phi=0., method=method)[0]
    return fd, optical, ref


@pytest.mark.gpu
@pytest.mark.timeout(300)
@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_512_texture_every_launch_shape(method, precision):
    """The 512 x 512 random texture, x and y polarisations, two wavelengths,
    through every (column lines, row rays) launch shape of the plan."""
    import torch
    from dtmm_amd import tune
    flush("case 512 texture %s %s" % (method, precision))
    fd, optical, ref = _c2_case(method)
    field, wl, ps = fd
    job = dm.transfer._prepare((field.astype(DT[precision]), wl, ps), optical, beta=0., phi=0.,
                               method=method, precision=precision)
    for c, r in tune.candidates(job.plan):
        tune.set_tiles(job.plan, c, r)
        flush("  tiles col_lines=%d row_rays=%d" % (c, r))
        job.plan.execute(job.stack, job.fin, job.fout, workspace=job.ws)
        torch.cuda.synchronize()
        out = job.fout.cpu().numpy().reshape(ref.shape)
        err = rel_err(out, ref)
        ierr = rel_err(dm.field2intensity(out), oracle.field2intensity(ref))
        flush("  rel field err %.3e  intensity err %.3e" % (err, ierr))
        assert err < TOL[precision], ((c, r), err)


@functools.lru_cache(maxsize=None)
def _batch_case():
    oracle.set_fft_workers(8)
    beta, phi = HEX_BEAMS
    samples = []
    for size, layers, wl, kind in ((128, 3, (450., 650.), "random"),
                                   (192, 2, (550.,), "droplet"),
                                   (384, 2, (500.,), "random")):
        if kind == "random":
            optical = [random_director_data((layers, size, size), seed=size, no=1.5, ne=1.75)]
        else:
            optical = [dm.nematic_droplet_data((layers, size, size), 30. * size / 96, profile="r",
                                               no=1.5, ne=1.7, nhost=1.5)]
        fd = dm.illumination_data((size, size), list(wl), pixelsize=100, beta=beta, phi=phi,
                                  jones=None, window=dm.blackman((size, size)))
        ref = oracle.transfer_field(fd, optical, beta=beta, phi=phi, method="4x4")[0]
        samples.append((fd, optical, ref))
    return samples


@pytest.mark.gpu
@pytest.mark.timeout(300)
@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_batch_128_192_384_hex_beams(precision):
    """One transfer_batch call with 128^2, 192^2 and 384^2 samples, seven
    hexagonal beams x two polarisations each."""
    flush("case batch 128/192/384 %s" % precision)
    cases = _batch_case()
    beta, phi = HEX_BEAMS
    batch = [{"field_data": (fd[0].astype(DT[precision]),) + tuple(fd[1:]), "optical_data": opt}
             for fd, opt, _ in cases]
    res = dm.transfer_batch(batch, beta=beta, phi=phi, method="4x4", precision=precision)
    for (fd, _, ref), (out, _, _) in zip(cases, res):
        err = rel_err(out, ref)
        flush("  sample %s err %.3e" % (out.shape, err))
        assert err < TOL[precision], err


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_256_three_wavelengths(method, precision):
    """256 x 256 with the gradient workload's wavelengths (450, 550, 650 nm),
    x polarisation."""
    flush("case 256 texture %s %s" % (method, precision))
    fd, optical = _texture(256, 3, (450., 550., 650.), jones=(1, 0), seed=4)
    ref = oracle.transfer_field(fd, optical, beta=0., phi=0., method=method)[0]
    field, wl, ps = fd
    out = dm.transfer_field((field.astype(DT[precision]), wl, ps), optical, beta=0., phi=0.,
                            method=method, precision=precision)[0]
    err = rel_err(out, ref)
    flush("  rel field err %.3e" % err)
    assert err < TOL[precision], err


@pytest.mark.gpu
@pytest.mark.timeout(300)
@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("method", ["4x4", "2x2"])
def test_1024_square(method, precision):
    """1024-point rows and columns together (two waves per line on both
    passes), an oblique beam."""
    flush("case 1024 square %s %s" % (method, precision))
    oracle.set_fft_workers(8)
    fd, optical = _texture(1024, 2, (550.,), jones=(1, 1j), beta=0.1, phi=0.5, seed=2)
    ref = oracle.transfer_field(fd, optical, beta=0.1, phi=0.5, method=method)[0]
    field, wl, ps = fd
    out = dm.transfer_field((field.astype(DT[precision]), wl, ps), optical, beta=0.1, phi=0.5,
                            method=method, precision=precision)[0]
    err = rel_err(out, ref)
    flush("  rel field err %.3e" % err)
    assert err < TOL[precision], err


@functools.lru_cache(maxsize=None)
def _c5_crop():
    oracle.set_fft_workers(8)
    n, layers, radius = 96, 8, 10.
    optical = [droplet_array_data((layers, n, n), radius, droplet_grid_centres(n, 3), no=1.5,
                                  ne=1.6, nhost=1.5)]
    beta, phi = C5_BEAMS
    fd = dm.illumination_data((n, n), [450., 650.], pixelsize=100, beta=beta, phi=phi,
                              jones=None, window=dm.blackman((n, n)), n=1.52)
    f_ref = np.array(fd[0], np.complex128)
    kw = dict(beta=beta, phi=phi, method="4x4", npass=3, diffraction=5, nin=1.52, nout=1.52)
    ref = oracle.transfer_field((f_ref,) + tuple(fd[1:]), optical, **kw)[0]
    return fd, optical, kw, ref, f_ref


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_droplet_array_glass_diffraction5_npass3(precision):
    """The reflection workload's physics on a 96 x 96 x 8 crop: output field
    and the input-side field (input plus reflection) against the oracle."""
    flush("case droplet array diffraction 5 npass 3 %s (oracle first)" % precision)
    fd, optical, kw, ref, f_ref = _c5_crop()
    f_gpu = np.array(fd[0], DT[precision])
    out = dm.transfer_field((f_gpu,) + tuple(fd[1:]), optical, precision=precision, **kw)[0]
    err = rel_err(out, ref)
    rerr = rel_err(f_gpu, f_ref)
    flush("  rel field err %.3e  input-side err %.3e" % (err, rerr))
    assert err < TOL[precision], err
    assert rerr < TOL[precision], rerr
