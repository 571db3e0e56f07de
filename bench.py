#!/usr/bin/env python3
"""Benchmark of the diffractive transfer-matrix path on one MI355X.

Workload (the 512x512x128 scaled-up plain-method case): smooth random
nematic texture (seed 0; tilt and twist as sums of 16 random Fourier modes,
|k| <= 8 cycles per box, amplitudes N(0, 0.3) rad, planar base tilt and a
90 degree linear twist), eps_o = 1.5^2, eps_e = 1.7^2, pixel 100 nm, layer
thickness 1 px, 9 wavelengths 380-780 nm, x and y polarisations at normal
incidence (18 field planes), diffraction = 1, npass = 1.

A step is one full transfer_field pass (all layers, all planes) with the
inputs resident in HBM.  value = voxel*wavelength*plane steps per second:
H*W*L*n_wavelengths*n_polarisations / seconds per pass.

--workload c5 measures the reflection workload instead: 1024x1024x200, a
3x3 grid of radial droplets (radius 80 px, no 1.5, ne 1.6, host 1.5) between
glass (nin = nout = 1.52), 5 wavelengths 450-650 nm, 3 Blackman-windowed
beams at beta 0.1, phi 0/120/240 deg, both polarisations (30 planes), 4x4,
npass = 3, diffraction = 5; value there counts npass passes.

--workload c3 measures a batch of 64 uneven samples in one batched engine
call: grids uniform over {128,192,256,384,512}^2, depth uniform in 32..160
layers, radial droplet (radius 30/96 of the grid) or smooth random texture,
ne^2 uniform in [1.6^2, 1.8^2] (no 1.5), 7 beams (beta 0, and 0.2 at
phi = k 60 deg) x 2 polarisations x 9 wavelengths 380-780 nm (126 planes),
4x4, c64.  Samples of one (kind, grid) share the director of a 160-layer
volume cut to their depth.

--workload c4 measures a gradient evaluation: 256x256x64 random texture
(the default workload's rules), 3 wavelengths 450/550/650 nm, x
polarisation at normal incidence, 4x4; loss = intensity after a 45 degree
analyser summed over the central 64x64; a step is the recording forward
transfer, the loss seed and the adjoint pass giving dL/dtheta and dL/dphi
of every voxel; value = voxel*wavelength*plane steps of the forward per
second of the whole gradient evaluation.

Prints ONE JSON line.  Options: --method 4x4|2x2, --precision c64|c128.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE, each
# scaled by the counter's reading for a 512 MiB device copy (FETCH x 2.000,
# WRITE x 1.000); profiles/transfer_c64_512x512_hbm_traffic.md.  Keyed by
# (kernel, H, W, planes, method, precision); other workloads report null.
MEASURED_TRAFFIC = {
    ("row_layer", 512, 512, 18, "4x4", "c64"): 312.8e6,
    ("col_kspace", 512, 512, 18, "4x4", "c64"): 303.0e6,
    ("row_layer", 512, 512, 18, "2x2", "c64"): 161.9e6,
    ("col_kspace", 512, 512, 18, "2x2", "c64"): 151.5e6,
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


C5_BEAMS = ([0.1, 0.1, 0.1], [0., 2 * np.pi / 3, 4 * np.pi / 3])
C3_BEAMS = ([0.] + [0.2] * 6, [0.] + [k * np.pi / 3 for k in range(6)])
C3_SIZES = (128, 192, 256, 384, 512)


def build_inputs_c3(nsamples, nlam, seed=0):
    """Samples of the uneven batch: list of (field_data, optical_data, info)."""
    import dtmm_amd as dm
    from dtmm_amd.samples import random_director
    t = time.time()
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nsamples):
        size = int(rng.choice(C3_SIZES))
        layers = int(rng.integers(32, 161))
        kind = str(rng.choice(["droplet", "random"]))
        ne = float(np.sqrt(rng.uniform(1.6 ** 2, 1.8 ** 2)))
        specs.append((size, layers, kind, ne))
    wl = np.linspace(380., 780., nlam)
    beta, phi = C3_BEAMS
    fields, dirs = {}, {}
    for size, _, kind, _ in specs:
        if size not in fields:
            f = dm.illumination_data((size, size), wl, pixelsize=100, beta=beta, phi=phi,
                                     jones=None, window=dm.blackman((size, size)))[0]
            fields[size] = f.astype(np.complex64)
        key = (kind, size)
        if key not in dirs:
            shape = (160, size, size)
            if kind == "random":
                dirs[key] = (random_director(shape, seed=size), None)
            else:
                dirs[key] = (dm.nematic_droplet_director(shape, 30. * size / 96, "r"), None)
    samples = []
    for size, layers, kind, ne in specs:
        director = dirs[(kind, size)][0][:layers]
        if kind == "random":
            optical = [dm.director2data(director, no=1.5, ne=ne)]
        else:
            mask = np.linalg.norm(director, axis=-1) > 0
            optical = [dm.director2data(director, mask=mask, no=1.5, ne=ne, nhost=1.5)]
        samples.append(((fields[size], wl, 100.), optical, (size, layers, kind, ne)))
    log("c3 inputs built in %.1f s: %d samples" % (time.time() - t, len(samples)))
    return samples


def run_c3(args):
    """The uneven batch through dtmm_amd.transfer_batch's engine call."""
    import torch
    from dtmm_amd import _lib
    from dtmm_amd.transfer import _prepare
    lib = _lib.load()
    samples = build_inputs_c3(args.samples, args.wavelengths)
    beta, phi = C3_BEAMS
    t = time.time()
    jobs = []
    steps = 0
    for fd, optical, (size, layers, kind, ne) in samples:
        jobs.append(_prepare(fd, optical, beta=beta, phi=phi, method=args.method,
                             precision=args.precision, nin=1.5, nout=1.5))
        steps += size * size * layers * int(np.prod(fd[0].shape[:-3]))
    torch.cuda.synchronize()
    log("c3 uploaded in %.1f s (%.3g voxel*lambda*plane steps per batch)"
        % (time.time() - t, steps))
    structs = [j.stack.c_struct() for j in jobs]
    arr = (_lib.Job * len(jobs))()
    for a, j, st in zip(arr, jobs, structs):
        a.plan = j.plan._h.value
        a.stack = ctypes.pointer(st)
        a.field_in = j.fin.data_ptr()
        a.field_out = j.fout.data_ptr()
        a.workspace = j.ws.data_ptr() if j.ws is not None else None
        a.workspace_bytes = j.plan.workspace_bytes
    stream = torch.cuda.current_stream()
    status = torch.zeros(4, dtype=torch.int32, device="cuda")

    def step():
        _lib.check(lib.dtmmx_transfer_batch(arr, len(jobs), status.data_ptr(),
                                            stream.cuda_stream))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    for j in jobs[:4]:
        if not torch.isfinite(torch.view_as_real(j.fout)).all():
            log("non-finite output")
            sys.exit(1)
    sizes = sorted(set(smp[2][0] for smp in samples))
    rec = {
        "metric": "voxel*lambda steps/s (transfer_field, 4x4 unless noted), c64 with c128 "
                  "reference, % HBM peak",
        "value": steps / (ms_per_step / 1e3),
        "unit": "voxel*lambda steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 complex, fp64 eigensolve)" if args.precision == "c64" else "c128",
        "data": "synthetic (seeded droplet / random nematic samples, plane-wave illumination)",
        "config": {"workload": "batch of %d uneven samples (grids %s, 32-160 layers), 7 beams x "
                               "2 pol x %d wavelengths, %s, one dtmmx_transfer_batch call"
                               % (len(jobs), sizes, args.wavelengths, args.method),
                   "samples": len(jobs), "planes_per_sample": 14 * args.wavelengths,
                   "method": args.method, "precision": args.precision},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(rec), flush=True)


def run_c4(args):
    """Forward (recording) + adjoint gradient of the analysed-intensity loss."""
    import torch
    import dtmm_amd as dm
    from dtmm_amd.samples import random_director_data
    size = args.size if args.size != 512 else 256
    layers = args.layers if args.layers != 128 else 64
    t = time.time()
    optical = [random_director_data((layers, size, size), seed=0, no=1.5, ne=1.7)]
    wl = np.array([450., 550., 650.])
    fd = dm.illumination_data((size, size), wl, pixelsize=100, beta=0., phi=0., jones=(1, 0))
    dt = np.complex64 if args.precision == "c64" else np.complex128
    fd = (fd[0].astype(dt),) + tuple(fd[1:])
    tr = dm.AdjointTransfer(fd, optical, beta=0., phi=0., precision=args.precision)
    log("c4 inputs in %.1f s, tape %.2f GB" % (time.time() - t, tr.tape_bytes / 1e9))
    roi = (size // 2 - 32, size // 2 + 32, size // 2 - 32, size // 2 + 32)
    # the whole gradient evaluation (recording forward, loss, adjoint) is
    # one HIP graph, replayed per step
    replay = tr.gradient_graph(
        lambda f: dm.analyser_intensity_loss(f, (1, 1), 1.5, roi, as_tensor=True))
    for _ in range(args.warmup):
        replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss_t, grad = replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss = float(loss_t.item())
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    planes = 3
    steps = size * size * layers * planes
    g = grad.cpu().numpy()
    if not np.all(np.isfinite(g)):
        log("non-finite gradient")
        sys.exit(1)
    log("loss %.6g, max |dL/dtheta| %.3e, max |dL/dphi| %.3e"
        % (loss, np.abs(g[0]).max(), np.abs(g[1]).max()))
    rec = {
        "metric": "voxel*lambda steps/s (transfer_field, 4x4 unless noted), c64 with c128 "
                  "reference, % HBM peak",
        "value": steps / (ms_per_step / 1e3),
        "unit": "voxel*lambda steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 complex, fp64 eigensolve)" if args.precision == "c64" else "c128",
        "data": "synthetic (seeded random nematic director, plane-wave illumination)",
        "config": {"workload": "gradient of the 45 degree analysed intensity over the central "
                               "64x64 w.r.t. theta, phi of %dx%dx%d voxels, 3 wavelengths, x "
                               "pol, 4x4 (recording forward + adjoint, one HIP graph)"
                               % (size, size, layers),
                   "height": size, "width": size, "layers": layers, "wavelengths": 3,
                   "planes": planes, "method": "4x4", "precision": args.precision},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(rec), flush=True)


def run_pom(args):
    """The polarizing-microscope step after the transfer on its output shape
    (2 polarisations x 9 wavelengths x 4 x 512 x 512): x polarizer, crossed
    analyser, refocus by 3 pixels in a medium of index 1.5."""
    import torch
    import dtmm_amd as dm
    from dtmm_amd import pom
    H = W = args.size
    nlam = args.wavelengths
    wl = np.linspace(400., 700., nlam)
    ks = dm.k0(wl, 100.)
    dt = torch.complex64 if args.precision == "c64" else torch.complex128
    g = torch.Generator(device="cuda").manual_seed(3)
    field = torch.randn((1, 2, nlam, 4, H, W), dtype=dt, device="cuda", generator=g)
    pmat = pom.jonesmat4x4(pom.polarizer([0., 1.]), 1.5)
    kw = dict(jvec=(1., 0.), pmat=pmat, focus=3., n=1.5, betamax=0.8)
    for _ in range(args.warmup):
        spec = pom.calculate_pom_specter(field, ks, **kw)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        spec = pom.calculate_pom_specter(field, ks, **kw)
    e1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    ev_ms = e0.elapsed_time(e1) / args.steps
    if not torch.isfinite(spec).all():
        log("non-finite specter")
        sys.exit(1)
    units = H * W * nlam
    # algorithmic bytes: both polarised fields read once, the specter written once
    csize = 8 if args.precision == "c64" else 16
    alg = 2 * nlam * 4 * H * W * csize + H * W * nlam * csize // 2
    log("pom step %.3f ms per call (events %.3f ms)" % (ms_per_step, ev_ms))
    cpu = None
    if not args.no_cpu:
        import oracle
        cores = len(os.sched_getaffinity(0))
        oracle.set_fft_workers(cores)
        nl = min(2, nlam)
        fn = field[:, :, :nl].cpu().numpy().astype(np.complex128)
        t = time.time()
        oracle.pom_specter(fn, ks[:nl], **kw)
        tc = time.time() - t
        cpu = {"value": H * W * nl / tc, "unit": "pixel*lambda/s", "cores": cores,
               "kind": "port",
               "sample": "oracle pom_specter (complex128) on %d of %d wavelengths of the same "
                         "field, %.2f s; scipy.fft on %d threads" % (nl, nlam, tc, cores)}
    rec = {
        "metric": "pixel*lambda/s of the polarizing-microscope specter (calculate_pom_specter)",
        "value": units / (ms_per_step / 1e3),
        "unit": "pixel*lambda/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 complex, fp64 operators)" if args.precision == "c64" else "c128",
        "data": "synthetic (seeded random field of the transfer's output shape)",
        "config": {"workload": "POM specter of a %dx%d field, %d wavelengths x 2 pol (x "
                               "polarizer, crossed analyser, focus 3 px in n 1.5)"
                               % (H, W, nlam),
                   "height": H, "width": W, "wavelengths": nlam, "precision": args.precision},
        "roofline": {"bound": "hbm",
                     "kernel": "whole step (3 launches: combination + forward x transforms, "
                               "y transforms with the field diffraction operator, inverse x "
                               "transforms + specter)",
                     "achieved": alg / (ms_per_step / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_launch": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


def run_tmm1d(args):
    """The non-iterative 1D transfer matrix method on the benchmark's material:
    one stack per pixel of the 512 x 512 x 128 random director, one
    wavelength, an x-polarised wave at normal incidence between media 1.0
    and 1.5."""
    import torch
    import dtmm_amd as dm
    from dtmm_amd import _lib
    size, layers = args.size, args.layers
    optical = build_inputs(size, layers, 1)[1]
    d, eps, ang = optical[0]
    n = size * size
    k = float(dm.k0(550., 100.))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")  # noqa: E731
    kd = dev(np.repeat((np.asarray(d, np.float64) * k)[:, None], n, 1))
    er = dev(np.moveaxis(np.asarray(eps, np.float64).reshape(layers, n, 3), -1, 0))
    an = dev(np.moveaxis(np.asarray(ang, np.float64).reshape(layers, n, 3), -1, 0))
    beta, phi = 0., 0.
    f0 = np.zeros(4, np.complex128)
    f0[0], f0[1] = 1., 1.  # x polarised unit wave in air at normal incidence
    fin0 = torch.from_numpy(np.broadcast_to(f0, (n, 4)).copy()).to("cuda")
    fin = torch.empty_like(fin0)
    fout = torch.empty_like(fin0)
    desc = _lib.TmmDesc()
    desc.n_elements, desc.n_layers = n, layers
    desc.kd, desc.eps_re, desc.eps_im, desc.angles = kd.data_ptr(), er.data_ptr(), None, \
        an.data_ptr()
    desc.beta, desc.phi, desc.n_in, desc.n_out = beta, phi, 1.0, 1.5
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        fin.copy_(fin0)
        _lib.check(lib.dtmmx_tmm_transfer4x4(ctypes.byref(desc), ctypes.c_void_p(fin.data_ptr()),
                                             ctypes.c_void_p(fout.data_ptr()), 0, None,
                                             ctypes.c_void_p(stream)))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    if not torch.isfinite(torch.view_as_real(fout)).all():
        log("non-finite output")
        sys.exit(1)
    units = n * layers
    alg = units * 7 * 8 + n * 4 * 16 * 3  # kd, eps, angles per layer; fields in/out
    log("tmm1d %.3f ms per call" % ms)
    cpu = None
    if not args.no_cpu:
        import oracle
        ns = 2048
        fo = oracle.transfer4x4(np.broadcast_to(f0, (ns, 4)), np.asarray(d, np.float64) * k,
                                np.asarray(eps, np.float64).reshape(layers, n, 3)[:, :ns],
                                np.asarray(ang, np.float64).reshape(layers, n, 3)[:, :ns],
                                beta, phi, nin=1.0, nout=1.5)
        t = time.time()
        oracle.transfer4x4(np.broadcast_to(f0, (ns, 4)), np.asarray(d, np.float64) * k,
                           np.asarray(eps, np.float64).reshape(layers, n, 3)[:, :ns],
                           np.asarray(ang, np.float64).reshape(layers, n, 3)[:, :ns],
                           beta, phi, nin=1.0, nout=1.5)
        tc = time.time() - t
        err = float(np.abs(fout[:ns].cpu().numpy() - fo[0]).max() / np.abs(fo[0]).max())
        log("GPU vs oracle on %d stacks: %.2e" % (ns, err))
        cpu = {"value": ns * layers / tc, "unit": "stack*layer/s", "cores": 1, "kind": "port",
               "sample": "oracle transfer4x4 (numpy, complex128) on %d of the %d stacks, "
                         "%.2f s" % (ns, n, tc)}
    rec = {
        "metric": "stack*layer/s of the non-iterative 1D transfer matrix method (transfer4x4)",
        "value": units / (ms / 1e3),
        "unit": "stack*layer/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic (seeded random nematic director, one stack per pixel)",
        "config": {"workload": "%d stacks of %d layers (one per pixel of the %dx%d random "
                               "director), 550 nm, normal incidence, 4x4 with reflections"
                               % (n, layers, size, size),
                   "stacks": n, "layers": layers},
        "roofline": {"bound": "hbm", "kernel": "tmm_kernel (fp64 compute, spilling)",
                     "achieved": alg / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


def run_tmm2d(args):
    """The 2D non-iterative method on the cholesteric grating of the
    reference's 2D example (examples/2D/cholesteric_grating.py): a tilted
    cholesteric helix, 180 pixels (one x pitch) by 72 layers (two z
    pitches), no 1.5 / ne 1.7, 19 wavelengths 380-780 nm, pixel size 10 nm,
    unpolarised normal incidence (2 rays) between media of index 1.5."""
    import torch
    import dtmm_amd as dm
    nlam, ps = 19, float(args.pixelsize)
    pitch_z, pitch_x = 36, 180
    size_x, size_z = pitch_x, 2 * pitch_z
    tilt = np.arctan(pitch_z / pitch_x)
    twist = (np.arange(0, 2 * np.pi * size_z / pitch_z, 2 * np.pi / pitch_z)[:, None] +
             np.arange(0, 2 * np.pi * size_x / pitch_x, 2 * np.pi / pitch_x)[None, :])
    n = np.stack([np.cos(twist), np.sin(twist), np.zeros_like(twist)], -1)
    ry = np.array([[np.cos(tilt), 0, np.sin(tilt)], [0, 1, 0], [-np.sin(tilt), 0, np.cos(tilt)]])
    n = n @ ry.T
    L, W = twist.shape
    epsa = dm.director2angles(n)
    epsv = np.empty((L, W, 3), np.complex128)
    epsv[..., :2] = 1.5 ** 2
    epsv[..., 2] = 1.7 ** 2
    d = np.ones(L)
    wl = np.linspace(380, 780, nlam)
    f3 = dm.illumination_data((W, 1), wl, jones=None, beta=0., phi=0., pixelsize=ps, n=1.5,
                              betamax=0.8)[0]
    f0 = np.ascontiguousarray(f3[..., 0])
    optical = (d, epsv, epsa)
    kw = dict(nin=1.5, nout=1.5)

    def step():
        f = f0.copy()
        return dm.tmm2d.transfer2d((f, wl, ps), optical, **kw)[0]
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    if not np.isfinite(out).all():
        log("non-finite output")
        sys.exit(1)
    ks = dm.k0(wl, ps)
    nmodes = [int(dm.tmm2d.eigenmask1(W, k).sum()) for k in ks]
    units = W * L * nlam
    # per wavelength: the n x W x 4x4 layer blocks, their DFT coupling and the
    # (4n)^3 complex stack product of every layer (8 real flops per MAC)
    flops = sum(L * (8. * (4 * m) ** 3 + 8. * (4 * m) ** 2 * W) for m in nmodes)
    log("tmm2d %.3f ms per call, modes %d..%d" % (ms, min(nmodes), max(nmodes)))
    cpu = None
    if not args.no_cpu:
        import oracle
        f = f0.copy()
        t = time.time()
        ref = oracle.transfer2d((f, wl, ps), optical, **kw)[0]
        tc = time.time() - t
        err = float(np.abs(out - ref).max() / np.abs(ref).max())
        log("GPU vs oracle: %.2e (oracle %.2f s)" % (err, tc))
        cpu = {"value": units / tc, "unit": "pixel*layer*lambda/s", "cores": 1, "kind": "port",
               "sample": "oracle transfer2d (numpy, complex128) on the whole workload, %.2f s"
                         % tc}
    rec = {
        "metric": "pixel*layer*lambda/s of the 2D non-iterative transfer (tmm2d.transfer2d)",
        "value": units / (ms / 1e3),
        "unit": "pixel*layer*lambda/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic (the tilted cholesteric grating of the reference's 2D example)",
        "config": {"workload": "cholesteric grating %d px x %d layers, %d wavelengths 380-780 nm, "
                               "pixel %.0f nm, 2 rays, %d-%d modes, host arrays in and out"
                               % (W, L, nlam, ps, min(nmodes), max(nmodes)),
                   "width": W, "layers": L, "wavelengths": nlam},
        "roofline": {"bound": "fp64", "kernel": "whole call (launch-bound at these mode counts)",
                     "achieved": flops / (ms / 1e3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": flops / (ms / 1e3) / 1e12 / 78.6, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


def build_inputs_c5(size, layers, nlam):
    import dtmm_amd as dm
    from dtmm_amd.samples import droplet_array_data, droplet_grid_centres
    t = time.time()
    optical = [droplet_array_data((layers, size, size), 80. * size / 1024,
                                  droplet_grid_centres(size, 3), no=1.5, ne=1.6, nhost=1.5)]
    wl = np.linspace(450., 650., nlam)
    beta, phi = C5_BEAMS
    fd = dm.illumination_data((size, size), wl, pixelsize=100, beta=beta, phi=phi, jones=None,
                              window=dm.blackman((size, size)), n=1.52)
    log("inputs built in %.1f s: field %s" % (time.time() - t, fd[0].shape))
    return fd, optical


def _layer_times(run_once, nlayers):
    """Per-layer wall times of one oracle transfer through `nlayers` layers,
    from its layer callbacks: the first layer (which also builds the cached
    k-space matrices) is the warm-up, the later nlayers - 1 are the samples."""
    stamps = [time.perf_counter()]
    run_once(lambda j: stamps.append(time.perf_counter()))
    return np.diff(stamps)[1:nlayers]


def _plane_groups(n_rays, nlam, ngroups):
    """(ray, first, last) wavelength ranges splitting the planes into about
    ngroups contiguous groups."""
    per_ray = max(1, min(nlam, ngroups // max(n_rays, 1)))
    groups = []
    for r in range(n_rays):
        edges = np.linspace(0, nlam, per_ray + 1).astype(int)
        groups += [(r, a, b) for a, b in zip(edges[:-1], edges[1:]) if b > a]
    return groups


def cpu_rates(field, wl, ps, opt, nlayers, kw):
    """Oracle rates (voxel*lambda steps/s) on the host.  Primary mode, as the
    reference runs by default: per-pixel math on one thread, scipy.fft on
    every core; the median of the per-layer times after a warm-up layer.
    All-cores mode: the planes split over one thread per core (NumPy's
    per-pixel loops release the GIL), FFTs single-threaded in each; the sum
    of the groups' rates from the median of their per-layer times."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    cores = len(os.sched_getaffinity(0))
    H, W = field.shape[-2:]
    lead = field.shape[:-4]
    n_rays = int(np.prod(lead)) if lead else 1
    nlam = len(wl)
    f = field.reshape((n_rays, nlam, 4, H, W))
    # per-beam beta, phi (beams x polarisations in the field's leading axes)
    # become per-ray values of the flattened ray axis
    beta = np.ravel(kw.get("beta", 0.))
    phi = np.ravel(kw.get("phi", 0.))
    rb, rp = np.repeat(beta, n_rays // beta.size), np.repeat(phi, n_rays // phi.size)
    rest = {k: v for k, v in kw.items() if k not in ("beta", "phi")}

    def transfer(ff, ww, cb, b, p):
        oracle.transfer_field((np.array(ff), ww, ps), opt, layer_callback=cb, beta=b, phi=p,
                              **rest)

    oracle.set_fft_workers(cores)
    b_all, p_all = (rb, rp) if beta.size > 1 else (float(beta[0]), float(phi[0]))
    dt = _layer_times(lambda cb: transfer(f, wl, cb, b_all, p_all), nlayers)
    primary = H * W * n_rays * nlam / float(np.median(dt))
    groups = _plane_groups(n_rays, nlam, cores)
    oracle.set_fft_workers(1)
    try:
        with ThreadPoolExecutor(len(groups)) as ex:
            res = list(ex.map(lambda g: (g, _layer_times(
                lambda cb: transfer(f[g[0], g[1]:g[2]], wl[g[1]:g[2]], cb, float(rb[g[0]]),
                                    float(rp[g[0]])), nlayers)), groups))
    finally:
        oracle.set_fft_workers(cores)
    allc = sum(H * W * (b - a) / float(np.median(t)) for (r, a, b), t in res)
    return primary, allc, dt, cores, len(groups)


def cpu_baseline_c5(fd, optical, crop, nlayers=4):
    """Oracle on a crop x crop window of the same workload (windows, all
    planes), npass = 1, from mid-stack layers."""
    field, wl, ps = fd
    d, eps, ang = optical[0]
    H, W = field.shape[-2:]
    y0, x0 = (H - crop) // 2, (W - crop) // 2
    f = np.ascontiguousarray(field[..., y0:y0 + crop, x0:x0 + crop])
    z0 = len(d) // 2
    opt = [(d[z0:z0 + nlayers], eps[z0:z0 + nlayers, y0:y0 + crop, x0:x0 + crop],
            ang[z0:z0 + nlayers, y0:y0 + crop, x0:x0 + crop])]
    beta, phi = C5_BEAMS
    kw = dict(beta=beta, phi=phi, method="4x4", npass=1, diffraction=5, nin=1.52, nout=1.52)
    primary, allc, dt, cores, ngroups = cpu_rates(f, wl, ps, opt, nlayers, kw)
    planes = int(np.prod(f.shape[:-3]))
    return {"value": primary, "unit": "voxel*lambda steps/s", "cores": cores, "kind": "port",
            "sample": "oracle 4x4 diffraction=5 transfer through %d mid-stack layers of a %dx%d "
                      "crop, all %d planes, npass=1; median of the layer times after the first "
                      "(%s s); per-pixel math 1 thread, scipy.fft on %d threads"
                      % (nlayers, crop, crop, planes, ", ".join("%.2f" % t for t in dt), cores),
            "all_cores_value": allc, "all_cores_threads": ngroups}


def build_inputs(size, layers, nlam, seed=0):
    import dtmm_amd as dm
    from dtmm_amd.samples import random_director_data
    t = time.time()
    optical = [random_director_data((layers, size, size), seed=seed, no=1.5, ne=1.7)]
    wl = np.linspace(380., 780., nlam)
    fd = dm.illumination_data((size, size), wl, pixelsize=100, beta=0., phi=0., jones=None)
    log("inputs built in %.1f s: field %s" % (time.time() - t, fd[0].shape))
    return fd, optical


def cpu_baseline(fd, optical, method, nlayers=4):
    """Oracle (complex128 NumPy restatement) through the first `nlayers`
    layers of the same workload, all planes (cpu_rates)."""
    field, wl, ps = fd
    d, eps, ang = optical[0]
    H, W = field.shape[-2:]
    opt = [(d[:nlayers], eps[:nlayers], ang[:nlayers])]
    primary, allc, dt, cores, ngroups = cpu_rates(field, wl, ps, opt, nlayers,
                                                  dict(beta=0., phi=0., method=method))
    planes = int(np.prod(field.shape[:-3]))
    return {"value": primary, "unit": "voxel*lambda steps/s", "cores": cores, "kind": "port",
            "sample": "oracle %s transfer through the first %d of %d layers, all %d planes of "
                      "%dx%d; median of the layer times after the first (%s s); per-pixel math "
                      "1 thread, scipy.fft on %d threads"
                      % (method, nlayers, len(d), planes, H, W,
                         ", ".join("%.2f" % t for t in dt), cores),
            "all_cores_value": allc, "all_cores_threads": ngroups}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c2",
                    choices=["c2", "c3", "c4", "c5", "pom", "tmm1d", "tmm2d"],
                    help="c2: 512^2 x 128 random texture (default); c3: batch of uneven "
                         "samples; c5: 1024^2 x 200 droplet grid with npass=3, diffraction=5; "
                         "pom: the polarizing-microscope specter of a 512^2 x 9 x 2 field; "
                         "tmm1d: the 1D transfer matrix method, one stack per pixel; "
                         "tmm2d: the 2D method on the reference's cholesteric grating")
    ap.add_argument("--pixelsize", type=float, default=10.,
                    help="tmm2d: pixel size in nm (more modes for larger pixels)")
    ap.add_argument("--samples", type=int, default=64, help="c3: batch size")
    ap.add_argument("--cpu-crop", type=int, default=96, help="c5: oracle crop size")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--method", default="4x4", choices=["4x4", "2x2"])
    ap.add_argument("--precision", default="c64", choices=["c64", "c128"])
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--layers", type=int, default=128)
    ap.add_argument("--wavelengths", type=int, default=9)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-2x2", action="store_true",
                    help="skip the 2x2 half of the default workload")
    ap.add_argument("--no-ref", action="store_true", help="skip the c128 comparison run")
    ap.add_argument("--cpu-layers", type=int, default=4,
                    help="oracle layers of the CPU baseline (the first is the warm-up)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the automatic launch shapes instead of timing the candidates "
                         "(dtmm_amd.tune) before the warm-up")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="time the steps without per-kernel HIP events")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 1048576 // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied %d MiB 5 times" % args.copy_calibration)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has one GPU and the "
              "path shards over independent field planes with no communication, so multi-GPU "
              "runs are not part of this build (got --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        log("no GPU")
        sys.exit(1)
    if args.workload == "c3":
        run_c3(args)
        return
    if args.workload == "c4":
        run_c4(args)
        return
    if args.workload == "pom":
        run_pom(args)
        return
    if args.workload == "tmm2d":
        run_tmm2d(args)
        return
    if args.workload == "tmm1d":
        run_tmm1d(args)
        return
    c5 = args.workload == "c5"
    if c5:
        if args.method != "4x4":
            log("the c5 workload is 4x4 only")
            sys.exit(2)
        size = args.size if args.size != 512 else 1024
        layers = args.layers if args.layers != 128 else 200
        nlam = args.wavelengths if args.wavelengths != 9 else 5
        fd, optical = build_inputs_c5(size, layers, nlam)
    else:
        fd, optical = build_inputs(args.size, args.layers, args.wavelengths)
    rec = measure(args, args.method, fd, optical, c5)
    if not c5 and args.method == "4x4" and not args.no_2x2:
        # the 2x2 half of the same workload (the plain method's other path)
        r2 = measure(args, "2x2", fd, optical, False)
        rec.update({"value_2x2": r2["value"], "ms_per_step_2x2": r2["ms_per_step"],
                    "roofline_2x2": r2["roofline"], "cpu_baseline_2x2": r2["cpu_baseline"]})
        for k in ("c128_value", "c64_over_c128", "rel_field_err_vs_c128",
                  "rel_intensity_err_vs_c128"):
            if k in r2:
                rec[k + "_2x2"] = r2[k]
    print(json.dumps(rec), flush=True)


def measure(args, method, fd, optical, c5):
    """One bench line's measurements of `method` on the workload."""
    import torch
    import dtmm_amd as dm
    from dtmm_amd import _lib
    from dtmm_amd.transfer import Plan, Stack
    lib = _lib.load()
    if c5:
        beams, phis = C5_BEAMS
        npass, diffraction, nin, nout = 3, 5, 1.52, 1.52
    else:
        beams, phis = [0.], [0.]
        npass, diffraction, nin, nout = 1, 1, 1.5, 1.5
    field, wl, ps = fd
    H, W = field.shape[-2:]
    nlam = len(wl)
    n_rays = int(np.prod(field.shape[:-4]))
    rays_per_beam = n_rays // len(beams)
    planes = n_rays * nlam
    ks = dm.k0(wl, ps)

    t = time.time()
    stack = Stack(optical, (H, W), nstep=1, nin=nin, nout=nout)
    L = stack.n_layers
    log("%s: stack uploaded in %.1f s (%d layers)" % (method, time.time() - t, L))

    def make(precision):
        plan = Plan((H, W), ks, beams, phis, n_rays, rays_per_beam, method, precision, 0.8,
                    diffraction=diffraction, npass=npass)
        host = np.ascontiguousarray(field.reshape((n_rays, nlam, 4, H, W)))
        fin = torch.from_numpy(host).to(device="cuda", dtype=plan.dtype).contiguous()
        fout = torch.empty_like(fin)
        ws = torch.empty(max(plan.workspace_bytes, 1), dtype=torch.uint8, device="cuda")
        return plan, fin, fout, ws

    plan, fin, fout, ws = make(args.precision)
    tiles = None
    if not args.no_tune and not c5:
        # launch shapes timed on this workload (the large windowed workload
        # keeps the automatic shapes: one pass takes seconds)
        from dtmm_amd import tune
        t = time.time()
        res = tune.tune_plan(plan, stack, fin, fout, ws, repeats=2, cache=False)
        tiles = {"col_lines": res["tiles"][0], "row_rays": res["tiles"][1],
                 "row_split": res["row_split"]}
        log("%s: tuned launch shapes in %.1f s: %s split %d (ms per pass %s; splits %s)"
            % (method, time.time() - t, res["tiles"], res["row_split"],
               ", ".join("%s %.2f" % (k, v) for k, v in sorted(res["ms"].items())),
               ", ".join("%s %.2f" % (k, v) for k, v in sorted(res["split_ms"].items()))))
    # npass > 1 overwrites the input with the input-side field: every step
    # starts from a fresh copy of the illumination
    fin0 = fin.clone() if npass > 1 else None
    nwin = 1
    if diffraction > 1:
        d = _lib.PlanDesc()
        nwin_c = ctypes.c_int()
        bb = np.asarray(beams, np.float64)
        pp = np.asarray(phis, np.float64)
        d.height, d.width, d.n_wavelengths = H, W, nlam
        d.n_rays, d.rays_per_beam, d.method, d.precision = n_rays, rays_per_beam, 4, 0
        d.betamax, d.k0, d.beta, d.phi = 0.8, _lib.dptr(ks), _lib.dptr(bb), _lib.dptr(pp)
        d.diffraction, d.npass, d.norm, d.smooth = diffraction, 1, 1, 0.1
        d.reflection = 1
        rc = lib.dtmmx_host_windows(ctypes.byref(d), diffraction * diffraction,
                                    ctypes.byref(nwin_c), None, None, None, None)
        _lib.check(rc)
        nwin = nwin_c.value
        log("%d non-empty k-space windows" % nwin)

    def step(p, fi, fo, w):
        if fin0 is not None:
            fi.copy_(fin0 if fi.dtype == fin0.dtype else fin0.to(fi.dtype))
        p.execute(stack, fi, fo, w, stream=stream)

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        step(plan, fin, fout, ws)
    torch.cuda.synchronize()

    # every 16th launch of each per-layer kernel is timed with HIP events on
    # its stream inside the timed region (timing every launch would add about
    # 8% to the pass through the extra stream commands)
    timing = not args.no_kernel_timing
    launches = 2 * nwin * (L + 4) * npass + 16  # per-layer kernel launches of one step
    if timing:
        _lib.check(lib.dtmmx_timing_begin(2 * args.steps * launches // 16 + 64, 16))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(plan, fin, fout, ws)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (ctypes.c_double * 4)()
    cnt = (ctypes.c_int * 4)()
    if timing:
        _lib.check(lib.dtmmx_timing_end(ms, cnt, 4))
    else:
        # per-kernel times from one further instrumented pass
        _lib.check(lib.dtmmx_timing_begin(2 * launches + 64, 1))
        step(plan, fin, fout, ws)
        _lib.check(lib.dtmmx_timing_end(ms, cnt, 4))
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    steps = H * W * L * planes * npass
    value = steps / (ms_per_step / 1e3)
    nc = 4 if method == "4x4" else 2
    cb = 8 if args.precision == "c64" else 16
    out_np = fout.cpu().numpy()
    if not np.all(np.isfinite(out_np)):
        log("non-finite output")
        sys.exit(1)

    # Roofline of the dominant per-layer kernel: algorithmic bytes per launch
    # = field read + write of every plane (2 * nc * cb per pixel per plane),
    # plus the row kernel's material read (eps0..2 + theta, phi as float64,
    # 40 B per pixel).
    field_bytes = 2 * nc * cb * H * W * planes
    # with k-space windows the row kernel also adds its result to the window
    # sum (read + write) except for the first window of a layer, and the
    # column kernel reads the field once and writes one window row per tent
    # row (diffraction of them)
    # The fused window row kernel (one launch per layer) reads the n window
    # rows once and writes the layer's output once; the per-window launches
    # (DTMMX_FUSED_WINDOWS=0) read a window row and read + write the sum
    fused = nwin > 1 and os.environ.get("DTMMX_FUSED_WINDOWS", "1") != "0"
    if fused:
        row_field = field_bytes * (diffraction + 1) / 2.
    elif nwin > 1:
        row_field = field_bytes * (1. + 0.5 * (nwin - 1) / nwin)
    else:
        row_field = field_bytes
    col_field = field_bytes * (1 + diffraction) / 2. if nwin > 1 else field_bytes
    kern = {
        "row_layer": {"ms": ms[0], "n": cnt[0], "bytes": row_field + 40 * H * W},
        "col_kspace": {"ms": ms[1], "n": cnt[1], "bytes": col_field},
    }
    dom = max(kern, key=lambda k: kern[k]["ms"])
    avg_ms = kern[dom]["ms"] / max(kern[dom]["n"], 1)
    achieved = kern[dom]["bytes"] / (avg_ms / 1e3) / 1e9
    log("%s kernels: " % method
        + ", ".join("%s %d x %.3f ms (%.2f of peak)"
                    % (k, v["n"], v["ms"] / max(v["n"], 1),
                       v["bytes"] / max(v["ms"] / max(v["n"], 1), 1e-12) / 1e6 / HBM_PEAK_GBS)
                    for k, v in kern.items())
        + ", kspace %d x %.3f ms" % (cnt[3], ms[3] / max(cnt[3], 1)))
    # launches per pass: row L (or L * windows), column L - 1 (or L window
    # splits), dense k-space 2
    per_pass = [ms[i] / max(cnt[i], 1) * n * npass
                for i, n in enumerate((L * (1 if fused else nwin), (L if nwin > 1 else max(L - 1, 0)),
                                       0, 2))]
    log("per pass (from launch averages): row %.2f ms, col %.2f ms, kspace %.2f ms, rest %.2f ms"
        % (per_pass[0], per_pass[1], per_pass[3], ms_per_step - sum(per_pass)))

    extra = {}
    if not args.no_ref and args.precision == "c64":
        plan2, fin2, fout2, ws2 = make("c128")
        step(plan2, fin2, fout2, ws2)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        step(plan2, fin2, fout2, ws2)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        ref = fout2.cpu().numpy()
        err = float(np.abs(out_np - ref).max() / np.abs(ref).max())
        i64 = dm.field2intensity(out_np.astype(np.complex128))
        iref = dm.field2intensity(ref)
        ierr = float(np.abs(i64 - iref).max() / np.abs(iref).max())
        extra = {"c128_value": steps / (t3 - t2), "c64_over_c128": (t3 - t2) * 1e3 / ms_per_step,
                 "rel_field_err_vs_c128": err, "rel_intensity_err_vs_c128": ierr}
        del plan2, fin2, fout2, ws2
        log("%s: c128 pass %.1f ms, c64 vs c128 rel field err %.2e"
            % (method, (t3 - t2) * 1e3, err))

    cpu = None
    if not args.no_cpu:
        if c5:
            cpu = cpu_baseline_c5(fd, optical, args.cpu_crop, args.cpu_layers)
        else:
            cpu = cpu_baseline(fd, optical, method, args.cpu_layers)
        log("%s cpu baseline: %.3g (all cores %.3g): %s"
            % (method, cpu["value"], cpu["all_cores_value"], cpu["sample"]))
    if c5:
        workload = ("%dx%dx%d 3x3 radial droplet grid, %d wavelengths x 3 beams x 2 pol, 4x4, "
                    "diffraction=5 (%d windows), npass=3, nin=nout=1.52" % (H, W, L, nlam, nwin))
    else:
        workload = ("%dx%dx%d random director, %d wavelengths x 2 pol, %s, diffraction=1, "
                    "npass=1" % (H, W, L, nlam, method))

    rec = {
        "metric": "voxel*lambda steps/s (transfer_field, 4x4 unless noted), c64 with c128 "
                  "reference, % HBM peak",
        "value": value,
        "unit": "voxel*lambda steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex, fp64 eigensolve)" if args.precision == "c64" else "c128",
        "data": "synthetic (seeded random nematic director, plane-wave illumination)",
        "config": {"workload": workload,
                   "height": H, "width": W, "layers": L, "wavelengths": nlam, "planes": planes,
                   "method": method, "precision": args.precision, "npass": npass,
                   "diffraction": diffraction, "launch_shapes": tiles or "automatic"},
        "hbm_frac_per_step": (steps * nwin * 2 * nc * cb / (ms_per_step / 1e3) / 1e9)
        / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((dom, H, W, planes, method,
                                                      args.precision)),
                     "avg_launch_ms": avg_ms, "bytes_per_launch": kern[dom]["bytes"]},
        "cpu_baseline": cpu,
    }
    # the fused FFT-diffraction kernel (forward y transforms, the k-space
    # diffraction operator, inverse y transforms) beside the dominant one
    ck = kern["col_kspace"]
    if ck["n"]:
        c_ms = ck["ms"] / ck["n"]
        c_ach = ck["bytes"] / (c_ms / 1e3) / 1e9
        rec["roofline_fft_diffraction"] = {
            "bound": "hbm", "kernel": "col_kspace", "achieved": c_ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": c_ach / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get(("col_kspace", H, W, planes, method, args.precision)),
            "avg_launch_ms": c_ms, "bytes_per_launch": ck["bytes"]}
    rec.update(extra)
    return rec


if __name__ == "__main__":
    main()
