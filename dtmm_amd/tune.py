"""Autotuner of the per-layer kernels' launch shapes.

The column kernel transforms ``col_lines`` field lines per work-group and
shares the k-space operator it computes once per work-group among them; the
row kernel takes ``row_rays`` rays per work-group and shares the per-pixel
fp64 solve among them.  More sharing means less arithmetic per line but
fewer, longer work-groups, and which shape is fastest depends on the grid,
the number of planes and the occupancy.  The row kernels can also split
each row's wavelengths over ``row_split`` work-groups (more work-groups in
flight, one more per-pixel solve per chunk).  ``tune_plan`` runs the
caller's own transfer once per candidate shape, then once per split with
the fastest shape, times each with HIP events and keeps the fastest in the
plan (``dtmmx_plan_set_tiles``, ``dtmmx_plan_set_row_split``).  The launch shape never
changes the result.  Winners are cached per workload signature for the
life of the process.
"""
import ctypes

from . import _lib

_CACHE = {}


def signature(plan):
    """What the launch-shape choice depends on."""
    return (plan.shape, len(plan.ks), plan.n_rays, plan.rays_per_beam, plan.method,
            plan.precision, plan.diffraction, plan.npass, plan.reflection, plan.norm,
            plan.betamax)


def candidates(plan):
    """Valid (col_lines, row_rays) pairs of a plan."""
    nc = 4 if plan.method == "4x4" else 2
    pair = plan.n_rays % 2 == 0 and plan.rays_per_beam % 2 == 0
    cols = [c for c in (1, 2, 4) if nc % c == 0] + ([2 * nc] if pair else [])
    rows = [1, 2] if (plan.precision == "c64" and plan.shape[1] <= 512 and pair) else [1]
    return [(c, r) for c in cols for r in rows]


def get_tiles(plan):
    c, r = ctypes.c_int(), ctypes.c_int()
    _lib.check(plan._lib.dtmmx_plan_get_tiles(plan._h, ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


def set_tiles(plan, col_lines, row_rays):
    _lib.check(plan._lib.dtmmx_plan_set_tiles(plan._h, int(col_lines), int(row_rays)))


def get_row_split(plan):
    v = ctypes.c_int()
    _lib.check(plan._lib.dtmmx_plan_get_row_split(plan._h, ctypes.byref(v)))
    return v.value


def set_row_split(plan, row_split):
    _lib.check(plan._lib.dtmmx_plan_set_row_split(plan._h, int(row_split)))


def split_candidates(plan):
    """Explicit wavelength splits of the row kernels tried after the shape
    (0, the automatic choice, is always a candidate)."""
    return [s for s in (1, 2, 4) if s <= len(plan.ks)]


def tune_plan(plan, stack, field_in, field_out, workspace=None, repeats=2, cache=True):
    """Times every candidate shape on this transfer and sets the fastest.

    ``field_in`` is restored before each run (npass > 1 writes the
    input-side field into it) and left as it was.  Returns
    ``{"tiles": (col_lines, row_rays), "row_split": n, "ms": {(c, r): ms},
    "split_ms": {n: ms}, "cached": bool}`` (``split_ms`` only when tuned).
    """
    import torch
    key = signature(plan)
    if cache and key in _CACHE:
        tiles, split = _CACHE[key]
        set_tiles(plan, *tiles)
        set_row_split(plan, split)
        return {"tiles": tiles, "row_split": split, "ms": {}, "cached": True}
    saved = field_in.clone()
    stream = torch.cuda.current_stream()

    def run():
        field_in.copy_(saved)
        plan.execute(stack, field_in, field_out, workspace=workspace)  # warm-up
        best = None
        for _ in range(max(1, repeats)):
            field_in.copy_(saved)
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record(stream)
            plan.execute(stack, field_in, field_out, workspace=workspace)
            t1.record(stream)
            t1.synchronize()
            ms = t0.elapsed_time(t1)
            best = ms if best is None else min(best, ms)
        return best

    # the shapes with the automatic wavelength split, then the explicit
    # splits with the fastest shape
    times = {}
    set_row_split(plan, 0)
    for c, r in candidates(plan):
        set_tiles(plan, c, r)
        times[(c, r)] = run()
    tiles = min(times, key=times.get)
    set_tiles(plan, *tiles)
    split_ms = {0: times[tiles]}
    for sp in split_candidates(plan):
        set_row_split(plan, sp)
        split_ms[sp] = run()
    split = min(split_ms, key=split_ms.get)
    set_row_split(plan, split)
    field_in.copy_(saved)
    if cache:
        _CACHE[key] = (tiles, split)
    return {"tiles": tiles, "row_split": split, "ms": times, "split_ms": split_ms,
            "cached": False}
