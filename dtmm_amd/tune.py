"""Autotuner of the per-layer kernels' launch shapes.

The column kernel transforms ``col_lines`` field lines per work-group and
shares the k-space operator it computes once per work-group among them; the
row kernel takes ``row_rays`` rays per work-group and shares the per-pixel
fp64 solve among them.  More sharing means less arithmetic per line but
fewer, longer work-groups, and which shape is fastest depends on the grid,
the number of planes and the occupancy.  ``tune_plan`` runs the caller's own
transfer once per candidate shape, times it with HIP events and keeps the
fastest in the plan (``dtmmx_plan_set_tiles``).  The launch shape never
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


def tune_plan(plan, stack, field_in, field_out, workspace=None, repeats=2, cache=True):
    """Times every candidate shape on this transfer and sets the fastest.

    ``field_in`` is restored before each run (npass > 1 writes the
    input-side field into it) and left as it was.  Returns
    ``{"tiles": (col_lines, row_rays), "ms": {(c, r): ms}, "cached": bool}``.
    """
    import torch
    key = signature(plan)
    if cache and key in _CACHE:
        set_tiles(plan, *_CACHE[key])
        return {"tiles": _CACHE[key], "ms": {}, "cached": True}
    saved = field_in.clone()
    stream = torch.cuda.current_stream()
    times = {}
    for c, r in candidates(plan):
        set_tiles(plan, c, r)
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
        times[(c, r)] = best
    field_in.copy_(saved)
    tiles = min(times, key=times.get)
    set_tiles(plan, *tiles)
    if cache:
        _CACHE[key] = tiles
    return {"tiles": tiles, "ms": times, "cached": False}
