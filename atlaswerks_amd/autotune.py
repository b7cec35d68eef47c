"""Autotuner for the launch-shape knobs of libawk_hip.so.

Each knob (``awk_tune_set``) selects between kernel variants that compute the
same result with a different amount of work per thread.  ``tune()`` times every
variant on the actual GPU at the workload's grid size, keeps the fastest, and
returns the choice.  FFT radix schedules are fixed per axis length by the
library's planner; the knobs here are the gather kernels' memory-level
parallelism.
"""
import ctypes

import torch

from . import _lib

KNOBS = {
    "apply_vpt": (1, 2, 4),
    "compose_vpt": (1, 2, 4),
    "xinv_unroll": (1, 2, 4),
}


def set_knob(name, value):
    _lib.call("awk_tune_set", name.encode(), int(value))


def get_knob(name):
    v = ctypes.c_int()
    _lib.call("awk_tune_get", name.encode(), ctypes.byref(v))
    return v.value


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune(size=(256, 256, 256), reps=10, device="cuda", verbose=False):
    """Time every variant of every knob at grid `size` (nx, ny, nz); keep the fastest."""
    from . import hfield
    from .greedy import GreedyDeformationDataGPU, GreedyIteratorGPU

    nx, ny, nz = size
    shape = (nz, ny, nx)
    g = torch.Generator(device=device).manual_seed(0)
    ident = torch.empty((3,) + shape, device=device)
    hfield.setToIdentity(ident)
    h = ident + 3.0 * torch.rand((3,) + shape, device=device, generator=g)
    h2 = ident + 3.0 * torch.rand((3,) + shape, device=device, generator=g)
    img = torch.rand(shape, device=device, generator=g)
    out = torch.empty_like(img)
    outv = torch.empty_like(h)
    chosen, timings = {}, {}

    def sweep(name, fn):
        best = None
        timings[name] = {}
        for v in KNOBS[name]:
            set_knob(name, v)
            ms = _time(fn, reps)
            timings[name][v] = ms
            if best is None or ms < best[1]:
                best = (v, ms)
        set_knob(name, best[0])
        chosen[name] = best[0]

    sweep("apply_vpt", lambda: hfield.apply(out, img, h))
    sweep("compose_vpt", lambda: hfield.compose(outv, h, h2, hfield.BACKGROUND_STRATEGY_PARTIAL_ID))
    it = GreedyIteratorGPU(size)
    data = GreedyDeformationDataGPU(img, torch.roll(img, 2, dims=2), max_iters=4 * reps + 8)
    it.UpdateStepSizeNextIteration()
    it.Iterate(data)
    sweep("xinv_unroll", lambda: it.Iterate(data))
    del it, data
    if verbose:
        print("autotune", chosen, timings)
    return chosen, timings
