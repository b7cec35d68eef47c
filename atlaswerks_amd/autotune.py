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
    """Mean GPU time of fn (ms).  The reps calls are captured in one graph and
    replayed, so host launch overhead, which exceeds these kernels' own time
    when they are launched one by one from Python, does not enter; eager timing
    if capture fails."""
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    try:
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(reps):
                fn()
        torch.cuda.current_stream().wait_stream(s)
        g.replay()
        torch.cuda.synchronize()
        a.record()
        g.replay()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps
    except Exception:
        torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _smooth_field(shape, amp, seed, device):
    """[3, nz, ny, nx] sum of 3 random sinusoidal modes per component, max ~amp."""
    import math
    g = torch.Generator().manual_seed(seed)
    nz, ny, nx = shape
    z = torch.arange(nz, device=device, dtype=torch.float32).view(nz, 1, 1)
    y = torch.arange(ny, device=device, dtype=torch.float32).view(1, ny, 1)
    x = torch.arange(nx, device=device, dtype=torch.float32).view(1, 1, nx)
    out = torch.zeros((3,) + tuple(shape), device=device)
    for c in range(3):
        for _ in range(3):
            k = torch.randint(1, 4, (3,), generator=g).tolist()
            ph = float(torch.rand(1, generator=g)) * 2 * math.pi
            out[c] += torch.sin(2 * math.pi * (k[0] * x / nx + k[1] * y / ny + k[2] * z / nz) + ph)
    return out * (amp / 3.0)


def tune(size=(256, 256, 256), reps=10, device="cuda", verbose=False):
    """Time every variant of every knob at grid `size` (nx, ny, nz); keep the fastest."""
    from . import hfield
    from .greedy import GreedyDeformationDataGPU, GreedyIteratorGPU

    nx, ny, nz = size
    shape = (nz, ny, nx)
    g = torch.Generator(device=device).manual_seed(0)
    ident = torch.empty((3,) + shape, device=device)
    hfield.setToIdentity(ident)
    # smooth displacements (a few sinusoidal modes, up to ~4 voxels), the access
    # pattern of registration fields; uncorrelated per-voxel noise would favour
    # other variants than the workloads do
    h = ident + _smooth_field(shape, 4.0, 1, device)
    h2 = ident + _smooth_field(shape, 4.0, 2, device)
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
