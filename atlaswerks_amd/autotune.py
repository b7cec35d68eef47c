"""Autotuner for the launch-shape knobs of libawk_hip.so.

Each knob (``awk_tune_set``) selects between kernel variants that compute the
same result with a different amount of work per thread.  ``tune()`` times every
variant on the actual GPU at the workload's grid size, keeps the fastest, and
returns the choice.  ``tune_atlas_batch()`` picks, per scale level of an atlas,
how many subjects each batched launch of the greedy step covers (bit-identical
results for every choice).
"""
import ctypes

import torch

from . import _lib

KNOBS = {
    "apply_vpt": (1, 2, 4),
    "compose_vpt": (1, 2, 4),
    "xinv_unroll": (1, 2, 4),
}
# Knobs left out of the sweep: "xinv_pairs" (the inverse x pass's corner pairs) ties
# on any proxy field here but measures 8 % slower in that pass on the 256^3
# registration (0.210-0.214 against 0.195-0.199 ms at every unroll), where more
# waves reach past the volume's edge; it stays settable through set_knob.


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
            if best is None or ms < best[1] * 0.99:  # a tie keeps the earlier (simpler) value
                best = (v, ms)
        set_knob(name, best[0])
        chosen[name] = best[0]

    sweep("apply_vpt", lambda: hfield.apply(out, img, h))
    sweep("compose_vpt", lambda: hfield.compose(outv, h, h2, hfield.BACKGROUND_STRATEGY_PARTIAL_ID))
    # the greedy iteration on a registration's state: smooth images, an hfield
    # displaced by up to n/32 voxels (8 at 256^3) as a registration's is, so that
    # as many gathers reach past the volume's edge as in the workloads (random
    # images from the identity keep every point inside and favour other variants)
    amp = max(2.0, max(size) / 32.0)
    sm = _smooth_field(shape, 1.0, 3, device)
    I0 = sm[0].contiguous()
    I1 = torch.roll(I0, 2, dims=2).contiguous()
    h0 = ident + _smooth_field(shape, amp, 4, device)
    it = GreedyIteratorGPU(size)
    data = GreedyDeformationDataGPU(I0, I1, max_iters=4 * reps + 8, h0=h0)
    it.UpdateStepSizeNextIteration()
    it.Iterate(data)
    # back-to-back iterations, as GreedyWarp runs them (the composition also forms
    # the next deformed image)
    sweep("xinv_unroll", lambda: it.IterateN(data, 4))
    del it, data
    if verbose:
        print("autotune", chosen, timings)
    return chosen, timings


def tune_atlas_batch(subjects, template_size, levels=((1, 0),), candidates=(0, 1, 2, 4, 8, 16),
                     reps=5, **atlas_kw):
    """Subjects per batched launch for each scale level factor of an atlas.

    For every level a scratch atlas of the subjects (identity start) runs its
    step-size iteration, then each candidate (0: one launch per subject on the
    stream lanes) is timed over ``reps`` hipGraph-replayed iterations; the
    fastest wins.  Returns ({factor: batch}, {factor: {batch: ms}})."""
    from . import atlas
    chosen, timings = {}, {}
    ns = len(subjects)
    for f, _ in levels:
        b = atlas.GreedyAtlasBuilderGPU(subjects, template_size, use_graph=True,
                                        scale_levels=[(f, 1)], **atlas_kw)
        b.Run()  # enters the level, sets the step sizes
        t = {}
        for c in candidates:
            if c > ns:
                continue
            _lib.call("awk_atlas_set_batch", b._h, int(c))
            b.Iterate(1)  # graph capture for this batch size
            t[c] = _events(lambda: b.Iterate(1), reps)
        del b
        best = min(t, key=t.get)
        chosen[f], timings[f] = best, t
    return chosen, timings


def _events(fn, reps):
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


PLAN_KNOBS = {
    "fft_order": (0, 1),
    "x_rt": (0, 2, 4, 6, 8),
    "y_tx": (0, 1, 2, 4, 8, 16),
    "z_tx": (0, 1, 2, 4, 8),
}


def tune_fft(size, reps=10, device="cuda", verbose=False):
    """FFT plan for one grid (nx, ny, nz): the radix order of the generic LDS
    stages and the x / y / z pass tiles, chosen one knob after the other by timing
    the operator L^-1 (DiffOperGPU::ApplyInverseOperator) through graph replays.
    Plans are read when an operator is created, so the chosen knobs stay set for
    the operators made afterwards.  They change results only at FP32 rounding
    level (another butterfly order or kernel).  Returns (chosen, timings)."""
    from .fluid import DiffOperGPU
    nx, ny, nz = size
    f = torch.randn((3, nz, ny, nx), device=device)
    chosen, timings = {}, {}
    op = DiffOperGPU((nx, ny, nz))  # warm-up: the first operator's first-use costs
    _time(lambda: op.ApplyInverseOperator(f), reps)
    del op
    for name, values in PLAN_KNOBS.items():
        timings[name] = {}
        best = None
        for v in values:
            set_knob(name, v)
            op = DiffOperGPU((nx, ny, nz))
            ms = _time(lambda: op.ApplyInverseOperator(f), reps)
            del op
            timings[name][v] = ms
            if best is None or ms < best[1] * 0.99:  # keep the earlier (simpler) value on a tie
                best = (v, ms)
        set_knob(name, best[0])
        chosen[name] = best[0]
    if verbose:
        print("tune_fft", size, chosen, timings)
    return chosen, timings
