"""Greedy-fluid iteration on the GPU vs the oracle and the reference's golden data.

Bars: per-iteration SSD within 0.1 % of the oracle; final hfield close in
relative L2; the reference's own FluidWarp fixtures (FluidWarpRoutineTests.cxx,
Sum |d|^2 < 10) reproduced far inside their tolerance."""
import os

import numpy as np
import pytest

import oracle
from atlaswerks_amd import synth
from atlaswerks_amd.metaimage import read_metaimage

pytestmark = pytest.mark.gpu


def _config0(shape=(64, 64, 64), seed=0):
    t, h, u = synth.pair(shape, seed=seed, max_disp=3.0)
    target = oracle.apply(t, np.moveaxis(h, 0, -1), dtype=np.float32)
    return t, target


@pytest.mark.parametrize("shape,n_iter", [((32, 32, 32), 20), ((64, 64, 64), 50), ((24, 40, 36), 30),
                                          ((16, 32, 256), 10), ((8, 24, 128), 10),
                                          ((12, 20, 67), 8), ((181, 14, 12), 6)])
def test_greedy_energy_trace(shape, n_iter):
    import torch
    from atlaswerks_amd.greedy import GreedyIteratorParam, GreedyWarp
    from gpu_util import aos, host, rel_l2, to_dev
    print("greedy", shape, n_iter, flush=True)
    I0, I1 = _config0(shape)
    sp = (1.0, 1.0, 1.0)
    h_ref, _, ef, ed, step = oracle.greedy_run(I0, I1, n_iter, sp=sp, alpha=0.01, beta=0.01,
                                               gamma=0.001, max_pert=0.1)
    warp = GreedyWarp(to_dev(I0), to_dev(I1), n_iter, sp,
                      GreedyIteratorParam(max_pert=0.1, alpha=0.01, beta=0.01, gamma=0.001))
    data = warp.RunWarp()
    e = data.energies()
    rel = np.abs(e - ed) / ed
    assert rel.max() <= 1e-3, rel.max()
    assert abs(warp.iterator.StepSize() - step) <= 1e-4 * step
    assert rel_l2(host(data.Def1To0), np.moveaxis(h_ref, -1, 0)) <= 1e-5


def test_greedy_inverse_and_inplace():
    """Def0To1 update (composeVHInv, PARTIAL_ZERO) and the in-place C-ABI path."""
    import ctypes
    import torch
    from atlaswerks_amd import _lib
    from atlaswerks_amd.device import planes_arr, ptr
    from atlaswerks_amd.greedy import GreedyIteratorParam, GreedyWarp
    from gpu_util import host, rel_l2, to_dev
    shape = (32, 32, 32)
    I0, I1 = _config0(shape, seed=3)
    h_ref, hinv_ref, _, ed, _ = oracle.greedy_run(I0, I1, 10, compute_inverse=True)
    warp = GreedyWarp(to_dev(I0), to_dev(I1), 10, compute_inverse=True)
    data = warp.RunWarp()
    assert rel_l2(host(data.Def1To0), np.moveaxis(h_ref, -1, 0)) <= 1e-5
    assert rel_l2(host(data.Def0To1), np.moveaxis(hinv_ref, -1, 0)) <= 1e-5
    # in place (h_out == h_in) through the raw C-ABI
    g = ctypes.c_void_p()
    _lib.call("awk_greedy_create", ctypes.byref(g), 32, 32, 32, 1.0, 1.0, 1.0, 0.01, 0.01, 0.001, 0.1)
    h = torch.from_numpy(synth.identity(shape)).cuda()
    hi = torch.from_numpy(synth.identity(shape)).cuda()
    e = torch.zeros(10, dtype=torch.float64, device="cuda")
    d0, d1 = to_dev(I0), to_dev(I1)
    for it in range(10):
        if it == 0:
            _lib.call("awk_greedy_update_step_next", g)
        _lib.call("awk_greedy_iterate", g, ptr(d0), ptr(d1), planes_arr(h), planes_arr(h),
                  planes_arr(hi), planes_arr(hi), ctypes.c_void_p(e[it:].data_ptr()), None)
    torch.cuda.synchronize()
    _lib.load().awk_greedy_destroy(g)
    assert rel_l2(host(h), np.moveaxis(h_ref, -1, 0)) <= 1e-5
    assert rel_l2(host(hi), np.moveaxis(hinv_ref, -1, 0)) <= 1e-5
    np.testing.assert_allclose(host(e), ed, rtol=1e-3)


def test_fluidwarp_fixture(golden_dir):
    """The reference's FluidWarp::computeHFieldAsymmetric baseline (32^3 bullseyes,
    alpha=beta=0.01, gamma=1e-4, MaxPert 0.5, 20 iterations, delta computed once)
    reproduced by the GPU greedy iteration: FluidWarpRoutineTests.cxx:178-218."""
    from atlaswerks_amd.greedy import GreedyIteratorParam, GreedyWarp
    from atlaswerks_amd import hfield
    from gpu_util import aos, host, to_dev
    import torch
    g = lambda n: read_metaimage(os.path.join(golden_dir, n)).data
    fixed, moving = g("BullseyeTest00.mha"), g("BullseyeTest01.mha")
    warp = GreedyWarp(to_dev(moving), to_dev(fixed), 20, (1.0, 1.0, 1.0),
                      GreedyIteratorParam(max_pert=0.5, alpha=0.01, beta=0.01, gamma=1e-4),
                      compute_inverse=True)
    data = warp.RunWarp()
    h = aos(host(data.Def1To0))
    hinv = aos(host(data.Def0To1))
    d_h = ((h - g("ComputeHFieldAsymmetricDeformation.mha")) ** 2).sum()
    d_hi = ((hinv - g("ComputeHFieldAsymmetricDeformationInv.mha")) ** 2).sum()
    out = torch.empty(moving.shape, dtype=torch.float32, device="cuda")
    hfield.apply(out, to_dev(moving), data.Def1To0)
    d_img = ((host(out) - g("ComputeHFieldAsymmetricDeformedImage.mha")) ** 2).sum()
    print("fixture sq-diff", d_h, d_hi, d_img)
    # reference tolerance is 10 (FluidWarpRoutineTests.cxx:28); hold the GPU to 1e-3
    assert d_h < 1e-3 and d_hi < 1e-3 and d_img < 1e-3


def test_inverse_x_pass_launch_variants_bit_identical():
    """The inverse x pass's launch variants (1, 2 or 4 voxels per epilogue step;
    gathers one corner at a time or in x-neighbour pairs) give the same hfields
    and energies bit for bit over 6 iterations (256-long rows, the specialised
    pass)."""
    import torch
    from atlaswerks_amd import autotune
    from atlaswerks_amd.greedy import GreedyIteratorParam, GreedyWarp
    from gpu_util import to_dev
    t, target = _config0((8, 16, 256), seed=3)
    print("xinv variants", flush=True)
    saved = {k: autotune.get_knob(k) for k in ("xinv_unroll", "xinv_pairs")}
    runs = []
    try:
        for xu in (1, 2, 4):
            for pr in (0, 1):
                autotune.set_knob("xinv_unroll", xu)
                autotune.set_knob("xinv_pairs", pr)
                w = GreedyWarp(to_dev(t), to_dev(target), 6, param=GreedyIteratorParam(max_pert=0.1))
                d = w.RunWarp()
                torch.cuda.synchronize()
                runs.append((d.Def1To0.cpu().numpy(), d.energies()))
    finally:
        for k, v in saved.items():
            autotune.set_knob(k, v)
    for h, e in runs[1:]:
        np.testing.assert_array_equal(h, runs[0][0])
        np.testing.assert_array_equal(e, runs[0][1])


@pytest.mark.parametrize("shape", [(32, 32, 32), (8, 16, 256), (8, 24, 128), (12, 20, 67), (17, 9, 5)])
def test_iterate_n_equals_single_iterations(shape):
    """awk_greedy_iterate_n (each composition also forms the next deformed image)
    gives the fields and energies of the same number of awk_greedy_iterate calls
    bit for bit: an odd and an even count, called after the iteration that sets
    the step or with that iteration inside the call, and calls of one iteration; at 256-long rows for every launch
    variant of the inverse x pass (unroll, corner pairs)."""
    import torch
    from atlaswerks_amd import autotune
    from atlaswerks_amd.greedy import GreedyDeformationDataGPU, GreedyIteratorGPU, GreedyIteratorParam
    from gpu_util import to_dev
    print("iterate_n", shape, flush=True)
    t, target = _config0(shape, seed=5)
    nz, ny, nx = shape
    variants = [(1, 0), (2, 0), (4, 0), (1, 1), (2, 1), (4, 1)] if nx == 256 else [None]
    saved = {k: autotune.get_knob(k) for k in ("xinv_unroll", "xinv_pairs")}

    def run(schedule):
        it = GreedyIteratorGPU((nx, ny, nz), (1.0, 1.0, 1.0), GreedyIteratorParam(max_pert=0.1))
        d = GreedyDeformationDataGPU(to_dev(t), to_dev(target), max_iters=16)
        it.UpdateStepSizeNextIteration()
        for kind, k in schedule:
            if kind == "one":
                for _ in range(k):
                    it.Iterate(d)
            else:
                it.IterateN(d, k)
        torch.cuda.synchronize()
        return d.Def1To0.cpu().numpy(), d.energies(), d.cur_iter

    try:
        for var in variants:
            if var:
                autotune.set_knob("xinv_unroll", var[0])
                autotune.set_knob("xinv_pairs", var[1])
            for n in (7, 4):
                h0, e0, c0 = run([("one", n + 1)])
                for sched in ([("one", 1), ("n", n)], [("n", n + 1)], [("n", 1), ("n", 1), ("n", n - 1)]):
                    h, e, c = run(sched)
                    assert c == c0 == n + 1
                    np.testing.assert_array_equal(h, h0)
                    np.testing.assert_array_equal(e, e0)
    finally:
        for k, v in saved.items():
            autotune.set_knob(k, v)


@pytest.mark.parametrize("shape", [(32, 32, 32), (16, 32, 256), (32, 64, 128), (64, 64, 64)])
def test_composition_band_order_bit_identical(shape):
    """The inverse and forward x passes taking their row groups band by band
    (knobs xinv_band, xfwd_band: 2, 4, 8 adjacent row groups per band) give the plane-order
    results bit for bit: GreedyWarp (one iteration, then the rest back to back,
    each composition also forming the next deformed image) over 8 iterations,
    and the per-iteration entry point over 3."""
    import torch
    from atlaswerks_amd import autotune
    from atlaswerks_amd.greedy import (GreedyDeformationDataGPU, GreedyIteratorGPU, GreedyIteratorParam,
                                       GreedyWarp)
    from gpu_util import to_dev
    print("xinv band", shape, flush=True)
    t, target = _config0(shape, seed=7)
    nz, ny, nx = shape
    saved = {k: autotune.get_knob(k) for k in ("xinv_band", "xfwd_band", "yfwd_rev")}

    def run(band):
        autotune.set_knob("xinv_band", band)
        autotune.set_knob("xfwd_band", band)
        autotune.set_knob("yfwd_rev", 1 if band else 0)
        w = GreedyWarp(to_dev(t), to_dev(target), 8, param=GreedyIteratorParam(max_pert=0.1))
        d = w.RunWarp()
        it = GreedyIteratorGPU((nx, ny, nz), (1.0, 1.0, 1.0), GreedyIteratorParam(max_pert=0.1))
        d2 = GreedyDeformationDataGPU(to_dev(t), to_dev(target), max_iters=3)
        it.UpdateStepSizeNextIteration()
        for _ in range(3):
            it.Iterate(d2)
        torch.cuda.synchronize()
        return d.Def1To0.cpu().numpy(), d.energies(), d2.Def1To0.cpu().numpy(), d2.energies()

    try:
        ref = run(0)
        for band in (2, 4, 8):
            got = run(band)
            for a, c in zip(got, ref):
                np.testing.assert_array_equal(a, c)
    finally:
        for k, v in saved.items():
            autotune.set_knob(k, v)


@pytest.mark.parametrize("shape", [(16, 32, 256), (8, 24, 128), (32, 64, 128), (64, 64, 64)])
@pytest.mark.parametrize("max_pert", [0.1, 1.5])
def test_composition_staged_rows_bit_identical(shape, max_pert):
    """The composing inverse x pass reading the field from rows staged in LDS
    (knob xinv_stage) against its gathers from memory: GreedyWarp over 8
    iterations (back to back, each composition also forming the next deformed
    image) and 3 per-iteration calls, bit for bit.  max_pert 1.5 moves many points
    more than a voxel, outside the staged rows, where the staged pass gathers from
    memory as well."""
    import torch
    from atlaswerks_amd import autotune
    from atlaswerks_amd.greedy import (GreedyDeformationDataGPU, GreedyIteratorGPU, GreedyIteratorParam,
                                       GreedyWarp)
    from gpu_util import to_dev
    print("xinv stage", shape, max_pert, flush=True)
    t, target = _config0(shape, seed=9)
    nz, ny, nx = shape
    saved = autotune.get_knob("xinv_stage")

    def run(stage):
        autotune.set_knob("xinv_stage", stage)
        w = GreedyWarp(to_dev(t), to_dev(target), 8, param=GreedyIteratorParam(max_pert=max_pert))
        d = w.RunWarp()
        it = GreedyIteratorGPU((nx, ny, nz), (1.0, 1.0, 1.0), GreedyIteratorParam(max_pert=max_pert))
        d2 = GreedyDeformationDataGPU(to_dev(t), to_dev(target), max_iters=3)
        it.UpdateStepSizeNextIteration()
        for _ in range(3):
            it.Iterate(d2)
        torch.cuda.synchronize()
        return d.Def1To0.cpu().numpy(), d.energies(), d2.Def1To0.cpu().numpy(), d2.energies()

    try:
        ref = run(0)
        got = run(1)
        for a, c in zip(got, ref):
            np.testing.assert_array_equal(a, c)
    finally:
        autotune.set_knob("xinv_stage", saved)


def test_iterate_n_rejects_overlapping_fields():
    import ctypes
    import torch
    from atlaswerks_amd import _lib
    from atlaswerks_amd.device import planes_arr, ptr
    from atlaswerks_amd.greedy import GreedyIteratorGPU
    it = GreedyIteratorGPU((8, 8, 8))
    I = torch.zeros((8, 8, 8), dtype=torch.float32, device="cuda")
    h = torch.zeros((3, 8, 8, 8), dtype=torch.float32, device="cuda")
    rc = _lib.load().awk_greedy_iterate_n(it._h, ptr(I), ptr(I), planes_arr(h), planes_arr(h), 2,
                                          ctypes.c_void_p(0), None)
    assert rc != 0
