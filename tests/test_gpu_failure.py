"""Failure detection on the GPU: AutoStepReduce, NaN energies, invalid steps.

The reference checks every iteration on the host: GreedyDeformationData::AddEnergy
logs a NaN energy and records an EnergyIncreaseEvent on an increase, halving the
step under AutoStepReduce (GreedyDeformationData.cxx:162-184); the LDMM iterator
throws on a NaN energy or max displacement (LDMMIteratorGPU.cxx:238-241, 283-285,
439-441).  Here the kernels raise sticky device flags and awk_*_get_flags reports
them (AWK_ERR_NUMERIC for NaN / invalid steps).  Every case runs ordinary kernels
on NaN data: all position-dependent loads are range-checked buffer loads, so no
kernel reads outside its buffers whatever the coordinates."""
import ctypes
import warnings

import numpy as np
import pytest

import oracle
from atlaswerks_amd import synth

pytestmark = pytest.mark.gpu


def _pair(shape, seed, disp=2.0):
    t, h, _ = synth.pair(shape, seed=seed, max_disp=disp)
    return t, oracle.apply(t, np.moveaxis(h, 0, -1), dtype=np.float32)


def test_auto_step_reduce_matches_oracle():
    # GPU twin of test_oracle_misc.test_auto_step_reduce_halves_on_increase
    from atlaswerks_amd import _lib
    from atlaswerks_amd.greedy import GreedyIteratorParam, GreedyWarp
    from gpu_util import to_dev
    print("AutoStepReduce 16^3, MaxPert 3, 12 iterations", flush=True)
    I0, I1 = _pair((16, 16, 16), seed=2)
    _, _, e_ref, _, step_ref = oracle.greedy_run(I0, I1, 12, max_pert=3.0, auto_step_reduce=True)
    _, _, _, _, step_plain = oracle.greedy_run(I0, I1, 12, max_pert=3.0)
    w = GreedyWarp(to_dev(I0), to_dev(I1), 12, param=GreedyIteratorParam(max_pert=3.0),
                   auto_step_reduce=True)
    data = w.RunWarp()
    e = data.energies()
    assert np.max(np.abs(e - e_ref) / e_ref) <= 1e-3
    inc = np.diff(e) > 0
    np.testing.assert_array_equal(inc, np.diff(e_ref) > 0)
    n_inc = int(inc.sum())
    assert n_inc >= 1
    step0 = w.step_history[0]
    assert abs(step0 - step_plain) <= 1e-5 * step_plain
    # halving is exact in float: the final step is the first one over 2^increases
    assert w.iterator.StepSize() == np.float32(step0) / np.float32(2 ** n_inc)
    assert abs(w.iterator.StepSize() - step_ref) <= 1e-5 * step_ref
    assert w.flags & _lib.FLAG_ENERGY_INCREASE
    assert not w.flags & (_lib.FLAG_NAN_ENERGY | _lib.FLAG_BAD_STEP)
    # the EnergyHistory: one EnergyIncrease after each rising iteration, followed by
    # the halved StepSizeChange; the first iteration's StepSizeChange
    ev = w.EnergyHistory().events
    kinds = [x[0] for x in ev]
    assert kinds.count("iter") == 12 and kinds.count("increase") == n_inc
    assert kinds.count("step") == 1 + n_inc
    steps = [x[1] for x in ev if x[0] == "step"]
    assert steps[0] == pytest.approx(step0) and steps[-1] == pytest.approx(step0 / 2 ** n_inc)


def test_plain_run_raises_no_flags():
    from atlaswerks_amd.greedy import GreedyWarp
    from gpu_util import to_dev
    I0, I1 = _pair((16, 16, 16), seed=3)
    w = GreedyWarp(to_dev(I0), to_dev(I1), 8)
    w.RunWarp()
    assert w.flags == 0
    assert "increase" not in [x[0] for x in w.EnergyHistory().events]


def test_greedy_nan_image_sets_flag_and_status():
    import torch
    from atlaswerks_amd import _lib
    from atlaswerks_amd.device import stream_handle
    from atlaswerks_amd.greedy import GreedyDeformationDataGPU, GreedyIteratorGPU, GreedyWarp
    from gpu_util import to_dev
    print("greedy with a NaN voxel in I1", flush=True)
    I0, I1 = _pair((16, 16, 16), seed=4)
    I1 = I1.copy()
    I1[5, 6, 7] = np.nan
    it = GreedyIteratorGPU((16, 16, 16))
    data = GreedyDeformationDataGPU(to_dev(I0), to_dev(I1))
    it.UpdateStepSizeNextIteration()
    it.Iterate(data)
    # through the C-ABI: the status is AWK_ERR_NUMERIC and the flags name the NaN
    f = ctypes.c_uint(0)
    code = _lib.load().awk_greedy_get_flags(it._h, ctypes.byref(f), 1, stream_handle(None))
    assert code == _lib.AWK_ERR_NUMERIC
    assert f.value & _lib.FLAG_NAN_ENERGY
    assert "NaN" in _lib.load().awk_last_error().decode()
    # cleared: nothing new since
    code = _lib.load().awk_greedy_get_flags(it._h, ctypes.byref(f), 1, stream_handle(None))
    assert code == _lib.AWK_OK and f.value == 0
    assert np.isnan(data.energies()[0])
    # the driver warns, as the reference logs, and keeps the flag
    w = GreedyWarp(to_dev(I0), to_dev(I1), 3)
    with pytest.warns(RuntimeWarning, match="NaN"):
        w.RunWarp()
    assert w.flags & _lib.FLAG_NAN_ENERGY
    torch.cuda.synchronize()


def test_greedy_identical_images_flag_invalid_step():
    # I0 = I1: the force and the velocity are exactly zero, max|v| = 0 and the step
    # MaxPert / 0 is infinite (the reference divides the same way)
    from atlaswerks_amd import _lib
    from atlaswerks_amd.greedy import GreedyDeformationDataGPU, GreedyIteratorGPU
    from gpu_util import to_dev
    I0 = synth.blobs((16, 16, 16), seed=5)
    it = GreedyIteratorGPU((16, 16, 16))
    data = GreedyDeformationDataGPU(to_dev(I0), to_dev(I0))
    it.UpdateStepSizeNextIteration()
    it.Iterate(data)
    flags, code = _lib.get_flags("awk_greedy_get_flags", it._h)
    assert code == _lib.AWK_ERR_NUMERIC and flags & _lib.FLAG_BAD_STEP
    assert np.isinf(it.StepSize())


@pytest.mark.parametrize("orig", [True, False])
def test_ldmm_nan_raises(orig):
    from atlaswerks_amd import _lib, ldmm
    from gpu_util import to_dev
    print("LDMM with a NaN voxel in I1, IterateOrig" if orig else "IterateNew", flush=True)
    I0, I1 = _pair((16, 20, 24), seed=6)
    I1 = I1.copy()
    I1[3, 4, 5] = np.nan
    w = ldmm.LDMMWarp(to_dev(I0), to_dev(I1), n_time_steps=3, n_iter=2,
                      param=ldmm.LDMMIteratorParam(sigma=0.1, max_pert=0.2, use_orig_gradient=orig))
    with pytest.raises(_lib.NumericalError):
        w.RunWarp()
    # a clean pair raises nothing
    I0, I1 = _pair((16, 20, 24), seed=6)
    w = ldmm.LDMMWarp(to_dev(I0), to_dev(I1), n_time_steps=3, n_iter=2,
                      param=ldmm.LDMMIteratorParam(sigma=0.1, max_pert=0.2, use_orig_gradient=orig))
    w.RunWarp()
    assert w.iterator.GetFlags() & (_lib.FLAG_NAN_ENERGY | _lib.FLAG_BAD_STEP) == 0


def test_atlases_report_nan():
    from atlaswerks_amd import _lib, atlas, ldmm
    from gpu_util import to_dev
    print("atlases with a NaN voxel in one subject", flush=True)
    shape = (16, 16, 16)
    t = synth.blobs(shape, seed=7)
    ims = []
    for i in range(3):
        u = np.moveaxis(synth.identity(shape) + synth.sinusoidal_displacement(shape, 1.5, seed=70 + i), 0, -1)
        ims.append(oracle.apply(t, u, dtype=np.float32))
    bad = [im.copy() for im in ims]
    bad[1][8, 8, 8] = np.nan
    b = atlas.GreedyAtlasBuilderGPU([atlas.AtlasSubject(to_dev(im)) for im in bad], (16, 16, 16),
                                    max_pert=0.2, scale_levels=[(1, 3)])
    with pytest.warns(RuntimeWarning, match="NaN"):
        b.Run()
    assert b.flags & _lib.FLAG_NAN_ENERGY
    good = atlas.GreedyAtlasBuilderGPU([atlas.AtlasSubject(to_dev(im)) for im in ims], (16, 16, 16),
                                       max_pert=0.2, scale_levels=[(1, 3)])
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        good.Run()
    assert good.flags & (_lib.FLAG_NAN_ENERGY | _lib.FLAG_BAD_STEP) == 0
    la = ldmm.LDMMAtlasBuilderGPU([to_dev(im) for im in bad], n_time_steps=3,
                                  param=ldmm.LDMMIteratorParam(sigma=0.1, max_pert=0.2),
                                  scale_levels=[(1, 2)])
    with pytest.raises(_lib.NumericalError):
        la.Run()
