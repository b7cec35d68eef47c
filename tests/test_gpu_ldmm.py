"""LDMM on the GPU vs the FP64 oracle, both gradient forms: IterateOrig
(LDMMIteratorCPU.cxx:379-527, the reference's default) and IterateNew
(LDMMIteratorCPU.cxx:529-631).

Bars: velocity fields relative L2 <= 1e-4 after several iterations (float GPU vs
FP64 restatement; IterateNew's splat is int64 fixed point, 2^-33 per
contribution), energies and step relative 1e-5, alpha images relative L2 <= 1e-5;
repeated runs bit-identical (deterministic splat and FP64 reductions)."""
import numpy as np
import pytest

import oracle
from atlaswerks_amd import synth

pytestmark = pytest.mark.gpu


def _pair(shape, seed, disp=2.0):
    t, h, _ = synth.pair(shape, seed=seed, max_disp=disp)
    return t, oracle.apply(t, np.moveaxis(h, 0, -1), dtype=np.float32)


def _soa_v(v_aos):
    return np.ascontiguousarray(np.moveaxis(v_aos, -1, 1))  # [T,z,y,x,3] -> [T,3,z,y,x]


CASES = [((16, 20, 24), 5, 4, (1.0, 1.0, 1.0)),
         ((32, 32, 32), 4, 3, (1.0, 1.2, 0.9)),
         ((64, 64, 64), 10, 2, (1.0, 1.0, 1.0))]


# (IterateOrig?, IterateNew's splat)
MODES = [(True, "fixed"), (False, "fixed"), (False, "sorted")]
GRADIENTS = [True, False]


def _oracle_run(orig):
    return oracle.ldmm_run_orig if orig else oracle.ldmm_run


@pytest.mark.parametrize("orig,splat", MODES)
@pytest.mark.parametrize("shape,nt,nit,sp", CASES)
def test_ldmm_run_vs_oracle(shape, nt, nit, sp, orig, splat):
    import torch
    from atlaswerks_amd import ldmm
    from gpu_util import rel_l2, to_dev
    print("ldmm", "orig" if orig else "new", splat, shape, nt, nit, sp, flush=True)
    I0, I1 = _pair(shape, seed=5)
    sigma, mp = 0.1, 0.2
    v_ref, ei, ev, st = _oracle_run(orig)(I0, I1, nt, nit, sp=sp, sigma=sigma, max_pert=mp)
    prm = ldmm.LDMMIteratorParam(sigma=sigma, max_pert=mp, use_orig_gradient=orig, splat=splat)
    w = ldmm.LDMMWarp(to_dev(I0), to_dev(I1), n_time_steps=nt, n_iter=nit, spacing=sp, param=prm)
    data = w.RunWarp(vec_energy=True)
    torch.cuda.synchronize()
    assert abs(w.iterator.StepSize() / st - 1) < 1e-5
    assert rel_l2(data.v.cpu().numpy(), _soa_v(v_ref)) <= 1e-4
    e = data.energies()
    np.testing.assert_allclose(e[:, 0], ei, rtol=1e-5)
    np.testing.assert_allclose(e[:, 1], ev, rtol=1e-5)
    np.testing.assert_allclose(e[:, 2], ei + ev, rtol=1e-5)


@pytest.mark.parametrize("orig,splat", MODES)
def test_ldmm_step_call_leaves_v_and_deterministic(orig, splat):
    import torch
    from atlaswerks_amd import ldmm
    from gpu_util import to_dev
    print("ldmm determinism", "orig" if orig else "new", splat, flush=True)
    shape = (24, 28, 20)
    I0, I1 = _pair(shape, seed=6)
    _, _, _, st = _oracle_run(orig)(I0, I1, 3, 0, sigma=0.1, max_pert=0.2)
    prm = ldmm.LDMMIteratorParam(sigma=0.1, max_pert=0.2, use_orig_gradient=orig, splat=splat)
    outs = []
    for rep in range(2):
        w = ldmm.LDMMWarp(to_dev(I0), to_dev(I1), n_time_steps=3, n_iter=0, param=prm)
        w.RunWarp()
        torch.cuda.synchronize()
        assert float(w.data.v.abs().max()) == 0.0
        assert abs(w.iterator.StepSize() / st - 1) < 1e-5
        for _ in range(3):
            w.iterator.Iterate(w.data, vec_energy=True)
        outs.append((w.data.v.cpu().numpy(), w.data.energies()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("orig,splat", MODES)
def test_ldmm_alphas_vs_oracle(orig, splat):
    import torch
    from atlaswerks_amd import ldmm
    from gpu_util import rel_l2, to_dev
    print("ldmm alphas", "orig" if orig else "new", splat, flush=True)
    shape = (20, 24, 28)
    nt = 4
    I0, I1 = _pair(shape, seed=7)
    rng = np.random.default_rng(7)
    v = np.stack([np.moveaxis(synth.sinusoidal_displacement(shape, max_disp=1.5, seed=10 + t), 0, -1)
                  for t in range(nt)]).astype(np.float64)
    want = (oracle.ldmm_alphas_orig if orig else oracle.ldmm_alphas)(I0, I1, v, sigma=0.1)
    prm = ldmm.LDMMIteratorParam(sigma=0.1, use_orig_gradient=orig, splat=splat)
    data = ldmm.LDMMDeformationDataGPU(to_dev(I0), to_dev(I1), nt)
    data.v.copy_(torch.from_numpy(_soa_v(v.astype(np.float32))))
    it = ldmm.LDMMIteratorGPU(shape[::-1], (1.0, 1.0, 1.0), nt, prm)
    a = it.ComputeAlphas(data)
    torch.cuda.synchronize()
    for t in range(nt):
        assert rel_l2(a[t].cpu().numpy(), want[t]) <= 1e-5, t
    del rng


def test_ldmm_warp_scale_levels_vs_oracle():
    """Two scale levels (LDMMWarp.txx:69-125, LDMMDeformationData.cxx:119-161): the
    oracle chain is a coarse run on the Gaussian-downsampled images, the velocity
    fields resampled (CLAMP, no rescaling) to the fine grid, then a fine run that
    starts with a new step-size call."""
    import torch
    from atlaswerks_amd import ldmm
    from gpu_util import rel_l2, to_dev
    print("ldmm scale levels", flush=True)
    shape = (24, 28, 32)
    sp = (1.0, 1.1, 0.9)
    I0, I1 = _pair(shape, seed=9)
    T, kw = 4, dict(sigma=0.1, max_pert=0.2)
    cshape = tuple(n // 2 for n in shape)
    c0, c1 = oracle.gaussian_downsample(I0, cshape), oracle.gaussian_downsample(I1, cshape)
    vc, eic, evc, _ = oracle.ldmm_run_orig(c0, c1, T, 3, sp=tuple(2 * s for s in sp), **kw)
    v0 = np.stack([oracle.resample_hfield(vc[t], shape) for t in range(T)])
    vf, eif, evf, stf = oracle.ldmm_run_orig(I0, I1, T, 2, v=v0, sp=sp, **kw)
    prm = ldmm.LDMMIteratorParam(**kw)
    w = ldmm.LDMMWarp(to_dev(I0), to_dev(I1), n_time_steps=T, spacing=sp, param=prm,
                      scale_levels=[(2, 3), (1, 2)])
    data = w.RunWarp(vec_energy=True)
    torch.cuda.synchronize()
    assert abs(w.iterator.StepSize() / stf - 1) < 1e-4
    assert rel_l2(data.v.cpu().numpy(), _soa_v(vf)) <= 2e-4
    np.testing.assert_allclose(w.energy_history[0][:, 0], eic, rtol=1e-4)
    np.testing.assert_allclose(w.energy_history[1][:, 0], eif, rtol=1e-4)
    np.testing.assert_allclose(w.energy_history[1][:, 1], evf, rtol=1e-4)
