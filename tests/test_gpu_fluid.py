"""FluidKernelFFT (L, L^-1) on the GPU against the FP64 oracle.

Bar: relative L2 <= 1e-5 against the FP64 restatement of
DiffOper::Apply(Inverse)OperatorOnTheFly (DiffOper.cxx:747-905)."""
import numpy as np
import pytest

import oracle
from atlaswerks_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5

SIZES = [(8, 8, 8), (6, 10, 12), (24, 40, 36), (32, 31, 36), (64, 64, 64), (17, 9, 5),
         (96, 48, 40), (1, 16, 16), (128, 128, 128), (32, 64, 256), (256, 256, 16),
         (128, 32, 64)]
# axes with a prime factor above 64 (Bluestein's transform): 67, 71, 127, 181 on each
# axis; 217 = 7 * 31 and 61 take direct prime stages
PRIME_SIZES = [(12, 20, 67), (16, 127, 10), (181, 12, 14), (61, 67, 71), (36, 217, 181),
               (1, 1, 181)]
# axes whose pass blocks stage more than 64 KiB of LDS (up to 159 KiB): 2500 along x,
# Bluestein primes 1021 along x and 3001 along y, 1499 along z
LONG_SIZES = [(3, 4, 2500), (4, 5, 1021), (2, 3001, 3), (1499, 3, 2)]


def _field(shape, seed):
    return synth.smooth_noise_field(shape, sigma=1.5, seed=seed)


@pytest.mark.parametrize("shape", SIZES + PRIME_SIZES + LONG_SIZES)
@pytest.mark.parametrize("inverse", [True, False])
def test_operator_matches_oracle(shape, inverse):
    import torch
    from atlaswerks_amd.fluid import DiffOperGPU, DiffOperParam
    from gpu_util import aos, host, rel_l2, soa, to_dev
    nz, ny, nx = shape
    print("fluid", shape, "inverse", inverse, flush=True)
    f = _field(shape, seed=nx * 7 + ny)
    sp = (1.0, 1.25, 0.8)
    p = DiffOperParam(alpha=0.01, beta=0.01, gamma=0.001)
    op = DiffOperGPU((nx, ny, nz), sp, p)
    d = to_dev(f)
    if inverse:
        op.ApplyInverseOperator(d)
    else:
        op.ApplyOperator(d)
    got = host(d)
    ref = oracle.diffoper(aos(f), sp=sp, alpha=0.01, beta=0.01, gamma=0.001, inverse=inverse)
    err = rel_l2(got, soa(ref))
    assert err <= TOL, err


@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 40, 36)])
@pytest.mark.parametrize("lpow,divfree", [(1.0, False), (0.5, False), (1.0, True), (2.0, True)])
def test_eigen_and_divfree(shape, lpow, divfree):
    from atlaswerks_amd.fluid import DiffOperGPU, DiffOperParam
    from gpu_util import aos, host, rel_l2, soa, to_dev
    nz, ny, nx = shape
    print("fluid eigen", shape, lpow, divfree, flush=True)
    f = _field(shape, seed=3)
    sp = (1.0, 1.0, 1.5)
    p = DiffOperParam(alpha=0.02, beta=0.01, gamma=0.01, lpow=lpow, use_eigen_lut=True,
                      divergence_free=divfree)
    op = DiffOperGPU((nx, ny, nz), sp, p)
    for inverse in (True, False):
        d = to_dev(f)
        (op.ApplyInverseOperator if inverse else op.ApplyOperator)(d)
        ref = oracle.diffoper(aos(f), sp=sp, alpha=0.02, beta=0.01, gamma=0.01, inverse=inverse,
                              lpow=lpow, eigen=True, divfree=divfree)
        err = rel_l2(host(d), soa(ref))
        assert err <= TOL, (inverse, err)


def test_round_trip_identity():
    """L(L^-1 f) = f on the GPU (DiffOperRoutineTests.cxx:120-158)."""
    from atlaswerks_amd.fluid import DiffOperGPU
    from gpu_util import host, rel_l2, to_dev
    shape = (32, 32, 32)
    f = _field(shape, seed=11)
    op = DiffOperGPU((32, 32, 32), (1.0, 1.25, 1.0))
    d = to_dev(f)
    op.ApplyInverseOperator(d)
    op.ApplyOperator(d)
    assert rel_l2(host(d), f) <= 1e-5


@pytest.mark.parametrize("knobs", [{"fft_order": 1}, {"x_rt": 2}, {"x_rt": 4, "y_tx": 2, "z_tx": 1},
                                   {"fft_order": 1, "x_rt": 6, "y_tx": 4, "z_tx": 2}])
@pytest.mark.parametrize("shape", [(24, 40, 36), (36, 48, 60), (12, 20, 67), (128, 128, 128)])
def test_tuned_fft_plans_match_oracle(shape, knobs):
    """The plans the autotuner chooses between (radix order of the generic stages,
    x / y / z pass tiles): L^-1 within the same 1e-5 of the FP64 oracle."""
    from atlaswerks_amd import autotune
    from atlaswerks_amd.fluid import DiffOperGPU, DiffOperParam
    from gpu_util import aos, host, rel_l2, soa, to_dev
    nz, ny, nx = shape
    print("fluid plan", shape, knobs, flush=True)
    f = _field(shape, seed=11)
    sp = (1.0, 1.25, 0.8)
    try:
        for k, v in knobs.items():
            autotune.set_knob(k, v)
        op = DiffOperGPU((nx, ny, nz), sp, DiffOperParam(alpha=0.01, beta=0.01, gamma=0.001))
    finally:
        for k in autotune.PLAN_KNOBS:
            autotune.set_knob(k, 0)
    d = to_dev(f)
    op.ApplyInverseOperator(d)
    ref = oracle.diffoper(aos(f), sp=sp, alpha=0.01, beta=0.01, gamma=0.001, inverse=True)
    assert rel_l2(host(d), soa(ref)) <= TOL
