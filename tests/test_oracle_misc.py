"""Gradient closed forms, splat adjoint identity and index keys, float/double
greedy agreement, MetaImage I/O."""
import os

import numpy as np
import pytest

import oracle
from atlaswerks_amd import synth
from atlaswerks_amd.metaimage import Image, read_metaimage, read_nrrd, write_metaimage

SHAPE = (5, 6, 7)


def test_gradient_ramp_exact_including_edges():
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in SHAPE], indexing="ij")
    img = (2.0 * x - 3.0 * y + 0.5 * z).astype(np.float32)
    sp = (2.0, 1.0, 0.5)
    g = oracle.gradient(img, sp)
    np.testing.assert_allclose(g[..., 0], 2.0 / sp[0], atol=1e-12)
    np.testing.assert_allclose(g[..., 1], -3.0 / sp[1], atol=1e-12)
    np.testing.assert_allclose(g[..., 2], 0.5 / sp[2], atol=1e-12)


def test_gradient_quadratic_interior_and_onesided_edges():
    x = np.arange(SHAPE[2], dtype=np.float64)
    img = np.broadcast_to(x * x, SHAPE).astype(np.float32)
    g = oracle.gradient(img)[..., 0]
    np.testing.assert_allclose(g[..., 1:-1], np.broadcast_to(2 * x[1:-1], g[..., 1:-1].shape), atol=1e-12)  # central: exact
    assert g[0, 0, 0] == 1.0                                           # (1 - 0)/1
    assert g[0, 0, -1] == x[-1] ** 2 - x[-2] ** 2                      # one-sided


def test_gradient_wrap():
    img = synth.blobs(SHAPE, n_blobs=2, seed=1)
    g = oracle.gradient(img, wrap=True)[..., 0]
    np.testing.assert_allclose(g, (np.roll(img, -1, 2).astype(np.float64) - np.roll(img, 1, 2)) / 2, atol=1e-6)


def test_gradient_planted_mistake_caught():
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in SHAPE], indexing="ij")
    img = (2.0 * x).astype(np.float32)
    oracle.set_mutation(5)
    try:
        g = oracle.gradient(img)
    finally:
        oracle.set_mutation(0)
    assert not np.allclose(g[..., 0], 2.0)


def test_splat_adjoint_identity_and_mass():
    shape = (12, 11, 10)
    h = np.moveaxis(synth.identity(shape) + synth.sinusoidal_displacement(shape, 2.0, seed=4), 0, -1)
    i1 = synth.blobs(shape, n_blobs=3, seed=5)
    i2 = synth.blobs(shape, n_blobs=3, seed=6)
    lhs = float((oracle.apply(i1, h) * i2).sum())
    rhs = float((i1.astype(np.float64) * oracle.splat(i2, h)).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # interior points deposit all their mass
    ones = np.zeros(shape, np.float32)
    ones[4:8, 4:8, 4:7] = 1.0
    assert abs(oracle.splat(ones, h).sum() - ones.sum()) < 1e-9
    # float forwardApply (the reference CPU path) agrees with the double splat
    fa = oracle.forward_apply(i2, h)
    assert np.abs(fa - oracle.splat(i2, h)).max() < 1e-5


def test_splat_planted_mistake_caught():
    shape = (8, 8, 8)
    h = np.moveaxis(synth.identity(shape) + 0.3, 0, -1)
    i1 = synth.blobs(shape, n_blobs=2, seed=7)
    i2 = synth.blobs(shape, n_blobs=2, seed=8)
    oracle.set_mutation(9)
    try:
        rhs = float((i1.astype(np.float64) * oracle.splat(i2, h)).sum())
    finally:
        oracle.set_mutation(0)
    lhs = float((oracle.apply(i1, h) * i2).sum())
    assert abs(lhs - rhs) > 1e-4 * abs(lhs)


def test_splat_index_brute_force():
    """createIndex (cudaSplat.cu:31-117) keys on a tiny input, enumerated by hand."""
    shape = (2, 2, 3)  # z, y, x: w=3, h=2, l=2, N=12
    src = np.arange(1, 13, dtype=np.float32).reshape(shape)
    v = np.zeros(shape + (3,), np.float32)
    v[0, 0, 0] = (0.5, 0.25, 0.0)     # voxel 0 -> (0.5, 0.25, 0)
    v[1, 1, 2] = (0.0, 0.0, 0.0)      # last voxel: only its own corner is inside
    keys, vals = oracle.splat_index(src, v)
    N = 12
    assert list(keys[0:4]) == [0, 1, 3, 4]          # z-floor corners of voxel 0
    assert list(keys[4 * N:4 * N + 4]) == [6, 7, 9, 10]  # z-ceil corners (weight 0 but inside)
    np.testing.assert_allclose(vals[0:4], [1 * 0.5 * 0.75, 1 * 0.5 * 0.75, 1 * 0.5 * 0.25, 1 * 0.5 * 0.25])
    last = 11
    assert keys[4 * last] == 11 and list(keys[4 * last + 1:4 * last + 4]) == [N, N, N]
    assert list(keys[4 * (N + last):4 * (N + last) + 4]) == [N, N, N, N]
    # truncation without the floor fix (the reference sort/scan path) differs for x in (-1, 0)
    v2 = np.zeros(shape + (3,), np.float32)
    v2[0, 0, 1] = (-1.5, 0, 0)
    k_fix, _ = oracle.splat_index(src, v2, floor_fix=True)
    k_trunc, _ = oracle.splat_index(src, v2, floor_fix=False)
    assert k_fix[4] == N and k_fix[5] == 0   # floor(-0.5) = -1: corner x=-1 out, x=0 in
    assert k_trunc[4] == 0                   # int(-0.5) = 0


def test_greedy_float_vs_double_and_decrease():
    shape = (24, 24, 24)
    t, h, _ = synth.pair(shape, seed=1, max_disp=2.0)
    target = oracle.apply(t, np.moveaxis(h, 0, -1), dtype=np.float32)
    hf, _, ef, ed, sf = oracle.greedy_run(t, target, 15)
    hd, _, _, edd, sd = oracle.greedy_run(t, target, 15, dtype=np.float64)
    assert np.max(np.abs(ed - edd) / edd) < 1e-5
    assert np.max(np.abs(ef - edd) / edd) < 1e-4
    assert abs(sf - sd) < 1e-6 * sd
    assert np.abs(hf - hd).max() < 1e-4
    assert ed[-1] < 0.5 * ed[0]


def test_auto_step_reduce_halves_on_increase():
    shape = (16, 16, 16)
    t, h, _ = synth.pair(shape, seed=2, max_disp=2.0)
    target = oracle.apply(t, np.moveaxis(h, 0, -1), dtype=np.float32)
    _, _, e, _, step_plain = oracle.greedy_run(t, target, 12, max_pert=3.0)
    _, _, e2, _, step_auto = oracle.greedy_run(t, target, 12, max_pert=3.0, auto_step_reduce=True)
    increases = int(np.sum(np.diff(e2) > 0))
    assert increases >= 1
    assert abs(step_auto - step_plain / 2 ** increases) < 1e-9 * step_plain


def _greedy_asr_from_primitives(I0, I1, n_iter, max_pert, late=False):
    """GreedyIteratorCPU::Iterate with AutoStepReduce written out from the pinned
    apply, gradient, L^-1 and composeHV: the energy is taken before the update and
    GreedyDeformationData::AddEnergy halves the step on an increase before
    updateDeformation uses it (GreedyIteratorCPU.cxx:120-136,
    GreedyDeformationData.cxx:175-183); late=True plants the mistake of halving
    after the update."""
    shape = I0.shape
    h = oracle.identity(shape, np.float64)
    step, prev, energies = None, None, []
    for it in range(n_iter):
        idef = oracle.apply(I0, h)
        r = I1.astype(np.float64) - idef
        e = float((r * r).sum())
        energies.append(e)
        v = oracle.diffoper(oracle.gradient(idef) * r[..., None], inverse=True)
        increase = prev is not None and e > prev
        if increase and not late:
            step /= 2
        if it == 0:
            step = max_pert / np.sqrt((v * v).sum(-1)).max()
        h = oracle.compose_hv(h, step * v)
        if increase and late:
            step /= 2
        prev = e
    return np.array(energies), step


def test_auto_step_reduce_applies_to_the_rising_iteration():
    shape = (16, 16, 16)
    t, h, _ = synth.pair(shape, seed=2, max_disp=2.0)
    target = oracle.apply(t, np.moveaxis(h, 0, -1), dtype=np.float32)
    _, _, _, ed, step = oracle.greedy_run(t, target, 12, max_pert=3.0, auto_step_reduce=True,
                                          dtype=np.float64)
    e_ref, s_ref = _greedy_asr_from_primitives(t, target, 12, 3.0)
    assert np.max(np.abs(ed - e_ref) / e_ref) < 1e-4
    assert abs(step - s_ref) < 1e-6 * s_ref
    e_late, _ = _greedy_asr_from_primitives(t, target, 12, 3.0, late=True)
    assert np.max(np.abs(ed - e_late) / e_late) > 0.1  # the planted mistake is caught


def test_metaimage_roundtrip(tmp_path, golden_dir):
    im = read_metaimage(os.path.join(golden_dir, "BullseyeTest00.mha"))
    assert im.data.shape == (32, 32, 32) and im.spacing == (1.0, 1.0, 1.0)
    p = str(tmp_path / "x.mha")
    write_metaimage(p, Image(im.data * 2, (1, 2, 3), (4, 5, 6)))
    back = read_metaimage(p)
    np.testing.assert_array_equal(back.data, im.data * 2)
    assert back.spacing == (1, 2, 3) and back.origin == (4, 5, 6)
    hf = read_metaimage(os.path.join(golden_dir, "ComputeHFieldAsymmetricDeformation.mha"))
    assert hf.data.shape == (32, 32, 32, 3)
    nr = read_nrrd(os.path.join(golden_dir, "BullseyeTestBlur00.nhdr"))
    assert nr.data.shape == (36, 31, 32)
    np.testing.assert_allclose(nr.spacing, (2.0, 2.06452, 1.77778))


def test_splat_sorted_pinned():
    """cplSplating (cudaSplat.cu:175-238) restated: stable sort + sequential sums.

    Pinned three ways: (1) with floored positions it equals cpuSplat3D's double
    accumulation to float rounding; (2) its sorted pairs equal numpy's stable
    argsort of the createIndex pairs; (3) for positions >= 0 truncation and floor
    agree bit for bit, and a negative position gives the extrapolated weights
    (1.5, -0.5) that int() truncation implies (by hand)."""
    rng = np.random.default_rng(0)
    shape = (12, 10, 14)
    src = rng.random(shape).astype(np.float32)
    v = (rng.random(shape + (3,)) * 3 - 1.5).astype(np.float32)
    ref = oracle.splat(src, oracle.identity(shape) + v)
    d, k, vals = oracle.splat_sorted(src, v, floor_fix=True)
    assert np.abs(d - ref).max() < 1e-5 * np.abs(ref).max()
    keys, vv = oracle.splat_index(src, v, floor_fix=True)
    o = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(keys[o], k)
    np.testing.assert_array_equal(vv[o], vals)
    vpos = np.abs(v)  # every position >= 0
    a = oracle.splat_sorted(src, vpos, floor_fix=True)
    b = oracle.splat_sorted(src, vpos, floor_fix=False)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    one = np.zeros((1, 1, 2), np.float32)
    one[0, 0, 0] = 2.0
    v1 = np.zeros((1, 1, 2, 3), np.float32)
    v1[0, 0, 0, 0] = -0.5
    d_t, _, _ = oracle.splat_sorted(one, v1, floor_fix=False)
    d_f, _, _ = oracle.splat_sorted(one, v1, floor_fix=True)
    np.testing.assert_allclose(d_t[0, 0], [3.0, -1.0])   # int(-0.5) = 0: dx = 1.5
    np.testing.assert_allclose(d_f[0, 0], [1.0, 0.0])    # floor: corner -1 dropped
