"""HField3DUtils apply / compose family / gradient / splat on the GPU vs the oracle.

Bar: relative L2 <= 1e-6 against the FP64 restatement (HField3DUtils.txx,
Array3DUtils.txx); splat against cpuSplat3D's double accumulation."""
import numpy as np
import pytest

import oracle
from atlaswerks_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-6
SHAPES = [(32, 32, 32), (24, 40, 36), (17, 9, 5), (64, 48, 40)]
VEC_BG = [oracle.BG_ID, oracle.BG_ZERO, oracle.BG_PARTIAL_ID, oracle.BG_PARTIAL_ZERO,
          oracle.BG_CLAMP, oracle.BG_WRAP]


def _hfield(shape, seed, amp=3.0):
    u = synth.sinusoidal_displacement(shape, max_disp=amp, seed=seed)
    return synth.identity(shape) + u, u


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bg", [oracle.BG_VAL, oracle.BG_ZERO, oracle.BG_CLAMP, oracle.BG_WRAP])
def test_apply(shape, bg):
    import torch
    from atlaswerks_amd import hfield
    from gpu_util import aos, host, rel_l2, to_dev
    print("apply", shape, bg, flush=True)
    img = synth.blobs(shape, seed=1) + 0.1
    h, _ = _hfield(shape, 2, amp=4.0)
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    hfield.apply(out, to_dev(img), to_dev(h), bg=bg, background=0.25)
    ref = oracle.apply(img, aos(h), bg=bg, background=0.25)
    assert rel_l2(host(out), ref) <= TOL


@pytest.mark.parametrize("shape", SHAPES[:3])
@pytest.mark.parametrize("bg", VEC_BG)
def test_compose_family(shape, bg):
    import torch
    from atlaswerks_amd import hfield
    from gpu_util import aos, host, rel_l2, soa, to_dev
    print("compose", shape, bg, flush=True)
    f, _ = _hfield(shape, 3, amp=4.0)
    g, _ = _hfield(shape, 4, amp=4.0)
    v = synth.sinusoidal_displacement(shape, max_disp=3.0, seed=5)
    sp = (1.0, 1.5, 0.75)
    out = torch.empty((3,) + shape, dtype=torch.float32, device="cuda")
    F, G, V = to_dev(f), to_dev(g), to_dev(v)
    hfield.compose(out, F, G, bg=bg)
    assert rel_l2(host(out), soa(oracle.compose(aos(f), aos(g), bg=bg))) <= TOL
    hfield.composeHV(out, G, V, sp, bg=bg)
    assert rel_l2(host(out), soa(oracle.compose_hv(aos(g), aos(v), sp, bg=bg))) <= TOL
    hfield.composeHVInv(out, G, V, sp, bg=bg)
    assert rel_l2(host(out), soa(oracle.compose_hv(aos(g), aos(v), sp, bg=bg, inv=True))) <= TOL
    hfield.composeVH(out, V, G, sp, bg=bg)
    assert rel_l2(host(out), soa(oracle.compose_vh(aos(v), aos(g), sp, bg=bg))) <= TOL
    hfield.composeVHInv(out, V, G, sp, bg=bg)
    assert rel_l2(host(out), soa(oracle.compose_vh(aos(v), aos(g), sp, bg=bg, inv=True))) <= TOL


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("wrap", [False, True])
def test_gradient(shape, wrap):
    import torch
    from atlaswerks_amd import hfield
    from gpu_util import host, rel_l2, soa, to_dev
    img = synth.blobs(shape, seed=6)
    sp = (1.0, 2.0, 0.5)
    out = torch.empty((3,) + shape, dtype=torch.float32, device="cuda")
    hfield.computeGradient(out, to_dev(img), sp, wrap)
    assert rel_l2(host(out), soa(oracle.gradient(img, sp, wrap))) <= TOL


def test_conversions():
    import torch
    from atlaswerks_amd import hfield
    from gpu_util import host, to_dev
    shape = (9, 10, 11)
    h, u = _hfield(shape, 7)
    sp = (1.0, 2.0, 0.5)
    t = to_dev(u * np.array(sp, dtype=np.float32)[:, None, None, None])
    hfield.velocityToH(t, sp)
    np.testing.assert_allclose(host(t), h, atol=1e-5)
    hfield.hToVelocity(t, sp)
    np.testing.assert_allclose(host(t), u * np.array(sp, dtype=np.float32)[:, None, None, None], atol=1e-4)
    ident = torch.empty((3,) + shape, dtype=torch.float32, device="cuda")
    hfield.setToIdentity(ident)
    np.testing.assert_array_equal(host(ident), synth.identity(shape))
    out = torch.empty_like(ident)
    hfield.computeInverseZerothOrder(to_dev(h), out)
    np.testing.assert_allclose(host(out), synth.identity(shape) - u, atol=1e-5)


@pytest.mark.parametrize("shape", SHAPES)
def test_splat_modes(shape):
    import torch
    from atlaswerks_amd import _lib, hfield
    from gpu_util import aos, host, rel_l2, to_dev
    print("splat", shape, flush=True)
    src = synth.blobs(shape, seed=8) + 0.05
    h, _ = _hfield(shape, 9, amp=3.0)
    ref = oracle.splat(src, aos(h))
    S, H = to_dev(src), to_dev(h)
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    hfield.splat(out, S, H, mode=_lib.SPLAT_FIXED64)
    a = host(out).copy()
    assert rel_l2(a, ref) <= TOL
    hfield.splat(out, S, H, mode=_lib.SPLAT_FIXED64)
    np.testing.assert_array_equal(host(out), a)  # deterministic
    hfield.splat(out, S, H, mode=_lib.SPLAT_FLOAT)
    assert rel_l2(host(out), ref) <= TOL
    hfield.splat(out, S, H, mode=_lib.SPLAT_FIXED20)
    assert rel_l2(host(out), ref) <= 1e-5


def test_splat_adjoint_identity():
    """<apply(I1,h), I2> = <I1, splat(I2,h)> (cudaHField3DUtilsRoutineTests.cxx:61-72)."""
    import torch
    from atlaswerks_amd import hfield
    from gpu_util import host, to_dev
    shape = (32, 32, 32)
    i1 = synth.blobs(shape, seed=10)
    i2 = synth.blobs(shape, seed=11)
    h, _ = _hfield(shape, 12, amp=4.0)
    a = torch.empty(shape, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    hfield.apply(a, to_dev(i1), to_dev(h))
    hfield.splat(b, to_dev(i2), to_dev(h))
    lhs = float((host(a).astype(np.float64) * i2).sum())
    rhs = float((i1.astype(np.float64) * host(b)).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


# (6, 7, 8): keys of 9 bits, one 9-bit radix pass; (32, 32, 64): N = 2^16, keys of
# 17 bits, an 8-bit then a 9-bit pass (the plan of every power-of-two volume
# whose exponent is a multiple of 8, 256^3 among them)
@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 15, 17), (4, 8, 8), (17, 9, 5), (6, 7, 8),
                                   (24, 40, 36), (32, 32, 64), (64, 64, 64)])
@pytest.mark.parametrize("trunc", [True, False])
def test_splat_sorted_bit_exact(shape, trunc):
    """Sort/scan splat: sorted keys, sorted contributions and cell sums equal the
    oracle's stable sort bit for bit (the sums run in the same sorted order)."""
    import torch
    from atlaswerks_amd import hfield
    from gpu_util import host, soa, to_dev
    print("splat_sorted", shape, trunc, flush=True)
    rng = np.random.default_rng(sum(shape))
    src = (rng.random(shape) + 0.05).astype(np.float32)
    v = (rng.random(shape + (3,)) * 5 - 2.5).astype(np.float32)
    ref_d, ref_k, ref_v = oracle.splat_sorted(src, v, floor_fix=not trunc)
    n8 = 8 * src.size
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    keys = torch.empty(n8, dtype=torch.int32, device="cuda")
    vals = torch.empty(n8, dtype=torch.float32, device="cuda")
    S, V = to_dev(src), to_dev(soa(v))
    hfield.splat_sorted(out, S, V, reference_truncation=trunc, keys=keys, vals=vals)
    np.testing.assert_array_equal(host(keys).view(np.uint32), ref_k)
    np.testing.assert_array_equal(host(vals), ref_v)
    d = host(out).copy()
    np.testing.assert_array_equal(d, ref_d)
    hfield.splat_sorted(out, S, V, reference_truncation=trunc)
    np.testing.assert_array_equal(host(out), d)  # deterministic, also without key export


@pytest.mark.parametrize("shape", [(16, 16, 16), (20, 24, 28)])
def test_splat_sorted_collapse(shape):
    """Every voxel displaced onto one point: 8 cells receive N contributions each,
    segments far longer than a segment-sum tile (2048 pairs), so each head's walk
    continues past several tile ends.  Bit-exact against the oracle."""
    import torch
    from atlaswerks_amd import hfield
    from gpu_util import host, soa, to_dev
    print("splat_sorted collapse", shape, flush=True)
    nz, ny, nx = shape
    rng = np.random.default_rng(7)
    src = (rng.random(shape) + 0.05).astype(np.float32)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    target = np.array([nx * 0.43, ny * 0.57, nz * 0.38], dtype=np.float32)
    v = np.stack([target[0] - x, target[1] - y, target[2] - z], axis=-1).astype(np.float32)
    ref_d, ref_k, ref_v = oracle.splat_sorted(src, v, floor_fix=False)
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    keys = torch.empty(8 * src.size, dtype=torch.int32, device="cuda")
    vals = torch.empty(8 * src.size, dtype=torch.float32, device="cuda")
    hfield.splat_sorted(out, to_dev(src), to_dev(soa(v)), keys=keys, vals=vals)
    np.testing.assert_array_equal(host(keys).view(np.uint32), ref_k)
    np.testing.assert_array_equal(host(vals), ref_v)
    np.testing.assert_array_equal(host(out), ref_d)
    assert np.count_nonzero(ref_d) <= 8


@pytest.mark.parametrize("shape,amp", [((17, 9, 5), 3.0), ((64, 48, 40), 3.0), ((40, 36, 100), 1.0),
                                       ((32, 32, 32), 40.0), ((256, 256, 256), 8.0)])
def test_splat_tiled_equals_direct(shape, amp):
    """The LDS-window splat and the direct atomic splat add the same int64
    contributions, so their FIXED64 results are bit-identical (40-voxel
    displacements overflow the window and take the direct path inside the tile)."""
    import torch
    from atlaswerks_amd import _lib, hfield
    from atlaswerks_amd.autotune import get_knob, set_knob
    from gpu_util import host, to_dev
    print("splat tiled", shape, amp, flush=True)
    src = synth.blobs(shape, seed=12) - 0.5
    h, _ = _hfield(shape, 13, amp=amp)
    S, H = to_dev(src), to_dev(h)
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    old = get_knob("splat_tile")
    try:
        set_knob("splat_tile", 0)
        hfield.splat(out, S, H, mode=_lib.SPLAT_FIXED64)
        direct = host(out).copy()
        set_knob("splat_tile", 1)
        hfield.splat(out, S, H, mode=_lib.SPLAT_FIXED64)
        np.testing.assert_array_equal(host(out), direct)
    finally:
        set_knob("splat_tile", old)
