"""HField3DUtils on the GPU (the reference's ``cudaHField3DUtils`` namespace,
CUDACommon/inc/cudaHField3DUtils.h).  All functions are asynchronous on the
current torch stream and write into caller-provided tensors.

Background strategies: BACKGROUND_STRATEGY_{ID, ZERO, PARTIAL_ID, PARTIAL_ZERO,
CLAMP, WRAP} for vector interpolation, VAL/ZERO/CLAMP/WRAP for images.
"""
import ctypes

import torch

from . import _lib
from .device import dims, planes, ptr, stream_handle

BACKGROUND_STRATEGY_ID = _lib.BG_ID
BACKGROUND_STRATEGY_ZERO = _lib.BG_ZERO
BACKGROUND_STRATEGY_PARTIAL_ID = _lib.BG_PARTIAL_ID
BACKGROUND_STRATEGY_PARTIAL_ZERO = _lib.BG_PARTIAL_ZERO
BACKGROUND_STRATEGY_CLAMP = _lib.BG_CLAMP
BACKGROUND_STRATEGY_WRAP = _lib.BG_WRAP
BACKGROUND_STRATEGY_VAL = _lib.BG_VAL


def _same(*ts):
    d = dims(ts[0])
    for t in ts[1:]:
        if dims(t) != d:
            raise ValueError("size mismatch: %s vs %s" % (d, dims(t)))
    return d


def setToIdentity(h, stream=None):
    nx, ny, nz = dims(h)
    _lib.call("awk_hfield_set_to_identity", *planes(h), nx, ny, nz, stream_handle(stream))


def velocityToH(v, spacing=(1.0, 1.0, 1.0), stream=None):
    nx, ny, nz = dims(v)
    _lib.call("awk_hfield_velocity_to_h", *planes(v), nx, ny, nz, *[float(s) for s in spacing],
              stream_handle(stream))


def hToVelocity(h, spacing=(1.0, 1.0, 1.0), stream=None):
    nx, ny, nz = dims(h)
    _lib.call("awk_hfield_h_to_velocity", *planes(h), nx, ny, nz, *[float(s) for s in spacing],
              stream_handle(stream))


def computeInverseZerothOrder(h, hinv, stream=None):
    nx, ny, nz = _same(h, hinv)
    _lib.call("awk_hfield_inverse_zeroth_order", *planes(hinv), *planes(h), nx, ny, nz,
              stream_handle(stream))


def apply(out, image, h, bg=BACKGROUND_STRATEGY_VAL, background=0.0, stream=None):
    """out(x) = image(h(x)).  The image may live on its own grid (h then holds
    coordinates in the image's voxels), as HField3DUtils::apply allows."""
    nx, ny, nz = _same(out, h)
    inx, iny, inz = dims(image)
    if (inx, iny, inz) == (nx, ny, nz):
        _lib.call("awk_hfield_apply", ptr(out), ptr(image), *planes(h), nx, ny, nz, int(bg),
                  float(background), stream_handle(stream))
    else:
        _lib.call("awk_hfield_apply_grid", ptr(out), ptr(image), inx, iny, inz, *planes(h), nx, ny,
                  nz, int(bg), float(background), stream_handle(stream))


def initializeFromAffine(h, m, stream=None):
    """h(x) = M [x, y, z, 1] for a 3x4 voxel-to-voxel affine M
    (HField3DUtils::initializeFromAffine with the grid conversions folded in)."""
    import numpy as np
    nx, ny, nz = dims(h)
    arr = (ctypes.c_float * 12)(*[float(v) for v in np.asarray(m, dtype=np.float64).reshape(12)])
    _lib.call("awk_hfield_init_affine", *planes(h), nx, ny, nz, arr, stream_handle(stream))


def grid_affine(src_origin, src_spacing, dst_origin, dst_spacing, world=None):
    """3x4 map from voxels of a source grid to voxels of a destination grid through
    world space: p = o_s + sp_s x; q = W p (W a 3x4 world affine, identity if None);
    result (q - o_d) / sp_d."""
    import numpy as np
    so, ss = np.asarray(src_origin, float), np.asarray(src_spacing, float)
    do, ds = np.asarray(dst_origin, float), np.asarray(dst_spacing, float)
    W = np.eye(4)[:3] if world is None else np.asarray(world, float).reshape(3, 4)
    A = np.zeros((3, 4))
    A[:, :3] = (W[:, :3] * ss[None, :]) / ds[:, None]
    A[:, 3] = (W[:, :3] @ so + W[:, 3] - do) / ds
    return A


def compose(out, f, g, bg=BACKGROUND_STRATEGY_ID, stream=None):
    """out(x) = f(g(x))."""
    nx, ny, nz = _same(out, f, g)
    _lib.call("awk_hfield_compose", *planes(out), *planes(f), *planes(g), nx, ny, nz, int(bg),
              stream_handle(stream))


def composeHV(out, g, v, spacing=(1.0, 1.0, 1.0), bg=BACKGROUND_STRATEGY_PARTIAL_ID, delta=1.0,
              stream=None):
    """out(x) = g(x + delta*v(x)/spacing)."""
    nx, ny, nz = _same(out, g, v)
    _lib.call("awk_hfield_compose_hv", *planes(out), *planes(g), *planes(v), nx, ny, nz,
              *[float(s) for s in spacing], float(delta), 0, int(bg), stream_handle(stream))


def composeHVInv(out, g, v, spacing=(1.0, 1.0, 1.0), bg=BACKGROUND_STRATEGY_PARTIAL_ID, delta=1.0,
                 stream=None):
    """out(x) = g(x - delta*v(x)/spacing)."""
    nx, ny, nz = _same(out, g, v)
    _lib.call("awk_hfield_compose_hv", *planes(out), *planes(g), *planes(v), nx, ny, nz,
              *[float(s) for s in spacing], float(delta), 1, int(bg), stream_handle(stream))


def composeVH(out, v, g, spacing=(1.0, 1.0, 1.0), bg=BACKGROUND_STRATEGY_PARTIAL_ZERO, stream=None):
    """out(x) = g(x) + v(g(x))/spacing."""
    nx, ny, nz = _same(out, v, g)
    _lib.call("awk_hfield_compose_vh", *planes(out), *planes(v), *planes(g), nx, ny, nz,
              *[float(s) for s in spacing], 0, int(bg), stream_handle(stream))


def composeVHInv(out, v, g, spacing=(1.0, 1.0, 1.0), bg=BACKGROUND_STRATEGY_PARTIAL_ZERO,
                 stream=None):
    """out(x) = g(x) - v(g(x))/spacing."""
    nx, ny, nz = _same(out, v, g)
    _lib.call("awk_hfield_compose_vh", *planes(out), *planes(v), *planes(g), nx, ny, nz,
              *[float(s) for s in spacing], 1, int(bg), stream_handle(stream))


def computeGradient(grad, image, spacing=(1.0, 1.0, 1.0), wrap=False, stream=None):
    """Array3DUtils::computeGradient / cplComputeGradient."""
    nx, ny, nz = _same(grad, image)
    _lib.call("awk_gradient", *planes(grad), ptr(image), nx, ny, nz, *[float(s) for s in spacing],
              int(bool(wrap)), stream_handle(stream))


def splat(dst, src, h, mode=_lib.SPLAT_FIXED64, stream=None):
    """cplSplat3DH / HField3DUtils::forwardApply (normalize=false): dst(h(x)) += src(x)."""
    nx, ny, nz = _same(dst, src, h)
    _lib.call("awk_splat", ptr(dst), ptr(src), *planes(h), nx, ny, nz, int(mode),
              stream_handle(stream))


cplSplat3DH = splat


def splat_sorted(dst, src, v, reference_truncation=True, keys=None, vals=None, stream=None):
    """cplSplating (cudaSplat.cu:175-238): sort/scan splat of src to voxel + v.

    Deterministic.  ``reference_truncation`` keeps the reference's int()
    truncation of negative positions (bit-exact keys); False floors them.
    ``keys`` (int32, 8*N, holding uint32 cells) / ``vals`` (float32, 8*N) receive
    the sorted (cell, contribution) pairs when given."""
    nx, ny, nz = _same(dst, src, v)
    n8 = 8 * nx * ny * nz
    kp = vp = None
    if keys is not None:
        if not keys.is_cuda or keys.dtype != torch.int32 or keys.numel() != n8 or not keys.is_contiguous():
            raise ValueError("keys must be a contiguous int32 GPU tensor of 8*N elements")
        kp = ctypes.c_void_p(keys.data_ptr())
    if vals is not None:
        if vals.numel() != n8:
            raise ValueError("vals must hold 8*N elements")
        vp = ptr(vals)
    _lib.call("awk_splat_sorted", ptr(dst), ptr(src), *planes(v), nx, ny, nz,
              int(bool(reference_truncation)), kp, vp, stream_handle(stream))


cplSplating = splat_sorted


def gaussianFilter(out, image, sigma, half, stream=None):
    """GaussianFilter3D: separable Gaussian (sigma, half sizes per axis x, y, z)."""
    nx, ny, nz = _same(out, image)
    _lib.call("awk_gaussian_filter", ptr(out), ptr(image), nx, ny, nz, *[float(v) for v in sigma],
              *[int(v) for v in half], stream_handle(stream))


def gaussianDownsample(out, image, stream=None):
    """ImageUtils::gaussianDownsample from image's grid to out's grid."""
    onx, ony, onz = dims(out)
    nx, ny, nz = dims(image)
    _lib.call("awk_gaussian_downsample", ptr(out), onx, ony, onz, ptr(image), nx, ny, nz,
              stream_handle(stream))


def resampleNew(out, field, bg=BACKGROUND_STRATEGY_CLAMP, rescale=False, stream=None):
    """HField3DUtils::resampleNew of a vector field onto out's grid."""
    onx, ony, onz = dims(out)
    nx, ny, nz = dims(field)
    _lib.call("awk_hfield_resample", *planes(out), onx, ony, onz, *planes(field), nx, ny, nz,
              int(bg), int(bool(rescale)), stream_handle(stream))


def addAffine(h, m, sign=1.0, stream=None):
    """h += sign * M [x, y, z, 1] (M a 3x4 voxel affine)."""
    import numpy as np
    nx, ny, nz = dims(h)
    arr = (ctypes.c_float * 12)(*[float(v) for v in np.asarray(m, dtype=np.float64).reshape(12)])
    _lib.call("awk_hfield_add_affine", *planes(h), nx, ny, nz, arr, float(sign), stream_handle(stream))


IDENTITY_AFFINE = ((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 0.0))


def affineDisplace(h, u, m=IDENTITY_AFFINE, stream=None):
    """h(x) = M [x + u(x), 1] (velocityToH with unit spacing, then the 3x4 voxel
    affine M); h may be u itself."""
    import numpy as np
    nx, ny, nz = _same(h, u)
    arr = (ctypes.c_float * 12)(*[float(v) for v in np.asarray(m, dtype=np.float64).reshape(12)])
    _lib.call("awk_hfield_affine_displace", *planes(h), *planes(u), nx, ny, nz, arr,
              stream_handle(stream))


def upsampleHField(h, new_size, rescale=True, base_in=IDENTITY_AFFINE, base_out=IDENTITY_AFFINE,
                   stream=None):
    """Scale-level change of an hfield (GreedyDeformationData::SetScaleLevel,
    GreedyDeformationData.cxx:208-219): v = h - base_in(x); v resampled to new_size
    with CLAMP (MultiscaleManager::UpsampleToLevel, scaled by 1/r when rescale);
    h' = base_out(x) + v.  The base maps are the identity in the reference."""
    nx, ny, nz = (int(v) for v in new_size)
    v = h.clone()
    addAffine(v, base_in, -1.0, stream)
    out = torch.empty((3, nz, ny, nx), dtype=torch.float32, device=h.device)
    resampleNew(out, v, BACKGROUND_STRATEGY_CLAMP, rescale, stream)
    addAffine(out, base_out, 1.0, stream)
    return out
