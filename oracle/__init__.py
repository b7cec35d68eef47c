"""CPU oracle for the greedy-fluid hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker the HIP path is compared against.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it,
and only as the checker; the product library never calls into it.

It wraps ``liboracle.so`` (built from ``awk_oracle.cpp`` by ``oracle/Makefile``),
a C++ restatement of the reference CPU path; every function there cites the
reference file:line it follows.  Arrays here are numpy, indexed ``[z, y, x]``
(x fastest, as ``Array3D`` stores them); vector fields are AoS ``[z, y, x, 3]``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BG_ID, BG_ZERO, BG_PARTIAL_ID, BG_PARTIAL_ZERO, BG_CLAMP, BG_WRAP, BG_VAL = range(7)
DO_EIGEN, DO_DIVFREE = 1, 2


class GreedyParams(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int * 3),
        ("sp", ctypes.c_double * 3),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("max_pert", ctypes.c_double),
        ("n_iter", ctypes.c_int),
        ("compute_inverse", ctypes.c_int),
        ("auto_step_reduce", ctypes.c_int),
    ]


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _LIB.orc_trilerp_scalar_d.restype = ctypes.c_double
        _LIB.orc_greedy_time_f.restype = ctypes.c_double
        _LIB.orc_greedy_time_threads_f.restype = ctypes.c_double
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _size(shape3):
    nz, ny, nx = shape3
    return (ctypes.c_int * 3)(nx, ny, nz)


def _sp(sp):
    return (ctypes.c_double * 3)(*[float(s) for s in sp])


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_mutation(m):
    lib().orc_set_mutation(int(m))


def fft(x, inverse=False, dtype=np.float64):
    """Unnormalised 1-D complex FFT of the oracle (forward = exp(-2 pi i jk/n))."""
    x = np.asarray(x)
    if dtype == np.float64:
        buf = np.ascontiguousarray(x, dtype=np.complex128).copy()
        lib().orc_fft_d(len(buf), _p(buf), int(inverse))
    else:
        buf = np.ascontiguousarray(x, dtype=np.complex64).copy()
        lib().orc_fft_f(len(buf), _p(buf), int(inverse))
    return buf


def lut(n, sp):
    c = np.zeros(n)
    s = np.zeros(n)
    lib().orc_lut(int(n), ctypes.c_double(sp), _p(c), _p(s))
    return c, s


def diffoper(vf, sp=(1, 1, 1), alpha=0.01, beta=0.01, gamma=0.001, inverse=True,
             lpow=1.0, eigen=False, divfree=False, dtype=np.float64):
    """DiffOper L or L^-1 of an AoS field vf[z,y,x,3]; returns a new array."""
    assert vf.ndim == 4 and vf.shape[3] == 3
    flags = (DO_EIGEN if eigen else 0) | (DO_DIVFREE if divfree else 0)
    out = np.ascontiguousarray(vf, dtype=dtype).copy()
    fn = lib().orc_diffoper_d if dtype == np.float64 else lib().orc_diffoper_f
    rc = fn(_size(vf.shape[:3]), _sp(sp), ctypes.c_double(alpha), ctypes.c_double(beta),
            ctypes.c_double(gamma), ctypes.c_double(lpow), int(flags), _p(out), int(inverse))
    if rc != 0:
        raise ValueError("oracle diffoper failed")
    return out


def apply(img, h, bg=BG_VAL, background=0.0, dtype=np.float64):
    img = _f32(img)
    h = _f32(h)
    out = np.zeros(img.shape, dtype=dtype)
    fn = lib().orc_apply_d if dtype == np.float64 else lib().orc_apply_f
    fn(_p(img), _p(h), _size(h.shape[:3]), _p(out), int(bg), ctypes.c_double(background))
    return out


def compose(f, g, bg=BG_PARTIAL_ID):
    f, g = _f32(f), _f32(g)
    out = np.zeros(g.shape, dtype=np.float64)
    lib().orc_compose_d(_p(f), _p(g), _size(g.shape[:3]), _p(out), int(bg))
    return out


def compose_hv(g, v, sp=(1, 1, 1), bg=BG_PARTIAL_ID, inv=False):
    g, v = _f32(g), _f32(v)
    out = np.zeros(g.shape, dtype=np.float64)
    lib().orc_compose_hv_d(_p(g), _p(v), _size(g.shape[:3]), _sp(sp), _p(out), int(bg), int(inv))
    return out


def compose_vh(v, g, sp=(1, 1, 1), bg=BG_PARTIAL_ZERO, inv=False):
    v, g = _f32(v), _f32(g)
    out = np.zeros(g.shape, dtype=np.float64)
    lib().orc_compose_vh_d(_p(v), _p(g), _size(g.shape[:3]), _sp(sp), _p(out), int(bg), int(inv))
    return out


def trilerp_vec(h, x, y, z, bg):
    h = _f32(h)
    out = np.zeros(3)
    lib().orc_trilerp_vec_d(_p(h), _size(h.shape[:3]), ctypes.c_double(x), ctypes.c_double(y),
                            ctypes.c_double(z), _p(out), int(bg))
    return out


def trilerp_scalar(img, x, y, z, bg=BG_VAL, background=0.0):
    img = _f32(img)
    return lib().orc_trilerp_scalar_d(_p(img), _size(img.shape), ctypes.c_double(x),
                                      ctypes.c_double(y), ctypes.c_double(z), int(bg),
                                      ctypes.c_double(background))


def forward_apply(img, h, background=0.0, normalize=False):
    img, h = _f32(img), _f32(h)
    out = np.zeros(img.shape, dtype=np.float32)
    lib().orc_forward_apply_f(_p(img), _p(h), _size(img.shape), _p(out),
                              ctypes.c_double(background), int(normalize))
    return out


def splat(src, h):
    """cpuSplat3D: double accumulation; h AoS positions."""
    src = _f32(src)
    h = _f32(h)
    px = np.ascontiguousarray(h[..., 0])
    py = np.ascontiguousarray(h[..., 1])
    pz = np.ascontiguousarray(h[..., 2])
    out = np.zeros(src.shape, dtype=np.float64)
    lib().orc_splat_d(_p(src), _p(px), _p(py), _p(pz), _size(src.shape), _p(out))
    return out


def splat_index(src, v, floor_fix=True):
    """createIndex keys/values (8N each) for positions voxel + v (AoS v)."""
    src = _f32(src)
    vx = np.ascontiguousarray(v[..., 0], dtype=np.float32)
    vy = np.ascontiguousarray(v[..., 1], dtype=np.float32)
    vz = np.ascontiguousarray(v[..., 2], dtype=np.float32)
    n = src.size
    keys = np.zeros(8 * n, dtype=np.uint32)
    vals = np.zeros(8 * n, dtype=np.float32)
    lib().orc_splat_index(_p(src), _p(vx), _p(vy), _p(vz), _size(src.shape), int(floor_fix),
                          _p(keys), _p(vals))
    return keys, vals


def splat_sorted(src, v, floor_fix=False):
    """cplSplating restated: (dst, sorted_keys, sorted_vals); positions voxel + v (AoS v)."""
    src = _f32(src)
    vx = np.ascontiguousarray(v[..., 0], dtype=np.float32)
    vy = np.ascontiguousarray(v[..., 1], dtype=np.float32)
    vz = np.ascontiguousarray(v[..., 2], dtype=np.float32)
    n = src.size
    dst = np.zeros(src.shape, dtype=np.float32)
    keys = np.zeros(8 * n, dtype=np.uint32)
    vals = np.zeros(8 * n, dtype=np.float32)
    lib().orc_splat_sorted(_p(src), _p(vx), _p(vy), _p(vz), _size(src.shape), int(floor_fix),
                           _p(dst), _p(keys), _p(vals))
    return dst, keys, vals


def gradient(img, sp=(1, 1, 1), wrap=False, dtype=np.float64):
    img = _f32(img)
    out = np.zeros(img.shape + (3,), dtype=dtype)
    fn = lib().orc_gradient_d if dtype == np.float64 else lib().orc_gradient_f
    fn(_p(img), _size(img.shape), _sp(sp), int(wrap), _p(out))
    return out


def identity(shape3, dtype=np.float32):
    nz, ny, nx = shape3
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([x, y, z], axis=-1).astype(dtype)


def greedy_run(I0, I1, n_iter, sp=(1, 1, 1), alpha=0.01, beta=0.01, gamma=0.001,
               max_pert=0.1, h0=None, compute_inverse=False, auto_step_reduce=False,
               dtype=np.float32):
    """GreedyWarp single scale level (GreedyWarp.txx:69-125) in float or double.

    Returns (h, hinv or None, energies (reference definition), energies summed in
    double, step)."""
    I0, I1 = _f32(I0), _f32(I1)
    p = GreedyParams()
    nz, ny, nx = I0.shape
    p.size[:] = [nx, ny, nz]
    p.sp[:] = [float(s) for s in sp]
    p.alpha, p.beta, p.gamma, p.max_pert = alpha, beta, gamma, max_pert
    p.n_iter = int(n_iter)
    p.compute_inverse = int(compute_inverse)
    p.auto_step_reduce = int(auto_step_reduce)
    h = identity(I0.shape, dtype) if h0 is None else np.ascontiguousarray(h0, dtype=dtype).copy()
    hinv = identity(I0.shape, dtype) if compute_inverse else None
    ef = np.zeros(n_iter)
    ed = np.zeros(n_iter)
    step = ctypes.c_double(0)
    if dtype == np.float32:
        lib().orc_greedy_run_f(ctypes.byref(p), _p(I0), _p(I1), _p(h),
                               _p(hinv) if hinv is not None else None, _p(ef), _p(ed),
                               ctypes.byref(step))
    else:
        lib().orc_greedy_run_d(ctypes.byref(p), _p(I0), _p(I1), _p(h),
                               _p(hinv) if hinv is not None else None, _p(ed),
                               ctypes.byref(step))
        ef = ed.copy()
    return h, hinv, ef, ed, step.value


class LdmmParams(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int * 3),
        ("sp", ctypes.c_double * 3),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("max_pert", ctypes.c_double),
        ("n_time_steps", ctypes.c_int),
        ("n_iter", ctypes.c_int),
    ]


def _ldmm_params(shape3, n_t, n_iter, sp, alpha, beta, gamma, sigma, max_pert):
    p = LdmmParams()
    nz, ny, nx = shape3
    p.size[:] = [nx, ny, nz]
    p.sp[:] = [float(s) for s in sp]
    p.alpha, p.beta, p.gamma, p.sigma, p.max_pert = alpha, beta, gamma, sigma, max_pert
    p.n_time_steps, p.n_iter = int(n_t), int(n_iter)
    return p


def ldmm_run(I0, I1, n_t, n_iter, v=None, step=0.0, sp=(1, 1, 1), alpha=0.01, beta=0.01,
             gamma=0.001, sigma=0.05, max_pert=0.1, dtype=np.float64):
    """LDMMWarp single scale: IterateNew x n_iter.  v: [T, z, y, x, 3] AoS (zeros if
    None).  step <= 0: one extra call first sets the step size.  Returns (v, e_img,
    e_vec, step)."""
    I0, I1 = _f32(I0), _f32(I1)
    p = _ldmm_params(I0.shape, n_t, n_iter, sp, alpha, beta, gamma, sigma, max_pert)
    v = (np.zeros((n_t,) + I0.shape + (3,), dtype) if v is None
         else np.ascontiguousarray(v, dtype=dtype).copy())
    ei = np.zeros(n_iter)
    ev = np.zeros(n_iter)
    st = ctypes.c_double(step)
    fn = lib().orc_ldmm_run_d if dtype == np.float64 else lib().orc_ldmm_run_f
    fn(ctypes.byref(p), _p(I0), _p(I1), _p(v), _p(ei), _p(ev), ctypes.byref(st))
    return v, ei, ev, st.value


def ldmm_run_orig(I0, I1, n_t, n_iter, v=None, step=0.0, sp=(1, 1, 1), alpha=0.01, beta=0.01,
                  gamma=0.001, sigma=0.05, max_pert=0.1):
    """LDMMWarp single scale with IterateOrig (FP64).  Returns (v, e_img, e_vec, step)."""
    I0, I1 = _f32(I0), _f32(I1)
    p = _ldmm_params(I0.shape, n_t, n_iter, sp, alpha, beta, gamma, sigma, max_pert)
    v = (np.zeros((n_t,) + I0.shape + (3,)) if v is None
         else np.ascontiguousarray(v, dtype=np.float64).copy())
    ei = np.zeros(n_iter)
    ev = np.zeros(n_iter)
    st = ctypes.c_double(step)
    lib().orc_ldmm_run_orig_d(ctypes.byref(p), _p(I0), _p(I1), _p(v), _p(ei), _p(ev),
                              ctypes.byref(st))
    return v, ei, ev, st.value


def jacobian_det(hw, sp=(1, 1, 1)):
    """det of the Jacobian of an AoS field [z, y, x, 3] in world units."""
    hw = np.ascontiguousarray(hw, dtype=np.float64)
    out = np.zeros(hw.shape[:3])
    lib().orc_jacobian_det_d(_p(hw), _size(hw.shape[:3]), _sp(sp), _p(out))
    return out


def ldmm_alphas(I0, I1, v, sp=(1, 1, 1), sigma=0.05):
    """IterateNew(computeAlphasOnly): alpha_t images [T, z, y, x]."""
    I0, I1 = _f32(I0), _f32(I1)
    v = np.ascontiguousarray(v, dtype=np.float64)
    n_t = v.shape[0]
    p = _ldmm_params(I0.shape, n_t, 1, sp, 0.01, 0.01, 0.001, sigma, 0.1)
    a = np.zeros((n_t,) + I0.shape, np.float32)
    lib().orc_ldmm_alphas_d(ctypes.byref(p), _p(I0), _p(I1), _p(v), _p(a))
    return a


def ldmm_alphas_orig(I0, I1, v, sp=(1, 1, 1), sigma=0.05):
    """IterateOrig(computeAlphasOnly): alpha_t images [T, z, y, x]."""
    I0, I1 = _f32(I0), _f32(I1)
    v = np.ascontiguousarray(v, dtype=np.float64)
    n_t = v.shape[0]
    p = _ldmm_params(I0.shape, n_t, 1, sp, 0.01, 0.01, 0.001, sigma, 0.1)
    a = np.zeros((n_t,) + I0.shape, np.float32)
    lib().orc_ldmm_alphas_orig_d(ctypes.byref(p), _p(I0), _p(I1), _p(v), _p(a))
    return a


def ldmm_atlas_run(images, n_t, n_iter, weights=None, v=None, steps=None, jacobian_scale=True,
                   sp=(1, 1, 1), alpha=0.01, beta=0.01, gamma=0.001, sigma=0.05, max_pert=0.1,
                   use_orig=True):
    """LDMM atlas, single scale (FP64).  images [S, z, y, x]; v [S, T, z, y, x, 3].
    Returns (mean, v, steps, energy [n_iter, 3])."""
    images = _f32(images)
    S = images.shape[0]
    shape = images.shape[1:]
    p = _ldmm_params(shape, n_t, n_iter, sp, alpha, beta, gamma, sigma, max_pert)
    w = np.full(S, 1.0 / S) if weights is None else np.ascontiguousarray(weights, np.float64)
    v = (np.zeros((S, n_t) + shape + (3,)) if v is None
         else np.ascontiguousarray(v, dtype=np.float64).copy())
    st = np.zeros(S) if steps is None else np.ascontiguousarray(steps, np.float64).copy()
    mean = np.zeros(shape)
    e = np.zeros((n_iter, 3))
    lib().orc_ldmm_atlas_run_d(ctypes.byref(p), S, _p(images), _p(w), int(bool(jacobian_scale)),
                               int(bool(use_orig)), _p(v), _p(st), _p(mean), _p(e))
    return mean, v, st, e


def ldmm_orig_terms(I1, v, sp=(1, 1, 1)):
    """IterateOrig's backward sweep: (JT, JacDet), each [T, z, y, x]."""
    I1 = _f32(I1)
    v = np.ascontiguousarray(v, dtype=np.float64)
    n_t = v.shape[0]
    p = _ldmm_params(I1.shape, n_t, 1, sp, 0.01, 0.01, 0.001, 0.05, 0.1)
    JT = np.zeros((n_t,) + I1.shape)
    JD = np.zeros((n_t,) + I1.shape)
    lib().orc_ldmm_orig_terms_d(ctypes.byref(p), _p(I1), _p(v), _p(JT), _p(JD))
    return JT, JD


def gaussian_filter(img, sigma, ksize):
    img = _f32(img)
    out = np.zeros(img.shape, np.float32)
    lib().orc_gaussian_filter(_p(img), _size(img.shape), _sp(sigma), (ctypes.c_int * 3)(*ksize),
                              _p(out))
    return out


def gaussian_downsample(img, new_shape):
    """ImageUtils::gaussianDownsample; new_shape as [z, y, x]."""
    img = _f32(img)
    out = np.zeros(new_shape, np.float32)
    lib().orc_gaussian_downsample(_p(img), _size(img.shape), _size(new_shape), _p(out))
    return out


def resample_hfield(h, new_shape, bg=BG_CLAMP, rescale=False):
    """HField3DUtils::resampleNew of an AoS field [z, y, x, 3]."""
    h = _f32(h)
    out = np.zeros(tuple(new_shape) + (3,))
    lib().orc_resample_hfield_d(_p(h), _size(h.shape[:3]), _size(new_shape), int(bg),
                                int(rescale), _p(out))
    return out


class AtlasParams(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int * 3),
        ("sp", ctypes.c_double * 3),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("max_pert", ctypes.c_double),
        ("n_iter", ctypes.c_int),
        ("n_subjects", ctypes.c_int),
    ]


def apply_grid(img, h, background=0.0):
    """apply with the image on its own grid: out has h's grid; h in image voxels."""
    img = _f32(img)
    out = np.zeros(h.shape[:3])
    for i, pt in enumerate(np.asarray(h, np.float64).reshape(-1, 3)):
        out.flat[i] = trilerp_scalar(img, *pt, bg=BG_VAL, background=background)
    return out


def atlas_run(images, h0, weights, n_iter, template_shape, sp=(1, 1, 1), alpha=0.01, beta=0.01,
              gamma=0.001, max_pert=0.1, hinv0=None):
    """GreedyAtlasBuilder, one scale level, FP64.  images: list of [z,y,x] arrays
    (any sizes); h0: list of AoS hfields on the template grid in each image's
    voxels; hinv0: None, or initial inverse fields (Def0To1) for every subject.
    Returns (h list, mean, energies), plus the hinv list when hinv0 is given."""
    p = AtlasParams()
    nz, ny, nx = template_shape
    p.size[:] = [nx, ny, nz]
    p.sp[:] = [float(s) for s in sp]
    p.alpha, p.beta, p.gamma, p.max_pert = alpha, beta, gamma, max_pert
    p.n_iter, p.n_subjects = int(n_iter), len(images)
    ims = [_f32(im) for im in images]
    hs = [np.ascontiguousarray(h, dtype=np.float64).copy() for h in h0]
    I0 = (ctypes.c_void_p * len(ims))(*[im.ctypes.data for im in ims])
    H = (ctypes.c_void_p * len(hs))(*[h.ctypes.data for h in hs])
    isz = np.array([[im.shape[2], im.shape[1], im.shape[0]] for im in ims], dtype=np.int32)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    mean = np.zeros(template_shape)
    energy = np.zeros(n_iter)
    his = None
    HI = None
    if hinv0 is not None:
        his = [np.ascontiguousarray(h, dtype=np.float64).copy() for h in hinv0]
        HI = (ctypes.c_void_p * len(his))(*[h.ctypes.data for h in his])
    lib().orc_atlas_run_d(ctypes.byref(p), I0, _p(isz), _p(w), H, HI, _p(mean), _p(energy))
    if his is not None:
        return hs, mean, energy, his
    return hs, mean, energy


def rescale_intensities(img, out_min=0.0, out_max=1.0):
    """IntensityWindow::Rescale with the image's own min/max as the input window
    (ImagePreprocessor.cxx:39-69, Array3DUtils::rescaleElements)."""
    img = np.asarray(img, np.float32)
    lo, hi = float(img.min()), float(img.max())
    if hi == lo:
        return np.full(img.shape, out_min, np.float32)
    return ((img - lo) * ((out_max - out_min) / (hi - lo)) + out_min).astype(np.float32)


def atlas_multiscale_run(images, weights, levels, sp=(1, 1, 1), alpha=0.01, beta=0.01,
                         gamma=0.001, max_pert=0.1):
    """GreedyAtlasBuilder over scale levels, FP64, every subject on the template
    grid (AtlasThread::AtlasThreadMain, AtlasThread.txx:95-181).  levels: list of
    (downsample factor, iterations), coarsest first.  At factor f the grid is
    size // f with spacing * f (MultiscaleManager::AddScaleLevel,
    MultiscaleManager.cxx:104-111) and the images are Gaussian-downsampled from the
    originals (GreedyDeformationData::SetScaleLevel, GreedyDeformationData.cxx:
    223-240); each hfield is carried to the next level as a voxel displacement
    resampled with CLAMP and rescaled (GreedyDeformationData.cxx:210-214,
    MultiscaleManager.cxx:327-341); the step sizes are recomputed at each level's
    first iteration (GreedyAtlasThreadCPU.cxx:84-91).  sp is the full-resolution
    spacing (x, y, z).  Returns (hfields, mean, deformed images, energies) at the
    last level."""
    shape = np.asarray(images[0]).shape
    hs = None
    energies = []
    for f, n in levels:
        ls = tuple(s // f for s in shape)
        ims = [np.asarray(im, np.float32) if f == 1 else gaussian_downsample(im, ls)
               for im in images]
        ident = identity(ls, np.float64)
        if hs is None:
            hs = [ident.copy() for _ in images]
        else:
            hs = [resample_hfield(h - identity(h.shape[:3], np.float64), ls, bg=BG_CLAMP,
                                  rescale=True) + ident for h in hs]
        spl = tuple(float(s) * f for s in sp)
        hs, mean, en = atlas_run(ims, hs, weights, int(n), ls, sp=spl, alpha=alpha, beta=beta,
                                 gamma=gamma, max_pert=max_pert)
        energies.append(en)
    defs = [apply(im, h) for im, h in zip(images, hs)]
    return hs, mean, defs, energies


def greedy_time(shape3, n_iter=1, seed=0):
    """Seconds per FP32 greedy iteration of the oracle on one host thread."""
    nz, ny, nx = shape3
    size = (ctypes.c_int * 3)(nx, ny, nz)
    return lib().orc_greedy_time_f(size, int(n_iter), int(seed))


def greedy_time_threads(shape3, n_threads, n_iter=1, seed=0):
    """Wall seconds of n_threads concurrent FP32 greedy registrations (one host
    thread each, n_iter iterations each), as the reference's CPU atlas runs one
    thread per image (GreedyAtlasThreadCPU.cxx:19-37)."""
    nz, ny, nx = shape3
    size = (ctypes.c_int * 3)(nx, ny, nz)
    return lib().orc_greedy_time_threads_f(size, int(n_iter), int(seed), int(n_threads))


def fluidwarp_asym(fixed, moving, n_iter=20, alpha=0.01, beta=0.01, gamma=1e-4, max_pert=0.5):
    fixed, moving = _f32(fixed), _f32(moving)
    h = identity(fixed.shape)
    hinv = identity(fixed.shape)
    lib().orc_fluidwarp_asym_f(_p(fixed), _p(moving), _size(fixed.shape), ctypes.c_double(alpha),
                               ctypes.c_double(beta), ctypes.c_double(gamma),
                               ctypes.c_double(max_pert), int(n_iter), _p(h), _p(hinv))
    return h, hinv
