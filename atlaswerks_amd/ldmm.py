"""LDMM registration on the GPU.

Mirrors ``LDMMIteratorGPU`` (Algorithms/LDMM/LDMMIteratorGPU.cxx), the
per-registration state of ``LDMMDeformationDataGPU`` (T velocity fields, I0, I1)
and the single-scale loop of ``LDMMWarp::RunWarp`` (Algorithms/LDMM/LDMMWarp.txx:
69-125).  The gradient is ``IterateOrig`` (LDMMIteratorGPU.cxx:295-441, the
Jacobian-determinant form, which the reference's constructor selects) unless
``use_orig_gradient=False`` picks ``IterateNew`` (LDMMIteratorGPU.cxx:159-293,
the adjoint splat).  Each ``Iterate`` is one ``awk_ldmm_iterate`` call: 2T fused
gather kernels, T Navier solves with the force and the velocity update fused
into the FFT x passes, no host synchronisation.
"""
import ctypes

import torch

from . import _lib, hfield
from .device import ptr, stream_handle


SPLAT_MODES = {"fixed": 0, "sorted": 1}


class LDMMIteratorParam:
    """LDMMIteratorParam (LDMMIteratorParam.h) + DiffOperParam."""

    def __init__(self, sigma=0.05, max_pert=0.1, alpha=0.01, beta=0.01, gamma=0.001,
                 use_adaptive_step_size=True, step_size=0.005, use_orig_gradient=True,
                 splat="fixed"):
        self.UseOrigGradientCalc = use_orig_gradient  # LDMMIteratorGPU.cxx:37
        if splat not in SPLAT_MODES:
            raise ValueError("splat must be one of %s" % sorted(SPLAT_MODES))
        self.Splat = splat  # IterateNew's splat: int64 fixed point or sort/scan
        self.Sigma = sigma
        self.MaxPert = max_pert
        self.alpha, self.beta, self.gamma = alpha, beta, gamma
        self.UseAdaptiveStepSize = use_adaptive_step_size
        self.StepSize = step_size


class LDMMDeformationDataGPU:
    """I0 (moving), I1 (fixed) and the velocity fields v_t, each [3, nz, ny, nx]."""

    def __init__(self, I0, I1, n_time_steps, spacing=(1.0, 1.0, 1.0), max_iters=4096):
        if I0.shape != I1.shape:
            raise ValueError("I0 and I1 must have the same size")
        self.I0, self.I1 = I0.contiguous(), I1.contiguous()
        self.spacing = tuple(float(s) for s in spacing)
        nz, ny, nx = I0.shape
        self.v = torch.zeros((n_time_steps, 3, nz, ny, nx), dtype=torch.float32, device=I0.device)
        self._energy = torch.zeros((max_iters, 3), dtype=torch.float64, device=I0.device)
        self.cur_iter = 0

    def planes(self):
        T = self.v.shape[0]
        arr = (ctypes.c_void_p * (3 * T))()
        for t in range(T):
            for c in range(3):
                arr[3 * t + c] = self.v[t, c].data_ptr()
        return arr

    def energy_slot(self):
        if self.cur_iter >= self._energy.shape[0]:
            grown = torch.zeros((2 * self._energy.shape[0], 3), dtype=torch.float64,
                                device=self._energy.device)
            grown[: self._energy.shape[0]] = self._energy
            self._energy = grown
        return self._energy[self.cur_iter]

    def energies(self):
        """[iterations, 3]: image, vector and total energy (host numpy, synchronises)."""
        return self._energy[: self.cur_iter].cpu().numpy()


class LDMMIteratorGPU:
    def __init__(self, size, spacing, n_time_steps, param=None):
        self.param = param or LDMMIteratorParam()
        self._h = None
        self.SetScaleLevel(size, spacing, n_time_steps, self.param)

    def SetScaleLevel(self, size, spacing, n_time_steps, param=None):
        if param is not None:
            self.param = param
        self._destroy()
        self.size = tuple(int(s) for s in size)
        self.spacing = tuple(float(s) for s in spacing)
        self.n_time_steps = int(n_time_steps)
        h = ctypes.c_void_p()
        p = self.param
        _lib.call("awk_ldmm_create", ctypes.byref(h), *self.size, *self.spacing, p.alpha, p.beta,
                  p.gamma, self.n_time_steps, p.Sigma, p.MaxPert)
        self._h = h
        _lib.call("awk_ldmm_set_gradient", self._h, int(bool(p.UseOrigGradientCalc)))
        _lib.call("awk_ldmm_set_splat", self._h, SPLAT_MODES[p.Splat])
        if not p.UseAdaptiveStepSize:
            _lib.call("awk_ldmm_set_step", self._h, float(p.StepSize), stream_handle(None))

    def UpdateStepSizeNextIteration(self):
        _lib.call("awk_ldmm_update_step_next", self._h)

    def StepSize(self, stream=None):
        out = ctypes.c_float()
        _lib.call("awk_ldmm_get_step", self._h, ctypes.byref(out), stream_handle(stream))
        return out.value

    def GetFlags(self, clear=True, stream=None):
        """Numerical flags since the last clear (FLAG_* of _lib); waits for the stream."""
        return _lib.get_flags("awk_ldmm_get_flags", self._h, clear, stream)[0]

    def CheckFinite(self, stream=None):
        """Raises _lib.NumericalError when an energy was NaN or a step size invalid
        since the last check, as the reference's iterator throws
        (LDMMIteratorGPU.cxx:238-241, 283-285, 439-441)."""
        _lib.call("awk_ldmm_get_flags", self._h, ctypes.byref(ctypes.c_uint()), 1,
                  stream_handle(stream))

    def _check(self, data):
        nz, ny, nx = data.I0.shape
        if (nx, ny, nz) != self.size or data.v.shape[0] != self.n_time_steps:
            raise ValueError("deformation data does not match the iterator's grid")

    def Iterate(self, data, vec_energy=False, record_energy=True, stream=None):
        self._check(data)
        e = data.energy_slot() if record_energy else None
        _lib.call("awk_ldmm_iterate", self._h, ptr(data.I0), ptr(data.I1), data.planes(),
                  ctypes.c_void_p(e.data_ptr()) if e is not None else None, int(bool(vec_energy)),
                  stream_handle(stream))
        if record_energy:
            data.cur_iter += 1

    def ComputeAlphas(self, data, stream=None):
        """Iterate(computeAlphasOnly=true): [T, nz, ny, nx] alpha images."""
        self._check(data)
        alpha = torch.empty((self.n_time_steps,) + tuple(data.I0.shape), dtype=torch.float32,
                            device=data.I0.device)
        arr = (ctypes.c_void_p * self.n_time_steps)(*[alpha[t].data_ptr() for t in range(self.n_time_steps)])
        _lib.call("awk_ldmm_alphas", self._h, ptr(data.I0), ptr(data.I1), data.planes(), arr,
                  stream_handle(stream))
        return alpha

    def DeformedImage(self, t, stream=None):
        """J_t = I0 o h_t of the last iterate."""
        nx, ny, nz = self.size
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_ldmm_deformed_image", self._h, int(t), ptr(out), stream_handle(stream))
        return out

    def _destroy(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("awk_ldmm_destroy", self._h)
        self._h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass


def _downsample(img, f):
    """MultiscaleManager::DownsampleToLevel (ImageUtils::gaussianDownsample); the
    level grid is the original size divided by f (MultiscaleManager.cxx:105-125)."""
    if f == 1:
        return img
    nz, ny, nx = img.shape
    out = torch.empty((nz // f, ny // f, nx // f), dtype=torch.float32, device=img.device)
    hfield.gaussianDownsample(out, img)
    return out


def _upsample_v(v, shape):
    """MultiscaleManager::UpsampleToLevel of every velocity field: resampleNew with
    CLAMP and, as the LDMM builders set SetScaleVectorFields(false)
    (LDMMAtlasBuilder.txx:39), no rescaling of the vectors (world units)."""
    nz, ny, nx = shape
    out = torch.empty(v.shape[:-3] + (nz, ny, nx), dtype=torch.float32, device=v.device)
    flat_in = v.reshape((-1, 3) + tuple(v.shape[-3:]))
    flat_out = out.reshape((-1, 3, nz, ny, nx))
    for k in range(flat_in.shape[0]):
        hfield.resampleNew(flat_out[k], flat_in[k], hfield.BACKGROUND_STRATEGY_CLAMP, False)
    return out


def _levels(scale_levels, n_iter, param):
    """[(factor, iterations, param)] in run order.  Entries (factor, iterations) take
    the common param."""
    if not scale_levels:
        return [(1, n_iter, param)]
    out = []
    for lv in scale_levels:
        f, n = int(lv[0]), int(lv[1])
        out.append((f, n, lv[2] if len(lv) > 2 and lv[2] is not None else param))
    return out


class LDMMWarp:
    """LDMMWarp::RunWarp (LDMMWarp.txx:69-125): per scale level (coarse to fine) an
    extra step-size call when the step is adaptive, then the level's iterations;
    between levels the images are downsampled from the originals and the velocity
    fields upsampled (LDMMDeformationData::SetScaleLevel, LDMMDeformationData.cxx:
    119-161).  scale_levels: [(factor, iterations[, LDMMIteratorParam])]."""

    def __init__(self, I0, I1, n_time_steps=10, n_iter=40, spacing=(1.0, 1.0, 1.0), param=None,
                 scale_levels=None):
        self.I0orig, self.I1orig = I0.contiguous(), I1.contiguous()
        self.n_time_steps = int(n_time_steps)
        self.spacing = tuple(float(s) for s in spacing)
        self.param = param or LDMMIteratorParam()
        self.levels = _levels(scale_levels, n_iter, self.param)
        self.n_iter = n_iter
        self.energy_history = []
        f0, _, p0 = self.levels[0]
        self._enter(f0, p0, None)

    def _enter(self, f, prm, v_prev):
        I0, I1 = _downsample(self.I0orig, f), _downsample(self.I1orig, f)
        nz, ny, nx = I0.shape
        sp = tuple(s * f for s in self.spacing)
        self.data = LDMMDeformationDataGPU(I0, I1, self.n_time_steps, sp)
        if v_prev is not None:
            self.data.v.copy_(_upsample_v(v_prev, (nz, ny, nx)))
        self.iterator = LDMMIteratorGPU((nx, ny, nz), sp, self.n_time_steps, prm)

    def RunWarp(self, vec_energy=True):
        for k, (f, n, prm) in enumerate(self.levels):
            if k > 0:
                self._enter(f, prm, self.data.v)
            if prm.UseAdaptiveStepSize:
                self.iterator.UpdateStepSizeNextIteration()
                self.iterator.Iterate(self.data, record_energy=False)
            for _ in range(n):
                self.iterator.Iterate(self.data, vec_energy=vec_energy)
            self.iterator.CheckFinite()
            self.energy_history.append(self.data.energies())
        return self.data


class LDMMAtlasBuilderGPU:
    """LDMMAtlasBuilder (Algorithms/LDMM/LDMMAtlasBuilder.txx) with LDMMAtlasThreadGPU
    (LDMMAtlasThreadGPU.cxx) on one GPU: every subject is an LDMM registration of
    the mean to the subject image (IterateOrig unless the param says otherwise);
    the mean is re-estimated after every iteration, scaled by the Jacobian
    determinants when ``jacobian_scale`` (LDMMAtlasParam JacobianScale, on by
    default).  ``awk_ldmm_atlas_*``, one handle per scale level.

    scale_levels: [(factor, iterations[, LDMMIteratorParam])] in run order (coarse
    to fine); ``Run()`` runs them, each starting with the subjects' step-size calls
    when adaptive, the images downsampled from the originals and the velocity
    fields upsampled in between (LDMMDeformationData::SetScaleLevel)."""

    def __init__(self, images, weights=None, n_time_steps=5, spacing=(1.0, 1.0, 1.0), param=None,
                 jacobian_scale=True, scale_levels=None):
        if not images:
            raise ValueError("need at least one image")
        shape = tuple(images[0].shape)
        if any(tuple(im.shape) != shape for im in images):
            raise ValueError("all images must be on the template grid")
        S = len(images)
        self.param = param or LDMMIteratorParam()
        self.weights = [1.0 / S] * S if weights is None else [float(w) for w in weights]
        if len(self.weights) != S:
            raise ValueError("one weight per image")
        self.spacing = tuple(float(x) for x in spacing)
        self.n_time_steps = int(n_time_steps)
        self.jacobian_scale = bool(jacobian_scale)
        self.images_orig = [im.contiguous() for im in images]
        self.levels = _levels(scale_levels, 0, self.param)
        self._h = None
        self._energy = []
        self._energy_level = []
        self._level = -1
        self._enter(0)

    def _enter(self, k):
        f, _, p = self.levels[k]
        ims = [_downsample(im, f) for im in self.images_orig]
        nz, ny, nx = ims[0].shape
        S = len(ims)
        v_prev = getattr(self, "v", None)
        self.v = torch.zeros((S, self.n_time_steps, 3, nz, ny, nx), dtype=torch.float32,
                             device=ims[0].device)
        if v_prev is not None:
            self.v.copy_(_upsample_v(v_prev, (nz, ny, nx)))
        self._destroy()
        self.images = ims
        self.size = (nx, ny, nz)
        sp = tuple(x * f for x in self.spacing)
        h = ctypes.c_void_p()
        _lib.call("awk_ldmm_atlas_create", ctypes.byref(h), nx, ny, nz, *sp, p.alpha, p.beta,
                  p.gamma, self.n_time_steps, p.Sigma, p.MaxPert, S, int(self.jacobian_scale))
        self._h = h
        _lib.call("awk_ldmm_atlas_set_gradient", self._h, int(bool(p.UseOrigGradientCalc)))
        _lib.call("awk_ldmm_atlas_set_splat", self._h, SPLAT_MODES[p.Splat])
        for s in range(S):
            arr = (ctypes.c_void_p * (3 * self.n_time_steps))()
            for t in range(self.n_time_steps):
                for c in range(3):
                    arr[3 * t + c] = self.v[s, t, c].data_ptr()
            _lib.call("awk_ldmm_atlas_set_subject", self._h, s, ptr(ims[s]), self.weights[s], arr,
                      None)
        if not p.UseAdaptiveStepSize:
            _lib.call("awk_ldmm_atlas_set_step", self._h, float(p.StepSize), stream_handle(None))
        self._level = k

    def Run(self, vec_energy=True):
        """Every scale level (AtlasThread::AtlasThreadMain)."""
        for k, (_, n, _) in enumerate(self.levels):
            if k != self._level:
                self._enter(k)
            self.Iterate(n, vec_energy=vec_energy)
            self.CheckFinite()

    def GetFlags(self, clear=True, stream=None):
        """Numerical flags of all subjects since the last clear (FLAG_* of _lib)."""
        return _lib.get_flags("awk_ldmm_atlas_get_flags", self._h, clear, stream)[0]

    def CheckFinite(self, stream=None):
        """Raises _lib.NumericalError on a NaN energy or an invalid step size."""
        _lib.call("awk_ldmm_atlas_get_flags", self._h, ctypes.byref(ctypes.c_uint()), 1,
                  stream_handle(stream))

    def Iterate(self, n_iter=1, vec_energy=True, stream=None):
        e = torch.zeros((n_iter, 3), dtype=torch.float64, device=self.v.device)
        _lib.call("awk_ldmm_atlas_iterate", self._h, int(n_iter), ctypes.c_void_p(e.data_ptr()),
                  int(bool(vec_energy)), stream_handle(stream))
        self._energy.append(e)
        self._energy_level.append(self._level)

    def Mean(self, stream=None):
        nx, ny, nz = self.size
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device=self.v.device)
        _lib.call("awk_ldmm_atlas_mean", self._h, ptr(out), stream_handle(stream))
        return out

    def DefToMean(self, s, t, stream=None):
        """Subject s at time t, I_s(h_t) (GetI1AtT; the DefToMeanTime images)."""
        nx, ny, nz = self.size
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device=self.v.device)
        _lib.call("awk_ldmm_atlas_deformed_image", self._h, int(s), int(t), ptr(out),
                  stream_handle(stream))
        return out

    def WriteOutputs(self, out_dir=".", prefix="LDMMAtlas", suffix="mha", names=None,
                     origin=(0.0, 0.0, 0.0), mean=True, def_to_mean_time=True, vel_fields=False,
                     energy=False):
        """LDMMAtlasBuilder's output files (its writers, LDMMAtlasBuilder.txx:566-775),
        without ITK: ``<prefix>MeanImage.<suffix>``, every subject's image at every
        time step ``<prefix><name>DefToMeanTime<tt>.<suffix>`` (WriteIntermediateImages,
        GetI1AtT), the velocity fields ``<prefix><name>VelFieldTime<tt>.<suffix>``
        (WriteVelFields: v as a 3-vector image) and ``<prefix>Energy.xml``, on the
        current level's grid.  suffix: mha / mhd (MetaImage) or nhdr / nrrd.  names:
        the reference's input file base names (default Image00, ...); origin: the
        images' origin (the iteration does not use it).  Returns the paths written."""
        import os

        from .metaimage import Image, write_metaimage, write_nrrd
        if suffix in ("nhdr", "nrrd"):
            write, suffix = write_nrrd, "nhdr"
        elif suffix in ("mha", "mhd"):
            write, suffix = write_metaimage, "mha"
        else:
            raise ValueError("suffix must be nhdr, nrrd, mha or mhd")
        S = len(self.images_orig)
        names = list(names) if names is not None else ["Image%02d" % i for i in range(S)]
        if len(names) != S:
            raise ValueError("one name per subject")
        f = self.levels[self._level][0]
        sp = tuple(x * f for x in self.spacing)
        os.makedirs(out_dir, exist_ok=True)
        paths = []

        def put(fname, arr):
            path = os.path.join(out_dir, fname)
            write(path, Image(arr, sp, origin))
            paths.append(path)

        if mean:
            put("%sMeanImage.%s" % (prefix, suffix), self.Mean().cpu().numpy())
        for s, name in enumerate(names):
            for t in range(self.n_time_steps):
                if def_to_mean_time:
                    put("%s%sDefToMeanTime%02d.%s" % (prefix, name, t, suffix),
                        self.DefToMean(s, t).cpu().numpy())
                if vel_fields:
                    put("%s%sVelFieldTime%02d.%s" % (prefix, name, t, suffix),
                        self.v[s, t].permute(1, 2, 3, 0).contiguous().cpu().numpy())
        if energy:
            path = os.path.join(out_dir, "%sEnergy.xml" % prefix)
            self.EnergyHistory().save_xml(path)
            paths.append(path)
        return paths

    def StepSize(self, s, stream=None):
        out = ctypes.c_float()
        _lib.call("awk_ldmm_atlas_get_step", self._h, int(s), ctypes.byref(out),
                  stream_handle(stream))
        return out.value

    def EnergyHistory(self):
        """The per-iteration weighted energies as the reference's EnergyHistory
        (LDMMAtlasBuilder::WriteEnergy, LDMMAtlasBuilder.txx:684-693); the
        VecEnergy element carries the mean over the time steps."""
        from .energy_history import EnergyHistory, LDMMEnergy
        h = EnergyHistory()
        count = {}
        for lvl, e in zip(self._energy_level, self._energy):
            for ei, ev, _ in e.cpu().numpy():
                h.add_iteration(lvl, count.get(lvl, 0), LDMMEnergy(ei, ev, self.n_time_steps))
                count[lvl] = count.get(lvl, 0) + 1
        return h

    def energies(self):
        """[iterations, 3]: weighted image, vector and total energy (synchronises)."""
        if not self._energy:
            return torch.zeros((0, 3), dtype=torch.float64).numpy()
        return torch.cat(self._energy).cpu().numpy()

    def _destroy(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("awk_ldmm_atlas_destroy", self._h)
        self._h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass
