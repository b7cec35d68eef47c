"""Greedy-fluid registration on the GPU.

Mirrors the reference driver layer for the greedy path:
``GreedyIteratorGPU`` (Algorithms/Greedy/GreedyIteratorGPU.cxx:24-135),
``GreedyDeformationDataGPU`` (per-registration device state) and
``GreedyWarp<IteratorT>::RunWarp`` (Algorithms/Greedy/GreedyWarp.txx:69-125).
Each ``Iterate`` is one call of ``awk_greedy_iterate`` (six HIP kernels, no host
sync); ``IterateN`` runs several iterations in one call of ``awk_greedy_iterate_n``,
where each composition also forms the next iteration's deformed image (five
kernels per iteration after the first, the same results bit for bit).  Energies
stay on the device until ``energies()`` is read.
"""
import ctypes
import warnings

import torch

from . import _lib
from .device import dims, planes_arr, ptr, stream_handle
from . import hfield


class GreedyIteratorParam:
    """GreedyIteratorParam (GreedyIteratorParam.h:20-40) + DiffOperParam."""

    def __init__(self, max_pert=0.1, alpha=0.01, beta=0.01, gamma=0.001):
        self.MaxPert = max_pert
        self.alpha, self.beta, self.gamma = alpha, beta, gamma


class GreedyDeformationDataGPU:
    """Device state of one registration: I0 (moving), I1 (fixed), Def1To0, Def0To1."""

    def __init__(self, I0, I1, spacing=(1.0, 1.0, 1.0), compute_inverse=False, max_iters=4096,
                 h0=None):
        if I0.shape != I1.shape:
            raise ValueError("I0 and I1 must have the same size")
        self.I0, self.I1 = I0, I1
        self.spacing = tuple(float(s) for s in spacing)
        nz, ny, nx = I0.shape
        dev = I0.device
        self.Def1To0 = torch.empty((3, nz, ny, nx), dtype=torch.float32, device=dev)
        self._hbuf = torch.empty_like(self.Def1To0)
        if h0 is None:
            hfield.setToIdentity(self.Def1To0)
        else:
            self.Def1To0.copy_(h0)
        self.compute_inverse = compute_inverse
        self.Def0To1 = None
        self._hinvbuf = None
        if compute_inverse:
            self.Def0To1 = torch.empty_like(self.Def1To0)
            self._hinvbuf = torch.empty_like(self.Def1To0)
            hfield.setToIdentity(self.Def0To1)
        self._energy = torch.zeros(max_iters, dtype=torch.float64, device=dev)
        self.cur_iter = 0

    def energy_slot(self, n=1):
        """Device slots for the energies of the next n iterations."""
        if self.cur_iter + n > self._energy.numel():
            size = max(2 * self._energy.numel(), self.cur_iter + n)
            grown = torch.zeros(size, dtype=torch.float64, device=self._energy.device)
            grown[: self._energy.numel()] = self._energy
            self._energy = grown
        return self._energy[self.cur_iter:self.cur_iter + n]

    def energies(self):
        """Per-iteration energies (host numpy, synchronises)."""
        return self._energy[: self.cur_iter].cpu().numpy()


class GreedyIteratorGPU:
    def __init__(self, size, spacing=(1.0, 1.0, 1.0), param=None):
        self.param = param or GreedyIteratorParam()
        self._h = None
        self.SetScaleLevel(size, spacing, self.param)

    def SetScaleLevel(self, size, spacing, param=None):
        if param is not None:
            self.param = param
        self._destroy()
        self.size = tuple(int(s) for s in size)
        self.spacing = tuple(float(s) for s in spacing)
        h = ctypes.c_void_p()
        p = self.param
        _lib.call("awk_greedy_create", ctypes.byref(h), *self.size, *self.spacing, p.alpha, p.beta,
                  p.gamma, p.MaxPert)
        self._h = h

    def UpdateStepSizeNextIteration(self):
        _lib.call("awk_greedy_update_step_next", self._h)

    def GetFlags(self, clear=True, stream=None):
        """Numerical flags since the last clear (FLAG_* of _lib); waits for the stream."""
        return _lib.get_flags("awk_greedy_get_flags", self._h, clear, stream)[0]

    def SetAutoStepReduce(self, on):
        _lib.call("awk_greedy_set_auto_step_reduce", self._h, int(bool(on)))

    def StepSize(self, stream=None):
        s = ctypes.c_float()
        _lib.call("awk_greedy_get_step", self._h, ctypes.byref(s), stream_handle(stream))
        return s.value

    def SetStepSize(self, step, stream=None):
        _lib.call("awk_greedy_set_step", self._h, float(step), stream_handle(stream))

    STAGES = ("apply", "xfwd_force", "y_fwd", "z_solve", "y_inv", "xinv_compose",
              "xinv_compose_apply")

    def SetProfiling(self, on):
        _lib.call("awk_greedy_set_profiling", self._h, int(bool(on)))

    def GetProfile(self, launches=False):
        """{stage: total ms} over the iterations recorded since the last read, and the
        iteration count; with launches=True also {stage: timed launches}."""
        ms = (ctypes.c_double * len(self.STAGES))()
        cnt = (ctypes.c_int * len(self.STAGES))()
        n = ctypes.c_int()
        _lib.call("awk_greedy_get_profile", self._h, ms, cnt, ctypes.byref(n))
        if launches:
            return dict(zip(self.STAGES, list(ms))), n.value, dict(zip(self.STAGES, list(cnt)))
        return dict(zip(self.STAGES, list(ms))), n.value

    def Iterate(self, data, stream=None):
        if dims(data.I0) != self.size:
            raise ValueError("deformation data size does not match the iterator")
        hinv_in = planes_arr(data.Def0To1) if data.compute_inverse else None
        hinv_out = planes_arr(data._hinvbuf) if data.compute_inverse else None
        e = data.energy_slot()
        _lib.call("awk_greedy_iterate", self._h, ptr(data.I0), ptr(data.I1),
                  planes_arr(data.Def1To0), planes_arr(data._hbuf), hinv_in, hinv_out,
                  ctypes.c_void_p(e.data_ptr()), stream_handle(stream))
        # ping-pong instead of copying the update back
        data.Def1To0, data._hbuf = data._hbuf, data.Def1To0
        if data.compute_inverse:
            data.Def0To1, data._hinvbuf = data._hinvbuf, data.Def0To1
        data.cur_iter += 1

    def IterateN(self, data, n, stream=None):
        """n iterations back to back (awk_greedy_iterate_n): the same fields and
        energies as n calls of Iterate.  Not with an inverse field."""
        if dims(data.I0) != self.size:
            raise ValueError("deformation data size does not match the iterator")
        if data.compute_inverse:
            raise ValueError("IterateN does not update Def0To1; use Iterate")
        if n <= 0:
            return
        e = data.energy_slot(n)
        _lib.call("awk_greedy_iterate_n", self._h, ptr(data.I0), ptr(data.I1),
                  planes_arr(data.Def1To0), planes_arr(data._hbuf), int(n),
                  ctypes.c_void_p(e.data_ptr()), stream_handle(stream))
        if n % 2:
            data.Def1To0, data._hbuf = data._hbuf, data.Def1To0
        data.cur_iter += n

    def _destroy(self):
        if self._h is not None:
            _lib.load().awk_greedy_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass


class GreedyWarp:
    """GreedyWarp::RunWarp (GreedyWarp.txx:69-125).

    ``scale_levels``: list of (downsample factor, iterations), coarsest first, as
    MultiscaleManager builds them (size /= f, spacing *= f); None = one level at full
    size with ``n_iterations``.  Between levels the hfield is upsampled as
    GreedyDeformationData::SetScaleLevel does (hfield -> displacement, resampleNew
    with CLAMP and vector rescaling, displacement -> hfield), and the images of every
    level but the finest are Gaussian-downsampled from the originals."""

    def __init__(self, I0, I1, n_iterations=50, spacing=(1.0, 1.0, 1.0), param=None,
                 compute_inverse=False, auto_step_reduce=False, h0=None, scale_levels=None):
        self.I0orig, self.I1orig = I0, I1
        self.spacing = tuple(float(v) for v in spacing)
        self.param = param
        self.compute_inverse = compute_inverse
        self.auto_step_reduce = auto_step_reduce
        self.h0 = h0
        self.levels = list(scale_levels) if scale_levels else [(1, n_iterations)]
        self.n_iterations = n_iterations
        self.energy_history = []
        self.step_history = []   # per level: the step size set at its first iteration
        self.flags = 0           # numerical flags of the run (FLAG_* of _lib)
        self.data = None

    def _level_images(self, f):
        if f == 1:
            return self.I0orig, self.I1orig
        nz, ny, nx = self.I0orig.shape
        out = []
        for im in (self.I0orig, self.I1orig):
            d = torch.empty((nz // f, ny // f, nx // f), dtype=torch.float32, device=im.device)
            hfield.gaussianDownsample(d, im)
            out.append(d)
        return out[0], out[1]

    def RunWarp(self, stream=None):
        prev = None
        for level, (f, n_iter) in enumerate(self.levels):
            I0, I1 = self._level_images(int(f))
            nz, ny, nx = I0.shape
            sp = tuple(s * f for s in self.spacing)
            h0 = self.h0 if level == 0 else hfield.upsampleHField(prev.Def1To0, (nx, ny, nz))
            data = GreedyDeformationDataGPU(I0, I1, sp, self.compute_inverse,
                                            max_iters=max(n_iter, 1), h0=h0)
            if level > 0 and self.compute_inverse:
                data.Def0To1.copy_(hfield.upsampleHField(prev.Def0To1, (nx, ny, nz)))
            it = GreedyIteratorGPU((nx, ny, nz), sp, self.param)
            it.SetAutoStepReduce(self.auto_step_reduce)
            if n_iter > 0:
                it.UpdateStepSizeNextIteration()
                it.Iterate(data, stream)
                self.step_history.append(it.StepSize(stream))
            if self.compute_inverse:
                for k in range(1, n_iter):
                    it.Iterate(data, stream)
            else:
                it.IterateN(data, n_iter - 1, stream)
            flags = it.GetFlags(stream=stream)
            self.flags |= flags
            if flags & _lib.FLAG_NAN_ENERGY:
                # GreedyDeformationData::AddEnergy logs and goes on (GreedyDeformationData.cxx:170-174)
                warnings.warn("NaN detected in deformation (scale level %d)" % level, RuntimeWarning)
            self.energy_history.append(data.energies())
            prev = data
            self.iterator = it
        self.data = prev
        return prev

    def EnergyHistory(self):
        """The run's EnergyHistory as the reference's deformation data collects it
        (``save_xml`` writes GreedyWarp::WriteEnergy's file, GreedyWarp.txx:220-230):
        one IterationEvent per iteration (DeformationDataInterface::AddEnergy), an
        EnergyIncreaseEvent after each iteration whose energy exceeds the previous
        one, across scale levels (GreedyDeformationData.cxx:162-184), with the
        AutoStepReduce halving as a StepSizeChange, and the StepSizeChange that each
        level's first iteration sets (GreedyIteratorCPU.cxx:129-134)."""
        from .energy_history import EnergyHistory
        h = EnergyHistory()
        prev = None
        step = None
        for level, energies in enumerate(self.energy_history):
            for k, e in enumerate(energies):
                h.add_iteration(level, k, e)
                if prev is not None and e - prev > 0:
                    h.add_energy_increase()
                    if self.auto_step_reduce and step is not None:
                        step = step / 2.0
                        h.add_step_size(step)
                prev = e
                if k == 0 and level < len(self.step_history):
                    step = self.step_history[level]
                    h.add_step_size(step)
        return h

    def WriteOutputs(self, name, out_dir=".", prefix="Greedy", suffix="mha", origin=(0.0, 0.0, 0.0),
                     def_image=True, inv_def_image=False, def_field=False, inv_def_field=False,
                     energy=False):
        """GreedyWarp's output files (GreedyWarp.txx:150-230) after RunWarp, without
        ITK: ``<prefix><name>DefImage.<suffix>`` (I0 at Def1To0),
        ``<prefix><name>InvDefImage.<suffix>`` (GetI1At0: I1 at Def0To1 with
        compute_inverse; without it the reference applies I0 through the zeroth-order
        inverse of Def1To0, GreedyDeformationData.cxx:66-79, and so does this),
        ``<prefix><name>DefField1To0`` / ``DefField0To1`` (voxel-coordinate hfields)
        and ``<prefix><name>Energy.xml``.  name: the reference's
        ``<moving base>_to_<static base>``.  suffix: mha / mhd or nhdr / nrrd."""
        import os

        import numpy as np

        from .metaimage import Image, write_metaimage, write_nrrd
        if self.data is None:
            raise ValueError("RunWarp first")
        if suffix in ("nhdr", "nrrd"):
            write, suffix = write_nrrd, "nhdr"
        elif suffix in ("mha", "mhd"):
            write, suffix = write_metaimage, "mha"
        else:
            raise ValueError("suffix must be nhdr, nrrd, mha or mhd")
        d = self.data
        os.makedirs(out_dir, exist_ok=True)
        paths = []

        def put(kind, arr):
            path = os.path.join(out_dir, "%s%s%s.%s" % (prefix, name, kind, suffix))
            write(path, Image(arr, self.spacing, origin))
            paths.append(path)

        if def_image:
            out = torch.empty_like(self.I0orig)
            hfield.apply(out, self.I0orig, d.Def1To0)
            put("DefImage", out.cpu().numpy())
        if inv_def_image:
            out = torch.empty_like(self.I0orig)
            if self.compute_inverse:
                hfield.apply(out, self.I1orig, d.Def0To1)
            else:
                hinv = torch.empty_like(d.Def1To0)
                hfield.computeInverseZerothOrder(d.Def1To0, hinv)
                hfield.apply(out, self.I0orig, hinv)
            put("InvDefImage", out.cpu().numpy())
        if def_field:
            put("DefField1To0", np.moveaxis(d.Def1To0.cpu().numpy(), 0, -1))
        if inv_def_field:
            if not self.compute_inverse:
                raise ValueError("DefField0To1 needs compute_inverse=True")
            put("DefField0To1", np.moveaxis(d.Def0To1.cpu().numpy(), 0, -1))
        if energy:
            path = os.path.join(out_dir, "%s%sEnergy.xml" % (prefix, name))
            self.EnergyHistory().save_xml(path)
            paths.append(path)
        return paths
