"""Greedy atlas building on the GPU.

Mirrors ``GreedyAtlasBuilder`` / ``AtlasThread`` / ``GreedyAtlasThreadGPU``
(Algorithms/Greedy/GreedyAtlasBuilder.txx:20-301, Algorithms/Common/
AtlasThread.txx:95-181, Algorithms/Greedy/GreedyAtlasThreadGPU.cxx:85-206) for the
subjects held by one device.  Subject images may have different sizes and
spacings: every hfield lives on the template (mean) grid and maps it into its
subject's voxels; ``initial_hfield`` builds that map from the grids' origins and
spacings (and an optional world affine).

Scale levels follow MultiscaleManager (Algorithms/Common/MultiscaleManager.cxx:
104-111, 300-366): at factor f the template grid is size // f with spacing * f,
each subject image is Gaussian-downsampled by the same factor, and the hfields are
carried to the next level as displacements from the grid-to-grid map, resampled
with CLAMP and rescaled (ScaleVectorFields, GreedyAtlasBuilder.txx:34).
"""
import ctypes
import os
import warnings

import numpy as np
import torch

from . import _lib, hfield
from .device import planes_arr, ptr, stream_handle


class AtlasSubject:
    """One input image with its grid (origin, spacing in world units) and weight."""

    def __init__(self, image, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), weight=None,
                 affine=None):
        self.image = image.contiguous()
        self.origin = tuple(float(v) for v in origin)
        self.spacing = tuple(float(v) for v in spacing)
        self.weight = weight
        self.affine = affine  # optional 3x4 world-space map template -> subject


class _AtlasLevel:
    """One awk_atlas handle: template grid, subjects' (level) images, hfields."""

    def __init__(self, images, weights, size, spacing, max_pert, alpha, beta, gamma, use_graph,
                 h0s, hinv0s=None, batch=None):
        self.size = tuple(int(v) for v in size)
        self.images = images  # kept alive: the library holds the pointers
        h = ctypes.c_void_p()
        _lib.call("awk_atlas_create", ctypes.byref(h), *self.size, *[float(s) for s in spacing],
                  alpha, beta, gamma, max_pert, len(images))
        self._h = h
        _lib.call("awk_atlas_set_graph", self._h, int(bool(use_graph)))
        if batch is not None:
            _lib.call("awk_atlas_set_batch", self._h, int(batch))
        for i, (im, w, h0) in enumerate(zip(images, weights, h0s)):
            iz, iy, ix = im.shape
            _lib.call("awk_atlas_set_subject", self._h, i, ptr(im), ix, iy, iz, float(w),
                      planes_arr(h0), stream_handle(None))
            if hinv0s is not None:
                _lib.call("awk_atlas_set_subject_inverse", self._h, i, planes_arr(hinv0s[i]),
                          stream_handle(None))
        torch.cuda.current_stream().synchronize()
        self.has_inverse = hinv0s is not None

    def field(self, i, stream=None):
        nx, ny, nz = self.size
        out = torch.empty((3, nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_atlas_get_hfield", self._h, int(i), planes_arr(out), stream_handle(stream))
        return out

    def inverse(self, i, stream=None):
        nx, ny, nz = self.size
        out = torch.empty((3, nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_atlas_get_hinv", self._h, int(i), planes_arr(out), stream_handle(stream))
        return out

    def close(self):
        if self._h is not None and self._h.value:
            _lib.call("awk_atlas_destroy", self._h)
        self._h = None


class GreedyAtlasBuilderGPU:
    def __init__(self, subjects, template_size, template_origin=(0.0, 0.0, 0.0),
                 template_spacing=(1.0, 1.0, 1.0), max_pert=0.1, alpha=0.01, beta=0.01,
                 gamma=0.001, use_graph=True, scale_levels=None, compute_inverse=False,
                 initial_hfields=None, batch=None):
        if not subjects:
            raise ValueError("an atlas needs at least one subject")
        self.subjects = list(subjects)
        self.size = tuple(int(v) for v in template_size)  # (nx, ny, nz)
        self.origin = tuple(float(v) for v in template_origin)
        self.spacing = tuple(float(v) for v in template_spacing)
        self.prm = (max_pert, alpha, beta, gamma, use_graph)
        # subjects per batched launch of the steady-state greedy step: None = the
        # library's default, 0 = one launch per subject on concurrent stream lanes,
        # an int, or {level factor: int} (autotune.tune_atlas_batch picks these)
        self.batch = batch
        ns = len(self.subjects)
        self.weights = [s.weight if s.weight is not None else 1.0 / ns for s in self.subjects]
        self.scale_levels = list(scale_levels) if scale_levels else None
        # Def0To1 per subject (GreedyDeformationData::ComputeInverseHField); the
        # subjects must share the template grid, as the reference's atlas requires
        self.compute_inverse = bool(compute_inverse)
        if self.compute_inverse:
            for s in self.subjects:
                iz, iy, ix = s.image.shape
                if ((ix, iy, iz) != self.size or s.spacing != self.spacing or
                        s.origin != self.origin or s.affine is not None):
                    raise ValueError("inverse fields need every subject on the template grid")
        # initial hfields on the full-resolution template grid (default: the grid
        # maps, i.e. identity starts as AtlasThread::AtlasThreadMain makes them)
        self.initial_hfields = list(initial_hfields) if initial_hfields is not None else None
        if self.initial_hfields is not None and len(self.initial_hfields) != len(self.subjects):
            raise ValueError("one initial hfield per subject")
        self._energy = []
        self._energy_level = []
        self.flags = 0  # numerical flags of the run (FLAG_* of _lib)
        self._level_index = -1
        self._level = None
        self._factor = None
        if self.scale_levels is None:
            self._enter_level(1)

    # ---- grids of a level
    def _template_grid(self, f):
        size = tuple(n // f for n in self.size)
        return size, tuple(s * f for s in self.spacing)

    def _subject_grid(self, s, f):
        iz, iy, ix = s.image.shape
        return (ix // f, iy // f, iz // f), tuple(v * f for v in s.spacing)

    def _base_map(self, s, f):
        _, tsp = self._template_grid(f)
        _, ssp = self._subject_grid(s, f)
        return hfield.grid_affine(self.origin, tsp, s.origin, ssp, s.affine)

    def initial_hfield(self, subject, out, f=1):
        hfield.initializeFromAffine(out, self._base_map(subject, f))
        return out

    def _enter_level(self, f):
        size, sp = self._template_grid(f)
        nx, ny, nz = size
        images, h0s, hinv0s = [], [], []
        for s in self.subjects:
            if f == 1:
                im = s.image
            else:
                (sx, sy, sz), _ = self._subject_grid(s, f)
                im = torch.empty((sz, sy, sx), dtype=torch.float32, device=s.image.device)
                hfield.gaussianDownsample(im, s.image)
            images.append(im)
        for i, s in enumerate(self.subjects):
            if self._level is None and self.initial_hfields is not None:
                g = self.initial_hfields[i]
                if tuple(g.shape) != (3,) + tuple(reversed(self.size)):
                    raise ValueError("initial hfields live on the template grid")
                h0 = g.clone() if f == 1 else hfield.upsampleHField(
                    g, size, rescale=True, base_in=self._base_map(s, 1), base_out=self._base_map(s, f))
                if self.compute_inverse:
                    raise ValueError("initial hfields with compute_inverse are not supported")
            elif self._level is None:
                h0 = torch.empty((3, nz, ny, nx), dtype=torch.float32, device="cuda")
                self.initial_hfield(s, h0, f)
                if self.compute_inverse:
                    hi = torch.empty_like(h0)
                    hfield.setToIdentity(hi)
                    hinv0s.append(hi)
            else:
                h0 = hfield.upsampleHField(self._level.field(i), size, rescale=True,
                                           base_in=self._base_map(s, self._factor),
                                           base_out=self._base_map(s, f))
                if self.compute_inverse:
                    # GreedyDeformationData::SetScaleLevel: Def0To1 upsampled as Def1To0
                    hinv0s.append(hfield.upsampleHField(self._level.inverse(i), size, rescale=True))
            h0s.append(h0)
        if self._level is not None:
            self._level.close()
        max_pert, alpha, beta, gamma, use_graph = self.prm
        self._level_index += 1
        self._level = _AtlasLevel(images, self.weights, size, sp, max_pert, alpha, beta, gamma,
                                  use_graph, h0s, hinv0s if self.compute_inverse else None,
                                  self.batch.get(f) if isinstance(self.batch, dict) else self.batch)
        self._factor = f

    @property
    def _h(self):
        return self._level._h

    def UpdateStepSizeNextIteration(self):
        _lib.call("awk_atlas_update_step_next", self._h)

    def Iterate(self, n_iter=1, stream=None):
        """n_iter atlas iterations; energies are kept on the device."""
        e = torch.empty(max(n_iter, 1), dtype=torch.float64, device="cuda")
        _lib.call("awk_atlas_iterate", self._h, int(n_iter), ctypes.c_void_p(e.data_ptr()),
                  stream_handle(stream))
        self._energy.append(e[:n_iter])
        self._energy_level.append(self._level_index)

    def RunScaleLevel(self, n_iter):
        """One scale level: step sizes reset, then n_iter iterations.  A NaN energy
        is reported as GreedyDeformationData::AddEnergy logs it
        (GreedyDeformationData.cxx:170-174): a warning, and the run goes on."""
        self.UpdateStepSizeNextIteration()
        self.Iterate(n_iter)
        flags = self.GetFlags()
        self.flags |= flags
        if flags & _lib.FLAG_NAN_ENERGY:
            warnings.warn("NaN detected in an atlas deformation (scale level %d)" % self._level_index,
                          RuntimeWarning)

    def GetFlags(self, clear=True, stream=None):
        """Numerical flags of all subjects since the last clear (FLAG_* of _lib)."""
        return _lib.get_flags("awk_atlas_get_flags", self._h, clear, stream)[0]

    def Run(self):
        """All scale levels (AtlasThread::AtlasThreadMain)."""
        if self.scale_levels is None:
            raise ValueError("no scale levels given; use RunScaleLevel")
        for f, n_iter in self.scale_levels:
            if self._factor != f or self._level is None:
                self._enter_level(int(f))
            self.RunScaleLevel(int(n_iter))

    def energies(self):
        if not self._energy:
            return torch.empty(0, dtype=torch.float64).numpy()
        return torch.cat(self._energy).cpu().numpy()

    def EnergyHistory(self):
        """The per-iteration atlas energies as the reference's EnergyHistory
        (GreedyAtlasBuilder::FinishIteration and WriteEnergy,
        GreedyAtlasBuilder.txx:340-348, 494-503)."""
        from .energy_history import EnergyHistory
        h = EnergyHistory()
        count = {}
        for lvl, e in zip(self._energy_level, self._energy):
            for v in e.cpu().numpy():
                h.add_iteration(lvl, count.get(lvl, 0), float(v))
                count[lvl] = count.get(lvl, 0) + 1
        return h

    def level_size(self):
        return self._level.size

    def Mean(self, stream=None):
        nx, ny, nz = self._level.size
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_atlas_mean", self._h, ptr(out), stream_handle(stream))
        return out

    def HField(self, i, stream=None):
        return self._level.field(i, stream)

    def HFieldInverse(self, i, stream=None):
        """Def0To1 of subject i (compute_inverse=True)."""
        if not self.compute_inverse:
            raise ValueError("the builder was created without compute_inverse")
        return self._level.inverse(i, stream)

    def DeformedImage(self, i, stream=None):
        nx, ny, nz = self._level.size
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_atlas_deformed_image", self._h, int(i), ptr(out), stream_handle(stream))
        return out

    def WriteOutputs(self, out_dir=".", prefix="Greedy", suffix="nhdr", names=None, mean=True,
                     def_images=True, def_fields_mean_to_im=False, def_fields_im_to_mean=False,
                     energy=False):
        """The reference's output files (GreedyAtlasBuilder::GenerateOutput and its
        writers, GreedyAtlasBuilder.txx:395-590), without ITK:
        ``<prefix>MeanImage.<suffix>``, ``<prefix><name>DefToMean.<suffix>``,
        ``<prefix><name>DefFieldMeanToIm.<suffix>`` (Def1To0) and
        ``<prefix><name>DefFieldImToMean.<suffix>`` (Def0To1, compute_inverse=True) as
        voxel-coordinate hfields on the template grid (SaveHFieldITK writes h as
        stored), and ``<prefix>Energy.xml``.  suffix: nhdr / nrrd (detached NRRD) or
        mha / mhd (MetaImage).  names: per-subject names, the reference's input file
        base names (default Image00, Image01, ...).  Returns the paths written."""
        from .metaimage import Image, write_metaimage, write_nrrd
        if suffix in ("nhdr", "nrrd"):
            write = write_nrrd
            suffix = "nhdr"
        elif suffix in ("mha", "mhd"):
            write = write_metaimage
            suffix = "mha"
        else:
            raise ValueError("suffix must be nhdr, nrrd, mha or mhd")
        ns = len(self.subjects)
        names = list(names) if names is not None else ["Image%02d" % i for i in range(ns)]
        if len(names) != ns:
            raise ValueError("one name per subject")
        f = self._factor or 1
        sp = tuple(s * f for s in self.spacing)
        os.makedirs(out_dir, exist_ok=True)
        paths = []

        def put(fname, arr):
            path = os.path.join(out_dir, fname)
            write(path, Image(arr, sp, self.origin))
            paths.append(path)

        if mean:
            put("%sMeanImage.%s" % (prefix, suffix), self.Mean().cpu().numpy())
        for i, name in enumerate(names):
            if def_images:
                put("%s%sDefToMean.%s" % (prefix, name, suffix), self.DeformedImage(i).cpu().numpy())
            if def_fields_mean_to_im:
                h = self.HField(i).cpu().numpy()
                put("%s%sDefFieldMeanToIm.%s" % (prefix, name, suffix), np.moveaxis(h, 0, -1))
            if def_fields_im_to_mean:
                h = self.HFieldInverse(i).cpu().numpy()
                put("%s%sDefFieldImToMean.%s" % (prefix, name, suffix), np.moveaxis(h, 0, -1))
        if energy:
            path = os.path.join(out_dir, "%sEnergy.xml" % prefix)
            self.EnergyHistory().save_xml(path)
            paths.append(path)
        return paths

    def __del__(self):
        try:
            if self._level is not None:
                self._level.close()
            self._level = None
        except Exception:
            pass
