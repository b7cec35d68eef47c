"""Greedy atlas building on the GPU.

Mirrors ``GreedyAtlasBuilder`` / ``AtlasThread`` / ``GreedyAtlasThreadGPU``
(Algorithms/Greedy/GreedyAtlasBuilder.txx:20-301, Algorithms/Common/
AtlasThread.txx:95-181, Algorithms/Greedy/GreedyAtlasThreadGPU.cxx:85-206) for the
subjects held by one device.  Subject images may have different sizes and
spacings: every hfield lives on the template (mean) grid and maps it into its
subject's voxels; ``initial_hfield`` builds that map from the grids' origins and
spacings (and an optional world affine).
"""
import ctypes

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


class GreedyAtlasBuilderGPU:
    def __init__(self, subjects, template_size, template_origin=(0.0, 0.0, 0.0),
                 template_spacing=(1.0, 1.0, 1.0), max_pert=0.1, alpha=0.01, beta=0.01,
                 gamma=0.001, use_graph=True):
        if not subjects:
            raise ValueError("an atlas needs at least one subject")
        self.subjects = list(subjects)
        self.size = tuple(int(v) for v in template_size)  # (nx, ny, nz)
        self.origin = tuple(float(v) for v in template_origin)
        self.spacing = tuple(float(v) for v in template_spacing)
        ns = len(self.subjects)
        h = ctypes.c_void_p()
        _lib.call("awk_atlas_create", ctypes.byref(h), *self.size, *self.spacing, alpha, beta,
                  gamma, max_pert, ns)
        self._h = h
        _lib.call("awk_atlas_set_graph", self._h, int(bool(use_graph)))
        nx, ny, nz = self.size
        self._energy = []
        dev = self.subjects[0].image.device
        h0 = torch.empty((3, nz, ny, nx), dtype=torch.float32, device=dev)
        for i, s in enumerate(self.subjects):
            w = s.weight if s.weight is not None else 1.0 / ns
            self.initial_hfield(s, h0)
            iz, iy, ix = s.image.shape
            _lib.call("awk_atlas_set_subject", self._h, i, ptr(s.image), ix, iy, iz, float(w),
                      planes_arr(h0), stream_handle(None))
        torch.cuda.current_stream().synchronize()

    def initial_hfield(self, subject, out):
        m = hfield.grid_affine(self.origin, self.spacing, subject.origin, subject.spacing,
                               subject.affine)
        hfield.initializeFromAffine(out, m)
        return out

    def UpdateStepSizeNextIteration(self):
        _lib.call("awk_atlas_update_step_next", self._h)

    def Iterate(self, n_iter=1, stream=None):
        """n_iter atlas iterations; energies are kept on the device."""
        e = torch.empty(max(n_iter, 1), dtype=torch.float64, device="cuda")
        _lib.call("awk_atlas_iterate", self._h, int(n_iter), ptr_f64(e), stream_handle(stream))
        self._energy.append(e[:n_iter])

    def RunScaleLevel(self, n_iter):
        """One scale level: step sizes reset, then n_iter iterations."""
        self.UpdateStepSizeNextIteration()
        self.Iterate(n_iter)

    def energies(self):
        if not self._energy:
            return torch.empty(0, dtype=torch.float64).numpy()
        return torch.cat(self._energy).cpu().numpy()

    def Mean(self, stream=None):
        nx, ny, nz = self.size
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_atlas_mean", self._h, ptr(out), stream_handle(stream))
        return out

    def HField(self, i, stream=None):
        nx, ny, nz = self.size
        out = torch.empty((3, nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_atlas_get_hfield", self._h, int(i), planes_arr(out), stream_handle(stream))
        return out

    def DeformedImage(self, i, stream=None):
        nx, ny, nz = self.size
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        _lib.call("awk_atlas_deformed_image", self._h, int(i), ptr(out), stream_handle(stream))
        return out

    def __del__(self):
        try:
            if self._h is not None and self._h.value:
                _lib.call("awk_atlas_destroy", self._h)
            self._h = None
        except Exception:
            pass


def ptr_f64(t):
    if not t.is_cuda or t.dtype != torch.float64:
        raise ValueError("expected a float64 GPU tensor")
    return ctypes.c_void_p(t.data_ptr())
