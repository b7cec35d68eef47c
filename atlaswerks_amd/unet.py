"""U-Net warm start (MFMA bf16) for the greedy and LDMM registrations.

The network (``awk_unet_*``; 4 levels, 16-32-64-128 channels, two 3x3x3
convolutions per level, 2x2x2 average pooling, nearest x2 upsampling with skip
concatenation, LeakyReLU(0.2), a 1x1x1 head) maps a (template, subject) pair to
an initial displacement u; the registration then starts from h = x + u.  It has
no counterpart in the reference.  Weights are a flat fp32 blob in the layer order
of ``LAYERS`` (PyTorch OIDHW layout, then the bias, per layer; the head last).
"""
import ctypes

import numpy as np
import torch

from . import _lib, hfield
from .device import planes_arr, ptr, stream_handle

# (cin, cout) of the 14 3x3x3 convolutions as the weight blob carries them
LAYERS = ((2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128),
          (192, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16))
HEAD = (16, 3)


def weight_count():
    return int(_lib.load().awk_unet_weight_count())


def init_weights(seed=0, head_std=1e-2):
    """He-normal convolutions (LeakyReLU gain), zero biases, a small head."""
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout in LAYERS:
        std = np.sqrt(2.0 / (1 + 0.2 ** 2) / (cin * 27))
        parts.append(rng.normal(0.0, std, size=(cout, cin, 3, 3, 3)).astype(np.float32).ravel())
        parts.append(np.zeros(cout, np.float32))
    parts.append(rng.normal(0.0, head_std, size=(3, 16)).astype(np.float32).ravel())
    parts.append(np.zeros(3, np.float32))
    return np.concatenate(parts)


def split_weights(blob):
    """[(W, b)] per convolution and (W_head, b_head) from a flat blob."""
    out, off = [], 0
    for cin, cout in LAYERS:
        n = cout * cin * 27
        out.append((blob[off:off + n].reshape(cout, cin, 3, 3, 3), blob[off + n:off + n + cout]))
        off += n + cout
    head = (blob[off:off + 48].reshape(3, 16), blob[off + 48:off + 51])
    return out, head


class UNetWarmStart:
    """The network on an (nx, ny, nz) grid, with activations for up to
    ``max_batch`` pairs per forward pass."""

    def __init__(self, size, weights=None, seed=0, max_batch=1):
        self.size = tuple(int(v) for v in size)  # (nx, ny, nz)
        w = init_weights(seed) if weights is None else np.ascontiguousarray(weights, np.float32)
        if w.size != weight_count():
            raise ValueError("weight blob has %d values, the network takes %d" % (w.size, weight_count()))
        self.weights = w
        self.max_batch = int(max_batch)
        h = ctypes.c_void_p()
        _lib.call("awk_unet_create_batch", ctypes.byref(h), *self.size, self.max_batch,
                  w.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(w.size))
        self._h = h

    def forward(self, template, subject, out=None, stream=None):
        """Displacement [3, nz, ny, nx] (fp32, voxels)."""
        nx, ny, nz = self.size
        if out is None:
            out = torch.empty((3, nz, ny, nx), dtype=torch.float32, device=template.device)
        _lib.call("awk_unet_forward", self._h, ptr(template), ptr(subject), planes_arr(out),
                  stream_handle(stream))
        return out

    def forward_batch(self, templates, subjects, out=None, stream=None):
        """Displacements [B, 3, nz, ny, nx] of B pairs in one pass (B <= max_batch).
        ``templates`` is one image shared by every pair or a list of B images."""
        subjects = list(subjects)
        B = len(subjects)
        if not isinstance(templates, (list, tuple)):
            templates = [templates] * B
        if len(templates) != B:
            raise ValueError("one template per subject (or a single shared one)")
        nx, ny, nz = self.size
        if out is None:
            out = torch.empty((B, 3, nz, ny, nx), dtype=torch.float32, device=subjects[0].device)
        for t in list(templates) + subjects:
            ptr(t)
            if tuple(t.shape) != (nz, ny, nx):
                raise ValueError("images must be [nz, ny, nx] = %s" % ((nz, ny, nx),))
        if tuple(out.shape) != (B, 3, nz, ny, nx) or not out.is_contiguous():
            raise ValueError("out must be a contiguous [B, 3, nz, ny, nx] tensor")
        tp = (ctypes.c_void_p * B)(*[t.data_ptr() for t in templates])
        sp = (ctypes.c_void_p * B)(*[t.data_ptr() for t in subjects])
        dp = (ctypes.c_void_p * (3 * B))(*[out[b, c].data_ptr() for b in range(B) for c in range(3)])
        _lib.call("awk_unet_forward_batch", self._h, B, tp, sp, dp, stream_handle(stream))
        return out

    def warm_start_hfield(self, template, subject, stream=None):
        """h = x + u (HField3DUtils::velocityToH with unit spacing)."""
        h = self.forward(template, subject, stream=stream)
        hfield.addAffine(h, hfield.IDENTITY_AFFINE, 1.0, stream)
        return h

    def __del__(self):
        try:
            if self._h is not None and self._h.value:
                _lib.call("awk_unet_destroy", self._h)
            self._h = None
        except Exception:
            pass


class UNetTrainer:
    """Adam training of a ``UNetWarmStart`` on (template, subject, target
    displacement) triples, all on the device (``awk_unet_trainer_*``)."""

    def __init__(self, net):
        self.net = net
        h = ctypes.c_void_p()
        w = net.weights
        _lib.call("awk_unet_trainer_create", ctypes.byref(h), net._h, w.ctypes.data_as(ctypes.c_void_p),
                  ctypes.c_size_t(w.size))
        self._h = h
        self.loss = torch.zeros(1, dtype=torch.float64, device="cuda")

    def step(self, template, subject, target, lr=1e-3, stream=None):
        """One Adam step; the step's loss is added to ``self.loss`` on the device."""
        _lib.call("awk_unet_trainer_step", self._h, ptr(template), ptr(subject), planes_arr(target),
                  float(lr), ctypes.c_void_p(self.loss.data_ptr()), stream_handle(stream))

    def gradients(self, template, subject, target):
        """(gradient blob, loss) of one pair without an update (synchronises)."""
        g = np.zeros(self.net.weights.size, np.float32)
        loss = ctypes.c_double()
        _lib.call("awk_unet_trainer_gradients", self._h, ptr(template), ptr(subject),
                  planes_arr(target), g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss))
        return g, loss.value

    def weights(self):
        w = np.zeros(self.net.weights.size, np.float32)
        _lib.call("awk_unet_trainer_get_weights", self._h, w.ctypes.data_as(ctypes.c_void_p),
                  ctypes.c_size_t(w.size))
        return w

    def __del__(self):
        try:
            if self._h is not None and self._h.value:
                _lib.call("awk_unet_trainer_destroy", self._h)
            self._h = None
        except Exception:
            pass
