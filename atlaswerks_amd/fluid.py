"""FluidKernelFFT: the Navier operator L = -alpha*Lap - beta*grad(div) + gamma on the GPU.

Mirrors the reference's GPU kernel operator, ``DiffOperGPU`` implementing
``KernelInterfaceGPU`` (Algorithms/Common/DiffOperGPU.h:27-135,
KernelInterfaceGPU.h:20-43): ``SetSize``, the parameter accessors and in-place
``ApplyOperator`` / ``ApplyInverseOperator`` on a device vector field.  Unlike the
reference GPU operator, LPow != 1, the eigen table path and the divergence-free
projection are supported (DiffOperGPU.cxx:127-161 throws for them).
"""
import ctypes

from . import _lib
from .device import dims, planes, stream_handle


class DiffOperParam:
    """DiffOperParam (Algorithms/Common/DiffOperParam.h:24-86) defaults."""

    def __init__(self, alpha=0.01, beta=0.01, gamma=0.001, lpow=1.0, use_eigen_lut=False,
                 divergence_free=False):
        self.alpha, self.beta, self.gamma, self.lpow = alpha, beta, gamma, lpow
        self.use_eigen_lut, self.divergence_free = use_eigen_lut, divergence_free


class DiffOperGPU:
    def __init__(self, size=None, spacing=(1.0, 1.0, 1.0), param=None):
        self._h = None
        self._size = None
        self._spacing = tuple(float(s) for s in spacing)
        self.param = param or DiffOperParam()
        if size is not None:
            self.SetSize(size, spacing, self.param)

    # --- KernelInterfaceGPU ---
    def SetSize(self, size, spacing, param=None):
        if param is not None:
            self.param = param
        self._destroy()
        size = tuple(int(s) for s in size)
        self._spacing = tuple(float(s) for s in spacing)
        h = ctypes.c_void_p()
        p = self.param
        _lib.call("awk_fluid_kernel_create", ctypes.byref(h), size[0], size[1], size[2],
                  self._spacing[0], self._spacing[1], self._spacing[2], p.alpha, p.beta, p.gamma,
                  p.lpow, self._flags())
        self._h = h
        self._size = size

    def ApplyOperator(self, vf, stream=None):
        self._apply(vf, 0, stream)

    def ApplyInverseOperator(self, vf, stream=None):
        self._apply(vf, 1, stream)

    # --- DiffOperGPU accessors ---
    def SetAlpha(self, a):
        self.param.alpha = a
        self._push()

    def GetAlpha(self):
        return self.param.alpha

    def SetBeta(self, b):
        self.param.beta = b
        self._push()

    def GetBeta(self):
        return self.param.beta

    def SetGamma(self, g):
        self.param.gamma = g
        self._push()

    def GetGamma(self):
        return self.param.gamma

    def SetLPow(self, p):
        self.param.lpow = p
        self._push()

    def GetLPow(self):
        return self.param.lpow

    def SetUseEigenLUT(self, b):
        self.param.use_eigen_lut = bool(b)
        self._push()

    def GetUseEigenLUT(self):
        return self.param.use_eigen_lut

    def SetDivergenceFree(self, b):
        self.param.divergence_free = bool(b)
        self._push()

    def GetDivergenceFree(self):
        return self.param.divergence_free

    def scratch_bytes(self):
        b = ctypes.c_ulonglong()
        _lib.call("awk_fluid_kernel_scratch_bytes", self._h, ctypes.byref(b))
        return b.value

    # --- internals ---
    def _flags(self):
        return ((_lib.FK_EIGEN if self.param.use_eigen_lut else 0) |
                (_lib.FK_DIVFREE if self.param.divergence_free else 0))

    def _push(self):
        if self._h is not None:
            p = self.param
            _lib.call("awk_fluid_kernel_set_params", self._h, p.alpha, p.beta, p.gamma, p.lpow,
                      self._flags())

    def _apply(self, vf, inverse, stream):
        if self._h is None:
            raise RuntimeError("DiffOperGPU: SetSize must be called first")
        if dims(vf) != self._size:
            raise ValueError("field size %s does not match operator size %s" % (dims(vf), self._size))
        x, y, z = planes(vf)
        _lib.call("awk_fluid_kernel_apply", self._h, x, y, z, int(inverse), stream_handle(stream))

    def _destroy(self):
        if self._h is not None:
            _lib.load().awk_fluid_kernel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass


FluidKernelFFT = DiffOperGPU
