"""Device-memory plumbing: torch tensors hold the fields, the HIP library computes.

A scalar image is a contiguous float32 tensor ``[nz, ny, nx]``; a vector field
(hfield or velocity) is a contiguous float32 tensor ``[3, nz, ny, nx]`` whose
three planes are the SoA x, y, z components the C-ABI takes.
"""
import ctypes

import torch


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    if t is None:
        return None
    _check(t)
    return ctypes.c_void_p(t.data_ptr())


def planes(v):
    if v.dim() != 4 or v.shape[0] != 3:
        raise ValueError("vector field must have shape [3, nz, ny, nx]")
    _check(v)
    return [ctypes.c_void_p(v[c].data_ptr()) for c in range(3)]


def planes_arr(v):
    if v is None:
        return None
    p = planes(v)
    return ctypes.pointer((ctypes.c_void_p * 3)(p[0].value, p[1].value, p[2].value))


def _check(t):
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if t.dtype != torch.float32:
        raise ValueError("tensor must be float32")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")


def dims(t):
    """(nx, ny, nz) of an image [nz,ny,nx] or a field [3,nz,ny,nx]."""
    s = t.shape[-3:]
    return int(s[2]), int(s[1]), int(s[0])


def new_field(size, device="cuda"):
    nx, ny, nz = size
    return torch.empty((3, nz, ny, nx), dtype=torch.float32, device=device)


def new_image(size, device="cuda"):
    nx, ny, nz = size
    return torch.empty((nz, ny, nx), dtype=torch.float32, device=device)
