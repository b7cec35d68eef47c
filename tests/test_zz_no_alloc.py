"""awk_splat_noalloc (the splat without allocations: its FIXED64 accumulator is a
file-scope device array) against awk_splat, two calls in a row, the second on a
smaller grid: FIXED64 and FIXED20 results are bit-identical (integer atomics)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(shape, seed):
    import torch
    from atlaswerks_amd import synth
    img = torch.from_numpy(synth.blobs(shape, seed=seed)).cuda()
    h = torch.from_numpy(synth.identity(shape) + synth.sinusoidal_displacement(shape, 3.0, seed=seed)).cuda()
    return img, h


def _splat(name, dst, img, h, mode):
    from atlaswerks_amd import _lib
    from atlaswerks_amd.device import planes, ptr, stream_handle
    nz, ny, nx = img.shape
    _lib.call(name, ptr(dst), ptr(img), *planes(h), nx, ny, nz, int(mode), stream_handle(None))


@pytest.mark.parametrize("mode", ["FIXED64", "FIXED20"])
def test_noalloc_splat_equals_splat_twice(mode):
    import torch
    from atlaswerks_amd import _lib
    m = getattr(_lib, "SPLAT_" + mode)
    for shape, seed in (((40, 48, 56), 1), ((17, 9, 30), 2)):
        print("noalloc", mode, shape, flush=True)
        img, h = _case(shape, seed)
        a = torch.full_like(img, 7.0)
        b = torch.full_like(img, -3.0)
        _splat("awk_splat", a, img, h, m)
        _splat("awk_splat_noalloc", b, img, h, m)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (shape, float((a - b).abs().max()))
        assert float(a.abs().sum()) > 0


def test_noalloc_splat_rejects_oversize():
    import torch
    from atlaswerks_amd import _lib
    from atlaswerks_amd.device import stream_handle
    lib = _lib.load()
    x = torch.zeros(8, device="cuda")
    p = ctypes.c_void_p(x.data_ptr())
    rc = lib.awk_splat_noalloc(p, p, p, p, p, 256, 256, 65, _lib.SPLAT_FIXED64, stream_handle(None))
    assert rc == _lib.AWK_ERR_INVALID
    assert b"2^22" in lib.awk_last_error()
