"""Helpers shared by the GPU parity tests."""
import numpy as np
import torch


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt((b ** 2).sum())
    return float(np.sqrt(((a - b) ** 2).sum()) / (den if den > 0 else 1.0))


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def soa(aos):
    """[z,y,x,3] -> [3,z,y,x]"""
    return np.ascontiguousarray(np.moveaxis(np.asarray(aos), -1, 0))


def aos(soa_arr):
    """[3,z,y,x] -> [z,y,x,3]"""
    return np.ascontiguousarray(np.moveaxis(np.asarray(soa_arr), 0, -1))


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()
