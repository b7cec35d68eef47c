"""atlaswerks_amd -- MI355X-native greedy-fluid / LDMM registration inner loop.

The compute path is libawk_hip.so (hand-written HIP for gfx950, C-ABI in
include/awk_hip.h).  Python mirrors the reference's operator and driver API on
top of it; torch tensors provide device memory and streams.
"""
from . import _lib  # noqa: F401

__version__ = "0.1.0"


def build(verbose=False):
    """Compile libawk_hip.so for gfx950 in place (hipcc)."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
    subprocess.check_call(["make", "-B", "-C", here] + ([] if verbose else ["-s"]))
