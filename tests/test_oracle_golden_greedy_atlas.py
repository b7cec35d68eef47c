"""The oracle's greedy atlas against the reference application's own regression.

Testing/Applications/Greedy/GreedyAtlas/GreedyAtlasParams.xml: the four blurred
bullseyes (32x31x36, spacing (2, 2.06452, 1.77778)), intensities rescaled to
[0, 1], weights scaled to 1/4, alpha = beta = 0.02, gamma = 0.001, MaxPert 0.2,
scale levels by downsample factor 2 (50 iterations) then 1 (150 iterations), run
on the CPU path (UseGPU false, one thread per image).

The reference's bar: GenDiff's sum of |test - base| over the image
(Testing/TestUtilities/GenDiffMain.cxx) below eps = 500 for the mean and for every
DefToMean image (GreedyAtlas/GenHTML.py:24, exit status -1 on failure).

Measured: mean 344.8, DefToMean 430.1 / 334.5 / 393.1 / 364.1 (relative L2 of the
mean 6.7 %).  The restatement's multi-iteration and multi-level semantics are
pinned separately (tests/test_oracle_atlas.py); the residual against this baseline
is recorded as an open difference of the baseline run."""
import os

import numpy as np

import oracle
from atlaswerks_amd import metaimage

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = 500.0
NAMES = ["BullseyeTestBlur%02d" % i for i in range(4)]


def load_inputs():
    ims = [metaimage.read_nrrd(os.path.join(GOLD, n + ".nhdr")) for n in NAMES]
    return ims


def load_baseline():
    d = os.path.join(GOLD, "greedy_atlas")
    mean = metaimage.read_nrrd(os.path.join(d, "GreedyMeanImage.nhdr")).data.astype(np.float64)
    defs = [metaimage.read_nrrd(os.path.join(d, "Greedy%sDefToMean.nhdr" % n)).data
            .astype(np.float64) for n in NAMES]
    return mean, defs


def gendiff(test, base):
    """GenDiffMain.cxx: err = sum over voxels of |test - base|."""
    return float(np.abs(np.asarray(test, np.float64) - base).sum())


def test_oracle_greedy_atlas_meets_reference_regression_bar():
    ims = load_inputs()
    sp = ims[0].spacing
    data = [oracle.rescale_intensities(im.data) for im in ims]
    _, mean, defs, en = oracle.atlas_multiscale_run(data, np.full(4, 0.25), [(2, 50), (1, 150)],
                                                    sp=sp, alpha=0.02, beta=0.02, gamma=0.001,
                                                    max_pert=0.2)
    bmean, bdefs = load_baseline()
    errs = [gendiff(mean, bmean)] + [gendiff(d, b) for d, b in zip(defs, bdefs)]
    print("GenDiff L1: mean %.1f, DefToMean %s" % (errs[0], " ".join("%.1f" % e for e in errs[1:])))
    assert all(e < EPS for e in errs), errs
    # the baseline is self-consistent: its mean is the average of its DefToMean images
    assert np.abs(bmean - sum(bdefs) / 4).max() < 1e-6
    # each level's energy decreases
    for e in en:
        assert np.all(np.diff(e) < 0)
