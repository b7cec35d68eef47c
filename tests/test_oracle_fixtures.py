"""The oracle reproduces the reference's own golden data: FluidWarp::
computeHFieldAsymmetric on the 32^3 bullseyes (FluidWarpRoutineTests.cxx:178-218,
tolerance Sum |d|^2 < 10 at :28), and planted mistakes break it."""
import os

import numpy as np
import pytest

import oracle
from atlaswerks_amd.metaimage import read_metaimage


def load(golden_dir, name):
    return read_metaimage(os.path.join(golden_dir, name)).data


@pytest.fixture(scope="module")
def fw(golden_dir):
    g = lambda n: load(golden_dir, n)
    return {k: g(v) for k, v in [("I0", "BullseyeTest00.mha"), ("I1", "BullseyeTest01.mha"),
                                 ("h", "ComputeHFieldAsymmetricDeformation.mha"),
                                 ("hinv", "ComputeHFieldAsymmetricDeformationInv.mha"),
                                 ("def", "ComputeHFieldAsymmetricDeformedImage.mha"),
                                 ("definv", "ComputeHFieldAsymmetricDeformedImageInv.mha")]}


def sqdiffs(fw):
    h, hinv = oracle.fluidwarp_asym(fw["I0"], fw["I1"])
    d = oracle.apply(fw["I1"], h, dtype=np.float32)
    di = oracle.apply(fw["I0"], hinv, dtype=np.float32)
    return [float(((h - fw["h"]) ** 2).sum()), float(((hinv - fw["hinv"]) ** 2).sum()),
            float(((d - fw["def"]) ** 2).sum()), float(((di - fw["definv"]) ** 2).sum())]


def test_fluidwarp_asymmetric_fixture(fw):
    # reference tolerance is 10; the restatement lands within 1e-5
    assert max(sqdiffs(fw)) < 1e-5


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 6])
def test_fixture_catches_planted_mistakes(fw, mutation):
    oracle.set_mutation(mutation)
    try:
        d = sqdiffs(fw)
    finally:
        oracle.set_mutation(0)
    assert not (max(d) <= 10.0)  # NaN also counts as caught


def test_ldmm_atlas_fixtures_load():
    """The LDMM atlas regression fixtures (the reference's Testing data) parse and
    share the inputs' grid."""
    import os
    from atlaswerks_amd import metaimage
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    ins = [metaimage.read_nrrd(os.path.join(g, "BullseyeTestBlur%02d.nhdr" % i)) for i in range(4)]
    m = metaimage.read_metaimage(os.path.join(g, "ldmm_atlas", "LDMMAtlasMeanImage.mha"))
    assert m.data.shape == (36, 31, 32) and all(im.data.shape == (36, 31, 32) for im in ins)
    assert np.allclose(m.spacing, ins[0].spacing, atol=1e-5)
    assert 0.0 <= float(m.data.min()) and float(m.data.max()) <= 1.0
