"""LDMM atlas against the reference application's own regression outputs.

Testing/Applications/LDMM/LDMMAtlas/LDMMAtlasParams.xml: the four blurred
bullseyes (Testing/Data/Input/Bullseyes/BullseyeTestBlur00-03, 32x31x36 at
spacing (2, 2.06452, 1.77778)), weights scaled to 1/4, T = 10, Sigma 3, the
default alpha = beta = 0.01, gamma = 0.001, adaptive steps, IterateOrig (the
reference default), JacobianScale on, scale levels by downsample factor: factor 2
with MaxPert 0.1, then factor 1 with MaxPert 0.01, 100 iterations each.  The
baselines (LDMMAtlas/Baseline) are the mean image and each subject's image
deformed to time t (DefToMeanTime, GetI1AtT); the fixtures keep the mean and
t = 0 and t = 9 of every subject.

Bar: max |difference| <= 2e-5 for the mean and 5e-5 for the deformed images, on
images in [0, 1] (measured: 2.7e-6 and 7.1e-6, float rounding after 200
iterations)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_ldmm_atlas_reproduces_reference_outputs(tmp_path):
    import torch
    from atlaswerks_amd import ldmm, metaimage
    print("golden LDMM atlas: 4 bullseyes, 2 levels x 100 iterations, T = 10", flush=True)
    raw = [metaimage.read_nrrd(os.path.join(GOLD, "BullseyeTestBlur%02d.nhdr" % i)) for i in range(4)]
    sp = raw[0].spacing
    ims = [torch.from_numpy(r.data.astype(np.float32)).cuda() for r in raw]

    def prm(max_pert):
        return ldmm.LDMMIteratorParam(sigma=3.0, max_pert=max_pert, alpha=0.01, beta=0.01,
                                      gamma=0.001)

    b = ldmm.LDMMAtlasBuilderGPU(ims, n_time_steps=10, spacing=sp, param=prm(0.1),
                                 jacobian_scale=True,
                                 scale_levels=[(2, 100, prm(0.1)), (1, 100, prm(0.01))])
    b.Run(vec_energy=True)
    mean = b.Mean().cpu().numpy().astype(np.float64)
    ref = metaimage.read_metaimage(os.path.join(GOLD, "ldmm_atlas", "LDMMAtlasMeanImage.mha")).data
    assert np.abs(mean - ref).max() <= 2e-5
    for s in range(4):
        for t in (0, 9):
            name = "LDMMAtlasBullseyeTestBlur%02dDefToMeanTime%02d.mha" % (s, t)
            want = metaimage.read_metaimage(os.path.join(GOLD, "ldmm_atlas", name)).data
            got = b.DefToMean(s, t).cpu().numpy().astype(np.float64)
            assert np.abs(got - want).max() <= 5e-5, (s, t)
    e = b.energies()
    assert e.shape == (200, 3) and np.all(np.isfinite(e))
    # the output files as the reference application writes them: every fixture is
    # one of them, under the same name, with the same contents to the same bars
    written = b.WriteOutputs(str(tmp_path), prefix="LDMMAtlas", suffix="mha",
                             names=["BullseyeTestBlur%02d" % i for i in range(4)],
                             origin=raw[0].origin, energy=True)
    names = {os.path.basename(p) for p in written}
    assert len(names) == 1 + 4 * 10 + 1
    for fx in sorted(os.listdir(os.path.join(GOLD, "ldmm_atlas"))):
        assert fx in names, fx
        got = metaimage.read_metaimage(str(tmp_path / fx))
        want = metaimage.read_metaimage(os.path.join(GOLD, "ldmm_atlas", fx))
        assert got.data.shape == want.data.shape
        np.testing.assert_allclose(got.spacing, want.spacing, rtol=1e-5)
        bar = 2e-5 if fx.endswith("MeanImage.mha") else 5e-5
        assert np.abs(got.data.astype(np.float64) - want.data).max() <= bar, fx
