"""The GPU greedy atlas against the reference application's own regression.

Testing/Applications/Greedy/GreedyAtlas/GreedyAtlasParams.xml: the four blurred
bullseyes (32x31x36, spacing (2, 2.06452, 1.77778), origin (-277.8, -198.6,
54.2)), intensities rescaled to [0, 1], weights 1/4, alpha = beta = 0.02,
gamma = 0.001, MaxPert 0.2, scale levels by downsample factor 2 (50 iterations)
then 1 (150 iterations).

Bars:
- the reference's own: GenDiff's sum of |test - base| (GenDiffMain.cxx) below
  500 for GreedyMeanImage and every GreedyBullseyeTestBlur0<i>DefToMean
  (GreedyAtlas/GenHTML.py:24);
- against the FP64 oracle's run of the same setup: relative L2 <= 1e-4 for the
  mean and the deformed images;
- the written files carry the reference's output names and hold what the
  builder returns."""
import os

import numpy as np
import pytest

from test_oracle_golden_greedy_atlas import EPS, NAMES, gendiff, load_baseline, load_inputs

pytestmark = pytest.mark.gpu


def test_greedy_atlas_reproduces_reference_regression(tmp_path):
    import torch
    import oracle
    from atlaswerks_amd import atlas, metaimage
    from gpu_util import rel_l2
    print("golden GreedyAtlas: 4 bullseyes 32x31x36, levels 2 -> 1, 50 + 150 iterations",
          flush=True)
    ims = load_inputs()
    sp, org = ims[0].spacing, ims[0].origin
    data = [oracle.rescale_intensities(im.data) for im in ims]
    subs = [atlas.AtlasSubject(torch.from_numpy(d).cuda(), origin=org, spacing=sp, weight=0.25)
            for d in data]
    b = atlas.GreedyAtlasBuilderGPU(subs, (32, 31, 36), template_origin=org, template_spacing=sp,
                                    max_pert=0.2, alpha=0.02, beta=0.02, gamma=0.001,
                                    scale_levels=[(2, 50), (1, 150)])
    b.Run()
    mean = b.Mean().cpu().numpy().astype(np.float64)
    defs = [b.DeformedImage(i).cpu().numpy().astype(np.float64) for i in range(4)]
    bmean, bdefs = load_baseline()
    errs = [gendiff(mean, bmean)] + [gendiff(d, bd) for d, bd in zip(defs, bdefs)]
    print("GenDiff L1: mean %.1f, DefToMean %s" % (errs[0], " ".join("%.1f" % e for e in errs[1:])),
          flush=True)
    assert all(e < EPS for e in errs), errs
    # the same setup through the FP64 oracle
    _, omean, odefs, _ = oracle.atlas_multiscale_run(data, np.full(4, 0.25), [(2, 50), (1, 150)],
                                                     sp=sp, alpha=0.02, beta=0.02, gamma=0.001,
                                                     max_pert=0.2)
    assert rel_l2(mean, omean) <= 1e-4
    for d, od in zip(defs, odefs):
        assert rel_l2(d, od) <= 1e-4
    # the reference application's output files, under its names
    paths = b.WriteOutputs(str(tmp_path), prefix="Greedy", suffix="nhdr", names=NAMES)
    base = {os.path.basename(p) for p in paths}
    assert "GreedyMeanImage.nhdr" in base
    for n, d in zip(NAMES, defs):
        fx = "Greedy%sDefToMean.nhdr" % n
        assert fx in base
        w = metaimage.read_nrrd(str(tmp_path / fx))
        assert np.array_equal(w.data.astype(np.float64), d)
        np.testing.assert_allclose(w.spacing, sp, rtol=1e-6)
