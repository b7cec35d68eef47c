"""GreedyWarp against the reference application's own regression output.

Testing/Applications/Greedy/GreedyWarp/GreedyWarpParams.xml: BullseyeTestBlur00
(moving) to BullseyeTestBlur01 (fixed), 32x31x36 at spacing (2, 2.06452,
1.77778), intensities rescaled to [0, 1] (already their range), alpha = beta =
0.01, gamma = 0.001, MaxPert 0.3, scale levels by downsample factor 2 then 1 with
50 iterations each, the inverse field computed.  The baseline
(GreedyWarp/Baseline/GreedyBullseyeTestBlur00_to_BullseyeTestBlur01DefImage.mha)
is I0 deformed by the final hfield.

Bar: relative L2 <= 1e-4 and max |difference| <= 2e-4 on an image in [0, 1]
(measured: 1e-5 and 4e-5)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_greedy_warp_reproduces_reference_output(tmp_path):
    import torch
    from atlaswerks_amd import greedy, hfield, metaimage
    print("golden GreedyWarp: bullseye 00 -> 01, levels 2 -> 1, 50 + 50 iterations", flush=True)
    a = metaimage.read_nrrd(os.path.join(GOLD, "BullseyeTestBlur00.nhdr"))
    b = metaimage.read_nrrd(os.path.join(GOLD, "BullseyeTestBlur01.nhdr"))
    want = metaimage.read_metaimage(os.path.join(
        GOLD, "greedy_warp", "GreedyBullseyeTestBlur00_to_BullseyeTestBlur01DefImage.mha")).data
    I0 = torch.from_numpy(a.data.astype(np.float32)).cuda()
    I1 = torch.from_numpy(b.data.astype(np.float32)).cuda()
    w = greedy.GreedyWarp(I0, I1, spacing=a.spacing,
                          param=greedy.GreedyIteratorParam(max_pert=0.3, alpha=0.01, beta=0.01,
                                                           gamma=0.001),
                          compute_inverse=True, scale_levels=[(2, 50), (1, 50)])
    data = w.RunWarp()
    out = torch.empty_like(I0)
    hfield.apply(out, I0, data.Def1To0)
    got = out.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-4
    assert np.abs(got - want).max() <= 2e-4
    # the reference application's output file, under its name, to the same bar
    paths = w.WriteOutputs("BullseyeTestBlur00_to_BullseyeTestBlur01", str(tmp_path), origin=a.origin,
                           inv_def_image=True, def_field=True, inv_def_field=True, energy=True)
    fx = "GreedyBullseyeTestBlur00_to_BullseyeTestBlur01DefImage.mha"
    assert fx in {os.path.basename(p) for p in paths} and len(paths) == 5
    written = metaimage.read_metaimage(str(tmp_path / fx))
    assert np.array_equal(written.data, out.cpu().numpy())
    np.testing.assert_allclose(written.spacing, a.spacing, rtol=1e-6)
