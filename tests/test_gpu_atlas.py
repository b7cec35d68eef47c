"""Greedy atlas on the GPU vs the FP64 oracle (AtlasThread.txx:95-181,
GreedyAtlasBuilder.txx:236-301).

Bars: hfields relative L2 <= 1e-4 after several iterations, the mean relative L2
<= 1e-5, energies relative 1e-4 (float mean against the FP64 mean); hipGraph
replays bit-identical to eager launches, replay after replay with host
synchronisation between them."""
import os

import numpy as np
import pytest

import oracle
from atlaswerks_amd import synth

pytestmark = pytest.mark.gpu


def _subjects(shape, n, seed):
    t = synth.blobs(shape, seed=seed)
    ims = []
    for i in range(n):
        u = synth.sinusoidal_displacement(shape, max_disp=1.5, seed=seed + 10 + i)
        h = np.moveaxis(synth.identity(shape) + u, 0, -1)
        ims.append(oracle.apply(t, h, dtype=np.float32))
    return ims


@pytest.mark.parametrize("shape", [(16, 20, 24), (32, 32, 32)])
def test_atlas_same_grid_vs_oracle(shape):
    import torch
    from atlaswerks_amd import atlas
    from gpu_util import aos, rel_l2, to_dev
    print("atlas", shape, flush=True)
    ims = _subjects(shape, 3, seed=40)
    w = [0.5, 0.3, 0.2]
    nz, ny, nx = shape
    subs = [atlas.AtlasSubject(to_dev(im), weight=wi) for im, wi in zip(ims, w)]
    b = atlas.GreedyAtlasBuilderGPU(subs, (nx, ny, nz), max_pert=0.2)
    b.RunScaleLevel(5)
    hs_ref, mean_ref, en_ref = oracle.atlas_run(ims, [oracle.identity(shape, np.float64)] * 3, w, 5,
                                                shape, max_pert=0.2)
    for i in range(3):
        assert rel_l2(aos(b.HField(i).cpu().numpy()), hs_ref[i]) <= 1e-4, i
    assert rel_l2(b.Mean().cpu().numpy(), mean_ref) <= 1e-5
    np.testing.assert_allclose(b.energies(), en_ref, rtol=1e-4)


def test_atlas_uneven_grids_vs_oracle():
    """Subjects of three different sizes and spacings over one template grid."""
    import torch
    from atlaswerks_amd import atlas
    from gpu_util import aos, rel_l2, to_dev
    # the same world-space blob template sampled on each subject's grid
    grids = [((32, 32, 32), (1.0, 1.0, 1.0)), ((24, 24, 24), (4 / 3, 4 / 3, 4 / 3)),
             ((20, 28, 16), (1.6, 8 / 7, 2.0))]
    base = synth.blobs((32, 32, 32), seed=50)
    ims = []
    for k, (shp, sp) in enumerate(grids):
        nz, ny, nx = shp
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        pts = np.stack([x * sp[0], y * sp[1], z * sp[2]], -1)  # world = base voxels
        u = np.moveaxis(synth.sinusoidal_displacement(shp, max_disp=1.0, seed=60 + k), 0, -1)
        ims.append(oracle.apply(base, pts + u, dtype=np.float32))
    tshape = (16, 16, 16)
    subs = [atlas.AtlasSubject(to_dev(im), spacing=sp) for im, (_, sp) in zip(ims, grids)]
    b = atlas.GreedyAtlasBuilderGPU(subs, (16, 16, 16), template_spacing=(2.0, 2.0, 2.0),
                                    max_pert=0.2)
    h0 = [aos(b.HField(i).cpu().numpy()).astype(np.float64) for i in range(3)]
    b.RunScaleLevel(4)
    hs_ref, mean_ref, en_ref = oracle.atlas_run(ims, h0, [1 / 3] * 3, 4, tshape, sp=(2, 2, 2),
                                                max_pert=0.2)
    for i in range(3):
        assert rel_l2(aos(b.HField(i).cpu().numpy()), hs_ref[i]) <= 1e-4, i
    assert rel_l2(b.Mean().cpu().numpy(), mean_ref) <= 1e-5
    np.testing.assert_allclose(b.energies(), en_ref, rtol=1e-4)


def test_atlas_graph_replays_match_eager():
    import torch
    from atlaswerks_amd import atlas
    from gpu_util import to_dev
    shape = (24, 32, 32)
    ims = _subjects(shape, 4, seed=70)
    nz, ny, nx = shape
    runs = []
    for use_graph in (False, True):
        subs = [atlas.AtlasSubject(to_dev(im)) for im in ims]
        b = atlas.GreedyAtlasBuilderGPU(subs, (nx, ny, nz), max_pert=0.2, use_graph=use_graph)
        b.UpdateStepSizeNextIteration()
        snaps = []
        for it in range(5):
            b.Iterate(1)
            torch.cuda.synchronize()
            snaps.append([b.HField(i).cpu().numpy() for i in range(4)] + [b.Mean().cpu().numpy()])
        runs.append((snaps, b.energies()))
    for it in range(5):
        for a, c in zip(runs[0][0][it], runs[1][0][it]):
            np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("use_graph", [False, True])
def test_atlas_inverse_fields_vs_oracle(use_graph):
    """Def0To1 per subject (composeVHInv of each step, GreedyIteratorCPU.cxx:82-95)."""
    import torch
    from atlaswerks_amd import atlas
    from gpu_util import aos, rel_l2, to_dev
    print("atlas inverse graph=%d" % use_graph, flush=True)
    shape = (16, 20, 24)
    sp = (1.0, 1.2, 0.9)
    ims = _subjects(shape, 3, seed=70)
    nz, ny, nx = shape
    subs = [atlas.AtlasSubject(to_dev(im), spacing=sp) for im in ims]
    b = atlas.GreedyAtlasBuilderGPU(subs, (nx, ny, nz), template_spacing=sp, max_pert=0.2,
                                    use_graph=use_graph, compute_inverse=True)
    b.RunScaleLevel(2)
    b.Iterate(4)  # steady-state iterations (captured and replayed with use_graph)
    ident = oracle.identity(shape, np.float64)
    hs_ref, mean_ref, en_ref, his_ref = oracle.atlas_run(ims, [ident] * 3, [1 / 3] * 3, 6, shape,
                                                         sp=sp, max_pert=0.2, hinv0=[ident] * 3)
    for i in range(3):
        assert rel_l2(aos(b.HField(i).cpu().numpy()), hs_ref[i]) <= 1e-4, i
        assert rel_l2(aos(b.HFieldInverse(i).cpu().numpy()), his_ref[i]) <= 1e-4, i
    assert rel_l2(b.Mean().cpu().numpy(), mean_ref) <= 1e-5
    np.testing.assert_allclose(b.energies(), en_ref, rtol=1e-4)


def test_atlas_inverse_multiscale_consistent():
    """Two scale levels with inverse fields: h(hinv(x)) stays near x in the interior."""
    import torch
    from atlaswerks_amd import atlas
    from gpu_util import aos, to_dev
    print("atlas inverse multiscale", flush=True)
    shape = (32, 32, 32)
    ims = _subjects(shape, 3, seed=80)
    subs = [atlas.AtlasSubject(to_dev(im)) for im in ims]
    b = atlas.GreedyAtlasBuilderGPU(subs, (32, 32, 32), max_pert=0.2, scale_levels=[(2, 10), (1, 10)],
                                    compute_inverse=True)
    b.Run()
    ident = oracle.identity(shape, np.float64)
    for i in range(3):
        h = aos(b.HField(i).cpu().numpy()).astype(np.float64)
        hi = aos(b.HFieldInverse(i).cpu().numpy()).astype(np.float64)
        err = np.sqrt(((oracle.compose(h, hi) - ident) ** 2).sum(-1))[4:-4, 4:-4, 4:-4]
        disp = np.sqrt(((h - ident) ** 2).sum(-1))[4:-4, 4:-4, 4:-4]
        assert disp.max() > 0.1
        assert err.max() < 0.15 * disp.max(), (err.max(), disp.max())
    # the energies as the reference's EnergyHistory file: one Iteration per level step
    import xml.etree.ElementTree as ET
    it = ET.fromstring(b.EnergyHistory().to_xml()).findall("Iteration")
    assert [(i.get("Scale"), i.get("Iter")) for i in it] == (
        [("0", str(k)) for k in range(10)] + [("1", str(k)) for k in range(10)])


def test_atlas_write_outputs(tmp_path, golden_dir):
    """The reference's GreedyAtlas output file set from the builder (NRRD, no ITK):
    names as GreedyAtlasBuilder writes them, contents equal to the builder's
    tensors, headers on the input images' grid, the energy history as XML."""
    import xml.etree.ElementTree as ET

    import torch
    from atlaswerks_amd import atlas, metaimage
    names = ["BullseyeTestBlur%02d" % i for i in range(4)]
    subs = []
    for n in names:
        im = metaimage.read_nrrd(os.path.join(golden_dir, n + ".nhdr"))
        subs.append(atlas.AtlasSubject(torch.from_numpy(im.data).cuda(), origin=im.origin,
                                       spacing=im.spacing))
    nz, ny, nx = subs[0].image.shape
    b = atlas.GreedyAtlasBuilderGPU(subs, (nx, ny, nz), template_origin=subs[0].origin,
                                    template_spacing=subs[0].spacing, max_pert=0.2, alpha=0.02,
                                    beta=0.02, gamma=0.001, scale_levels=[(2, 3), (1, 3)],
                                    compute_inverse=True)
    b.Run()
    paths = b.WriteOutputs(str(tmp_path), names=names, def_fields_mean_to_im=True,
                           def_fields_im_to_mean=True, energy=True)
    want = ["GreedyMeanImage.nhdr"]
    for n in names:
        want += ["Greedy%sDefToMean.nhdr" % n, "Greedy%sDefFieldMeanToIm.nhdr" % n,
                 "Greedy%sDefFieldImToMean.nhdr" % n]
    want.append("GreedyEnergy.xml")
    assert sorted(os.path.basename(p) for p in paths) == sorted(want)
    mean = metaimage.read_nrrd(str(tmp_path / "GreedyMeanImage.nhdr"))
    assert np.array_equal(mean.data, b.Mean().cpu().numpy())
    d1 = metaimage.read_nrrd(str(tmp_path / ("Greedy%sDefToMean.nhdr" % names[1])))
    assert np.array_equal(d1.data, b.DeformedImage(1).cpu().numpy())
    h2 = metaimage.read_nrrd(str(tmp_path / ("Greedy%sDefFieldMeanToIm.nhdr" % names[2])))
    assert np.array_equal(h2.data, np.moveaxis(b.HField(2).cpu().numpy(), 0, -1))
    g3 = metaimage.read_nrrd(str(tmp_path / ("Greedy%sDefFieldImToMean.nhdr" % names[3])))
    assert np.array_equal(g3.data, np.moveaxis(b.HFieldInverse(3).cpu().numpy(), 0, -1))

    def fields(path):
        out = {}
        for line in open(path):
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                out[k.strip()] = v.strip()
        return out
    ref = fields(os.path.join(golden_dir, names[0] + ".nhdr"))
    got = fields(str(tmp_path / "GreedyMeanImage.nhdr"))
    for k in ("sizes", "space directions", "space origin", "type", "endian", "encoding"):
        assert got[k] == ref[k], (k, got[k], ref[k])
    root = ET.parse(str(tmp_path / "GreedyEnergy.xml")).getroot()
    assert root.tag == "EnergyHistory" and len(root.findall("Iteration")) == 6


@pytest.mark.parametrize("shape,batches", [((24, 32, 32), (1, 3, None)), ((16, 20, 24), (2, None)),
                                           ((64, 64, 64), (None,))])
def test_atlas_batched_launches_bit_identical(shape, batches):
    """Steady-state iterations with each pass launched once over a batch of
    subjects (blockIdx.y = subject; batches of 1, 3, 2, all) against one launch per
    subject and pass on the stream lanes (batch 0): hfields, mean and energies bit
    for bit, eager and through hipGraph replays, over 6 iterations; the first
    iteration (step sizes) takes the per-subject path in both."""
    import torch
    from atlaswerks_amd import atlas
    from gpu_util import to_dev
    print("batched atlas", shape, batches, flush=True)
    ims = _subjects(shape, 5, seed=90)
    nz, ny, nx = shape
    subs = [atlas.AtlasSubject(to_dev(im), weight=w) for im, w in zip(ims, [0.3, 0.1, 0.2, 0.25, 0.15])]

    def run(batch, use_graph):
        b = atlas.GreedyAtlasBuilderGPU(subs, (nx, ny, nz), max_pert=0.2, use_graph=use_graph,
                                        batch=batch)
        b.UpdateStepSizeNextIteration()
        b.Iterate(6)
        torch.cuda.synchronize()
        return [b.HField(i).cpu().numpy() for i in range(5)] + [b.Mean().cpu().numpy()], b.energies()

    ref_f, ref_e = run(0, False)
    for batch in batches:
        for use_graph in (False, True):
            f, e = run(batch, use_graph)
            for a, c in zip(f, ref_f):
                np.testing.assert_array_equal(a, c)
            np.testing.assert_array_equal(e, ref_e)


def test_atlas_batched_uneven_grids_bit_identical():
    """Subjects on three different image grids (the gather of stage A per subject,
    stage B batched on the template grid) against the per-subject launches."""
    import torch
    from atlaswerks_amd import atlas
    from gpu_util import to_dev
    grids = [((32, 32, 32), (1.0, 1.0, 1.0)), ((24, 24, 24), (4 / 3, 4 / 3, 4 / 3)),
             ((20, 28, 16), (1.6, 8 / 7, 2.0))]
    base = synth.blobs((32, 32, 32), seed=51)
    subs = []
    for k, (shp, sp) in enumerate(grids):
        nz, ny, nx = shp
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        pts = np.stack([x * sp[0], y * sp[1], z * sp[2]], -1)
        u = np.moveaxis(synth.sinusoidal_displacement(shp, max_disp=1.0, seed=61 + k), 0, -1)
        subs.append(atlas.AtlasSubject(to_dev(oracle.apply(base, pts + u, dtype=np.float32)), spacing=sp))
    out = []
    for batch in (0, None):
        b = atlas.GreedyAtlasBuilderGPU(subs, (16, 16, 16), template_spacing=(2.0, 2.0, 2.0),
                                        max_pert=0.2, batch=batch)
        b.RunScaleLevel(5)
        torch.cuda.synchronize()
        out.append(([b.HField(i).cpu().numpy() for i in range(3)] + [b.Mean().cpu().numpy()], b.energies()))
    for a, c in zip(out[0][0], out[1][0]):
        np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("grids", ["same", "uneven"])
def test_atlas_fused_mean_stage_bit_identical(grids):
    """The deformed images and the mean in one launch (knob atlas_mean = 1, 2 or 4
    subjects per group; 5 subjects, so groups of 2 and 4 leave a remainder) against
    one apply launch per subject followed by the mean (atlas_mean = 0): hfields,
    mean, every deformed image and the energies bit for bit over 5 iterations,
    eager and through hipGraph replays."""
    import torch
    from atlaswerks_amd import atlas, autotune
    from gpu_util import to_dev
    print("fused atlas mean", grids, flush=True)
    if grids == "same":
        shape = (20, 24, 28)
        ims = _subjects(shape, 5, seed=140)
        subs = [atlas.AtlasSubject(to_dev(im), weight=w) for im, w in zip(ims, [0.3, 0.1, 0.2, 0.25, 0.15])]
        kw = dict(max_pert=0.2)
        nz, ny, nx = shape
        tgrid = (nx, ny, nz)
    else:
        shps = [((32, 32, 32), (1.0, 1.0, 1.0)), ((24, 24, 24), (4 / 3, 4 / 3, 4 / 3)),
                ((20, 28, 16), (1.6, 8 / 7, 2.0)), ((16, 16, 16), (2.0, 2.0, 2.0)),
                ((30, 18, 22), (16 / 15, 16 / 9, 16 / 11))]
        base = synth.blobs((32, 32, 32), seed=151)
        subs = []
        for k, (shp, sp) in enumerate(shps):
            nz, ny, nx = shp
            z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            pts = np.stack([x * sp[0], y * sp[1], z * sp[2]], -1)
            u = np.moveaxis(synth.sinusoidal_displacement(shp, max_disp=1.0, seed=161 + k), 0, -1)
            subs.append(atlas.AtlasSubject(to_dev(oracle.apply(base, pts + u, dtype=np.float32)), spacing=sp))
        kw = dict(template_spacing=(2.0, 2.0, 2.0), max_pert=0.2)
        tgrid = (16, 16, 16)
    saved = autotune.get_knob("atlas_mean")

    def run(su, use_graph):
        autotune.set_knob("atlas_mean", su)
        b = atlas.GreedyAtlasBuilderGPU(subs, tgrid, use_graph=use_graph, **kw)
        b.UpdateStepSizeNextIteration()
        b.Iterate(5)
        torch.cuda.synchronize()
        f = [b.HField(i).cpu().numpy() for i in range(5)] + [b.Mean().cpu().numpy()]
        f += [b.DeformedImage(i).cpu().numpy() for i in range(5)]
        return f, b.energies()

    try:
        ref_f, ref_e = run(0, False)
        for su in (1, 2, 4):
            for use_graph in (False, True):
                f, e = run(su, use_graph)
                for a, c in zip(f, ref_f):
                    np.testing.assert_array_equal(a, c)
                np.testing.assert_array_equal(e, ref_e)
    finally:
        autotune.set_knob("atlas_mean", saved)


def test_atlas_staged_composition_bit_identical():
    """The batched composing pass (one launch over a batch of subjects) reading the
    fields from rows staged in LDS (knob xinv_stage) against its gathers from
    memory: hfields, mean and energies bit for bit over 6 iterations."""
    import torch
    from atlaswerks_amd import atlas, autotune
    from gpu_util import to_dev
    print("staged atlas", flush=True)
    shape = (16, 32, 64)
    ims = _subjects(shape, 4, seed=170)
    nz, ny, nx = shape
    subs = [atlas.AtlasSubject(to_dev(im)) for im in ims]
    saved = autotune.get_knob("xinv_stage")

    def run(stage):
        autotune.set_knob("xinv_stage", stage)
        b = atlas.GreedyAtlasBuilderGPU(subs, (nx, ny, nz), max_pert=0.2, batch=None)
        b.UpdateStepSizeNextIteration()
        b.Iterate(6)
        torch.cuda.synchronize()
        return [b.HField(i).cpu().numpy() for i in range(4)] + [b.Mean().cpu().numpy()], b.energies()

    try:
        ref_f, ref_e = run(0)
        f, e = run(1)
        for a, c in zip(f, ref_f):
            np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(e, ref_e)
    finally:
        autotune.set_knob("xinv_stage", saved)
