"""Learned warm start: batched inference, the warm-started atlas, the fallback.

- A batch of pairs in one forward pass gives bit for bit the single-pair
  forwards (the batch stacks volumes along z; split-K is chosen per volume).
- A warm start that misses its SSD target falls back to the plain method, and
  the result is bit for bit a plain run's (hfields, mean and energies).
- The trained network (weights trained on the device by
  ``atlaswerks_amd.train_warmstart``) reaches the plain atlas's final SSD within
  1 % on a held-out atlas group, without falling back, in fewer full-resolution
  iterations than the plain method spends there; an untrained (zero-output)
  network from the same start does not.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "atlaswerks_amd", "weights", "unet_atlas_128.npy")


def _zero_head(seed=0):
    from atlaswerks_amd.unet import init_weights
    w = init_weights(seed)
    w[-51:] = 0.0  # head weights and biases: the prediction is 0
    return w


@pytest.mark.parametrize("shape,batch", [((16, 16, 16), 16), ((64, 64, 64), 3), ((24, 16, 8), 5)])
def test_unet_batch_equals_single(shape, batch):
    import torch
    from atlaswerks_amd.unet import UNetWarmStart, init_weights
    print("batch", shape, batch, flush=True)
    nz, ny, nx = shape
    w = init_weights(3)
    w[-51:] = np.random.default_rng(1).normal(0, 0.2, 51).astype(np.float32)
    g = torch.Generator(device="cuda").manual_seed(5)
    ts = [torch.rand(shape, device="cuda", generator=g) for _ in range(batch)]
    ss = [torch.rand(shape, device="cuda", generator=g) for _ in range(batch)]
    net = UNetWarmStart((nx, ny, nz), weights=w, max_batch=batch)
    out = net.forward_batch(ts, ss)
    single = UNetWarmStart((nx, ny, nz), weights=w)
    for b in range(batch):
        one = single.forward(ts[b], ss[b])
        torch.cuda.synchronize()
        assert torch.equal(out[b], one), (b, float((out[b] - one).abs().max()))
    assert float(out.abs().max()) > 0


def test_unet_batch_rejects_oversize():
    import torch
    from atlaswerks_amd import _lib
    from atlaswerks_amd.unet import UNetWarmStart
    net = UNetWarmStart((8, 8, 8), weights=_zero_head(), max_batch=2)
    im = torch.zeros((8, 8, 8), device="cuda")
    with pytest.raises(_lib.AwkError):
        net.forward_batch([im] * 3, [im] * 3)


def test_atlas_warm_start_fallback_is_the_plain_run():
    import torch
    from atlaswerks_amd import atlas, warmstart
    from atlaswerks_amd.unet import UNetWarmStart
    n = 32
    _, subs = warmstart.atlas_group((n, n, n), 4, seed=77, max_disp=2.0)
    S = [atlas.AtlasSubject(s) for s in subs]
    levels = ((2, 10), (1, 6))
    net = UNetWarmStart((n, n, n), weights=_zero_head(), max_batch=4)
    ws = warmstart.WarmStartGreedyAtlas(net, S, (n, n, n), max_pert=0.2, plain_levels=levels,
                                        max_refine=3)
    res = ws.Run(target_energy=0.0)  # unreachable: the judge rejects the warm start
    assert res.fell_back and res.refine_iterations == 3
    plain = atlas.GreedyAtlasBuilderGPU(S, (n, n, n), max_pert=0.2, scale_levels=levels)
    plain.Run()
    assert np.array_equal(res.energies, plain.energies())
    assert torch.equal(res.builder.Mean(), plain.Mean())
    for i in range(len(S)):
        assert torch.equal(res.builder.HField(i), plain.HField(i))


def test_atlas_warm_start_accepts_a_reached_target():
    """A target the refinement reaches is accepted at the first iteration whose
    energy is at or below it, with no fallback."""
    from atlaswerks_amd import atlas, warmstart
    from atlaswerks_amd.unet import UNetWarmStart
    n = 32
    _, subs = warmstart.atlas_group((n, n, n), 3, seed=78, max_disp=2.0)
    S = [atlas.AtlasSubject(s) for s in subs]
    net = UNetWarmStart((n, n, n), weights=_zero_head(), max_batch=2)  # two batches of the 3
    ws = warmstart.WarmStartGreedyAtlas(net, S, (n, n, n), max_pert=0.2, max_refine=20)
    # the zero prediction starts at the identity: iterations 0..19 from there
    ref = atlas.GreedyAtlasBuilderGPU(S, (n, n, n), max_pert=0.2)
    ref.UpdateStepSizeNextIteration()
    ref.Iterate(20)
    e = ref.energies()
    target = float(e[7])
    res = ws.Run(target_energy=target)
    k = int(np.argmax(e <= target))
    assert not res.fell_back
    assert res.refine_iterations == k + 1
    assert np.array_equal(res.energies, e[:k + 1])


def test_pair_warm_start_fallback_is_the_plain_run():
    import torch
    from atlaswerks_amd import warmstart
    from atlaswerks_amd.greedy import GreedyIteratorParam, GreedyWarp
    from atlaswerks_amd.unet import UNetWarmStart
    n = 32
    t, subs = warmstart.atlas_group((n, n, n), 1, seed=79, max_disp=2.0, noise=0.0)
    prm = GreedyIteratorParam(max_pert=0.1)
    net = UNetWarmStart((n, n, n), weights=_zero_head())
    res = warmstart.warm_start_warp(net, t, subs[0], target_energy=0.0, max_refine=2, param=prm,
                                    n_plain=15)
    assert res.fell_back
    p = GreedyWarp(t, subs[0], 15, param=prm)
    p.RunWarp()
    assert np.array_equal(res.energies, p.data.energies())
    assert torch.equal(res.builder.data.Def1To0, p.data.Def1To0)


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained weights")
def test_trained_network_reaches_the_plain_ssd_sooner():
    """Held-out atlas group (seed outside the training groups), 16 subjects at
    128^3: the plain atlas over 1/4, 1/2, 1 with 100 / 50 / 25 iterations sets the
    target 1.01 x its final SSD.  The trained warm start reaches it without
    falling back in fewer than 25 full-resolution iterations; the zero network
    (identity start at full resolution) does not within 25."""
    from atlaswerks_amd import atlas, warmstart
    from atlaswerks_amd.unet import UNetWarmStart
    n = 128
    _, subs = warmstart.atlas_group((n, n, n), 16, seed=999)
    S = [atlas.AtlasSubject(s) for s in subs]
    ws_plain = warmstart.WarmStartGreedyAtlas(UNetWarmStart((n, n, n), weights=_zero_head(), max_batch=16),
                                              S, (n, n, n), max_refine=25)
    plain = ws_plain.RunPlain()
    e_plain = plain.energies()
    target = 1.01 * float(e_plain[-1])
    del plain
    net = UNetWarmStart((n, n, n), weights=np.load(WEIGHTS), max_batch=16)
    res = warmstart.WarmStartGreedyAtlas(net, S, (n, n, n), max_refine=25).Run(target)
    print("trained: iterations %d, e0 %.3f, target %.3f" % (res.refine_iterations, res.energies[0], target))
    assert not res.fell_back
    assert res.refine_iterations < 25
    zero = ws_plain.Run(target)
    print("zero net: fell back %s after %d" % (zero.fell_back, zero.refine_iterations))
    assert zero.fell_back or zero.refine_iterations > res.refine_iterations
