"""Learned warm start for greedy atlas building and pairwise greedy registration.

No reference counterpart.  The reference starts every registration from the
identity hfield or from an initial affine (GreedyDeformationData::SetScaleLevel,
GreedyDeformationData.cxx:186-210, mInitialAffine) and every atlas subject from
the identity (AtlasThread::AtlasThreadMain, AtlasThread.txx:95-181).  This module
replaces that initial hfield by a prediction and keeps plain greedy as the judge
and the fallback:

1. predict: the U-Net (``unet.UNetWarmStart``, bf16 MFMA) maps a pair (moving
   image I0, fixed image I1) on the template grid to a displacement u with
   I0(x + u(x)) ~ I1(x), up to 16 pairs per forward pass;
2. start: h0(x) = M [x + u(x), 1], with M the subject's grid map (the identity
   when the subject lies on the template grid), ``hfield.affineDisplace``;
3. refine: greedy iterations at full resolution from h0 until the energy is at
   most the target (``target_energy``), at most ``max_refine`` iterations;
4. judge and fall back: if the target is missed, the plain method runs from the
   identity with its own scale levels, and its result is returned (bit for bit
   what a plain run gives).

For an atlas, subject i is the moving image and the mean the fixed one
(I_i(h_i(x)) ~ mean(x), GreedyAtlasThreadCPU.cxx:115-143); the network sees the
subject resampled onto the template grid and the initial mean (the weighted
average of those resampled subjects, the mean the plain method starts from).

Training data (``make_atlas_training_set``) follow the atlas workload: groups of
subjects, each a common 8-blob template warped by its own sinusoidal displacement
(max 5 voxels) plus Rician noise 0.02, generated on the device; the targets are
the plain greedy atlas's final displacements u_i = h_i - x (HField3DUtils::
HtoDisplacement, HField3DUtils.txx:112-128) after its full schedule.
"""
import math
import time

import numpy as np
import torch

from . import atlas, hfield
from .greedy import GreedyDeformationDataGPU, GreedyIteratorGPU, GreedyIteratorParam, GreedyWarp

PLAIN_LEVELS = ((4, 100), (2, 50), (1, 25))


# ---------------------------------------------------------------- synthetic data
def _grid(shape, dev):
    nz, ny, nx = shape
    z = torch.arange(nz, device=dev, dtype=torch.float32).view(nz, 1, 1)
    y = torch.arange(ny, device=dev, dtype=torch.float32).view(1, ny, 1)
    x = torch.arange(nx, device=dev, dtype=torch.float32).view(1, 1, nx)
    return z, y, x


def blobs_device(shape, rng, n_blobs=8, sigma=(4.0, 10.0), amp=(0.5, 1.0), dev="cuda"):
    """Zero background plus n_blobs Gaussian blobs (centres uniform in the inner
    80 % of the grid, sigma U(4, 10) voxels, amplitude U(0.5, 1)); rng is a
    numpy Generator for the blob parameters, the image is computed on the device."""
    nz, ny, nx = shape
    z, y, x = _grid(shape, dev)
    img = torch.zeros(shape, device=dev, dtype=torch.float32)
    for _ in range(n_blobs):
        cx, cy, cz = rng.uniform(0.1 * nx, 0.9 * nx), rng.uniform(0.1 * ny, 0.9 * ny), rng.uniform(0.1 * nz, 0.9 * nz)
        s, a = rng.uniform(*sigma), rng.uniform(*amp)
        k = -1.0 / (2.0 * s * s)
        img += a * (torch.exp(k * (z - cz) ** 2) * torch.exp(k * (y - cy) ** 2) * torch.exp(k * (x - cx) ** 2))
    return img


def sinusoid_device(shape, max_disp, rng, n_modes=4, dev="cuda"):
    """u[3, z, y, x]: per component a sum of n_modes sinusoids with integer
    wavenumbers 1-3 per axis and random phases, scaled to max |u| = max_disp."""
    nz, ny, nx = shape
    z, y, x = _grid(shape, dev)
    u = torch.zeros((3,) + tuple(shape), device=dev, dtype=torch.float32)
    for c in range(3):
        for _ in range(n_modes):
            k = rng.integers(1, 4, size=3)
            phi, a = rng.uniform(0, 2 * math.pi), rng.uniform(0.5, 1.0)
            u[c] += a * torch.sin(2 * math.pi * (float(k[0]) * x / nx + float(k[1]) * y / ny
                                                 + float(k[2]) * z / nz) + phi)
    return u * (max_disp / u.norm(dim=0).max())


def rician_device(img, sigma, gen):
    """|img + n1 + i n2| with n1, n2 ~ N(0, sigma^2) (gen: a device torch.Generator)."""
    n1 = torch.randn(img.shape, device=img.device, generator=gen) * sigma
    n2 = torch.randn(img.shape, device=img.device, generator=gen) * sigma
    return torch.sqrt((img + n1) ** 2 + n2 ** 2)


def atlas_group(shape, n_subjects, seed, max_disp=5.0, noise=0.02, template=None, dev="cuda"):
    """(template, [subjects]): a common 8-blob template (seeded) and n_subjects
    images template(x + u_i(x)) with their own sinusoidal u_i plus Rician noise."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev).manual_seed(int(seed))
    if template is None:
        template = blobs_device(shape, rng, dev=dev)
    nz, ny, nx = shape
    h = torch.empty((3, nz, ny, nx), device=dev, dtype=torch.float32)
    subs = []
    for _ in range(n_subjects):
        hfield.affineDisplace(h, sinusoid_device(shape, max_disp, rng, dev=dev))
        im = torch.empty(shape, device=dev, dtype=torch.float32)
        hfield.apply(im, template, h)
        subs.append(rician_device(im, noise, gen) if noise > 0 else im)
    return template, subs


def weighted_mean(images, weights=None):
    w = weights or [1.0 / len(images)] * len(images)
    m = torch.zeros_like(images[0])
    for wi, im in zip(w, images):
        m += wi * im
    return m


def make_atlas_training_set(n_pairs, shape=(128, 128, 128), group=16, seed=4, max_pert=0.1,
                            levels=PLAIN_LEVELS, dev="cuda", log=None):
    """Training triples (moving, fixed, target displacement) from plain atlases.

    Groups of ``group`` subjects (``atlas_group``, group g seeded seed * 100000 + g);
    each group's plain greedy atlas runs its full schedule; subject i gives
    (I_i, m0, h_i - x) with m0 the group's initial mean.  Returns the list of
    triples and a summary dict."""
    nz, ny, nx = shape
    ident = torch.empty((3, nz, ny, nx), device=dev, dtype=torch.float32)
    hfield.setToIdentity(ident)
    out, t0, g = [], time.time(), 0
    final_energy = []
    while len(out) < n_pairs:
        k = min(group, n_pairs - len(out))
        k = max(k, 2)
        _, subs = atlas_group(shape, k, seed * 100000 + g)
        m0 = weighted_mean(subs)
        b = atlas.GreedyAtlasBuilderGPU([atlas.AtlasSubject(s) for s in subs], (nx, ny, nz),
                                        max_pert=max_pert, scale_levels=levels)
        b.Run()
        final_energy.append(float(b.energies()[-1]))
        for i, s in enumerate(subs):
            if len(out) < n_pairs:
                out.append((s, m0, (b.HField(i) - ident).contiguous()))
        del b
        g += 1
        if log and g % 10 == 0:
            log("training set: %d groups, %d pairs, %.1f s" % (g, len(out), time.time() - t0))
    torch.cuda.synchronize()
    return out, {"pairs": len(out), "groups": g, "group_size": group, "size": [nx, ny, nz],
                 "generation_s": round(time.time() - t0, 2),
                 "plain_final_energy_mean": round(float(np.mean(final_energy)), 4)}


def train(net, triples, steps, lr=3e-4, lr_final=3e-5, seed=0, log=None):
    """Adam on the device, one triple per step in a seeded shuffled order, the
    learning rate decayed geometrically from lr to lr_final.  Returns the trainer
    and the mean loss per epoch."""
    from .unet import UNetTrainer
    tr = UNetTrainer(net)
    rng = np.random.default_rng(seed)
    n = len(triples)
    order = []
    losses = []
    t0 = time.time()
    last = 0.0
    for k in range(steps):
        if not order:
            order = list(rng.permutation(n))
        i = order.pop()
        mov, fix, tgt = triples[i]
        rate = lr * (lr_final / lr) ** (k / max(steps - 1, 1))
        tr.step(mov, fix, tgt, lr=rate)
        if (k + 1) % n == 0 or k + 1 == steps:
            tot = float(tr.loss.item())
            cnt = (k % n) + 1
            losses.append((tot - last) / cnt)
            last = tot
            if log:
                log("train: step %d, epoch loss %.5f, %.1f s" % (k + 1, losses[-1], time.time() - t0))
    return tr, losses


# ---------------------------------------------------------------- product path
class WarmStartResult:
    def __init__(self, builder, fell_back, refine_iterations, energies, predict_ms):
        self.builder = builder              # the atlas (or warp data) that produced the result
        self.fell_back = fell_back          # True: the plain method's own result
        self.refine_iterations = refine_iterations
        self.energies = energies            # energies of the returned run (numpy)
        self.predict_ms = predict_ms


class WarmStartGreedyAtlas:
    """Greedy atlas started from U-Net predictions, with plain greedy as the
    judge and the fallback.

    subjects, template grid and parameters are as for
    ``atlas.GreedyAtlasBuilderGPU``; ``net`` is a ``unet.UNetWarmStart`` on the
    template grid.  ``Run(target_energy)`` returns a ``WarmStartResult``."""

    def __init__(self, net, subjects, template_size, template_origin=(0.0, 0.0, 0.0),
                 template_spacing=(1.0, 1.0, 1.0), max_pert=0.1, alpha=0.01, beta=0.01,
                 gamma=0.001, use_graph=True, plain_levels=PLAIN_LEVELS, max_refine=50,
                 batch=None):
        self.net = net
        self.subjects = list(subjects)
        self.size = tuple(int(v) for v in template_size)
        if tuple(net.size) != self.size:
            raise ValueError("the network's grid %s is not the template grid %s" % (net.size, self.size))
        self.kw = dict(template_origin=template_origin, template_spacing=template_spacing,
                       max_pert=max_pert, alpha=alpha, beta=beta, gamma=gamma, use_graph=use_graph,
                       batch=batch)
        self.plain_levels = tuple(plain_levels)
        self.max_refine = int(max_refine)

    def _plain(self):
        return atlas.GreedyAtlasBuilderGPU(self.subjects, self.size, scale_levels=self.plain_levels,
                                           **self.kw)

    def RunPlain(self):
        """The plain method: identity start, its scale levels."""
        b = self._plain()
        b.Run()
        return b

    def _base_map(self, s):
        o, sp = self.kw["template_origin"], self.kw["template_spacing"]
        # (subjects' grid maps at full resolution, as the builder forms them)
        return hfield.grid_affine(o, sp, s.origin, s.spacing, s.affine)

    def predict(self):
        """Initial hfields h_i = M_i [x + u_i, 1] on the template grid: every
        subject resampled through its grid map M_i, the initial mean of those
        (weights as the atlas takes them), then U-Net(subject_i, mean) in batches
        of up to net.max_batch pairs."""
        nx, ny, nz = self.size
        ns = len(self.subjects)
        weights = [s.weight if s.weight is not None else 1.0 / ns for s in self.subjects]
        h = torch.empty((3, nz, ny, nx), dtype=torch.float32, device="cuda")
        bases, moving = [], []
        for s in self.subjects:
            M = self._base_map(s)
            hfield.initializeFromAffine(h, M)
            im = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
            hfield.apply(im, s.image, h)
            moving.append(im)
            bases.append(M)
        del h
        m0 = weighted_mean(moving, weights)
        fields = []
        B = self.net.max_batch
        for k in range(0, ns, B):
            chunk = moving[k:k + B]
            u = self.net.forward_batch(chunk, [m0] * len(chunk))
            for j in range(len(chunk)):
                f = u[j]
                hfield.affineDisplace(f, f, bases[k + j])
                fields.append(f)
        return fields

    def Run(self, target_energy):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        h0s = self.predict()
        t1.record()
        b = atlas.GreedyAtlasBuilderGPU(self.subjects, self.size, initial_hfields=h0s, **self.kw)
        del h0s
        b.UpdateStepSizeNextIteration()
        reached, it = False, 0
        while it < self.max_refine:
            b.Iterate(1)
            it += 1
            if float(b._energy[-1][0].item()) <= target_energy:
                reached = True
                break
        t1.synchronize()
        predict_ms = t0.elapsed_time(t1)
        if reached:
            return WarmStartResult(b, False, it, b.energies(), predict_ms)
        del b
        p = self.RunPlain()
        return WarmStartResult(p, True, it, p.energies(), predict_ms)


def warm_start_warp(net, I0, I1, target_energy, max_refine=50, param=None, n_plain=50,
                    plain_levels=None):
    """Pairwise greedy registration from h0 = x + U-Net(I0, I1): iterations at full
    resolution until the energy is at most target_energy (at most max_refine); if
    missed, the plain GreedyWarp from the identity (n_plain iterations, or
    plain_levels).  Returns a ``WarmStartResult``: its builder is the
    GreedyDeformationDataGPU of the refinement, or the plain GreedyWarp."""
    param = param or GreedyIteratorParam()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    h0 = net.forward(I0, I1)
    hfield.affineDisplace(h0, h0)
    t1.record()
    nz, ny, nx = I0.shape
    data = GreedyDeformationDataGPU(I0, I1, max_iters=max(max_refine, 1), h0=h0)
    del h0
    it_ = GreedyIteratorGPU((nx, ny, nz), (1.0, 1.0, 1.0), param)
    it_.UpdateStepSizeNextIteration()
    reached, it = False, 0
    while it < max_refine:
        it_.Iterate(data)
        it += 1
        if float(data._energy[data.cur_iter - 1].item()) <= target_energy:
            reached = True
            break
    t1.synchronize()
    predict_ms = t0.elapsed_time(t1)
    if reached:
        return WarmStartResult(data, False, it, data.energies(), predict_ms)
    del data
    p = GreedyWarp(I0, I1, n_plain, param=param, scale_levels=plain_levels)
    p.RunWarp()
    return WarmStartResult(p, True, it, p.data.energies(), predict_ms)
