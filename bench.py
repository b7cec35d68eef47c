"""Greedy-fluid iteration throughput at 256^3 on one MI355X.

One step = one greedy-fluid iteration (GreedyIteratorGPU::Iterate semantics:
apply, gradient/residual/force/SSD, FluidKernelFFT L^-1, step-scaled composeHV)
of a 256^3 float32 pairwise registration, inputs resident in HBM.  The K timed
iterations run back to back as GreedyWarp's inner loop does (one call of
awk_greedy_iterate_n: from the second iteration on, each composition also forms
the next iteration's deformed image, bit-identical to separate iterations); the
same K iterations as K calls of the per-iteration entry point (awk_greedy_iterate)
are reported beside them ("per_iteration_calls").  Prints one
JSON line with the iteration rate, the roofline of the dominant kernel (measured
live with HIP events on the iteration's stream), per-stage times, the kernel
microbenchmarks of the 256^3 workload (L^-1, apply, compose, splat) and the CPU
baseline (the oracle's FP32 restatement of the reference CPU iteration, timed on
a bounded sample: all host cores with one registration per thread, and one
thread alone).  The second half of the metric, time-to-SSD of the 16-subject
atlas, plain vs U-Net warm start, is the "warm_start_atlas_16" object.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def stage_bytes(nx, ny, nz):
    """Algorithmic HBM bytes of one launch of each stage of the fused iteration.

    N voxels, spectrum = 3 components x (nx/2+1) * ny * nz complex64."""
    N = nx * ny * nz
    spec = 3 * (nx // 2 + 1) * ny * nz * 8
    return {
        "apply": 20 * N,                         # read h 12N + I0 4N, write Idef 4N
        "xfwd_force": 8 * N + spec,              # read Idef + I1, write spectrum
        "y_fwd": 2 * spec,                       # read + write spectrum
        "z_solve": 2 * spec,
        "y_inv": 2 * spec,
        "xinv_compose": spec + 24 * N,           # read spectrum + h, write h
        # read spectrum + h + I0, write h + the next deformed image
        "xinv_compose_apply": spec + 32 * N,
    }


def _events_ms(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_ldmm(dev, n=128, T=10, subjects=8, reps=5, orig=True):
    """One LDMM iteration at n^3 with T velocity fields (IterateOrig, the
    reference's default gradient, or IterateNew); subjects run back to back."""
    import torch
    from atlaswerks_amd import hfield, ldmm, synth
    shape = (n, n, n)
    t = torch.from_numpy(synth.blobs(shape, seed=3)).to(dev)
    warps = []
    for s in range(subjects):
        u = torch.from_numpy(synth.identity(shape) + synth.sinusoidal_displacement(shape, 3.0, seed=300 + s)).to(dev)
        tgt = torch.empty_like(t)
        hfield.apply(tgt, t, u)
        prm = ldmm.LDMMIteratorParam(sigma=0.05, use_orig_gradient=orig)
        w = ldmm.LDMMWarp(t, tgt, n_time_steps=T, n_iter=0, param=prm)
        w.RunWarp()  # step-size call
        warps.append(w)

    def one():
        for w in warps:
            w.iterator.Iterate(w.data, record_energy=False)
    ms = _events_ms(one, reps)
    N = n ** 3
    return {"gradient": "IterateOrig" if orig else "IterateNew",
            "ms_per_iteration_all_subjects": round(ms, 3), "subjects": subjects, "T": T,
            "subject_iter_per_s": round(1000.0 * subjects / ms, 1),
            "ms_per_subject_iteration": round(ms / subjects, 3),
            "size": [n, n, n], "N": N}


def bench_ldmm_atlas(dev, n=128, S=8, T=10, n_iter=40):
    """LDMM atlas: S subjects at n^3 (a common 8-blob template warped by its own
    sinusoidal displacement, max 5 voxels, plus Rician noise 0.02), T = 10 velocity
    fields, sigma 0.05, alpha 0.01, beta 0.01, gamma 0.001, n_iter iterations; each
    gradient / splat variant timed as one whole build (step-size calls, n_iter
    iterations and the final mean)."""
    import torch
    from atlaswerks_amd import hfield, ldmm, synth
    shape = (n, n, n)
    base = synth.blobs(shape, seed=2)
    ims = []
    for s in range(S):
        h = synth.identity(shape) + synth.sinusoidal_displacement(shape, 5.0, seed=3000 + s)
        ht = torch.from_numpy(h).to(dev)
        im = torch.empty(shape, dtype=torch.float32, device=dev)
        hfield.apply(im, torch.from_numpy(base).to(dev), ht)
        ims.append(torch.from_numpy(synth.rician(im.cpu().numpy(), 0.02, seed=4000 + s)).to(dev))
        del ht
    out = {"subjects": S, "size": [n, n, n], "T": T, "iterations": n_iter}
    for name, orig, splat in (("IterateOrig", True, "fixed"), ("IterateNew_fixed64", False, "fixed"),
                              ("IterateNew_sortscan", False, "sorted")):
        prm = ldmm.LDMMIteratorParam(sigma=0.05, max_pert=0.1, alpha=0.01, beta=0.01, gamma=0.001,
                                     use_orig_gradient=orig, splat=splat)
        b = ldmm.LDMMAtlasBuilderGPU(ims, n_time_steps=T, param=prm)
        b.Mean()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        b.Iterate(n_iter, vec_energy=False)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        en = b.energies()
        out[name] = {"build_s": round(ms / 1000.0, 3), "ms_per_iteration": round(ms / n_iter, 3),
                     "subject_iter_per_s": round(1000.0 * S * n_iter / ms, 1),
                     "image_energy_first_last": [round(float(en[0, 0]), 3), round(float(en[-1, 0]), 3)]}
        del b
        torch.cuda.empty_cache()
    return out


def atlas_subjects(dev):
    """The 16-subject atlas workload: 6 x 96^3, 6 x 128^3, 4 x 160x192x144 (x, y,
    z) of one world-space 8-blob template (128 world units), each on its own grid
    warped by its own sinusoidal displacement (max 5 voxels) plus Rician noise
    0.02."""
    import torch
    from atlaswerks_amd import atlas, hfield, synth
    world = 128.0
    grids = [(96, 96, 96)] * 6 + [(128, 128, 128)] * 6 + [(144, 192, 160)] * 4  # (z, y, x)
    base = torch.from_numpy(synth.blobs((128, 128, 128), seed=2)).to(dev)
    subs = []
    for k, g in enumerate(grids):
        nz, ny, nx = g
        sp = (world / nx, world / ny, world / nz)
        h = torch.empty((3, nz, ny, nx), dtype=torch.float32, device=dev)
        hfield.initializeFromAffine(h, hfield.grid_affine((0, 0, 0), sp, (0, 0, 0), (1, 1, 1)))
        u = torch.from_numpy(synth.sinusoidal_displacement(g, 5.0, seed=2000 + k)).to(dev)
        h += u
        im = torch.empty(g, dtype=torch.float32, device=dev)
        hfield.apply(im, base, h)
        im = torch.from_numpy(synth.rician(im.cpu().numpy(), 0.02, seed=5000 + k)).to(dev)
        subs.append(atlas.AtlasSubject(im, spacing=sp))
    return subs


def bench_atlas(dev, subs, reps=10):
    """The 16-subject atlas on a 128^3 template grid: per-iteration time at full
    resolution (batched launches eager and through hipGraph replays, and one launch
    per subject and pass through graph replays), then one whole 3-level build (1/4,
    1/2, 1 with 100 / 50 / 25 iterations)."""
    import torch
    from atlaswerks_amd import atlas
    from atlaswerks_amd import autotune
    out = {}
    levels = [(4, 100), (2, 50), (1, 25)]
    # subjects per batched launch of the greedy step, per level (autotuned on a
    # scratch atlas; 0 = one launch per subject and pass on 3 stream lanes)
    batch, bt_ms = autotune.tune_atlas_batch(subs, (128, 128, 128), levels, max_pert=0.1)
    out["batch_tuning"] = {"chosen": {str(f): v for f, v in batch.items()},
                           "ms": {str(f): {str(c): round(m, 4) for c, m in t.items()}
                                  for f, t in bt_ms.items()}}
    for name, use_graph, bsz in (("eager", False, batch[1]), ("graph", True, batch[1]),
                                 ("graph_per_subject", True, 0)):
        b = atlas.GreedyAtlasBuilderGPU(subs, (128, 128, 128), max_pert=0.1, use_graph=use_graph,
                                        batch=bsz)
        b.RunScaleLevel(2)  # step sizes + graph capture
        ms = _events_ms(lambda: b.Iterate(1), reps)
        out[name] = {"ms_per_atlas_iteration": round(ms, 3)}
        del b
    out["subjects"] = len(subs)
    out["template"] = [128, 128, 128]
    out["subject_iter_per_s_graph"] = round(1000.0 * len(subs) / out["graph"]["ms_per_atlas_iteration"], 1)
    for name, bsz in (("multiscale_build", batch), ("multiscale_build_per_subject", 0)):
        b = atlas.GreedyAtlasBuilderGPU(subs, (128, 128, 128), max_pert=0.1, use_graph=True,
                                        scale_levels=levels, batch=bsz)
        torch.cuda.synchronize()
        t0 = time.time()
        b.Run()
        torch.cuda.synchronize()
        e = b.energies()
        out[name] = {"levels": [list(x) for x in levels], "build_s": round(time.time() - t0, 4),
                     "energy_first_last": [round(float(e[0]), 4), round(float(e[-1]), 4)]}
        del b
    return out


def bench_splat(dev, n=256, reps=20):
    """The fixed-point (FIXED64) splat of the 256^3 workload alone: an 8-blob image
    splatted along identity + sinusoidal displacement (max 8 voxels)."""
    import torch
    from atlaswerks_amd import _lib, hfield, synth
    shape = (n, n, n)
    img = torch.from_numpy(synth.blobs(shape, seed=1)).to(dev)
    h = torch.from_numpy(synth.identity(shape) + synth.sinusoidal_displacement(shape, 8.0, seed=1)).to(dev)
    out = torch.empty_like(img)
    ms = _events_ms(lambda: hfield.splat(out, img, h, mode=_lib.SPLAT_FIXED64), reps)
    N = n ** 3
    return {"ms": round(ms, 4), "gbs_20N": round(20 * N / (ms * 1e-3) / 1e9, 1), "size": [n, n, n]}


def bench_unet(dev, n=128, reps=10, train_reps=5):
    """U-Net warm start at n^3 (2 input channels -> 3-channel displacement): one
    forward pass, and one training step (forward, backward, Adam)."""
    import torch
    from atlaswerks_amd import synth
    from atlaswerks_amd.unet import LAYERS, UNetTrainer, UNetWarmStart
    shape = (n, n, n)
    t = torch.from_numpy(synth.blobs(shape, seed=1)).to(dev)
    s = torch.from_numpy(synth.blobs(shape, seed=2)).to(dev)
    net = UNetWarmStart((n, n, n), seed=0)
    out = torch.empty((3, n, n, n), dtype=torch.float32, device=dev)
    ms = _events_ms(lambda: net.forward(t, s, out=out), reps)
    N0 = n ** 3
    res = [N0, N0, N0 // 8, N0 // 8, N0 // 64, N0 // 64, N0 // 512, N0 // 512,
           N0 // 64, N0 // 64, N0 // 8, N0 // 8, N0, N0]
    flops = sum(2.0 * N * 27 * cin * cout for N, (cin, cout) in zip(res, LAYERS)) + 2.0 * N0 * 16 * 3
    r = {"ms": round(ms, 3), "gflop": round(flops / 1e9, 1),
         "tflops": round(flops / (ms * 1e-3) / 1e12, 1), "peak_bf16_tflops": 2500.0,
         "frac": round(flops / (ms * 1e-3) / 2.5e15, 4), "size": [n, n, n]}
    if train_reps:
        tr = UNetTrainer(net)
        tgt = torch.zeros((3, n, n, n), device=dev)
        tms = _events_ms(lambda: tr.step(t, s, tgt, lr=1e-6), train_reps)
        # forward + input-gradient (dgrad) + weight-gradient (wgrad) convolutions: 3x the forward
        r["train_step_ms"] = round(tms, 3)
        r["train_tflops"] = round(3 * flops / (tms * 1e-3) / 1e12, 1)
    return r


def bench_warm_start(dev, subs, max_refine=50, batch=None):
    """Time-to-SSD of the 16-subject atlas, plain vs U-Net warm start.

    Plain: the greedy atlas from the identity over scale levels 1/4, 1/2, 1 with
    100 / 50 / 25 iterations (hipGraph replays); its final energy E sets the target
    1.01 E.  Warm start (atlaswerks_amd.warmstart.WarmStartGreedyAtlas): the U-Net,
    trained on the device by atlaswerks_amd.train_warmstart (2000 pairs at 128^3,
    weights in atlaswerks_amd/weights), predicts all 16 initial hfields in one
    batch-16 pass over (subject resampled onto the 128^3 template grid, initial
    mean); greedy atlas iterations at full resolution then run until the energy
    reaches the target (judged after every iteration), else the plain build runs as
    the fallback.  Both are timed whole (wall clock between device
    synchronisations) after one untimed run of each."""
    import numpy as np
    import torch
    from atlaswerks_amd import warmstart
    from atlaswerks_amd.unet import UNetTrainer, UNetWarmStart
    path = os.path.join(ROOT, "atlaswerks_amd", "weights", "unet_atlas_128.npy")
    if not os.path.exists(path):
        return {"skipped": "no trained weights at %s" % os.path.relpath(path, ROOT)}
    n = 128
    net = UNetWarmStart((n, n, n), weights=np.load(path), max_batch=16)
    ws = warmstart.WarmStartGreedyAtlas(net, subs, (n, n, n), max_pert=0.1,
                                        plain_levels=warmstart.PLAIN_LEVELS, max_refine=max_refine,
                                        batch=batch)

    def wall(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, 1000.0 * (time.perf_counter() - t)

    b = ws.RunPlain()  # untimed: first-use costs (allocator pools, code objects)
    ws.Run(1.01 * float(b.energies()[-1]))
    del b
    plain, plain_ms = wall(ws.RunPlain)
    e_plain = plain.energies()
    target = 1.01 * float(e_plain[-1])
    del plain
    res, warm_ms = wall(lambda: ws.Run(target))
    out = {"subjects": len(subs), "template": [n, n, n],
           "plain_levels": [list(v) for v in warmstart.PLAIN_LEVELS],
           "plain_ms": round(plain_ms, 2), "plain_iterations": len(e_plain),
           "plain_final_energy": round(float(e_plain[-1]), 4), "target_energy": round(target, 4),
           "warm_ms": round(warm_ms, 2), "warm_fell_back": bool(res.fell_back),
           "warm_refine_iterations": int(res.refine_iterations),
           "warm_energy_first_last": [round(float(res.energies[0]), 4), round(float(res.energies[-1]), 4)],
           "warm_predict_ms": round(res.predict_ms, 3),
           "warm_over_plain": round(warm_ms / plain_ms, 4)}
    del res
    # batch-16 inference against 16 single passes, and the training step
    moving = [torch.rand((n, n, n), device=dev) for _ in range(16)]
    m0 = torch.rand((n, n, n), device=dev)
    outb = torch.empty((16, 3, n, n, n), device=dev)
    out["unet_batch16_ms"] = round(_events_ms(lambda: net.forward_batch(moving, [m0] * 16, out=outb), 3), 3)
    single = UNetWarmStart((n, n, n), weights=np.load(path))
    o1 = torch.empty((3, n, n, n), device=dev)
    out["unet_16_singles_ms"] = round(_events_ms(lambda: [single.forward(mv, m0, out=o1) for mv in moving], 3), 3)
    tr = UNetTrainer(single)
    tgt = torch.zeros((3, n, n, n), device=dev)
    out["unet_train_ms_per_step"] = round(_events_ms(lambda: tr.step(moving[0], m0, tgt, lr=1e-6), 10), 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-microbench", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--only", choices=("ldmm", "ldmm_new", "ldmm_atlas", "atlas", "unet", "warm_start",
                                       "splat"),
                    help="run one secondary workload alone (for profiling) and print its JSON")
    ap.add_argument("--copy-ref", action="store_true",
                    help="also run a plain 1 GiB device copy (counter scaling reference)")
    ap.add_argument("--tune", action="append", default=[], metavar="KNOB=VALUE",
                    help="set a library knob (awk_tune_set) before running")
    args = ap.parse_args()
    if args.tune:
        from atlaswerks_amd import autotune
        for kv in args.tune:
            k, v = kv.split("=")
            autotune.set_knob(k, int(v))
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU only (--gpus %d requested); across GPUs it "
              "would shard an atlas by subject, one process per GPU, with one all-reduce of the "
              "mean image and energy per atlas iteration" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.only:
        import torch
        dev = torch.device("cuda:0")
        fn = {"ldmm": lambda: bench_ldmm(dev, subjects=2, reps=3),
              "ldmm_new": lambda: bench_ldmm(dev, subjects=2, reps=3, orig=False),
              "ldmm_atlas": lambda: bench_ldmm_atlas(dev),
              "atlas": lambda: bench_atlas(dev, atlas_subjects(dev), reps=3),
              "unet": lambda: bench_unet(dev),
              "warm_start": lambda: bench_warm_start(dev, atlas_subjects(dev)),
              "splat": lambda: bench_splat(dev)}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return

    import numpy as np
    import torch

    from atlaswerks_amd import _lib, hfield, synth
    from atlaswerks_amd.fluid import DiffOperGPU
    from atlaswerks_amd.greedy import GreedyDeformationDataGPU, GreedyIteratorGPU, GreedyIteratorParam

    _lib.load()
    dev = torch.device("cuda:0")
    n = args.size
    shape = (n, n, n)
    nx = ny = nz = n
    N = n ** 3
    # ---- synthetic workload: 8-blob template, target = template(x + u), |u| <= 8
    t0 = time.time()
    tmpl = synth.blobs(shape, seed=1)
    u = synth.sinusoidal_displacement(shape, max_disp=8.0, seed=1)
    I0 = torch.from_numpy(tmpl).to(dev)
    h = torch.from_numpy(synth.identity(shape) + u).to(dev)
    I1 = torch.empty_like(I0)
    hfield.apply(I1, I0, h)
    del h
    torch.cuda.synchronize()
    gen_s = time.time() - t0

    tuning = None
    if not args.no_autotune:
        from atlaswerks_amd import autotune
        # FFT plan of the grid (radix order, pass tiles), then the gather knobs
        plan, plan_ms = autotune.tune_fft((nx, ny, nz), reps=5)
        tuning, tune_ms = autotune.tune((nx, ny, nz), reps=5)
        tuning = {"chosen": dict(tuning, **plan),
                  "ms": {k: {str(a): round(b, 4) for a, b in v.items()}
                         for k, v in list(tune_ms.items()) + list(plan_ms.items())}}
    param = GreedyIteratorParam(max_pert=0.1, alpha=0.01, beta=0.01, gamma=0.001)
    it = GreedyIteratorGPU((nx, ny, nz), (1.0, 1.0, 1.0), param)
    data = GreedyDeformationDataGPU(I0, I1, max_iters=args.warmup + args.steps + 1)
    it.UpdateStepSizeNextIteration()
    it.Iterate(data)  # sets the step (MaxPert / max |v|)
    it.IterateN(data, args.warmup)
    torch.cuda.synchronize()
    # timed region: K iterations back to back in one awk_greedy_iterate_n call
    # (GreedyWarp's inner loop: from the second iteration on each composition also
    # forms the next deformed image), nothing else on the stream
    ts = time.perf_counter()
    it.IterateN(data, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - ts
    energies = data.energies()
    # per-kernel attribution: a second pass with start/stop events attached to
    # each kernel dispatch (kernel durations, as rocprof reports them)
    n_att = min(args.steps, 50)
    it.SetProfiling(True)
    it.IterateN(data, n_att)
    prof, n_prof, launches = it.GetProfile(launches=True)
    it.SetProfiling(False)
    ms_per_step = 1e3 * elapsed / args.steps
    iters_per_s = args.steps / elapsed
    # the same K iterations as K calls of the per-iteration entry point
    # (awk_greedy_iterate, the reference's GreedyIteratorGPU::Iterate), and their
    # kernel attribution
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        it.Iterate(data)
    torch.cuda.synchronize()
    elapsed_1 = time.perf_counter() - ts
    it.SetProfiling(True)
    for _ in range(n_att):
        it.Iterate(data)
    prof_1, n_prof_1, launches_1 = it.GetProfile(launches=True)
    it.SetProfiling(False)

    # per iteration: each stage's share; per launch: the kernel's mean duration
    stage_avg = {k: v / max(n_prof, 1) for k, v in prof.items()}
    launch_avg = {k: prof[k] / launches[k] for k in prof if launches[k]}
    sb = stage_bytes(nx, ny, nz)
    dom = max(stage_avg, key=lambda k: stage_avg[k])
    achieved = sb[dom] / (launch_avg[dom] * 1e-3) / 1e9
    stage_avg_1 = {k: v / max(n_prof_1, 1) for k, v in prof_1.items()}
    launch_avg_1 = {k: prof_1[k] / launches_1[k] for k in prof_1 if launches_1[k]}
    dom_1 = max(stage_avg_1, key=lambda k: stage_avg_1[k])
    achieved_1 = sb[dom_1] / (launch_avg_1[dom_1] * 1e-3) / 1e9
    # HBM bytes per launch of the dominant kernel, measured by rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE (scaled by the copy calibration) and recorded in
    # profiles/pmc_traffic_<n>.txt; null when no measurement exists for this size.
    def traffic_of(stage):
        tpath = os.path.join(ROOT, "profiles", "pmc_traffic_%d.txt" % n)
        if os.path.exists(tpath):
            with open(tpath) as f:
                for line in f:
                    parts = line.split()
                    if len(parts) == 4 and parts[0] == stage:
                        return int(parts[2]) + int(parts[3])
        return None
    traffic = traffic_of(dom)
    yard = 196 * N  # fixed yardstick bytes of one greedy iteration
    out = {
        "metric": "greedy-fluid iter/s @256^3 (% HBM peak); time-to-SSD plain vs warm-start",
        "value": round(iters_per_s, 3),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "greedy-fluid pairwise iteration %d^3 float32: 8-blob template, target "
                        "warped by sinusoidal displacement (max 8 voxels), alpha=beta=0.01, "
                        "gamma=0.001, MaxPert 0.1" % n,
            "size": [nx, ny, nz],
        },
        "hbm_frac_yardstick": round(yard * iters_per_s / (HBM_PEAK_GBS * 1e9), 4),
        "yardstick_bytes_per_iter": yard,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": sb[dom],
            "avg_launch_ms": round(launch_avg[dom], 5),
        },
        "stages_ms": {k: round(v, 5) for k, v in stage_avg.items()},
        "stages_launches": {k: launches[k] for k in launches},
        "stages_sum_ms": round(sum(stage_avg.values()), 5),
        "stages_gbs": {k: round(sb[k] / (v * 1e-3) / 1e9, 1) for k, v in launch_avg.items() if sb[k]},
        "per_iteration_calls": {
            "entry_point": "awk_greedy_iterate",
            "iter_per_s": round(args.steps / elapsed_1, 3),
            "ms_per_step": round(1e3 * elapsed_1 / args.steps, 4),
            "roofline": {"bound": "hbm", "kernel": dom_1, "achieved": round(achieved_1, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved_1 / HBM_PEAK_GBS, 4),
                         "traffic": traffic_of(dom_1), "alg_bytes_per_launch": sb[dom_1],
                         "avg_launch_ms": round(launch_avg_1[dom_1], 5)},
            "stages_ms": {k: round(v, 5) for k, v in stage_avg_1.items() if launches_1[k]},
        },
        "energy_first_last": [float(energies[0]), float(energies[-1])],
        "input_gen_s": round(gen_s, 2),
        "tuning": tuning,
    }

    # ---- kernel microbenchmarks of the 256^3 workload (config-1 operations)
    if not args.no_microbench:
        def timeit(fn, reps=20):
            fn()
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                fn()
            b.record()
            torch.cuda.synchronize()
            return a.elapsed_time(b) / reps

        mb = {}
        f = torch.randn((3, nz, ny, nx), device=dev)
        op = DiffOperGPU((nx, ny, nz))
        ms = timeit(lambda: op.ApplyInverseOperator(f))
        mb["L_inverse"] = {"ms": round(ms, 4), "gbs_120N": round(120 * N / (ms * 1e-3) / 1e9, 1)}
        g1 = torch.from_numpy(synth.identity(shape) + u).to(dev)
        g2 = torch.from_numpy(synth.identity(shape) + synth.sinusoidal_displacement(shape, 8.0, seed=2)).to(dev)
        outi = torch.empty_like(I0)
        ms = timeit(lambda: hfield.apply(outi, I0, g1))
        mb["apply"] = {"ms": round(ms, 4), "gbs_20N": round(20 * N / (ms * 1e-3) / 1e9, 1)}
        outv = torch.empty_like(g1)
        ms = timeit(lambda: hfield.compose(outv, g1, g2, hfield.BACKGROUND_STRATEGY_PARTIAL_ID))
        mb["compose"] = {"ms": round(ms, 4), "gbs_36N": round(36 * N / (ms * 1e-3) / 1e9, 1)}
        for name, mode in (("splat_fixed64", _lib.SPLAT_FIXED64), ("splat_float", _lib.SPLAT_FLOAT)):
            ms = timeit(lambda: hfield.splat(outi, I0, g1, mode=mode))
            mb[name] = {"ms": round(ms, 4), "gbs_20N": round(20 * N / (ms * 1e-3) / 1e9, 1)}
        vel = torch.from_numpy(synth.sinusoidal_displacement(shape, 8.0, seed=3)).to(dev)
        ms = timeit(lambda: hfield.splat_sorted(outi, I0, vel), reps=5)
        mb["splat_sorted"] = {"ms": round(ms, 4), "gbs_20N": round(20 * N / (ms * 1e-3) / 1e9, 1)}
        # FFT plan autotuning on a grid of mixed radices (the largest subject grid
        # of the atlas workload), then back to the default plan
        if not args.no_autotune:
            from atlaswerks_amd import autotune
            saved = {k: autotune.get_knob(k) for k in autotune.PLAN_KNOBS}
            g = (160, 192, 144)
            fm = torch.randn((3, g[2], g[1], g[0]), device=dev)
            for k in autotune.PLAN_KNOBS:
                autotune.set_knob(k, 0)
            op = DiffOperGPU(g)
            d_ms = timeit(lambda: op.ApplyInverseOperator(fm))
            del op
            plan, _ = autotune.tune_fft(g, reps=5)
            op = DiffOperGPU(g)
            t_ms = timeit(lambda: op.ApplyInverseOperator(fm))
            del op, fm
            Ng = g[0] * g[1] * g[2]
            mb["L_inverse_160x192x144"] = {"default_plan_ms": round(d_ms, 4), "tuned_plan_ms": round(t_ms, 4),
                                           "tuned_plan": plan,
                                           "gbs_120N_tuned": round(120 * Ng / (t_ms * 1e-3) / 1e9, 1)}
            for k, v in saved.items():
                autotune.set_knob(k, v)
        out["microbench"] = mb
        del f, g1, g2, outi, outv, vel

    # ---- secondary workloads: LDMM (8 subjects x 128^3, T = 10) and a 16-subject
    #      greedy atlas over uneven grids (template 128^3); not part of the headline
    if not args.no_secondary:
        out["ldmm_128_T10"] = bench_ldmm(dev)
        out["ldmm_new_128_T10"] = bench_ldmm(dev, orig=False)
        out["ldmm_atlas_8x128_T10"] = bench_ldmm_atlas(dev)
        subs = atlas_subjects(dev)
        out["atlas_16_uneven"] = bench_atlas(dev, subs)
        out["unet_128"] = bench_unet(dev)
        tuned = out["atlas_16_uneven"]["batch_tuning"]["chosen"]
        out["warm_start_atlas_16"] = bench_warm_start(dev, subs, batch={int(f): v for f, v in tuned.items()})
        del subs

    if args.copy_ref:
        # plain 1 GiB device copy: the calibration point for the HBM byte counters
        a = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 20
        out["copy_1GiB"] = {"ms": round(ms, 4), "gbs_rw": round(2 * a.numel() * 4 / (ms * 1e-3) / 1e9, 1)}
        del a, b

    # ---- CPU baseline: the oracle's FP32 restatement, one host thread, bounded sample
    if not args.no_cpu_baseline:
        import oracle
        # (i) one registration on one host thread, as GreedyWarp's CPU path runs
        sec = oracle.greedy_time(shape, n_iter=1, seed=0)
        # (ii) every host core: one registration per thread, as the CPU atlas runs
        #      one thread per image (the GPU box grants this process 16 cores)
        cores = max(1, min(16, os.cpu_count() or 1))
        wall = oracle.greedy_time_threads(shape, cores, n_iter=1, seed=0)
        out["cpu_baseline"] = {
            "value": round(cores / wall, 5),
            "unit": "iter/s",
            "cores": cores,
            "kind": "port",
            "sample": "%d concurrent FP32 greedy registrations at %d^3, one iteration each, one "
                      "host thread each (C++ restatement of GreedyIteratorCPU::Iterate, own "
                      "mixed-radix FFT in place of FFTW) = %.2f s wall; host %s, nproc %d"
                      % (cores, n, wall, _cpu_model(), os.cpu_count() or 0),
        }
        out["cpu_baseline_1core"] = {"value": round(1.0 / sec, 5), "unit": "iter/s", "cores": 1,
                                     "sample": "1 FP32 greedy iteration at %d^3 on one thread = %.2f s"
                                               % (n, sec)}
        out["gpu_vs_cpu_1core"] = round(iters_per_s * sec, 1)
        out["gpu_vs_cpu_all_cores"] = round(iters_per_s * wall / cores, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
