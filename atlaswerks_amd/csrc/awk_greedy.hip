// awk_greedy.hip -- one greedy-fluid iteration on gfx950.
//
// GreedyIteratorGPU::Iterate (Algorithms/Greedy/GreedyIteratorGPU.cxx:92-135) runs
// about fifteen kernels and two host syncs per iteration.  Here one iteration is
// five kernels and no host sync:
//   1. apply          Idef = I0(h)                                  (gather)
//   2. X  (force)     f = (I1 - Idef) grad(Idef), r2c along x, FP64 SSD partials
//   3. Y, Z+solve, Y' the Navier solve in the spectrum; block 0 of Y also sums
//                     the SSD partials in index order (deterministic) and applies
//                     AutoStepReduce
//   4. X' (compose)   c2r along x and h' = h(x + step v / sp) in the epilogue
// On the first iteration of a scale level (step not yet known) step 5 writes v
// and a per-block max |v|; the step MaxPert / max|v| is then formed on the device
// and a separate composeHV applies it.  The same unfused tail serves the inverse
// hfield update.  Iterations run back to back (awk_greedy_iterate_n) fold step 1
// of the next iteration into step 4: the composition samples I0 at the new field
// it holds in registers.
#include <array>
#include <utility>
#include <cstring>
#include <new>
#include <vector>

#include "awk_fluid.h"
#include "awk_ops.h"

namespace awk {

// step = MaxPert / max |v| (GreedyIteratorCPU.cxx:65-79, 129-134)
__global__ void __launch_bounds__(256)
k_step_from_max(const float *__restrict__ maxpart, int n, float max_pert, float *__restrict__ step,
                unsigned *__restrict__ flags) {
  __shared__ float s[256];
  float m = 0.f;
  m = strided_fold(maxpart, n, (int)threadIdx.x, 256, m, FoldMax());
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float mx = s[0];
    *step = max_pert / mx;
    if (!(mx > 0.f) || !(mx <= 3.402823466e38f)) raise_flags(flags, AWK_FLAG_BAD_STEP);
  }
}

awk_status launch_step_from_max(const float *maxpart, int n, float max_pert, float *step,
                                unsigned *flags, hipStream_t s) {
  AWK_LAUNCH(k_step_from_max, dim3(1), dim3(256), 0, s, maxpart, n, max_pert, step, flags);
  AWK_LAUNCH_CHECK("k_step_from_max");
  return AWK_OK;
}

}  // namespace awk

struct awk_greedy {
  awk::FluidKernel k;
  int nx = 0, ny = 0, nz = 0;
  float sp[3] = {1, 1, 1};
  float max_pert = 0.1f;
  bool update_step = true;
  int auto_reduce = 0;
  float *idef = nullptr;
  float *v = nullptr;     // 3 planes (first iteration / inverse update)
  float *hbuf = nullptr;  // 3 planes (in-place update staging)
  double *partials = nullptr;
  float *maxpart = nullptr;
  float *step = nullptr;  // device scalar
  double *prev = nullptr;
  int *have_prev = nullptr;
  unsigned *flags = nullptr;  // sticky AWK_FLAG_* word
  int nblk = 0;
  // optional per-kernel timing (awk_greedy_set_profiling): start/stop events
  // attached to each stage's kernel dispatch
  bool prof = false;
  std::vector<std::array<awk::LaunchEvents, AWK_GREEDY_NSTAGES>> ev;
  std::vector<unsigned> ev_mask;  // stages launched in each recorded iteration
  int ev_used = 0;
  void stage(int i) {
    if (prof) {
      awk::launch_events() = &ev[ev_used][i];
      ev_mask[ev_used] |= 1u << i;
    }
  }
};

using namespace awk;

static void greedy_free(awk_greedy *g) {
  for (auto &set : g->ev)
    for (LaunchEvents &e : set) {
      (void)hipEventDestroy(e.start);
      (void)hipEventDestroy(e.stop);
    }
  g->ev.clear();
  g->ev_mask.clear();
  fk_release(g->k);
  if (g->idef) (void)hipFree(g->idef);
  if (g->v) (void)hipFree(g->v);
  if (g->hbuf) (void)hipFree(g->hbuf);
  if (g->partials) (void)hipFree(g->partials);
  if (g->maxpart) (void)hipFree(g->maxpart);
  if (g->step) (void)hipFree(g->step);
  if (g->prev) (void)hipFree(g->prev);
  if (g->have_prev) (void)hipFree(g->have_prev);
  if (g->flags) (void)hipFree(g->flags);
}

extern "C" {

awk_status awk_greedy_create(awk_greedy_t *out, int nx, int ny, int nz, float spx, float spy,
                             float spz, float alpha, float beta, float gamma, float max_pert) {
  if (!out) return fail(AWK_ERR_INVALID, "awk_greedy_create: null handle pointer");
  *out = nullptr;
  if (!(spx > 0 && spy > 0 && spz > 0)) return fail(AWK_ERR_INVALID, "awk_greedy_create: spacing must be > 0");
  if (!(max_pert > 0)) return fail(AWK_ERR_INVALID, "awk_greedy_create: MaxPert must be > 0");
  awk_greedy *g = new (std::nothrow) awk_greedy();
  if (!g) return fail(AWK_ERR_ALLOC, "awk_greedy_create: host allocation failed");
  const float sp[3] = {spx, spy, spz};
  awk_status st = fk_init(g->k, nx, ny, nz, sp);
  if (st) {
    greedy_free(g);
    delete g;
    return st;
  }
  g->k.alpha = alpha;
  g->k.beta = beta;
  g->k.gamma = gamma;
  g->k.lpow = 1.f;
  g->k.flags = 0;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  for (int i = 0; i < 3; ++i) g->sp[i] = sp[i];
  g->max_pert = max_pert;
  const size_t N = (size_t)nx * ny * nz;
  g->nblk = fk_xblocks(g->k);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&g->idef, N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->v, 3 * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->hbuf, 3 * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->partials, (size_t)g->nblk * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&g->maxpart, (size_t)g->nblk * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->step, sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->prev, sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&g->have_prev, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&g->flags, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(g->flags, 0, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(g->step, 0, sizeof(float));
  if (e == hipSuccess) e = hipMemset(g->prev, 0, sizeof(double));
  if (e == hipSuccess) e = hipMemset(g->have_prev, 0, sizeof(int));
  if (e != hipSuccess) {
    st = hip_check(e, "awk_greedy_create: allocation");
    greedy_free(g);
    delete g;
    return st;
  }
  *out = g;
  return AWK_OK;
}

awk_status awk_greedy_destroy(awk_greedy_t g) {
  if (!g) return AWK_OK;
  greedy_free(g);
  delete g;
  return AWK_OK;
}

awk_status awk_greedy_update_step_next(awk_greedy_t g) {
  if (!g) return fail(AWK_ERR_INVALID, "awk_greedy_update_step_next: null handle");
  g->update_step = true;
  return AWK_OK;
}

awk_status awk_greedy_set_step(awk_greedy_t g, float step, awk_stream_t stream) {
  if (!g) return fail(AWK_ERR_INVALID, "awk_greedy_set_step: null handle");
  AWK_HIP(store_scalar_async(g->step, step, as_stream(stream)));
  g->update_step = false;
  return AWK_OK;
}

awk_status awk_greedy_get_step(awk_greedy_t g, float *step, awk_stream_t stream) {
  if (!g || !step) return fail(AWK_ERR_INVALID, "awk_greedy_get_step: null argument");
  AWK_HIP(load_scalar_sync(step, g->step, as_stream(stream)));
  return AWK_OK;
}

awk_status awk_greedy_get_flags(awk_greedy_t g, unsigned *flags, int clear, awk_stream_t stream) {
  if (!g) return fail(AWK_ERR_INVALID, "awk_greedy_get_flags: null handle");
  return read_flags(&g->flags, 1, flags, clear, as_stream(stream), "awk_greedy_get_flags");
}

awk_status awk_greedy_set_auto_step_reduce(awk_greedy_t g, int on) {
  if (!g) return fail(AWK_ERR_INVALID, "awk_greedy_set_auto_step_reduce: null handle");
  g->auto_reduce = on ? 1 : 0;
  return AWK_OK;
}

}  // extern "C"

// One iteration.  have_idef: g->idef already holds I0(h_in), written by the
// previous iteration's composition.  fuse_apply: a steady-state composition (step
// known, no inverse field) also writes I0(h_out) into g->idef, and *fused says
// that it did; h_out must not overlap h_in then.
static awk_status iterate_one(awk_greedy *g, const float *I0, const float *I1,
                              const float *const h_in[3], float *const h_out[3],
                              const float *const hinv_in[3], float *const hinv_out[3],
                              double *energy_dev, hipStream_t s, bool have_idef, bool fuse_apply,
                              bool *fused) {
  *fused = false;
  const bool inv = hinv_in && hinv_out;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const size_t N = (size_t)nx * ny * nz;
  awk_status st;
  if (g->prof) {
    if (g->ev_used == (int)g->ev.size()) {
      std::array<LaunchEvents, AWK_GREEDY_NSTAGES> set;
      for (auto &e : set) {
        AWK_HIP(hipEventCreate(&e.start));
        AWK_HIP(hipEventCreate(&e.stop));
      }
      g->ev.push_back(set);
      g->ev_mask.push_back(0u);
    }
    g->ev_mask[g->ev_used] = 0u;
  }
  // the launches of this call are timed only while launch_events() is set
  struct Clear {
    ~Clear() { launch_events() = nullptr; }
  } clear_events;
  if (!have_idef) {
    g->stage(0);
    // 1. deformed image (VAL background 0, HField3DUtils::apply defaults)
    if ((st = launch_apply(g->idef, I0, h_in[0], h_in[1], h_in[2], nx, ny, nz, AWK_BG_VAL, 0.f, s))) return st;
  }
  g->stage(1);
  XFwdArgs fa;
  std::memset(&fa, 0, sizeof(fa));
  fa.mode = XLOAD_FORCE;
  fa.idef = g->idef;
  fa.i1 = I1;
  xfwd_spacing(fa, g->sp);
  fa.partials = g->partials;
  int nb = 0;
  if ((st = fk_xfwd(g->k, fa, s, &nb))) return st;
  g->stage(2);
  // 3. solve; block 0 of the forward y pass also reduces the energy
  ReduceArgs ra;
  ra.partials = g->partials;
  ra.n = (ny * nz + g->k.rt - 1) / g->k.rt;
  ra.scale = (double)g->sp[0] * g->sp[1] * g->sp[2];  // ImageUtils::l2NormSqr (ImageUtils.txx:55-58)
  ra.energy_out = energy_dev;
  ra.prev = g->prev;
  ra.have_prev = g->have_prev;
  ra.step = g->step;
  ra.auto_reduce = g->auto_reduce;
  ra.flags = g->flags;
  if ((st = fk_ypass(g->k, false, s, &ra))) return st;
  g->stage(3);
  if ((st = fk_zsolve(g->k, true, s))) return st;
  g->stage(4);
  if ((st = fk_ypass(g->k, true, s, nullptr))) return st;
  // 5. velocity -> hfield update
  const bool in_place = planes_alias(h_out, h_in, N);
  float *dst[3];
  for (int c = 0; c < 3; ++c) dst[c] = in_place ? g->hbuf + c * N : h_out[c];
  XInvArgs ia;
  std::memset(&ia, 0, sizeof(ia));
  for (int c = 0; c < 3; ++c) ia.sp[c] = g->sp[c];
  if (g->update_step || inv) {
    g->stage(5);
    ia.mode = g->update_step ? XSTORE_MAXNORM : XSTORE_PLAIN;
    for (int c = 0; c < 3; ++c) ia.out[c] = g->v + c * N;
    ia.maxpart = g->maxpart;
    if ((st = fk_xinv(g->k, ia, s, &nb))) return st;
    if (g->update_step) {
      if ((st = launch_step_from_max(g->maxpart, nb, g->max_pert, g->step, g->flags, s))) return st;
      g->update_step = false;
    }
    float *vx = g->v, *vy = g->v + N, *vz = g->v + 2 * N;
    if ((st = launch_compose_hv(dst[0], dst[1], dst[2], h_in[0], h_in[1], h_in[2], vx, vy, vz, nx,
                                ny, nz, g->sp, 1.f, g->step, 0, AWK_BG_PARTIAL_ID, s)))
      return st;
    if (in_place)
      for (int c = 0; c < 3; ++c)
        AWK_HIP(hipMemcpyAsync(h_out[c], g->hbuf + c * N, N * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (inv) {
      // composeVHInv(v * step, hinv, PARTIAL_ZERO) (GreedyIteratorCPU.cxx:89-93)
      const bool inv_in_place = planes_alias(hinv_out, hinv_in, N);
      float *idst[3];
      for (int c = 0; c < 3; ++c) idst[c] = inv_in_place ? g->hbuf + c * N : hinv_out[c];
      if ((st = launch_compose_vh(idst[0], idst[1], idst[2], vx, vy, vz, hinv_in[0], hinv_in[1],
                                  hinv_in[2], nx, ny, nz, g->sp, g->step, 1, AWK_BG_PARTIAL_ZERO, s)))
        return st;
      if (inv_in_place)
        for (int c = 0; c < 3; ++c)
          AWK_HIP(hipMemcpyAsync(hinv_out[c], g->hbuf + c * N, N * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    if (g->prof) ++g->ev_used;
    return AWK_OK;
  } else {
    ia.mode = XSTORE_COMPOSE;
    for (int c = 0; c < 3; ++c) {
      ia.h[c] = h_in[c];
      ia.hnew[c] = dst[c];
    }
    ia.step = g->step;
    if (fuse_apply && !in_place) {
      ia.i0 = I0;
      ia.idef = g->idef;
      ia.i0_bg = 0.f;
      *fused = true;
    }
    g->stage(*fused ? 6 : 5);
    if ((st = fk_xinv(g->k, ia, s, &nb))) return st;
  }
  if (in_place)
    for (int c = 0; c < 3; ++c)
      AWK_HIP(hipMemcpyAsync(h_out[c], g->hbuf + c * N, N * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (g->prof) ++g->ev_used;
  return AWK_OK;
}

extern "C" {

awk_status awk_greedy_iterate(awk_greedy_t g, const float *I0, const float *I1,
                              const float *const h_in[3], float *const h_out[3],
                              const float *const hinv_in[3], float *const hinv_out[3],
                              double *energy_dev, awk_stream_t stream) {
  if (!g || !I0 || !I1 || !h_in || !h_out) return fail(AWK_ERR_INVALID, "awk_greedy_iterate: null argument");
  for (int c = 0; c < 3; ++c)
    if (!h_in[c] || !h_out[c]) return fail(AWK_ERR_INVALID, "awk_greedy_iterate: null hfield plane");
  if (hinv_in && hinv_out)
    for (int c = 0; c < 3; ++c)
      if (!hinv_in[c] || !hinv_out[c]) return fail(AWK_ERR_INVALID, "awk_greedy_iterate: null hinv plane");
  bool fused = false;
  return iterate_one(g, I0, I1, h_in, h_out, hinv_in, hinv_out, energy_dev, as_stream(stream), false,
                     false, &fused);
}

awk_status awk_greedy_iterate_n(awk_greedy_t g, const float *I0, const float *I1, float *const h_a[3],
                                float *const h_b[3], int n, double *energies_dev,
                                awk_stream_t stream) {
  if (!g || !I0 || !I1 || !h_a || !h_b) return fail(AWK_ERR_INVALID, "awk_greedy_iterate_n: null argument");
  if (n < 0) return fail(AWK_ERR_INVALID, "awk_greedy_iterate_n: negative iteration count");
  for (int c = 0; c < 3; ++c)
    if (!h_a[c] || !h_b[c]) return fail(AWK_ERR_INVALID, "awk_greedy_iterate_n: null hfield plane");
  const size_t N = (size_t)g->nx * g->ny * g->nz;
  if (planes_alias(h_b, h_a, N) || planes_alias(h_a, h_b, N))
    return fail(AWK_ERR_INVALID, "awk_greedy_iterate_n: the two hfields must not overlap");
  hipStream_t s = as_stream(stream);
  bool have_idef = false;
  for (int i = 0; i < n; ++i) {
    float *const *src = (i & 1) ? h_b : h_a;
    float *const *dst = (i & 1) ? h_a : h_b;
    bool fused = false;
    awk_status st = iterate_one(g, I0, I1, src, dst, nullptr, nullptr,
                                energies_dev ? energies_dev + i : nullptr, s, have_idef, i + 1 < n,
                                &fused);
    if (st) return st;
    have_idef = fused;
  }
  return AWK_OK;
}

awk_status awk_greedy_set_profiling(awk_greedy_t g, int on) {
  if (!g) return fail(AWK_ERR_INVALID, "awk_greedy_set_profiling: null handle");
  g->prof = on != 0;
  g->ev_used = 0;
  return AWK_OK;
}

awk_status awk_greedy_get_profile(awk_greedy_t g, double *stage_ms, int *stage_launches,
                                  int *iterations) {
  if (!g || !stage_ms || !iterations) return fail(AWK_ERR_INVALID, "awk_greedy_get_profile: null argument");
  for (int i = 0; i < AWK_GREEDY_NSTAGES; ++i) {
    stage_ms[i] = 0.0;
    if (stage_launches) stage_launches[i] = 0;
  }
  for (int it = 0; it < g->ev_used; ++it) {
    for (int i = 0; i < AWK_GREEDY_NSTAGES; ++i) {
      if (!(g->ev_mask[it] >> i & 1u)) continue;  // not launched in this iteration
      AWK_HIP(hipEventSynchronize(g->ev[it][i].stop));
      float ms = 0.f;
      AWK_HIP(hipEventElapsedTime(&ms, g->ev[it][i].start, g->ev[it][i].stop));
      stage_ms[i] += ms;
      if (stage_launches) ++stage_launches[i];
    }
  }
  *iterations = g->ev_used;
  g->ev_used = 0;
  return AWK_OK;
}

}  // extern "C"
