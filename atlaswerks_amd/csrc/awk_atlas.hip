// awk_atlas.hip -- greedy atlas building on one MI355X.
//
// GreedyAtlasBuilder + AtlasThread + GreedyAtlasThreadGPU (Algorithms/Common/
// AtlasThread.txx:95-181, Algorithms/Greedy/GreedyAtlasBuilder.txx:236-301,
// GreedyAtlasThreadGPU.cxx:161-206).  Every subject's hfield lives on the template
// (mean) grid and holds coordinates in that subject's own image grid, so subject
// images may have different sizes.  One atlas iteration:
//   A  Idef_s = I0_s(h_s) for every subject                    (gather on s's grid)
//   M  mean = sum_s w_s Idef_s in subject order                 (A and M: one kernel)
//   B  per subject the greedy step against I1 = mean: force + FFT passes + Navier
//      solve + composeHV in the inverse x pass, energy E_s on the device
//   E  energy = sum_s w_s E_s (FP64, subject order) into a device history
// The mean that the reference forms after iteration k from the updated hfields
// (ComputeThreadMean) is the same sum of the Idef_s that iteration k+1 needs, so
// stage A serves both.  Steady-state iterations (step sizes known) are captured
// once into a hipGraph per hfield parity and replayed: one graph launch per atlas
// iteration for all subjects.  In those iterations stage B runs batched: each of its
// five passes is one launch over all subjects (blockIdx.y = subject), every
// subject's hfields, deformed image, step and energy state lying at a fixed stride
// in per-atlas slabs, each with its own spectrum.  The kernels and their
// block-to-data mapping are those of the per-subject launches, so the results are
// bit-identical; iterations that set step sizes or track inverse fields take the
// per-subject path on the stream lanes.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "awk_fluid.h"
#include "awk_ops.h"

namespace awk {

__global__ void __launch_bounds__(256)
k_atlas_mean(float *__restrict__ mean, const float *const *__restrict__ idef,
             const float *__restrict__ w, int ns, int N) {
  const int i = (int)(xcd_block() * 256 + threadIdx.x);
  if (i >= N) return;
  // GreedyAtlasThreadGPU::ComputeThreadMean: mean = 0; mean += Idef_s * w_s, s in order
  // (one explicit fma per subject, as in k_atlas_apply_mean)
  float m = 0.f;
  for (int s = 0; s < ns; ++s) m = __builtin_fmaf(idef[s][i], w[s], m);
  mean[i] = m;
}

// One subject's image as the fused mean stage samples it.
struct AtlasSrc {
  const float *i0;  // image on the subject's own grid
  int nx, ny, nz;
  float w;          // the subject's weight in the mean
};

// Stages A and M in one pass over the template grid: for every subject s in
// order, Idef_s = I0_s(h_s) (trilinear, background 0; the gather of k_apply) and
// m = fma(Idef_s, w_s, m), then mean = m.  Same samples and the same summation order
// as one apply launch per subject followed by k_atlas_mean, so the results are bit
// for bit those; the mean pass's re-read of every Idef_s and ns - 1 launches go.
// h_s and Idef_s lie at fixed strides (3N and N floats) in the atlas slabs.  SU
// subjects per group: their field loads are all issued before the gathers.
template <int SU>
__global__ void __launch_bounds__(256)
k_atlas_apply_mean(float *__restrict__ mean, float *__restrict__ idef, const float *__restrict__ h,
                   const AtlasSrc *__restrict__ tab, int ns, int N) {
  const int i = (int)(xcd_block() * 256 + threadIdx.x);
  if (i >= N) return;
  const size_t n = (size_t)N;
  float m = 0.f;
  int s = 0;
  for (; s + SU <= ns; s += SU) {
    float px[SU], py[SU], pz[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      // each field is read once: streamed past the memory-side cache
      const float *hs = h + (size_t)(s + u) * 3 * n + i;
      px[u] = __builtin_nontemporal_load(hs);
      py[u] = __builtin_nontemporal_load(hs + n);
      pz[u] = __builtin_nontemporal_load(hs + 2 * n);
    }
    float r[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const AtlasSrc t = tab[s + u];
      const BufLoad ib = make_buf(t.i0, t.nx * t.ny * t.nz);
      r[u] = trilerp1g<AWK_BG_VAL>(ib, Dims{t.nx, t.ny, t.nz}, px[u], py[u], pz[u], 0.f);
    }
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      idef[(size_t)(s + u) * n + i] = r[u];
      m = __builtin_fmaf(r[u], tab[s + u].w, m);
    }
  }
  for (; s < ns; ++s) {
    const float *hs = h + (size_t)s * 3 * n + i;
    const float px = __builtin_nontemporal_load(hs);
    const float py = __builtin_nontemporal_load(hs + n);
    const float pz = __builtin_nontemporal_load(hs + 2 * n);
    const AtlasSrc t = tab[s];
    const BufLoad ib = make_buf(t.i0, t.nx * t.ny * t.nz);
    const float r = trilerp1g<AWK_BG_VAL>(ib, Dims{t.nx, t.ny, t.nz}, px, py, pz, 0.f);
    idef[(size_t)s * n + i] = r;
    m = __builtin_fmaf(r, t.w, m);
  }
  mean[i] = m;
}

__global__ void k_atlas_energy(const double *__restrict__ esub, const float *__restrict__ w, int ns,
                               double *__restrict__ hist, int *__restrict__ cursor, int cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double e = 0.0;
  for (int s = 0; s < ns; ++s) e += esub[s] * (double)w[s];
  const int c = *cursor;
  hist[c % cap] = e;
  *cursor = c + 1;
}

}  // namespace awk

namespace {

struct Subject {
  const float *I0 = nullptr;
  int dims[3] = {0, 0, 0};
  float w = 0.f;
  float *h[2] = {nullptr, nullptr};  // 3 planes each, ping-pong
  float *hinv[2] = {nullptr, nullptr};  // Def0To1, 3 planes each, when tracked
  float *idef = nullptr;
  float *step = nullptr;
  double *prev = nullptr;
  int *have_prev = nullptr;
  bool update_step = true;
  bool set = false;
};

constexpr int kHistCap = 1 << 16;

// Subjects' updates after the mean are independent: they run on up to kLanes
// streams (lane 0 is the iteration's own stream), each lane with its own operator
// spectrum and reduction scratch; the operator tables are shared.
constexpr int kLanes = 3;
struct Lane {
  awk::FluidKernel k;  // copy of the atlas operator with its own spectrum
  hipStream_t st = nullptr;
  hipEvent_t join = nullptr;
  double *partials = nullptr;
  float *maxpart = nullptr;
  float *vbuf = nullptr;
};

}  // namespace

struct awk_atlas {
  awk::FluidKernel k;
  int nx = 0, ny = 0, nz = 0, ns = 0;
  float sp[3] = {1, 1, 1};
  float max_pert = 0.1f;
  std::vector<Subject> subj;
  int cur = 0;  // which hfield buffer holds the current hfields
  float *mean = nullptr;
  const float **idef_tab = nullptr;  // device table of Idef pointers
  float *w_dev = nullptr;
  awk::AtlasSrc *src_tab = nullptr;  // device table of the subjects' images and weights
  double *esub = nullptr;
  double *partials = nullptr;
  float *maxpart = nullptr;
  double *hist = nullptr;
  int *cursor = nullptr;
  unsigned *flags = nullptr;  // sticky AWK_FLAG_* word of all subjects
  long long host_cursor = 0;
  float *vbuf = nullptr;  // 3 planes: velocity of a step-size iteration
  hipStream_t as = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_fork = nullptr;
  std::vector<Lane> lanes;  // lanes 1..; lane 0 is the atlas operator itself
  bool use_graph = true;
  bool tables_dirty = true;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  // slabs holding every subject's buffers at fixed strides (batched stage B)
  float *hslab[2] = {nullptr, nullptr};  // ns x 3 planes
  float *idef_slab = nullptr;            // ns x N
  float *step_slab = nullptr;            // ns
  double *prev_slab = nullptr;           // ns
  int *have_prev_slab = nullptr;         // ns
  int batch = 0;                         // subjects per batched launch (0: per-subject path)
  int batch_cap = 0;                     // subjects the batch spectra hold
  float2 *bspec = nullptr;               // batch_cap spectra
  double *bpart = nullptr;               // batch_cap x x-pass partials
};

using namespace awk;

static void atlas_drop_graphs(awk_atlas *a) {
  for (int i = 0; i < 2; ++i)
    if (a->exec[i]) {
      (void)hipGraphExecDestroy(a->exec[i]);
      a->exec[i] = nullptr;
    }
}

static void atlas_free(awk_atlas *a) {
  atlas_drop_graphs(a);
  fk_release(a->k);
  for (Subject &s : a->subj) {
    void *b[] = {s.hinv[0], s.hinv[1]};
    for (void *p : b)
      if (p) (void)hipFree(p);
  }
  void *b[] = {a->mean, (void *)a->idef_tab, a->w_dev, a->src_tab, a->esub, a->partials, a->maxpart,
               a->hist, a->cursor, a->vbuf, a->flags, a->hslab[0], a->hslab[1], a->idef_slab,
               a->step_slab, a->prev_slab, a->have_prev_slab, a->bspec, a->bpart};
  for (void *p : b)
    if (p) (void)hipFree(p);
  for (Lane &L : a->lanes) {
    void *lb[] = {L.k.spec, L.partials, L.maxpart, L.vbuf};
    for (void *p : lb)
      if (p) (void)hipFree(p);
    if (L.join) (void)hipEventDestroy(L.join);
    if (L.st) (void)hipStreamDestroy(L.st);
  }
  if (a->ev_fork) (void)hipEventDestroy(a->ev_fork);
  if (a->ev_in) (void)hipEventDestroy(a->ev_in);
  if (a->ev_out) (void)hipEventDestroy(a->ev_out);
  if (a->as) (void)hipStreamDestroy(a->as);
}

// Stage A + M: Idef_s = I0_s(h_s), mean = sum_s w_s Idef_s.  One fused launch
// (knob atlas_mean = subjects per group, 1, 2 or 4), or with atlas_mean = 0 one
// apply launch per subject and then the mean; the results are the same bits.
static awk_status atlas_mean_stage(awk_atlas *a, hipStream_t s) {
  awk_status st;
  const int N = a->nx * a->ny * a->nz;
  const int su = tuning().atlas_mean;
  if (su) {
    const dim3 grid(grid_for(N, 256)), block(256);
    const float *h = a->hslab[a->cur];
    if (su == 4)
      AWK_LAUNCH(k_atlas_apply_mean<4>, grid, block, 0, s, a->mean, a->idef_slab, h, a->src_tab, a->ns, N);
    else if (su == 2)
      AWK_LAUNCH(k_atlas_apply_mean<2>, grid, block, 0, s, a->mean, a->idef_slab, h, a->src_tab, a->ns, N);
    else
      AWK_LAUNCH(k_atlas_apply_mean<1>, grid, block, 0, s, a->mean, a->idef_slab, h, a->src_tab, a->ns, N);
    AWK_LAUNCH_CHECK("k_atlas_apply_mean");
    return AWK_OK;
  }
  for (Subject &sb : a->subj) {
    const float *h = sb.h[a->cur];
    if ((st = launch_apply_grid(sb.idef, sb.I0, sb.dims[0], sb.dims[1], sb.dims[2], h, h + N,
                                h + 2 * (size_t)N, a->nx, a->ny, a->nz, AWK_BG_VAL, 0.f, s)))
      return st;
  }
  AWK_LAUNCH(k_atlas_mean, dim3(grid_for(N, 256)), dim3(256), 0, s, a->mean, a->idef_tab, a->w_dev,
             a->ns, N);
  AWK_LAUNCH_CHECK("k_atlas_mean");
  return AWK_OK;
}

// Stage B batched: subjects [s0, s0 + n) of the slabs in one launch per pass.
static awk_status atlas_stage_b_batched(awk_atlas *a, int s0, int n, hipStream_t s) {
  awk_status st;
  const long long N = (long long)a->nx * a->ny * a->nz;
  const int nxt = 1 - a->cur;
  FluidKernel k = a->k;  // the atlas operator's tables, the batch's spectra
  k.spec = a->bspec;
  Batch b;
  b.n = n;
  b.spec = 3 * (long long)a->k.spec_plane();
  b.img = N;
  b.field = 3 * N;
  b.part = fk_xblocks(a->k);
  b.scalar = 1;
  const Subject &sb = a->subj[s0];
  XFwdArgs fa;
  std::memset(&fa, 0, sizeof(fa));
  fa.mode = XLOAD_FORCE;
  fa.idef = sb.idef;
  fa.i1 = a->mean;
  xfwd_spacing(fa, a->sp);
  fa.partials = a->bpart;
  int nb = 0;
  if ((st = fk_xfwd(k, fa, s, &nb, &b))) return st;
  ReduceArgs ra;
  ra.partials = a->bpart;
  ra.n = (a->ny * a->nz + k.rt - 1) / k.rt;
  ra.scale = (double)a->sp[0] * a->sp[1] * a->sp[2];
  ra.energy_out = a->esub + s0;
  ra.prev = sb.prev;
  ra.have_prev = sb.have_prev;
  ra.step = sb.step;
  ra.auto_reduce = 0;
  ra.flags = a->flags;
  if ((st = fk_ypass(k, false, s, &ra, &b))) return st;
  if ((st = fk_zsolve(k, true, s, &b))) return st;
  if ((st = fk_ypass(k, true, s, nullptr, &b))) return st;
  XInvArgs ia;
  std::memset(&ia, 0, sizeof(ia));
  for (int c = 0; c < 3; ++c) ia.sp[c] = a->sp[c];
  ia.mode = XSTORE_COMPOSE;
  for (int c = 0; c < 3; ++c) {
    ia.h[c] = sb.h[a->cur] + c * N;
    ia.hnew[c] = sb.h[nxt] + c * N;
  }
  ia.step = sb.step;
  return fk_xinv(k, ia, s, &nb, &b);
}

// One atlas iteration on stream s (stages A, M, B, E).  Leaves the new hfields in
// buffer 1 - cur; the caller flips cur.
static awk_status atlas_iteration(awk_atlas *a, hipStream_t s) {
  awk_status st;
  const int N = a->nx * a->ny * a->nz;
  const int nxt = 1 - a->cur;
  if ((st = atlas_mean_stage(a, s))) return st;
  bool batched = a->batch > 0;
  for (const Subject &sb : a->subj) batched = batched && !sb.update_step && !sb.hinv[0];
  if (batched) {
    for (int s0 = 0; s0 < a->ns; s0 += a->batch)
      if ((st = atlas_stage_b_batched(a, s0, std::min(a->batch, a->ns - s0), s))) return st;
    AWK_LAUNCH(k_atlas_energy, dim3(1), dim3(64), 0, s, a->esub, a->w_dev, a->ns, a->hist,
               a->cursor, kHistCap);
    AWK_LAUNCH_CHECK("k_atlas_energy");
    return AWK_OK;
  }
  const int nl = 1 + (int)a->lanes.size();
  if (nl > 1) {
    AWK_HIP(hipEventRecord(a->ev_fork, s));
    for (Lane &L : a->lanes) AWK_HIP(hipStreamWaitEvent(L.st, a->ev_fork, 0));
  }
  for (int si = 0; si < a->ns; ++si) {
    Subject &sb = a->subj[si];
    const float *h = sb.h[a->cur];
    float *hn = sb.h[nxt];
    // this subject's lane: stream, operator spectrum, reduction scratch
    const int li = si % nl;
    FluidKernel &k = li ? a->lanes[li - 1].k : a->k;
    const hipStream_t ls = li ? a->lanes[li - 1].st : s;
    double *partials = li ? a->lanes[li - 1].partials : a->partials;
    float *maxpart = li ? a->lanes[li - 1].maxpart : a->maxpart;
    float *vbuf = li ? a->lanes[li - 1].vbuf : a->vbuf;
    XFwdArgs fa;
    std::memset(&fa, 0, sizeof(fa));
    fa.mode = XLOAD_FORCE;
    fa.idef = sb.idef;
    fa.i1 = a->mean;
    xfwd_spacing(fa, a->sp);
    fa.partials = partials;
    int nb = 0;
    if ((st = fk_xfwd(k, fa, ls, &nb))) return st;
    ReduceArgs ra;
    ra.partials = partials;
    ra.n = (a->ny * a->nz + k.rt - 1) / k.rt;
    ra.scale = (double)a->sp[0] * a->sp[1] * a->sp[2];
    ra.energy_out = a->esub + si;
    ra.prev = sb.prev;
    ra.have_prev = sb.have_prev;
    ra.step = sb.step;
    ra.auto_reduce = 0;
    ra.flags = a->flags;
    if ((st = fk_ypass(k, false, ls, &ra))) return st;
    if ((st = fk_zsolve(k, true, ls))) return st;
    if ((st = fk_ypass(k, true, ls, nullptr))) return st;
    XInvArgs ia;
    std::memset(&ia, 0, sizeof(ia));
    for (int c = 0; c < 3; ++c) ia.sp[c] = a->sp[c];
    const bool inv = sb.hinv[0] != nullptr;
    if (sb.update_step || inv) {
      // v into vbuf (iteration 0 of a scale level: step = MaxPert / max|v|), then
      // composeHV and, when the inverse is tracked, composeVHInv
      // (GreedyIteratorCPU::updateDeformation, GreedyIteratorCPU.cxx:82-95)
      ia.mode = sb.update_step ? XSTORE_MAXNORM : XSTORE_PLAIN;
      for (int c = 0; c < 3; ++c) ia.out[c] = vbuf + (size_t)c * N;
      ia.maxpart = maxpart;
      if ((st = fk_xinv(k, ia, ls, &nb))) return st;
      if (sb.update_step) {
        if ((st = launch_step_from_max(maxpart, nb, a->max_pert, sb.step, a->flags, ls))) return st;
      }
      const float *vx = vbuf, *vy = vbuf + N, *vz = vbuf + 2 * (size_t)N;
      if ((st = launch_compose_hv(hn, hn + N, hn + 2 * (size_t)N, h, h + N, h + 2 * (size_t)N, vx,
                                  vy, vz, a->nx, a->ny, a->nz, a->sp, 1.f, sb.step, 0,
                                  AWK_BG_PARTIAL_ID, ls)))
        return st;
      if (inv) {
        const float *hi = sb.hinv[a->cur];
        float *hin = sb.hinv[nxt];
        if ((st = launch_compose_vh(hin, hin + N, hin + 2 * (size_t)N, vx, vy, vz, hi, hi + N,
                                    hi + 2 * (size_t)N, a->nx, a->ny, a->nz, a->sp, sb.step, 1,
                                    AWK_BG_PARTIAL_ZERO, ls)))
          return st;
      }
    } else {
      ia.mode = XSTORE_COMPOSE;
      for (int c = 0; c < 3; ++c) {
        ia.h[c] = h + (size_t)c * N;
        ia.hnew[c] = hn + (size_t)c * N;
      }
      ia.step = sb.step;
      if ((st = fk_xinv(k, ia, ls, &nb))) return st;
    }
  }
  if (nl > 1)
    for (Lane &L : a->lanes) {
      AWK_HIP(hipEventRecord(L.join, L.st));
      AWK_HIP(hipStreamWaitEvent(s, L.join, 0));
    }
  AWK_LAUNCH(k_atlas_energy, dim3(1), dim3(64), 0, s, a->esub, a->w_dev, a->ns, a->hist, a->cursor,
             kHistCap);
  AWK_LAUNCH_CHECK("k_atlas_energy");
  return AWK_OK;
}

static awk_status atlas_sync_tables(awk_atlas *a) {
  if (!a->tables_dirty) return AWK_OK;
  for (int i = 0; i < a->ns; ++i)
    if (!a->subj[i].set) return fail(AWK_ERR_INVALID, "awk_atlas: subject " + std::to_string(i) + " not set");
  std::vector<const float *> tab(a->ns);
  std::vector<float> w(a->ns);
  std::vector<AtlasSrc> src(a->ns);
  for (int i = 0; i < a->ns; ++i) {
    const Subject &sb = a->subj[i];
    tab[i] = sb.idef;
    w[i] = sb.w;
    src[i] = AtlasSrc{sb.I0, sb.dims[0], sb.dims[1], sb.dims[2], sb.w};
  }
  AWK_HIP(hipMemcpyAsync((void *)a->idef_tab, tab.data(), a->ns * sizeof(float *), hipMemcpyHostToDevice, a->as));
  AWK_HIP(hipMemcpyAsync(a->w_dev, w.data(), a->ns * sizeof(float), hipMemcpyHostToDevice, a->as));
  AWK_HIP(hipMemcpyAsync(a->src_tab, src.data(), a->ns * sizeof(AtlasSrc), hipMemcpyHostToDevice, a->as));
  AWK_HIP(hipStreamSynchronize(a->as));
  a->tables_dirty = false;
  return AWK_OK;
}

extern "C" {

awk_status awk_atlas_create(awk_atlas_t *out, int nx, int ny, int nz, float spx, float spy, float spz,
                            float alpha, float beta, float gamma, float max_pert, int n_subjects) {
  if (!out) return fail(AWK_ERR_INVALID, "awk_atlas_create: null handle pointer");
  *out = nullptr;
  if (!(spx > 0 && spy > 0 && spz > 0)) return fail(AWK_ERR_INVALID, "awk_atlas_create: spacing must be > 0");
  if (!(max_pert > 0)) return fail(AWK_ERR_INVALID, "awk_atlas_create: MaxPert must be > 0");
  if (n_subjects < 1 || n_subjects > 65536) return fail(AWK_ERR_INVALID, "awk_atlas_create: bad subject count");
  awk_atlas *a = new (std::nothrow) awk_atlas();
  if (!a) return fail(AWK_ERR_ALLOC, "awk_atlas_create: host allocation failed");
  const float sp[3] = {spx, spy, spz};
  awk_status st = fk_init(a->k, nx, ny, nz, sp);
  if (st) {
    atlas_free(a);
    delete a;
    return st;
  }
  a->k.alpha = alpha;
  a->k.beta = beta;
  a->k.gamma = gamma;
  a->k.lpow = 1.f;
  a->k.flags = 0;
  a->nx = nx;
  a->ny = ny;
  a->nz = nz;
  a->ns = n_subjects;
  for (int i = 0; i < 3; ++i) a->sp[i] = sp[i];
  a->max_pert = max_pert;
  a->subj.resize(n_subjects);
  const size_t N = (size_t)nx * ny * nz;
  const int nblk = fk_xblocks(a->k);
  hipError_t e = hipSuccess;
  const size_t ns = (size_t)n_subjects;
  if (e == hipSuccess) e = hipMalloc(&a->hslab[0], ns * 3 * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&a->hslab[1], ns * 3 * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&a->idef_slab, ns * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&a->step_slab, ns * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&a->prev_slab, ns * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->have_prev_slab, ns * sizeof(int));
  if (e == hipSuccess) e = hipMemset(a->step_slab, 0, ns * sizeof(float));
  if (e == hipSuccess) e = hipMemset(a->have_prev_slab, 0, ns * sizeof(int));
  if (e == hipSuccess)
    for (size_t i = 0; i < ns; ++i) {
      Subject &s = a->subj[i];
      s.h[0] = a->hslab[0] + i * 3 * N;
      s.h[1] = a->hslab[1] + i * 3 * N;
      s.idef = a->idef_slab + i * N;
      s.step = a->step_slab + i;
      s.prev = a->prev_slab + i;
      s.have_prev = a->have_prev_slab + i;
    }
  // batch spectra for up to every subject (at most 8 GiB of them); by default a
  // launch takes as many subjects as keep their spectra within 128 MiB, half the
  // memory-side cache, so each pass still reads what the previous one wrote from
  // that cache (16 subjects at 128^3 in one launch: 1.81 ms per iteration against
  // 1.70 ms one launch per subject on 3 lanes, measured)
  {
    const size_t sb = 3 * a->k.spec_plane() * sizeof(float2);
    size_t cap = std::min(ns, std::max<size_t>(1, (8ull << 30) / sb));
    if (e == hipSuccess) e = hipMalloc(&a->bspec, cap * sb);
    if (e == hipSuccess) e = hipMalloc(&a->bpart, cap * (size_t)nblk * sizeof(double));
    a->batch_cap = (int)cap;
    a->batch = (int)std::min(cap, std::max<size_t>(1, (128ull << 20) / sb));
  }
  if (e == hipSuccess) e = hipMalloc(&a->mean, N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&a->idef_tab, n_subjects * sizeof(float *));
  if (e == hipSuccess) e = hipMalloc(&a->w_dev, n_subjects * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&a->src_tab, n_subjects * sizeof(AtlasSrc));
  if (e == hipSuccess) e = hipMalloc(&a->esub, n_subjects * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->partials, (size_t)nblk * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->maxpart, (size_t)nblk * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&a->hist, kHistCap * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->cursor, sizeof(int));
  if (e == hipSuccess) e = hipMemset(a->cursor, 0, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&a->flags, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(a->flags, 0, sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc(&a->vbuf, 3 * N * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&a->as, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a->ev_out, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a->ev_fork, hipEventDisableTiming);
  const int extra = std::min(n_subjects, kLanes) - 1;
  for (int l = 0; l < extra && e == hipSuccess; ++l) {
    a->lanes.emplace_back();
    Lane &L = a->lanes.back();
    L.k = a->k;  // shares the operator tables
    L.k.spec = nullptr;
    L.k.eig = nullptr;
    e = hipMalloc(&L.k.spec, 3 * a->k.spec_plane() * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&L.partials, (size_t)nblk * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&L.maxpart, (size_t)nblk * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&L.vbuf, 3 * N * sizeof(float));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.join, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    st = hip_check(e, "awk_atlas_create: allocation");
    atlas_free(a);
    delete a;
    return st;
  }
  *out = a;
  return AWK_OK;
}

awk_status awk_atlas_destroy(awk_atlas_t a) {
  if (!a) return AWK_OK;
  (void)hipStreamSynchronize(a->as);
  atlas_free(a);
  delete a;
  return AWK_OK;
}

awk_status awk_atlas_set_subject(awk_atlas_t a, int s, const float *I0, int nx_img, int ny_img,
                                 int nz_img, float weight, const float *const h0[3],
                                 awk_stream_t stream) {
  if (!a || !I0 || !h0 || !h0[0] || !h0[1] || !h0[2]) return fail(AWK_ERR_INVALID, "awk_atlas_set_subject: null argument");
  if (s < 0 || s >= a->ns) return fail(AWK_ERR_INVALID, "awk_atlas_set_subject: subject index out of range");
  if (nx_img < 1 || ny_img < 1 || nz_img < 1 || (long long)nx_img * ny_img * nz_img > kBufMaxElems)
    return fail(AWK_ERR_INVALID, "awk_atlas_set_subject: bad image size");
  Subject &sb = a->subj[s];
  for (int b = 0; b < 2; ++b)  // a new subject starts without a tracked inverse
    if (sb.hinv[b]) {
      AWK_HIP(hipStreamSynchronize(a->as));
      AWK_HIP(hipFree(sb.hinv[b]));
      sb.hinv[b] = nullptr;
    }
  sb.I0 = I0;
  sb.dims[0] = nx_img;
  sb.dims[1] = ny_img;
  sb.dims[2] = nz_img;
  sb.w = weight;
  sb.update_step = true;
  const size_t N = (size_t)a->nx * a->ny * a->nz;
  hipStream_t cs = as_stream(stream);
  for (int c = 0; c < 3; ++c)
    AWK_HIP(hipMemcpyAsync(sb.h[a->cur] + c * N, h0[c], N * sizeof(float), hipMemcpyDeviceToDevice, cs));
  sb.set = true;
  a->tables_dirty = true;
  atlas_drop_graphs(a);  // the image pointer is baked into captured graphs
  return AWK_OK;
}

awk_status awk_atlas_get_hfield(awk_atlas_t a, int s, float *const h_out[3], awk_stream_t stream) {
  if (!a || !h_out || !h_out[0] || !h_out[1] || !h_out[2]) return fail(AWK_ERR_INVALID, "awk_atlas_get_hfield: null argument");
  if (s < 0 || s >= a->ns) return fail(AWK_ERR_INVALID, "awk_atlas_get_hfield: subject index out of range");
  const size_t N = (size_t)a->nx * a->ny * a->nz;
  hipStream_t cs = as_stream(stream);
  AWK_HIP(hipEventRecord(a->ev_out, a->as));
  AWK_HIP(hipStreamWaitEvent(cs, a->ev_out, 0));
  for (int c = 0; c < 3; ++c)
    AWK_HIP(hipMemcpyAsync(h_out[c], a->subj[s].h[a->cur] + c * N, N * sizeof(float), hipMemcpyDeviceToDevice, cs));
  return AWK_OK;
}

awk_status awk_atlas_set_subject_inverse(awk_atlas_t a, int s, const float *const hinv0[3],
                                         awk_stream_t stream) {
  if (!a || !hinv0 || !hinv0[0] || !hinv0[1] || !hinv0[2])
    return fail(AWK_ERR_INVALID, "awk_atlas_set_subject_inverse: null argument");
  if (s < 0 || s >= a->ns) return fail(AWK_ERR_INVALID, "awk_atlas_set_subject_inverse: subject index out of range");
  Subject &sb = a->subj[s];
  if (!sb.set) return fail(AWK_ERR_INVALID, "awk_atlas_set_subject_inverse: subject not set");
  if (sb.dims[0] != a->nx || sb.dims[1] != a->ny || sb.dims[2] != a->nz)
    return fail(AWK_ERR_INVALID, "awk_atlas_set_subject_inverse: the subject's image must be on the template grid");
  const size_t N = (size_t)a->nx * a->ny * a->nz;
  for (int b = 0; b < 2; ++b)
    if (!sb.hinv[b]) AWK_HIP(hipMalloc(&sb.hinv[b], 3 * N * sizeof(float)));
  hipStream_t cs = as_stream(stream);
  for (int c = 0; c < 3; ++c)
    AWK_HIP(hipMemcpyAsync(sb.hinv[a->cur] + c * N, hinv0[c], N * sizeof(float), hipMemcpyDeviceToDevice, cs));
  atlas_drop_graphs(a);  // the update path of s changes
  return AWK_OK;
}

awk_status awk_atlas_get_hinv(awk_atlas_t a, int s, float *const out[3], awk_stream_t stream) {
  if (!a || !out || !out[0] || !out[1] || !out[2]) return fail(AWK_ERR_INVALID, "awk_atlas_get_hinv: null argument");
  if (s < 0 || s >= a->ns) return fail(AWK_ERR_INVALID, "awk_atlas_get_hinv: subject index out of range");
  if (!a->subj[s].hinv[0]) return fail(AWK_ERR_INVALID, "awk_atlas_get_hinv: the subject's inverse is not tracked");
  const size_t N = (size_t)a->nx * a->ny * a->nz;
  hipStream_t cs = as_stream(stream);
  AWK_HIP(hipEventRecord(a->ev_out, a->as));
  AWK_HIP(hipStreamWaitEvent(cs, a->ev_out, 0));
  for (int c = 0; c < 3; ++c)
    AWK_HIP(hipMemcpyAsync(out[c], a->subj[s].hinv[a->cur] + c * N, N * sizeof(float), hipMemcpyDeviceToDevice, cs));
  return AWK_OK;
}

awk_status awk_atlas_update_step_next(awk_atlas_t a) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_atlas_update_step_next: null handle");
  for (Subject &s : a->subj) s.update_step = true;
  return AWK_OK;
}

awk_status awk_atlas_get_flags(awk_atlas_t a, unsigned *flags, int clear, awk_stream_t stream) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_atlas_get_flags: null handle");
  hipStream_t cs = as_stream(stream);
  AWK_HIP(hipEventRecord(a->ev_out, a->as));
  AWK_HIP(hipStreamWaitEvent(cs, a->ev_out, 0));
  return read_flags(&a->flags, 1, flags, clear, cs, "awk_atlas_get_flags");
}

awk_status awk_atlas_set_graph(awk_atlas_t a, int on) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_atlas_set_graph: null handle");
  a->use_graph = on != 0;
  if (!a->use_graph) atlas_drop_graphs(a);
  return AWK_OK;
}

awk_status awk_atlas_set_batch(awk_atlas_t a, int n) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_atlas_set_batch: null handle");
  if (n < 0) return fail(AWK_ERR_INVALID, "awk_atlas_set_batch: negative batch");
  const int b = std::min(n, a->batch_cap);
  if (b != a->batch) atlas_drop_graphs(a);
  a->batch = b;
  return AWK_OK;
}

awk_status awk_atlas_iterate(awk_atlas_t a, int n_iter, double *energy_dev, awk_stream_t stream) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_atlas_iterate: null handle");
  if (n_iter < 0 || n_iter > kHistCap) return fail(AWK_ERR_INVALID, "awk_atlas_iterate: bad iteration count");
  awk_status st;
  if ((st = atlas_sync_tables(a))) return st;
  hipStream_t cs = as_stream(stream);
  AWK_HIP(hipEventRecord(a->ev_in, cs));
  AWK_HIP(hipStreamWaitEvent(a->as, a->ev_in, 0));
  const long long first = a->host_cursor;
  for (int it = 0; it < n_iter; ++it) {
    bool any_step = false;
    for (const Subject &s : a->subj) any_step = any_step || s.update_step;
    if (any_step || !a->use_graph) {
      if ((st = atlas_iteration(a, a->as))) return st;
      for (Subject &s : a->subj) s.update_step = false;
    } else {
      hipGraphExec_t &ex = a->exec[a->cur];
      if (!ex) {
        hipGraph_t g = nullptr;
        AWK_HIP(hipStreamBeginCapture(a->as, hipStreamCaptureModeThreadLocal));
        st = atlas_iteration(a, a->as);
        hipError_t ce = hipStreamEndCapture(a->as, &g);
        if (st) {
          if (g) (void)hipGraphDestroy(g);
          return st;
        }
        if (ce != hipSuccess) return hip_check(ce, "awk_atlas_iterate: graph capture");
        hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
          ex = nullptr;
          return hip_check(ie, "awk_atlas_iterate: graph instantiate");
        }
      }
      AWK_HIP(hipGraphLaunch(ex, a->as));
    }
    a->cur = 1 - a->cur;
    ++a->host_cursor;
  }
  if (energy_dev && n_iter > 0) {
    // history slots [first, first + n_iter) of the ring
    const int c0 = (int)(first % kHistCap);
    const int n1 = std::min(n_iter, kHistCap - c0);
    AWK_HIP(hipMemcpyAsync(energy_dev, a->hist + c0, n1 * sizeof(double), hipMemcpyDeviceToDevice, a->as));
    if (n1 < n_iter)
      AWK_HIP(hipMemcpyAsync(energy_dev + n1, a->hist, (n_iter - n1) * sizeof(double),
                             hipMemcpyDeviceToDevice, a->as));
  }
  AWK_HIP(hipEventRecord(a->ev_out, a->as));
  AWK_HIP(hipStreamWaitEvent(cs, a->ev_out, 0));
  return AWK_OK;
}

awk_status awk_atlas_mean(awk_atlas_t a, float *mean_out, awk_stream_t stream) {
  if (!a || !mean_out) return fail(AWK_ERR_INVALID, "awk_atlas_mean: null argument");
  awk_status st;
  if ((st = atlas_sync_tables(a))) return st;
  hipStream_t cs = as_stream(stream);
  AWK_HIP(hipEventRecord(a->ev_in, cs));
  AWK_HIP(hipStreamWaitEvent(a->as, a->ev_in, 0));
  if ((st = atlas_mean_stage(a, a->as))) return st;
  const size_t N = (size_t)a->nx * a->ny * a->nz;
  AWK_HIP(hipMemcpyAsync(mean_out, a->mean, N * sizeof(float), hipMemcpyDeviceToDevice, a->as));
  AWK_HIP(hipEventRecord(a->ev_out, a->as));
  AWK_HIP(hipStreamWaitEvent(cs, a->ev_out, 0));
  return AWK_OK;
}

awk_status awk_atlas_deformed_image(awk_atlas_t a, int s, float *out, awk_stream_t stream) {
  if (!a || !out) return fail(AWK_ERR_INVALID, "awk_atlas_deformed_image: null argument");
  if (s < 0 || s >= a->ns || !a->subj[s].set) return fail(AWK_ERR_INVALID, "awk_atlas_deformed_image: bad subject");
  const Subject &sb = a->subj[s];
  const size_t N = (size_t)a->nx * a->ny * a->nz;
  hipStream_t cs = as_stream(stream);
  AWK_HIP(hipEventRecord(a->ev_out, a->as));
  AWK_HIP(hipStreamWaitEvent(cs, a->ev_out, 0));
  const float *h = sb.h[a->cur];
  return launch_apply_grid(out, sb.I0, sb.dims[0], sb.dims[1], sb.dims[2], h, h + N, h + 2 * N,
                           a->nx, a->ny, a->nz, AWK_BG_VAL, 0.f, cs);
}

}  // extern "C"
