// awk_ldmm_atlas.hip -- LDMM atlas estimation on gfx950.
//
// LDMMAtlasBuilder (Algorithms/LDMM/LDMMAtlasBuilder.txx:270-340),
// LDMMAtlasThreadGPU (LDMMAtlasThreadGPU.cxx ComputeThreadMean, BeginImageIteration)
// and the loop of AtlasThread::AtlasThreadMain (Algorithms/Common/AtlasThread.txx):
// every subject s is one LDMM registration (IterateOrig, the reference's default,
// or IterateNew) of the mean (I0) to the subject image I_s (I1), and the mean is
//   sum_s w_s I_s(h_0^s) JacDet_0^s / sum_s w_s JacDet_0^s   (JacobianScale)
//   sum_s w_s I_s(h_0^s)                                      (otherwise)
// with h_0 = GetDefTTo1(0) and JacDet_0 = ComputeJacDet of the current fields.
// Those two are exactly IterateOrig's backward-sweep terms at t = 0, and the
// backward sweep does not depend on the mean, so with IterateOrig one backward
// sweep per subject and iteration serves both the new mean and the next
// iteration's update (the reference runs it twice).  Subjects run in stream order on one shared iterator
// state; each keeps its own JT_t / JacDet_t, step and energy.  No host
// synchronisation; the mean and the energies are summed in subject order.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "awk_ldmm.h"
#include "awk_ops.h"

namespace awk {

constexpr int kLaBlock = 256;

// ComputeThreadMean (Add_MulMulC_I / Add_MulC_I per subject) followed by
// LDMMAtlasBuilder::ComputeMean's division by the JacDet sum; float, subject order.
__global__ void __launch_bounds__(kLaBlock)
k_la_mean(float *__restrict__ mean, const float *const *__restrict__ jt0,
          const float *const *__restrict__ jd0, const float *__restrict__ w, int S, int N,
          int jscale) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  float num = 0.f, den = 0.f;
  for (int s = 0; s < S; ++s) {
    const float j = jt0[s][i];
    if (jscale) {
      const float dd = jd0[s][i];
      num += j * dd * w[s];
      den += dd * w[s];
    } else {
      num += j * w[s];
    }
  }
  mean[i] = jscale ? num / den : num;
}

// out[c] = sum_s w_s e_s[c] (LDMMEnergy operator* and +=, LDMMEnergy.cxx:138-192)
__global__ void k_la_energy(const double *__restrict__ esub, const float *__restrict__ w, int S,
                            double *__restrict__ out) {
  const int c = (int)threadIdx.x;
  if (c >= 3) return;
  double a = 0.0;
  for (int s = 0; s < S; ++s) a += (double)w[s] * esub[3 * s + c];
  out[c] = a;
}

}  // namespace awk

using namespace awk;

namespace {

struct LaSubject {
  float *I = nullptr;          // the subject image (I1), N floats, owned
  std::vector<float *> v;      // 3T caller planes
  float *jt = nullptr, *jd = nullptr;  // [T][N] each, owned
  float *step = nullptr;       // device float, owned
  bool update_step = true;
  bool set = false;
};

// Subjects are independent within a sweep: they run on up to kLaLanes streams,
// lane 0 on the caller's stream with the atlas engine, the others each with an
// engine (operator, hfields, scratch) of their own.
constexpr int kLaLanes = 3;
struct LaLane {
  awk_ldmm eng;
  hipStream_t st = nullptr;
  hipEvent_t join = nullptr;
};

}  // namespace

struct awk_ldmm_atlas {
  awk_ldmm eng;  // lane 0's iterator state (operator, hfields, scratch), no JT/JacDet
  std::vector<LaLane> lanes;  // lanes 1..
  hipEvent_t fork = nullptr;
  int S = 0;
  bool jscale = true;
  std::vector<LaSubject> subs;
  std::vector<float> w_host;
  float *mean = nullptr;
  float *w = nullptr;                // S weights
  const float **jt0 = nullptr, **jd0 = nullptr;  // S device pointers each
  double *esub = nullptr;            // [S][3]
  double *energy = nullptr;          // 3 doubles when the caller passes none
  bool current = false;              // mean and the JT / JacDet terms match the fields
  bool orig = true;                  // IterateOrig (mUseOrigGradientCalc) or IterateNew
};

static void la_free(awk_ldmm_atlas *a) {
  ldmm_release(&a->eng);
  for (LaLane &L : a->lanes) {
    ldmm_release(&L.eng);
    if (L.join) (void)hipEventDestroy(L.join);
    if (L.st) (void)hipStreamDestroy(L.st);
  }
  if (a->fork) (void)hipEventDestroy(a->fork);
  for (LaSubject &s : a->subs) {
    void *b[] = {s.I, s.jt, s.jd, s.step};
    for (void *p : b)
      if (p) (void)hipFree(p);
  }
  void *bufs[] = {a->mean, a->w, a->jt0, a->jd0, a->esub, a->energy};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
}

// lane of subject k: engine and stream
static awk_ldmm *lane_eng(awk_ldmm_atlas *a, int k) {
  const int li = k % (1 + (int)a->lanes.size());
  return li ? &a->lanes[li - 1].eng : &a->eng;
}
static hipStream_t lane_stream(awk_ldmm_atlas *a, int k, hipStream_t s) {
  const int li = k % (1 + (int)a->lanes.size());
  return li ? a->lanes[li - 1].st : s;
}
static awk_status lanes_fork(awk_ldmm_atlas *a, hipStream_t s) {
  if (a->lanes.empty()) return AWK_OK;
  AWK_HIP(hipEventRecord(a->fork, s));
  for (LaLane &L : a->lanes) AWK_HIP(hipStreamWaitEvent(L.st, a->fork, 0));
  return AWK_OK;
}
static awk_status lanes_join(awk_ldmm_atlas *a, hipStream_t s) {
  for (LaLane &L : a->lanes) {
    AWK_HIP(hipEventRecord(L.join, L.st));
    AWK_HIP(hipStreamWaitEvent(s, L.join, 0));
  }
  return AWK_OK;
}

// backward sweep of every subject, then the mean
static awk_status la_refresh(awk_ldmm_atlas *a, hipStream_t s) {
  const int N = a->eng.nx * a->eng.ny * a->eng.nz;
  awk_status st = lanes_fork(a, s);
  if (st) return st;
  for (int k = 0; k < a->S; ++k) {
    LaSubject &sb = a->subs[k];
    st = ldmm_orig_backward(lane_eng(a, k), sb.I, sb.v.data(), sb.jt, sb.jd, lane_stream(a, k, s));
    if (st) return st;
  }
  if ((st = lanes_join(a, s))) return st;
  AWK_LAUNCH(k_la_mean, dim3((N + kLaBlock - 1) / kLaBlock), dim3(kLaBlock), 0, s, a->mean, a->jt0,
             a->jd0, a->w, a->S, N, a->jscale ? 1 : 0);
  AWK_LAUNCH_CHECK("k_la_mean");
  a->current = true;
  return AWK_OK;
}

static awk_status la_check(awk_ldmm_atlas *a, const char *fn) {
  if (!a) return fail(AWK_ERR_INVALID, std::string(fn) + ": null handle");
  for (int k = 0; k < a->S; ++k)
    if (!a->subs[k].set)
      return fail(AWK_ERR_INVALID, std::string(fn) + ": subject " + std::to_string(k) + " not set");
  return AWK_OK;
}

extern "C" {

awk_status awk_ldmm_atlas_create(awk_ldmm_atlas_t *out, int nx, int ny, int nz, float spx,
                                 float spy, float spz, float alpha, float beta, float gamma,
                                 int n_time_steps, float sigma, float max_pert, int n_subjects,
                                 int jacobian_scale) {
  if (!out) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_create: null handle pointer");
  *out = nullptr;
  if (!(spx > 0 && spy > 0 && spz > 0)) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_create: spacing must be > 0");
  if (n_time_steps < 1) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_create: need at least one time step");
  if (!(sigma > 0) || !(max_pert > 0)) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_create: sigma and MaxPert must be > 0");
  if (n_subjects < 1 || n_subjects > 4096) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_create: 1 to 4096 subjects");
  if ((long long)nx * ny * nz > (1LL << 28)) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_create: at most 2^28 voxels");
  awk_ldmm_atlas *a = new (std::nothrow) awk_ldmm_atlas();
  if (!a) return fail(AWK_ERR_ALLOC, "awk_ldmm_atlas_create: host allocation failed");
  const float sp[3] = {spx, spy, spz};
  awk_status st = ldmm_init(&a->eng, nx, ny, nz, sp, alpha, beta, gamma, n_time_steps, sigma,
                            max_pert, false);
  hipError_t le = hipSuccess;
  if (!st) le = hipEventCreateWithFlags(&a->fork, hipEventDisableTiming);
  for (int l = 0; !st && le == hipSuccess && l < std::min(n_subjects, kLaLanes) - 1; ++l) {
    a->lanes.emplace_back();
    LaLane &L = a->lanes.back();
    st = ldmm_init(&L.eng, nx, ny, nz, sp, alpha, beta, gamma, n_time_steps, sigma, max_pert, false);
    if (!st) le = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking);
    if (!st && le == hipSuccess) le = hipEventCreateWithFlags(&L.join, hipEventDisableTiming);
  }
  if (!st && le != hipSuccess) st = hip_check(le, "awk_ldmm_atlas_create: lanes");
  const size_t N = (size_t)nx * ny * nz;
  const int S = n_subjects;
  a->S = S;
  a->jscale = jacobian_scale != 0;
  a->subs.resize(S);
  a->w_host.assign(S, 0.f);
  hipError_t e = hipSuccess;
  if (!st) {
    e = hipMalloc(&a->mean, N * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&a->w, S * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&a->jt0, S * sizeof(float *));
    if (e == hipSuccess) e = hipMalloc(&a->jd0, S * sizeof(float *));
    if (e == hipSuccess) e = hipMalloc(&a->esub, 3 * (size_t)S * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&a->energy, 3 * sizeof(double));
    std::vector<const float *> p0(S), p1(S);
    for (int k = 0; k < S && e == hipSuccess; ++k) {
      LaSubject &sb = a->subs[k];
      sb.v.assign(3 * (size_t)n_time_steps, nullptr);
      e = hipMalloc(&sb.I, N * sizeof(float));
      if (e == hipSuccess) e = hipMalloc(&sb.jt, (size_t)n_time_steps * N * sizeof(float));
      if (e == hipSuccess) e = hipMalloc(&sb.jd, (size_t)n_time_steps * N * sizeof(float));
      if (e == hipSuccess) e = hipMalloc(&sb.step, sizeof(float));
      if (e == hipSuccess) e = hipMemset(sb.step, 0, sizeof(float));
      p0[k] = sb.jt;  // t = 0 slots
      p1[k] = sb.jd;
    }
    if (e == hipSuccess) e = hipMemcpy(a->jt0, p0.data(), S * sizeof(float *), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(a->jd0, p1.data(), S * sizeof(float *), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(a->esub, 0, 3 * (size_t)S * sizeof(double));
    if (e != hipSuccess) st = hip_check(e, "awk_ldmm_atlas_create: allocation");
  }
  if (st) {
    la_free(a);
    delete a;
    return st;
  }
  *out = a;
  return AWK_OK;
}

awk_status awk_ldmm_atlas_destroy(awk_ldmm_atlas_t a) {
  if (!a) return AWK_OK;
  la_free(a);
  delete a;
  return AWK_OK;
}

awk_status awk_ldmm_atlas_set_subject(awk_ldmm_atlas_t a, int s, const float *image, float weight,
                                      float *const v[], awk_stream_t stream) {
  if (!a || !image || !v) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_set_subject: null argument");
  if (s < 0 || s >= a->S) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_set_subject: subject out of range");
  if (!(weight >= 0.f)) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_set_subject: weight must be >= 0");
  const int T = a->eng.nt;
  for (int i = 0; i < 3 * T; ++i)
    if (!v[i]) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_set_subject: null velocity plane");
  LaSubject &sb = a->subs[s];
  const size_t N = (size_t)a->eng.nx * a->eng.ny * a->eng.nz;
  hipStream_t st = as_stream(stream);
  AWK_HIP(hipMemcpyAsync(sb.I, image, N * sizeof(float), hipMemcpyDeviceToDevice, st));
  for (int i = 0; i < 3 * T; ++i) sb.v[i] = v[i];
  a->w_host[s] = weight;
  AWK_HIP(hipMemcpyAsync(a->w, a->w_host.data(), a->S * sizeof(float), hipMemcpyHostToDevice, st));
  AWK_HIP(hipStreamSynchronize(st));  // w_host is read by the copy
  sb.set = true;
  a->current = false;
  return AWK_OK;
}

awk_status awk_ldmm_atlas_set_gradient(awk_ldmm_atlas_t a, int use_orig) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_set_gradient: null handle");
  a->orig = use_orig != 0;
  return AWK_OK;
}

awk_status awk_ldmm_atlas_set_splat(awk_ldmm_atlas_t a, int mode) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_set_splat: null handle");
  awk_status st = awk_ldmm_set_splat(&a->eng, mode);
  for (LaLane &L : a->lanes)
    if (!st) st = awk_ldmm_set_splat(&L.eng, mode);
  return st;
}

awk_status awk_ldmm_atlas_update_step_next(awk_ldmm_atlas_t a) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_update_step_next: null handle");
  for (LaSubject &sb : a->subs) sb.update_step = true;
  return AWK_OK;
}

awk_status awk_ldmm_atlas_get_flags(awk_ldmm_atlas_t a, unsigned *flags, int clear,
                                    awk_stream_t stream) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_get_flags: null handle");
  // the iterate calls join their lanes back into the caller's stream
  std::vector<unsigned *> words{a->eng.flags};
  for (LaLane &L : a->lanes) words.push_back(L.eng.flags);
  return read_flags(words.data(), (int)words.size(), flags, clear, as_stream(stream),
                    "awk_ldmm_atlas_get_flags");
}

awk_status awk_ldmm_atlas_set_step(awk_ldmm_atlas_t a, float step, awk_stream_t stream) {
  if (!a) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_set_step: null handle");
  for (LaSubject &sb : a->subs) {
    AWK_HIP(store_scalar_async(sb.step, step, as_stream(stream)));
    sb.update_step = false;
  }
  return AWK_OK;
}

awk_status awk_ldmm_atlas_get_step(awk_ldmm_atlas_t a, int s, float *step, awk_stream_t stream) {
  if (!a || !step) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_get_step: null argument");
  if (s < 0 || s >= a->S) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_get_step: subject out of range");
  AWK_HIP(load_scalar_sync(step, a->subs[s].step, as_stream(stream)));
  return AWK_OK;
}

awk_status awk_ldmm_atlas_iterate(awk_ldmm_atlas_t a, int n_iter, double *energy_dev,
                                  int vec_energy, awk_stream_t stream) {
  awk_status st = la_check(a, "awk_ldmm_atlas_iterate");
  if (st) return st;
  if (n_iter < 0) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_iterate: n_iter must be >= 0");
  hipStream_t s = as_stream(stream);
  for (int it = 0; it < n_iter; ++it) {
    if (!a->current && (st = la_refresh(a, s))) return st;
    if ((st = lanes_fork(a, s))) return st;
    for (int k = 0; k < a->S; ++k) {
      LaSubject &sb = a->subs[k];
      awk_ldmm *eng = lane_eng(a, k);
      const hipStream_t ls = lane_stream(a, k, s);
      for (int pass = sb.update_step ? 0 : 1; pass < 2; ++pass) {
        // pass 0: LDMMAtlasThreadGPU::BeginImageIteration's step-size call (its
        // energies go to the lane's own unused slot)
        const bool step_call = pass == 0;
        double *e = step_call ? eng->energy : a->esub + 3 * (size_t)k;
        if (a->orig)
          st = ldmm_orig_forward(eng, a->mean, sb.I, sb.v.data(), sb.jt, sb.jd, sb.step, step_call,
                                 nullptr, e, step_call ? 0 : vec_energy, ls);
        else
          st = ldmm_new_iterate(eng, a->mean, sb.I, sb.v.data(), sb.step, step_call, nullptr, e,
                                step_call ? 0 : vec_energy, ls);
        if (st) return st;
      }
      sb.update_step = false;
    }
    if ((st = lanes_join(a, s))) return st;
    a->current = false;
    AWK_LAUNCH(k_la_energy, dim3(1), dim3(64), 0, s, a->esub, a->w, a->S,
               energy_dev ? energy_dev + 3 * (size_t)it : a->energy);
    AWK_LAUNCH_CHECK("k_la_energy");
  }
  if (!a->current && (st = la_refresh(a, s))) return st;
  return AWK_OK;
}

awk_status awk_ldmm_atlas_mean(awk_ldmm_atlas_t a, float *mean_out, awk_stream_t stream) {
  awk_status st = la_check(a, "awk_ldmm_atlas_mean");
  if (st) return st;
  if (!mean_out) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_mean: null output");
  hipStream_t s = as_stream(stream);
  if (!a->current && (st = la_refresh(a, s))) return st;
  const size_t N = (size_t)a->eng.nx * a->eng.ny * a->eng.nz;
  AWK_HIP(hipMemcpyAsync(mean_out, a->mean, N * sizeof(float), hipMemcpyDeviceToDevice, s));
  return AWK_OK;
}

awk_status awk_ldmm_atlas_deformed_image(awk_ldmm_atlas_t a, int s, int t, float *out,
                                         awk_stream_t stream) {
  awk_status st = la_check(a, "awk_ldmm_atlas_deformed_image");
  if (st) return st;
  if (!out) return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_deformed_image: null output");
  if (s < 0 || s >= a->S || t < 0 || t >= a->eng.nt)
    return fail(AWK_ERR_INVALID, "awk_ldmm_atlas_deformed_image: subject or time step out of range");
  hipStream_t hs = as_stream(stream);
  if (!a->current && (st = la_refresh(a, hs))) return st;
  const size_t N = (size_t)a->eng.nx * a->eng.ny * a->eng.nz;
  AWK_HIP(hipMemcpyAsync(out, a->subs[s].jt + (size_t)t * N, N * sizeof(float),
                         hipMemcpyDeviceToDevice, hs));
  return AWK_OK;
}

}  // extern "C"
