// awk_ldmm.h -- internal: the LDMM iterator state and the IterateOrig sweeps,
// shared by the single-registration iterator and the LDMM atlas.
#pragma once

#include "awk_fluid.h"

struct awk_ldmm {
  awk::FluidKernel k;
  int nx = 0, ny = 0, nz = 0, nt = 0;
  float sp[3] = {1, 1, 1};
  float sigma = 0.05f, max_pert = 0.1f;
  bool update_step = true;
  float *h = nullptr, *hs = nullptr;  // 3 planes each
  float *J = nullptr;                  // nt images
  float *diff = nullptr;               // IterateOrig: the force multiplier m
  bool orig = true;                    // mUseOrigGradientCalc (LDMMIteratorGPU.cxx:37)
  float *jt = nullptr, *jd = nullptr;  // IterateOrig: JT_t, JacDet_t, nt images each
  long long *acc = nullptr;
  double *partials = nullptr;  // max(forward grid, dot grid) doubles
  float *maxpart = nullptr;    // nt * nblk
  float *step = nullptr;
  double *energy = nullptr;    // 3 doubles (internal when the caller passes none)
  unsigned *flags = nullptr;   // sticky AWK_FLAG_* word
  int nblk = 0, egrid = 0;
  int splat_mode = 0;          // IterateNew's splat: 0 int64 fixed point (LDS tiles), 1 sort/scan
  void *sort_work = nullptr;   // sort/scan scratch
  float *msplat = nullptr;     // sort/scan splat output (N floats)
};

namespace awk {

// Allocates l's device state; terms = also IterateOrig's JT_t / JacDet_t buffers.
awk_status ldmm_init(awk_ldmm *l, int nx, int ny, int nz, const float sp[3], float alpha,
                     float beta, float gamma, int n_time_steps, float sigma, float max_pert,
                     bool terms);
void ldmm_release(awk_ldmm *l);

// IterateOrig's backward sweep (LDMMIteratorGPU.cxx:312-345): JT_t = I1(h_t) and
// JacDet_t into jt, jd ([T][N] each) for the fields v.
awk_status ldmm_orig_backward(awk_ldmm *l, const float *I1, float *const v[], float *jt, float *jd,
                              hipStream_t s);
// IterateOrig's forward sweep (LDMMIteratorGPU.cxx:347-441) from the backward
// terms: updates v (or, step_call, writes MaxPert / max|...| to *step), or, with
// alpha, writes the alphas only.  energy: 3 device doubles (image, vector, total).
awk_status ldmm_orig_forward(awk_ldmm *l, const float *I0, const float *I1, float *const v[],
                             const float *jt, const float *jd, float *step, bool step_call,
                             float *const alpha[], double *energy, int vec_energy, hipStream_t s);

// IterateNew (LDMMIteratorGPU.cxx:159-293): forward gathers, adjoint splat,
// update of v (or, step_call, MaxPert / max|...| into *step), or the alphas only.
awk_status ldmm_new_iterate(awk_ldmm *l, const float *I0, const float *I1, float *const v[],
                            float *step, bool step_call, float *const alpha[], double *energy,
                            int vec_energy, hipStream_t s);

}  // namespace awk
