// LDMMIteratorGPU.h -- LDMM registration and atlas behind the reference's iterator
// API (Algorithms/LDMM/LDMMIteratorGPU.h, LDMMDeformationDataGPU.h,
// LDMMAtlasThreadGPU.h): SetScaleLevel, UpdateStepSizeNextIteration,
// Iterate(data, computeAlphasOnly), GetAlpha(t).  Header-only over the C-ABI.
#ifndef AWK_LDMMITERATORGPU_H
#define AWK_LDMMITERATORGPU_H

#include <vector>

#include "cplVector3DArray.h"
#include "DiffOperGPU.h"

// LDMMIteratorParam (Algorithms/LDMM/LDMMIteratorParam.h:28-62) + its kernel's
// DiffOperParam; the reference's defaults.
struct LDMMIteratorParam {
  float sigma = 5.f;
  float stepSize = 0.005f;
  bool useAdaptiveStepSize = true;
  float maxPert = 0.1f;
  DiffOperParam kernel;
  float &Sigma() { return sigma; }
  float &StepSize() { return stepSize; }
  bool &UseAdaptiveStepSize() { return useAdaptiveStepSize; }
  float &MaxPert() { return maxPert; }
  DiffOperParam &Kernel() { return kernel; }
};

// Per-registration device state (LDMMDeformationDataGPU): I0, I1 and the T velocity
// fields, caller-owned device memory; v is updated in place.
class LDMMDeformationDataGPU {
 public:
  LDMMDeformationDataGPU(float *dI0, float *dI1, const std::vector<cplVector3DArray> &v)
      : mI0(dI0), mI1(dI1), mV(v) {
    for (cplVector3DArray &f : mV) {
      mPlanes.push_back(f.x);
      mPlanes.push_back(f.y);
      mPlanes.push_back(f.z);
    }
  }
  float *dI0() { return mI0; }
  float *dI1() { return mI1; }
  cplVector3DArray &dV(unsigned t) { return mV[t]; }
  unsigned NTimeSteps() const { return (unsigned)mV.size(); }
  // (image, vector, total) energy of each iteration run so far (host copy, synchronises)
  std::vector<double> Energies() {
    std::vector<double> e(3 * mEnergyCount);
    if (mEnergyCount)
      (void)hipMemcpy(e.data(), mEnergy, e.size() * sizeof(double), hipMemcpyDeviceToHost);
    return e;
  }
  ~LDMMDeformationDataGPU() {
    if (mEnergy) (void)hipFree(mEnergy);
  }

 private:
  friend class LDMMIteratorGPU;
  double *energySlot() {
    if (mEnergyCount == mEnergyCap) {
      double *grown = nullptr;
      const unsigned cap = mEnergyCap ? 2 * mEnergyCap : 256;
      if (hipMalloc(&grown, 3 * cap * sizeof(double)) != hipSuccess)
        throw AtlasWerksException(__FILE__, __LINE__, "energy buffer allocation failed");
      if (mEnergy) {
        (void)hipMemcpy(grown, mEnergy, 3 * mEnergyCount * sizeof(double), hipMemcpyDeviceToDevice);
        (void)hipFree(mEnergy);
      }
      mEnergy = grown;
      mEnergyCap = cap;
    }
    return mEnergy + 3 * mEnergyCount++;
  }
  float *mI0, *mI1;
  std::vector<cplVector3DArray> mV;
  std::vector<float *> mPlanes;
  double *mEnergy = nullptr;
  unsigned mEnergyCount = 0, mEnergyCap = 0;
};

class LDMMIteratorGPU {
 public:
  // LDMMIteratorGPU(size, origin, spacing, nTimeSteps, debug) (LDMMIteratorGPU.cxx:26-60);
  // the origin does not enter the iteration.
  LDMMIteratorGPU(const SizeType &size, const SpacingType &spacing, unsigned nTimeSteps,
                  const LDMMIteratorParam &p, bool debug = true)
      : mNT(nTimeSteps), mDebug(debug) {
    SetScaleLevel(size, spacing, p);
  }
  ~LDMMIteratorGPU() { release(); }
  LDMMIteratorGPU(const LDMMIteratorGPU &) = delete;
  LDMMIteratorGPU &operator=(const LDMMIteratorGPU &) = delete;

  void SetScaleLevel(const SizeType &size, const SpacingType &spacing, const LDMMIteratorParam &p) {
    release();
    mSize = size;
    const DiffOperParam &k = p.kernel;
    AWK_CHECK(awk_ldmm_create(&mL, (int)size.x, (int)size.y, (int)size.z, spacing.x, spacing.y,
                              spacing.z, k.alpha, k.beta, k.gamma, (int)mNT, p.sigma, p.maxPert));
    AWK_CHECK(awk_ldmm_set_gradient(mL, mOrig ? 1 : 0));
    if (!p.useAdaptiveStepSize) AWK_CHECK(awk_ldmm_set_step(mL, p.stepSize, nullptr));
    const size_t N = (size_t)size.x * size.y * size.z;
    freeAlphas();
    for (unsigned t = 0; t < mNT; ++t) {
      float *a = nullptr;
      if (hipMalloc(&a, N * sizeof(float)) != hipSuccess)
        throw AtlasWerksException(__FILE__, __LINE__, "alpha allocation failed");
      mAlpha.push_back(a);
    }
  }
  // mUseOrigGradientCalc (LDMMIteratorGPU.cxx:37): IterateOrig (true, the default)
  // or IterateNew.
  void SetUseOrigGradientCalc(bool on) {
    mOrig = on;
    AWK_CHECK(awk_ldmm_set_gradient(mL, on ? 1 : 0));
  }
  void UpdateStepSizeNextIteration() { AWK_CHECK(awk_ldmm_update_step_next(mL)); }
  float StepSize() {
    float s = 0.f;
    AWK_CHECK(awk_ldmm_get_step(mL, &s, nullptr));
    return s;
  }

  // Iterate (LDMMIteratorGPU.cxx:147-156): updates the data's velocity fields, or
  // with computeAlphasOnly leaves them and fills GetAlpha(t).
  void Iterate(LDMMDeformationDataGPU &d, bool computeAlphasOnly = false, awkStream_t s = nullptr) {
    if (d.NTimeSteps() != mNT)
      throw AtlasWerksException(__FILE__, __LINE__, "deformation data has a different number of time steps");
    if (computeAlphasOnly) {
      AWK_CHECK(awk_ldmm_alphas(mL, d.dI0(), d.dI1(), d.mPlanes.data(), mAlpha.data(), s));
    } else {
      AWK_CHECK(awk_ldmm_iterate(mL, d.dI0(), d.dI1(), d.mPlanes.data(),
                                 mDebug ? d.energySlot() : nullptr, mDebug ? 1 : 0, s));
    }
  }
  // alpha_t of the last computeAlphasOnly iterate (device, N floats)
  float *GetAlpha(unsigned t) { return mAlpha.at(t); }

 private:
  void freeAlphas() {
    for (float *a : mAlpha) (void)hipFree(a);
    mAlpha.clear();
  }
  void release() {
    if (mL) awk_ldmm_destroy(mL);
    mL = nullptr;
    freeAlphas();
  }
  awk_ldmm_t mL = nullptr;
  SizeType mSize;
  unsigned mNT;
  bool mDebug;
  bool mOrig = true;
  std::vector<float *> mAlpha;
};

// The per-device part of LDMMAtlasBuilder / LDMMAtlasThreadGPU (LDMMAtlasBuilder.txx,
// LDMMAtlasThreadGPU.cxx), single scale: subjects' images and weights, the caller's
// velocity fields per subject, Iterate(n), the (Jacobian-scaled) mean.
class LDMMAtlasGPU {
 public:
  LDMMAtlasGPU(const SizeType &size, const SpacingType &spacing, unsigned nTimeSteps,
               unsigned nSubjects, const LDMMIteratorParam &p, bool jacobianScale = true) {
    const DiffOperParam &k = p.kernel;
    AWK_CHECK(awk_ldmm_atlas_create(&mA, (int)size.x, (int)size.y, (int)size.z, spacing.x,
                                    spacing.y, spacing.z, k.alpha, k.beta, k.gamma, (int)nTimeSteps,
                                    p.sigma, p.maxPert, (int)nSubjects, jacobianScale ? 1 : 0));
    if (!p.useAdaptiveStepSize) AWK_CHECK(awk_ldmm_atlas_set_step(mA, p.stepSize, nullptr));
  }
  ~LDMMAtlasGPU() {
    if (mA) awk_ldmm_atlas_destroy(mA);
  }
  LDMMAtlasGPU(const LDMMAtlasGPU &) = delete;
  LDMMAtlasGPU &operator=(const LDMMAtlasGPU &) = delete;

  void SetUseOrigGradientCalc(bool on) { AWK_CHECK(awk_ldmm_atlas_set_gradient(mA, on ? 1 : 0)); }
  void SetSubject(unsigned s, const float *dImage, float weight, std::vector<cplVector3DArray> &v,
                  awkStream_t st = nullptr) {
    std::vector<float *> planes;
    for (cplVector3DArray &f : v) {
      planes.push_back(f.x);
      planes.push_back(f.y);
      planes.push_back(f.z);
    }
    AWK_CHECK(awk_ldmm_atlas_set_subject(mA, (int)s, dImage, weight, planes.data(), st));
  }
  // n iterations; energies (3 per iteration) into the caller's device buffer or none
  void Iterate(unsigned n, double *dEnergy = nullptr, awkStream_t st = nullptr) {
    AWK_CHECK(awk_ldmm_atlas_iterate(mA, (int)n, dEnergy, 1, st));
  }
  void GetMean(float *dMean, awkStream_t st = nullptr) { AWK_CHECK(awk_ldmm_atlas_mean(mA, dMean, st)); }
  // subject s at time t (GetI1AtT)
  void GetDefToMean(unsigned s, unsigned t, float *dOut, awkStream_t st = nullptr) {
    AWK_CHECK(awk_ldmm_atlas_deformed_image(mA, (int)s, (int)t, dOut, st));
  }

 private:
  awk_ldmm_atlas_t mA = nullptr;
};

#endif
