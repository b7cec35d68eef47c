// LDMMIteratorGPU.h -- LDMM registration and atlas behind the reference's driver
// API, so that LDMMWarp<LDMMIteratorGPU> binds it at compile time
// (LDMM/LDMMIteratorGPU.h:28-80, LDMMDeformationDataGPU.h:25-95,
// Common/DeformationIteratorInterface.h:84-92):
//   LDMMDeformationDataGPU: the host images and velocity fields of
//     LDMMDeformationData plus their device copies; InitializeWarp /
//     SetScaleLevel / FinalizeWarp, the scale-level work on the device (velocity
//     fields zeroed at the first level, resampled with CLAMP afterwards,
//     LDMMDeformationData.cxx:116-161).
//   LDMMIteratorGPU: DeformationDataType / ParamType, SetScaleLevel(
//     MultiscaleManager, LDMMIteratorParam), UpdateStepSizeNextIteration,
//     Iterate(data[, computeAlphasOnly]), GetAlpha(t).
// With debug on, each iteration checks the numerical flags and throws on a NaN
// energy or an invalid step, as the reference's iterator throws
// (LDMMIteratorGPU.cxx:238-241, 283-285, 439-441).  Header-only over the C-ABI.
#ifndef AWK_LDMMITERATORGPU_H
#define AWK_LDMMITERATORGPU_H

#include <vector>

#include "CUDAUtilities.h"
#include "DiffOperGPU.h"
#include "GreedyIteratorGPU.h"
#include "awk_drivers.h"

class LDMMDeformationDataGPU : public LDMMDeformationData {
 public:
  LDMMDeformationDataGPU(const RealImage *I0, const RealImage *I1, const LDMMParam &param)
      : LDMMDeformationData(I0, I1, param), mEnergies(3) {}
  ~LDMMDeformationDataGPU() override { FreeDeviceData(); }

  void InitializeWarp() override { InitializeDeviceData(); }
  void FinalizeWarp() override {
    SyncEnergies();
    syncFields();
    FreeDeviceData();
  }
  void InitializeDeviceData() {
    if (mDeviceDataInitialized) return;
    const size_t n = (size_t)mI0Orig->getNumElements();
    alloc(&mdI0Orig, n);
    alloc(&mdI1Orig, n);
    alloc(&mdI0Scaled, n);
    alloc(&mdI1Scaled, n);
    alloc(&mdV, 3 * n * mNTimeSteps);
    alloc(&mdScratch, 3 * n);
    if (hipMemcpy(mdI0Orig, mI0Orig->getDataPointer(), n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(mdI1Orig, mI1Orig->getDataPointer(), n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      throw AtlasWerksException(__FILE__, __LINE__, "image upload failed");
    mdI0Ptr = mdI0Orig;
    mdI1Ptr = mdI1Orig;
    mNFull = n;
    mDeviceDataInitialized = true;
  }
  void FreeDeviceData() {
    float **bufs[] = {&mdI0Orig, &mdI1Orig, &mdI0Scaled, &mdI1Scaled, &mdV, &mdScratch};
    for (float **b : bufs) {
      if (*b) (void)hipFree(*b);
      *b = nullptr;
    }
    mdI0Ptr = mdI1Ptr = nullptr;
    mdVt.clear();
    mDeviceDataInitialized = false;
  }

  void SetScaleLevel(const MultiscaleManager &scaleManager) override {
    InitializeDeviceData();
    SyncEnergies();
    const SizeType prev = mCurSize;
    const bool initial = scaleManager.InitialScaleLevel() || !mHaveLevel;
    DeformationDataInterface::SetScaleLevel(scaleManager);
    const SizeType cur = mCurSize;
    const size_t n = cur.productOfElements();
    if (initial) {
      // v_t = 0 (LDMMDeformationData.cxx:121-130)
      if (hipMemset(mdV, 0, 3 * n * mNTimeSteps * sizeof(float)) != hipSuccess)
        throw AtlasWerksException(__FILE__, __LINE__, "velocity initialisation failed");
      layout(cur);
    } else {
      // UpsampleToLevel(v_t) (LDMMDeformationData.cxx:131-138), each field through
      // the scratch planes; the new layout packs the fields at the new size
      // (moving up in size, so the fields are processed last to first)
      const int rescale = scaleManager.GetScaleVectorFields() ? 1 : 0;
      const size_t np = prev.productOfElements();
      for (int t = (int)mNTimeSteps - 1; t >= 0; --t) {
        float *src = mdV + 3 * np * t;
        AWK_CHECK(awk_hfield_resample(mdScratch, mdScratch + n, mdScratch + 2 * n, (int)cur.x,
                                      (int)cur.y, (int)cur.z, src, src + np, src + 2 * np,
                                      (int)prev.x, (int)prev.y, (int)prev.z, AWK_BG_CLAMP, rescale,
                                      nullptr));
        if (hipMemcpy(mdV + 3 * n * t, mdScratch, 3 * n * sizeof(float), hipMemcpyDeviceToDevice) != hipSuccess)
          throw AtlasWerksException(__FILE__, __LINE__, "velocity upsampling failed");
      }
      layout(cur);
    }
    if (scaleManager.FinalScaleLevel()) {
      mdI0Ptr = mdI0Orig;
      mdI1Ptr = mdI1Orig;
    } else {
      const SizeType o = mI0Orig->getSize();
      if (mScaleI0) {
        AWK_CHECK(awk_gaussian_downsample(mdI0Scaled, (int)cur.x, (int)cur.y, (int)cur.z, mdI0Orig,
                                          (int)o.x, (int)o.y, (int)o.z, nullptr));
        mdI0Ptr = mdI0Scaled;
      }
      if (mScaleI1) {
        AWK_CHECK(awk_gaussian_downsample(mdI1Scaled, (int)cur.x, (int)cur.y, (int)cur.z, mdI1Orig,
                                          (int)o.x, (int)o.y, (int)o.z, nullptr));
        mdI1Ptr = mdI1Scaled;
      }
    }
    mHaveLevel = true;
  }

  // I0 at the last time step's hfield is the iterator's; these host views give
  // the velocity fields, brought up to date from the device
  std::vector<VectorField *> &v() override {
    syncFields();
    return mV;
  }
  VectorField &v(int i) override {
    syncFields();
    return *mV[i];
  }
  void GetI0At1(RealImage &) override {
    throw AtlasWerksException(__FILE__, __LINE__, "LDMMDeformationDataGPU::GetI0At1: use the iterator's deformed images");
  }
  void GetI1At0(RealImage &) override {
    throw AtlasWerksException(__FILE__, __LINE__, "LDMMDeformationDataGPU::GetI1At0: use the iterator's deformed images");
  }
  EnergyHistory &GetEnergyHistory() override {
    SyncEnergies();
    return mEnergyHistory;
  }

  float *dI0() { return mdI0Ptr; }
  float *dI1() { return mdI1Ptr; }
  std::vector<cplVector3DArray> &dV() { return mdVt; }
  cplVector3DArray &dV(int i) { return mdVt[i]; }
  // the 3T plane pointers the C-ABI takes
  float *const *Planes() { return mPlanes.data(); }

  double *NextEnergySlot() { return mEnergies.Next(mCurIter); }
  // AddEnergy with the total energy (image + vector) of each recorded iteration
  void SyncEnergies() {
    const unsigned keep = mCurIter;
    for (auto &rec : mEnergies.Take()) {
      SetCurIter(rec.first);
      Energy e;
      e.SetEnergy((Real)rec.second[2]);
      AddEnergy(e);
    }
    SetCurIter(keep);
  }

 private:
  static void alloc(float **p, size_t n) {
    if (hipMalloc(p, n * sizeof(float)) != hipSuccess)
      throw AtlasWerksException(__FILE__, __LINE__, "device allocation failed");
  }
  void layout(const SizeType &s) {
    const unsigned n = s.productOfElements();
    mdVt.clear();
    mPlanes.clear();
    for (unsigned t = 0; t < mNTimeSteps; ++t) {
      float *b = mdV + 3 * (size_t)n * t;
      mdVt.push_back(cplVector3DArray(b, b + n, b + 2 * (size_t)n, n));
      mPlanes.push_back(b);
      mPlanes.push_back(b + n);
      mPlanes.push_back(b + 2 * (size_t)n);
    }
  }
  void syncFields() {
    if (!mDeviceDataInitialized || !mHaveLevel) return;
    for (unsigned t = 0; t < mNTimeSteps; ++t) {
      mV[t]->resize(mCurSize);
      CUDAUtilities::CopyVectorFieldFromDevice(mdVt[t], *mV[t], true);
    }
  }

  bool mDeviceDataInitialized = false, mHaveLevel = false;
  size_t mNFull = 0;
  float *mdI0Orig = nullptr, *mdI1Orig = nullptr, *mdI0Scaled = nullptr, *mdI1Scaled = nullptr;
  float *mdI0Ptr = nullptr, *mdI1Ptr = nullptr;
  float *mdV = nullptr, *mdScratch = nullptr;
  std::vector<cplVector3DArray> mdVt;
  std::vector<float *> mPlanes;
  AwkDeviceEnergies mEnergies;
};

class LDMMIteratorGPU : public DeformationIteratorInterface {
 public:
  typedef LDMMDeformationDataGPU DeformationDataType;
  typedef LDMMIteratorParam ParamType;

  // LDMMIteratorGPU(size, origin, spacing, nTimeSteps, debug) (LDMMIteratorGPU.cxx:26-60)
  LDMMIteratorGPU(const SizeType &size, const OriginType &origin, const SpacingType &spacing,
                  const unsigned int &nTimeSteps, bool debug = true)
      : DeformationIteratorInterface(size, origin, spacing), mNT(nTimeSteps), mDebug(debug) {}
  ~LDMMIteratorGPU() override { release(); }
  LDMMIteratorGPU(const LDMMIteratorGPU &) = delete;
  LDMMIteratorGPU &operator=(const LDMMIteratorGPU &) = delete;

  // LDMMIteratorGPU::SetScaleLevel (LDMMIteratorGPU.cxx:62-120): the level's grid,
  // operator, sigma, MaxPert; a fixed step when the step is not adaptive
  using ScaleLevelDataInfo::SetScaleLevel;
  void SetScaleLevel(const MultiscaleManager &scaleManager, const LDMMIteratorParam &param) {
    ScaleLevelDataInfo::SetScaleLevel(scaleManager);
    const KernelParam &kp = param.Kernel();
    if (!kp.IsDiffOperParam())
      throw AtlasWerksException(__FILE__, __LINE__, "LDMMIteratorGPU: only the DiffOper kernel is built");
    const DiffOperParam &k = *kp.AsDiffOperParam();
    release();
    AWK_CHECK(awk_ldmm_create(&mL, (int)mCurSize.x, (int)mCurSize.y, (int)mCurSize.z,
                              mCurSpacing.x, mCurSpacing.y, mCurSpacing.z, k.Alpha(), k.Beta(),
                              k.Gamma(), (int)mNT, param.Sigma(), param.MaxPert()));
    AWK_CHECK(awk_ldmm_set_gradient(mL, mOrig ? 1 : 0));
    if (!param.UseAdaptiveStepSize()) AWK_CHECK(awk_ldmm_set_step(mL, param.StepSize(), nullptr));
    const size_t N = mCurVox;
    for (unsigned t = 0; t < mNT; ++t) {
      float *a = nullptr;
      if (hipMalloc(&a, N * sizeof(float)) != hipSuccess)
        throw AtlasWerksException(__FILE__, __LINE__, "alpha allocation failed");
      mAlpha.push_back(a);
    }
    mUpdateStep = false;
  }
  // mUseOrigGradientCalc (LDMMIteratorGPU.cxx:37): IterateOrig (true, the default)
  // or IterateNew.
  void SetUseOrigGradientCalc(bool on) {
    mOrig = on;
    if (mL) AWK_CHECK(awk_ldmm_set_gradient(mL, on ? 1 : 0));
  }
  void UpdateStepSizeNextIteration() {
    check();
    mUpdateStep = true;
    AWK_CHECK(awk_ldmm_update_step_next(mL));
  }
  void SetStream(awkStream_t s) { mStream = s; }

  void Iterate(DeformationDataInterface &defData) override { Iterate(defData, false); }
  // Iterate (LDMMIteratorGPU.cxx:147-156): updates the data's velocity fields, or
  // with computeAlphasOnly leaves them and fills GetAlpha(t).  A step-size
  // iteration only sets the step, which goes to the data (StepSize).
  void Iterate(DeformationDataInterface &defData, bool computeAlphasOnly) {
    check();
    LDMMDeformationDataGPU &d = dynamic_cast<LDMMDeformationDataGPU &>(defData);
    if (d.dV().size() != mNT || d.dV(0).size() != mCurVox)
      throw AtlasWerksException(__FILE__, __LINE__, "deformation data does not match the iterator's level");
    if (computeAlphasOnly) {
      AWK_CHECK(awk_ldmm_alphas(mL, d.dI0(), d.dI1(), d.Planes(), mAlpha.data(), mStream));
      return;
    }
    const bool stepCall = mUpdateStep;
    AWK_CHECK(awk_ldmm_iterate(mL, d.dI0(), d.dI1(), d.Planes(),
                               (mDebug && !stepCall) ? d.NextEnergySlot() : nullptr,
                               mDebug ? 1 : 0, mStream));
    if (mDebug) {
      unsigned f = 0;
      awk_status st = awk_ldmm_get_flags(mL, &f, 1, mStream);
      if (st == AWK_ERR_NUMERIC)
        throw AtlasWerksException(__FILE__, __LINE__,
                                  (f & AWK_FLAG_NAN_ENERGY) ? "Error, NaN encountered"
                                                            : "Error, max displacement is NaN");
      AWK_CHECK(st);
    }
    if (stepCall) {
      mUpdateStep = false;
      float s = 0.f;
      AWK_CHECK(awk_ldmm_get_step(mL, &s, mStream));
      d.SyncEnergies();
      d.StepSize(s);
    }
  }
  // alpha_t of the last computeAlphasOnly iterate (device, N floats)
  float *GetAlpha(unsigned int t) { return mAlpha.at(t); }

 private:
  void check() const {
    if (!mL) throw AtlasWerksException(__FILE__, __LINE__, "LDMMIteratorGPU: SetScaleLevel must be called first");
  }
  void release() {
    if (mL) awk_ldmm_destroy(mL);
    mL = nullptr;
    for (float *a : mAlpha) (void)hipFree(a);
    mAlpha.clear();
  }
  awk_ldmm_t mL = nullptr;
  unsigned mNT;
  bool mDebug;
  bool mOrig = true;
  bool mUpdateStep = false;
  awkStream_t mStream = nullptr;
  std::vector<float *> mAlpha;
};

// The per-device part of LDMMAtlasBuilder / LDMMAtlasThreadGPU (LDMMAtlasBuilder.txx,
// LDMMAtlasThreadGPU.cxx), single scale: subjects' images and weights, the caller's
// velocity fields per subject, Iterate(n), the (Jacobian-scaled) mean.
class LDMMAtlasGPU {
 public:
  LDMMAtlasGPU(const SizeType &size, const SpacingType &spacing, unsigned nTimeSteps,
               unsigned nSubjects, const LDMMIteratorParam &p, bool jacobianScale = true) {
    const DiffOperParam &k = *p.Kernel().AsDiffOperParam();
    AWK_CHECK(awk_ldmm_atlas_create(&mA, (int)size.x, (int)size.y, (int)size.z, spacing.x,
                                    spacing.y, spacing.z, k.Alpha(), k.Beta(), k.Gamma(),
                                    (int)nTimeSteps, p.Sigma(), p.MaxPert(), (int)nSubjects,
                                    jacobianScale ? 1 : 0));
    if (!p.UseAdaptiveStepSize()) AWK_CHECK(awk_ldmm_atlas_set_step(mA, p.StepSize(), nullptr));
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
