// awk_drivers.h -- the reference's CPU-side driver vocabulary that the GPU
// iterators bind at compile time, for the standalone build:
//   parameters   DiffOperParam, KernelParam (Algorithms/Common/DiffOperParam.h,
//                KernelParam.h), GreedyIteratorParam / GreedyScaleLevelParam /
//                GreedyParam (Greedy/GreedyIteratorParam.h, GreedyParam.h),
//                LDMMIteratorParam / LDMMScaleLevelParam / LDMMParam
//                (LDMM/LDMMIteratorParam.h, LDMMParam.h), with the reference's
//                accessor names and defaults (no XML parsing);
//   scale levels MultiscaleManager (Common/MultiscaleManager.h:240-366): the
//                level sizes and spacings; the GPU data classes resample on the
//                device themselves;
//   interfaces   ScaleLevelDataInfo, DeformationDataInterface,
//                DeformationIteratorInterface (Common/DeformationIteratorInterface.h),
//                Energy / EnergyHistory (Common/Energy.h, EnergyHistory.h);
//   host data    GreedyDeformationData (Greedy/GreedyDeformationData.h) and
//                LDMMDeformationData (LDMM/LDMMDeformationData.h): the host
//                images, fields and bookkeeping the GPU subclasses extend.
// With -DAWK_USE_REFERENCE_TYPES the reference's own headers are used instead.
#ifndef AWK_DRIVERS_H
#define AWK_DRIVERS_H

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "awk_types.h"

#ifdef AWK_USE_REFERENCE_TYPES
#include "DeformationIteratorInterface.h"
#include "Energy.h"
#include "EnergyHistory.h"
#include "GreedyDeformationData.h"
#include "GreedyParam.h"
#include "KernelParam.h"
#include "LDMMDeformationData.h"
#include "LDMMParam.h"
#include "MultiscaleManager.h"
#else

// ------------------------------------------------------------------ parameters
// DiffOperParam (DiffOperParam.h:24-86): alpha 0.01, beta 0.01, gamma 0.001,
// LPow 1, UseEigenLUT false, DivergenceFree false.
class DiffOperParam {
 public:
  Real &Alpha() { return mAlpha; }
  Real &Beta() { return mBeta; }
  Real &Gamma() { return mGamma; }
  Real &LPow() { return mLPow; }
  bool &UseEigenLUT() { return mEigen; }
  bool &DivergenceFree() { return mDivFree; }
  const Real &Alpha() const { return mAlpha; }
  const Real &Beta() const { return mBeta; }
  const Real &Gamma() const { return mGamma; }
  const Real &LPow() const { return mLPow; }
  const bool &UseEigenLUT() const { return mEigen; }
  const bool &DivergenceFree() const { return mDivFree; }

 private:
  Real mAlpha = 0.01f, mBeta = 0.01f, mGamma = 0.001f, mLPow = 1.f;
  bool mEigen = false, mDivFree = false;
};

// KernelParam (KernelParam.h:20-90): an XOR of DiffOper and MultiGaussian; this
// build has the DiffOper kernel only.
class KernelParam {
 public:
  bool IsDiffOperParam() const { return true; }
  bool IsMultiGaussianKernelParam() const { return false; }
  DiffOperParam *AsDiffOperParam() { return &mDiffOper; }
  const DiffOperParam *AsDiffOperParam() const { return &mDiffOper; }
  DiffOperParam &DiffOper() { return mDiffOper; }

 private:
  DiffOperParam mDiffOper;
};

// GreedyIteratorParam (GreedyIteratorParam.h:20-40): Kernel, MaxPert 0.1.
class GreedyIteratorParam {
 public:
  KernelParam &Kernel() { return mKernel; }
  const KernelParam &Kernel() const { return mKernel; }
  Real &MaxPert() { return mMaxPert; }
  const Real &MaxPert() const { return mMaxPert; }

 private:
  KernelParam mKernel;
  Real mMaxPert = 0.1f;
};

// ScaleLevelParam (MultiscaleManager.h:60-100): DownsampleFactor 1.
class ScaleLevelParam {
 public:
  unsigned int &DownsampleFactor() { return mFactor; }
  const unsigned int &DownsampleFactor() const { return mFactor; }

 private:
  unsigned int mFactor = 1;
};

// MultiscaleParamBase (MultiscaleManager.h:133-215): the scale levels ordered by
// downsample factor, largest first (ScaleLevelParamOrderingConstraint), unique.
template <class LevelT>
class MultiscaleParamBase {
 public:
  unsigned int GetNumberOfScaleLevels() const { return (unsigned int)mLevels.size(); }
  LevelT &GetScaleLevel(int i) { return mLevels.at(i); }
  const LevelT &GetScaleLevel(int i) const { return mLevels.at(i); }
  void AddScaleLevel(const LevelT &level) {
    const unsigned f = level.ScaleLevel().DownsampleFactor();
    auto it = mLevels.begin();
    for (; it != mLevels.end(); ++it) {
      if (it->ScaleLevel().DownsampleFactor() == f)
        throw AtlasWerksException(__FILE__, __LINE__,
                                  "Trying to add multiple scale levels with the same downsample value");
      if (f > it->ScaleLevel().DownsampleFactor()) break;
    }
    mLevels.insert(it, level);
  }

 private:
  std::vector<LevelT> mLevels;
};

// GreedyScaleLevelParam (GreedyParam.h:20-44): Iterator, NIterations 50, ScaleLevel.
class GreedyScaleLevelParam {
 public:
  GreedyIteratorParam &Iterator() { return mIter; }
  const GreedyIteratorParam &Iterator() const { return mIter; }
  unsigned int &NIterations() { return mNIter; }
  const unsigned int &NIterations() const { return mNIter; }
  ScaleLevelParam &ScaleLevel() { return mLevel; }
  const ScaleLevelParam &ScaleLevel() const { return mLevel; }

 private:
  GreedyIteratorParam mIter;
  unsigned int mNIter = 50;
  ScaleLevelParam mLevel;
};

// GreedyParam (GreedyParam.h:51-74): the levels plus AutoStepReduce (false).
class GreedyParam : public MultiscaleParamBase<GreedyScaleLevelParam> {
 public:
  bool &AutoStepReduce() { return mAutoStepReduce; }
  const bool &AutoStepReduce() const { return mAutoStepReduce; }

 private:
  bool mAutoStepReduce = false;
};

// LDMMIteratorParam (LDMMIteratorParam.h:20-64): Kernel, Sigma 5, StepSize 0.005,
// UseAdaptiveStepSize true, MaxPert 0.1.
class LDMMIteratorParam {
 public:
  KernelParam &Kernel() { return mKernel; }
  const KernelParam &Kernel() const { return mKernel; }
  Real &Sigma() { return mSigma; }
  Real &StepSize() { return mStep; }
  bool &UseAdaptiveStepSize() { return mAdaptive; }
  Real &MaxPert() { return mMaxPert; }
  const Real &Sigma() const { return mSigma; }
  const Real &StepSize() const { return mStep; }
  const bool &UseAdaptiveStepSize() const { return mAdaptive; }
  const Real &MaxPert() const { return mMaxPert; }

 private:
  KernelParam mKernel;
  Real mSigma = 5.f, mStep = 0.005f, mMaxPert = 0.1f;
  bool mAdaptive = true;
};

// LDMMScaleLevelParam (LDMMParam.h:22-46): Iterator, NIterations 50, ScaleLevel.
class LDMMScaleLevelParam {
 public:
  LDMMIteratorParam &Iterator() { return mIter; }
  const LDMMIteratorParam &Iterator() const { return mIter; }
  unsigned int &NIterations() { return mNIter; }
  const unsigned int &NIterations() const { return mNIter; }
  ScaleLevelParam &ScaleLevel() { return mLevel; }
  const ScaleLevelParam &ScaleLevel() const { return mLevel; }

 private:
  LDMMIteratorParam mIter;
  unsigned int mNIter = 50;
  ScaleLevelParam mLevel;
};

// LDMMParam (LDMMParam.h:50-82): the levels plus NTimeSteps (10), AutoStepReduce.
class LDMMParam : public MultiscaleParamBase<LDMMScaleLevelParam> {
 public:
  unsigned int &NTimeSteps() { return mNT; }
  const unsigned int &NTimeSteps() const { return mNT; }
  bool &AutoStepReduce() { return mAutoStepReduce; }
  const bool &AutoStepReduce() const { return mAutoStepReduce; }

 private:
  unsigned int mNT = 10;
  bool mAutoStepReduce = false;
};

// ------------------------------------------------------------- scale levels
// MultiscaleManager: level f has size orig / f (integer division per axis) and
// spacing orig * f (MultiscaleManager::AddScaleLevel, MultiscaleManager.cxx:104-127),
// levels kept largest factor first.
class MultiscaleManager {
 public:
  MultiscaleManager(const SizeType &origSize, const SpacingType &origSpacing,
                    const OriginType &origin = OriginType(0, 0, 0))
      : mOrigSize(origSize), mOrigSpacing(origSpacing), mOrigin(origin) {}
  template <class LevelT>
  MultiscaleManager(const SizeType &origSize, const SpacingType &origSpacing,
                    const OriginType &origin, const MultiscaleParamBase<LevelT> &param)
      : MultiscaleManager(origSize, origSpacing, origin) {
    for (unsigned i = 0; i < param.GetNumberOfScaleLevels(); ++i)
      AddScaleLevel(param.GetScaleLevel(i).ScaleLevel().DownsampleFactor());
  }
  virtual ~MultiscaleManager() {}
  void SetScaleVectorFields(bool scale) { mScaleVectorFields = scale; }
  bool GetScaleVectorFields() const { return mScaleVectorFields; }
  void AddScaleLevel(unsigned int f) {
    if (f < 1) throw AtlasWerksException(__FILE__, __LINE__, "downsample factor must be >= 1");
    Level l;
    l.factor = f;
    l.size = SizeType(mOrigSize.x / f, mOrigSize.y / f, mOrigSize.z / f);
    l.spacing = SpacingType(mOrigSpacing.x * f, mOrigSpacing.y * f, mOrigSpacing.z * f);
    auto it = mLevels.begin();
    for (; it != mLevels.end(); ++it) {
      if (it->factor == f)
        throw AtlasWerksException(__FILE__, __LINE__,
                                  "Error, cannot add multiple scale levels with the same downsample factor");
      if (it->factor < f) break;
    }
    mLevels.insert(it, l);
  }
  unsigned int NumberOfScaleLevels() const { return (unsigned int)mLevels.size(); }
  virtual void SetScaleLevel(unsigned int level) {
    if (level >= mLevels.size()) throw AtlasWerksException(__FILE__, __LINE__, "invalid scale level");
    mCur = (int)level;
  }
  int CurScaleLevel() const { return mCur; }
  bool InitialScaleLevel() const { return mCur == (int)mInitial; }
  bool FinalScaleLevel() const { return mCur == (int)NumberOfScaleLevels() - 1; }
  SizeType CurScaleSize() const { return mLevels.at(mCur).size; }
  SpacingType CurScaleSpacing() const { return mLevels.at(mCur).spacing; }
  unsigned int CurDownsampleFactor() const { return mLevels.at(mCur).factor; }
  SizeType ScaleLevelSize(unsigned int level) const { return mLevels.at(level).size; }
  const SizeType &OrigSize() const { return mOrigSize; }
  const OriginType &Origin() const { return mOrigin; }
  unsigned int GetInitialScaleLevel() const { return mInitial; }
  void SetInitialScaleLevel(unsigned int init) { mInitial = init; }

 private:
  struct Level {
    unsigned int factor = 1;
    SizeType size;
    SpacingType spacing;
  };
  SizeType mOrigSize;
  SpacingType mOrigSpacing;
  OriginType mOrigin;
  std::vector<Level> mLevels;
  int mCur = 0;
  unsigned int mInitial = 0;
  bool mScaleVectorFields = false;
};

// ------------------------------------------------------------ energy records
// Energy (Common/Energy.h): the total energy of one iteration.
class Energy {
 public:
  Energy() {}
  explicit Energy(Real e) : mEnergy(e) {}
  void SetEnergy(Real e) { mEnergy = e; }
  Real GetEnergy() const { return mEnergy; }

 private:
  Real mEnergy = 0;
};

// EnergyHistory (Common/EnergyHistory.h): iteration energies with their scale
// level and iteration, energy-increase events and step-size events in order.
class EnergyHistory {
 public:
  enum EventType { ITERATION, ENERGY_INCREASE, STEP_SIZE };
  struct Event {
    EventType type;
    unsigned int scale, iter;
    double value;  // the energy or the step size
  };
  void AddIteration(unsigned int scale, unsigned int iter, double e) {
    mEvents.push_back({ITERATION, scale, iter, e});
  }
  void AddEnergyIncrease() { mEvents.push_back({ENERGY_INCREASE, 0, 0, 0.0}); }
  void AddStepSize(double s) { mEvents.push_back({STEP_SIZE, 0, 0, s}); }
  const std::vector<Event> &Events() const { return mEvents; }
  bool HasEnergy() const { return LastEnergyIndex() >= 0; }
  double LastEnergy() const { return mEvents.at(LastEnergyIndex()).value; }
  // the last energy minus the one before it (0 with fewer than two)
  double LastEnergyChange() const {
    int last = -1, prev = -1;
    for (int i = (int)mEvents.size() - 1; i >= 0; --i)
      if (mEvents[i].type == ITERATION) {
        if (last < 0) last = i;
        else { prev = i; break; }
      }
    return prev < 0 ? 0.0 : mEvents[last].value - mEvents[prev].value;
  }

 private:
  int LastEnergyIndex() const {
    for (int i = (int)mEvents.size() - 1; i >= 0; --i)
      if (mEvents[i].type == ITERATION) return i;
    return -1;
  }
  std::vector<Event> mEvents;
};

// ---------------------------------------------------------------- interfaces
// ScaleLevelDataInfo / DeformationDataInterface / DeformationIteratorInterface
// (Common/DeformationIteratorInterface.h, DeformationIteratorInterface.cxx:17-101).
class ScaleLevelDataInfo {
 public:
  ScaleLevelDataInfo(SizeType size, OriginType origin, SpacingType spacing)
      : mImSize(size), mImOrigin(origin), mImSpacing(spacing), mNVox(size.productOfElements()),
        mCurSize(size), mCurSpacing(spacing), mCurVox(size.productOfElements()) {}
  virtual ~ScaleLevelDataInfo() {}
  virtual void SetScaleLevel(const MultiscaleManager &scaleManager) {
    mCurScaleLevel = (unsigned int)scaleManager.CurScaleLevel();
    mCurSize = scaleManager.CurScaleSize();
    mCurSpacing = scaleManager.CurScaleSpacing();
    mCurVox = mCurSize.productOfElements();
  }

 protected:
  SizeType mImSize;
  OriginType mImOrigin;
  SpacingType mImSpacing;
  unsigned int mNVox;
  unsigned int mCurScaleLevel = 0;
  SizeType mCurSize;
  SpacingType mCurSpacing;
  unsigned int mCurVox;
};

class DeformationDataInterface : public ScaleLevelDataInfo {
 public:
  DeformationDataInterface(const RealImage &imTemplate)
      : ScaleLevelDataInfo(imTemplate.getSize(), imTemplate.getOrigin(), imTemplate.getSpacing()) {}
  virtual ~DeformationDataInterface() {}
  virtual const RealImage &I0() = 0;
  virtual const RealImage &I1() = 0;
  virtual void GetI0At1(RealImage &iDef) = 0;
  virtual void GetI1At0(RealImage &iDef) = 0;
  virtual EnergyHistory &GetEnergyHistory() { return mEnergyHistory; }
  // DeformationDataInterface::AddEnergy: an IterationEvent at (scale, iteration)
  virtual void AddEnergy(const Energy &e) {
    mEnergyHistory.AddIteration(mCurScaleLevel, mCurIter, e.GetEnergy());
  }
  virtual bool HasEnergy() { return mEnergyHistory.HasEnergy(); }
  const std::string &GetName() { return mName; }
  void SetName(const std::string &name) { mName = name; }
  unsigned int GetCurIter() { return mCurIter; }
  void SetCurIter(unsigned int iter) { mCurIter = iter; }
  Real StepSize() { return mStepSize; }
  void StepSize(Real step) {
    mStepSize = step;
    mEnergyHistory.AddStepSize(step);
  }

 protected:
  std::string mName;
  Real mStepSize = 0;
  unsigned int mCurIter = 0;
  EnergyHistory mEnergyHistory;
};

class DeformationIteratorInterface : public ScaleLevelDataInfo {
 public:
  DeformationIteratorInterface(SizeType size, OriginType origin, SpacingType spacing)
      : ScaleLevelDataInfo(size, origin, spacing) {}
  virtual ~DeformationIteratorInterface() {}
  virtual void Iterate(DeformationDataInterface &deformationData) = 0;
};

// ------------------------------------------------------------------ host data
// GreedyDeformationData (GreedyDeformationData.h:27-110): I0 (moving), I1
// (static), the hfields Def1To0 / Def0To1 on the host, and AddEnergy's checks
// (GreedyDeformationData.cxx:162-184).  The GPU subclass keeps the working copies
// on the device and does the scale-level resampling there.
class GreedyDeformationData : public DeformationDataInterface {
 public:
  GreedyDeformationData(const RealImage *I0, const RealImage *I1, const GreedyParam &param)
      : DeformationDataInterface(*I0), mParam(param), mI0Orig(I0), mI1Orig(I1), mI0Ptr(I0),
        mI1Ptr(I1), mDef1To0(new VectorField()) {}
  virtual ~GreedyDeformationData() {
    delete mDef1To0;
    delete mDef0To1;
  }
  GreedyDeformationData(const GreedyDeformationData &) = delete;
  GreedyDeformationData &operator=(const GreedyDeformationData &) = delete;
  const RealImage &I0() override { return *mI0Ptr; }
  const RealImage &I1() override { return *mI1Ptr; }
  const RealImage &I0Orig() { return *mI0Orig; }
  const RealImage &I1Orig() { return *mI1Orig; }
  virtual VectorField &Def0To1() { return *mDef0To1; }
  virtual VectorField &Def1To0() { return *mDef1To0; }
  virtual void InitializeWarp() {}
  virtual void FinalizeWarp() {}
  void ScaleI0(bool scale) { mScaleI0 = scale; }
  bool ScaleI0() { return mScaleI0; }
  void ScaleI1(bool scale) { mScaleI1 = scale; }
  bool ScaleI1() { return mScaleI1; }
  bool ComputeInverseHField() { return mComputeInverseHField; }
  void ComputeInverseHField(bool computeInv) {
    mComputeInverseHField = computeInv;
    if (computeInv && !mDef0To1) mDef0To1 = new VectorField();
  }
  void AddEnergy(const Energy &e) override {
    DeformationDataInterface::AddEnergy(e);
    const double change = mEnergyHistory.LastEnergyChange();
    if (change > 0) {
      mEnergyHistory.AddEnergyIncrease();
      if (mParam.AutoStepReduce()) StepSize(StepSize() / 2.0f);
    }
  }

 protected:
  const GreedyParam &mParam;
  const RealImage *mI0Orig, *mI1Orig;
  const RealImage *mI0Ptr, *mI1Ptr;
  VectorField *mDef1To0;
  VectorField *mDef0To1 = nullptr;
  bool mComputeInverseHField = false;
  bool mScaleI0 = true, mScaleI1 = true;
};

// LDMMDeformationData (LDMMDeformationData.h:30-140): I0, I1 and the T velocity
// fields on the host (LDMMParam::NTimeSteps), the alphas of the last
// computeAlphasOnly iteration.
class LDMMDeformationData : public DeformationDataInterface {
 public:
  LDMMDeformationData(const RealImage *I0, const RealImage *I1, const LDMMParam &param)
      : DeformationDataInterface(*I0), mParam(param), mNTimeSteps(param.NTimeSteps()),
        mI0Orig(I0), mI1Orig(I1), mI0Ptr(I0), mI1Ptr(I1) {
    for (unsigned t = 0; t < mNTimeSteps; ++t) mV.push_back(new VectorField());
  }
  virtual ~LDMMDeformationData() {
    for (VectorField *v : mV) delete v;
  }
  LDMMDeformationData(const LDMMDeformationData &) = delete;
  LDMMDeformationData &operator=(const LDMMDeformationData &) = delete;
  const RealImage &I0() override { return *mI0Ptr; }
  const RealImage &I1() override { return *mI1Ptr; }
  virtual std::vector<VectorField *> &v() { return mV; }
  virtual VectorField &v(int i) { return *mV[i]; }
  virtual void InitializeWarp() {}
  virtual void FinalizeWarp() {}
  unsigned int NTimeSteps() const { return mNTimeSteps; }

 protected:
  const LDMMParam &mParam;
  unsigned int mNTimeSteps;
  const RealImage *mI0Orig, *mI1Orig;
  const RealImage *mI0Ptr, *mI1Ptr;
  std::vector<VectorField *> mV;
  bool mScaleI0 = true, mScaleI1 = true;
};

#endif  // AWK_USE_REFERENCE_TYPES

#endif
