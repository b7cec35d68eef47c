// GreedyIteratorGPU.h -- the greedy-fluid iteration behind the reference's driver
// API, so that GreedyWarp<GreedyIteratorGPU> and the atlas threads bind it at
// compile time (Greedy/GreedyWarp.h:34-35, Common/DeformationIteratorInterface.h:
// 84-92):
//   GreedyDeformationDataGPU (Greedy/GreedyDeformationDataGPU.h:20-70): host
//     images and fields of GreedyDeformationData plus their device copies;
//     InitializeWarp / SetScaleLevel / FinalizeWarp as the reference calls them.
//     The scale-level work (hfield upsampling, image downsampling) runs on the
//     device, where the reference copies the fields to the host, resamples there
//     and copies back (GreedyDeformationDataGPU.cxx:75-116).
//   GreedyIteratorGPU (Greedy/GreedyIteratorGPU.h:29-70): DeformationDataType /
//     ParamType, SetScaleLevel(MultiscaleManager, GreedyIteratorParam),
//     UpdateStepSizeNextIteration, Iterate(DeformationDataInterface&).
// Energies stay on the device and reach the EnergyHistory (AddEnergy, with its
// increase events) when the data is synchronised: at the step-size iteration, at
// scale-level changes, at FinalizeWarp and on GetEnergyHistory().
#ifndef AWK_GREEDYITERATORGPU_H
#define AWK_GREEDYITERATORGPU_H

#include <utility>
#include <vector>

#include "CUDAUtilities.h"
#include "DiffOperGPU.h"
#include "awk_drivers.h"

// Device energies of one deformation (one or three doubles per iteration) and
// the iteration number each was recorded at.
class AwkDeviceEnergies {
 public:
  explicit AwkDeviceEnergies(unsigned per) : mPer(per) {}
  ~AwkDeviceEnergies() {
    if (mDev) (void)hipFree(mDev);
  }
  AwkDeviceEnergies(const AwkDeviceEnergies &) = delete;
  AwkDeviceEnergies &operator=(const AwkDeviceEnergies &) = delete;
  double *Next(unsigned iter) {
    if (mCount == mCap) {
      const unsigned cap = mCap ? 2 * mCap : 256;
      double *grown = nullptr;
      if (hipMalloc(&grown, (size_t)mPer * cap * sizeof(double)) != hipSuccess)
        throw AtlasWerksException(__FILE__, __LINE__, "energy buffer allocation failed");
      if (mDev) {
        (void)hipDeviceSynchronize();  // iterations on any stream may still write here
        (void)hipMemcpy(grown, mDev, (size_t)mPer * mCount * sizeof(double), hipMemcpyDeviceToDevice);
        (void)hipFree(mDev);
      }
      mDev = grown;
      mCap = cap;
    }
    mIter.push_back(iter);
    return mDev + (size_t)mPer * mCount++;
  }
  // (iteration, energies) recorded since the last call; waits for the device
  std::vector<std::pair<unsigned, std::vector<double> > > Take() {
    std::vector<std::pair<unsigned, std::vector<double> > > out;
    if (mTaken == mCount) return out;
    std::vector<double> h((size_t)mPer * (mCount - mTaken));
    (void)hipDeviceSynchronize();
    if (hipMemcpy(h.data(), mDev + (size_t)mPer * mTaken, h.size() * sizeof(double),
                  hipMemcpyDeviceToHost) != hipSuccess)
      throw AtlasWerksException(__FILE__, __LINE__, "energy read-back failed");
    for (unsigned k = mTaken; k < mCount; ++k)
      out.emplace_back(mIter[k], std::vector<double>(h.begin() + (size_t)mPer * (k - mTaken),
                                                     h.begin() + (size_t)mPer * (k - mTaken + 1)));
    mTaken = mCount;
    return out;
  }

 private:
  unsigned mPer;
  double *mDev = nullptr;
  unsigned mCount = 0, mCap = 0, mTaken = 0;
  std::vector<unsigned> mIter;
};

class GreedyDeformationDataGPU : public GreedyDeformationData {
 public:
  GreedyDeformationDataGPU(const RealImage *I0, const RealImage *I1, const GreedyParam &param)
      : GreedyDeformationData(I0, I1, param), mEnergies(1) {}
  ~GreedyDeformationDataGPU() override { FreeDeviceData(); }

  // GreedyDeformationDataGPU::InitializeWarp / FinalizeWarp
  // (GreedyDeformationDataGPU.cxx:56-72)
  void InitializeWarp() override { InitializeDeviceData(); }
  void FinalizeWarp() override {
    SyncEnergies();
    syncFields();
    FreeDeviceData();
  }

  // Device buffers at the full size (every scale level fits them); the original
  // images copied once.
  void InitializeDeviceData() {
    if (mDeviceDataInitialized) return;
    const size_t n = (size_t)mI0Orig->getNumElements();
    mNFull = n;
    alloc(&mdI0Orig, n);
    alloc(&mdI1Orig, n);
    alloc(&mdI0Scaled, n);
    alloc(&mdI1Scaled, n);
    alloc(&mdH, 3 * n);
    alloc(&mdHs, 3 * n);
    if (mComputeInverseHField) {
      alloc(&mdHi, 3 * n);
      alloc(&mdHis, 3 * n);
    }
    copyIn(mdI0Orig, mI0Orig->getDataPointer(), n);
    copyIn(mdI1Orig, mI1Orig->getDataPointer(), n);
    mdI0Ptr = mdI0Orig;
    mdI1Ptr = mdI1Orig;
    mDeviceDataInitialized = true;
  }
  void FreeDeviceData() {
    float **bufs[] = {&mdI0Orig, &mdI1Orig, &mdI0Scaled, &mdI1Scaled, &mdH, &mdHs, &mdHi, &mdHis};
    for (float **b : bufs) {
      if (*b) (void)hipFree(*b);
      *b = nullptr;
    }
    mdI0Ptr = mdI1Ptr = nullptr;
    mdDef1To0 = mdDef0To1 = cplVector3DArray();
    mDeviceDataInitialized = false;
  }

  // GreedyDeformationData::SetScaleLevel (GreedyDeformationData.cxx:186-247) on the
  // device: the initial level starts from the identity; later levels carry the
  // hfield over as a voxel displacement resampled with CLAMP (and rescaled when
  // the manager scales vector fields); images are the originals at the final
  // level, Gaussian-downsampled from them before.
  void SetScaleLevel(const MultiscaleManager &scaleManager) override {
    InitializeDeviceData();
    SyncEnergies();
    const SizeType prev = mCurSize;
    const bool initial = scaleManager.InitialScaleLevel() || !mHaveLevel;
    DeformationDataInterface::SetScaleLevel(scaleManager);
    const SizeType cur = mCurSize;
    if (initial) {
      mdDef1To0 = planes(mdH, cur);
      AWK_CHECK(awk_hfield_set_to_identity(mdDef1To0.x, mdDef1To0.y, mdDef1To0.z, (int)cur.x,
                                           (int)cur.y, (int)cur.z, nullptr));
      if (mComputeInverseHField) {
        mdDef0To1 = planes(mdHi, cur);
        AWK_CHECK(awk_hfield_set_to_identity(mdDef0To1.x, mdDef0To1.y, mdDef0To1.z, (int)cur.x,
                                             (int)cur.y, (int)cur.z, nullptr));
      }
    } else {
      const int rescale = scaleManager.GetScaleVectorFields() ? 1 : 0;
      upsample(mdH, mdHs, mdDef1To0, prev, cur, rescale);
      if (mComputeInverseHField) upsample(mdHi, mdHis, mdDef0To1, prev, cur, rescale);
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

  // host views of the fields, brought up to date from the device
  // (GreedyDeformationDataGPU::Def1To0 / Def0To1)
  VectorField &Def1To0() override {
    syncFields();
    return *mDef1To0;
  }
  VectorField &Def0To1() override {
    if (!mComputeInverseHField)
      throw AtlasWerksException(__FILE__, __LINE__, "Error, inverse HField not computed");
    syncFields();
    return *mDef0To1;
  }
  // I0 at Def1To0 (GreedyDeformationData::GetI0At1, GreedyDeformationData.cxx:56-64)
  void GetI0At1(RealImage &iDef) override { deform(iDef, mdI0Ptr, mdDef1To0, false); }
  // GreedyDeformationData::GetI1At0 (GreedyDeformationData.cxx:66-80): I1 at
  // Def0To1 when the inverse is tracked, else I0 at the zeroth-order inverse of
  // Def1To0, as the reference does
  void GetI1At0(RealImage &iDef) override {
    if (mComputeInverseHField) deform(iDef, mdI1Ptr, mdDef0To1, false);
    else deform(iDef, mdI0Ptr, mdDef1To0, true);
  }
  EnergyHistory &GetEnergyHistory() override {
    SyncEnergies();
    return mEnergyHistory;
  }

  float *dI0() { return mdI0Ptr; }
  float *dI1() { return mdI1Ptr; }
  cplVector3DArray &dDef1To0() { return mdDef1To0; }
  cplVector3DArray &dDef0To1() { return mdDef0To1; }
  bool AutoStepReduceParam() { return const_cast<GreedyParam &>(mParam).AutoStepReduce(); }

  // the iterator's energy slot for the current iteration
  double *NextEnergySlot() { return mEnergies.Next(mCurIter); }
  // AddEnergy for every energy recorded on the device since the last sync, at
  // the iteration it belongs to (IterationEvent, EnergyIncreaseEvent, the
  // AutoStepReduce halving of the host step)
  void SyncEnergies() {
    const unsigned keep = mCurIter;
    for (auto &rec : mEnergies.Take()) {
      SetCurIter(rec.first);
      Energy e;
      e.SetEnergy((Real)rec.second[0]);
      AddEnergy(e);
    }
    SetCurIter(keep);
  }

 private:
  static void alloc(float **p, size_t n) {
    if (hipMalloc(p, n * sizeof(float)) != hipSuccess)
      throw AtlasWerksException(__FILE__, __LINE__, "device allocation failed");
  }
  static void copyIn(float *d, const float *h, size_t n) {
    if (hipMemcpy(d, h, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      throw AtlasWerksException(__FILE__, __LINE__, "image upload failed");
  }
  static cplVector3DArray planes(float *base, const SizeType &s) {
    const unsigned n = s.productOfElements();
    return cplVector3DArray(base, base + n, base + 2 * (size_t)n, n);
  }
  void upsample(float *&buf, float *&spare, cplVector3DArray &f, const SizeType &from,
                const SizeType &to, int rescale) {
    AWK_CHECK(awk_hfield_h_to_velocity(f.x, f.y, f.z, (int)from.x, (int)from.y, (int)from.z, 1.f,
                                       1.f, 1.f, nullptr));
    cplVector3DArray out = planes(spare, to);
    AWK_CHECK(awk_hfield_resample(out.x, out.y, out.z, (int)to.x, (int)to.y, (int)to.z, f.x, f.y,
                                  f.z, (int)from.x, (int)from.y, (int)from.z, AWK_BG_CLAMP, rescale,
                                  nullptr));
    AWK_CHECK(awk_hfield_velocity_to_h(out.x, out.y, out.z, (int)to.x, (int)to.y, (int)to.z, 1.f,
                                       1.f, 1.f, nullptr));
    std::swap(buf, spare);
    f = out;
  }
  void syncFields() {
    if (!mDeviceDataInitialized || !mHaveLevel) return;
    mDef1To0->resize(mCurSize);
    CUDAUtilities::CopyVectorFieldFromDevice(mdDef1To0, *mDef1To0, true);
    if (mComputeInverseHField) {
      mDef0To1->resize(mCurSize);
      CUDAUtilities::CopyVectorFieldFromDevice(mdDef0To1, *mDef0To1, true);
    }
  }
  void deform(RealImage &iDef, const float *img, const cplVector3DArray &h, bool inverse) {
    iDef.resize(mCurSize);
    iDef.setOrigin(mImOrigin);
    iDef.setSpacing(mCurSpacing);
    const size_t n = mCurVox;
    float *out = nullptr;
    float *hi = nullptr;
    alloc(&out, n);
    const float *hx = h.x, *hy = h.y, *hz = h.z;
    awk_status st = AWK_OK;
    if (inverse) {
      alloc(&hi, 3 * n);
      st = awk_hfield_inverse_zeroth_order(hi, hi + n, hi + 2 * n, h.x, h.y, h.z, (int)mCurSize.x,
                                           (int)mCurSize.y, (int)mCurSize.z, nullptr);
      hx = hi;
      hy = hi + n;
      hz = hi + 2 * n;
    }
    if (st == AWK_OK)
      st = awk_hfield_apply(out, img, hx, hy, hz, (int)mCurSize.x, (int)mCurSize.y, (int)mCurSize.z,
                            AWK_BG_VAL, 0.f, nullptr);
    hipError_t e = st == AWK_OK ? hipMemcpy(iDef.getDataPointer(), out, n * sizeof(float),
                                            hipMemcpyDeviceToHost)
                                : hipSuccess;
    (void)hipFree(out);
    if (hi) (void)hipFree(hi);
    AWK_CHECK(st);
    if (e != hipSuccess) throw AtlasWerksException(__FILE__, __LINE__, hipGetErrorString(e));
  }

  bool mDeviceDataInitialized = false, mHaveLevel = false;
  size_t mNFull = 0;
  float *mdI0Orig = nullptr, *mdI1Orig = nullptr, *mdI0Scaled = nullptr, *mdI1Scaled = nullptr;
  float *mdI0Ptr = nullptr, *mdI1Ptr = nullptr;
  float *mdH = nullptr, *mdHs = nullptr, *mdHi = nullptr, *mdHis = nullptr;
  cplVector3DArray mdDef1To0, mdDef0To1;
  AwkDeviceEnergies mEnergies;
};

class GreedyIteratorGPU : public DeformationIteratorInterface {
 public:
  typedef GreedyDeformationDataGPU DeformationDataType;
  typedef GreedyIteratorParam ParamType;

  // GreedyIteratorGPU(size, origin, spacing, debug) (GreedyIteratorGPU.cxx:24-42)
  GreedyIteratorGPU(const SizeType &size, const OriginType &origin, const SpacingType &spacing,
                    bool debug = true)
      : DeformationIteratorInterface(size, origin, spacing), mDebug(debug) {}
  ~GreedyIteratorGPU() override { release(); }
  GreedyIteratorGPU(const GreedyIteratorGPU &) = delete;
  GreedyIteratorGPU &operator=(const GreedyIteratorGPU &) = delete;

  // GreedyIteratorGPU::SetScaleLevel (GreedyIteratorGPU.cxx:44-66): the level's
  // grid, its operator (a DiffOper KernelParam) and MaxPert
  using ScaleLevelDataInfo::SetScaleLevel;
  void SetScaleLevel(const MultiscaleManager &scaleManager, const GreedyIteratorParam &param) {
    ScaleLevelDataInfo::SetScaleLevel(scaleManager);
    const KernelParam &kp = param.Kernel();
    if (!kp.IsDiffOperParam())
      throw AtlasWerksException(__FILE__, __LINE__, "GreedyIteratorGPU: only the DiffOper kernel is built");
    const DiffOperParam &k = *kp.AsDiffOperParam();
    release();
    AWK_CHECK(awk_greedy_create(&mG, (int)mCurSize.x, (int)mCurSize.y, (int)mCurSize.z,
                                mCurSpacing.x, mCurSpacing.y, mCurSpacing.z, k.Alpha(), k.Beta(),
                                k.Gamma(), param.MaxPert()));
    mUpdateStep = false;
  }
  void SetDebug(bool debug) { mDebug = debug; }
  bool GetDebug() { return mDebug; }
  void UpdateStepSizeNextIteration() {
    check();
    mUpdateStep = true;
    AWK_CHECK(awk_greedy_update_step_next(mG));
  }
  // the stream later iterations run on (the reference uses the null stream)
  void SetStream(awkStream_t s) { mStream = s; }
  // sticky numerical flags since the last read (AWK_FLAG_*); waits for the stream
  unsigned GetFlags(bool clear = true) {
    check();
    unsigned f = 0;
    awk_status st = awk_greedy_get_flags(mG, &f, clear ? 1 : 0, mStream);
    if (st != AWK_OK && st != AWK_ERR_NUMERIC) AWK_CHECK(st);
    return f;
  }

  // GreedyIteratorGPU::Iterate (GreedyIteratorGPU.cxx:92-135), in place on the
  // data's device hfields; a step-size iteration hands the step to the data
  // (DeformationDataInterface::StepSize) after its energy.
  void Iterate(DeformationDataInterface &deformationData) override {
    check();
    GreedyDeformationDataGPU &d = dynamic_cast<GreedyDeformationDataGPU &>(deformationData);
    cplVector3DArray &h = d.dDef1To0();
    if (h.size() != mCurVox)
      throw AtlasWerksException(__FILE__, __LINE__, "deformation data is not at the iterator's scale level");
    AWK_CHECK(awk_greedy_set_auto_step_reduce(mG, d.AutoStepReduceParam() ? 1 : 0));
    const float *hin[3] = {h.x, h.y, h.z};
    float *hout[3] = {h.x, h.y, h.z};
    const float *iin[3] = {nullptr, nullptr, nullptr};
    float *iout[3] = {nullptr, nullptr, nullptr};
    const bool inv = d.ComputeInverseHField();
    if (inv) {
      cplVector3DArray &hi = d.dDef0To1();
      iin[0] = iout[0] = hi.x;
      iin[1] = iout[1] = hi.y;
      iin[2] = iout[2] = hi.z;
    }
    AWK_CHECK(awk_greedy_iterate(mG, d.dI0(), d.dI1(), hin, hout, inv ? iin : nullptr,
                                 inv ? iout : nullptr, mDebug ? d.NextEnergySlot() : nullptr,
                                 mStream));
    if (mUpdateStep) {
      mUpdateStep = false;
      float s = 0.f;
      AWK_CHECK(awk_greedy_get_step(mG, &s, mStream));
      d.SyncEnergies();
      d.StepSize(s);
    }
  }

 private:
  void check() const {
    if (!mG) throw AtlasWerksException(__FILE__, __LINE__, "GreedyIteratorGPU: SetScaleLevel must be called first");
  }
  void release() {
    if (mG) awk_greedy_destroy(mG);
    mG = nullptr;
  }
  awk_greedy_t mG = nullptr;
  bool mDebug;
  bool mUpdateStep = false;
  awkStream_t mStream = nullptr;
};

#endif
