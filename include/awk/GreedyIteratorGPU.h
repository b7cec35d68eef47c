// GreedyIteratorGPU.h -- one greedy-fluid registration step behind the reference's
// iterator API (Algorithms/Greedy/GreedyIteratorGPU.h; DeformationIteratorInterface.h:84-92):
// SetScaleLevel, UpdateStepSizeNextIteration, Iterate(deformation data).
#ifndef AWK_GREEDYITERATORGPU_H
#define AWK_GREEDYITERATORGPU_H

#include <vector>

#include "cplVector3DArray.h"
#include "DiffOperGPU.h"

// GreedyIteratorParam (Algorithms/Greedy/GreedyIteratorParam.h:20-40)
struct GreedyIteratorParam {
  float maxPert = 0.1f;
  DiffOperParam kernel;
  float &MaxPert() { return maxPert; }
  DiffOperParam &Kernel() { return kernel; }
};

// Per-registration device state (GreedyDeformationDataGPU): I0, I1, Def1To0 and
// optionally Def0To1, all caller-owned device memory, updated in place.
class GreedyDeformationDataGPU {
 public:
  GreedyDeformationDataGPU(float *dI0, float *dI1, cplVector3DArray def1to0,
                           cplVector3DArray *def0to1 = nullptr)
      : mI0(dI0), mI1(dI1), mH(def1to0), mHasInv(def0to1 != nullptr) {
    if (def0to1) mHinv = *def0to1;
  }
  float *dI0() { return mI0; }
  float *dI1() { return mI1; }
  cplVector3DArray &dDef1To0() { return mH; }
  cplVector3DArray &dDef0To1() { return mHinv; }
  bool ComputeInverseHField() const { return mHasInv; }
  // energies of the iterations run so far (host copy, synchronises)
  std::vector<double> Energies() {
    std::vector<double> e(mEnergyCount);
    if (mEnergyCount) (void)hipMemcpy(e.data(), mEnergy, mEnergyCount * sizeof(double), hipMemcpyDeviceToHost);
    return e;
  }
  ~GreedyDeformationDataGPU() {
    if (mEnergy) (void)hipFree(mEnergy);
  }

 private:
  friend class GreedyIteratorGPU;
  double *energySlot() {
    if (mEnergyCount == mEnergyCap) {
      double *grown = nullptr;
      const unsigned cap = mEnergyCap ? 2 * mEnergyCap : 256;
      if (hipMalloc(&grown, cap * sizeof(double)) != hipSuccess)
        throw AtlasWerksException(__FILE__, __LINE__, "energy buffer allocation failed");
      if (mEnergy) {
        (void)hipMemcpy(grown, mEnergy, mEnergyCount * sizeof(double), hipMemcpyDeviceToDevice);
        (void)hipFree(mEnergy);
      }
      mEnergy = grown;
      mEnergyCap = cap;
    }
    return mEnergy + mEnergyCount++;
  }
  float *mI0, *mI1;
  cplVector3DArray mH, mHinv;
  bool mHasInv;
  double *mEnergy = nullptr;
  unsigned mEnergyCount = 0, mEnergyCap = 0;
};

class GreedyIteratorGPU {
 public:
  GreedyIteratorGPU(const SizeType &size, const SpacingType &spacing, const GreedyIteratorParam &p) {
    SetScaleLevel(size, spacing, p);
  }
  ~GreedyIteratorGPU() { release(); }
  GreedyIteratorGPU(const GreedyIteratorGPU &) = delete;
  GreedyIteratorGPU &operator=(const GreedyIteratorGPU &) = delete;

  void SetScaleLevel(const SizeType &size, const SpacingType &spacing, const GreedyIteratorParam &p) {
    release();
    mSize = size;
    const DiffOperParam &k = p.kernel;
    AWK_CHECK(awk_greedy_create(&mG, (int)size.x, (int)size.y, (int)size.z, spacing.x, spacing.y,
                                spacing.z, k.alpha, k.beta, k.gamma, p.maxPert));
  }
  void UpdateStepSizeNextIteration() { AWK_CHECK(awk_greedy_update_step_next(mG)); }
  float StepSize() {
    float s = 0.f;
    AWK_CHECK(awk_greedy_get_step(mG, &s, nullptr));
    return s;
  }
  void SetAutoStepReduce(bool on) { AWK_CHECK(awk_greedy_set_auto_step_reduce(mG, on ? 1 : 0)); }

  // In place on the data's hfields (the update goes through the library's
  // scratch, as GreedyIteratorGPU::updateDeformation copies back,
  // GreedyIteratorGPU.cxx:72-88).
  void Iterate(GreedyDeformationDataGPU &d, awkStream_t s = nullptr) {
    cplVector3DArray &h = d.dDef1To0();
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
                                 inv ? iout : nullptr, d.energySlot(), s));
  }

 private:
  void release() {
    if (mG) awk_greedy_destroy(mG);
    mG = nullptr;
  }
  awk_greedy_t mG = nullptr;
  SizeType mSize;
};

#endif
