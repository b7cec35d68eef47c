// DiffOperGPU.h -- FluidKernelFFT behind the reference's GPU kernel-operator API:
// KernelInterfaceGPU (Algorithms/Common/KernelInterfaceGPU.h:20-43) and
// DiffOperGPU (Algorithms/Common/DiffOperGPU.h:27-135), over awk_fluid_kernel_*.
// Unlike the reference GPU operator, LPow != 1, the eigen table path and the
// divergence-free projection are supported rather than throwing
// (DiffOperGPU.cxx:127-161).
#ifndef AWK_DIFFOPERGPU_H
#define AWK_DIFFOPERGPU_H

#include "awk_drivers.h"
#include "cplVector3DArray.h"

// KernelInterfaceGPU (KernelInterfaceGPU.h:20-43): SetSize with the KernelParam
// the iterators hold; ApplyOperator / ApplyInverseOperator overwrite their input.
class KernelInterfaceGPU {
 public:
  virtual ~KernelInterfaceGPU() {}
  virtual void SetSize(const SizeType &logicalSize, const SpacingType &spacing,
                       const KernelParam &params) = 0;
  virtual void ApplyOperator(cplVector3DArray &dV) = 0;         // f = L v, in place
  virtual void ApplyInverseOperator(cplVector3DArray &dF) = 0;  // v = L^-1 f, in place
};

class DiffOperGPU : public KernelInterfaceGPU {
 public:
  DiffOperGPU() {}
  ~DiffOperGPU() override { release(); }
  DiffOperGPU(const DiffOperGPU &) = delete;
  DiffOperGPU &operator=(const DiffOperGPU &) = delete;

  // DiffOperGPU::SetSize (DiffOperGPU.cxx:32-60): the KernelParam must hold a
  // DiffOper (KernelFactory pairs a MultiGaussian param with another operator).
  void SetSize(const SizeType &logicalSize, const SpacingType &spacing,
               const KernelParam &params) override {
    if (!params.IsDiffOperParam())
      throw AtlasWerksException(__FILE__, __LINE__, "DiffOperGPU::SetSize: not a DiffOper parameter");
    SetSize(logicalSize, spacing, *params.AsDiffOperParam());
  }
  void SetSize(const SizeType &logicalSize, const SpacingType &spacing, const DiffOperParam &p) {
    release();
    mSize = logicalSize;
    mSpacing = spacing;
    copyParams(p);
    AWK_CHECK(awk_fluid_kernel_create(&mH, (int)logicalSize.x, (int)logicalSize.y,
                                      (int)logicalSize.z, spacing.x, spacing.y, spacing.z, mAlpha,
                                      mBeta, mGamma, mLPow, flags()));
  }
  void SetParams(const DiffOperParam &p) {
    copyParams(p);
    push();
  }
  void SetAlpha(Real a) { mAlpha = a; push(); }
  Real GetAlpha() const { return mAlpha; }
  void SetBeta(Real b) { mBeta = b; push(); }
  Real GetBeta() const { return mBeta; }
  void SetGamma(Real g) { mGamma = g; push(); }
  Real GetGamma() const { return mGamma; }
  void SetLPow(Real p) { mLPow = p; push(); }
  Real GetLPow() const { return mLPow; }
  void SetUseEigenLUT(bool b) { mEigen = b; push(); }
  bool GetUseEigenLUT() const { return mEigen; }
  void SetDivergenceFree(bool b) { mDivFree = b; push(); }
  bool GetDivergenceFree() const { return mDivFree; }

  void ApplyOperator(float *dVx, float *dVy, float *dVz, awkStream_t s = nullptr) {
    check();
    AWK_CHECK(awk_fluid_kernel_apply(mH, dVx, dVy, dVz, 0, s));
  }
  void ApplyOperator(cplVector3DArray &dV) override { ApplyOperator(dV.x, dV.y, dV.z); }
  void ApplyInverseOperator(float *dFx, float *dFy, float *dFz, awkStream_t s = nullptr) {
    check();
    AWK_CHECK(awk_fluid_kernel_apply(mH, dFx, dFy, dFz, 1, s));
  }
  void ApplyInverseOperator(cplVector3DArray &dF) override { ApplyInverseOperator(dF.x, dF.y, dF.z); }

 private:
  void copyParams(const DiffOperParam &p) {
    mAlpha = p.Alpha();
    mBeta = p.Beta();
    mGamma = p.Gamma();
    mLPow = p.LPow();
    mEigen = p.UseEigenLUT();
    mDivFree = p.DivergenceFree();
  }
  int flags() const { return (mEigen ? AWK_FK_EIGEN : 0) | (mDivFree ? AWK_FK_DIVFREE : 0); }
  void push() {
    if (mH) AWK_CHECK(awk_fluid_kernel_set_params(mH, mAlpha, mBeta, mGamma, mLPow, flags()));
  }
  void check() const {
    if (!mH) throw AtlasWerksException(__FILE__, __LINE__, "DiffOperGPU: SetSize must be called first");
  }
  void release() {
    if (mH) awk_fluid_kernel_destroy(mH);
    mH = nullptr;
  }
  awk_fk_t mH = nullptr;
  SizeType mSize;
  SpacingType mSpacing;
  Real mAlpha = 0.01f, mBeta = 0.01f, mGamma = 0.001f, mLPow = 1.f;
  bool mEigen = false, mDivFree = false;
};

// FluidKernelFFT is the name this operator goes by in the registration drivers.
typedef DiffOperGPU FluidKernelFFT;

#endif
