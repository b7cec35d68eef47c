// DiffOperGPU.h -- FluidKernelFFT behind the reference's GPU kernel-operator API:
// KernelInterfaceGPU (Algorithms/Common/KernelInterfaceGPU.h:20-43) and
// DiffOperGPU (Algorithms/Common/DiffOperGPU.h:27-135), over awk_fluid_kernel_*.
// Unlike the reference GPU operator, LPow != 1, the eigen table path and the
// divergence-free projection are supported rather than throwing.
#ifndef AWK_DIFFOPERGPU_H
#define AWK_DIFFOPERGPU_H

#include "cplVector3DArray.h"

// DiffOperParam (Algorithms/Common/DiffOperParam.h:24-86): defaults alpha =
// 0.01, beta = 0.01, gamma = 0.001, LPow = 1.
struct DiffOperParam {
  float alpha = 0.01f, beta = 0.01f, gamma = 0.001f, lpow = 1.0f;
  bool useEigenLUT = false, divergenceFree = false;
  float &Alpha() { return alpha; }
  float &Beta() { return beta; }
  float &Gamma() { return gamma; }
  float &LPow() { return lpow; }
  bool &UseEigenLUT() { return useEigenLUT; }
  bool &DivergenceFree() { return divergenceFree; }
};

class KernelInterfaceGPU {
 public:
  virtual ~KernelInterfaceGPU() {}
  virtual void SetSize(const SizeType &logicalSize, const SpacingType &spacing,
                       const DiffOperParam &params) = 0;
  virtual void ApplyOperator(cplVector3DArray &dV) = 0;         // f = L v, in place
  virtual void ApplyInverseOperator(cplVector3DArray &dF) = 0;  // v = L^-1 f, in place
};

class DiffOperGPU : public KernelInterfaceGPU {
 public:
  DiffOperGPU() {}
  ~DiffOperGPU() override { release(); }
  DiffOperGPU(const DiffOperGPU &) = delete;
  DiffOperGPU &operator=(const DiffOperGPU &) = delete;

  void SetSize(const SizeType &logicalSize, const SpacingType &spacing,
               const DiffOperParam &params) override {
    release();
    mSize = logicalSize;
    mSpacing = spacing;
    mParam = params;
    AWK_CHECK(awk_fluid_kernel_create(&mH, (int)logicalSize.x, (int)logicalSize.y,
                                      (int)logicalSize.z, spacing.x, spacing.y, spacing.z,
                                      mParam.alpha, mParam.beta, mParam.gamma, mParam.lpow, flags()));
  }
  void SetParams(const DiffOperParam &p) {
    mParam = p;
    push();
  }
  DiffOperParam GetParams() const { return mParam; }
  void SetAlpha(float a) { mParam.alpha = a; push(); }
  float GetAlpha() const { return mParam.alpha; }
  void SetBeta(float b) { mParam.beta = b; push(); }
  float GetBeta() const { return mParam.beta; }
  void SetGamma(float g) { mParam.gamma = g; push(); }
  float GetGamma() const { return mParam.gamma; }
  void SetLPow(float p) { mParam.lpow = p; push(); }
  float GetLPow() const { return mParam.lpow; }
  void SetUseEigenLUT(bool b) { mParam.useEigenLUT = b; push(); }
  bool GetUseEigenLUT() const { return mParam.useEigenLUT; }
  void SetDivergenceFree(bool b) { mParam.divergenceFree = b; push(); }
  bool GetDivergenceFree() const { return mParam.divergenceFree; }

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
  int flags() const {
    return (mParam.useEigenLUT ? AWK_FK_EIGEN : 0) | (mParam.divergenceFree ? AWK_FK_DIVFREE : 0);
  }
  void push() {
    if (mH) AWK_CHECK(awk_fluid_kernel_set_params(mH, mParam.alpha, mParam.beta, mParam.gamma, mParam.lpow, flags()));
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
  DiffOperParam mParam;
};

// FluidKernelFFT is the name this operator goes by in the registration drivers.
typedef DiffOperGPU FluidKernelFFT;

#endif
