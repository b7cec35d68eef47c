// C++ drop-in API check: the reference's own GPU tests restated against the shim
// headers in include/awk/ (built by __graft_entry__.build(); run by
// tests/test_gpu_cxx.py on the GPU).
//  * DiffOperRoutineTests.cxx:161-196 -- GPU L(L^-1 g) = g, Sum |d|^2 < 0.1
//  * cudaHField3DUtilsRoutineTests.cxx:61-72 -- <apply(I1,h), I2> = <I1, splat(I2,h)>
//  * GreedyIteratorGPU drives a few iterations with decreasing energy.
//  * LDMMIteratorGPU (IterateOrig) drives a few iterations with decreasing total
//    energy; computeAlphasOnly fills GetAlpha; LDMMAtlasGPU builds a finite mean.
#include <cmath>
#include <cstdio>
#include <vector>

#include "awk/DiffOperGPU.h"
#include "awk/GreedyIteratorGPU.h"
#include "awk/LDMMIteratorGPU.h"
#include "awk/cudaHField3DUtils.h"
#include "awk/cudaSplat.h"

static int fails = 0;
#define EXPECT(cond, ...)                  \
  do {                                     \
    if (!(cond)) {                         \
      std::printf("FAIL: " __VA_ARGS__);   \
      std::printf("\n");                   \
      ++fails;                             \
    }                                      \
  } while (0)

static std::vector<float> wavy(int nx, int ny, int nz, int c) {
  std::vector<float> v((size_t)nx * ny * nz);
  size_t i = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x, ++i)
        v[i] = std::sin(0.3f * x + 0.7f * c) * std::cos(0.2f * y) + 0.5f * std::sin(0.4f * z + c);
  return v;
}

int main() {
  const int nx = 32, ny = 31, nz = 36;  // the reference's odd-sized bullseye grid
  const unsigned N = (unsigned)nx * ny * nz;
  const Vector3Di n(nx, ny, nz);
  // ---- L(L^-1 g) == g
  {
    cplVector3DArray d;
    allocateDeviceVector3DArray(d, N);
    std::vector<float> g[3];
    for (int c = 0; c < 3; ++c) {
      g[c] = wavy(nx, ny, nz, c);
      (void)hipMemcpy(d.elementArray(c), g[c].data(), N * 4, hipMemcpyHostToDevice);
    }
    DiffOperGPU op;
    DiffOperParam p;
    op.SetSize(SizeType(nx, ny, nz), SpacingType(1.f, 1.25f, 1.f), p);
    op.ApplyInverseOperator(d);
    op.ApplyOperator(d);
    double sq = 0;
    std::vector<float> back(N);
    for (int c = 0; c < 3; ++c) {
      (void)hipMemcpy(back.data(), d.elementArray(c), N * 4, hipMemcpyDeviceToHost);
      for (unsigned i = 0; i < N; ++i) sq += (back[i] - g[c][i]) * (double)(back[i] - g[c][i]);
    }
    EXPECT(sq < 1e-6, "L(L^-1 g) != g: sum sq %g", sq);
    std::printf("round trip sum sq diff %g\n", sq);
    // configuration errors surface as AtlasWerksException
    bool threw = false;
    try {
      DiffOperGPU bad;
      bad.ApplyOperator(d);
    } catch (const AtlasWerksException &) {
      threw = true;
    }
    EXPECT(threw, "unsized operator did not throw");
    freeDeviceVector3DArray(d);
  }
  // ---- adjoint identity of apply and splat on a wavy hfield
  {
    cplVector3DArray h;
    allocateDeviceVector3DArray(h, N);
    cudaHField3DUtils::setToIdentity(h, n);
    std::vector<float> hx(N), i1 = wavy(nx, ny, nz, 5), i2 = wavy(nx, ny, nz, 9);
    for (int c = 0; c < 3; ++c) {
      (void)hipMemcpy(hx.data(), h.elementArray(c), N * 4, hipMemcpyDeviceToHost);
      std::vector<float> w = wavy(nx, ny, nz, 11 + c);
      for (unsigned i = 0; i < N; ++i) hx[i] += 1.5f * w[i];
      (void)hipMemcpy(h.elementArray(c), hx.data(), N * 4, hipMemcpyHostToDevice);
    }
    float *d1, *d2, *o1, *o2;
    (void)hipMalloc(&d1, N * 4); (void)hipMalloc(&d2, N * 4);
    (void)hipMalloc(&o1, N * 4); (void)hipMalloc(&o2, N * 4);
    (void)hipMemcpy(d1, i1.data(), N * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(d2, i2.data(), N * 4, hipMemcpyHostToDevice);
    cudaHField3DUtils::apply(o1, d1, h, n);
    cplSplat3DH(o2, d2, h, n);
    std::vector<float> a(N), b(N);
    (void)hipMemcpy(a.data(), o1, N * 4, hipMemcpyDeviceToHost);
    (void)hipMemcpy(b.data(), o2, N * 4, hipMemcpyDeviceToHost);
    double lhs = 0, rhs = 0;
    for (unsigned i = 0; i < N; ++i) {
      lhs += (double)a[i] * i2[i];
      rhs += (double)i1[i] * b[i];
    }
    EXPECT(std::fabs(lhs - rhs) < 1e-5 * std::fabs(lhs), "adjoint identity %g vs %g", lhs, rhs);
    std::printf("adjoint <apply(I1,h),I2> = %.9g, <I1,splat(I2,h)> = %.9g\n", lhs, rhs);
    // ---- greedy iterations: energy decreases
    cudaHField3DUtils::setToIdentity(h, n);
    (void)hipMemcpy(d2, b.data(), N * 4, hipMemcpyHostToDevice);  // target: something splatted
    cudaHField3DUtils::apply(d2, d1, h, n);
    cplVector3DArray hw;
    allocateDeviceVector3DArray(hw, N);
    cudaHField3DUtils::setToIdentity(hw, n);
    for (int c = 0; c < 3; ++c) {  // target = I1 warped by a smooth field
      (void)hipMemcpy(hx.data(), hw.elementArray(c), N * 4, hipMemcpyDeviceToHost);
      std::vector<float> w = wavy(nx, ny, nz, 21 + c);
      for (unsigned i = 0; i < N; ++i) hx[i] += 0.8f * w[i];
      (void)hipMemcpy(hw.elementArray(c), hx.data(), N * 4, hipMemcpyHostToDevice);
    }
    cudaHField3DUtils::apply(d2, d1, hw, n);
    GreedyIteratorParam gp;
    GreedyIteratorGPU it(SizeType(nx, ny, nz), SpacingType(1, 1, 1), gp);
    GreedyDeformationDataGPU data(d1, d2, h);
    for (int k = 0; k < 10; ++k) {
      if (k == 0) it.UpdateStepSizeNextIteration();
      it.Iterate(data);
    }
    std::vector<double> e = data.Energies();
    EXPECT(e.size() == 10 && e.back() < e.front(), "energy did not decrease");
    std::printf("greedy energy %g -> %g, step %g\n", e.front(), e.back(), it.StepSize());
    // ---- LDMM (IterateOrig): total energy decreases; alphas; a two-subject atlas
    {
      const unsigned T = 4;
      std::vector<cplVector3DArray> v(T), v2(T);
      for (unsigned t = 0; t < T; ++t) {
        allocateDeviceVector3DArray(v[t], N);
        allocateDeviceVector3DArray(v2[t], N);
        for (int c = 0; c < 3; ++c) {
          (void)hipMemset(v[t].elementArray(c), 0, N * 4);
          (void)hipMemset(v2[t].elementArray(c), 0, N * 4);
        }
      }
      LDMMIteratorParam lp;
      lp.Sigma() = 0.5f;
      lp.MaxPert() = 0.2f;
      LDMMIteratorGPU li(SizeType(nx, ny, nz), SpacingType(1, 1, 1), T, lp);
      LDMMDeformationDataGPU ld(d1, d2, v);
      li.UpdateStepSizeNextIteration();
      li.Iterate(ld);  // step-size call (energies recorded from here on)
      for (int k = 0; k < 5; ++k) li.Iterate(ld);
      std::vector<double> le = ld.Energies();  // 6 x (image, vector, total)
      EXPECT(le.size() == 18 && le[17] < le[5], "LDMM total energy did not decrease");
      std::printf("LDMM total energy %g -> %g, step %g\n", le[5], le[17], li.StepSize());
      li.Iterate(ld, true);
      std::vector<float> al(N);
      (void)hipMemcpy(al.data(), li.GetAlpha(0), N * 4, hipMemcpyDeviceToHost);
      bool finite = true;
      for (float q : al) finite = finite && std::isfinite(q);
      EXPECT(finite, "LDMM alphas not finite");
      LDMMAtlasGPU atlas(SizeType(nx, ny, nz), SpacingType(1, 1, 1), T, 2, lp);
      atlas.SetSubject(0, d1, 0.5f, v);
      atlas.SetSubject(1, d2, 0.5f, v2);
      atlas.Iterate(3);
      atlas.GetMean(o1);
      (void)hipMemcpy(al.data(), o1, N * 4, hipMemcpyDeviceToHost);
      double msum = 0.0;
      for (float q : al) msum += q;
      EXPECT(std::isfinite(msum), "LDMM atlas mean not finite");
      std::printf("LDMM atlas mean sum %g\n", msum);
      for (unsigned t = 0; t < T; ++t) {
        freeDeviceVector3DArray(v[t]);
        freeDeviceVector3DArray(v2[t]);
      }
    }
    freeDeviceVector3DArray(hw);
    freeDeviceVector3DArray(h);
    (void)hipFree(d1); (void)hipFree(d2); (void)hipFree(o1); (void)hipFree(o2);
  }
  std::printf(fails ? "CXX API TEST FAILED (%d)\n" : "CXX API TEST OK\n", fails);
  return fails ? 1 : 0;
}
