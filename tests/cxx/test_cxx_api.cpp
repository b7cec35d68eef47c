// C++ drop-in API check against the shim headers in include/awk/ (built by
// __graft_entry__.build(); run by tests/test_gpu_cxx.py on the GPU with the
// golden-data directory as its argument).
//  * KernelFactory::NewGPUKernel + KernelInterfaceGPU::SetSize(KernelParam), then
//    the reference's DiffOperRoutineTests.cxx:161-196 GPU round trip
//    L(L^-1 g) = g (Sum |d|^2 < 0.1 there; far smaller here);
//  * CUDAUtilities::CopyVectorFieldToDevice / FromDevice round trip (exact);
//  * cudaHField3DUtilsRoutineTests.cxx:61-72: <apply(I1,h), I2> = <I1, splat(I2,h)>;
//  * a GreedyWarp-shaped template over GreedyIteratorGPU, following the calls of
//    GreedyWarp.txx:19-125 (typedefs DeformationDataType / ParamType, the
//    MultiscaleManager levels, SetScaleLevel, UpdateStepSizeNextIteration,
//    Iterate(DeformationDataInterface&), InitializeWarp / FinalizeWarp), on the
//    reference's GreedyWarp regression (BullseyeTestBlur00 -> 01, MaxPert 0.3,
//    levels 2 -> 1, 50 + 50 iterations): the deformed image equals the
//    reference's GreedyBullseyeTestBlur00_to_BullseyeTestBlur01DefImage.mha to
//    max |d| <= 2e-4 and relative L2 <= 1e-4;
//  * an LDMMWarp-shaped template over LDMMIteratorGPU (LDMMWarp.txx:69-125): the
//    step-size call, decreasing total energy, alphas; LDMMAtlasGPU builds a
//    finite mean.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "awk/CUDAUtilities.h"
#include "awk/DiffOperGPU.h"
#include "awk/GreedyIteratorGPU.h"
#include "awk/KernelFactory.h"
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

// ---- the two test-data formats the reference's regression files use
static bool readNrrd(const std::string &path, RealImage &im) {
  std::ifstream f(path);
  if (!f) return false;
  std::string line, data;
  unsigned sz[3] = {0, 0, 0};
  float sp[3] = {1, 1, 1}, org[3] = {0, 0, 0};
  while (std::getline(f, line)) {
    if (line.rfind("sizes:", 0) == 0) std::sscanf(line.c_str(), "sizes: %u %u %u", sz, sz + 1, sz + 2);
    else if (line.rfind("space directions:", 0) == 0)
      std::sscanf(line.c_str(), "space directions: (%f,%*f,%*f) (%*f,%f,%*f) (%*f,%*f,%f)", sp, sp + 1, sp + 2);
    else if (line.rfind("space origin:", 0) == 0)
      std::sscanf(line.c_str(), "space origin: (%f,%f,%f)", org, org + 1, org + 2);
    else if (line.rfind("data file:", 0) == 0) data = line.substr(line.find(':') + 2);
  }
  const std::string dir = path.substr(0, path.find_last_of('/') + 1);
  std::ifstream r(dir + data, std::ios::binary);
  if (!r) return false;
  im = RealImage(SizeType(sz[0], sz[1], sz[2]), OriginType(org[0], org[1], org[2]),
                 SpacingType(sp[0], sp[1], sp[2]));
  r.read(reinterpret_cast<char *>(im.getDataPointer()), (std::streamsize)im.getNumElements() * 4);
  return (bool)r;
}
static bool readMeta(const std::string &path, std::vector<float> &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::string line;
  unsigned n = 0;
  while (std::getline(f, line)) {
    unsigned a, b, c;
    if (std::sscanf(line.c_str(), "DimSize = %u %u %u", &a, &b, &c) == 3) n = a * b * c;
    if (line.rfind("ElementDataFile", 0) == 0) break;
  }
  out.resize(n);
  f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)n * 4);
  return n > 0 && (bool)f;
}

// ---- GreedyWarp<GreedyIteratorT>: the driver's calls, as GreedyWarp.txx:19-125
template <class GreedyIteratorT>
class GreedyWarpShaped {
 public:
  typedef typename GreedyIteratorT::DeformationDataType DeformationDataType;
  typedef typename GreedyIteratorT::ParamType IteratorParamType;

  GreedyWarpShaped(const RealImage *I0, const RealImage *I1, const GreedyParam &param,
                   bool writeInvDefField)
      : mImSize(I0->getSize()), mImOrigin(I0->getOrigin()), mImSpacing(I0->getSpacing()),
        mNScaleLevels(param.GetNumberOfScaleLevels()), mDeformationData(I0, I1, param),
        mParam(param) {
    mScaleManager = new MultiscaleManager(mImSize, mImSpacing, mImOrigin, mParam);
    mScaleManager->SetScaleVectorFields(true);
    mDeformationData.ComputeInverseHField(writeInvDefField);
    mIterator = new GreedyIteratorT(mImSize, mImOrigin, mImSpacing, true);
  }
  ~GreedyWarpShaped() {
    delete mIterator;
    delete mScaleManager;
  }
  void SetScaleLevel(const MultiscaleManager &scaleManager) {
    const int scale = scaleManager.CurScaleLevel();
    const IteratorParamType &p = mParam.GetScaleLevel(scale).Iterator();
    mIterator->SetScaleLevel(scaleManager, p);
    mDeformationData.SetScaleLevel(scaleManager);
  }
  void RunWarp() {
    mDeformationData.InitializeWarp();
    for (unsigned scaleLevel = 0; scaleLevel < mNScaleLevels; scaleLevel++) {
      mScaleManager->SetScaleLevel(scaleLevel);
      this->SetScaleLevel(*mScaleManager);
      const unsigned nIters = mParam.GetScaleLevel(scaleLevel).NIterations();
      for (unsigned iter = 0; iter < nIters; iter++) {
        mDeformationData.SetCurIter(iter);
        if (iter == 0) mIterator->UpdateStepSizeNextIteration();
        mIterator->Iterate(mDeformationData);
        if (iter == 0) std::printf("  scale %u step size %g\n", scaleLevel, mDeformationData.StepSize());
      }
    }
    mDeformationData.FinalizeWarp();
  }
  DeformationDataType &Data() { return mDeformationData; }

 private:
  SizeType mImSize;
  OriginType mImOrigin;
  SpacingType mImSpacing;
  unsigned mNScaleLevels;
  DeformationDataType mDeformationData;
  const GreedyParam &mParam;
  MultiscaleManager *mScaleManager = nullptr;
  GreedyIteratorT *mIterator = nullptr;
};

// ---- LDMMWarp<LDMMIteratorT>: LDMMWarp.txx:69-125 with the adaptive step
template <class LDMMIteratorT>
class LDMMWarpShaped {
 public:
  typedef typename LDMMIteratorT::DeformationDataType DeformationDataType;
  typedef typename LDMMIteratorT::ParamType IteratorParamType;
  LDMMWarpShaped(const RealImage *I0, const RealImage *I1, const LDMMParam &param)
      : mDeformationData(I0, I1, param), mParam(param), mNT(param.NTimeSteps()) {
    SizeType size = I0->getSize();
    OriginType origin = I0->getOrigin();
    SpacingType spacing = I0->getSpacing();
    mScaleManager = new MultiscaleManager(size, spacing, origin, mParam);
    mScaleManager->SetScaleVectorFields(false);
    mIterator = new LDMMIteratorT(size, origin, spacing, mNT, true);
  }
  ~LDMMWarpShaped() {
    delete mIterator;
    delete mScaleManager;
  }
  void RunWarp() {
    mDeformationData.InitializeWarp();
    for (unsigned l = 0; l < mParam.GetNumberOfScaleLevels(); ++l) {
      mScaleManager->SetScaleLevel(l);
      const IteratorParamType &p = mParam.GetScaleLevel(l).Iterator();
      mIterator->SetScaleLevel(*mScaleManager, p);
      mDeformationData.SetScaleLevel(*mScaleManager);
      if (p.UseAdaptiveStepSize()) {
        mIterator->UpdateStepSizeNextIteration();
        mIterator->Iterate(mDeformationData);
      }
      for (unsigned iter = 0; iter < mParam.GetScaleLevel(l).NIterations(); ++iter) {
        mDeformationData.SetCurIter(iter);
        mIterator->Iterate(mDeformationData);
      }
    }
    mDeformationData.FinalizeWarp();
  }
  DeformationDataType &Data() { return mDeformationData; }
  LDMMIteratorT &Iterator() { return *mIterator; }

 private:
  DeformationDataType mDeformationData;
  const LDMMParam &mParam;
  unsigned int mNT;
  MultiscaleManager *mScaleManager = nullptr;
  LDMMIteratorT *mIterator = nullptr;
};

static std::vector<double> historyEnergies(EnergyHistory &h) {
  std::vector<double> e;
  for (const auto &ev : h.Events())
    if (ev.type == EnergyHistory::ITERATION) e.push_back(ev.value);
  return e;
}

int main(int argc, char **argv) {
  const std::string gold = argc > 1 ? argv[1] : "tests/golden";
  const int nx = 32, ny = 31, nz = 36;  // the reference's odd-sized bullseye grid
  const unsigned N = (unsigned)nx * ny * nz;
  const Vector3Di n(nx, ny, nz);
  // ---- KernelFactory + KernelInterfaceGPU::SetSize(KernelParam); L(L^-1 g) == g
  {
    cplVector3DArray d;
    allocateDeviceVector3DArray(d, N);
    std::vector<float> g[3];
    for (int c = 0; c < 3; ++c) {
      g[c] = wavy(nx, ny, nz, c);
      (void)hipMemcpy(d.elementArray(c), g[c].data(), N * 4, hipMemcpyHostToDevice);
    }
    KernelParam kp;
    KernelInterfaceGPU *op = KernelFactory::NewGPUKernel(kp);
    EXPECT(dynamic_cast<DiffOperGPU *>(op) != nullptr, "NewGPUKernel did not give a DiffOperGPU");
    op->SetSize(SizeType(nx, ny, nz), SpacingType(1.f, 1.25f, 1.f), kp);
    op->ApplyInverseOperator(d);
    op->ApplyOperator(d);
    delete op;
    double sq = 0;
    std::vector<float> back(N);
    for (int c = 0; c < 3; ++c) {
      (void)hipMemcpy(back.data(), d.elementArray(c), N * 4, hipMemcpyDeviceToHost);
      for (unsigned i = 0; i < N; ++i) sq += (back[i] - g[c][i]) * (double)(back[i] - g[c][i]);
    }
    EXPECT(sq < 1e-6, "L(L^-1 g) != g: sum sq %g", sq);
    std::printf("round trip sum sq diff %g\n", sq);
    bool threw = false;
    try {
      DiffOperGPU bad;
      bad.ApplyOperator(d);
    } catch (const AtlasWerksException &) {
      threw = true;
    }
    EXPECT(threw, "unsized operator did not throw");
    // ---- host AoS <-> device SoA
    VectorField hv(SizeType(nx, ny, nz)), back2(SizeType(nx, ny, nz));
    for (unsigned i = 0; i < N; ++i) hv(i) = Vector3D<float>(0.5f * i, -1.f * i, 3.f + i);
    CUDAUtilities::CopyVectorFieldToDevice(hv, d, true);
    std::vector<float> py(N);
    (void)hipMemcpy(py.data(), d.y, N * 4, hipMemcpyDeviceToHost);
    bool ok = true;
    for (unsigned i = 0; i < N; ++i) ok = ok && py[i] == -1.f * i;
    CUDAUtilities::CopyVectorFieldFromDevice(d, back2);
    for (unsigned i = 0; i < N; ++i)
      ok = ok && back2(i).x == hv(i).x && back2(i).y == hv(i).y && back2(i).z == hv(i).z;
    EXPECT(ok, "CopyVectorField round trip differs");
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
    cplSplat3DH(o2, d2, h, n, nullptr, AWK_SPLAT_FIXED64);
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
    // the reference's S2p20 splat (the shim default) agrees to its quantisation
    cplSplat3DH(o1, d2, h, n);
    (void)hipMemcpy(a.data(), o1, N * 4, hipMemcpyDeviceToHost);
    double md = 0;
    for (unsigned i = 0; i < N; ++i) md = std::fmax(md, std::fabs(a[i] - b[i]));
    EXPECT(md < 1e-4, "S2p20 splat differs from the 64-bit one by %g", md);
    // ---- LDMMAtlasGPU on two wavy subjects
    {
      const unsigned T = 4;
      std::vector<cplVector3DArray> v(T), v2(T);
      for (unsigned t = 0; t < T; ++t) {
        allocateDeviceVector3DArray(v[t], N);
        allocateDeviceVector3DArray(v2[t], N);
        (void)hipMemset(v[t].x, 0, 3 * (size_t)N * 4);
        (void)hipMemset(v2[t].x, 0, 3 * (size_t)N * 4);
      }
      LDMMIteratorParam lp;
      lp.Sigma() = 0.5f;
      lp.MaxPert() = 0.2f;
      LDMMAtlasGPU atlas(SizeType(nx, ny, nz), SpacingType(1, 1, 1), T, 2, lp);
      atlas.SetSubject(0, d1, 0.5f, v);
      atlas.SetSubject(1, d2, 0.5f, v2);
      atlas.Iterate(3);
      atlas.GetMean(o1);
      std::vector<float> al(N);
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
    freeDeviceVector3DArray(h);
    (void)hipFree(d1); (void)hipFree(d2); (void)hipFree(o1); (void)hipFree(o2);
  }
  // ---- GreedyWarp<GreedyIteratorGPU> on the reference's GreedyWarp regression
  {
    RealImage I0, I1;
    std::vector<float> want;
    const bool have = readNrrd(gold + "/BullseyeTestBlur00.nhdr", I0) &&
                      readNrrd(gold + "/BullseyeTestBlur01.nhdr", I1) &&
                      readMeta(gold + "/greedy_warp/GreedyBullseyeTestBlur00_to_BullseyeTestBlur01DefImage.mha", want);
    EXPECT(have, "golden data not found under %s", gold.c_str());
    if (have) {
      GreedyParam gp;
      for (unsigned f : {1u, 2u}) {  // the parameter file lists factor 1 first
        GreedyScaleLevelParam lvl;
        lvl.ScaleLevel().DownsampleFactor() = f;
        lvl.NIterations() = 50;
        lvl.Iterator().MaxPert() = 0.3f;
        gp.AddScaleLevel(lvl);
      }
      EXPECT(gp.GetScaleLevel(0).ScaleLevel().DownsampleFactor() == 2, "levels not ordered coarse first");
      GreedyWarpShaped<GreedyIteratorGPU> warp(&I0, &I1, gp, true);
      warp.RunWarp();
      // the final hfield on the host (FinalizeWarp copied it), back to the device
      VectorField &h = warp.Data().Def1To0();
      EXPECT(h.getSize() == I0.getSize(), "final hfield has the wrong size");
      cplVector3DArray dh;
      allocateDeviceVector3DArray(dh, N);
      CUDAUtilities::CopyVectorFieldToDevice(h, dh);
      float *dI0, *dOut;
      (void)hipMalloc(&dI0, N * 4);
      (void)hipMalloc(&dOut, N * 4);
      (void)hipMemcpy(dI0, I0.getDataPointer(), N * 4, hipMemcpyHostToDevice);
      cudaHField3DUtils::apply(dOut, dI0, dh, n);
      std::vector<float> got(N);
      (void)hipMemcpy(got.data(), dOut, N * 4, hipMemcpyDeviceToHost);
      double num = 0, den = 0, mx = 0;
      for (unsigned i = 0; i < N; ++i) {
        const double dd = (double)got[i] - want[i];
        num += dd * dd;
        den += (double)want[i] * want[i];
        mx = std::fmax(mx, std::fabs(dd));
      }
      const double rel = std::sqrt(num / den);
      std::printf("GreedyWarp<GreedyIteratorGPU>: rel L2 %.3g, max |d| %.3g vs the reference output\n", rel, mx);
      EXPECT(rel <= 1e-4 && mx <= 2e-4, "GreedyWarp-shaped run differs from the reference output");
      // the energy history: 100 iteration events, a step-size event per level
      std::vector<double> e = historyEnergies(warp.Data().GetEnergyHistory());
      unsigned steps = 0;
      for (const auto &ev : warp.Data().GetEnergyHistory().Events()) steps += ev.type == EnergyHistory::STEP_SIZE;
      EXPECT(e.size() == 100 && steps == 2, "energy history has %zu energies, %u step events", e.size(), steps);
      EXPECT(e.size() == 100 && e[49] < e[0] && e[99] < e[50], "energies did not decrease within the levels");
      freeDeviceVector3DArray(dh);
      (void)hipFree(dI0);
      (void)hipFree(dOut);
    }
  }
  // ---- LDMMWarp<LDMMIteratorGPU>: adaptive step, decreasing energy, alphas
  {
    RealImage I0(SizeType(nx, ny, nz)), I1(SizeType(nx, ny, nz));
    std::vector<float> a = wavy(nx, ny, nz, 3), b = wavy(nx, ny, nz, 4);
    std::memcpy(I0.getDataPointer(), a.data(), N * 4);
    std::memcpy(I1.getDataPointer(), b.data(), N * 4);
    LDMMParam lp;
    lp.NTimeSteps() = 4;
    for (unsigned f : {2u, 1u}) {
      LDMMScaleLevelParam lvl;
      lvl.ScaleLevel().DownsampleFactor() = f;
      lvl.NIterations() = 4;
      lvl.Iterator().Sigma() = 0.5f;
      lvl.Iterator().MaxPert() = 0.2f;
      lp.AddScaleLevel(lvl);
    }
    LDMMWarpShaped<LDMMIteratorGPU> w(&I0, &I1, lp);
    w.RunWarp();
    std::vector<double> e = historyEnergies(w.Data().GetEnergyHistory());
    EXPECT(e.size() == 8, "LDMM energy count %zu", e.size());
    if (e.size() == 8)
      EXPECT(e[3] < e[0] && e[7] < e[4], "LDMM energy did not decrease: %g %g / %g %g", e[0], e[3], e[4], e[7]);
    std::vector<VectorField *> &vv = w.Data().v();
    double vs = 0;
    for (VectorField *f : vv)
      for (unsigned i = 0; i < f->getNumElements(); ++i) vs += std::fabs((*f)(i).x);
    EXPECT(vv.size() == 4 && std::isfinite(vs) && vs > 0, "LDMM velocity fields not set");
    std::printf("LDMMWarp<LDMMIteratorGPU> energies %g -> %g (level 1), %g -> %g (level 2)\n",
                e.size() == 8 ? e[0] : 0.0, e.size() == 8 ? e[3] : 0.0, e.size() == 8 ? e[4] : 0.0,
                e.size() == 8 ? e[7] : 0.0);
  }
  std::printf(fails ? "CXX API TEST FAILED (%d)\n" : "CXX API TEST OK\n", fails);
  return fails ? 1 : 0;
}
