// awk_ldmm.hip -- one LDMM iteration on gfx950, in both gradient forms of the
// reference (LDMMIteratorGPU.cxx / LDMMIteratorCPU.cxx; mUseOrigGradientCalc,
// true by default, selects IterateOrig).
//
// IterateOrig (LDMMIteratorGPU.cxx:295-441, LDMMIteratorCPU.cxx:379-527):
//   backward  t = T-1..0, one fused gather kernel per t: h <- h(x + v_t/sp),
//             JT_t = I1(h), JacDet_t = JacDet_{t+1}(x + v_t/sp) |D(x sp + v_t)|
//   forward   t = 0..T-1: the Navier solve whose x pass forms grad(J_t) m_t,
//             m_t = (J_t - JT_t) JacDet_t, and whose inverse x pass updates v_t
//             in place; then one fused kernel h <- h(x - v_t/sp), J_{t+1} = I0(h)
//             and m_{t+1}; the last one forms the FP64 SSD partials.
// IterateNew (LDMMIteratorGPU.cxx:159-293, LDMMIteratorCPU.cxx:529-631):
//   forward   h <- h(x - v_t(x)/sp) (PARTIAL_ID) and J_t = I0(h), one fused kernel
//             per t; the last one also forms diff = J_{T-1} - I1 and FP64 SSD partials
//   backward  h <- h - v_t(h)/sp (PARTIAL_ZERO) and the splat of diff to h(x) into an
//             int64 fixed-point accumulator, one fused kernel per t; then the
//             Navier solve whose forward x pass forms grad(J_{t-1}) * alpha_t
//             straight from the accumulator (and clears it), and whose inverse x
//             pass updates v_t = (1 - 2 eps) v_t + (2 eps / sigma^2) u in place
//             (or, on the step-size call, takes max |2 v_t - 2/sigma^2 u|).
// No host synchronisation; energies stay on the device (FP64, fixed order).
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "awk_ldmm.h"
#include "awk_fluid.h"
#include "awk_ops.h"
#include "awk_splat_tile.h"

namespace awk {

constexpr int kLdmmBlock = 256;

__device__ __forceinline__ void block_sum_store(double v, double *out, int lb) {
  __shared__ double red[kLdmmBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kLdmmBlock / 64; ++i) s += red[i];
    out[lb] = s;
  }
}

__global__ void __launch_bounds__(kLdmmBlock)
k_ldmm_identity(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz, Dims d) {
  const int N = d.nx * d.ny * d.nz;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const int x = i % d.nx, r = i / d.nx;
  hx[i] = (float)x;
  hy[i] = (float)(r % d.ny);
  hz[i] = (float)(r / d.ny);
}

// Forward step t: composeHVInv(h, v_t, PARTIAL_ID) (HField3DUtils.txx:457-480) and
// apply(I0, h) (VAL background 0); the last step also diff = J - I1.
__global__ void __launch_bounds__(kLdmmBlock)
k_ldmm_fwd(float *__restrict__ nx_, float *__restrict__ ny_, float *__restrict__ nz_,
           const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
           const float *__restrict__ vx, const float *__restrict__ vy, const float *__restrict__ vz,
           const float *__restrict__ I0, float *__restrict__ J, const float *__restrict__ I1,
           float *__restrict__ diff, double *__restrict__ partials, Dims d, float sx, float sy,
           float sz) {
  const int N = d.nx * d.ny * d.nz;
  const int lb = (int)xcd_block();
  const int i = lb * kLdmmBlock + (int)threadIdx.x;
  double ssd = 0.0;
  if (i < N) {
    const int x = i % d.nx, r = i / d.nx, y = r % d.ny, z = r / d.ny;
    const float px = (float)x - vx[i] / sx, py = (float)y - vy[i] / sy, pz = (float)z - vz[i] / sz;
    const float3 hn = trilerp3g<AWK_BG_PARTIAL_ID>(make_buf(hx, N), make_buf(hy, N), make_buf(hz, N), d,
                                                   px, py, pz);
    nx_[i] = hn.x;
    ny_[i] = hn.y;
    nz_[i] = hn.z;
    const float j = trilerp1g<AWK_BG_VAL>(make_buf(I0, N), d, hn.x, hn.y, hn.z, 0.f);
    J[i] = j;
    if (diff) {
      const float dv = j - I1[i];
      diff[i] = dv;
      ssd = (double)dv * (double)dv;
    }
  }
  if (partials) block_sum_store(ssd, partials, lb);
}

// Backward step t: composeVHInv(v_t, h, PARTIAL_ZERO) (HField3DUtils.txx:400-430)
// and the splat of diff to h(x) (cplSplat3DH, cudaSplatAtomic.cu:185-247) as
// int64 fixed point (2^-32), which is order-independent; tiled through LDS
// windows (awk_splat_tile.h).
__global__ void __launch_bounds__(kSplatTileThreads)
k_ldmm_bwd(float *__restrict__ nx_, float *__restrict__ ny_, float *__restrict__ nz_,
           const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
           const float *__restrict__ vx, const float *__restrict__ vy, const float *__restrict__ vz,
           const float *__restrict__ diff, unsigned long long *__restrict__ acc, Dims d, float sx,
           float sy, float sz) {
  const int N = d.nx * d.ny * d.nz;
  SplatPosLdmm p{hx, hy, hz, make_buf(vx, N), make_buf(vy, N), make_buf(vz, N), d, sx, sy, sz,
                 nx_, ny_, nz_};
  splat_tile<SplatShapeLdmm>(p, diff, acc, d);
}

// composeVHInv(v_t, h, PARTIAL_ZERO) alone, for the sort/scan splat: the same
// positions as the fused tiled kernel's (SplatPosLdmm).
__global__ void __launch_bounds__(kLdmmBlock)
k_ldmm_vhinv(float *__restrict__ nx_, float *__restrict__ ny_, float *__restrict__ nz_,
             const float *__restrict__ hx, const float *__restrict__ hy,
             const float *__restrict__ hz, const float *__restrict__ vx,
             const float *__restrict__ vy, const float *__restrict__ vz, Dims d, float sx, float sy,
             float sz) {
  const int N = d.nx * d.ny * d.nz;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const float a = hx[i], b = hy[i], c = hz[i];
  const float3 o = trilerp3g<AWK_BG_PARTIAL_ZERO>(make_buf(vx, N), make_buf(vy, N), make_buf(vz, N), d,
                                                  a, b, c);
  nx_[i] = o.x / -sx + a;
  ny_[i] = o.y / -sy + b;
  nz_[i] = o.z / -sz + c;
}

__global__ void __launch_bounds__(kLdmmBlock)
k_ldmm_scale(float *__restrict__ out, const float *__restrict__ in, int N, float k) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < N) out[i] = in[i] * k;
}

// alpha_t = splat / sigma^2 (LDMMIteratorGPU.cxx:197-204); clears the accumulator
__global__ void __launch_bounds__(kLdmmBlock)
k_ldmm_alpha(float *__restrict__ alpha, long long *__restrict__ acc, int N, float inv_sig2) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const long long q = acc[i];
  acc[i] = 0;
  alpha[i] = splat_unq(q) * inv_sig2;
}

__global__ void __launch_bounds__(kLdmmBlock) k_ldmm_zero(long long *__restrict__ acc, int N) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < N) acc[i] = 0;
}

// FP64 dot product partials of two 3-plane fields: sum_c a_c . b_c
__global__ void __launch_bounds__(kLdmmBlock)
k_ldmm_dot(const float *__restrict__ a, const float *__restrict__ b0, const float *__restrict__ b1,
           const float *__restrict__ b2, int N, double *__restrict__ partials) {
  const int lb = (int)xcd_block();
  const int i = lb * kLdmmBlock + (int)threadIdx.x;
  double s = 0.0;
  if (i < N) {
    s = (double)a[i] * b0[i];
    s += (double)a[N + i] * b1[i];
    s += (double)a[2 * N + i] * b2[i];
  }
  block_sum_store(s, partials, lb);
}

// energy[0] = Prod(sp) * sum(diff^2) / sigma^2 (LDMMIteratorGPU.cxx:276-283,
// LDMMEnergy.h:42); energy[1] = 0; energy[2] = energy[0].  Fixed order.
__global__ void k_ldmm_image_energy(const double *__restrict__ partials, int n, double scale,
                                    double *__restrict__ energy, unsigned *__restrict__ flags) {
  __shared__ double s[256];
  double a = 0.0;
  a = strided_fold(partials, n, (int)threadIdx.x, 256, a, FoldAdd());
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    energy[0] = scale * s[0];
    energy[1] = 0.0;
    energy[2] = energy[0];
    if (energy[0] != energy[0]) raise_flags(flags, AWK_FLAG_NAN_ENERGY);
  }
}

// energy[1] += Prod(sp) <L v_t, v_t> / T (LDMMIteratorGPU.cxx:250-256,
// LDMMEnergy.cxx:57-64); energy[2] = energy[0] + energy[1].  Fixed order.
__global__ void k_ldmm_vec_energy(const double *__restrict__ partials, int n, double scale,
                                  double *__restrict__ energy, unsigned *__restrict__ flags) {
  __shared__ double s[256];
  double a = 0.0;
  a = strided_fold(partials, n, (int)threadIdx.x, 256, a, FoldAdd());
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    energy[1] += scale * s[0];
    energy[2] = energy[0] + energy[1];
    if (energy[1] != energy[1]) raise_flags(flags, AWK_FLAG_NAN_ENERGY);
  }
}

// ---- IterateOrig (LDMMIteratorGPU.cxx:295-441, LDMMIteratorCPU.cxx:379-527) ----

// Gradient along one axis of f(q) = b(q) + a[q] (computeGradient, wrap = false;
// Array3DUtils.txx:783-901); bm, b0, bp = the identity part at p - 1, p, p + 1.
// r1 = 1 / sp and r2 = 1 / (2 sp) for div_const.
__device__ __forceinline__ float hgrad(const float *__restrict__ a, int i, int p, int n, int stride,
                                       float sp, float r1, float r2, float bm, float b0, float bp) {
  if (n == 1) return 0.f;
  if (p > 0 && p < n - 1) return div_const((bp + a[i + stride]) - (bm + a[i - stride]), 2.f * sp, r2);
  if (p == 0) return div_const((bp + a[i + stride]) - (b0 + a[i]), sp, r1);
  return div_const((b0 + a[i]) - (bm + a[i - stride]), sp, r1);
}

// |D(x sp + v)| at voxel i: HField3DUtils::jacobian (HField3DUtils.txx:2370-2420) of
// velocityToH_US (cudaHField3DUtilsSp.cu:137-153), one row per component.
__device__ __forceinline__ float jac_det_xsp_v(const float *__restrict__ vx,
                                               const float *__restrict__ vy,
                                               const float *__restrict__ vz, int i, int x, int y,
                                               int z, Dims d, const Spacing3 &S) {
  const int sxy = d.nx * d.ny;
  const float *vc[3] = {vx, vy, vz};
  const int pos[3] = {x, y, z}, n[3] = {d.nx, d.ny, d.nz}, stride[3] = {1, d.nx, sxy};
  const float *sp = S.s;
  float g[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float own = (float)pos[c] * sp[c];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      float bm = own, b0 = own, bp = own;
      if (c == e) {
        bm = (float)(pos[c] - 1) * sp[c];
        bp = (float)(pos[c] + 1) * sp[c];
      }
      g[c][e] = hgrad(vc[c], i, pos[e], n[e], stride[e], sp[e], S.r[e], S.r2[e], bm, b0, bp);
    }
  }
  const float t1 = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1]);
  const float t2 = g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0]);
  const float t3 = g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
  return t1 - t2 + t3;
}

// Backward step t: h <- h(x + v_t/sp) (composeHV, PARTIAL_ID), JT_t = I1(h) (VAL 0),
// JacDet_t = JacDet_{t+1}(x + v_t/sp) (applyU, VAL 0) * |D(x sp + v_t)|, or the
// determinant alone at t = T-1 (jdn == nullptr).
__global__ void __launch_bounds__(kLdmmBlock)
k_orig_bwd(float *__restrict__ nx_, float *__restrict__ ny_, float *__restrict__ nz_,
           const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
           const float *__restrict__ vx, const float *__restrict__ vy, const float *__restrict__ vz,
           const float *__restrict__ I1, const float *__restrict__ jdn, float *__restrict__ jt,
           float *__restrict__ jd, Dims d, Spacing3 S) {
  const int N = d.nx * d.ny * d.nz;
  const int i = (int)xcd_block() * kLdmmBlock + (int)threadIdx.x;
  if (i >= N) return;
  const int x = i % d.nx, r = i / d.nx, y = r % d.ny, z = r / d.ny;
  const float px = (float)x + div_const(vx[i], S.s[0], S.r[0]);
  const float py = (float)y + div_const(vy[i], S.s[1], S.r[1]);
  const float pz = (float)z + div_const(vz[i], S.s[2], S.r[2]);
  const float3 hn = trilerp3g<AWK_BG_PARTIAL_ID>(make_buf(hx, N), make_buf(hy, N), make_buf(hz, N), d,
                                                 px, py, pz);
  nx_[i] = hn.x;
  ny_[i] = hn.y;
  nz_[i] = hn.z;
  jt[i] = trilerp1g<AWK_BG_VAL>(make_buf(I1, N), d, hn.x, hn.y, hn.z, 0.f);
  const float det = jac_det_xsp_v(vx, vy, vz, i, x, y, z, d, S);
  jd[i] = jdn ? trilerp1g<AWK_BG_VAL>(make_buf(jdn, N), d, px, py, pz, 0.f) * det : det;
}

// Forward start: h = identity and m_0 = (I0 - JT_0) JacDet_0 * mscale.
__global__ void __launch_bounds__(kLdmmBlock)
k_orig_init(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
            const float *__restrict__ I0, const float *__restrict__ jt, const float *__restrict__ jd,
            float *__restrict__ m, float mscale, Dims d) {
  const int N = d.nx * d.ny * d.nz;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const int x = i % d.nx, r = i / d.nx;
  hx[i] = (float)x;
  hy[i] = (float)(r % d.ny);
  hz[i] = (float)(r / d.ny);
  m[i] = (I0[i] - jt[i]) * jd[i] * mscale;
}

// Forward step t after the update of v_t: h <- h(x - v_t/sp) (composeHVInv,
// PARTIAL_ID), J = I0(h) (VAL 0) into J (may be null); with the next step's terms,
// m = (J - JT_{t+1}) JacDet_{t+1} * mscale; at the last step diff = J - I1 into
// FP64 SSD partials.
__global__ void __launch_bounds__(kLdmmBlock)
k_orig_fwd(float *__restrict__ nx_, float *__restrict__ ny_, float *__restrict__ nz_,
           const float *__restrict__ hx, const float *__restrict__ hy, const float *__restrict__ hz,
           const float *__restrict__ vx, const float *__restrict__ vy, const float *__restrict__ vz,
           const float *__restrict__ I0, float *__restrict__ J, const float *__restrict__ jt,
           const float *__restrict__ jd, float *__restrict__ m, float mscale,
           const float *__restrict__ I1, double *__restrict__ partials, Dims d, Spacing3 S) {
  const int N = d.nx * d.ny * d.nz;
  const int lb = (int)xcd_block();
  const int i = lb * kLdmmBlock + (int)threadIdx.x;
  double ssd = 0.0;
  if (i < N) {
    const int x = i % d.nx, r = i / d.nx, y = r % d.ny, z = r / d.ny;
    const float px = (float)x - div_const(vx[i], S.s[0], S.r[0]);
    const float py = (float)y - div_const(vy[i], S.s[1], S.r[1]);
    const float pz = (float)z - div_const(vz[i], S.s[2], S.r[2]);
    const float3 hn = trilerp3g<AWK_BG_PARTIAL_ID>(make_buf(hx, N), make_buf(hy, N), make_buf(hz, N), d,
                                                   px, py, pz);
    nx_[i] = hn.x;
    ny_[i] = hn.y;
    nz_[i] = hn.z;
    const float j = trilerp1g<AWK_BG_VAL>(make_buf(I0, N), d, hn.x, hn.y, hn.z, 0.f);
    if (J) J[i] = j;
    if (jt) m[i] = (j - jt[i]) * jd[i] * mscale;
    if (I1) {
      const float dv = j - I1[i];
      ssd = (double)dv * (double)dv;
    }
  }
  if (partials) block_sum_store(ssd, partials, lb);
}

__global__ void k_orig_energy_zero(double *__restrict__ energy) {
  if (threadIdx.x < 3) energy[threadIdx.x] = 0.0;
}

// energy[0] = Prod(sp) sum(diff^2) / sigma^2 measured after the sweep
// (LDMMIteratorGPU.cxx:430-441); energy[2] = energy[0] + energy[1].  Fixed order.
__global__ void k_orig_image_energy(const double *__restrict__ partials, int n, double scale,
                                    double *__restrict__ energy, unsigned *__restrict__ flags) {
  __shared__ double s[256];
  double a = 0.0;
  a = strided_fold(partials, n, (int)threadIdx.x, 256, a, FoldAdd());
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    energy[0] = scale * s[0];
    energy[2] = energy[0] + energy[1];
    if (energy[0] != energy[0]) raise_flags(flags, AWK_FLAG_NAN_ENERGY);
  }
}

}  // namespace awk

using namespace awk;

void awk::ldmm_release(awk_ldmm *l) {
  fk_release(l->k);
  void *bufs[] = {l->h,      l->hs,     l->J,  l->diff,      l->acc,   l->partials, l->maxpart,
                  l->step,   l->energy, l->jt, l->jd, l->sort_work, l->msplat, l->flags};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
}

awk_status awk::ldmm_init(awk_ldmm *l, int nx, int ny, int nz, const float sp[3], float alpha,
                          float beta, float gamma, int n_time_steps, float sigma, float max_pert,
                          bool terms) {
  awk_status st = fk_init(l->k, nx, ny, nz, sp);
  if (st) return st;
  l->k.alpha = alpha;
  l->k.beta = beta;
  l->k.gamma = gamma;
  l->k.lpow = 1.f;
  l->k.flags = 0;
  l->nx = nx;
  l->ny = ny;
  l->nz = nz;
  l->nt = n_time_steps;
  for (int i = 0; i < 3; ++i) l->sp[i] = sp[i];
  l->sigma = sigma;
  l->max_pert = max_pert;
  const size_t N = (size_t)nx * ny * nz;
  l->nblk = fk_xblocks(l->k);
  l->egrid = (int)grid_for((long long)N, kLdmmBlock);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&l->h, 3 * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&l->hs, 3 * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&l->J, (size_t)l->nt * N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&l->diff, N * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&l->acc, N * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc(&l->partials, (size_t)l->egrid * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&l->maxpart, (size_t)l->nt * l->nblk * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&l->step, sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&l->energy, 3 * sizeof(double));
  if (terms && e == hipSuccess) e = hipMalloc(&l->jt, (size_t)l->nt * N * sizeof(float));
  if (terms && e == hipSuccess) e = hipMalloc(&l->jd, (size_t)l->nt * N * sizeof(float));
  if (e == hipSuccess) e = hipMemset(l->acc, 0, N * sizeof(long long));
  if (e == hipSuccess) e = hipMemset(l->step, 0, sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&l->flags, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(l->flags, 0, sizeof(unsigned));
  return e == hipSuccess ? AWK_OK : hip_check(e, "LDMM iterator allocation");
}

// energy[1] += Prod(sp) <L v_t, v_t> / T, with L applied to a copy of v_t in lv (3N).
static awk_status vec_energy_step(awk_ldmm *l, float *const v[], int t, float *lv, double *energy,
                                  hipStream_t s) {
  const int N = l->nx * l->ny * l->nz;
  const double vol = (double)l->sp[0] * l->sp[1] * l->sp[2];
  awk_status st;
  int nb = 0;
  for (int c = 0; c < 3; ++c)
    AWK_HIP(hipMemcpyAsync(lv + (size_t)c * N, v[3 * t + c], (size_t)N * sizeof(float),
                           hipMemcpyDeviceToDevice, s));
  XFwdArgs pa;
  std::memset(&pa, 0, sizeof(pa));
  pa.mode = XLOAD_PLAIN;
  for (int c = 0; c < 3; ++c) pa.in[c] = lv + (size_t)c * N;
  if ((st = fk_xfwd(l->k, pa, s, &nb))) return st;
  if ((st = fk_ypass(l->k, false, s, nullptr))) return st;
  if ((st = fk_zsolve(l->k, false, s))) return st;
  if ((st = fk_ypass(l->k, true, s, nullptr))) return st;
  XInvArgs pi;
  std::memset(&pi, 0, sizeof(pi));
  pi.mode = XSTORE_PLAIN;
  for (int c = 0; c < 3; ++c) pi.out[c] = lv + (size_t)c * N;
  if ((st = fk_xinv(l->k, pi, s, &nb))) return st;
  AWK_LAUNCH(k_ldmm_dot, dim3(l->egrid), dim3(kLdmmBlock), 0, s, lv, v[3 * t], v[3 * t + 1],
             v[3 * t + 2], N, l->partials);
  AWK_LAUNCH(k_ldmm_vec_energy, dim3(1), dim3(256), 0, s, l->partials, l->egrid, vol / l->nt,
             energy, l->flags);
  AWK_LAUNCH_CHECK("vector energy");
  return AWK_OK;
}

// u_t = L^-1 (grad(src) * m) and the update of v_t (or the step-size maximum),
// the force and the update fused into the FFT x passes.
static awk_status solve_update(awk_ldmm *l, float *const v[], int t, const float *src,
                               const float *m, const long long *acc, float *step, bool step_call,
                               hipStream_t s) {
  awk_status st;
  int nb = 0;
  XFwdArgs fa;
  std::memset(&fa, 0, sizeof(fa));
  fa.mode = acc ? XLOAD_GRADMUL_FX : XLOAD_GRADMUL;
  fa.idef = src;
  fa.i1 = m;
  fa.acc = const_cast<long long *>(acc);
  xfwd_spacing(fa, l->sp);
  if ((st = fk_xfwd(l->k, fa, s, &nb))) return st;
  if ((st = fk_ypass(l->k, false, s, nullptr))) return st;
  if ((st = fk_zsolve(l->k, true, s))) return st;
  if ((st = fk_ypass(l->k, true, s, nullptr))) return st;
  XInvArgs ia;
  std::memset(&ia, 0, sizeof(ia));
  ia.mode = step_call ? XSTORE_MAXAXPY : XSTORE_AXPBY;
  for (int c = 0; c < 3; ++c) ia.out[c] = v[3 * t + c];
  ia.maxpart = l->maxpart + (size_t)t * l->nblk;
  ia.step = step;
  ia.sigma2 = l->sigma * l->sigma;
  return fk_xinv(l->k, ia, s, &nb);
}

awk_status awk::ldmm_orig_backward(awk_ldmm *l, const float *I1, float *const v[], float *jt,
                                   float *jd, hipStream_t s) {
  const int nx = l->nx, ny = l->ny, nz = l->nz, T = l->nt;
  const int N = nx * ny * nz;
  const Dims d{nx, ny, nz};
  const unsigned lgrid = (unsigned)((N + kLdmmBlock - 1) / kLdmmBlock);
  float *h = l->h, *hs = l->hs;
  AWK_LAUNCH(k_ldmm_identity, dim3(lgrid), dim3(kLdmmBlock), 0, s, h, h + N, h + 2 * N, d);
  for (int t = T - 1; t >= 0; --t) {
    AWK_LAUNCH(k_orig_bwd, dim3(l->egrid), dim3(kLdmmBlock), 0, s, hs, hs + N, hs + 2 * N, h,
               h + N, h + 2 * N, v[3 * t], v[3 * t + 1], v[3 * t + 2], I1,
               t == T - 1 ? nullptr : jd + (size_t)(t + 1) * N, jt + (size_t)t * N,
               jd + (size_t)t * N, d, spacing3(l->sp));
    AWK_LAUNCH_CHECK("k_orig_bwd");
    std::swap(h, hs);
  }
  return AWK_OK;
}

awk_status awk::ldmm_orig_forward(awk_ldmm *l, const float *I0, const float *I1, float *const v[],
                                  const float *jt, const float *jd, float *step, bool step_call,
                                  float *const alpha[], double *energy, int vec_energy,
                                  hipStream_t s) {
  const int nx = l->nx, ny = l->ny, nz = l->nz, T = l->nt;
  const int N = nx * ny * nz;
  const Dims d{nx, ny, nz};
  const float sx = l->sp[0], sy = l->sp[1], sz = l->sp[2];
  const unsigned egrid = (unsigned)l->egrid;
  const unsigned lgrid = (unsigned)((N + kLdmmBlock - 1) / kLdmmBlock);
  const double vol = (double)sx * sy * sz;
  const float inv_sig2 = 1.f / (l->sigma * l->sigma);
  if (alpha) step_call = false;
  awk_status st;
  float *h = l->h, *hs = l->hs;
  if (!alpha) AWK_LAUNCH(k_orig_energy_zero, dim3(1), dim3(64), 0, s, energy);
  AWK_LAUNCH(k_orig_init, dim3(lgrid), dim3(kLdmmBlock), 0, s, h, h + N, h + 2 * N, I0, jt, jd,
             alpha ? alpha[0] : l->diff, alpha ? inv_sig2 : 1.f, d);
  AWK_LAUNCH_CHECK("k_orig_init");
  for (int t = 0; t < T; ++t) {
    const bool last = (t == T - 1);
    if (!alpha &&
        (st = solve_update(l, v, t, t > 0 ? l->J + (size_t)(t - 1) * N : I0, l->diff, nullptr,
                           step, step_call, s)))
      return st;
    AWK_LAUNCH(k_orig_fwd, dim3(egrid), dim3(kLdmmBlock), 0, s, hs, hs + N, hs + 2 * N, h, h + N,
               h + 2 * N, v[3 * t], v[3 * t + 1], v[3 * t + 2], I0,
               alpha ? nullptr : l->J + (size_t)t * N, last ? nullptr : jt + (size_t)(t + 1) * N,
               last ? nullptr : jd + (size_t)(t + 1) * N,
               alpha ? (last ? nullptr : alpha[t + 1]) : l->diff, alpha ? inv_sig2 : 1.f,
               (last && !alpha) ? I1 : nullptr, (last && !alpha) ? l->partials : nullptr, d,
               spacing3(l->sp));
    AWK_LAUNCH_CHECK("k_orig_fwd");
    std::swap(h, hs);
    if (last && !alpha) {
      // reduce the SSD partials before the vector energy reuses the buffer
      AWK_LAUNCH(k_orig_image_energy, dim3(1), dim3(256), 0, s, l->partials, (int)egrid,
                 vol * inv_sig2, energy, l->flags);
      AWK_LAUNCH_CHECK("k_orig_image_energy");
    }
    if (!alpha && !step_call && vec_energy && (st = vec_energy_step(l, v, t, hs, energy, s)))
      return st;
  }
  if (step_call) {
    if ((st = launch_step_from_max(l->maxpart, T * l->nblk, l->max_pert, step, l->flags, s)))
      return st;
  }
  return AWK_OK;
}

// IterateOrig: backward sweep for JT_t and JacDet_t, then the forward sweep that
// updates each v_t and composes h with the updated field.
static awk_status ldmm_run_orig(awk_ldmm *l, const float *I0, const float *I1, float *const v[],
                                float *const alpha[], double *energy_dev, int vec_energy,
                                hipStream_t s) {
  const bool step_call = !alpha && l->update_step;
  awk_status st = ldmm_orig_backward(l, I1, v, l->jt, l->jd, s);
  if (st) return st;
  st = ldmm_orig_forward(l, I0, I1, v, l->jt, l->jd, l->step, step_call, alpha,
                         energy_dev ? energy_dev : l->energy, vec_energy, s);
  if (st) return st;
  if (step_call) l->update_step = false;
  return AWK_OK;
}

awk_status awk::ldmm_new_iterate(awk_ldmm *l, const float *I0, const float *I1, float *const v[],
                                 float *step, bool step_call, float *const alpha[], double *energy,
                                 int vec_energy, hipStream_t s) {
  const int nx = l->nx, ny = l->ny, nz = l->nz, T = l->nt;
  const int N = nx * ny * nz;
  const Dims d{nx, ny, nz};
  const float sx = l->sp[0], sy = l->sp[1], sz = l->sp[2];
  const unsigned egrid = (unsigned)l->egrid;
  const unsigned lgrid = (unsigned)((N + kLdmmBlock - 1) / kLdmmBlock);
  const double vol = (double)sx * sy * sz;
  if (alpha) step_call = false;
  awk_status st;
  float *h = l->h, *hs = l->hs;
  // ---- forward: J_t = I0(h_t), h_t = h_{t-1}(x - v_t / sp)
  AWK_LAUNCH(k_ldmm_identity, dim3(lgrid), dim3(kLdmmBlock), 0, s, h, h + N, h + 2 * N, d);
  for (int t = 0; t < T; ++t) {
    const bool last = (t == T - 1);
    AWK_LAUNCH(k_ldmm_fwd, dim3(egrid), dim3(kLdmmBlock), 0, s, hs, hs + N, hs + 2 * N, h, h + N,
               h + 2 * N, v[3 * t], v[3 * t + 1], v[3 * t + 2], I0, l->J + (size_t)t * N,
               last ? I1 : nullptr, last ? l->diff : nullptr, last ? l->partials : nullptr, d, sx, sy,
               sz);
    AWK_LAUNCH_CHECK("k_ldmm_fwd");
    std::swap(h, hs);
  }
  AWK_LAUNCH(k_ldmm_image_energy, dim3(1), dim3(256), 0, s, l->partials, (int)egrid,
             vol / ((double)l->sigma * l->sigma), energy, l->flags);
  AWK_LAUNCH_CHECK("k_ldmm_image_energy");
  // ---- backward: h_t = h_{t+1} - v_t(h_{t+1}) / sp; alpha_t = splat(diff, h_t)
  AWK_LAUNCH(k_ldmm_zero, dim3(lgrid), dim3(kLdmmBlock), 0, s, l->acc, N);
  AWK_LAUNCH(k_ldmm_identity, dim3(lgrid), dim3(kLdmmBlock), 0, s, h, h + N, h + 2 * N, d);
  const bool sorted = l->splat_mode == 1;
  for (int t = T - 1; t >= 0; --t) {
    if (sorted) {
      // composeVHInv, then the sort/scan splat of diff to the new h (floor, as the
      // atomic splat the reference's iterator uses)
      AWK_LAUNCH(k_ldmm_vhinv, dim3(lgrid), dim3(kLdmmBlock), 0, s, hs, hs + N, hs + 2 * N, h,
                 h + N, h + 2 * N, v[3 * t], v[3 * t + 1], v[3 * t + 2], d, sx, sy, sz);
      AWK_LAUNCH_CHECK("k_ldmm_vhinv");
      std::swap(h, hs);
      if ((st = launch_splat_sorted(l->msplat, l->diff, h, h + N, h + 2 * N, nx, ny, nz, 0, nullptr,
                                    nullptr, s, 1, l->sort_work)))
        return st;
    } else {
      AWK_LAUNCH(k_ldmm_bwd, dim3(splat_tile_grid<SplatShapeLdmm>(nx, ny, nz)), dim3(kSplatTileThreads), 0, s, hs,
                 hs + N, hs + 2 * N, h, h + N, h + 2 * N, v[3 * t], v[3 * t + 1], v[3 * t + 2],
                 l->diff, reinterpret_cast<unsigned long long *>(l->acc), d, sx, sy, sz);
      AWK_LAUNCH_CHECK("k_ldmm_bwd");
      std::swap(h, hs);
    }
    if (alpha) {
      if (sorted)
        AWK_LAUNCH(k_ldmm_scale, dim3(lgrid), dim3(kLdmmBlock), 0, s, alpha[t], l->msplat, N,
                   1.f / (l->sigma * l->sigma));
      else
        AWK_LAUNCH(k_ldmm_alpha, dim3(lgrid), dim3(kLdmmBlock), 0, s, alpha[t], l->acc, N,
                   1.f / (l->sigma * l->sigma));
      AWK_LAUNCH_CHECK("alpha");
      continue;
    }
    // u = L^-1 (grad(J_{t-1}) * alpha_t), grad(I0) at t = 0
    if ((st = solve_update(l, v, t, t > 0 ? l->J + (size_t)(t - 1) * N : I0,
                           sorted ? l->msplat : nullptr, sorted ? nullptr : l->acc, step, step_call,
                           s)))
      return st;
    // <L v_t, v_t> with L applied to a copy of v_t in the free hfield buffer
    if (!step_call && vec_energy && (st = vec_energy_step(l, v, t, hs, energy, s))) return st;
  }
  if (step_call) {
    if ((st = launch_step_from_max(l->maxpart, T * l->nblk, l->max_pert, step, l->flags, s)))
      return st;
  }
  return AWK_OK;
}

static awk_status ldmm_run(awk_ldmm *l, const float *I0, const float *I1, float *const v[],
                           float *const alpha[], double *energy_dev, int vec_energy, hipStream_t s) {
  const bool step_call = !alpha && l->update_step;
  awk_status st = ldmm_new_iterate(l, I0, I1, v, l->step, step_call, alpha,
                                   energy_dev ? energy_dev : l->energy, vec_energy, s);
  if (st) return st;
  if (step_call) l->update_step = false;
  return AWK_OK;
}

extern "C" {

awk_status awk_ldmm_create(awk_ldmm_t *out, int nx, int ny, int nz, float spx, float spy, float spz,
                           float alpha, float beta, float gamma, int n_time_steps, float sigma,
                           float max_pert) {
  if (!out) return fail(AWK_ERR_INVALID, "awk_ldmm_create: null handle pointer");
  *out = nullptr;
  if (!(spx > 0 && spy > 0 && spz > 0)) return fail(AWK_ERR_INVALID, "awk_ldmm_create: spacing must be > 0");
  if (n_time_steps < 1) return fail(AWK_ERR_INVALID, "awk_ldmm_create: need at least one time step");
  if (!(sigma > 0) || !(max_pert > 0)) return fail(AWK_ERR_INVALID, "awk_ldmm_create: sigma and MaxPert must be > 0");
  if ((long long)nx * ny * nz > (1LL << 28)) return fail(AWK_ERR_INVALID, "awk_ldmm_create: at most 2^28 voxels");
  awk_ldmm *l = new (std::nothrow) awk_ldmm();
  if (!l) return fail(AWK_ERR_ALLOC, "awk_ldmm_create: host allocation failed");
  const float sp[3] = {spx, spy, spz};
  awk_status st = ldmm_init(l, nx, ny, nz, sp, alpha, beta, gamma, n_time_steps, sigma, max_pert, true);
  if (st) {
    ldmm_release(l);
    delete l;
    return st;
  }
  *out = l;
  return AWK_OK;
}

awk_status awk_ldmm_destroy(awk_ldmm_t l) {
  if (!l) return AWK_OK;
  ldmm_release(l);
  delete l;
  return AWK_OK;
}

awk_status awk_ldmm_set_gradient(awk_ldmm_t l, int use_orig) {
  if (!l) return fail(AWK_ERR_INVALID, "awk_ldmm_set_gradient: null handle");
  l->orig = use_orig != 0;
  return AWK_OK;
}

awk_status awk_ldmm_set_splat(awk_ldmm_t l, int mode) {
  if (!l) return fail(AWK_ERR_INVALID, "awk_ldmm_set_splat: null handle");
  if (mode != 0 && mode != 1) return fail(AWK_ERR_INVALID, "awk_ldmm_set_splat: mode must be 0 or 1");
  if (mode == 1 && !l->sort_work) {
    const size_t N = (size_t)l->nx * l->ny * l->nz;
    if (8 * N > (size_t)((1LL << 31) - 1)) return fail(AWK_ERR_INVALID, "awk_ldmm_set_splat: sort/scan splat takes at most 2^28 voxels");
    AWK_HIP(hipMalloc(&l->sort_work, splat_sorted_scratch_bytes(l->nx, l->ny, l->nz)));
    AWK_HIP(hipMalloc(&l->msplat, N * sizeof(float)));
  }
  l->splat_mode = mode;
  return AWK_OK;
}

awk_status awk_ldmm_update_step_next(awk_ldmm_t l) {
  if (!l) return fail(AWK_ERR_INVALID, "awk_ldmm_update_step_next: null handle");
  l->update_step = true;
  return AWK_OK;
}

awk_status awk_ldmm_get_flags(awk_ldmm_t l, unsigned *flags, int clear, awk_stream_t stream) {
  if (!l) return fail(AWK_ERR_INVALID, "awk_ldmm_get_flags: null handle");
  return read_flags(&l->flags, 1, flags, clear, as_stream(stream), "awk_ldmm_get_flags");
}

awk_status awk_ldmm_set_step(awk_ldmm_t l, float step, awk_stream_t stream) {
  if (!l) return fail(AWK_ERR_INVALID, "awk_ldmm_set_step: null handle");
  AWK_HIP(store_scalar_async(l->step, step, as_stream(stream)));
  l->update_step = false;
  return AWK_OK;
}

awk_status awk_ldmm_get_step(awk_ldmm_t l, float *step, awk_stream_t stream) {
  if (!l || !step) return fail(AWK_ERR_INVALID, "awk_ldmm_get_step: null argument");
  AWK_HIP(load_scalar_sync(step, l->step, as_stream(stream)));
  return AWK_OK;
}

awk_status awk_ldmm_iterate(awk_ldmm_t l, const float *I0, const float *I1, float *const v[],
                            double *energy_dev, int vec_energy, awk_stream_t stream) {
  if (!l || !I0 || !I1 || !v) return fail(AWK_ERR_INVALID, "awk_ldmm_iterate: null argument");
  for (int i = 0; i < 3 * l->nt; ++i)
    if (!v[i]) return fail(AWK_ERR_INVALID, "awk_ldmm_iterate: null velocity plane");
  if (l->orig) return ldmm_run_orig(l, I0, I1, v, nullptr, energy_dev, vec_energy, as_stream(stream));
  return ldmm_run(l, I0, I1, v, nullptr, energy_dev, vec_energy, as_stream(stream));
}

awk_status awk_ldmm_alphas(awk_ldmm_t l, const float *I0, const float *I1, float *const v[],
                           float *const alpha[], awk_stream_t stream) {
  if (!l || !I0 || !I1 || !v || !alpha) return fail(AWK_ERR_INVALID, "awk_ldmm_alphas: null argument");
  for (int i = 0; i < 3 * l->nt; ++i)
    if (!v[i]) return fail(AWK_ERR_INVALID, "awk_ldmm_alphas: null velocity plane");
  for (int t = 0; t < l->nt; ++t)
    if (!alpha[t]) return fail(AWK_ERR_INVALID, "awk_ldmm_alphas: null alpha image");
  if (l->orig) return ldmm_run_orig(l, I0, I1, v, alpha, nullptr, 0, as_stream(stream));
  return ldmm_run(l, I0, I1, v, alpha, nullptr, 0, as_stream(stream));
}

awk_status awk_ldmm_deformed_image(awk_ldmm_t l, int t, float *out, awk_stream_t stream) {
  if (!l || !out) return fail(AWK_ERR_INVALID, "awk_ldmm_deformed_image: null argument");
  if (t < 0 || t >= l->nt) return fail(AWK_ERR_INVALID, "awk_ldmm_deformed_image: time step out of range");
  const size_t N = (size_t)l->nx * l->ny * l->nz;
  AWK_HIP(hipMemcpyAsync(out, l->J + (size_t)t * N, N * sizeof(float), hipMemcpyDeviceToDevice,
                         as_stream(stream)));
  return AWK_OK;
}

}  // extern "C"
