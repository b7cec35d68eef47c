// awk_common.h -- shared device/host helpers for the MI355X (gfx950) hot path.
//
// Layout conventions (device):
//   * scalar image: float[nz][ny][nx], x fastest (the reference Array3D order).
//   * vector field: three separate float planes x, y, z of the same shape (SoA,
//     as the reference cplVector3DArray, CUDACommon/src/cudaVector3DArray.cu:46-57).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/awk_hip.h"

namespace awk {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
awk_status fail(awk_status code, const std::string &msg);
awk_status hip_check(hipError_t e, const char *what);

#define AWK_HIP(call)                                          \
  do {                                                         \
    hipError_t _e = (call);                                    \
    if (_e != hipSuccess) return ::awk::hip_check(_e, #call);  \
  } while (0)

#define AWK_LAUNCH_CHECK(name)                                 \
  do {                                                         \
    hipError_t _e = hipGetLastError();                         \
    if (_e != hipSuccess) return ::awk::hip_check(_e, name);   \
  } while (0)

inline hipStream_t as_stream(awk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// One device scalar written in stream order (a kernel argument carries the value,
// so nothing on the host has to outlive the call and a graph may capture it).
template <class T>
__global__ void k_store_scalar(T *p, T v) {
  *p = v;
}
template <class T>
inline hipError_t store_scalar_async(T *dev, T v, hipStream_t s) {
  hipLaunchKernelGGL(k_store_scalar<T>, dim3(1), dim3(1), 0, s, dev, v);
  return hipGetLastError();
}
// Splat accumulators: int64 fixed point with 32 fractional bits, formed in FP32: scaling by 2^32 is
// exact, and the nearest-even rounding of that exact value to an integer is the
// same in any precision (|w| < 2^31).
__device__ __forceinline__ unsigned long long splat_q(float w) {
  return (unsigned long long)(long long)rintf(w * 4294967296.0f);
}
// Back to float: one correctly rounded int64 -> float conversion and an exact
// scaling by 2^-32.
__device__ __forceinline__ float splat_unq(long long q) {
  return (float)q * (1.0f / 4294967296.0f);
}

// Sticky numerical flags (AWK_FLAG_*): a vector atomic OR into the handle's word.
__device__ __forceinline__ void raise_flags(unsigned *flags, unsigned f) {
  if (flags && f) atomicOr(flags, f);
}
// Reads (and optionally clears) n device flag words after the work queued on s;
// AWK_ERR_NUMERIC when a NaN energy or an invalid step is among them.
awk_status read_flags(unsigned *const *dev, int n, unsigned *out, int clear, hipStream_t s,
                      const char *who);

// True when any of the three output planes (n floats each) overlaps any of the
// three input planes: the update must then go through a staging buffer.
inline bool planes_alias(const float *const out[3], const float *const in[3], size_t n) {
  const uintptr_t bytes = n * sizeof(float);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const uintptr_t o = reinterpret_cast<uintptr_t>(out[a]);
      const uintptr_t i = reinterpret_cast<uintptr_t>(in[b]);
      if (o < i + bytes && i < o + bytes) return true;
    }
  return false;
}
// One device scalar read back after the work queued on s.
template <class T>
inline hipError_t load_scalar_sync(T *host, const T *dev, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(host, dev, sizeof(T), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

// Kernel-attached timing events: while set, launches go through
// hipExtLaunchKernelGGL with these start/stop events, which the dispatch itself
// timestamps (kernel duration only, the figure rocprof reports).
struct LaunchEvents {
  hipEvent_t start, stop;
};
LaunchEvents *&launch_events();

#define AWK_LAUNCH(K, G, B, L, S, ...)                                                       \
  do {                                                                                     \
    if (::awk::launch_events())                                                            \
      hipExtLaunchKernelGGL(K, G, B, L, S, ::awk::launch_events()->start,                  \
                            ::awk::launch_events()->stop, 0, __VA_ARGS__);                 \
    else                                                                                   \
      hipLaunchKernelGGL(K, G, B, L, S, __VA_ARGS__);                                      \
  } while (0)

// x / c for a launch constant c, given rc = RN(1 / c): the product, then one
// fma correction by the exact remainder (Markstein's division by a constant).
// Three instructions where an IEEE division takes about ten; over 4e8 samples at
// the spacings 0.5 ... 3 and their doubles it agrees with x / c in every bit.
__device__ __forceinline__ float div_const(float x, float c, float rc) {
  const float q = x * rc;
  return fmaf(fmaf(-q, c, x), rc, q);
}

// A grid spacing with the reciprocals of its two divisors (sp and 2 sp), for
// div_const; built on the host.
struct Spacing3 {
  float s[3], r[3], r2[3];
};
inline Spacing3 spacing3(const float sp[3]) {
  Spacing3 S;
  for (int c = 0; c < 3; ++c) {
    S.s[c] = sp[c];
    S.r[c] = 1.f / sp[c];
    S.r2[c] = 1.f / (2.f * sp[c]);
  }
  return S;
}

// One thread's part of a block-level fold: op over p[start], p[start + stride], ...
// in that order (so a sum rounds exactly as the plain loop would), with the loads
// issued 16 at a time.  A block that folds thousands of partials while the rest of
// the grid saturates HBM then waits a few memory latencies, not one per element.
template <class T, class Op>
__device__ __forceinline__ T strided_fold(const T *__restrict__ p, int n, int start, int stride,
                                          T acc, Op op) {
  int i = start;
  for (; i + 15 * stride < n; i += 16 * stride) {
    T v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = p[i + k * stride];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc = op(acc, v[k]);
  }
  for (; i < n; i += stride) acc = op(acc, p[i]);
  return acc;
}
struct FoldAdd {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct FoldMax {
  __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
};

// Launch-shape knobs (set by the autotuner through awk_tune_set).
struct Tuning {
  int apply_vpt = 1;    // voxels per thread, apply
  int compose_vpt = 1;  // voxels per thread, compose family
  int xinv_unroll = 1;  // voxels per epilogue step of the inverse x pass
  int xinv_pairs = 0;   // inverse x pass: x-neighbour corners as 8-byte pairs inside the volume
  int xinv_stage = 2;   // composing inverse x pass reads the field's rows staged in LDS:
                        // 0 never, 1 always, 2 rows of at most 128 (where it measured faster)
  int xinv_band = 8;
  int xfwd_band = 4;
  int yfwd_rev = 0;     // forward y pass visits its planes in reverse (0: in order)    // forward x pass: row groups per band of its block order (0: planes)    // composing inverse x pass: row groups per band of its block order (0: planes)
  int atlas_mean = 4;   // greedy atlas deformed images + mean: subjects per group of one fused
                        // launch (1, 2, 4), or 0: one apply launch per subject, then the mean
  int splat_tile = 1;   // fixed-point splat through per-block LDS windows (0: direct atomics)
  int unet_lds = 1;     // U-Net convolutions on grids >= 32^3 through LDS-staged input bricks
  int unet_pipe = 2;    // ... K steps' fragment loads software-pipelined: 0 never, 1 always, 2 CC >= 32
  int unet_cc = 0;      // ... cap on the channels staged per chunk (0: all up to 48, else 32)
  // FFT plan knobs, read when an operator is created; they change results at
  // FP32 rounding level (another butterfly order or kernel), never beyond
  int fft_order = 0;    // radix order of the generic LDS stages: 0 largest first, 1 reversed
  int x_rt = 0;         // rows per x-pass block (0: the largest even <= 8 that fits LDS)
  int y_tx = 0;         // kx columns per generic y-pass block (0: the largest <= 16 that fits)
  int z_tx = 0;         // kx columns per generic z-pass block (0: the largest <= 8 that fits)
};
Tuning &tuning();

// Grid size rounded up to a multiple of the XCD count so that xcd_block() is a
// bijection; blocks past the end exit early.
constexpr int kNumXcd = 8;
inline unsigned grid_for(long long items, int per_block) {
  long long b = (items + per_block - 1) / per_block;
  b = (b + kNumXcd - 1) / kNumXcd * kNumXcd;
  return (unsigned)b;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap so that XCD k works through one contiguous 1/8 of the logical blocks in
// order: spatial neighbours then share an L2.  Requires gridDim.x % 8 == 0.
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  return (b % kNumXcd) * (nb / kNumXcd) + b / kNumXcd;
}

// The band block order of a grid of G blocks per plane (the x passes: one block
// per RT rows): where whole planes fill each XCD's share of the grid (the host
// checks it), XCD k takes its planes one band of `band` adjacent blocks at a time,
// so its resident blocks span a narrow y range over many planes and their
// gathers' footprint in its L2 shrinks; band 0 keeps the plane order.  A
// bijection of the logical block index.
__device__ __forceinline__ int band_block(int lb, int band, int G) {
  if (!band) return lb;
  const int per = (int)gridDim.x / kNumXcd, Zc = per / G, nbb = Zc * band;
  const int c = lb / per, j = lb - c * per;
  const int bnd = j / nbb, rem = j - bnd * nbb;
  return (c * Zc + rem / band) * G + bnd * band + rem % band;
}

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// ---------------------------------------------------------------- loads
// Gathers go through a loader: plain pointers, or raw buffer resources (32-bit
// byte offsets, hardware range check: an offset past num_records reads 0).  The
// buffer form saves the 64-bit address arithmetic of every gather.
// pair(i) loads elements i and i + 1 with one 8-byte access (4-byte aligned, which
// multi-dword loads allow): the two x-neighbour corners of a trilinear gather.
typedef float f2a4 __attribute__((ext_vector_type(2), aligned(4)));
struct PtrLoad {
  const float *__restrict__ p;
  __device__ __forceinline__ float operator()(int i) const { return p[i]; }
  __device__ __forceinline__ float2 pair(int i) const {
    const f2a4 v = *reinterpret_cast<const f2a4 *>(p + i);
    return make_float2(v.x, v.y);
  }
};
struct BufLoad {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ float operator()(int i) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, i * 4, 0, 0));
  }
  __device__ __forceinline__ float2 pair(int i) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, i * 4, 0, 0);
    return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
  }
};
// Buffer descriptor word 3 for gfx9-family raw buffers (32-bit data format).
constexpr int kBufDword3 = 0x00020000;
__device__ __forceinline__ BufLoad make_buf(const float *p, int n) {
  return BufLoad{__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, n * 4, kBufDword3)};
}
// volumes whose byte size fits a 31-bit offset can use BufLoad
constexpr long long kBufMaxElems = (1LL << 29);

// A row-major 3x4 affine map of voxel coordinates (kernel argument).
struct Affine34f {
  float m[12];
};

// ---------------------------------------------------------------- trilerp
// Background strategies (numbering of include/awk_hip.h):
//   ID, ZERO, PARTIAL_ID, PARTIAL_ZERO, CLAMP, WRAP, VAL(scalar)
struct Dims {
  int nx, ny, nz;
};

__device__ __forceinline__ int wrap_mod(int r, int b) {
  int m = r % b;
  return m < 0 ? m + b : m;
}

// The nested trilinear lerp of corners c0..c7 (HField3DUtils::trilerp's corner
// order: c0 (x0,y0,z0), c1 (x1,y0,z0), c2 (x1,y1,z0), c3 (x0,y1,z0), c4 (x0,y1,z1),
// c5 (x1,y1,z1), c6 (x1,y0,z1), c7 (x0,y0,z1)): z first, then y, then x, each as
// one multiply and one explicit fma, so every caller rounds identically whatever
// contraction the compiler would choose at its site.
__device__ __forceinline__ float lerp8(float c0, float c1, float c2, float c3, float c4, float c5,
                                       float c6, float c7, float t, float u, float v) {
  const float omt = 1.f - t, omu = 1.f - u, omv = 1.f - v;
  const float a00 = fmaf(c7, v, c0 * omv), a01 = fmaf(c4, v, c3 * omv);
  const float a10 = fmaf(c6, v, c1 * omv), a11 = fmaf(c5, v, c2 * omv);
  return fmaf(t, fmaf(u, a11, omu * a10), omt * fmaf(u, a01, omu * a00));
}

// Vector trilerp of a SoA field at (x, y, z) in voxel coordinates, following
// HField3DUtils::trilerp (UtilitiesDataTypes/HField3DUtils.txx:1969-2300):
// corner order v0..v7, nested lerp, CPU semantics for PARTIAL_ID (out-of-volume
// corners take their own integer coordinates).  FP32 weights.  FAST: a
// wave-uniform path for waves whose points are all inside the volume (fewer
// instructions, a few more registers; off where registers bound occupancy).
template <int BG, class L, bool FAST = true>
__device__ __forceinline__ float3 trilerp3g(const L &fx, const L &fy, const L &fz, Dims d, float x,
                                            float y, float z) {
  const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
  int x0 = (int)flx, y0 = (int)fly, z0 = (int)flz;
  int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  const float t = x - flx, u = y - fly, v = z - flz;
  float c[8][3];
  if (BG == AWK_BG_WRAP || BG == AWK_BG_CLAMP) {
    if (BG == AWK_BG_WRAP) {
      x0 = wrap_mod(x0, d.nx); x1 = wrap_mod(x1, d.nx);
      y0 = wrap_mod(y0, d.ny); y1 = wrap_mod(y1, d.ny);
      z0 = wrap_mod(z0, d.nz); z1 = wrap_mod(z1, d.nz);
    } else {
      x0 = min(max(x0, 0), d.nx - 1); x1 = min(max(x1, 0), d.nx - 1);
      y0 = min(max(y0, 0), d.ny - 1); y1 = min(max(y1, 0), d.ny - 1);
      z0 = min(max(z0, 0), d.nz - 1); z1 = min(max(z1, 0), d.nz - 1);
    }
    const int sxy = d.nx * d.ny;
    const int b00 = z0 * sxy + y0 * d.nx, b10 = z0 * sxy + y1 * d.nx;
    const int b01 = z1 * sxy + y0 * d.nx, b11 = z1 * sxy + y1 * d.nx;
    const int idx[8] = {b00 + x0, b00 + x1, b10 + x1, b10 + x0, b11 + x0, b11 + x1, b01 + x1, b01 + x0};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      c[i][0] = fx(idx[i]);
      c[i][1] = fy(idx[i]);
      c[i][2] = fz(idx[i]);
    }
  } else {
    // ID, ZERO, PARTIAL_ID, PARTIAL_ZERO, branch-free: every corner is loaded from a
    // clamped (always valid) index, so all 24 loads are issued together; corners
    // outside the volume are replaced afterwards.  Per-axis in-range flags of the
    // floor and ceil corners (x0 < n - 1 rather than x1 < n: stays correct when a
    // huge coordinate saturates the float->int conversion).
    const bool xin0 = (x0 >= 0) && (x0 < d.nx), xin1 = (x0 >= -1) && (x0 < d.nx - 1);
    const bool yin0 = (y0 >= 0) && (y0 < d.ny), yin1 = (y0 >= -1) && (y0 < d.ny - 1);
    const bool zin0 = (z0 >= 0) && (z0 < d.nz), zin1 = (z0 >= -1) && (z0 < d.nz - 1);
    const bool inside = xin0 && xin1 && yin0 && yin1 && zin0 && zin1;
    if (FAST && __all(inside)) {
      // every lane of the wave interpolates inside the volume (the common case):
      // corner indices are the base plus fixed offsets, nothing to substitute;
      // the same loads and values as the general path below
      const int b0 = (z0 * d.ny + y0) * d.nx + x0;
      const int b1 = b0 + d.nx, b2 = b0 + d.nx * d.ny, b3 = b2 + d.nx;
      // x-neighbour corners in pairs: (c0, c1) at b0, (c3, c2) at b1, (c4, c5) at b3,
      // (c7, c6) at b2
      const int pb[4] = {b0, b1, b3, b2};
      const int lo[4] = {0, 3, 4, 7}, hi[4] = {1, 2, 5, 6};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float2 qx = fx.pair(pb[k]), qy = fy.pair(pb[k]), qz = fz.pair(pb[k]);
        c[lo[k]][0] = qx.x; c[hi[k]][0] = qx.y;
        c[lo[k]][1] = qy.x; c[hi[k]][1] = qy.y;
        c[lo[k]][2] = qz.x; c[hi[k]][2] = qz.y;
      }
      float r[3];
#pragma unroll
      for (int q = 0; q < 3; ++q)
        r[q] = lerp8(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], c[5][q], c[6][q], c[7][q], t, u, v);
      return make_float3(r[0], r[1], r[2]);
    }
    const int xc0 = min(max(x0, 0), d.nx - 1), xc1 = xin1 ? x0 + 1 : xc0;
    const int yc0 = min(max(y0, 0), d.ny - 1), yc1 = yin1 ? y0 + 1 : yc0;
    const int zc0 = min(max(z0, 0), d.nz - 1), zc1 = zin1 ? z0 + 1 : zc0;
    const int sxy = d.nx * d.ny;
    const int b00 = zc0 * sxy + yc0 * d.nx, b10 = zc0 * sxy + yc1 * d.nx;
    const int b01 = zc1 * sxy + yc0 * d.nx, b11 = zc1 * sxy + yc1 * d.nx;
    const int idx[8] = {b00 + xc0, b00 + xc1, b10 + xc1, b10 + xc0, b11 + xc0, b11 + xc1, b01 + xc1, b01 + xc0};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      c[i][0] = fx(idx[i]);
      c[i][1] = fy(idx[i]);
      c[i][2] = fz(idx[i]);
    }
    if (BG == AWK_BG_ID || BG == AWK_BG_ZERO) {
      if (!inside) return BG == AWK_BG_ID ? make_float3(x, y, z) : make_float3(0.f, 0.f, 0.f);
    } else {
      const bool xin[2] = {xin0, xin1}, yin[2] = {yin0, yin1}, zin[2] = {zin0, zin1};
      const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0};
      const int oy[8] = {0, 0, 1, 1, 1, 1, 0, 0};
      const int oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const bool in = xin[ox[i]] && yin[oy[i]] && zin[oz[i]];
        if (BG == AWK_BG_PARTIAL_ID) {
          // the corner's own integer coordinates (taken from the float floor so
          // that they stay exact for any finite coordinate)
          c[i][0] = in ? c[i][0] : flx + (float)ox[i];
          c[i][1] = in ? c[i][1] : fly + (float)oy[i];
          c[i][2] = in ? c[i][2] : flz + (float)oz[i];
        } else {
          c[i][0] = in ? c[i][0] : 0.f;
          c[i][1] = in ? c[i][1] : 0.f;
          c[i][2] = in ? c[i][2] : 0.f;
        }
      }
    }
  }
  float r[3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    r[q] = lerp8(c[0][q], c[1][q], c[2][q], c[3][q], c[4][q], c[5][q], c[6][q], c[7][q], t, u, v);
  return make_float3(r[0], r[1], r[2]);
}

template <int BG>
__device__ __forceinline__ float3 trilerp3(const float *__restrict__ fx,
                                           const float *__restrict__ fy,
                                           const float *__restrict__ fz, Dims d, float x,
                                           float y, float z) {
  return trilerp3g<BG>(PtrLoad{fx}, PtrLoad{fy}, PtrLoad{fz}, d, x, y, z);
}

// Scalar trilerp (Array3DUtils::trilerp, UtilitiesDataTypes/Array3DUtils.txx:59-232).
// BG: AWK_BG_VAL / AWK_BG_ZERO (background value), AWK_BG_CLAMP, AWK_BG_WRAP.
template <int BG, class L>
__device__ __forceinline__ float trilerp1g(const L &a, Dims d, float x, float y, float z,
                                           float background) {
  const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
  int x0 = (int)flx, y0 = (int)fly, z0 = (int)flz;
  int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  const float t = x - flx, u = y - fly, v = z - flz;
  const float omt = 1.f - t, omu = 1.f - u, omv = 1.f - v;
  float c[8];
  if (BG == AWK_BG_WRAP || BG == AWK_BG_CLAMP) {
    if (BG == AWK_BG_WRAP) {
      x0 = wrap_mod(x0, d.nx); x1 = wrap_mod(x1, d.nx);
      y0 = wrap_mod(y0, d.ny); y1 = wrap_mod(y1, d.ny);
      z0 = wrap_mod(z0, d.nz); z1 = wrap_mod(z1, d.nz);
    } else {
      x0 = min(max(x0, 0), d.nx - 1); x1 = min(max(x1, 0), d.nx - 1);
      y0 = min(max(y0, 0), d.ny - 1); y1 = min(max(y1, 0), d.ny - 1);
      z0 = min(max(z0, 0), d.nz - 1); z1 = min(max(z1, 0), d.nz - 1);
    }
    const int sxy = d.nx * d.ny;
    const int b00 = z0 * sxy + y0 * d.nx, b10 = z0 * sxy + y1 * d.nx;
    const int b01 = z1 * sxy + y0 * d.nx, b11 = z1 * sxy + y1 * d.nx;
    c[0] = a(b00 + x0); c[1] = a(b00 + x1); c[2] = a(b10 + x1); c[3] = a(b10 + x0);
    c[4] = a(b11 + x0); c[5] = a(b11 + x1); c[6] = a(b01 + x1); c[7] = a(b01 + x0);
  } else {
    const float bgv = (BG == AWK_BG_ZERO) ? 0.f : background;
    const bool inside = (x0 >= 0) & (x0 < d.nx - 1) & (y0 >= 0) & (y0 < d.ny - 1) & (z0 >= 0) &
                        (z0 < d.nz - 1);
    if (inside) {
      const int sxy = d.nx * d.ny;
      const int b00 = z0 * sxy + y0 * d.nx, b10 = z0 * sxy + y1 * d.nx;
      const int b01 = z1 * sxy + y0 * d.nx, b11 = z1 * sxy + y1 * d.nx;
      c[0] = a(b00 + x0); c[1] = a(b00 + x1); c[2] = a(b10 + x1); c[3] = a(b10 + x0);
      c[4] = a(b11 + x0); c[5] = a(b11 + x1); c[6] = a(b01 + x1); c[7] = a(b01 + x0);
    } else {
      const bool xin[2] = {(x0 >= 0) && (x0 < d.nx), (x0 >= -1) && (x0 < d.nx - 1)};
      const bool yin[2] = {(y0 >= 0) && (y0 < d.ny), (y0 >= -1) && (y0 < d.ny - 1)};
      const bool zin[2] = {(z0 >= 0) && (z0 < d.nz), (z0 >= -1) && (z0 < d.nz - 1)};
      const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0};
      const int oy[8] = {0, 0, 1, 1, 1, 1, 0, 0};
      const int oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const bool in = xin[ox[i]] & yin[oy[i]] & zin[oz[i]];
        c[i] = in ? a(((z0 + oz[i]) * d.ny + (y0 + oy[i])) * d.nx + (x0 + ox[i])) : bgv;
      }
    }
  }
  return omt * (omu * (c[0] * omv + c[7] * v) + u * (c[3] * omv + c[4] * v)) +
         t * (omu * (c[1] * omv + c[6] * v) + u * (c[2] * omv + c[5] * v));
}

template <int BG>
__device__ __forceinline__ float trilerp1(const float *__restrict__ a, Dims d, float x, float y,
                                          float z, float background) {
  return trilerp1g<BG>(PtrLoad{a}, d, x, y, z, background);
}

}  // namespace awk
