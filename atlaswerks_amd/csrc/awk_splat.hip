// awk_splat.hip -- splatting (forward apply) on gfx950.
//
// dst(y) = sum_x w(h(x), y) * src(x) over the 8 trilinear corners of h(x)
// (cplSplat3DH, CUDACommon/src/cudaSplatAtomic.cu:185-247 and 569-599, with the
// floor fix ADJUST_INT of cudaSplatAtomic.cu:27).  Three accumulation modes:
//   FIXED64  int64 fixed point with 32 fractional bits: integer adds commute, so
//            the result is bit-identical for any scheduling (deterministic);
//            per-contribution rounding error <= 2^-33, |sum| < 2^31.  By default
//            accumulated per block in LDS windows first (awk_splat_tile.h), then
//            converted to float by a pass that also clears the accumulator.
//   FIXED20  the reference's int32 fixed point S2p20 = int(a * 2^20 + 0.5f)
//            (CUDACommon/inc/cplMacro.h:17, 67-91), accumulated in the dst buffer
//            itself and converted in place (FixedToFloating); |sum| < 2048.
//   FLOAT    float atomics (fast; summation order, hence rounding, varies).
// Buffers are cleared by fill kernels rather than memset nodes, so a captured
// graph replays them correctly.
#include <mutex>

#include "awk_ops.h"
#include "awk_splat_tile.h"

namespace awk {

constexpr int kSplatBlock = 256;

template <int MODE>
__global__ void __launch_bounds__(kSplatBlock)
k_splat(void *__restrict__ acc, const float *__restrict__ src, const float *__restrict__ hx,
        const float *__restrict__ hy, const float *__restrict__ hz, int nx, int ny, int nz) {
  const int N = nx * ny * nz;
  const int i = (int)(xcd_block() * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const float x = hx[i], y = hy[i], z = hz[i];
  const float mass = src[i];
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const int xi = (int)fx, yi = (int)fy, zi = (int)fz;
  const float dx = 1.f - (x - fx), dy = 1.f - (y - fy), dz = 1.f - (z - fz);
  const float wx[2] = {dx, 1.f - dx}, wy[2] = {dy, 1.f - dy}, wz[2] = {dz, 1.f - dz};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
    const int X = xi + ox, Y = yi + oy, Z = zi + oz;
    if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
    const int id = (Z * ny + Y) * nx + X;
    // same factor order as the reference: mass * wx * wy * wz
    const float w = mass * wx[ox] * wy[oy] * wz[oz];
    if (MODE == AWK_SPLAT_FIXED64) {
      atomicAdd(reinterpret_cast<unsigned long long *>(acc) + id, splat_q(w));
    } else if (MODE == AWK_SPLAT_FIXED20) {
      atomicAdd(reinterpret_cast<int *>(acc) + id, (int)(w * 1048576.f + 0.5f));
    } else {
      atomicAdd(reinterpret_cast<float *>(acc) + id, w);
    }
  }
}

__global__ void __launch_bounds__(kSplatTileThreads)
k_splat_tiled(unsigned long long *__restrict__ acc, const float *__restrict__ src,
              const float *__restrict__ hx, const float *__restrict__ hy,
              const float *__restrict__ hz, Dims d) {
  splat_tile<SplatShapeEntry>(SplatPosH{hx, hy, hz}, src, acc, d);
}

template <class T>
__global__ void k_fill(T *__restrict__ p, long long n, T v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// dst = acc * 2^-32 in float; the accumulator is left cleared for its next use.
__global__ void k_fixed64_to_float(float *__restrict__ dst, long long *__restrict__ acc, int N) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const long long q = acc[i];
  acc[i] = 0;
  dst[i] = splat_unq(q);
}

__global__ void k_fixed20_to_float(float *__restrict__ dst, int N) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < N) dst[i] = (float)reinterpret_cast<const int *>(dst)[i] / 1048576.f;
}

template <class T>
awk_status fill_async(T *p, long long n, T v, hipStream_t s) {
  if (n <= 0) return AWK_OK;
  AWK_LAUNCH(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  AWK_LAUNCH_CHECK("k_fill");
  return AWK_OK;
}

awk_status clear_splat_acc(unsigned long long *acc, long long n, hipStream_t s) {
  return fill_async(acc, n, 0ull, s);
}

awk_status launch_splat_acc(float *dst, const float *src, const float *hx, const float *hy,
                            const float *hz, int nx, int ny, int nz, unsigned long long *acc,
                            hipStream_t s) {
  const long long N = (long long)nx * ny * nz;
  if (tuning().splat_tile)
    AWK_LAUNCH(k_splat_tiled, dim3(splat_tile_grid<SplatShapeEntry>(nx, ny, nz)),
               dim3(kSplatTileThreads), 0, s, acc, src, hx, hy, hz, Dims{nx, ny, nz});
  else
    AWK_LAUNCH(k_splat<AWK_SPLAT_FIXED64>, dim3(grid_for(N, kSplatBlock)), dim3(kSplatBlock), 0,
               s, (void *)acc, src, hx, hy, hz, nx, ny, nz);
  AWK_LAUNCH_CHECK("k_splat<FIXED64>");
  AWK_LAUNCH(k_fixed64_to_float, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, dst,
             reinterpret_cast<long long *>(acc), (int)N);
  AWK_LAUNCH_CHECK("k_fixed64_to_float");
  return AWK_OK;
}

awk_status launch_splat(float *dst, const float *src, const float *hx, const float *hy,
                        const float *hz, int nx, int ny, int nz, int mode, hipStream_t s) {
  const long long N = (long long)nx * ny * nz;
  const unsigned grid = grid_for(N, kSplatBlock);
  const unsigned cgrid = (unsigned)((N + 255) / 256);
  if (mode == AWK_SPLAT_FIXED64) {
    // a stream-ordered allocation: after the first call the pool hands the same
    // block back without a device allocation
    void *acc = nullptr;
    AWK_HIP(hipMallocAsync(&acc, (size_t)N * 8, s));
    awk_status st = clear_splat_acc(reinterpret_cast<unsigned long long *>(acc), N, s);
    if (st == AWK_OK)
      st = launch_splat_acc(dst, src, hx, hy, hz, nx, ny, nz,
                            reinterpret_cast<unsigned long long *>(acc), s);
    hipError_t e = hipFreeAsync(acc, s);
    if (st != AWK_OK) return st;
    AWK_HIP(e);
  } else if (mode == AWK_SPLAT_FIXED20) {
    if ((const void *)dst == (const void *)src) return fail(AWK_ERR_INVALID, "awk_splat: dst aliases src");
    awk_status st = fill_async(reinterpret_cast<int *>(dst), N, 0, s);
    if (st != AWK_OK) return st;
    AWK_LAUNCH(k_splat<AWK_SPLAT_FIXED20>, dim3(grid), dim3(kSplatBlock), 0, s, (void *)dst,
                       src, hx, hy, hz, nx, ny, nz);
    AWK_LAUNCH_CHECK("k_splat<FIXED20>");
    AWK_LAUNCH(k_fixed20_to_float, dim3(cgrid), dim3(256), 0, s, dst, (int)N);
    AWK_LAUNCH_CHECK("k_fixed20_to_float");
  } else if (mode == AWK_SPLAT_FLOAT) {
    if ((const void *)dst == (const void *)src) return fail(AWK_ERR_INVALID, "awk_splat: dst aliases src");
    awk_status st = fill_async(dst, N, 0.f, s);
    if (st != AWK_OK) return st;
    AWK_LAUNCH(k_splat<AWK_SPLAT_FLOAT>, dim3(grid), dim3(kSplatBlock), 0, s, (void *)dst,
                       src, hx, hy, hz, nx, ny, nz);
    AWK_LAUNCH_CHECK("k_splat<FLOAT>");
  } else {
    return fail(AWK_ERR_INVALID, "awk_splat: unknown mode");
  }
  return AWK_OK;
}

}  // namespace awk

using namespace awk;

// ---- the splat without allocations (awk_splat_noalloc)
//
// FIXED64 needs an int64 accumulator beside dst.  Callers that cannot allocate
// (code under graph capture, serving loops) get it from this file-scope array:
// one copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel: its device address is fetched once per device with
// hipGetSymbolAddress and kept in `g_scratch`.  Every call on a device shares that
// device's array, so calls that overlap on one GPU race.  Each call's conversion
// pass leaves the part it used cleared; a device's array is cleared once when it
// is first set up and again after any call that failed.
constexpr long long kNoAllocMaxVoxels = 1LL << 22;  // 32 MiB of int64 accumulators
static __device__ __align__(256) unsigned long long scratch[kNoAllocMaxVoxels];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long *g_scratch[kMaxDevices] = {};
bool g_scratch_dirty[kMaxDevices] = {};
}  // namespace

static awk_status noalloc_scratch(hipStream_t s, unsigned long long **out, int *dev_out) {
  int dev = -1;
  AWK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(AWK_ERR_UNSUPPORTED, "awk_splat_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_check(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                "awk_splat_noalloc: scratch symbol address");
    g_scratch[dev] = reinterpret_cast<unsigned long long *>(p);
    g_scratch_dirty[dev] = true;
  }
  if (g_scratch_dirty[dev]) {
    awk_status st = clear_splat_acc(g_scratch[dev], kNoAllocMaxVoxels, s);
    if (st != AWK_OK) return st;
    g_scratch_dirty[dev] = false;
  }
  *out = g_scratch[dev];
  *dev_out = dev;
  return AWK_OK;
}

extern "C" awk_status awk_splat_noalloc(float *dst, const float *src, const float *hx,
                                        const float *hy, const float *hz, int nx, int ny, int nz,
                                        int mode, awk_stream_t stream) {
  if (nx < 1 || ny < 1 || nz < 1)
    return fail(AWK_ERR_INVALID, "awk_splat_noalloc: bad sizes");
  if ((long long)nx * ny * nz > kNoAllocMaxVoxels)
    return fail(AWK_ERR_INVALID, "awk_splat_noalloc: more than 2^22 voxels (the scratch holds 2^22 accumulators)");
  if (!dst || !src || !hx || !hy || !hz) return fail(AWK_ERR_INVALID, "awk_splat_noalloc: null pointer");
  if (mode != AWK_SPLAT_FIXED64 && mode != AWK_SPLAT_FIXED20 && mode != AWK_SPLAT_FLOAT)
    return fail(AWK_ERR_INVALID, "awk_splat_noalloc: unknown mode");
  hipStream_t s = as_stream(stream);
  if (mode != AWK_SPLAT_FIXED64)  // these accumulate in dst itself and allocate nothing
    return launch_splat(dst, src, hx, hy, hz, nx, ny, nz, mode, s);
  unsigned long long *acc = nullptr;
  int dev = -1;
  awk_status st = noalloc_scratch(s, &acc, &dev);
  if (st != AWK_OK) return st;
  st = launch_splat_acc(dst, src, hx, hy, hz, nx, ny, nz, acc, s);
  if (st != AWK_OK) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    g_scratch_dirty[dev] = true;
  }
  return st;
}

extern "C" awk_status awk_splat(float *dst, const float *src, const float *hx, const float *hy,
                                const float *hz, int nx, int ny, int nz, int mode,
                                awk_stream_t stream) {
  if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * ny * nz > kBufMaxElems)
    return fail(AWK_ERR_INVALID, "awk_splat: bad sizes");
  if (!dst || !src || !hx || !hy || !hz) return fail(AWK_ERR_INVALID, "awk_splat: null pointer");
  return launch_splat(dst, src, hx, hy, hz, nx, ny, nz, mode, as_stream(stream));
}
