// awk_sortsplat.hip -- deterministic sort/scan splatting on gfx950.
//
// cplSplating (CUDACommon/src/cudaSplat.cu:175-238): every source voxel emits
// 8 (cell, weight*mass) pairs in the createIndex order (cudaSplat.cu:31-117,
// sentinel cell = N for corners outside the volume); the pairs are stably
// sorted by cell; each cell's contributions are summed.  Here:
//   1. k_create_index  keys/values, two uint4/float4 stores per voxel
//   2. stable LSD radix sort, 8-bit digits (a last 9-bit one where that saves a
//      pass), about log2(N+1)/8 passes:
//        k_radix_hist  per-tile digit counts (LDS), digit-major layout
//        k_scan_*      exclusive scan of the counts (3 kernels, integer: exact)
//        k_radix_scatter  stable ranks from wave ballots + per-wave digit counts
//   3. k_segment_sum   one thread per segment head, sequential sum in sorted
//      order (float) over LDS tiles of the sorted pairs: the result is a pure
//      function of the inputs, identical run to run and to a CPU sequential sum
//      in stable-sorted order.
// The reference sums with a parallel CUDPP segmented scan; the keys are the
// bit-exact target, the sums match to rounding.
#include <algorithm>
#include <cstring>
#include <vector>

#include "awk_ops.h"

namespace awk {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;                     // elements per thread per tile
constexpr int kTile = kSortBlock * kSortItems;     // 4096 elements per tile
constexpr int kScanChunk = kSortBlock * kSortItems;

// ---------------------------------------------------------------- keys
__global__ void __launch_bounds__(256)
k_create_index(uint4 *__restrict__ keys, float4 *__restrict__ vals, const float *__restrict__ src,
               const float *__restrict__ vx, const float *__restrict__ vy,
               const float *__restrict__ vz, int w, int h, int l, int truncate, int absolute) {
  const int N = w * h * l;
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= N) return;
  const int i = id % w, rest = id / w, j = rest % h, k = rest / h;
  const float mass = src[id];
  // displacements (cplSplating) or, absolute, the target positions themselves
  const float x = absolute ? vx[id] : (float)i + vx[id];
  const float y = absolute ? vy[id] : (float)j + vy[id];
  const float z = absolute ? vz[id] : (float)k + vz[id];
  int xi = (int)x, yi = (int)y, zi = (int)z;  // the reference truncates (cudaSplat.cu:50-52)
  if (!truncate) {
    if (x < 0.f && x != (float)xi) --xi;
    if (y < 0.f && y != (float)yi) --yi;
    if (z < 0.f && z != (float)zi) --zi;
  }
  const float dx = 1.f - (x - xi), dy = 1.f - (y - yi), dz = 1.f - (z - zi);
  const unsigned sent = (unsigned)N;
  auto in = [&](int a, int b, int c) { return a >= 0 && a < w && b >= 0 && b < h && c >= 0 && c < l; };
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int zz = zi + half;
    const float wz = half ? (1.f - dz) : dz;
    const unsigned base = (unsigned)((zz * h + yi) * w + xi);
    uint4 c;
    float4 d;
    // corner order x, x+1, x+w, x+w+1 and weight products as cudaSplat.cu:64-112
    c.x = in(xi, yi, zz) ? base : sent;
    d.x = in(xi, yi, zz) ? mass * dx * dy * wz : 0.f;
    c.y = in(xi + 1, yi, zz) ? base + 1 : sent;
    d.y = in(xi + 1, yi, zz) ? mass * (1.f - dx) * dy * wz : 0.f;
    c.z = in(xi, yi + 1, zz) ? base + w : sent;
    d.z = in(xi, yi + 1, zz) ? mass * dx * (1.f - dy) * wz : 0.f;
    c.w = in(xi + 1, yi + 1, zz) ? base + w + 1 : sent;
    d.w = in(xi + 1, yi + 1, zz) ? mass * (1.f - dx) * (1.f - dy) * wz : 0.f;
    keys[(size_t)half * N + id] = c;
    vals[(size_t)half * N + id] = d;
  }
}

// ---------------------------------------------------------------- radix sort
// Per-tile digit counts of one pass, RB-bit digits, digit-major layout
// (hist[d * ntiles + tile]).
template <int RB>
__global__ void __launch_bounds__(kSortBlock)
k_radix_hist(const unsigned *__restrict__ keys, long long M, int shift, int ntiles,
             unsigned *__restrict__ hist) {
  constexpr int R = 1 << RB;
  __shared__ unsigned cnt[R];
  const int t = blockIdx.x;
  for (int q = threadIdx.x; q < R; q += kSortBlock) cnt[q] = 0;
  __syncthreads();
  const long long base = (long long)t * kTile;
  const unsigned *__restrict__ kt = keys + base;
  const int n = M - base < kTile ? (int)(M - base) : kTile;
  const int lane = threadIdx.x & 63;
  // counts do not depend on the order: each thread takes 4 groups of 4 consecutive
  // keys, 16-byte loads where the group lies inside the pairs
  unsigned kr[kSortItems];
#pragma unroll
  for (int g = 0; g < kSortItems / 4; ++g) {
    const int e = 4 * (g * kSortBlock + (int)threadIdx.x);
    if (e + 3 < n) {
      const uint4 q = *reinterpret_cast<const uint4 *>(kt + e);
      kr[4 * g] = q.x; kr[4 * g + 1] = q.y; kr[4 * g + 2] = q.z; kr[4 * g + 3] = q.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) kr[4 * g + j] = e + j < n ? kt[e + j] : 0u;
    }
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int e = 4 * ((r / 4) * kSortBlock + (int)threadIdx.x) + (r % 4);
    const bool valid = e < n;
    const unsigned d = (kr[r] >> shift) & (R - 1);
    // a wave whose keys share one digit (the common case past the first pass:
    // cells of neighbouring voxels) adds once; otherwise each lane adds, with
    // little contention when the digits differ.  Valid lanes are a prefix, so
    // lane 0 holds a valid digit whenever any lane does.
    const unsigned d0 = __builtin_amdgcn_readfirstlane(d);
    if (__all(!valid || d == d0)) {
      const unsigned long long vm = __ballot(valid);
      if (lane == 0 && vm) atomicAdd(&cnt[d0], (unsigned)__popcll(vm));
    } else if (valid) {
      atomicAdd(&cnt[d], 1u);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < R; q += kSortBlock) hist[(size_t)q * ntiles + t] = cnt[q];
}

// exclusive scan of n unsigned values, in three steps (exact integer arithmetic)
__global__ void __launch_bounds__(kSortBlock)
k_scan_reduce(const unsigned *__restrict__ a, long long n, unsigned *__restrict__ sums) {
  __shared__ unsigned s[kSortBlock];
  const long long base = (long long)blockIdx.x * kScanChunk;
  unsigned acc = 0;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const long long e = base + r * kSortBlock + threadIdx.x;
    if (e < n) acc += a[e];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kSortBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned *s, unsigned &total) {
  // Hillis-Steele over the block (kSortBlock threads)
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kSortBlock; off <<= 1) {
    const unsigned add = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  total = s[kSortBlock - 1];
  const unsigned incl = s[threadIdx.x];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(kSortBlock)
k_scan_top(unsigned *__restrict__ sums, int n) {
  __shared__ unsigned s[kSortBlock];
  unsigned carry = 0;
  for (int base = 0; base < n; base += kSortBlock) {
    const int i = base + threadIdx.x;
    const unsigned v = i < n ? sums[i] : 0u;
    unsigned total;
    const unsigned ex = block_exclusive_scan(v, s, total);
    if (i < n) sums[i] = carry + ex;
    carry += total;
  }
}

__global__ void __launch_bounds__(kSortBlock)
k_scan_down(unsigned *__restrict__ a, long long n, const unsigned *__restrict__ sums) {
  __shared__ unsigned s[kSortBlock];
  // thread t owns kSortItems consecutive entries
  const long long base = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kSortItems;
  unsigned v[kSortItems];
  unsigned acc = 0;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    v[r] = (base + r < n) ? a[base + r] : 0u;
    acc += v[r];
  }
  unsigned total;
  unsigned run = sums[blockIdx.x] + block_exclusive_scan(acc, s, total);
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (base + r < n) a[base + r] = run;
    run += v[r];
  }
}

// Stable scatter of one pass.  Wave w owns the tile's elements
// [w * 1024, (w + 1) * 1024), loaded in 16 coalesced rounds of 64; a stable order
// is index order, i.e. (wave, round, lane).  Within a wave the rank among equal
// digits comes from RB ballots plus the wave's running count of each digit, kept
// in the wave's own LDS row (a wave's LDS operations complete in program order,
// so the rounds need no barrier).  One barrier later the per-wave totals give each
// wave's start within each digit's run of the tile, every element goes to its
// tile-local sorted position in LDS (digit-major), and the tile is written out in
// that order: consecutive threads store consecutive addresses of one digit's run.
// The tile's digit counts are differences of consecutive entries of the exclusive
// scan (digit-major: entry d * ntiles + t), the last against M.
template <int RB>
__global__ void __launch_bounds__(kSortBlock)
k_radix_scatter(const unsigned *__restrict__ kin, const float *__restrict__ vin,
                unsigned *__restrict__ kout, float *__restrict__ vout, long long M, int shift,
                int ntiles, const unsigned *__restrict__ offs) {
  constexpr int kWaves = kSortBlock / 64;
  constexpr int kSpan = kTile / kWaves;  // elements per wave
  constexpr int kRounds = kSpan / 64;
  constexpr int R = 1 << RB, DPT = R / kSortBlock;  // digits per thread
  static_assert(DPT * kSortBlock == R && kRounds * 64 * kWaves == kTile, "tile shape");
  __shared__ unsigned gbase[R];   // global start of this tile's run of each digit
  __shared__ unsigned lstart[R];  // tile-local start of each digit
  __shared__ unsigned wrun[kWaves][R];
  __shared__ unsigned lk[kTile];
  __shared__ float lv[kTile];
  __shared__ unsigned scan_s[kSortBlock];
  const int t = blockIdx.x;
  const long long nh = (long long)R * ntiles;
  // thread t owns digits t * DPT ... t * DPT + DPT - 1
  unsigned c[DPT], tsum = 0;
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const int d = (int)threadIdx.x * DPT + k;
    const long long hi = (long long)d * ntiles + t;
    const unsigned g0 = offs[hi];
    c[k] = (hi + 1 < nh ? offs[hi + 1] : (unsigned)M) - g0;
    gbase[d] = g0;
    tsum += c[k];
  }
  for (int q = threadIdx.x; q < kWaves * R; q += kSortBlock) (&wrun[0][0])[q] = 0u;
  // exclusive scan of the tile's digit counts (its barriers also order the zeroing)
  unsigned total;
  unsigned ex = block_exclusive_scan(tsum, scan_s, total);
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    lstart[(int)threadIdx.x * DPT + k] = ex;
    ex += c[k];
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // tile-local 32-bit indices: element w0 + r * 64 + lane of the tile
  const long long tile0 = (long long)t * kTile;
  const unsigned *__restrict__ kt = kin + tile0;
  const float *__restrict__ vt = vin + tile0;
  const int n = M - tile0 < kTile ? (int)(M - tile0) : kTile;
  const int w0 = wave * kSpan + lane;
  // the wave's loads first: one memory latency per tile, not per round
  unsigned kr[kRounds], dr[kRounds], pr[kRounds];
  float vr[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int e = w0 + r * 64;
    kr[r] = e < n ? kt[e] : 0u;
    vr[r] = e < n ? vt[e] : 0.f;
  }
  unsigned *run = wrun[wave];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const bool valid = w0 + r * 64 < n;
    const unsigned d = (kr[r] >> shift) & (R - 1);
    unsigned long long same = __ballot(valid);
    // lanes with this lane's digit: all valid lanes when the wave shares one digit
    // (valid lanes are a prefix, so lane 0's digit is a valid one), else RB ballots
    const unsigned d0 = __builtin_amdgcn_readfirstlane(d);
    if (!__all(!valid || d == d0)) {
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long m = __ballot(bit);
        same &= bit ? m : ~m;
      }
    }
    const unsigned before = valid ? run[d] : 0u;
    // the lowest lane of each digit group advances the wave's count
    if (valid && (same & lt_mask) == 0ull) run[d] = before + (unsigned)__popcll(same);
    dr[r] = d;
    pr[r] = before + (unsigned)__popcll(same & lt_mask);
  }
  __syncthreads();
  // each wave's start within each digit's run: exclusive prefix over the waves
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const int d = (int)threadIdx.x * DPT + k;
    unsigned acc = lstart[d];
#pragma unroll
    for (int w2 = 0; w2 < kWaves; ++w2) {
      const unsigned n = wrun[w2][d];
      wrun[w2][d] = acc;
      acc += n;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    if (w0 + r * 64 < n) {
      const unsigned pos = run[dr[r]] + pr[r];
      lk[pos] = kr[r];
      lv[pos] = vr[r];
    }
  }
  __syncthreads();
  // the tile in digit-major order: slot p of digit d goes to gbase[d] + p - lstart[d]
  for (int p = threadIdx.x; p < n; p += kSortBlock) {
    const unsigned key = lk[p];
    const unsigned d = (key >> shift) & (R - 1);
    const unsigned g = gbase[d] + ((unsigned)p - lstart[d]);
    kout[g] = key;
    vout[g] = lv[p];
  }
}

// ---------------------------------------------------------------- segment sums
__global__ void __launch_bounds__(256) k_zero_image(float *__restrict__ dst, long long N) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) dst[i] = 0.f;
}

// Segment sums over the sorted pairs: the thread at a segment's head adds the
// segment's values in sorted order.  A block stages a tile of kSegTile pairs in
// LDS with coalesced loads, so the heads walk their segments in LDS; a segment
// that runs past the tile's end continues in global memory.  The same additions
// in the same order as one walk over global memory.
constexpr int kSegItems = 8;
constexpr int kSegTile = 256 * kSegItems;
__global__ void __launch_bounds__(256)
k_segment_sum(const unsigned *__restrict__ keys, const float *__restrict__ vals, long long M,
              unsigned N, float *__restrict__ dst) {
  __shared__ unsigned sk[kSegTile + 1];  // sk[0]: the key before the tile
  __shared__ float sv[kSegTile];
  const long long t0 = (long long)blockIdx.x * kSegTile;
  // past the pairs: a key no cell has (cells are < N <= 2^31), which ends every walk
  constexpr unsigned kNone = 0xffffffffu;
#pragma unroll
  for (int q = 0; q < kSegItems; ++q) {
    const int l = q * 256 + (int)threadIdx.x;
    const long long e = t0 + l;
    sk[l + 1] = e < M ? keys[e] : kNone;
    sv[l] = e < M ? vals[e] : 0.f;
  }
  if (threadIdx.x == 0) sk[0] = t0 > 0 ? keys[t0 - 1] : kNone;
  __syncthreads();
#pragma unroll 1
  for (int q = 0; q < kSegItems; ++q) {
    const int l = q * 256 + (int)threadIdx.x;
    const unsigned k = sk[l + 1];
    if (k >= N || sk[l] == k) continue;  // past the pairs, the outside-volume cell, or not a head
    float acc = 0.f;
    int j = l;
    // four pairs per LDS round trip, added one by one in order while the key holds
    bool open = true;
    while (open && j + 4 <= kSegTile) {
      const unsigned k0 = sk[j + 1], k1 = sk[j + 2], k2 = sk[j + 3], k3 = sk[j + 4];
      const float v0 = sv[j], v1 = sv[j + 1], v2 = sv[j + 2], v3 = sv[j + 3];
      if (k0 != k) { open = false; break; }
      acc += v0;
      if (k1 != k) { j += 1; open = false; break; }
      acc += v1;
      if (k2 != k) { j += 2; open = false; break; }
      acc += v2;
      if (k3 != k) { j += 3; open = false; break; }
      acc += v3;
      j += 4;
    }
    if (open)
      while (j < kSegTile && sk[j + 1] == k) acc += sv[j++];
    if (j == kSegTile)
      for (long long g = t0 + kSegTile; g < M && keys[g] == k; ++g) acc += vals[g];
    dst[k] = acc;
  }
}

size_t splat_sorted_scratch_bytes(int nx, int ny, int nz) {
  const long long M = 8LL * nx * ny * nz;
  const int ntiles = (int)((M + kTile - 1) / kTile);
  const long long nh = 2LL * kRadix * ntiles;  // room for a 9-bit pass
  const int nchunks = (int)((nh + kScanChunk - 1) / kScanChunk);
  // two key/value buffers, histogram, chunk sums
  return 2 * (size_t)M * 4 + 2 * (size_t)M * 4 + (size_t)nh * 4 + (size_t)nchunks * 4 + 256;
}

awk_status launch_splat_sorted(float *dst, const float *src, const float *vx, const float *vy,
                               const float *vz, int nx, int ny, int nz, int truncate,
                               unsigned *keys_out, float *vals_out, hipStream_t s, int absolute,
                               void *work) {
  const long long N = (long long)nx * ny * nz;
  const long long M = 8 * N;
  int bits = 1;
  while ((1LL << bits) < N + 1) ++bits;
  // 8-bit passes; 9 remaining bits go in one pass (a power-of-two volume's keys,
  // cells 0 .. N - 1 and the outside-volume cell N, need one bit past a multiple of 8)
  std::vector<int> plan;
  for (int rem = bits; rem > 0;) {
    const int b = rem == 9 ? 9 : std::min(rem, kRadixBits);
    plan.push_back(b);
    rem -= b;
  }
  const int ntiles = (int)((M + kTile - 1) / kTile);
  const long long nh_max = 2LL * kRadix * ntiles;
  // scratch: the caller's work buffer (splat_sorted_scratch_bytes) or stream-ordered
  char *scratch = static_cast<char *>(work);
  if (!scratch) AWK_HIP(hipMallocAsync((void **)&scratch, splat_sorted_scratch_bytes(nx, ny, nz), s));
  unsigned *ka = reinterpret_cast<unsigned *>(scratch);
  unsigned *kb = ka + M;
  float *va = reinterpret_cast<float *>(kb + M);
  float *vb = va + M;
  unsigned *hist = reinterpret_cast<unsigned *>(vb + M);
  unsigned *sums = hist + nh_max;
  AWK_LAUNCH(k_create_index, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s,
             reinterpret_cast<uint4 *>(ka), reinterpret_cast<float4 *>(va), src, vx, vy, vz, nx, ny,
             nz, truncate, absolute);
  AWK_LAUNCH_CHECK("k_create_index");
  int shift = 0;
  for (const int b : plan) {
    const long long nh = (long long)(b == 9 ? 2 * kRadix : kRadix) * ntiles;
    const int nchunks = (int)((nh + kScanChunk - 1) / kScanChunk);
    if (b == 9) AWK_LAUNCH(k_radix_hist<9>, dim3(ntiles), dim3(kSortBlock), 0, s, ka, M, shift, ntiles, hist);
    else AWK_LAUNCH(k_radix_hist<8>, dim3(ntiles), dim3(kSortBlock), 0, s, ka, M, shift, ntiles, hist);
    AWK_LAUNCH(k_scan_reduce, dim3(nchunks), dim3(kSortBlock), 0, s, hist, nh, sums);
    AWK_LAUNCH(k_scan_top, dim3(1), dim3(kSortBlock), 0, s, sums, nchunks);
    AWK_LAUNCH(k_scan_down, dim3(nchunks), dim3(kSortBlock), 0, s, hist, nh, sums);
    if (b == 9)
      AWK_LAUNCH(k_radix_scatter<9>, dim3(ntiles), dim3(kSortBlock), 0, s, ka, va, kb, vb, M, shift,
                 ntiles, hist);
    else
      AWK_LAUNCH(k_radix_scatter<8>, dim3(ntiles), dim3(kSortBlock), 0, s, ka, va, kb, vb, M, shift,
                 ntiles, hist);
    AWK_LAUNCH_CHECK("radix pass");
    shift += b;
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  AWK_LAUNCH(k_zero_image, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, dst, N);
  AWK_LAUNCH(k_segment_sum, dim3((unsigned)((M + kSegTile - 1) / kSegTile)), dim3(256), 0, s, ka, va, M,
             (unsigned)N, dst);
  AWK_LAUNCH_CHECK("k_segment_sum");
  if (keys_out) AWK_HIP(hipMemcpyAsync(keys_out, ka, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
  if (vals_out) AWK_HIP(hipMemcpyAsync(vals_out, va, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
  if (!work) AWK_HIP(hipFreeAsync(scratch, s));
  return AWK_OK;
}

}  // namespace awk

using namespace awk;

extern "C" awk_status awk_splat_sorted(float *dst, const float *src, const float *vx,
                                       const float *vy, const float *vz, int nx, int ny, int nz,
                                       int reference_truncation, unsigned int *keys_out,
                                       float *vals_out, awk_stream_t stream) {
  if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * ny * nz * 8 > (1LL << 31) - 1)
    return fail(AWK_ERR_INVALID, "awk_splat_sorted: bad sizes (at most 2^28 voxels)");
  if (!dst || !src || !vx || !vy || !vz) return fail(AWK_ERR_INVALID, "awk_splat_sorted: null pointer");
  return launch_splat_sorted(dst, src, vx, vy, vz, nx, ny, nz, reference_truncation ? 1 : 0, keys_out,
                             vals_out, as_stream(stream), 0, nullptr);
}
