// awk_splat_tile.h -- tiled fixed-point splatting shared by the splat entry point
// and the LDMM backward sweep.
//
// A block takes a TX x TY x TZ tile of source voxels.  Pass A forms every voxel's
// destination and the bounding box of the 8-corner footprints; when the box
// (clipped to the volume) fits the LDS window of CAP cells, pass B accumulates the
// int64 fixed-point contributions (2^-32 units, the same quantisation as the
// direct kernel) with LDS atomics and one flush adds the non-zero window cells to
// the global accumulator; otherwise pass B adds straight to global memory.
// Integer sums commute, so the result is bit-identical to the direct splat for
// any tiling or scheduling.  Each thread handles the same voxels in both passes,
// so a position it stored in pass A is read back by the same thread.
//
// Tile shapes: a tile's footprint is the tile itself widened by the variation of
// the displacement across it.  For smooth fields of up to 8 voxels at 256^3 a
// 16 x 8 x 8 tile's box stays under 4,100 cells (4,032 at most, 2,173 on average,
// over all tiles of the sinusoid workload), where a 64 x 8 x 8 tile's exceeds
// 7,168 in 99 % of tiles; the flush then issues about 1.3 global atomics per
// source voxel instead of 8.
#pragma once

#include <climits>

#include "awk_common.h"

namespace awk {

constexpr int kSplatTileThreads = 256;

template <int TX, int TY, int TZ, int CAP>
struct SplatShape {
  static constexpr int tx = TX, ty = TY, tz = TZ, cap = CAP;
  static constexpr int vpt = TX * TY * TZ / kSplatTileThreads;
  static_assert(TX * TY * TZ % kSplatTileThreads == 0, "tile must be a multiple of the block");
  static_assert((TX & (TX - 1)) == 0 && (TY & (TY - 1)) == 0, "power-of-two tile rows");
};
// the splat entry points (1,024 voxels, 32 KiB window: five blocks per CU)
using SplatShapeEntry = SplatShape<16, 8, 8, 4096>;
// the LDMM backward sweep (4,096 voxels, 56 KiB window), as measured there
using SplatShapeLdmm = SplatShape<64, 8, 8, 7168>;

template <class S>
inline unsigned splat_tile_grid(int nx, int ny, int nz) {
  return (unsigned)(((nx + S::tx - 1) / S::tx) * ((ny + S::ty - 1) / S::ty) *
                    ((nz + S::tz - 1) / S::tz));
}

// Positions given by an hfield.
struct SplatPosH {
  const float *hx, *hy, *hz;
  __device__ __forceinline__ float3 first(int i) const { return make_float3(hx[i], hy[i], hz[i]); }
  __device__ __forceinline__ float3 again(int i) const { return make_float3(hx[i], hy[i], hz[i]); }
};

// LDMM backward step: h' = h - v(h)/sp (composeVHInv, PARTIAL_ZERO), stored to
// (ox, oy, oz) in pass A and read back by the same thread in pass B.
struct SplatPosLdmm {
  const float *gx, *gy, *gz;
  BufLoad vx, vy, vz;
  Dims d;
  float sx, sy, sz;
  float *ox, *oy, *oz;
  __device__ __forceinline__ float3 first(int i) const {
    const float a = gx[i], b = gy[i], c = gz[i];
    const float3 o = trilerp3g<AWK_BG_PARTIAL_ZERO>(vx, vy, vz, d, a, b, c);
    const float x = o.x / -sx + a, y = o.y / -sy + b, z = o.z / -sz + c;
    ox[i] = x;
    oy[i] = y;
    oz[i] = z;
    return make_float3(x, y, z);
  }
  __device__ __forceinline__ float3 again(int i) const { return make_float3(ox[i], oy[i], oz[i]); }
};

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// One block's tile of shape S: contributions in the direct kernel's factor order
// (mass * wx * wy * wz) and rounding (splat_q).
template <class S, class P>
__device__ void splat_tile(const P &pos, const float *__restrict__ mass,
                           unsigned long long *__restrict__ acc, Dims d) {
  constexpr int kSTX = S::tx, kSTY = S::ty, kSTZ = S::tz, kSplatTileVPT = S::vpt;
  __shared__ unsigned long long win[S::cap];
  __shared__ int red[6][kSplatTileThreads / 64];
  const int ntx = (d.nx + kSTX - 1) / kSTX, nty = (d.ny + kSTY - 1) / kSTY;
  const int b = (int)blockIdx.x;
  const int x0 = (b % ntx) * kSTX, y0 = ((b / ntx) % nty) * kSTY, z0 = (b / (ntx * nty)) * kSTZ;
  const int t = (int)threadIdx.x;
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
  // pass A: destinations and their bounding box
#pragma unroll 4
  for (int v = 0; v < kSplatTileVPT; ++v) {
    const int lin = v * kSplatTileThreads + t;
    const int x = x0 + (lin & (kSTX - 1)), y = y0 + ((lin / kSTX) & (kSTY - 1)), z = z0 + lin / (kSTX * kSTY);
    if (x >= d.nx || y >= d.ny || z >= d.nz) continue;
    const float3 p = pos.first((z * d.ny + y) * d.nx + x);
    const int xi = (int)floorf(p.x), yi = (int)floorf(p.y), zi = (int)floorf(p.z);
    mn[0] = min(mn[0], xi); mx[0] = max(mx[0], xi);
    mn[1] = min(mn[1], yi); mx[1] = max(mx[1], yi);
    mn[2] = min(mn[2], zi); mx[2] = max(mx[2], zi);
  }
  const int wave = t >> 6, lane = t & 63;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int lo = wave_min(mn[a]), hi = wave_max(mx[a]);
    if (lane == 0) {
      red[a][wave] = lo;
      red[3 + a][wave] = hi;
    }
  }
  __syncthreads();
  int lo[3], W[3];
  bool empty = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int l = red[a][0], h = red[3 + a][0];
    for (int w = 1; w < kSplatTileThreads / 64; ++w) {
      l = min(l, red[a][w]);
      h = max(h, red[3 + a][w]);
    }
    const int n = a == 0 ? d.nx : a == 1 ? d.ny : d.nz;
    // corners span [floor, floor + 1]; keep the part inside [0, n - 1]
    const int clo = l < 0 ? 0 : l;
    const int chi = h >= n - 1 ? n - 1 : h + 1;
    lo[a] = clo;
    W[a] = chi - clo + 1;
    if (clo > chi) empty = true;
  }
  if (empty) return;  // uniform over the block: nothing lands inside the volume
  const long long cells = (long long)W[0] * W[1] * W[2];
  const bool use_lds = cells <= S::cap;
  if (use_lds) {
    for (int c = t; c < (int)cells; c += kSplatTileThreads) win[c] = 0ull;
    __syncthreads();
  }
  // pass B: contributions
#pragma unroll 2
  for (int v = 0; v < kSplatTileVPT; ++v) {
    const int lin = v * kSplatTileThreads + t;
    const int x = x0 + (lin & (kSTX - 1)), y = y0 + ((lin / kSTX) & (kSTY - 1)), z = z0 + lin / (kSTX * kSTY);
    if (x >= d.nx || y >= d.ny || z >= d.nz) continue;
    const int i = (z * d.ny + y) * d.nx + x;
    const float3 p = pos.again(i);
    const float m = mass[i];
    const float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    const int xi = (int)fx, yi = (int)fy, zi = (int)fz;
    const float dx = 1.f - (p.x - fx), dy = 1.f - (p.y - fy), dz = 1.f - (p.z - fz);
    const float wx[2] = {dx, 1.f - dx}, wy[2] = {dy, 1.f - dy}, wz[2] = {dz, 1.f - dz};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
      const int X = xi + ox, Y = yi + oy, Z = zi + oz;
      if (X < 0 || X >= d.nx || Y < 0 || Y >= d.ny || Z < 0 || Z >= d.nz) continue;
      const unsigned long long q = splat_q(m * wx[ox] * wy[oy] * wz[oz]);
      if (use_lds)
        atomicAdd(&win[((Z - lo[2]) * W[1] + (Y - lo[1])) * W[0] + (X - lo[0])], q);
      else
        atomicAdd(acc + (Z * d.ny + Y) * d.nx + X, q);
    }
  }
  if (use_lds) {
    __syncthreads();
    for (int c = t; c < (int)cells; c += kSplatTileThreads) {
      const unsigned long long q = win[c];
      if (q == 0ull) continue;
      const int cx = c % W[0], r = c / W[0], cy = r % W[1], cz = r / W[1];
      atomicAdd(acc + ((lo[2] + cz) * d.ny + (lo[1] + cy)) * d.nx + (lo[0] + cx), q);
    }
  }
}

}  // namespace awk
