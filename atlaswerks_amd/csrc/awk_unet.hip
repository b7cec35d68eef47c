// awk_unet.hip -- 3-D U-Net warm start on MFMA (gfx950 bf16 matrix cores).
//
// Predicts an initial displacement u (3 x nz x ny x nx, fp32) from a (template,
// subject) pair; the greedy or LDMM loop then starts from h = x + u.  No
// reference counterpart: the network is fixed at 4 levels with 16-32-64-128
// channels, two 3x3x3 convolutions per level, 2x2x2 average pooling, nearest x2
// upsampling concatenated with the skip tensor, LeakyReLU(0.2), a 1x1x1 head.
//
// Activations are channels-last bf16 (NDHWC: act[((z*ny + y)*nx + x)*C + c]), C a
// multiple of 8.  A 3x3x3 convolution is an implicit GEMM D[co][p] = sum_k W[co][k]
// X[k][p] on v_mfma_f32_16x16x32_bf16 with fp32 accumulation.  K is tap-major,
// k = tap*C + ci, in chunks of 8 channels, so a lane's 8 consecutive k are 8
// channels of one tap (one 16-byte load at the position's base address plus the
// tap's constant offset, guarded by a 27-bit in-volume mask computed once per
// position) and the 4 lane groups of a K step read neighbouring channels of the
// same position wherever C >= 32.  A wave owns MT 16-channel output tiles x PT
// 16-position tiles; bias, LeakyReLU (or, for gradients, the derivative of the
// layer's activation) and bf16 rounding are fused into the epilogue, which stores
// 4 consecutive channels per lane.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "awk_unet.h"

namespace awk {

// blockIdx.y selects the wave's group of MT output-channel tiles.  SK = 1: the 4
// waves of a block take consecutive position groups; SK = 4: they share one
// position group and split the K steps (ks = wave mod 4), and wave 0 adds the
// other waves' partial sums from LDS before the epilogue; there the fragments of
// the next step are loaded before the MFMAs of the current one.
template <int MT, int PT, int SK>
__global__ void __launch_bounds__(kConvBlock)
k_conv3(ConvArgs g) {
  __shared__ f32x4 red[SK > 1 ? (SK - 1) * MT * PT * 64 : 1];
  const int mbase = (int)blockIdx.y * MT * 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int N = g.nx * g.ny * g.nz * g.batch;
  const int slot = SK > 1 ? (int)blockIdx.x : (int)blockIdx.x * kConvWaves + wave;
  const int p0 = slot * (kPosPerWave * PT);
  if (p0 >= N) return;  // uniform over the block when SK > 1
  const int col = lane & 15, grp = lane >> 4;
  // per position tile: base address and the mask of taps inside the volume
  long long base[PT];
  unsigned mask[PT];
#pragma unroll
  for (int t = 0; t < PT; ++t) {
    const int p = p0 + t * 16 + col;
    unsigned m = 0u;
    if (p < N) {
      const int px = p % g.nx, pr = p / g.nx, py = pr % g.ny, pz = (pr / g.ny) % g.nz;
      const unsigned xm = (px > 0 ? 1u : 0u) | 2u | (px < g.nx - 1 ? 4u : 0u);
      const unsigned ym = (py > 0 ? 1u : 0u) | 2u | (py < g.ny - 1 ? 4u : 0u);
      const unsigned zm = (pz > 0 ? 1u : 0u) | 2u | (pz < g.nz - 1 ? 4u : 0u);
#pragma unroll
      for (int tap = 0; tap < 27; ++tap)
        if (((xm >> (tap % 3)) & (ym >> ((tap / 3) % 3)) & (zm >> (tap / 9))) & 1u) m |= 1u << tap;
    }
    mask[t] = m;
    base[t] = (long long)p * g.C;
  }
  const int sy = g.nx, sz = g.nx * g.ny;
  f32x4 acc[MT][PT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < PT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ksteps = g.kpad >> 5;
  bf16x8 bv[PT], av[MT];
  auto load = [&](int ks, bf16x8 *b, bf16x8 *a) {
    const int c = ks * 4 + grp;
    const int tap = g.C == 8 ? c : (int)__umulhi((unsigned)c, g.c8magic), cg = c - tap * (g.C >> 3);
    const int tz = tap / 9, tr = tap - tz * 9, ty = tr / 3, tx = tr - ty * 3;
    const long long off = (long long)((tz - 1) * sz + (ty - 1) * sy + (tx - 1)) * g.C + cg * 8;
    const bool live = c < g.nchunks;
#pragma unroll
    for (int t = 0; t < PT; ++t) {
      b[t] = zero8();
      if (live && ((mask[t] >> tap) & 1u)) b[t] = *reinterpret_cast<const bf16x8 *>(g.x + base[t] + off);
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
      a[m] = *reinterpret_cast<const bf16x8 *>(g.w + (size_t)(mbase + m * 16 + col) * g.kpad + ks * 32 + grp * 8);
  };
  if constexpr (SK > 1) {
    // split K: few waves, so the next step's loads are issued before these MFMAs
    if (wave < ksteps) load(wave, bv, av);
    for (int ks = wave; ks < ksteps; ks += SK) {
      bf16x8 bn[PT], an[MT];
      const bool more = ks + SK < ksteps;
      if (more) load(ks + SK, bn, an);
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < PT; ++t)
          acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[t], acc[m][t], 0, 0, 0);
      if (more) {
#pragma unroll
        for (int t = 0; t < PT; ++t) bv[t] = bn[t];
#pragma unroll
        for (int m = 0; m < MT; ++m) av[m] = an[m];
      }
    }
  } else {
    // full grids: occupancy hides the load latency
    for (int ks = 0; ks < ksteps; ++ks) {
      load(ks, bv, av);
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < PT; ++t)
          acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[t], acc[m][t], 0, 0, 0);
    }
  }
  if (SK > 1) {
    if (wave > 0) {
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < PT; ++t) red[(((wave - 1) * MT + m) * PT + t) * 64 + lane] = acc[m][t];
    }
    __syncthreads();
    if (wave > 0) return;
#pragma unroll
    for (int w = 0; w < SK - 1; ++w)
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < PT; ++t) acc[m][t] += red[((w * MT + m) * PT + t) * 64 + lane];
  }
  // epilogue: lane holds channels m*16 + 4*grp + 0..3 of its position in tile t
#pragma unroll
  for (int t = 0; t < PT; ++t) {
    const int p = p0 + t * 16 + col;
    if (p >= N) continue;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int c0 = mbase + m * 16 + grp * 4;
      const size_t o0 = (size_t)p * g.cout + c0;
      bf16x4 o;
      if (g.dact) {
        const bf16x4 a = *reinterpret_cast<const bf16x4 *>(g.dact + o0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float v = acc[m][t][i];
          o[i] = (bf16)((float)a[i] > 0.f ? v : v * g.slope);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float v = acc[m][t][i] + (g.bias ? g.bias[c0 + i] : 0.f);
          v = v >= 0.f ? v : v * g.slope;
          o[i] = (bf16)v;
        }
      }
      *reinterpret_cast<bf16x4 *>(g.out + o0) = o;
    }
  }
}

// LDS-staged form for grids of at least 32^3 per volume (nx % 16 == 0, ny and
// nz % 4 == 0).  A block owns a 16 x 4 x 4 brick of output positions (wave w: the
// brick's plane z0 + w, its 4 rows as the 4 position tiles) and MT 16-channel
// output tiles.  The brick's 18 x 6 x 6 input halo is staged in LDS one chunk of
// CC channels at a time (positions outside the volume as zeros: the convolution's
// zero padding), so every tap's fragment is an LDS read and each input value is
// fetched from memory once per brick instead of once per tap.  K runs chunk by
// chunk, and within a chunk over (tap, 8-channel group) in the weights' tap-major
// order.  Positions p = ((b*nz + z)*ny + y)*nx + x of the stacked volumes; a brick
// never straddles two volumes.
constexpr int kLdsBX = 16, kLdsBY = 4, kLdsBZ = 4;
constexpr int kLdsHX = kLdsBX + 2, kLdsHY = kLdsBY + 2, kLdsHZ = kLdsBZ + 2;
constexpr int kLdsHalo = kLdsHX * kLdsHY * kLdsHZ;  // 648 positions

template <int MT, bool PIPE>
__global__ void __launch_bounds__(kConvBlock)
k_conv3_lds(ConvArgs g, int CC) {
  extern __shared__ bf16 tile[];  // kLdsHalo x PS, PS = CC + 8, then the K-group table
  const int PS = CC + 8;          // the 8 padding channels are zeros (dead K groups read them)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int mbase = (int)blockIdx.y * MT * 16;
  const int nbx = g.nx / kLdsBX, nby = g.ny / kLdsBY;
  const int b = (int)blockIdx.x;
  const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
  const int x0 = bx * kLdsBX, y0 = by * kLdsBY, zs = bz * kLdsBZ;  // zs: stacked z
  const int vol = zs / g.nz, z0 = zs - vol * g.nz;
  const size_t vbase = (size_t)vol * g.nz * g.ny * g.nx;
  f32x4 acc[MT][kLdsBY];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < kLdsBY; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int groups = CC >> 3;         // 8-channel groups per position in a chunk
  const int cgroups = 27 * groups;    // (tap, group) pairs per chunk
  const int ksteps = (cgroups + 3) >> 2;
  const int nchunk = g.C / CC;
  // c / groups as a multiply-high (groups > 1): exact for c < 2^32 / groups^2
  const unsigned gmagic = groups > 1 ? (unsigned)(4294967296ull / (unsigned)groups) + 1u : 0u;
  auto divg = [&](int c) { return groups == 1 ? c : (int)__umulhi((unsigned)c, gmagic); };
  // K-group table: for c = 4 ks + grp, the LDS offset of its fragment relative to
  // the lane's own position and the weight offset within a chunk; past the last
  // group, the lane's own zero padding (and weight offset 0: finite weights times
  // zeros add nothing)
  int2 *koff = reinterpret_cast<int2 *>(tile + kLdsHalo * PS);
  for (int c = threadIdx.x; c < ksteps * 4; c += kConvBlock) {
    int2 o = make_int2(CC, 0);
    if (c < cgroups) {
      const int tap = divg(c), gi = c - tap * groups;
      const int tz = tap / 9, tr = tap - tz * 9, ty = tr / 3, tx = tr - ty * 3;
      o = make_int2(((tz * kLdsHY + ty) * kLdsHX + tx) * PS + gi * 8, tap * g.C + gi * 8);
    }
    koff[c] = o;
  }
  for (int hp = threadIdx.x; hp < kLdsHalo; hp += kConvBlock)
    *reinterpret_cast<bf16x8 *>(tile + hp * PS + CC) = zero8();
  // the lane's positions (4 rows of its plane) and weight rows
  int bpos[kLdsBY];
#pragma unroll
  for (int t = 0; t < kLdsBY; ++t) bpos[t] = ((wave * kLdsHY + t) * kLdsHX + col) * PS;
  const bf16 *wrow[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) wrow[m] = g.w + (size_t)(mbase + m * 16 + col) * g.kpad;
  for (int ch = 0; ch < nchunk; ++ch) {
    if (ch) __syncthreads();  // the previous chunk's reads are done
    // one halo position per thread and step: its coordinates, bounds test and
    // addresses once, then its CC / 8 channel groups
    for (int hp = threadIdx.x; hp < kLdsHalo; hp += kConvBlock) {
      const int hx = hp % kLdsHX, hr = hp / kLdsHX, hy = hr % kLdsHY, hz = hr / kLdsHY;
      const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = z0 - 1 + hz;
      bf16 *dst = tile + hp * PS;
      if (x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz) {
        const size_t pos = vbase + ((size_t)z * g.ny + y) * g.nx + x;
        if (!g.xa) {
          const bf16 *src = g.x + pos * g.C + ch * CC;
          for (int gi = 0; gi < groups; ++gi)
            *reinterpret_cast<bf16x8 *>(dst + gi * 8) = *reinterpret_cast<const bf16x8 *>(src + gi * 8);
        } else {
          // [the upsampled coarse tensor (ca channels), the skip tensor]
          const int hx2 = g.nx >> 1, hy2 = g.ny >> 1, hz2 = g.nz >> 1;
          const size_t q = (size_t)vol * hz2 * hy2 * hx2 + ((size_t)(z >> 1) * hy2 + (y >> 1)) * hx2 + (x >> 1);
          const bf16 *sa = g.xa + q * g.ca;
          const bf16 *sb = g.x + pos * (g.C - g.ca);
          for (int gi = 0; gi < groups; ++gi) {
            const int cidx = ch * CC + gi * 8;
            const bf16 *src = cidx < g.ca ? sa + cidx : sb + (cidx - g.ca);
            *reinterpret_cast<bf16x8 *>(dst + gi * 8) = *reinterpret_cast<const bf16x8 *>(src);
          }
        }
      } else {
        for (int gi = 0; gi < groups; ++gi) *reinterpret_cast<bf16x8 *>(dst + gi * 8) = zero8();
      }
    }
    __syncthreads();
    const int cw = ch * CC;
    auto load = [&](int ks, bf16x8 *av, bf16x8 *bv) {
      const int2 o = koff[ks * 4 + grp];
#pragma unroll
      for (int t = 0; t < kLdsBY; ++t) bv[t] = *reinterpret_cast<const bf16x8 *>(tile + bpos[t] + o.x);
#pragma unroll
      for (int m = 0; m < MT; ++m) av[m] = *reinterpret_cast<const bf16x8 *>(wrow[m] + cw + o.y);
    };
    if constexpr (!PIPE) {
      // the weight fragments (global, L1/L2 hits) of the next step are requested
      // before this step's MFMAs; the input fragments come from LDS within the step
      bf16x8 an[MT];
      {
        const int oy = koff[grp].y;
#pragma unroll
        for (int m = 0; m < MT; ++m) an[m] = *reinterpret_cast<const bf16x8 *>(wrow[m] + cw + oy);
      }
      for (int ks = 0; ks < ksteps; ++ks) {
        bf16x8 av[MT], bv[kLdsBY];
#pragma unroll
        for (int m = 0; m < MT; ++m) av[m] = an[m];
        if (ks + 1 < ksteps) {
          const int oy = koff[(ks + 1) * 4 + grp].y;
#pragma unroll
          for (int m = 0; m < MT; ++m) an[m] = *reinterpret_cast<const bf16x8 *>(wrow[m] + cw + oy);
        }
        const int ox = koff[ks * 4 + grp].x;
#pragma unroll
        for (int t = 0; t < kLdsBY; ++t) bv[t] = *reinterpret_cast<const bf16x8 *>(tile + bpos[t] + ox);
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int t = 0; t < kLdsBY; ++t)
            acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[t], acc[m][t], 0, 0, 0);
      }
    } else {
      // the next step's fragments are loaded before this step's MFMAs, into the
      // other of two register sets (two steps per trip: no copies)
      bf16x8 a0[MT], b0[kLdsBY], a1[MT], b1[kLdsBY];
      auto mma = [&](const bf16x8 *av, const bf16x8 *bv) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int t = 0; t < kLdsBY; ++t)
            acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[t], acc[m][t], 0, 0, 0);
      };
      load(0, a0, b0);
      for (int ks = 0; ks < ksteps; ks += 2) {  // ks + 1, ks + 2 < ksteps: uniform
        if (ks + 1 < ksteps) load(ks + 1, a1, b1);
        mma(a0, b0);
        if (ks + 1 < ksteps) {
          if (ks + 2 < ksteps) load(ks + 2, a0, b0);
          mma(a1, b1);
        }
      }
    }
  }
  // epilogue: lane holds channels m*16 + 4*grp + 0..3 of position (x0 + col, y0 + t, zs + wave)
#pragma unroll
  for (int t = 0; t < kLdsBY; ++t) {
    const size_t p = ((size_t)(zs + wave) * g.ny + (y0 + t)) * g.nx + x0 + col;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int c0 = mbase + m * 16 + grp * 4;
      const size_t o0 = p * g.cout + c0;
      bf16x4 o;
      if (g.dact) {
        const bf16x4 a = *reinterpret_cast<const bf16x4 *>(g.dact + o0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float v = acc[m][t][i];
          o[i] = (bf16)((float)a[i] > 0.f ? v : v * g.slope);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float v = acc[m][t][i] + (g.bias ? g.bias[c0 + i] : 0.f);
          v = v >= 0.f ? v : v * g.slope;
          o[i] = (bf16)v;
        }
      }
      *reinterpret_cast<bf16x4 *>(g.out + o0) = o;
    }
  }
}

// two fp32 images -> bf16 NDHWC with 8 channels (template, subject, 6 zeros)
__global__ void __launch_bounds__(256)
k_pack_input(bf16 *__restrict__ out, const float *__restrict__ t, const float *__restrict__ s, int N) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  bf16x8 v = zero8();
  v[0] = (bf16)t[i];
  v[1] = (bf16)s[i];
  *reinterpret_cast<bf16x8 *>(out + (size_t)i * 8) = v;
}

// 2x2x2 average pooling, C channels (multiple of 8), fp32 sum of the 8 bf16 values
__global__ void __launch_bounds__(256)
k_pool2(bf16 *__restrict__ out, const bf16 *__restrict__ in, int nx, int ny, int nz, int C) {
  const int ox = nx >> 1, oy = ny >> 1, oz = nz >> 1;
  const int groups = C >> 3;
  const long long total = (long long)ox * oy * oz * groups;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  const int gi = (int)(id % groups);
  const long long q = id / groups;
  const int x = (int)(q % ox), r = (int)(q / ox), y = r % oy, z = r / oy;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const int X = 2 * x + (d & 1), Y = 2 * y + ((d >> 1) & 1), Z = 2 * z + (d >> 2);
    const bf16x8 v = *reinterpret_cast<const bf16x8 *>(in + ((size_t)(Z * ny + Y) * nx + X) * C + gi * 8);
#pragma unroll
    for (int c = 0; c < 8; ++c) s[c] += (float)v[c];
  }
  bf16x8 o;
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = (bf16)(s[c] * 0.125f);
  *reinterpret_cast<bf16x8 *>(out + (size_t)q * C + gi * 8) = o;
}

// out = [nearest x2 upsample of a (ca channels, half resolution), b (cb channels)]
__global__ void __launch_bounds__(256)
k_upcat(bf16 *__restrict__ out, const bf16 *__restrict__ a, int ca, const bf16 *__restrict__ b,
        int cb, int nx, int ny, int nz) {
  const int C = ca + cb, groups = C >> 3;
  const long long total = (long long)nx * ny * nz * groups;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  const int gi = (int)(id % groups);
  const long long p = id / groups;
  const int x = (int)(p % nx), r = (int)(p / nx), y = r % ny, z = r / ny;
  const int c = gi * 8;
  bf16x8 v;
  if (c < ca) {
    const long long q = ((long long)(z >> 1) * (ny >> 1) + (y >> 1)) * (nx >> 1) + (x >> 1);
    v = *reinterpret_cast<const bf16x8 *>(a + q * ca + c);
  } else {
    v = *reinterpret_cast<const bf16x8 *>(b + p * cb + (c - ca));
  }
  *reinterpret_cast<bf16x8 *>(out + p * C + c) = v;
}

// 1x1x1 head: u_c = b_c + sum_ci W[c][ci] d[ci] (fp32), written as 3 planes
__global__ void __launch_bounds__(256)
k_head(float *__restrict__ ux, float *__restrict__ uy, float *__restrict__ uz,
       const bf16 *__restrict__ d, const float *__restrict__ w, const float *__restrict__ b, int N,
       int C) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  float a0 = b[0], a1 = b[1], a2 = b[2];
  // the position's C channels (a multiple of 8) as 16-byte loads, summed in order
  const bf16 *dp = d + (size_t)i * C;
  for (int c0 = 0; c0 < C; c0 += 8) {
    const bf16x8 v8 = *reinterpret_cast<const bf16x8 *>(dp + c0);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int c = c0 + j;
      const float v = (float)v8[j];
      a0 += w[c] * v;
      a1 += w[C + c] * v;
      a2 += w[2 * C + c] * v;
    }
  }
  ux[i] = a0;
  uy[i] = a1;
  uz[i] = a2;
}

// Lane-layout probe: D = A B with A[m][k] = ((3m + k) % 5) - 2 and B[k][n] =
// ((7k + n) % 5) - 2 (exact in bf16 and fp32); lane l writes its 4 outputs.
__global__ void k_mfma_probe(float *__restrict__ out) {
  const int l = threadIdx.x;
  bf16x8 a, b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * (l >> 4) + j;
    a[j] = (bf16)(float)((3 * (l & 15) + k) % 5 - 2);
    b[j] = (bf16)(float)((7 * k + (l & 15)) % 5 - 2);
  }
  const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) out[l * 4 + i] = d[i];
}

// Launch one 3x3x3 convolution.  Full grids: the most reuse per wave (MT * PT <=
// 16) that still gives at least 8192 waves.  Grids too small for that keep all
// output channels per wave (MT up to 4) and split K over the block's 4 waves,
// splitting the channels over blockIdx.y only if that is still too few.
// Channels staged per chunk.  Chunks of 16 keep 5 blocks per CU resident where the
// output tile is narrow; 32-channel chunks (software-pipelined) pay off from 64
// outputs or more than 48 inputs on (measured per layer at 128^3: 48 -> 16
// 302 -> 225 us, 32 -> 32 37 -> 32 us with 16-channel chunks; 192 -> 64 70 -> 85 us,
// 64 -> 64 27 -> 32 us the other way).  unet_cc, when set, caps the chunk.
static int lds_chunk(const ConvArgs &g) {
  int CC = g.C <= 16 ? g.C : (g.C > 48 || (g.cout >= 64 && g.C % 32 == 0)) ? 32 : 16;
  const int ccap = tuning().unet_cc;
  if (ccap && CC > ccap && g.C % ccap == 0) CC = ccap;
  return CC;
}

bool conv3_lds_ok(const ConvArgs &g) {
  const long long N1 = (long long)g.nx * g.ny * g.nz;
  const int CC = lds_chunk(g);
  const int mt_all = g.cout / 16;
  const int mt = mt_all % 4 == 0 ? 4 : mt_all % 2 == 0 ? 2 : 1;
  return tuning().unet_lds && N1 >= 32768 && g.nx % kLdsBX == 0 && g.ny % kLdsBY == 0 &&
         g.nz % kLdsBZ == 0 && g.C % CC == 0 && g.cout % 16 == 0 && (mt_all / mt) * mt * 16 == g.cout &&
         (!g.xa || (g.ca % 8 == 0 && g.ca < g.C && g.nx % 2 == 0 && g.ny % 2 == 0 && g.nz % 2 == 0));
}

awk_status launch_conv3(const ConvArgs &g, hipStream_t s) {
  const long long N1 = (long long)g.nx * g.ny * g.nz, N = N1 * (g.batch < 1 ? 1 : g.batch);
  const int mt_all = g.cout / 16;
  if (g.cout % 16 || mt_all < 1 || mt_all > 16 || g.C % 8 || g.C < 8)
    return fail(AWK_ERR_INVALID, "unet: channels must be multiples of 16 (outputs, at most 256) and 8 (inputs)");
  if (g.batch < 1 || N > (1LL << 30)) return fail(AWK_ERR_INVALID, "unet: bad batch");
  // the LDS-staged form where one volume is at least 32^3 (chosen per volume, so a
  // batch rounds as its single volumes do); CC channels staged at a time
  if (g.xa && !conv3_lds_ok(g)) return fail(AWK_ERR_INVALID, "unet: in-place concatenation needs the LDS form");
  const int CC = lds_chunk(g);
  if (conv3_lds_ok(g)) {
    // output-channel tiles per wave: the largest of 4, 2, 1 dividing the tile
    // count (so that blockIdx.y * MT * 16 covers exactly cout) that still gives
    // 1024 blocks, 4 per CU, where the grid allows (MT changes no result)
    const long long bricks = N / (kLdsBX * kLdsBY * kLdsBZ);
    int mt = mt_all % 4 == 0 ? 4 : mt_all % 2 == 0 ? 2 : 1;
    while (mt > 1 && bricks * (mt_all / mt) < 1024) mt >>= 1;
    if ((mt_all / mt) * mt * 16 == g.cout) {
      const int ksteps = (27 * (CC / 8) + 3) / 4;
      const size_t lds = (size_t)kLdsHalo * (CC + 8) * sizeof(bf16) + (size_t)ksteps * 4 * sizeof(int2);
      static bool attr_set = false;  // dynamic LDS above 64 KiB (CC = 48: 71 KiB)
      if (!attr_set) {
        const int mx = kLdsHalo * (48 + 8) * (int)sizeof(bf16) + 164 * (int)sizeof(int2);
        const void *fns[] = {reinterpret_cast<const void *>(&k_conv3_lds<1, false>),
                             reinterpret_cast<const void *>(&k_conv3_lds<2, false>),
                             reinterpret_cast<const void *>(&k_conv3_lds<4, false>),
                             reinterpret_cast<const void *>(&k_conv3_lds<1, true>),
                             reinterpret_cast<const void *>(&k_conv3_lds<2, true>),
                             reinterpret_cast<const void *>(&k_conv3_lds<4, true>)};
        for (const void *f : fns)
          AWK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        attr_set = true;
      }
      const dim3 grid((unsigned)(N / (kLdsBX * kLdsBY * kLdsBZ)), (unsigned)(mt_all / mt));
#define AWK_CL(M, P) AWK_LAUNCH((k_conv3_lds<M, P>), grid, dim3(kConvBlock), lds, s, g, CC)
      // software pipelining pays where a K step is long enough to hide the
      // prefetch registers' occupancy cost: chunks of 32 or more channels
      // (measured per layer at 128^3: 96 -> 32 150 -> 118 us, 16 -> 16 109 -> 123 us)
      const bool pipe = tuning().unet_pipe == 2 ? CC >= 32 : tuning().unet_pipe == 1;
      if (pipe) {
        if (mt == 1) AWK_CL(1, true); else if (mt == 2) AWK_CL(2, true); else AWK_CL(4, true);
      } else {
        if (mt == 1) AWK_CL(1, false); else if (mt == 2) AWK_CL(2, false); else AWK_CL(4, false);
      }
#undef AWK_CL
      AWK_LAUNCH_CHECK("k_conv3_lds");
      return AWK_OK;
    }
  }
  // channel tiles per wave: the largest power of two (<= 8) dividing the tile count
  int mt_max = 1;
  while (mt_max < 8 && mt_all % (mt_max * 2) == 0) mt_max *= 2;
  const long long target = 8192;
  auto tiles = [&](long long n, int p) { return (n + kPosPerWave * p - 1) / (kPosPerWave * p); };
  // Split K or not is decided from ONE volume's size: it fixes each output's
  // summation order, so a batch gives bit for bit the outputs of single volumes.
  // The tile shape (which outputs a wave holds) is then sized for the whole batch.
  int pt = mt_max >= 8 ? 2 : 4;
  while (tiles(N1, pt) * (mt_all / mt_max) < target && pt > 1) pt >>= 1;
  const int sk = tiles(N1, pt) * (mt_all / mt_max) < target ? 4 : 1;
  int mt = mt_max;
  if (sk == 1) {
    pt = mt_max >= 8 ? 2 : 4;
    while (tiles(N, pt) * (mt_all / mt) < target && pt > 1) pt >>= 1;
  } else {
    pt = 1;
    mt = mt_max > 4 ? 4 : mt_max;
    while (tiles(N, 1) * 4 * (mt_all / mt) < target && mt > 1) mt >>= 1;
  }
  const long long per_block = sk > 1 ? (long long)kPosPerWave * pt : (long long)kConvWaves * kPosPerWave * pt;
  const dim3 grid((unsigned)((N + per_block - 1) / per_block), (unsigned)(mt_all / mt));
#define AWK_CV(M, P, S) AWK_LAUNCH((k_conv3<M, P, S>), grid, dim3(kConvBlock), 0, s, g)
  if (sk == 1) {
    if (mt == 1) { if (pt == 4) AWK_CV(1, 4, 1); else if (pt == 2) AWK_CV(1, 2, 1); else AWK_CV(1, 1, 1); }
    else if (mt == 2) { if (pt == 4) AWK_CV(2, 4, 1); else if (pt == 2) AWK_CV(2, 2, 1); else AWK_CV(2, 1, 1); }
    else if (mt == 4) { if (pt == 4) AWK_CV(4, 4, 1); else if (pt == 2) AWK_CV(4, 2, 1); else AWK_CV(4, 1, 1); }
    else { if (pt == 2) AWK_CV(8, 2, 1); else AWK_CV(8, 1, 1); }
  } else {
    if (mt == 1) AWK_CV(1, 1, 4);
    else if (mt == 2) AWK_CV(2, 1, 4);
    else AWK_CV(4, 1, 4);
  }
#undef AWK_CV
  AWK_LAUNCH_CHECK("k_conv3");
  return AWK_OK;
}

}  // namespace awk

using namespace awk;

static void unet_free(awk_unet *u) {
  for (int l = 0; l < 14; ++l) {
    if (u->w[l]) (void)hipFree(u->w[l]);
    if (u->bias[l]) (void)hipFree(u->bias[l]);
  }
  void *b[] = {u->head_w, u->head_b, u->x0, u->t0, u->s0, u->p0, u->t1, u->s1, u->p1, u->t2, u->s2,
               u->p2, u->t3, u->s3, u->c2, u->d2a, u->d2, u->c1, u->d1a, u->d1, u->c0, u->d0a, u->d0};
  for (void *p : b)
    if (p) (void)hipFree(p);
}

// repack OIDHW fp32 (blob_cin inputs) into [cout][kpad] bf16, k = tap*cin + ci
static std::vector<bf16> pack_weights(const float *w, int cout, int cin, int blob_cin, int kpad) {
  std::vector<bf16> wt((size_t)cout * kpad, (bf16)0.f);
  for (int co = 0; co < cout; ++co)
    for (int ci = 0; ci < blob_cin; ++ci)
      for (int tap = 0; tap < 27; ++tap)
        wt[(size_t)co * kpad + tap * cin + ci] = (bf16)w[((size_t)co * blob_cin + ci) * 27 + tap];
  return wt;
}

static ConvArgs conv_args(const awk_unet *u, int l, const bf16 *x, bf16 *out, int nx, int ny,
                          int nz, int nb) {
  ConvArgs g;
  g.x = x;
  g.w = u->w[l];
  g.bias = u->bias[l];
  g.dact = nullptr;
  g.out = out;
  g.C = kUnetLayers[l].cin;
  g.cout = kUnetLayers[l].cout;
  g.kpad = u->kpad[l];
  g.nchunks = 27 * (g.C / 8);
  g.c8magic = g.C > 8 ? (unsigned)(4294967296ull / (unsigned)(g.C / 8)) + 1u : 0u;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.batch = nb;
  g.slope = kUnetSlope;
  g.xa = nullptr;
  g.ca = 0;
  return g;
}

static awk_status conv(const awk_unet *u, int l, const bf16 *x, bf16 *out, int nx, int ny, int nz,
                       int nb, hipStream_t s) {
  return launch_conv3(conv_args(u, l, x, out, nx, ny, nz, nb), s);
}

// Decoder layer l on [nearest x2 upsampling of a (ca channels), b (cb channels)]:
// read in place by the LDS form unless the concatenation must be kept (training)
// or the grid takes the direct form; else written to cat first.
static awk_status conv_cat(const awk_unet *u, int l, const bf16 *a, int ca, const bf16 *b, int cb,
                           bf16 *cat, bf16 *out, int nx, int ny, int nz, int nb, bool keep,
                           hipStream_t s) {
  ConvArgs g = conv_args(u, l, b, out, nx, ny, nz, nb);
  g.xa = a;
  g.ca = ca;
  if (!keep && conv3_lds_ok(g)) return launch_conv3(g, s);
  awk_status st = unet_upcat(cat, a, ca, b, cb, nx, ny, nz * nb, s);
  if (st) return st;
  return conv(u, l, cat, out, nx, ny, nz, nb, s);
}

awk_status awk::unet_pool(const bf16 *in, bf16 *out, int nx, int ny, int nz, int C, hipStream_t s) {
  const long long total = (long long)(nx / 2) * (ny / 2) * (nz / 2) * (C / 8);
  AWK_LAUNCH(k_pool2, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, out, in, nx, ny, nz, C);
  AWK_LAUNCH_CHECK("k_pool2");
  return AWK_OK;
}

awk_status awk::unet_upcat(bf16 *out, const bf16 *a, int ca, const bf16 *b, int cb, int nx, int ny,
                           int nz, hipStream_t s) {
  const long long total = (long long)nx * ny * nz * ((ca + cb) / 8);
  AWK_LAUNCH(k_upcat, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, out, a, ca, b, cb,
             nx, ny, nz);
  AWK_LAUNCH_CHECK("k_upcat");
  return AWK_OK;
}

awk_status awk::unet_forward(awk_unet *u, int nb, const float *const *templ,
                             const float *const *subject, float *const *disp, hipStream_t s,
                             bool keep_concat) {
  const int nx = u->nx, ny = u->ny, nz = u->nz;
  const int N0 = nx * ny * nz;
  if (nb < 1 || nb > u->max_batch) return fail(AWK_ERR_INVALID, "unet: batch exceeds the network's");
  awk_status st;
  for (int b = 0; b < nb; ++b) {
    AWK_LAUNCH(k_pack_input, dim3((N0 + 255) / 256), dim3(256), 0, s, u->x0 + (size_t)b * N0 * 8,
               templ[b], subject[b], N0);
    AWK_LAUNCH_CHECK("k_pack_input");
  }
  // the nb volumes are stacked along z: every pooling and upsampling pairs planes
  // of one volume, since each level's nz is even
  const int a1 = nx / 2, b1 = ny / 2, c1 = nz / 2, a2 = nx / 4, b2 = ny / 4, c2 = nz / 4,
            a3 = nx / 8, b3 = ny / 8, c3 = nz / 8;
  if ((st = conv(u, 0, u->x0, u->t0, nx, ny, nz, nb, s))) return st;
  if ((st = conv(u, 1, u->t0, u->s0, nx, ny, nz, nb, s))) return st;
  if ((st = unet_pool(u->s0, u->p0, nx, ny, nz * nb, 16, s))) return st;
  if ((st = conv(u, 2, u->p0, u->t1, a1, b1, c1, nb, s))) return st;
  if ((st = conv(u, 3, u->t1, u->s1, a1, b1, c1, nb, s))) return st;
  if ((st = unet_pool(u->s1, u->p1, a1, b1, c1 * nb, 32, s))) return st;
  if ((st = conv(u, 4, u->p1, u->t2, a2, b2, c2, nb, s))) return st;
  if ((st = conv(u, 5, u->t2, u->s2, a2, b2, c2, nb, s))) return st;
  if ((st = unet_pool(u->s2, u->p2, a2, b2, c2 * nb, 64, s))) return st;
  if ((st = conv(u, 6, u->p2, u->t3, a3, b3, c3, nb, s))) return st;
  if ((st = conv(u, 7, u->t3, u->s3, a3, b3, c3, nb, s))) return st;
  const bool kc = keep_concat;
  if ((st = conv_cat(u, 8, u->s3, 128, u->s2, 64, u->c2, u->d2a, a2, b2, c2, nb, kc, s))) return st;
  if ((st = conv(u, 9, u->d2a, u->d2, a2, b2, c2, nb, s))) return st;
  if ((st = conv_cat(u, 10, u->d2, 64, u->s1, 32, u->c1, u->d1a, a1, b1, c1, nb, kc, s))) return st;
  if ((st = conv(u, 11, u->d1a, u->d1, a1, b1, c1, nb, s))) return st;
  if ((st = conv_cat(u, 12, u->d1, 32, u->s0, 16, u->c0, u->d0a, nx, ny, nz, nb, kc, s))) return st;
  if ((st = conv(u, 13, u->d0a, u->d0, nx, ny, nz, nb, s))) return st;
  for (int b = 0; b < nb; ++b) {
    AWK_LAUNCH(k_head, dim3((N0 + 255) / 256), dim3(256), 0, s, disp[3 * b], disp[3 * b + 1],
               disp[3 * b + 2], u->d0 + (size_t)b * N0 * 16, u->head_w, u->head_b, N0, 16);
    AWK_LAUNCH_CHECK("k_head");
  }
  return AWK_OK;
}


extern "C" {

size_t awk_unet_weight_count(void) {
  size_t n = 0;
  for (int l = 0; l < 14; ++l) n += (size_t)kUnetLayers[l].cout * kUnetBlobCin[l] * 27 + kUnetLayers[l].cout;
  return n + 3 * 16 + 3;
}

awk_status awk_unet_create(awk_unet_t *out, int nx, int ny, int nz, const float *weights,
                           size_t n_weights) {
  return awk_unet_create_batch(out, nx, ny, nz, 1, weights, n_weights);
}

awk_status awk_unet_create_batch(awk_unet_t *out, int nx, int ny, int nz, int max_batch,
                                 const float *weights, size_t n_weights) {
  if (!out || !weights) return fail(AWK_ERR_INVALID, "awk_unet_create: null argument");
  *out = nullptr;
  if (max_batch < 1 || max_batch > 64) return fail(AWK_ERR_INVALID, "awk_unet_create: batch must be 1..64");
  if (nx < 8 || ny < 8 || nz < 8 || nx % 8 || ny % 8 || nz % 8)
    return fail(AWK_ERR_INVALID, "awk_unet_create: every axis must be a multiple of 8 (three poolings)");
  if ((long long)nx * ny * nz > (1LL << 24)) return fail(AWK_ERR_INVALID, "awk_unet_create: at most 2^24 voxels");
  if ((long long)nx * ny * nz * max_batch > (1LL << 26))
    return fail(AWK_ERR_INVALID, "awk_unet_create: at most 2^26 voxels over the batch");
  if (n_weights != awk_unet_weight_count())
    return fail(AWK_ERR_INVALID, "awk_unet_create: weight count does not match the network");
  awk_unet *u = new (std::nothrow) awk_unet();
  if (!u) return fail(AWK_ERR_ALLOC, "awk_unet_create: host allocation failed");
  u->nx = nx;
  u->ny = ny;
  u->nz = nz;
  u->max_batch = max_batch;
  hipError_t e = hipSuccess;
  size_t off = 0;
  for (int l = 0; l < 14 && e == hipSuccess; ++l) {
    const LayerSpec &L = kUnetLayers[l];
    const int bc = kUnetBlobCin[l];
    u->kpad[l] = (27 * L.cin + 31) / 32 * 32;
    const std::vector<bf16> wt = pack_weights(weights + off, L.cout, L.cin, bc, u->kpad[l]);
    off += (size_t)L.cout * bc * 27;
    if (e == hipSuccess) e = hipMalloc(&u->w[l], wt.size() * sizeof(bf16));
    if (e == hipSuccess) e = hipMemcpy(u->w[l], wt.data(), wt.size() * sizeof(bf16), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&u->bias[l], L.cout * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(u->bias[l], weights + off, L.cout * sizeof(float), hipMemcpyHostToDevice);
    off += L.cout;
  }
  if (e == hipSuccess) e = hipMalloc(&u->head_w, 48 * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(u->head_w, weights + off, 48 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&u->head_b, 3 * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(u->head_b, weights + off + 48, 3 * sizeof(float), hipMemcpyHostToDevice);
  const size_t N0 = (size_t)nx * ny * nz * max_batch, N1 = N0 / 8, N2 = N1 / 8, N3 = N2 / 8;
  struct {
    bf16 **p;
    size_t n;
  } bufs[] = {{&u->x0, N0 * 8},   {&u->t0, N0 * 16},  {&u->s0, N0 * 16},  {&u->p0, N1 * 16},
              {&u->t1, N1 * 32},  {&u->s1, N1 * 32},  {&u->p1, N2 * 32},  {&u->t2, N2 * 64},
              {&u->s2, N2 * 64},  {&u->p2, N3 * 64},  {&u->t3, N3 * 128}, {&u->s3, N3 * 128},
              {&u->c2, N2 * 192}, {&u->d2a, N2 * 64}, {&u->d2, N2 * 64},  {&u->c1, N1 * 96},
              {&u->d1a, N1 * 32}, {&u->d1, N1 * 32},  {&u->c0, N0 * 48},  {&u->d0a, N0 * 16},
              {&u->d0, N0 * 16}};
  for (auto &b : bufs)
    if (e == hipSuccess) e = hipMalloc((void **)b.p, b.n * sizeof(bf16));
  if (e != hipSuccess) {
    awk_status st = hip_check(e, "awk_unet_create: allocation");
    unet_free(u);
    delete u;
    return st;
  }
  *out = u;
  return AWK_OK;
}

awk_status awk_unet_destroy(awk_unet_t u) {
  if (!u) return AWK_OK;
  unet_free(u);
  delete u;
  return AWK_OK;
}

awk_status awk_unet_forward(awk_unet_t u, const float *templ, const float *subject,
                            float *const disp[3], awk_stream_t stream) {
  if (!u || !templ || !subject || !disp || !disp[0] || !disp[1] || !disp[2])
    return fail(AWK_ERR_INVALID, "awk_unet_forward: null argument");
  return unet_forward(u, 1, &templ, &subject, disp, as_stream(stream));
}

awk_status awk_unet_forward_batch(awk_unet_t u, int batch, const float *const templ[],
                                  const float *const subject[], float *const disp[],
                                  awk_stream_t stream) {
  if (!u || !templ || !subject || !disp) return fail(AWK_ERR_INVALID, "awk_unet_forward_batch: null argument");
  if (batch < 1 || batch > u->max_batch)
    return fail(AWK_ERR_INVALID, "awk_unet_forward_batch: batch must be 1..max_batch of awk_unet_create_batch");
  for (int b = 0; b < batch; ++b)
    if (!templ[b] || !subject[b] || !disp[3 * b] || !disp[3 * b + 1] || !disp[3 * b + 2])
      return fail(AWK_ERR_INVALID, "awk_unet_forward_batch: null volume pointer");
  return unet_forward(u, batch, templ, subject, disp, as_stream(stream));
}

awk_status awk_unet_mfma_probe(float *out, awk_stream_t stream) {
  if (!out) return fail(AWK_ERR_INVALID, "awk_unet_mfma_probe: null pointer");
  AWK_LAUNCH(k_mfma_probe, dim3(1), dim3(64), 0, as_stream(stream), out);
  AWK_LAUNCH_CHECK("k_mfma_probe");
  return AWK_OK;
}

}  // extern "C"
