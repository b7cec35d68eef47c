// awk_unet_train.hip -- training of the U-Net warm start on the device.
//
// One step = forward (awk_unet.hip, activations kept), loss L = mean over voxels
// and components of (u - u_target)^2, backward, Adam on fp32 master weights, and
// a repack of the bf16 weights the convolutions read.  The backward pass:
//   dgrad  the same implicit-GEMM MFMA kernel as the forward convolution, run on
//          the output gradient with the weights flipped in space and transposed
//          in channels, with the previous activation's LeakyReLU derivative
//          folded into its epilogue;
//   wgrad  dW[co][ci][tap] = sum_p dZ[p][co] X[p + tap][ci]: MFMAs with K =
//          positions over 3-D bricks whose dZ and X halo are staged once for all
//          27 taps (k_wgrad below); block partial sums are added in a fixed order;
//   pooling, upsampling + concatenation, the head and the loss have their own
//   small kernels.
// Gradients are stored in bf16 between layers and accumulated in fp32.  Every
// reduction runs in a fixed order, so a training step is reproducible bit for bit.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "awk_unet.h"

namespace awk {

// Weight gradient.  A block takes a brick of 16 x 4 x 4 output positions and
// stages, once per brick, dZ for its output-channel tiles ([tile][256 positions][16])
// and the 18 x 6 x 6 halo of one 16-channel tile of X ([648 positions][16]), both
// channels-last as in memory.  The MFMA operands need 8 consecutive positions of one
// channel per lane; ds_read_b64_tr_b16 delivers exactly that from the channels-last
// rows (within each group of 16 lanes, lane l receives element l % 4 of the 8 bytes
// addressed by lanes 4j + l / 4, j = 0..3), and because every lane supplies its own
// row address, a tap's shift is only a different row: no staging per tap.
// Wave w accumulates taps w, w + 4, ... for all positions of the brick; each block
// walks a strided list of bricks and stores its partial sums once, and
// k_wgrad_reduce adds the blocks' partials in a fixed order.
constexpr int kWbX = 16, kWbY = 4, kWbZ = 4;
constexpr int kWbP = kWbX * kWbY * kWbZ;  // 256 positions = 8 K steps of 32
constexpr int kWhX = kWbX + 2, kWhY = kWbY + 2, kWhZ = kWbZ + 2;
constexpr int kWhP = kWhX * kWhY * kWhZ;  // 648 halo positions
constexpr long long kWgWsFloats = 8LL << 20;  // partial sums per layer (32 MiB)
constexpr int kWgMaxBlocks = 2048;        // position blocks per group
constexpr int kWgSumRows = 32;           // partial rows per thread in the first reduction stage

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;

// 8 positions (rows r and r + 4 of the lane's K group) of one channel, transposed
// out of channels-last LDS rows of 16 channels (32 bytes).
__device__ __forceinline__ bf16x8 tr_rows(const bf16 *p) {
  const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(p));
  const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(p + 4 * 16));
  const short __attribute__((ext_vector_type(8))) v =
      __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, v);
}

struct WgradArgs {
  const bf16 *dz;  // [N][CO]
  const bf16 *x;   // [N][CI]
  float *ws;       // partials [nb][G][E], then bias partials [nb][CO]
  int CO, CI;
  int nx, ny, nz;
  int bxn, byn;    // bricks along x and y
  int nbricks;
  int citiles;     // G = (CO / (16 COT)) * citiles; blockIdx.y = cog * citiles + cit
};

template <int COT>
__global__ void __launch_bounds__(256)
k_wgrad(WgradArgs g) {
  extern __shared__ bf16 lds[];
  bf16 *xl = lds;               // [kWhP][16]
  bf16 *zl = lds + kWhP * 16;   // [COT][kWbP][16]
  constexpr int E = 27 * COT * 16 * 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane & 15, grp = lane >> 4;
  const int cit = (int)blockIdx.y % g.citiles, cog = (int)blockIdx.y / g.citiles;
  const int ci0 = cit * 16, co0 = cog * COT * 16;
  const int ntap = wave < 3 ? 7 : 6;
  // per-lane row addresses (elements): K index k = grp*8 + h*4 + (t>>2), channel quad t&3
  const int a_lane = (grp * 8 + (t >> 2)) * 16 + 4 * (t & 3);
  const int b_lane = ((1 * kWhY + (grp >> 1) + 1) * kWhX + (grp & 1) * 8 + (t >> 2) + 1) * 16 + 4 * (t & 3);
  int boff[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const int tap = wave + 4 * i < 27 ? wave + 4 * i : 13;
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    boff[i] = b_lane + ((dz * kWhY + dy) * kWhX + dx) * 16;
  }
  f32x4 acc[7][COT];
#pragma unroll
  for (int i = 0; i < 7; ++i)
#pragma unroll
    for (int c = 0; c < COT; ++c) acc[i][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum[COT];
#pragma unroll
  for (int c = 0; c < COT; ++c) bsum[c] = 0.f;
  const bool do_bias = cit == 0 && wave == 0;
  for (int br = (int)blockIdx.x; br < g.nbricks; br += (int)gridDim.x) {
    const int bxi = br % g.bxn, rr = br / g.bxn, byi = rr % g.byn, bzi = rr / g.byn;
    const int x0 = bxi * kWbX, y0 = byi * kWbY, z0 = bzi * kWbZ;
    __syncthreads();  // the previous brick's reads are done
    for (int w = threadIdx.x; w < kWhP * 2; w += 256) {
      const int hp = w >> 1, h8 = w & 1;
      const int hx = hp % kWhX, r = hp / kWhX, hy = r % kWhY, hz = r / kWhY;
      const int X = x0 + hx - 1, Y = y0 + hy - 1, Z = z0 + hz - 1, c = ci0 + h8 * 8;
      bf16x8 v = zero8();
      if (c < g.CI && X >= 0 && X < g.nx && Y >= 0 && Y < g.ny && Z >= 0 && Z < g.nz)
        v = *reinterpret_cast<const bf16x8 *>(g.x + ((size_t)(Z * g.ny + Y) * g.nx + X) * g.CI + c);
      *reinterpret_cast<bf16x8 *>(xl + hp * 16 + h8 * 8) = v;
    }
    for (int w = threadIdx.x; w < kWbP * COT * 2; w += 256) {
      const int p = w / (COT * 2), rc = w - p * (COT * 2), c = rc >> 1, h8 = rc & 1;
      const int X = x0 + (p & 15), Y = y0 + ((p >> 4) & 3), Z = z0 + (p >> 6);
      bf16x8 v = zero8();
      if (X < g.nx && Y < g.ny && Z < g.nz)
        v = *reinterpret_cast<const bf16x8 *>(g.dz + ((size_t)(Z * g.ny + Y) * g.nx + X) * g.CO + co0 +
                                               c * 16 + h8 * 8);
      *reinterpret_cast<bf16x8 *>(zl + (c * kWbP + p) * 16 + h8 * 8) = v;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      // K step s: positions 32 s + k, i.e. z = s/2, y = 2 (s%2) + k/16, x = k%16
      bf16x8 a[COT];
#pragma unroll
      for (int c = 0; c < COT; ++c) a[c] = tr_rows(zl + c * kWbP * 16 + s * 32 * 16 + a_lane);
      if (do_bias) {
#pragma unroll
        for (int c = 0; c < COT; ++c)
#pragma unroll
          for (int j = 0; j < 8; ++j) bsum[c] += (float)a[c][j];
      }
      const int soff = ((s >> 1) * kWhY + 2 * (s & 1)) * kWhX * 16;
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        if (i >= ntap) break;
        const bf16x8 b = tr_rows(xl + boff[i] + soff);
#pragma unroll
        for (int c = 0; c < COT; ++c)
          acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c], b, acc[i][c], 0, 0, 0);
      }
    }
  }
  // partials: element (tap, co_l, ci_l) at (tap * 16 COT + co_l) * 16 + ci_l; lane (t, grp)
  // holds co_l = c*16 + grp*4 + r, ci_l = t
  const int G = (int)gridDim.y;
  float *ws = g.ws + ((size_t)blockIdx.x * G + blockIdx.y) * E;
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (i >= ntap) break;
    const int tap = wave + 4 * i;
#pragma unroll
    for (int c = 0; c < COT; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) ws[(tap * COT * 16 + c * 16 + grp * 4 + r) * 16 + t] = acc[i][c][r];
  }
  if (do_bias) {
    float *wb = g.ws + (size_t)gridDim.x * G * E + (size_t)blockIdx.x * g.CO + co0;
#pragma unroll
    for (int c = 0; c < COT; ++c) {
      float v = bsum[c];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (grp == 0) wb[c * 16 + t] = v;
    }
  }
}

// Partial sums, nb rows of ne weight elements then nb rows of CO bias elements,
// reduced in fixed order.  k_wgrad_sum adds row ranges [y*per, (y+1)*per) into
// row y of dst (same layout, gridDim.y rows); k_wgrad_reduce adds all rows into
// dW[co][ci][tap] and db.
__global__ void __launch_bounds__(256)
k_wgrad_sum(const float *__restrict__ src, float *__restrict__ dst, int nb, int per, int ne, int CO) {
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= ne + CO) return;
  const int y = (int)blockIdx.y, nr = (int)gridDim.y;
  const int b0 = y * per, b1 = min(nb, b0 + per);
  float s = 0.f;
  if (e < ne) {
    for (int b = b0; b < b1; ++b) s += src[(size_t)b * ne + e];
    dst[(size_t)y * ne + e] = s;
  } else {
    const int co = e - ne;
    for (int b = b0; b < b1; ++b) s += src[(size_t)nb * ne + (size_t)b * CO + co];
    dst[(size_t)nr * ne + (size_t)y * CO + co] = s;
  }
}

__global__ void __launch_bounds__(256)
k_wgrad_reduce(const float *__restrict__ ws, float *__restrict__ dw, float *__restrict__ db, int nb,
               int G, int cot, int citiles, int CO, int CIb) {
  const int E = 27 * cot * 256;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ne = (long long)G * E;
  if (id < ne) {
    const int gi = (int)(id / E), e = (int)(id - (long long)gi * E);
    const int tap = e / (cot * 256), r = e - tap * cot * 256, co_l = r >> 4, ci_l = r & 15;
    const int ci = (gi % citiles) * 16 + ci_l, co = (gi / citiles) * cot * 16 + co_l;
    if (ci >= CIb) return;
    float s = 0.f;
    for (int b = 0; b < nb; ++b) s += ws[(size_t)b * ne + id];
    dw[((size_t)co * CIb + ci) * 27 + tap] += s;
  } else if (id < ne + CO) {
    const int co = (int)(id - ne);
    const float *wb = ws + (size_t)nb * ne;
    float s = 0.f;
    for (int b = 0; b < nb; ++b) s += wb[(size_t)b * CO + co];
    db[co] += s;
  }
}

// Loss and head backward: du = 2 (u - t) / (3N); dZ13 = (Wh^T du) * f'(d0).  The
// 48 + 3 head-gradient sums and the loss are reduced per block (shuffles, then
// LDS) into part[block][52]; k_head_reduce adds the blocks in order.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(256)
k_head_bwd(bf16 *__restrict__ dz13, float *__restrict__ part, const float *__restrict__ u0,
           const float *__restrict__ u1, const float *__restrict__ u2, const float *__restrict__ t0,
           const float *__restrict__ t1, const float *__restrict__ t2, const bf16 *__restrict__ d0,
           const float *__restrict__ wh, int N, float slope) {
  __shared__ float sred[4][52];
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const bool live = i < N;
  float l = 0.f;
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  float a[16];
  if (live) {
    const float inv = 2.f / (3.f * (float)N);
    const float e0 = u0[i] - t0[i], e1 = u1[i] - t1[i], e2 = u2[i] - t2[i];
    l = (e0 * e0 + e1 * e1 + e2 * e2) / (3.f * (float)N);
    g0 = e0 * inv;
    g1 = e1 * inv;
    g2 = e2 * inv;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bf16x8 dv = *reinterpret_cast<const bf16x8 *>(d0 + (size_t)i * 16 + h * 8);
      bf16x8 o;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int c = h * 8 + j;
        a[c] = (float)dv[j];
        const float gd = wh[c] * g0 + wh[16 + c] * g1 + wh[32 + c] * g2;
        o[j] = (bf16)(a[c] > 0.f ? gd : gd * slope);
      }
      *reinterpret_cast<bf16x8 *>(dz13 + (size_t)i * 16 + h * 8) = o;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 16; ++c) a[c] = 0.f;
  }
  const int wave = threadIdx.x >> 6;
  const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const float s0 = wave_sum(g0 * a[c]), s1 = wave_sum(g1 * a[c]), s2 = wave_sum(g2 * a[c]);
    if (lead) {
      sred[wave][c] = s0;
      sred[wave][16 + c] = s1;
      sred[wave][32 + c] = s2;
    }
  }
  const float b0 = wave_sum(g0), b1 = wave_sum(g1), b2 = wave_sum(g2), ls = wave_sum(l);
  if (lead) {
    sred[wave][48] = b0;
    sred[wave][49] = b1;
    sred[wave][50] = b2;
    sred[wave][51] = ls;
  }
  __syncthreads();
  if (threadIdx.x < 52)
    part[(size_t)blockIdx.x * 52 + threadIdx.x] =
        sred[0][threadIdx.x] + sred[1][threadIdx.x] + sred[2][threadIdx.x] + sred[3][threadIdx.x];
}

// One block per column c of the partials: a fixed-order FP64 tree over the
// blocks; dWh, dbh into the gradient, the loss into *loss.
__global__ void __launch_bounds__(256)
k_head_reduce(float *__restrict__ dwh, float *__restrict__ dbh, double *__restrict__ loss,
              const float *__restrict__ part, int nblocks) {
  __shared__ double sm[256];
  const int c = (int)blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) s += part[(size_t)b * 52 + c];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (c < 48) dwh[c] += (float)sm[0];
  else if (c < 51) dbh[c - 48] += (float)sm[0];
  else *loss += sm[0];
}

// Backward of [up(A), B]: dA[q][c] = (sum over the 8 children of dCat[child][c]) *
// f'(A[q][c]) for the upsampled part (the skip part is read in place later).
__global__ void __launch_bounds__(256)
k_upcat_bwd(bf16 *__restrict__ dza, const bf16 *__restrict__ dcat, const bf16 *__restrict__ a,
            int ca, int C, int nx, int ny, int nz, float slope) {
  const int hx = nx >> 1, hy = ny >> 1, hz = nz >> 1, groups = ca >> 3;
  const long long total = (long long)hx * hy * hz * groups;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  const int gi = (int)(id % groups);
  const long long q = id / groups;
  const int x = (int)(q % hx), r = (int)(q / hx), y = r % hy, z = r / hy;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const int X = 2 * x + (d & 1), Y = 2 * y + ((d >> 1) & 1), Z = 2 * z + (d >> 2);
    const bf16x8 v = *reinterpret_cast<const bf16x8 *>(dcat + ((size_t)(Z * ny + Y) * nx + X) * C + gi * 8);
#pragma unroll
    for (int c = 0; c < 8; ++c) s[c] += (float)v[c];
  }
  const bf16x8 av = *reinterpret_cast<const bf16x8 *>(a + (size_t)q * ca + gi * 8);
  bf16x8 o;
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = (bf16)((float)av[c] > 0.f ? s[c] : s[c] * slope);
  *reinterpret_cast<bf16x8 *>(dza + (size_t)q * ca + gi * 8) = o;
}

// Backward of pooling plus the skip path: dZ[p][c] = (dP[parent][c] / 8 +
// skip[p * sstride + soff + c]) * f'(Y[p][c]).
__global__ void __launch_bounds__(256)
k_pool_bwd(bf16 *__restrict__ dz, const bf16 *__restrict__ dp, const bf16 *__restrict__ skip,
           int sstride, int soff, const bf16 *__restrict__ y, int C, int nx, int ny, int nz,
           float slope) {
  const int groups = C >> 3;
  const long long total = (long long)nx * ny * nz * groups;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  const int gi = (int)(id % groups);
  const long long p = id / groups;
  const int x = (int)(p % nx), r = (int)(p / nx), yy = r % ny, z = r / ny;
  const long long parent = ((long long)(z >> 1) * (ny >> 1) + (yy >> 1)) * (nx >> 1) + (x >> 1);
  const bf16x8 g = *reinterpret_cast<const bf16x8 *>(dp + parent * C + gi * 8);
  const bf16x8 k = *reinterpret_cast<const bf16x8 *>(skip + p * sstride + soff + gi * 8);
  const bf16x8 a = *reinterpret_cast<const bf16x8 *>(y + p * C + gi * 8);
  bf16x8 o;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float v = (float)g[c] * 0.125f + (float)k[c];
    o[c] = (bf16)((float)a[c] > 0.f ? v : v * slope);
  }
  *reinterpret_cast<bf16x8 *>(dz + p * C + gi * 8) = o;
}

// Adam (fp32 master weights); the gradient is cleared for the next step.
__global__ void __launch_bounds__(256)
k_adam(float *__restrict__ w, float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
       long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  g[i] = 0.f;
}

// fp32 OIDHW weights -> the forward layout W[co][tap*cin + ci] and the gradient
// layout Wd[ci][(26 - tap)*cout + co], both bf16.
__global__ void __launch_bounds__(256)
k_repack(bf16 *__restrict__ wf, bf16 *__restrict__ wd, const float *__restrict__ w, int cout,
         int cin, int cinb, int kpad, int kpad_d) {
  const int n = cout * cinb * 27;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const int tap = i % 27, r = i / 27, ci = r % cinb, co = r / cinb;
  const bf16 v = (bf16)w[i];
  wf[(size_t)co * kpad + tap * cin + ci] = v;
  if (wd) wd[(size_t)ci * kpad_d + (26 - tap) * cout + co] = v;
}

}  // namespace awk

using namespace awk;

struct awk_unet_trainer {
  awk_unet *net = nullptr;
  size_t nw = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32, blob layout
  size_t woff[14] = {}, boff[14] = {}, hoff = 0;
  bf16 *wd[14] = {};   // gradient-layout weights (layers 1..13)
  int kpad_d[14] = {};
  long long t = 0;
  // gradient tensors
  bf16 *dz13 = nullptr, *dz12 = nullptr, *dc0 = nullptr, *dz11 = nullptr, *dz10 = nullptr,
       *dc1 = nullptr, *dz9 = nullptr, *dz8 = nullptr, *dc2 = nullptr, *dz7 = nullptr,
       *dz6 = nullptr, *dp2 = nullptr, *dz5 = nullptr, *dz4 = nullptr, *dp1 = nullptr,
       *dz3 = nullptr, *dz2 = nullptr, *dp0 = nullptr, *dz1 = nullptr, *dz0 = nullptr;
  float *u = nullptr;     // 3 planes: the forward output
  float *hpart = nullptr; // head-gradient block partials [blocks][52]
  float *ws = nullptr;    // weight-gradient block partials (kWgWsFloats + bias partials)
  float *ws2 = nullptr;   // their first reduction stage
};

static void trainer_free(awk_unet_trainer *t) {
  void *b[] = {t->w, t->g, t->m, t->v, t->dz13, t->dz12, t->dc0, t->dz11, t->dz10, t->dc1, t->dz9,
               t->dz8, t->dc2, t->dz7, t->dz6, t->dp2, t->dz5, t->dz4, t->dp1, t->dz3, t->dz2,
               t->dp0, t->dz1, t->dz0, t->u, t->hpart, t->ws, t->ws2};
  for (void *p : b)
    if (p) (void)hipFree(p);
  for (int l = 0; l < 14; ++l)
    if (t->wd[l]) (void)hipFree(t->wd[l]);
}

static awk_status repack_all(awk_unet_trainer *t, hipStream_t s) {
  awk_unet *u = t->net;
  for (int l = 0; l < 14; ++l) {
    const LayerSpec &L = kUnetLayers[l];
    const int n = L.cout * kUnetBlobCin[l] * 27;
    AWK_LAUNCH(k_repack, dim3((n + 255) / 256), dim3(256), 0, s, u->w[l], t->wd[l], t->w + t->woff[l],
               L.cout, L.cin, kUnetBlobCin[l], u->kpad[l], t->kpad_d[l]);
    AWK_HIP(hipMemcpyAsync(u->bias[l], t->w + t->boff[l], L.cout * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  AWK_HIP(hipMemcpyAsync(u->head_w, t->w + t->hoff, 48 * sizeof(float), hipMemcpyDeviceToDevice, s));
  AWK_HIP(hipMemcpyAsync(u->head_b, t->w + t->hoff + 48, 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
  AWK_LAUNCH_CHECK("k_repack");
  return AWK_OK;
}

// dX = conv^T(dZ) of layer l: the output has cin_l channels, scaled by the
// derivative of `act` (the layer's input activation) when given.
static awk_status dgrad(awk_unet_trainer *t, int l, const bf16 *dz, const bf16 *act, bf16 *out,
                        int nx, int ny, int nz, hipStream_t s) {
  const LayerSpec &L = kUnetLayers[l];
  ConvArgs g;
  g.x = dz;
  g.w = t->wd[l];
  g.bias = nullptr;
  g.dact = act;
  g.out = out;
  g.C = L.cout;
  g.cout = L.cin;
  g.kpad = t->kpad_d[l];
  g.nchunks = 27 * (L.cout / 8);
  g.c8magic = L.cout > 8 ? (unsigned)(4294967296ull / (unsigned)(L.cout / 8)) + 1u : 0u;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.batch = 1;
  g.slope = act ? kUnetSlope : 1.f;
  g.xa = nullptr;
  g.ca = 0;
  return launch_conv3(g, s);
}

static awk_status wgrad(awk_unet_trainer *t, int l, const bf16 *dz, const bf16 *x, int nx, int ny,
                        int nz, hipStream_t s) {
  const LayerSpec &L = kUnetLayers[l];
  const int cot = L.cout >= 64 ? 4 : L.cout >= 32 ? 2 : 1;
  WgradArgs g;
  g.dz = dz;
  g.x = x;
  g.ws = t->ws;
  g.CO = L.cout;
  g.CI = L.cin;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.bxn = (nx + kWbX - 1) / kWbX;
  g.byn = (ny + kWbY - 1) / kWbY;
  const long long nbr = (long long)g.bxn * g.byn * ((nz + kWbZ - 1) / kWbZ);
  g.nbricks = (int)nbr;
  g.citiles = (L.cin + 15) / 16;
  const int G = L.cout / (16 * cot) * g.citiles;
  const long long E = 27LL * cot * 256;
  // position blocks: up to kWgMaxBlocks in all, one brick at least each, and the
  // partial sums within the workspace
  long long nb = kWgMaxBlocks / G;
  if (nb > nbr) nb = nbr;
  if (nb > kWgWsFloats / (G * E)) nb = kWgWsFloats / (G * E);
  if (nb < 1) nb = 1;
  const dim3 grid((unsigned)nb, (unsigned)G);
  const size_t lds = (size_t)(kWhP + cot * kWbP) * 16 * sizeof(bf16);
  if (cot == 4) AWK_LAUNCH(k_wgrad<4>, grid, dim3(256), lds, s, g);
  else if (cot == 2) AWK_LAUNCH(k_wgrad<2>, grid, dim3(256), lds, s, g);
  else AWK_LAUNCH(k_wgrad<1>, grid, dim3(256), lds, s, g);
  AWK_LAUNCH_CHECK("k_wgrad");
  // fixed-order reduction: rows of at most kWgSumRows partials, then the rows
  const long long ne = G * E, nr = ne + L.cout;
  const float *src = t->ws;
  int rows = (int)nb;
  if (nb > kWgSumRows) {
    rows = (int)((nb + kWgSumRows - 1) / kWgSumRows);
    AWK_LAUNCH(k_wgrad_sum, dim3((unsigned)((nr + 255) / 256), (unsigned)rows), dim3(256), 0, s, t->ws,
               t->ws2, (int)nb, kWgSumRows, (int)ne, L.cout);
    src = t->ws2;
  }
  AWK_LAUNCH(k_wgrad_reduce, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, src,
             t->g + t->woff[l], t->g + t->boff[l], rows, G, cot, g.citiles, L.cout, kUnetBlobCin[l]);
  AWK_LAUNCH_CHECK("k_wgrad_reduce");
  return AWK_OK;
}

// Forward and backward of one pair: the gradient is added to t->g, the loss to
// *loss_dev.
static awk_status train_backward(awk_unet_trainer *t, const float *templ, const float *subject,
                                 const float *const target[3], double *loss_dev, hipStream_t s) {
  awk_unet *u = t->net;
  const int nx = u->nx, ny = u->ny, nz = u->nz;
  const int a1 = nx / 2, b1 = ny / 2, c1 = nz / 2, a2 = nx / 4, b2 = ny / 4, c2 = nz / 4,
            a3 = nx / 8, b3 = ny / 8, c3 = nz / 8;
  const int N0 = nx * ny * nz;
  const size_t n0 = (size_t)N0;
  float *up[3] = {t->u, t->u + n0, t->u + 2 * n0};
  awk_status st;
  if ((st = unet_forward(u, 1, &templ, &subject, up, s, true))) return st;
  const int hb = (N0 + 255) / 256;
  AWK_LAUNCH(k_head_bwd, dim3(hb), dim3(256), 0, s, t->dz13, t->hpart, up[0], up[1], up[2],
             target[0], target[1], target[2], u->d0, u->head_w, N0, kUnetSlope);
  AWK_LAUNCH(k_head_reduce, dim3(52), dim3(256), 0, s, t->g + t->hoff, t->g + t->hoff + 48, loss_dev,
             t->hpart, hb);
  AWK_LAUNCH_CHECK("k_head_bwd");
  const float sl = kUnetSlope;
  auto blocks = [](long long n) { return dim3((unsigned)((n + 255) / 256)); };
  // decoder level 0
  if ((st = wgrad(t, 13, t->dz13, u->d0a, nx, ny, nz, s))) return st;
  if ((st = dgrad(t, 13, t->dz13, u->d0a, t->dz12, nx, ny, nz, s))) return st;
  if ((st = wgrad(t, 12, t->dz12, u->c0, nx, ny, nz, s))) return st;
  if ((st = dgrad(t, 12, t->dz12, nullptr, t->dc0, nx, ny, nz, s))) return st;
  AWK_LAUNCH(k_upcat_bwd, blocks((long long)(n0 / 8) * 4), dim3(256), 0, s, t->dz11, t->dc0, u->d1, 32, 48,
             nx, ny, nz, sl);
  // decoder level 1
  if ((st = wgrad(t, 11, t->dz11, u->d1a, a1, b1, c1, s))) return st;
  if ((st = dgrad(t, 11, t->dz11, u->d1a, t->dz10, a1, b1, c1, s))) return st;
  if ((st = wgrad(t, 10, t->dz10, u->c1, a1, b1, c1, s))) return st;
  if ((st = dgrad(t, 10, t->dz10, nullptr, t->dc1, a1, b1, c1, s))) return st;
  AWK_LAUNCH(k_upcat_bwd, blocks((long long)(n0 / 64) * 8), dim3(256), 0, s, t->dz9, t->dc1, u->d2, 64, 96,
             a1, b1, c1, sl);
  // decoder level 2
  if ((st = wgrad(t, 9, t->dz9, u->d2a, a2, b2, c2, s))) return st;
  if ((st = dgrad(t, 9, t->dz9, u->d2a, t->dz8, a2, b2, c2, s))) return st;
  if ((st = wgrad(t, 8, t->dz8, u->c2, a2, b2, c2, s))) return st;
  if ((st = dgrad(t, 8, t->dz8, nullptr, t->dc2, a2, b2, c2, s))) return st;
  AWK_LAUNCH(k_upcat_bwd, blocks((long long)(n0 / 512) * 16), dim3(256), 0, s, t->dz7, t->dc2, u->s3, 128,
             192, a2, b2, c2, sl);
  // bottom
  if ((st = wgrad(t, 7, t->dz7, u->t3, a3, b3, c3, s))) return st;
  if ((st = dgrad(t, 7, t->dz7, u->t3, t->dz6, a3, b3, c3, s))) return st;
  if ((st = wgrad(t, 6, t->dz6, u->p2, a3, b3, c3, s))) return st;
  if ((st = dgrad(t, 6, t->dz6, nullptr, t->dp2, a3, b3, c3, s))) return st;
  // encoder level 2: pooled gradient + skip from the level-2 concatenation
  AWK_LAUNCH(k_pool_bwd, blocks((long long)(n0 / 64) * 8), dim3(256), 0, s, t->dz5, t->dp2, t->dc2, 192,
             128, u->s2, 64, a2, b2, c2, sl);
  if ((st = wgrad(t, 5, t->dz5, u->t2, a2, b2, c2, s))) return st;
  if ((st = dgrad(t, 5, t->dz5, u->t2, t->dz4, a2, b2, c2, s))) return st;
  if ((st = wgrad(t, 4, t->dz4, u->p1, a2, b2, c2, s))) return st;
  if ((st = dgrad(t, 4, t->dz4, nullptr, t->dp1, a2, b2, c2, s))) return st;
  // encoder level 1
  AWK_LAUNCH(k_pool_bwd, blocks((long long)(n0 / 8) * 4), dim3(256), 0, s, t->dz3, t->dp1, t->dc1, 96, 64,
             u->s1, 32, a1, b1, c1, sl);
  if ((st = wgrad(t, 3, t->dz3, u->t1, a1, b1, c1, s))) return st;
  if ((st = dgrad(t, 3, t->dz3, u->t1, t->dz2, a1, b1, c1, s))) return st;
  if ((st = wgrad(t, 2, t->dz2, u->p0, a1, b1, c1, s))) return st;
  if ((st = dgrad(t, 2, t->dz2, nullptr, t->dp0, a1, b1, c1, s))) return st;
  // encoder level 0
  AWK_LAUNCH(k_pool_bwd, blocks((long long)n0 * 2), dim3(256), 0, s, t->dz1, t->dp0, t->dc0, 48, 32,
             u->s0, 16, nx, ny, nz, sl);
  if ((st = wgrad(t, 1, t->dz1, u->t0, nx, ny, nz, s))) return st;
  if ((st = dgrad(t, 1, t->dz1, u->t0, t->dz0, nx, ny, nz, s))) return st;
  if ((st = wgrad(t, 0, t->dz0, u->x0, nx, ny, nz, s))) return st;
  AWK_LAUNCH_CHECK("backward");
  return AWK_OK;
}

extern "C" {

awk_status awk_unet_trainer_create(awk_unet_trainer_t *out, awk_unet_t net, const float *weights,
                                   size_t n_weights) {
  if (!out || !net || !weights) return fail(AWK_ERR_INVALID, "awk_unet_trainer_create: null argument");
  *out = nullptr;
  if (n_weights != awk_unet_weight_count())
    return fail(AWK_ERR_INVALID, "awk_unet_trainer_create: weight count does not match the network");
  // the gradient kernel's shapes: 16-channel output tiles, 8-channel input groups,
  // at most 128 outputs (bias partials), one block's partial sums within the workspace
  for (int l = 0; l < 14; ++l)
    if (kUnetLayers[l].cout % 16 || kUnetLayers[l].cout > 128 || kUnetLayers[l].cin % 8 ||
        27LL * 256 * (kUnetLayers[l].cout / 16) * ((kUnetLayers[l].cin + 15) / 16) > kWgWsFloats)
      return fail(AWK_ERR_UNSUPPORTED, "awk_unet_trainer_create: layer shape not supported by the gradient kernel");
  awk_unet_trainer *t = new (std::nothrow) awk_unet_trainer();
  if (!t) return fail(AWK_ERR_ALLOC, "awk_unet_trainer_create: host allocation failed");
  t->net = net;
  t->nw = n_weights;
  size_t off = 0;
  for (int l = 0; l < 14; ++l) {
    t->woff[l] = off;
    off += (size_t)kUnetLayers[l].cout * kUnetBlobCin[l] * 27;
    t->boff[l] = off;
    off += kUnetLayers[l].cout;
  }
  t->hoff = off;
  const size_t N0 = (size_t)net->nx * net->ny * net->nz, N1 = N0 / 8, N2 = N1 / 8, N3 = N2 / 8;
  hipError_t e = hipSuccess;
  const size_t wb = n_weights * sizeof(float);
  float **fb[] = {&t->w, &t->g, &t->m, &t->v};
  for (float **p : fb) {
    if (e == hipSuccess) e = hipMalloc((void **)p, wb);
    if (e == hipSuccess) e = hipMemset(*p, 0, wb);
  }
  if (e == hipSuccess) e = hipMemcpy(t->w, weights, wb, hipMemcpyHostToDevice);
  for (int l = 1; l < 14 && e == hipSuccess; ++l) {
    const LayerSpec &L = kUnetLayers[l];
    t->kpad_d[l] = (27 * L.cout + 31) / 32 * 32;
    const size_t n = (size_t)L.cin * t->kpad_d[l];
    e = hipMalloc(&t->wd[l], n * sizeof(bf16));
    if (e == hipSuccess) e = hipMemset(t->wd[l], 0, n * sizeof(bf16));
  }
  struct {
    bf16 **p;
    size_t n;
  } bufs[] = {{&t->dz13, N0 * 16}, {&t->dz12, N0 * 16}, {&t->dc0, N0 * 48},  {&t->dz11, N1 * 32},
              {&t->dz10, N1 * 32}, {&t->dc1, N1 * 96},  {&t->dz9, N2 * 64},  {&t->dz8, N2 * 64},
              {&t->dc2, N2 * 192}, {&t->dz7, N3 * 128}, {&t->dz6, N3 * 128}, {&t->dp2, N3 * 64},
              {&t->dz5, N2 * 64},  {&t->dz4, N2 * 64},  {&t->dp1, N2 * 32},  {&t->dz3, N1 * 32},
              {&t->dz2, N1 * 32},  {&t->dp0, N1 * 16},  {&t->dz1, N0 * 16},  {&t->dz0, N0 * 16}};
  for (auto &b : bufs)
    if (e == hipSuccess) e = hipMalloc((void **)b.p, b.n * sizeof(bf16));
  if (e == hipSuccess) e = hipMalloc(&t->u, 3 * N0 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&t->hpart, (N0 + 255) / 256 * 52 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&t->ws, (kWgWsFloats + (size_t)kWgMaxBlocks * 128) * sizeof(float));
  // rows = ceil(nb / kWgSumRows) <= nb / kWgSumRows + 1, so rows * ne <= ws / kWgSumRows + ne
  size_t max_ne = 0;
  for (int l = 0; l < 14; ++l) {
    const size_t ne = (size_t)27 * 256 * (kUnetLayers[l].cout / 16) * ((kUnetLayers[l].cin + 15) / 16);
    if (ne > max_ne) max_ne = ne;
  }
  if (e == hipSuccess)
    e = hipMalloc(&t->ws2, (kWgWsFloats / kWgSumRows + max_ne + (size_t)kWgMaxBlocks * 128) * sizeof(float));
  awk_status st = AWK_OK;
  if (e != hipSuccess) st = hip_check(e, "awk_unet_trainer_create: allocation");
  if (!st) st = repack_all(t, nullptr);
  if (!st) {
    hipError_t se = hipDeviceSynchronize();
    if (se != hipSuccess) st = hip_check(se, "awk_unet_trainer_create: repack");
  }
  if (st) {
    trainer_free(t);
    delete t;
    return st;
  }
  *out = t;
  return AWK_OK;
}

awk_status awk_unet_trainer_destroy(awk_unet_trainer_t t) {
  if (!t) return AWK_OK;
  trainer_free(t);
  delete t;
  return AWK_OK;
}

awk_status awk_unet_trainer_step(awk_unet_trainer_t t, const float *templ, const float *subject,
                                 const float *const target[3], float lr, double *loss_dev,
                                 awk_stream_t stream) {
  if (!t || !templ || !subject || !target || !target[0] || !target[1] || !target[2] || !loss_dev)
    return fail(AWK_ERR_INVALID, "awk_unet_trainer_step: null argument");
  if (!(lr > 0.f)) return fail(AWK_ERR_INVALID, "awk_unet_trainer_step: learning rate must be > 0");
  hipStream_t s = as_stream(stream);
  awk_status st;
  if ((st = train_backward(t, templ, subject, target, loss_dev, s))) return st;
  // Adam, then the bf16 copies the convolutions read
  ++t->t;
  const float b1c = 0.9f, b2c = 0.999f;
  const float c1c = 1.f - std::pow(b1c, (float)t->t), c2c = 1.f - std::pow(b2c, (float)t->t);
  AWK_LAUNCH(k_adam, dim3((unsigned)((t->nw + 255) / 256)), dim3(256), 0, s, t->w, t->g, t->m, t->v,
             (long long)t->nw, lr, b1c, b2c, 1e-8f, c1c, c2c);
  AWK_LAUNCH_CHECK("k_adam");
  return repack_all(t, s);
}

awk_status awk_unet_trainer_gradients(awk_unet_trainer_t t, const float *templ, const float *subject,
                                      const float *const target[3], float *grad_host,
                                      double *loss_host) {
  if (!t || !templ || !subject || !target || !target[0] || !target[1] || !target[2] || !grad_host ||
      !loss_host)
    return fail(AWK_ERR_INVALID, "awk_unet_trainer_gradients: null argument");
  double *ld = nullptr;
  AWK_HIP(hipMalloc(&ld, sizeof(double)));
  awk_status st = AWK_OK;
  hipError_t e = hipMemset(ld, 0, sizeof(double));
  if (e == hipSuccess) e = hipMemset(t->g, 0, t->nw * sizeof(float));
  if (e != hipSuccess) st = hip_check(e, "awk_unet_trainer_gradients: clear");
  if (!st) st = train_backward(t, templ, subject, target, ld, nullptr);
  if (!st) {
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(grad_host, t->g, t->nw * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(loss_host, ld, sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemset(t->g, 0, t->nw * sizeof(float));
    if (e != hipSuccess) st = hip_check(e, "awk_unet_trainer_gradients");
  }
  (void)hipFree(ld);
  return st;
}

awk_status awk_unet_trainer_get_weights(awk_unet_trainer_t t, float *host_out, size_t n) {
  if (!t || !host_out || n != t->nw) return fail(AWK_ERR_INVALID, "awk_unet_trainer_get_weights: bad argument");
  AWK_HIP(hipDeviceSynchronize());
  AWK_HIP(hipMemcpy(host_out, t->w, n * sizeof(float), hipMemcpyDeviceToHost));
  return AWK_OK;
}

}  // extern "C"
