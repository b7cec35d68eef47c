// awk_unet_train.hip -- training of the U-Net warm start on the device.
//
// One step = forward (awk_unet.hip, activations kept), loss L = mean over voxels
// and components of (u - u_target)^2, backward, Adam on fp32 master weights, and
// a repack of the bf16 weights the convolutions read.  The backward pass:
//   dgrad  the same implicit-GEMM MFMA kernel as the forward convolution, run on
//          the output gradient with the weights flipped in space and transposed
//          in channels, with the previous activation's LeakyReLU derivative
//          folded into its epilogue;
//   wgrad  dW[co][ci][tap] = sum_p dZ[p][co] X[p + tap][ci]: blocks take one tap
//          and a range of positions, stage 64 positions of dZ and of the shifted
//          X in LDS transposed (channel-major), and run MFMAs with K = positions;
//          partial sums are added to the fp32 gradient with atomics;
//   pooling, upsampling + concatenation, the head and the loss have their own
//   small kernels.
// Gradients are stored in bf16 between layers and accumulated in fp32.  The
// gradient sums use float atomics, so training is not bit-reproducible run to
// run (inference is).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "awk_unet.h"

namespace awk {

constexpr int kWgStage = 128;     // positions staged per LDS round (4 K steps of 32)
constexpr int kWgPitch = 136;     // LDS row pitch (bf16): 272 B keeps 16-byte reads aligned
constexpr int kWgMaxCi = 192;
constexpr int kWgMaxCo = 128;
constexpr int kWgMaxTiles = 16;   // accumulator tiles per wave

struct WgradArgs {
  const bf16 *dz;  // [N][CO]
  const bf16 *x;   // [N][CI]
  float *dw;       // [CO][CIb][27] (blob OIDHW), accumulated
  float *db;       // [CO], accumulated by the centre-tap blocks; may be null
  int CO, CI, CIb;
  int nx, ny, nz;
  int chunk;       // positions per block (multiple of kWgStage)
};

// Work split: with at least 4 output tiles, wave w owns tiles w, w+4, ... over all
// K steps; with fewer, every wave owns all tiles and the K step ks = w of each
// staged round, and the waves' partial sums are added through LDS at the end.
__global__ void __launch_bounds__(256)
k_wgrad(WgradArgs g) {
  extern __shared__ bf16 lds[];
  bf16 *zt = lds;
  bf16 *xt = lds + g.CO * kWgPitch;
  const int tap = (int)blockIdx.y;
  const int N = g.nx * g.ny * g.nz;
  const int pbeg = (int)blockIdx.x * g.chunk;
  const int pend = min(pbeg + g.chunk, N);
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int CIt = (g.CI + 15) / 16, COt = g.CO / 16;
  const int ntiles = COt * CIt;
  const bool split_k = ntiles < 4;
  const bool do_bias = g.db && tap == 13;
  float bsum = 0.f;
  f32x4 acc[kWgMaxTiles];
#pragma unroll
  for (int j = 0; j < kWgMaxTiles; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int zchunks = g.CO / 8, xchunks = CIt * 2;  // 8-channel chunks per position
  for (int s0 = pbeg; s0 < pend; s0 += kWgStage) {
    // stage dZ and the tap-shifted X, transposed to [channel][position]
    for (int w = threadIdx.x; w < kWgStage * zchunks; w += 256) {
      const int q = w / zchunks, c8 = w - q * zchunks;
      const int p = s0 + q;
      bf16x8 v = zero8();
      if (p < pend) v = *reinterpret_cast<const bf16x8 *>(g.dz + (size_t)p * g.CO + c8 * 8);
#pragma unroll
      for (int j = 0; j < 8; ++j) zt[(c8 * 8 + j) * kWgPitch + q] = v[j];
    }
    for (int w = threadIdx.x; w < kWgStage * xchunks; w += 256) {
      const int q = w / xchunks, c8 = w - q * xchunks;
      const int p = s0 + q;
      bf16x8 v = zero8();
      if (p < pend && c8 * 8 < g.CI) {
        const int x = p % g.nx, r = p / g.nx, y = r % g.ny, z = r / g.ny;
        const int X = x + dx, Y = y + dy, Z = z + dz;
        if (X >= 0 && X < g.nx && Y >= 0 && Y < g.ny && Z >= 0 && Z < g.nz)
          v = *reinterpret_cast<const bf16x8 *>(g.x + ((size_t)(Z * g.ny + Y) * g.nx + X) * g.CI + c8 * 8);
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) xt[(c8 * 8 + j) * kWgPitch + q] = v[j];
    }
    __syncthreads();
    if (do_bias && (int)threadIdx.x < g.CO)
      for (int q = 0; q < kWgStage; ++q) bsum += (float)zt[threadIdx.x * kWgPitch + q];
    if (split_k) {
      const int ks = wave;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (j >= ntiles) break;
        const int m = j / CIt, n = j - m * CIt;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&zt[(m * 16 + col) * kWgPitch + ks * 32 + grp * 8]);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&xt[(n * 16 + col) * kWgPitch + ks * 32 + grp * 8]);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < kWgStage / 32; ++ks) {
#pragma unroll
        for (int j = 0; j < kWgMaxTiles; ++j) {
          const int tile = wave + 4 * j;
          if (tile >= ntiles) break;
          const int m = tile / CIt, n = tile - m * CIt;
          const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&zt[(m * 16 + col) * kWgPitch + ks * 32 + grp * 8]);
          const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&xt[(n * 16 + col) * kWgPitch + ks * 32 + grp * 8]);
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  if (do_bias && (int)threadIdx.x < g.CO) atomicAdd(&g.db[threadIdx.x], bsum);
  if (split_k) {
    // add the 4 waves' partial tiles (the staging space is free after the last sync)
    f32x4 *red = reinterpret_cast<f32x4 *>(lds);
    if (wave > 0)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (j < ntiles) red[((wave - 1) * 3 + j) * 64 + lane] = acc[j];
    __syncthreads();
    if (wave > 0) return;
#pragma unroll
    for (int w = 0; w < 3; ++w)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (j < ntiles) acc[j] += red[(w * 3 + j) * 64 + lane];
  }
  // D[co][ci]: lane holds co = m*16 + 4*grp + i, ci = n*16 + col
#pragma unroll
  for (int j = 0; j < kWgMaxTiles; ++j) {
    const int tile = split_k ? j : wave + 4 * j;
    if (tile >= ntiles) break;
    const int m = tile / CIt, n = tile - m * CIt;
    const int ci = n * 16 + col;
    if (ci >= g.CIb) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = m * 16 + grp * 4 + i;
      atomicAdd(&g.dw[((size_t)co * g.CIb + ci) * 27 + tap], acc[j][i]);
    }
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
};

static void trainer_free(awk_unet_trainer *t) {
  void *b[] = {t->w, t->g, t->m, t->v, t->dz13, t->dz12, t->dc0, t->dz11, t->dz10, t->dc1, t->dz9,
               t->dz8, t->dc2, t->dz7, t->dz6, t->dp2, t->dz5, t->dz4, t->dp1, t->dz3, t->dz2,
               t->dp0, t->dz1, t->dz0, t->u, t->hpart};
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
  WgradArgs g;
  g.dz = dz;
  g.x = x;
  g.dw = t->g + t->woff[l];
  g.CO = L.cout;
  g.CI = L.cin;
  g.CIb = kUnetBlobCin[l];
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.db = t->g + t->boff[l];
  const long long N = (long long)nx * ny * nz;
  // blocks per tap: up to 256, but few enough that the layer's partial-sum
  // atomics (blocks x cout x cin) stay around 4M
  long long per_tap = (4LL << 20) / (27LL * L.cout * kUnetBlobCin[l]);
  per_tap = per_tap < 1 ? 1 : (per_tap > 256 ? 256 : per_tap);
  long long chunk = (N + per_tap - 1) / per_tap;
  chunk = (chunk + kWgStage - 1) / kWgStage * kWgStage;
  g.chunk = (int)chunk;
  const dim3 grid((unsigned)((N + chunk - 1) / chunk), 27);
  const int CIt = (L.cin + 15) / 16;
  size_t lds = (size_t)(L.cout + CIt * 16) * kWgPitch * sizeof(bf16);
  const size_t red = (size_t)3 * 3 * 64 * sizeof(f32x4);  // split-K partials reuse the staging space
  if (lds < red) lds = red;
  AWK_LAUNCH(k_wgrad, grid, dim3(256), lds, s, g);
  AWK_LAUNCH_CHECK("k_wgrad");
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
  for (int l = 0; l < 14; ++l)
    if (kUnetLayers[l].cout > kWgMaxCo || kUnetLayers[l].cin > kWgMaxCi ||
        (kUnetLayers[l].cout / 16) * ((kUnetLayers[l].cin + 15) / 16) > 4 * kWgMaxTiles)
      return fail(AWK_ERR_UNSUPPORTED, "awk_unet_trainer_create: layer too wide for the gradient kernel");
  // the widest layers stage (128 + 128) x 136 bf16 = 68 KiB of LDS per wgrad block
  AWK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_wgrad),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (kWgMaxCo + kWgMaxCi) * kWgPitch * (int)sizeof(bf16)));
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
