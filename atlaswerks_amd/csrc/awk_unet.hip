// awk_unet.hip -- 3-D U-Net warm start on MFMA (gfx950 bf16 matrix cores).
//
// Predicts an initial displacement u (3 x nz x ny x nx, fp32) from a (template,
// subject) pair; the greedy or LDMM loop then starts from h = x + u.  No
// reference counterpart: the network is fixed at 4 levels with 16-32-64-128
// channels, two 3x3x3 convolutions per level, 2x2x2 average pooling, nearest x2
// upsampling with skip concatenation, LeakyReLU(0.2), and a 1x1x1 head.
//
// Activations are channels-last bf16 (NDHWC: act[((z*ny + y)*nx + x)*C + c]).  A
// 3x3x3 convolution is an implicit GEMM: D[co][p] = sum_k W[co][k] X[k][p] with
// k = tap * Cin + ci, tap = kz*9 + ky*3 + kx, on v_mfma_f32_16x16x32_bf16 with
// fp32 accumulation.  A wave owns 16 output positions and all output channels;
// lane l supplies, per 32-wide K step, 8 consecutive k (one tap, 8 channels: a
// single 16-byte load) of position l%16 and of output channel l%16 of every
// 16-channel tile.  The decoder's nearest upsampling and skip concatenation are
// folded into the B-operand load (source A read at half resolution, channels
// past A's count read from source B), and bias + LeakyReLU + bf16 rounding into
// the epilogue, which stores 4 consecutive channels per lane.
#include <cstring>
#include <new>
#include <vector>

#include "awk_ops.h"

namespace awk {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kConvWaves = 4;
constexpr int kConvBlock = 64 * kConvWaves;
constexpr int kPosPerWave = 16;

struct ConvArgs {
  const bf16 *a;  // source A: ca channels, at half resolution when a_half
  const bf16 *b;  // source B (skip connection): cb channels, full resolution; may be null
  const bf16 *w;  // [cout][kpad]
  const float *bias;
  bf16 *out;      // [N][cout]
  int ca, cb, a_half, cout, kpad;
  int nx, ny, nz;
  float slope;
};

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (bf16)0.f;
  return z;
}

// MT: 16-channel output tiles (cout / 16)
template <int MT>
__global__ void __launch_bounds__(kConvBlock)
k_conv3(ConvArgs g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int N = g.nx * g.ny * g.nz;
  const int p0 = (blockIdx.x * kConvWaves + wave) * kPosPerWave;
  if (p0 >= N) return;
  const int col = lane & 15, grp = lane >> 4;
  // this lane's position (B-operand column)
  const int p = p0 + col;
  const bool valid = p < N;  // grids smaller than 16 voxels leave lanes idle
  const int px = p % g.nx, pr = p / g.nx, py = pr % g.ny, pz = pr / g.ny;
  const int cin = g.ca + g.cb;
  const int hx = g.nx >> 1;
  f32x4 acc[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ksteps = g.kpad >> 5;
  for (int ks = 0; ks < ksteps; ++ks) {
    const int k0 = ks * 32 + grp * 8;
    const int tap = k0 / cin, ci = k0 - tap * cin;
    bf16x8 bv = zero8();
    if (valid && tap < 27) {
      const int qx = px + (tap % 3) - 1, qy = py + (tap / 3) % 3 - 1, qz = pz + tap / 9 - 1;
      if (qx >= 0 && qx < g.nx && qy >= 0 && qy < g.ny && qz >= 0 && qz < g.nz) {
        if (ci < g.ca) {
          const long long q = g.a_half ? ((long long)((qz >> 1) * (g.ny >> 1) + (qy >> 1)) * hx + (qx >> 1))
                                       : ((long long)(qz * g.ny + qy) * g.nx + qx);
          bv = *reinterpret_cast<const bf16x8 *>(g.a + q * g.ca + ci);
        } else {
          const long long q = (long long)(qz * g.ny + qy) * g.nx + qx;
          bv = *reinterpret_cast<const bf16x8 *>(g.b + q * g.cb + (ci - g.ca));
        }
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const bf16x8 av = *reinterpret_cast<const bf16x8 *>(g.w + (size_t)(m * 16 + col) * g.kpad + k0);
      acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[m], 0, 0, 0);
    }
  }
  // epilogue: lane holds channels m*16 + 4*grp + 0..3 of position p
  if (!valid) return;
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const int c0 = m * 16 + grp * 4;
    bf16x4 o;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v = acc[m][i] + g.bias[c0 + i];
      v = v >= 0.f ? v : v * g.slope;
      o[i] = (bf16)v;
    }
    *reinterpret_cast<bf16x4 *>(g.out + (size_t)p * g.cout + c0) = o;
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

// 1x1x1 head: u_c = b_c + sum_ci W[c][ci] d[ci] (fp32), written as 3 planes
__global__ void __launch_bounds__(256)
k_head(float *__restrict__ ux, float *__restrict__ uy, float *__restrict__ uz,
       const bf16 *__restrict__ d, const float *__restrict__ w, const float *__restrict__ b, int N,
       int C) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  float a0 = b[0], a1 = b[1], a2 = b[2];
  for (int c = 0; c < C; ++c) {
    const float v = (float)d[(size_t)i * C + c];
    a0 += w[c] * v;
    a1 += w[C + c] * v;
    a2 += w[2 * C + c] * v;
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

}  // namespace awk

namespace {

// layer table: (cin of source A, cin of source B, cout, A at half resolution)
struct LayerSpec {
  int ca, cb, cout, a_half;
};
// encoder: 2 convs per level at 16/32/64/128; decoder: up + skip concat, 2 convs
constexpr LayerSpec kLayers[14] = {
    {8, 0, 16, 0},   {16, 0, 16, 0},                     // level 0 (input padded 2 -> 8)
    {16, 0, 32, 0},  {32, 0, 32, 0},                     // level 1
    {32, 0, 64, 0},  {64, 0, 64, 0},                     // level 2
    {64, 0, 128, 0}, {128, 0, 128, 0},                   // level 3 (bottom)
    {128, 64, 64, 1}, {64, 0, 64, 0},                    // decoder level 2
    {64, 32, 32, 1},  {32, 0, 32, 0},                    // decoder level 1
    {32, 16, 16, 1},  {16, 0, 16, 0},                    // decoder level 0
};
// input channels the weight blob carries per layer (the first layer has 2)
constexpr int kBlobCin[14] = {2, 16, 16, 32, 32, 64, 64, 128, 192, 64, 96, 32, 48, 16};

}  // namespace

struct awk_unet {
  int nx = 0, ny = 0, nz = 0;
  awk::bf16 *w[14] = {};
  float *bias[14] = {};
  int kpad[14] = {};
  float *head_w = nullptr, *head_b = nullptr;
  // activations
  awk::bf16 *x0 = nullptr, *t0 = nullptr, *s0 = nullptr, *p0 = nullptr;
  awk::bf16 *t1 = nullptr, *s1 = nullptr, *p1 = nullptr;
  awk::bf16 *t2 = nullptr, *s2 = nullptr, *p2 = nullptr;
  awk::bf16 *t3 = nullptr, *s3 = nullptr;
  awk::bf16 *d2a = nullptr, *d2 = nullptr, *d1a = nullptr, *d1 = nullptr, *d0a = nullptr,
            *d0 = nullptr;
};

using namespace awk;

static void unet_free(awk_unet *u) {
  for (int l = 0; l < 14; ++l) {
    if (u->w[l]) (void)hipFree(u->w[l]);
    if (u->bias[l]) (void)hipFree(u->bias[l]);
  }
  void *b[] = {u->head_w, u->head_b, u->x0, u->t0, u->s0, u->p0, u->t1, u->s1, u->p1, u->t2,
               u->s2, u->p2, u->t3, u->s3, u->d2a, u->d2, u->d1a, u->d1, u->d0a, u->d0};
  for (void *p : b)
    if (p) (void)hipFree(p);
}

static awk_status conv(const awk_unet *u, int l, const bf16 *a, const bf16 *b, bf16 *out, int nx,
                       int ny, int nz, hipStream_t s) {
  const LayerSpec &L = kLayers[l];
  ConvArgs g;
  g.a = a;
  g.b = b;
  g.w = u->w[l];
  g.bias = u->bias[l];
  g.out = out;
  g.ca = L.ca;
  g.cb = L.cb;
  g.a_half = L.a_half;
  g.cout = L.cout;
  g.kpad = u->kpad[l];
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.slope = 0.2f;
  const long long N = (long long)nx * ny * nz;
  const unsigned grid = (unsigned)((N + kConvWaves * kPosPerWave - 1) / (kConvWaves * kPosPerWave));
  switch (L.cout / 16) {
    case 1: AWK_LAUNCH(k_conv3<1>, dim3(grid), dim3(kConvBlock), 0, s, g); break;
    case 2: AWK_LAUNCH(k_conv3<2>, dim3(grid), dim3(kConvBlock), 0, s, g); break;
    case 4: AWK_LAUNCH(k_conv3<4>, dim3(grid), dim3(kConvBlock), 0, s, g); break;
    case 8: AWK_LAUNCH(k_conv3<8>, dim3(grid), dim3(kConvBlock), 0, s, g); break;
    default: return fail(AWK_ERR_INVALID, "unet: unsupported channel count");
  }
  AWK_LAUNCH_CHECK("k_conv3");
  return AWK_OK;
}

static awk_status pool(const bf16 *in, bf16 *out, int nx, int ny, int nz, int C, hipStream_t s) {
  const long long total = (long long)(nx / 2) * (ny / 2) * (nz / 2) * (C / 8);
  AWK_LAUNCH(k_pool2, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, out, in, nx, ny, nz, C);
  AWK_LAUNCH_CHECK("k_pool2");
  return AWK_OK;
}

extern "C" {

size_t awk_unet_weight_count(void) {
  size_t n = 0;
  for (int l = 0; l < 14; ++l) n += (size_t)kLayers[l].cout * kBlobCin[l] * 27 + kLayers[l].cout;
  return n + 3 * 16 + 3;
}

awk_status awk_unet_create(awk_unet_t *out, int nx, int ny, int nz, const float *weights,
                           size_t n_weights) {
  if (!out || !weights) return fail(AWK_ERR_INVALID, "awk_unet_create: null argument");
  *out = nullptr;
  if (nx < 8 || ny < 8 || nz < 8 || nx % 8 || ny % 8 || nz % 8)
    return fail(AWK_ERR_INVALID, "awk_unet_create: every axis must be a multiple of 8 (three poolings)");
  if ((long long)nx * ny * nz > (1LL << 27)) return fail(AWK_ERR_INVALID, "awk_unet_create: at most 2^27 voxels");
  if (n_weights != awk_unet_weight_count())
    return fail(AWK_ERR_INVALID, "awk_unet_create: weight count does not match the network");
  awk_unet *u = new (std::nothrow) awk_unet();
  if (!u) return fail(AWK_ERR_ALLOC, "awk_unet_create: host allocation failed");
  u->nx = nx;
  u->ny = ny;
  u->nz = nz;
  hipError_t e = hipSuccess;
  size_t off = 0;
  for (int l = 0; l < 14 && e == hipSuccess; ++l) {
    const LayerSpec &L = kLayers[l];
    const int cin = L.ca + L.cb, bc = kBlobCin[l];
    const int K = 27 * cin;
    u->kpad[l] = (K + 31) / 32 * 32;
    // repack OIDHW fp32 -> [cout][kpad] bf16 with k = tap*cin + ci (zeros for padding)
    std::vector<bf16> wt((size_t)L.cout * u->kpad[l], (bf16)0.f);
    for (int co = 0; co < L.cout; ++co)
      for (int ci = 0; ci < bc; ++ci)
        for (int tap = 0; tap < 27; ++tap)
          wt[(size_t)co * u->kpad[l] + tap * cin + ci] =
              (bf16)weights[off + ((size_t)co * bc + ci) * 27 + tap];
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
  const size_t N0 = (size_t)nx * ny * nz, N1 = N0 / 8, N2 = N1 / 8, N3 = N2 / 8;
  struct {
    bf16 **p;
    size_t n;
  } bufs[] = {{&u->x0, N0 * 8},  {&u->t0, N0 * 16},  {&u->s0, N0 * 16},  {&u->p0, N1 * 16},
              {&u->t1, N1 * 32}, {&u->s1, N1 * 32},  {&u->p1, N2 * 32},  {&u->t2, N2 * 64},
              {&u->s2, N2 * 64}, {&u->p2, N3 * 64},  {&u->t3, N3 * 128}, {&u->s3, N3 * 128},
              {&u->d2a, N2 * 64}, {&u->d2, N2 * 64}, {&u->d1a, N1 * 32}, {&u->d1, N1 * 32},
              {&u->d0a, N0 * 16}, {&u->d0, N0 * 16}};
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
  hipStream_t s = as_stream(stream);
  const int nx = u->nx, ny = u->ny, nz = u->nz;
  const int N0 = nx * ny * nz;
  awk_status st;
  AWK_LAUNCH(k_pack_input, dim3((N0 + 255) / 256), dim3(256), 0, s, u->x0, templ, subject, N0);
  AWK_LAUNCH_CHECK("k_pack_input");
  const int n1[3] = {nx / 2, ny / 2, nz / 2}, n2[3] = {nx / 4, ny / 4, nz / 4},
            n3[3] = {nx / 8, ny / 8, nz / 8};
  if ((st = conv(u, 0, u->x0, nullptr, u->t0, nx, ny, nz, s))) return st;
  if ((st = conv(u, 1, u->t0, nullptr, u->s0, nx, ny, nz, s))) return st;
  if ((st = pool(u->s0, u->p0, nx, ny, nz, 16, s))) return st;
  if ((st = conv(u, 2, u->p0, nullptr, u->t1, n1[0], n1[1], n1[2], s))) return st;
  if ((st = conv(u, 3, u->t1, nullptr, u->s1, n1[0], n1[1], n1[2], s))) return st;
  if ((st = pool(u->s1, u->p1, n1[0], n1[1], n1[2], 32, s))) return st;
  if ((st = conv(u, 4, u->p1, nullptr, u->t2, n2[0], n2[1], n2[2], s))) return st;
  if ((st = conv(u, 5, u->t2, nullptr, u->s2, n2[0], n2[1], n2[2], s))) return st;
  if ((st = pool(u->s2, u->p2, n2[0], n2[1], n2[2], 64, s))) return st;
  if ((st = conv(u, 6, u->p2, nullptr, u->t3, n3[0], n3[1], n3[2], s))) return st;
  if ((st = conv(u, 7, u->t3, nullptr, u->s3, n3[0], n3[1], n3[2], s))) return st;
  if ((st = conv(u, 8, u->s3, u->s2, u->d2a, n2[0], n2[1], n2[2], s))) return st;
  if ((st = conv(u, 9, u->d2a, nullptr, u->d2, n2[0], n2[1], n2[2], s))) return st;
  if ((st = conv(u, 10, u->d2, u->s1, u->d1a, n1[0], n1[1], n1[2], s))) return st;
  if ((st = conv(u, 11, u->d1a, nullptr, u->d1, n1[0], n1[1], n1[2], s))) return st;
  if ((st = conv(u, 12, u->d1, u->s0, u->d0a, nx, ny, nz, s))) return st;
  if ((st = conv(u, 13, u->d0a, nullptr, u->d0, nx, ny, nz, s))) return st;
  AWK_LAUNCH(k_head, dim3((N0 + 255) / 256), dim3(256), 0, s, disp[0], disp[1], disp[2], u->d0,
             u->head_w, u->head_b, N0, 16);
  AWK_LAUNCH_CHECK("k_head");
  return AWK_OK;
}

awk_status awk_unet_mfma_probe(float *out, awk_stream_t stream) {
  if (!out) return fail(AWK_ERR_INVALID, "awk_unet_mfma_probe: null pointer");
  AWK_LAUNCH(k_mfma_probe, dim3(1), dim3(64), 0, as_stream(stream), out);
  AWK_LAUNCH_CHECK("k_mfma_probe");
  return AWK_OK;
}

}  // extern "C"
