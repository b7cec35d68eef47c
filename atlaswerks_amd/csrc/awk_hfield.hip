// awk_hfield.hip -- hfield utilities on gfx950: identity / conversions, apply
// (trilinear gather of an image), the compose family, gradient, and the error
// plumbing of the C-ABI.
//
// Every per-voxel kernel uses one thread per voxel, 256 threads per block, and
// the XCD-aware block order of awk_common.h so that the 8-corner gathers of
// neighbouring blocks share one XCD's L2.
#include <cstring>
#include <string>

#include "awk_ops.h"

namespace awk {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

awk_status fail(awk_status code, const std::string &msg) {
  g_err = msg;
  return code;
}

awk_status hip_check(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return AWK_ERR_HIP;
}

awk_status read_flags(unsigned *const *dev, int n, unsigned *out, int clear, hipStream_t s,
                      const char *who) {
  if (!out) return fail(AWK_ERR_INVALID, std::string(who) + ": null flags pointer");
  unsigned acc = 0;
  for (int i = 0; i < n; ++i) {
    unsigned f = 0;
    AWK_HIP(load_scalar_sync(&f, dev[i], s));
    acc |= f;
    if (clear) AWK_HIP(store_scalar_async(dev[i], 0u, s));
  }
  *out = acc;
  if (acc & (AWK_FLAG_NAN_ENERGY | AWK_FLAG_BAD_STEP))
    return fail(AWK_ERR_NUMERIC, std::string(who) + ((acc & AWK_FLAG_NAN_ENERGY)
                                                         ? ": NaN encountered in the energy"
                                                         : ": max displacement is NaN, infinite or zero"));
  return AWK_OK;
}

constexpr int kBlock = 256;

LaunchEvents *&launch_events() {
  static thread_local LaunchEvents *ev = nullptr;
  return ev;
}

Tuning &tuning() {
  static Tuning t;
  return t;
}

static awk_status check_dims(int nx, int ny, int nz, const char *who) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(AWK_ERR_INVALID, std::string(who) + ": sizes must be >= 1");
  if ((long long)nx * ny * nz > kBufMaxElems) return fail(AWK_ERR_INVALID, std::string(who) + ": volume too large (at most 2^29 voxels)");
  return AWK_OK;
}

__device__ __forceinline__ bool voxel(int N, int nx, int ny, int &i, int &x, int &y, int &z) {
  i = (int)(xcd_block() * blockDim.x + threadIdx.x);
  if (i >= N) return false;
  x = i % nx;
  const int r = i / nx;
  y = r % ny;
  z = r / ny;
  return true;
}

// ------------------------------------------------------------ identity etc.
__global__ void k_identity(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
                           int nx, int ny, int nz) {
  int i, x, y, z;
  if (!voxel(nx * ny * nz, nx, ny, i, x, y, z)) return;
  hx[i] = (float)x;
  hy[i] = (float)y;
  hz[i] = (float)z;
}

// h(x) = M [x y z 1]^T, M a 3x4 voxel-to-voxel affine (row major)
__global__ void k_affine(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
                         int nx, int ny, int nz, Affine34f a) {
  int i, x, y, z;
  if (!voxel(nx * ny * nz, nx, ny, i, x, y, z)) return;
  const float fx = (float)x, fy = (float)y, fz = (float)z;
  hx[i] = a.m[0] * fx + a.m[1] * fy + a.m[2] * fz + a.m[3];
  hy[i] = a.m[4] * fx + a.m[5] * fy + a.m[6] * fz + a.m[7];
  hz[i] = a.m[8] * fx + a.m[9] * fy + a.m[10] * fz + a.m[11];
}

// mode 0: v -> h (x + v/sp), 1: h -> v ((h - x)*sp)
__global__ void k_hv_convert(float *__restrict__ a, float *__restrict__ b, float *__restrict__ c,
                             int nx, int ny, int nz, float spx, float spy, float spz, int mode) {
  int i, x, y, z;
  if (!voxel(nx * ny * nz, nx, ny, i, x, y, z)) return;
  if (mode == 0) {
    a[i] = (float)x + a[i] / spx;
    b[i] = (float)y + b[i] / spy;
    c[i] = (float)z + c[i] / spz;
  } else {
    a[i] = (a[i] - (float)x) * spx;
    b[i] = (b[i] - (float)y) * spy;
    c[i] = (c[i] - (float)z) * spz;
  }
}

__global__ void k_inv0(float *__restrict__ ox, float *__restrict__ oy, float *__restrict__ oz,
                       const float *__restrict__ hx, const float *__restrict__ hy,
                       const float *__restrict__ hz, int nx, int ny, int nz) {
  int i, x, y, z;
  if (!voxel(nx * ny * nz, nx, ny, i, x, y, z)) return;
  ox[i] = (float)(x + x) - hx[i];
  oy[i] = (float)(y + y) - hy[i];
  oz[i] = (float)(z + z) - hz[i];
}

// ------------------------------------------------------------ apply
// VPT voxels per thread (strided by the block size, so every load instruction
// stays coalesced): all h loads are issued before the gathers, raising the bytes
// in flight per wave.
template <int BG, int VPT>
__global__ void __launch_bounds__(kBlock)
k_apply(float *__restrict__ out, const float *__restrict__ img, const float *__restrict__ hx,
        const float *__restrict__ hy, const float *__restrict__ hz, int N, Dims d,
        float background) {
  // N: voxels of the hfield / output; d: the image's own grid
  const int i0 = (int)(xcd_block() * (kBlock * VPT) + threadIdx.x);
  float px[VPT], py[VPT], pz[VPT];
#pragma unroll
  for (int v = 0; v < VPT; ++v) {
    const int i = i0 + v * kBlock;
    if (i < N) {
      // h is read once: streamed past the memory-side cache, which keeps the
      // output for its reader
      px[v] = __builtin_nontemporal_load(hx + i);
      py[v] = __builtin_nontemporal_load(hy + i);
      pz[v] = __builtin_nontemporal_load(hz + i);
    }
  }
  float r[VPT];
  const BufLoad ib = make_buf(img, d.nx * d.ny * d.nz);
#pragma unroll
  for (int v = 0; v < VPT; ++v)
    if (i0 + v * kBlock < N) r[v] = trilerp1g<BG>(ib, d, px[v], py[v], pz[v], background);
#pragma unroll
  for (int v = 0; v < VPT; ++v) {
    const int i = i0 + v * kBlock;
    if (i < N) out[i] = r[v];
  }
}

// ------------------------------------------------------------ compose family
// op 0: h = f(g(x))
// op 1: h = g(x + s*delta*v(x)/sp)                 (composeHV / composeHVInv)
// op 2: h = g(x) + s*v(g(x))/sp                    (composeVH / composeVHInv)
template <int BG, int OP, int VPT>
__global__ void __launch_bounds__(kBlock)
k_compose(float *__restrict__ ox, float *__restrict__ oy, float *__restrict__ oz,
          const float *__restrict__ fx, const float *__restrict__ fy, const float *__restrict__ fz,
          const float *__restrict__ gx, const float *__restrict__ gy, const float *__restrict__ gz,
          int nx, int ny, int nz, float spx, float spy, float spz, float delta_h,
          const float *__restrict__ delta_dev, float sgn) {
  const int N = nx * ny * nz;
  const int i0 = (int)(xcd_block() * (kBlock * VPT) + threadIdx.x);
  const Dims d{nx, ny, nz};
  const float delta = delta_dev ? *delta_dev : delta_h;
  const BufLoad bx = make_buf(fx, N), by = make_buf(fy, N), bz = make_buf(fz, N);
  float ax[VPT], ay[VPT], az[VPT];
#pragma unroll
  for (int v = 0; v < VPT; ++v) {
    const int i = i0 + v * kBlock;
    if (i < N) { ax[v] = gx[i]; ay[v] = gy[i]; az[v] = gz[i]; }
  }
  float3 r[VPT];
#pragma unroll
  for (int v = 0; v < VPT; ++v) {
    const int i = i0 + v * kBlock;
    if (i >= N) continue;
    if (OP == 0) {
      r[v] = trilerp3g<BG>(bx, by, bz, d, ax[v], ay[v], az[v]);
    } else if (OP == 1) {
      // f = g (field gathered), (gx,gy,gz) = v.  The reference scales v by the step
      // first (GreedyIteratorCPU.cxx:84), then divides by the spacing.
      const int x = i % nx, rr = i / nx, y = rr % ny, z = rr / ny;
      const float vx = ax[v] * delta, vy = ay[v] * delta, vz = az[v] * delta;
      const float px = (float)x + sgn * (vx / spx);
      const float py = (float)y + sgn * (vy / spy);
      const float pz = (float)z + sgn * (vz / spz);
      r[v] = trilerp3g<BG>(bx, by, bz, d, px, py, pz);
    } else {
      // f = v (field gathered), (gx,gy,gz) = g
      float3 q = trilerp3g<BG>(bx, by, bz, d, ax[v], ay[v], az[v]);
      if (delta_dev || delta_h != 1.f) {  // v scaled by the step (greedy inverse update)
        q.x *= delta; q.y *= delta; q.z *= delta;
      }
      r[v] = make_float3(q.x / (sgn * spx) + ax[v], q.y / (sgn * spy) + ay[v], q.z / (sgn * spz) + az[v]);
    }
  }
#pragma unroll
  for (int v = 0; v < VPT; ++v) {
    const int i = i0 + v * kBlock;
    if (i < N) { ox[i] = r[v].x; oy[i] = r[v].y; oz[i] = r[v].z; }
  }
}

// ------------------------------------------------------------ gradient
__global__ void __launch_bounds__(kBlock)
k_gradient(float *__restrict__ gx, float *__restrict__ gy, float *__restrict__ gz,
           const float *__restrict__ a, int nx, int ny, int nz, float spx, float spy, float spz,
           int wrap) {
  int i, x, y, z;
  if (!voxel(nx * ny * nz, nx, ny, i, x, y, z)) return;
  const float c = a[i];
  const int s1 = nx, s2 = nx * ny;
  auto comp = [&](int p, int n, int stride, float sp) -> float {
    if (n == 1) return 0.f;
    if (p > 0 && p < n - 1) return (a[i + stride] - a[i - stride]) / (2.f * sp);
    if (wrap) {
      const int ip = (p + 1 < n) ? i + stride : i - (n - 1) * stride;
      const int im = (p > 0) ? i - stride : i + (n - 1) * stride;
      return (a[ip] - a[im]) / (2.f * sp);
    }
    if (p == 0) return (a[i + stride] - c) / sp;
    return (c - a[i - stride]) / sp;
  };
  gx[i] = comp(x, nx, 1, spx);
  gy[i] = comp(y, ny, s1, spy);
  gz[i] = comp(z, nz, s2, spz);
}

template <int OP, int VPT>
static void launch_compose_v(unsigned grid, float *ox, float *oy, float *oz, const float *fx,
                             const float *fy, const float *fz, const float *gx, const float *gy,
                             const float *gz, int nx, int ny, int nz, float spx, float spy,
                             float spz, float delta, const float *delta_dev, float sgn, int bg,
                             hipStream_t s) {
#define AWK_CMP(B)                                                                            \
  AWK_LAUNCH((k_compose<B, OP, VPT>), dim3(grid), dim3(kBlock), 0, s, ox, oy, oz, fx, fy, \
                     fz, gx, gy, gz, nx, ny, nz, spx, spy, spz, delta, delta_dev, sgn)
  switch (bg) {
    case AWK_BG_ID: AWK_CMP(AWK_BG_ID); break;
    case AWK_BG_ZERO: AWK_CMP(AWK_BG_ZERO); break;
    case AWK_BG_PARTIAL_ID: AWK_CMP(AWK_BG_PARTIAL_ID); break;
    case AWK_BG_PARTIAL_ZERO: AWK_CMP(AWK_BG_PARTIAL_ZERO); break;
    case AWK_BG_CLAMP: AWK_CMP(AWK_BG_CLAMP); break;
    default: AWK_CMP(AWK_BG_WRAP); break;
  }
#undef AWK_CMP
}

template <int OP>
static awk_status launch_compose(float *ox, float *oy, float *oz, const float *fx, const float *fy,
                                 const float *fz, const float *gx, const float *gy,
                                 const float *gz, int nx, int ny, int nz, float spx, float spy,
                                 float spz, float delta, const float *delta_dev, float sgn, int bg,
                                 hipStream_t s) {
  if (bg < AWK_BG_ID || bg > AWK_BG_WRAP) return fail(AWK_ERR_INVALID, "compose: invalid background strategy");
  const long long N = (long long)nx * ny * nz;
  const int vpt = tuning().compose_vpt;
  if (vpt == 4)
    launch_compose_v<OP, 4>(grid_for(N, kBlock * 4), ox, oy, oz, fx, fy, fz, gx, gy, gz, nx, ny, nz,
                            spx, spy, spz, delta, delta_dev, sgn, bg, s);
  else if (vpt == 2)
    launch_compose_v<OP, 2>(grid_for(N, kBlock * 2), ox, oy, oz, fx, fy, fz, gx, gy, gz, nx, ny, nz,
                            spx, spy, spz, delta, delta_dev, sgn, bg, s);
  else
    launch_compose_v<OP, 1>(grid_for(N, kBlock), ox, oy, oz, fx, fy, fz, gx, gy, gz, nx, ny, nz,
                            spx, spy, spz, delta, delta_dev, sgn, bg, s);
  AWK_LAUNCH_CHECK("k_compose");
  return AWK_OK;
}

template <int VPT>
static void launch_apply_v(float *out, const float *img, Dims di, const float *hx,
                           const float *hy, const float *hz, int N, int bg, float background,
                           hipStream_t s) {
  const unsigned grid = grid_for((long long)N, kBlock * VPT);
#define AWK_AP(B)                                                                            \
  AWK_LAUNCH((k_apply<B, VPT>), dim3(grid), dim3(kBlock), 0, s, out, img, hx, hy, hz, N, di, \
                     background)
  switch (bg) {
    case AWK_BG_VAL: AWK_AP(AWK_BG_VAL); break;
    case AWK_BG_ZERO: AWK_AP(AWK_BG_ZERO); break;
    case AWK_BG_CLAMP: AWK_AP(AWK_BG_CLAMP); break;
    default: AWK_AP(AWK_BG_WRAP); break;
  }
#undef AWK_AP
}

awk_status launch_apply_grid(float *out, const float *img, int inx, int iny, int inz,
                             const float *hx, const float *hy, const float *hz, int nx, int ny,
                             int nz, int bg, float background, hipStream_t s) {
  if (bg != AWK_BG_VAL && bg != AWK_BG_ZERO && bg != AWK_BG_CLAMP && bg != AWK_BG_WRAP)
    return fail(AWK_ERR_INVALID, "apply: background must be VAL, ZERO, CLAMP or WRAP");
  const int vpt = tuning().apply_vpt;
  const Dims di{inx, iny, inz};
  const int N = nx * ny * nz;
  if (vpt == 4) launch_apply_v<4>(out, img, di, hx, hy, hz, N, bg, background, s);
  else if (vpt == 2) launch_apply_v<2>(out, img, di, hx, hy, hz, N, bg, background, s);
  else launch_apply_v<1>(out, img, di, hx, hy, hz, N, bg, background, s);
  AWK_LAUNCH_CHECK("k_apply");
  return AWK_OK;
}

awk_status launch_apply(float *out, const float *img, const float *hx, const float *hy,
                        const float *hz, int nx, int ny, int nz, int bg, float background,
                        hipStream_t s) {
  return launch_apply_grid(out, img, nx, ny, nz, hx, hy, hz, nx, ny, nz, bg, background, s);
}

awk_status launch_compose_hv(float *ox, float *oy, float *oz, const float *gx, const float *gy,
                             const float *gz, const float *vx, const float *vy, const float *vz,
                             int nx, int ny, int nz, const float sp[3], float delta,
                             const float *delta_dev, int inverse, int bg, hipStream_t s) {
  return launch_compose<1>(ox, oy, oz, gx, gy, gz, vx, vy, vz, nx, ny, nz, sp[0], sp[1], sp[2],
                           delta, delta_dev, inverse ? -1.f : 1.f, bg, s);
}

awk_status launch_compose_vh(float *ox, float *oy, float *oz, const float *vx, const float *vy,
                             const float *vz, const float *gx, const float *gy, const float *gz,
                             int nx, int ny, int nz, const float sp[3], const float *delta_dev,
                             int inverse, int bg, hipStream_t s) {
  return launch_compose<2>(ox, oy, oz, vx, vy, vz, gx, gy, gz, nx, ny, nz, sp[0], sp[1], sp[2],
                           1.f, delta_dev, inverse ? -1.f : 1.f, bg, s);
}

}  // namespace awk

using namespace awk;

// AoS (x0 y0 z0 x1 ...) <-> SoA planes, one voxel per thread: the three planes
// are read or written coalesced, the AoS side as 12-byte strides that the vector
// L1 combines into whole lines.
__global__ void k_aos_to_soa(const float *__restrict__ aos, float *__restrict__ x,
                             float *__restrict__ y, float *__restrict__ z, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = aos[3 * i];
  y[i] = aos[3 * i + 1];
  z[i] = aos[3 * i + 2];
}
__global__ void k_soa_to_aos(const float *__restrict__ x, const float *__restrict__ y,
                             const float *__restrict__ z, float *__restrict__ aos, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  aos[3 * i] = x[i];
  aos[3 * i + 1] = y[i];
  aos[3 * i + 2] = z[i];
}

extern "C" {

const char *awk_last_error(void) { return g_err.c_str(); }

int awk_version(void) { return 100; }

awk_status awk_tune_set(const char *name, int value) {
  if (!name) return fail(AWK_ERR_INVALID, "awk_tune_set: null name");
  const std::string n(name);
  Tuning &t = tuning();
  if (n == "apply_vpt" || n == "compose_vpt" || n == "xinv_unroll") {
    if (value != 1 && value != 2 && value != 4) return fail(AWK_ERR_INVALID, "awk_tune_set: value must be 1, 2 or 4");
    if (n == "apply_vpt") t.apply_vpt = value;
    else if (n == "compose_vpt") t.compose_vpt = value;
    else t.xinv_unroll = value;
    return AWK_OK;
  }
  if (n == "unet_cc") {
    if (value != 0 && value != 8 && value != 16 && value != 32)
      return fail(AWK_ERR_INVALID, "awk_tune_set: unet_cc must be 0, 8, 16 or 32");
    t.unet_cc = value;
    return AWK_OK;
  }
  if (n == "xinv_band" || n == "xfwd_band") {
    if (value != 0 && value != 2 && value != 4 && value != 8 && value != 16 && value != 32)
      return fail(AWK_ERR_INVALID, "awk_tune_set: " + n + " must be 0, 2, 4, 8, 16 or 32");
    (n == "xinv_band" ? t.xinv_band : t.xfwd_band) = value;
    return AWK_OK;
  }
  if (n == "atlas_mean") {
    if (value != 0 && value != 1 && value != 2 && value != 4)
      return fail(AWK_ERR_INVALID, "awk_tune_set: atlas_mean must be 0, 1, 2 or 4");
    t.atlas_mean = value;
    return AWK_OK;
  }
  if (n == "unet_pipe") {
    if (value < 0 || value > 2) return fail(AWK_ERR_INVALID, "awk_tune_set: unet_pipe must be 0, 1 or 2");
    t.unet_pipe = value;
    return AWK_OK;
  }
  if (n == "splat_tile" || n == "fft_order" || n == "unet_lds" || n == "xinv_pairs" || n == "yfwd_rev") {
    if (value != 0 && value != 1) return fail(AWK_ERR_INVALID, "awk_tune_set: " + n + " must be 0 or 1");
    (n == "splat_tile"   ? t.splat_tile
     : n == "fft_order"  ? t.fft_order
     : n == "unet_lds"   ? t.unet_lds
     : n == "xinv_pairs" ? t.xinv_pairs
                         : t.yfwd_rev) = value;
    return AWK_OK;
  }
  if (n == "xinv_stage") {
    if (value < 0 || value > 2) return fail(AWK_ERR_INVALID, "awk_tune_set: xinv_stage must be 0, 1 or 2");
    t.xinv_stage = value;
    return AWK_OK;
  }
  if (n == "x_rt") {
    if (value < 0 || value > 8 || value % 2) return fail(AWK_ERR_INVALID, "awk_tune_set: x_rt must be 0, 2, 4, 6 or 8");
    t.x_rt = value;
    return AWK_OK;
  }
  if (n == "y_tx" || n == "z_tx") {
    const int mx = n == "y_tx" ? 16 : 8;
    if (value < 0 || value > mx || (value & (value - 1)))
      return fail(AWK_ERR_INVALID, "awk_tune_set: " + n + " must be 0 or a power of two up to " + std::to_string(mx));
    (n == "y_tx" ? t.y_tx : t.z_tx) = value;
    return AWK_OK;
  }
  return fail(AWK_ERR_INVALID, "awk_tune_set: unknown knob " + n);
}

awk_status awk_tune_get(const char *name, int *value) {
  if (!name || !value) return fail(AWK_ERR_INVALID, "awk_tune_get: null argument");
  const std::string n(name);
  const Tuning &t = tuning();
  if (n == "apply_vpt") *value = t.apply_vpt;
  else if (n == "compose_vpt") *value = t.compose_vpt;
  else if (n == "xinv_unroll") *value = t.xinv_unroll;
  else if (n == "splat_tile") *value = t.splat_tile;
  else if (n == "fft_order") *value = t.fft_order;
  else if (n == "unet_lds") *value = t.unet_lds;
  else if (n == "xinv_pairs") *value = t.xinv_pairs;
  else if (n == "unet_pipe") *value = t.unet_pipe;
  else if (n == "unet_cc") *value = t.unet_cc;
  else if (n == "atlas_mean") *value = t.atlas_mean;
  else if (n == "xinv_band") *value = t.xinv_band;
  else if (n == "xfwd_band") *value = t.xfwd_band;
  else if (n == "yfwd_rev") *value = t.yfwd_rev;
  else if (n == "xinv_stage") *value = t.xinv_stage;
  else if (n == "x_rt") *value = t.x_rt;
  else if (n == "y_tx") *value = t.y_tx;
  else if (n == "z_tx") *value = t.z_tx;
  else return fail(AWK_ERR_INVALID, "awk_tune_get: unknown knob " + n);
  return AWK_OK;
}

awk_status awk_vf_aos_to_soa(const float *aos, float *x, float *y, float *z, long long n,
                             awk_stream_t stream) {
  if (n < 0) return fail(AWK_ERR_INVALID, "awk_vf_aos_to_soa: negative size");
  if (n == 0) return AWK_OK;
  if (!aos || !x || !y || !z) return fail(AWK_ERR_INVALID, "awk_vf_aos_to_soa: null pointer");
  AWK_LAUNCH(k_aos_to_soa, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
             as_stream(stream), aos, x, y, z, n);
  AWK_LAUNCH_CHECK("k_aos_to_soa");
  return AWK_OK;
}

awk_status awk_vf_soa_to_aos(const float *x, const float *y, const float *z, float *aos, long long n,
                             awk_stream_t stream) {
  if (n < 0) return fail(AWK_ERR_INVALID, "awk_vf_soa_to_aos: negative size");
  if (n == 0) return AWK_OK;
  if (!aos || !x || !y || !z) return fail(AWK_ERR_INVALID, "awk_vf_soa_to_aos: null pointer");
  AWK_LAUNCH(k_soa_to_aos, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
             as_stream(stream), x, y, z, aos, n);
  AWK_LAUNCH_CHECK("k_soa_to_aos");
  return AWK_OK;
}

awk_status awk_hfield_set_to_identity(float *hx, float *hy, float *hz, int nx, int ny, int nz,
                                      awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_set_to_identity")) return st;
  if (!hx || !hy || !hz) return fail(AWK_ERR_INVALID, "awk_hfield_set_to_identity: null pointer");
  AWK_LAUNCH(k_identity, dim3(grid_for((long long)nx * ny * nz, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), hx, hy, hz, nx, ny, nz);
  AWK_LAUNCH_CHECK("k_identity");
  return AWK_OK;
}

awk_status awk_hfield_velocity_to_h(float *x, float *y, float *z, int nx, int ny, int nz, float spx,
                                    float spy, float spz, awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_velocity_to_h")) return st;
  if (!x || !y || !z) return fail(AWK_ERR_INVALID, "awk_hfield_velocity_to_h: null pointer");
  AWK_LAUNCH(k_hv_convert, dim3(grid_for((long long)nx * ny * nz, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), x, y, z, nx, ny, nz, spx, spy, spz, 0);
  AWK_LAUNCH_CHECK("k_hv_convert");
  return AWK_OK;
}

awk_status awk_hfield_h_to_velocity(float *x, float *y, float *z, int nx, int ny, int nz, float spx,
                                    float spy, float spz, awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_h_to_velocity")) return st;
  if (!x || !y || !z) return fail(AWK_ERR_INVALID, "awk_hfield_h_to_velocity: null pointer");
  AWK_LAUNCH(k_hv_convert, dim3(grid_for((long long)nx * ny * nz, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), x, y, z, nx, ny, nz, spx, spy, spz, 1);
  AWK_LAUNCH_CHECK("k_hv_convert");
  return AWK_OK;
}

awk_status awk_hfield_inverse_zeroth_order(float *ox, float *oy, float *oz, const float *hx,
                                           const float *hy, const float *hz, int nx, int ny, int nz,
                                           awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_inverse_zeroth_order")) return st;
  if (!ox || !oy || !oz || !hx || !hy || !hz)
    return fail(AWK_ERR_INVALID, "awk_hfield_inverse_zeroth_order: null pointer");
  AWK_LAUNCH(k_inv0, dim3(grid_for((long long)nx * ny * nz, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), ox, oy, oz, hx, hy, hz, nx, ny, nz);
  AWK_LAUNCH_CHECK("k_inv0");
  return AWK_OK;
}

awk_status awk_hfield_apply(float *out, const float *img, const float *hx, const float *hy,
                            const float *hz, int nx, int ny, int nz, int bg, float background,
                            awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_apply")) return st;
  if (!out || !img || !hx || !hy || !hz) return fail(AWK_ERR_INVALID, "awk_hfield_apply: null pointer");
  return launch_apply(out, img, hx, hy, hz, nx, ny, nz, bg, background, as_stream(stream));
}

awk_status awk_hfield_apply_grid(float *out, const float *img, int inx, int iny, int inz,
                                 const float *hx, const float *hy, const float *hz, int nx, int ny,
                                 int nz, int bg, float background, awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_apply_grid")) return st;
  if (awk_status st = check_dims(inx, iny, inz, "awk_hfield_apply_grid (image)")) return st;
  if (!out || !img || !hx || !hy || !hz) return fail(AWK_ERR_INVALID, "awk_hfield_apply_grid: null pointer");
  return launch_apply_grid(out, img, inx, iny, inz, hx, hy, hz, nx, ny, nz, bg, background,
                           as_stream(stream));
}

awk_status awk_hfield_init_affine(float *hx, float *hy, float *hz, int nx, int ny, int nz,
                                  const float m[12], awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_init_affine")) return st;
  if (!hx || !hy || !hz || !m) return fail(AWK_ERR_INVALID, "awk_hfield_init_affine: null pointer");
  Affine34f a;
  for (int i = 0; i < 12; ++i) a.m[i] = m[i];
  AWK_LAUNCH(k_affine, dim3(grid_for((long long)nx * ny * nz, kBlock)), dim3(kBlock), 0,
             as_stream(stream), hx, hy, hz, nx, ny, nz, a);
  AWK_LAUNCH_CHECK("k_affine");
  return AWK_OK;
}

awk_status awk_hfield_compose(float *ox, float *oy, float *oz, const float *fx, const float *fy,
                              const float *fz, const float *gx, const float *gy, const float *gz,
                              int nx, int ny, int nz, int bg, awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_compose")) return st;
  if (!ox || !oy || !oz || !fx || !fy || !fz || !gx || !gy || !gz)
    return fail(AWK_ERR_INVALID, "awk_hfield_compose: null pointer");
  if (ox == fx || oy == fy || oz == fz) return fail(AWK_ERR_INVALID, "awk_hfield_compose: output aliases f");
  return launch_compose<0>(ox, oy, oz, fx, fy, fz, gx, gy, gz, nx, ny, nz, 1.f, 1.f, 1.f, 1.f,
                           nullptr, 1.f, bg, as_stream(stream));
}

awk_status awk_hfield_compose_hv(float *ox, float *oy, float *oz, const float *gx, const float *gy,
                                 const float *gz, const float *vx, const float *vy, const float *vz,
                                 int nx, int ny, int nz, float spx, float spy, float spz,
                                 float delta, int inverse, int bg, awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_compose_hv")) return st;
  if (!ox || !oy || !oz || !gx || !gy || !gz || !vx || !vy || !vz)
    return fail(AWK_ERR_INVALID, "awk_hfield_compose_hv: null pointer");
  if (ox == gx || oy == gy || oz == gz) return fail(AWK_ERR_INVALID, "awk_hfield_compose_hv: output aliases g");
  const float sp[3] = {spx, spy, spz};
  return launch_compose_hv(ox, oy, oz, gx, gy, gz, vx, vy, vz, nx, ny, nz, sp, delta, nullptr,
                           inverse, bg, as_stream(stream));
}

awk_status awk_hfield_compose_vh(float *ox, float *oy, float *oz, const float *vx, const float *vy,
                                 const float *vz, const float *gx, const float *gy, const float *gz,
                                 int nx, int ny, int nz, float spx, float spy, float spz, int inverse,
                                 int bg, awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_hfield_compose_vh")) return st;
  if (!ox || !oy || !oz || !gx || !gy || !gz || !vx || !vy || !vz)
    return fail(AWK_ERR_INVALID, "awk_hfield_compose_vh: null pointer");
  if (ox == vx || oy == vy || oz == vz) return fail(AWK_ERR_INVALID, "awk_hfield_compose_vh: output aliases v");
  const float sp[3] = {spx, spy, spz};
  return launch_compose_vh(ox, oy, oz, vx, vy, vz, gx, gy, gz, nx, ny, nz, sp, nullptr, inverse,
                           bg, as_stream(stream));
}

awk_status awk_gradient(float *gx, float *gy, float *gz, const float *img, int nx, int ny, int nz,
                        float spx, float spy, float spz, int wrap, awk_stream_t stream) {
  if (awk_status st = check_dims(nx, ny, nz, "awk_gradient")) return st;
  if (!gx || !gy || !gz || !img) return fail(AWK_ERR_INVALID, "awk_gradient: null pointer");
  AWK_LAUNCH(k_gradient, dim3(grid_for((long long)nx * ny * nz, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), gx, gy, gz, img, nx, ny, nz, spx, spy, spz, wrap);
  AWK_LAUNCH_CHECK("k_gradient");
  return AWK_OK;
}

}  // extern "C"
