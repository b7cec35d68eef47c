// awk_multiscale.hip -- scale-level transitions on gfx950: Gaussian downsampling
// of images and resampling of vector fields.
//
// ImageUtils::gaussianDownsample (UtilitiesDataTypes/ImageUtils.txx:255-290) =
// GaussianFilter3D::filter_symmetric_full (Utilities/DownsampleFilter/
// GaussianFilter3D.cxx:159-330, separable, renormalised one-sided kernels within
// k of an edge) followed by resampleNew (ImageUtils.txx:320-347) at the centred
// points f x + (f - 1)/2; HField3DUtils::resampleNew (HField3DUtils.txx:1814-1885)
// for hfields and velocities between scale levels.  These run once per scale
// level, so the kernels are plain one-voxel-per-thread passes.
#include <cmath>
#include <vector>

#include "awk_ops.h"

namespace awk {

constexpr int kMaxHalf = 15;

struct GaussTab {
  int k;
  float f[kMaxHalf + 1][kMaxHalf + 1];  // row 0: kernel; row j: renormalised for j missing taps
};

// The kernel table: normalised half Gaussian (float, as GaussianFilter3D.cxx:92-140
// forms it) and its renormalisations for elements near an edge.
static GaussTab make_tab(float sigma, int k) {
  GaussTab t;
  t.k = k;
  const float es = (float)(1.0 / (2.0 * sigma * sigma));
  float sum = 1.0f;
  t.f[0][0] = 1.0f;
  for (int i = 1; i <= k; ++i) {
    t.f[0][i] = std::exp(-(float)(i * i) * es);
    sum += 2.0f * t.f[0][i];
  }
  const float inv = 1.0f / sum;
  for (int i = 0; i <= k; ++i) t.f[0][i] *= inv;
  float missing = 0.0f;
  for (int j = 1; j <= k; ++j) {
    missing += t.f[0][k + 1 - j];
    for (int i = 0; i <= k; ++i) t.f[j][i] = (float)(t.f[0][i] / (1.0 - missing));
  }
  return t;
}

// One axis of the separable filter: out(p) from in along `axis` (stride s, length n).
__global__ void __launch_bounds__(256)
k_gauss_axis(float *__restrict__ out, const float *__restrict__ in, int nx, int ny, int nz,
             int axis, GaussTab t) {
  const int N = nx * ny * nz;
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= N) return;
  const int x = id % nx, r = id / nx, y = r % ny, z = r / ny;
  const int p = axis == 0 ? x : axis == 1 ? y : z;
  const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
  const int s = axis == 0 ? 1 : axis == 1 ? nx : nx * ny;
  const float *c = in + id;
  const int k = t.k;
  float acc;
  if (p >= k && p < n - k) {
    acc = c[0] * t.f[0][0];
    for (int i = 1; i <= k; ++i) acc += (c[-i * s] + c[i * s]) * t.f[0][i];
  } else if (p < k) {
    // left edge: taps beyond the edge are missing, the kernel is renormalised
    const float *g = t.f[k - p];
    acc = c[0] * g[0];
    for (int i = 1; i <= p; ++i) acc += (c[-i * s] + c[i * s]) * g[i];
    for (int i = p + 1; i <= k; ++i) acc += c[i * s] * g[i];
  } else {
    const int q = n - 1 - p;
    const float *g = t.f[k - q];
    acc = c[0] * g[0];
    for (int i = 1; i <= q; ++i) acc += (c[i * s] + c[-i * s]) * g[i];
    for (int i = q + 1; i <= k; ++i) acc += c[-i * s] * g[i];
  }
  out[id] = acc;
}

// resampleNew of the filtered image: output voxel x reads the source at
// f x + (f - 1)/2 (trilinear, background 0).
__global__ void __launch_bounds__(256)
k_resample_image(float *__restrict__ out, int onx, int ony, int onz, const float *__restrict__ in,
                 Dims d, double fx, double fy, double fz) {
  const int N = onx * ony * onz;
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= N) return;
  const int x = id % onx, r = id / onx, y = r % ony, z = r / ony;
  const float px = (float)(fx * x + (fx - 1.0) / 2.0);
  const float py = (float)(fy * y + (fy - 1.0) / 2.0);
  const float pz = (float)(fz * z + (fz - 1.0) / 2.0);
  out[id] = trilerp1g<AWK_BG_VAL>(make_buf(in, d.nx * d.ny * d.nz), d, px, py, pz, 0.f);
}

// HField3DUtils::resampleNew: output x samples the input at (r - 1)/2 + x r, then
// optionally scales the vectors by 1/r.
template <int BG>
__global__ void __launch_bounds__(256)
k_resample_field(float *__restrict__ ox, float *__restrict__ oy, float *__restrict__ oz, int onx,
                 int ony, int onz, const float *__restrict__ ix, const float *__restrict__ iy,
                 const float *__restrict__ iz, Dims d, double rx, double ry, double rz, int rescale) {
  const int N = onx * ony * onz;
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= N) return;
  const int x = id % onx, r = id / onx, y = r % ony, z = r / ony;
  const float px = (float)((rx - 1.0) / 2.0 + x * rx);
  const float py = (float)((ry - 1.0) / 2.0 + y * ry);
  const float pz = (float)((rz - 1.0) / 2.0 + z * rz);
  const int n = d.nx * d.ny * d.nz;
  float3 v = trilerp3g<BG>(make_buf(ix, n), make_buf(iy, n), make_buf(iz, n), d, px, py, pz);
  if (rescale) {
    v.x *= (float)(1.0 / rx);
    v.y *= (float)(1.0 / ry);
    v.z *= (float)(1.0 / rz);
  }
  ox[id] = v.x;
  oy[id] = v.y;
  oz[id] = v.z;
}

// h += sign * M [x y z 1]
__global__ void __launch_bounds__(256)
k_add_affine(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz, int nx, int ny,
             int nz, Affine34f m, float sign) {
  const int N = nx * ny * nz;
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= N) return;
  const float x = (float)(id % nx), y = (float)((id / nx) % ny), z = (float)(id / (nx * ny));
  hx[id] += sign * (m.m[0] * x + m.m[1] * y + m.m[2] * z + m.m[3]);
  hy[id] += sign * (m.m[4] * x + m.m[5] * y + m.m[6] * z + m.m[7]);
  hz[id] += sign * (m.m[8] * x + m.m[9] * y + m.m[10] * z + m.m[11]);
}

// h = M [x + u(x), 1]: a displacement on a grid mapped through an affine base map
// (h may alias u)
__global__ void __launch_bounds__(256)
k_affine_displace(float *hx, float *hy, float *hz, const float *ux, const float *uy, const float *uz,
                  int nx, int ny, int nz, Affine34f m) {
  const int N = nx * ny * nz;
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= N) return;
  const float x = (float)(id % nx) + ux[id], y = (float)((id / nx) % ny) + uy[id],
              z = (float)(id / (nx * ny)) + uz[id];
  const float a = m.m[0] * x + m.m[1] * y + m.m[2] * z + m.m[3];
  const float b = m.m[4] * x + m.m[5] * y + m.m[6] * z + m.m[7];
  const float c = m.m[8] * x + m.m[9] * y + m.m[10] * z + m.m[11];
  hx[id] = a;
  hy[id] = b;
  hz[id] = c;
}

static awk_status gauss_passes(float *out, const float *in, float *tmp, int nx, int ny, int nz,
                               const float sigma[3], const int half[3], hipStream_t s) {
  const int N = nx * ny * nz;
  const int dims[3] = {nx, ny, nz};
  const unsigned grid = (unsigned)((N + 255) / 256);
  // x: in -> tmp, y: tmp -> out, z: out -> tmp, then copy tmp -> out
  const float *src[3] = {in, tmp, out};
  float *dst[3] = {tmp, out, tmp};
  for (int a = 0; a < 3; ++a) {
    // GaussianFilter3D.cxx:178-184 clips the half size to n/2 - 1; below 2 voxels
    // (where the reference's unsigned arithmetic skips the clip) no filtering
    const int kmax = dims[a] / 2 - 1 > 0 ? dims[a] / 2 - 1 : 0;
    const int k = half[a] < 0 ? 0 : (half[a] > kmax ? kmax : half[a]);
    if (k > kMaxHalf) return fail(AWK_ERR_UNSUPPORTED, "gaussian filter: half kernel size above 15");
    const GaussTab t = make_tab(sigma[a], k);
    AWK_LAUNCH(k_gauss_axis, dim3(grid), dim3(256), 0, s, dst[a], src[a], nx, ny, nz, a, t);
    AWK_LAUNCH_CHECK("k_gauss_axis");
  }
  AWK_HIP(hipMemcpyAsync(out, tmp, (size_t)N * sizeof(float), hipMemcpyDeviceToDevice, s));
  return AWK_OK;
}

}  // namespace awk

using namespace awk;

static awk_status check_grid(int nx, int ny, int nz, const char *who) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(AWK_ERR_INVALID, std::string(who) + ": sizes must be >= 1");
  if ((long long)nx * ny * nz > kBufMaxElems) return fail(AWK_ERR_INVALID, std::string(who) + ": volume too large");
  return AWK_OK;
}

extern "C" {

awk_status awk_gaussian_filter(float *out, const float *in, int nx, int ny, int nz, float sigma_x,
                               float sigma_y, float sigma_z, int half_x, int half_y, int half_z,
                               awk_stream_t stream) {
  if (awk_status st = check_grid(nx, ny, nz, "awk_gaussian_filter")) return st;
  if (!out || !in) return fail(AWK_ERR_INVALID, "awk_gaussian_filter: null pointer");
  if (out == in) return fail(AWK_ERR_INVALID, "awk_gaussian_filter: out aliases in");
  if (!(sigma_x > 0 && sigma_y > 0 && sigma_z > 0)) return fail(AWK_ERR_INVALID, "awk_gaussian_filter: sigma must be > 0");
  hipStream_t s = as_stream(stream);
  const size_t N = (size_t)nx * ny * nz;
  float *tmp = nullptr;
  AWK_HIP(hipMallocAsync((void **)&tmp, N * sizeof(float), s));
  const float sg[3] = {sigma_x, sigma_y, sigma_z};
  const int hk[3] = {half_x, half_y, half_z};
  awk_status st = gauss_passes(out, in, tmp, nx, ny, nz, sg, hk, s);
  (void)hipFreeAsync(tmp, s);
  return st;
}

awk_status awk_gaussian_downsample(float *out, int onx, int ony, int onz, const float *in, int nx,
                                   int ny, int nz, awk_stream_t stream) {
  if (awk_status st = check_grid(nx, ny, nz, "awk_gaussian_downsample")) return st;
  if (awk_status st = check_grid(onx, ony, onz, "awk_gaussian_downsample (output)")) return st;
  if (!out || !in) return fail(AWK_ERR_INVALID, "awk_gaussian_downsample: null pointer");
  hipStream_t s = as_stream(stream);
  const size_t N = (size_t)nx * ny * nz;
  const double f[3] = {(double)nx / onx, (double)ny / ony, (double)nz / onz};
  float sg[3];
  int hk[3];
  for (int a = 0; a < 3; ++a) {
    const double sd = std::sqrt(f[a] / 2.0);
    sg[a] = (float)sd;
    hk[a] = 2 * (int)std::ceil(sd);
  }
  float *buf = nullptr;
  AWK_HIP(hipMallocAsync((void **)&buf, 2 * N * sizeof(float), s));
  awk_status st = gauss_passes(buf, in, buf + N, nx, ny, nz, sg, hk, s);
  if (!st) {
    const unsigned grid = (unsigned)(((long long)onx * ony * onz + 255) / 256);
    AWK_LAUNCH(k_resample_image, dim3(grid), dim3(256), 0, s, out, onx, ony, onz, (const float *)buf,
               Dims{nx, ny, nz}, f[0], f[1], f[2]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) st = hip_check(e, "k_resample_image");
  }
  (void)hipFreeAsync(buf, s);
  return st;
}

awk_status awk_hfield_resample(float *ox, float *oy, float *oz, int onx, int ony, int onz,
                               const float *ix, const float *iy, const float *iz, int nx, int ny,
                               int nz, int bg, int rescale, awk_stream_t stream) {
  if (awk_status st = check_grid(nx, ny, nz, "awk_hfield_resample")) return st;
  if (awk_status st = check_grid(onx, ony, onz, "awk_hfield_resample (output)")) return st;
  if (!ox || !oy || !oz || !ix || !iy || !iz) return fail(AWK_ERR_INVALID, "awk_hfield_resample: null pointer");
  if (ox == ix || oy == iy || oz == iz) return fail(AWK_ERR_INVALID, "awk_hfield_resample: output aliases input");
  hipStream_t s = as_stream(stream);
  const double r[3] = {(double)nx / onx, (double)ny / ony, (double)nz / onz};
  const unsigned grid = (unsigned)(((long long)onx * ony * onz + 255) / 256);
  const Dims d{nx, ny, nz};
#define AWK_RS(B)                                                                                \
  AWK_LAUNCH((k_resample_field<B>), dim3(grid), dim3(256), 0, s, ox, oy, oz, onx, ony, onz, ix, \
             iy, iz, d, r[0], r[1], r[2], rescale ? 1 : 0)
  switch (bg) {
    case AWK_BG_ID: AWK_RS(AWK_BG_ID); break;
    case AWK_BG_ZERO: AWK_RS(AWK_BG_ZERO); break;
    case AWK_BG_PARTIAL_ID: AWK_RS(AWK_BG_PARTIAL_ID); break;
    case AWK_BG_PARTIAL_ZERO: AWK_RS(AWK_BG_PARTIAL_ZERO); break;
    case AWK_BG_CLAMP: AWK_RS(AWK_BG_CLAMP); break;
    case AWK_BG_WRAP: AWK_RS(AWK_BG_WRAP); break;
    default: return fail(AWK_ERR_INVALID, "awk_hfield_resample: unknown background strategy");
  }
#undef AWK_RS
  AWK_LAUNCH_CHECK("k_resample_field");
  return AWK_OK;
}

awk_status awk_hfield_affine_displace(float *hx, float *hy, float *hz, const float *ux,
                                      const float *uy, const float *uz, int nx, int ny, int nz,
                                      const float m[12], awk_stream_t stream) {
  if (awk_status st = check_grid(nx, ny, nz, "awk_hfield_affine_displace")) return st;
  if (!hx || !hy || !hz || !ux || !uy || !uz || !m)
    return fail(AWK_ERR_INVALID, "awk_hfield_affine_displace: null pointer");
  Affine34f a;
  for (int i = 0; i < 12; ++i) a.m[i] = m[i];
  const unsigned grid = (unsigned)(((long long)nx * ny * nz + 255) / 256);
  AWK_LAUNCH(k_affine_displace, dim3(grid), dim3(256), 0, as_stream(stream), hx, hy, hz, ux, uy, uz,
             nx, ny, nz, a);
  AWK_LAUNCH_CHECK("k_affine_displace");
  return AWK_OK;
}

awk_status awk_hfield_add_affine(float *hx, float *hy, float *hz, int nx, int ny, int nz,
                                 const float m[12], float sign, awk_stream_t stream) {
  if (awk_status st = check_grid(nx, ny, nz, "awk_hfield_add_affine")) return st;
  if (!hx || !hy || !hz || !m) return fail(AWK_ERR_INVALID, "awk_hfield_add_affine: null pointer");
  Affine34f a;
  for (int i = 0; i < 12; ++i) a.m[i] = m[i];
  const unsigned grid = (unsigned)(((long long)nx * ny * nz + 255) / 256);
  AWK_LAUNCH(k_add_affine, dim3(grid), dim3(256), 0, as_stream(stream), hx, hy, hz, nx, ny, nz, a,
             sign);
  AWK_LAUNCH_CHECK("k_add_affine");
  return AWK_OK;
}

}  // extern "C"
