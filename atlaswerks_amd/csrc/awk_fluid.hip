// awk_fluid.hip -- FluidKernelFFT: L and L^-1 of the Navier operator
//   L = -alpha*Lap - beta*grad(div) + gamma
// applied through a 3-D real FFT, on gfx950.
//
// Semantics follow DiffOper::ApplyInverseOperatorOnTheFly / ApplyOperatorOnTheFly
// (Algorithms/Common/DiffOper.cxx:747-905): scale by 1/N, forward r2c of each
// component, per-frequency 3x3 L built from the LUT of DiffOper.cxx:587-639,
// Cholesky solve (DiffOper.cxx:1002-1083) or multiply (1085-1116), c2r.
//
// Pass structure (each pass is one kernel, one HBM round trip of the spectrum):
//   X  r2c along x of row triples (x, y, z components of the same rows), two real
//      lines packed per complex line; writes the half spectrum [c][z][y][kx].
//   Y  complex FFT along y for a tile of kx columns.
//   Z  forward FFT along z for all 3 components of a kx tile, the per-frequency
//      solve (with the 1/N scale folded in), inverse FFT along z.
//   Y' inverse along y.   X' c2r along x with an epilogue (plain store, max-norm,
//      or the fused greedy hfield update).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <new>
#include <vector>

#include "awk_fluid.h"

namespace awk {

Axis AxisPlan::dev() const {
  Axis a;
  std::memset(&a, 0, sizeof(a));
  a.tw = tw;
  a.n = n;
  a.nst = (int)rad.size();
  for (int i = 0; i < a.nst; ++i) a.rad[i] = rad[i];
  a.m = m;
  a.chirp = chirp;
  a.bhat = bhat;
  a.mtw = mtw;
  a.mst = (int)mrad.size();
  for (int i = 0; i < a.mst; ++i) a.mrad[i] = mrad[i];
  return a;
}

// Unnormalised forward DFT of a power-of-two length in double (iterative
// radix 2), for the Bluestein kernel spectrum.
static void host_fft_pow2(std::vector<std::complex<double>> &a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = -2.0 * M_PI / (double)len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const std::complex<double> w(cos(ang * k), sin(ang * k));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

// Bluestein tables of one axis appended to the host table `tw`: the chirp
// c_j = exp(-i pi j^2 / n) (the angle from j^2 mod 2n, exact in integers), the
// spectrum of b_l = conj(c_|l|) laid cyclically over m and scaled by 1/m, and
// the m twiddles.  Returns the offsets of the three tables.
static void bluestein_tables(int n, int m, std::vector<float2> &tw, size_t off[3]) {
  off[0] = tw.size();
  std::vector<std::complex<double>> c(n), b(m, 0.0);
  for (int j = 0; j < n; ++j) {
    const long long r = ((long long)j * j) % (2LL * n);
    const double a = -M_PI * (double)r / (double)n;
    c[j] = std::complex<double>(cos(a), sin(a));
    tw.push_back(make_float2((float)c[j].real(), (float)c[j].imag()));
  }
  for (int j = 0; j < n; ++j) {
    b[j] = std::conj(c[j]);
    if (j > 0) b[m - j] = std::conj(c[j]);
  }
  host_fft_pow2(b);
  off[1] = tw.size();
  for (int k = 0; k < m; ++k)
    tw.push_back(make_float2((float)(b[k].real() / m), (float)(b[k].imag() / m)));
  off[2] = tw.size();
  for (int k = 0; k < m; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)m;
    tw.push_back(make_float2((float)cos(a), (float)sin(a)));
  }
}

static std::vector<int> factor(int n) {
  std::vector<int> r;
  int m = n;
  while (m % 16 == 0) { r.push_back(16); m /= 16; }
  if (m % 8 == 0) { r.push_back(8); m /= 8; }
  if (m % 4 == 0) { r.push_back(4); m /= 4; }
  if (m % 2 == 0) { r.push_back(2); m /= 2; }
  for (int p = 3; m > 1; p += 2)
    while (m % p == 0) { r.push_back(p); m /= p; }
  return r;
}

// ------------------------------------------------------------------ kernels

// X forward: r2c of RT spatial rows (3 components each) per block.
// Complex line layout per row pair (q = 2p, 2p+1):
//   line 3p   = (c0[2p],   c1[2p])
//   line 3p+1 = (c2[2p],   c2[2p+1])
//   line 3p+2 = (c0[2p+1], c1[2p+1])
__device__ __forceinline__ void line_targets(int line, int r0, int &rowA, int &cA, int &rowB,
                                             int &cB) {
  const int p = line / 3, e = line - 3 * p;
  const int ra = r0 + 2 * p;
  if (e == 0) { rowA = ra; cA = 0; rowB = ra; cB = 1; }
  else if (e == 1) { rowA = ra; cA = 2; rowB = ra + 1; cB = 2; }
  else { rowA = ra + 1; cA = 0; rowB = ra + 1; cB = 1; }
}

// Batched launches: subject blockIdx.y's reduction arguments.
__device__ __forceinline__ void batch_reduce(ReduceArgs &ra, const Batch &bt) {
  const long long sb = blockIdx.y;
  ra.partials += sb * bt.part;
  if (bt.scalar) {
    if (ra.energy_out) ra.energy_out += sb;
    ra.prev += sb;
    ra.have_prev += sb;
    ra.step += sb;
  }
}

// Energy bookkeeping of one greedy iteration, run by logical block 0 of the
// forward y pass (the kernel boundary after the x pass makes every partial
// visible): fixed-order FP64 sum of the x-pass block partials, then
// GreedyDeformationData::AddEnergy (GreedyDeformationData.cxx:162-184).
__device__ __forceinline__ void energy_block(const ReduceArgs &ra, double *tree, int bdim) {
  double acc = 0.0;
  acc = strided_fold(ra.partials, ra.n, (int)threadIdx.x, bdim, acc, FoldAdd());
  tree[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && (int)threadIdx.x + w < bdim) tree[threadIdx.x] += tree[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double e = ra.scale * tree[0];
    if (ra.energy_out) *ra.energy_out = e;
    const bool have = *ra.have_prev != 0;
    const double change = have ? e - *ra.prev : 0.0;
    unsigned f = 0;
    if (e != e || change != change) f |= AWK_FLAG_NAN_ENERGY;
    if (change > 0.0) f |= AWK_FLAG_ENERGY_INCREASE;
    raise_flags(ra.flags, f);
    if (ra.auto_reduce && change > 0.0) *ra.step = *ra.step * 0.5f;
    *ra.prev = e;
    *ra.have_prev = 1;
  }
}


// Force multiplier of the GRADMUL modes: the image i1, or an int64 fixed-point
// splat accumulator (2^-32 units), which is cleared as it is read (every voxel is
// read by exactly one thread of the x pass).  The image is read once per
// iteration: a non-temporal load, so that it does not displace in the memory-side
// cache the spectrum this pass writes and the y pass reads next.
__device__ __forceinline__ float load_mul(const XFwdArgs &a, size_t i) {
  if (a.mode == XLOAD_GRADMUL_FX) {
    const long long q = a.acc[i];
    a.acc[i] = 0;
    return splat_unq(q);
  }
  return __builtin_nontemporal_load(a.i1 + i);
}

// NC, RTC, BDC: compile-time length, rows per block and block size (0 = runtime).
template <int NC, int RTC, int BDC>
__global__ void __launch_bounds__(BDC ? BDC : 1024)
k_xfwd(XFwdArgs a, Axis ax, int nx_rt, int ny, int nz, int nxh_rt, int nxh_pad, int RT_rt,
       float2 *__restrict__ spec, size_t plane, Batch bt) {
  extern __shared__ float2 sm[];
  __shared__ double red[16];
  {  // subject blockIdx.y of a batched launch
    const long long sb = blockIdx.y;
    spec += sb * bt.spec;
    if (a.idef) a.idef += sb * bt.img;
    if (a.partials) a.partials += sb * bt.part;
  }
  const int nx = NC ? NC : nx_rt, nxh = NC ? NC / 2 + 1 : nxh_rt, RT = RTC ? RTC : RT_rt;
  const int bdim = BDC ? BDC : (int)blockDim.x;
  const int NL = 3 * RT / 2, LS = NL + 1;
  const int NR = ny * nz;
  const int lb = (int)xcd_block();
  const int g = band_block(lb, a.band, ny / RT);  // the row group (its SSD partial's index)
  const int r0 = g * RT;
  if (r0 >= NR) return;
  double ssd = 0.0;
  float *smf = reinterpret_cast<float *>(sm);
  auto put = [&](int q, int x, float f0, float f1, float f2) {
    const int p = q >> 1;
    if ((q & 1) == 0) {
      sm[x * LS + 3 * p] = make_float2(f0, f1);
      smf[2 * (x * LS + 3 * p + 1)] = f2;
    } else {
      sm[x * LS + 3 * p + 2] = make_float2(f0, f1);
      smf[2 * (x * LS + 3 * p + 1) + 1] = f2;
    }
  };
  bool done = false;
  if constexpr (NC != 0 && NC == BDC) {
    // Column mode: thread x owns column x of the block's RT rows, which are RT
    // consecutive y at one z.  Each Idef value is loaded once: the column with a
    // one-row halo in y, the two neighbouring z rows, and x +- 1 from the
    // neighbouring lanes (a load only at the wave edges).  Idef and I1 are loaded
    // non-temporally (the halo rows other blocks re-read come from L2): with the
    // apply's hfield loads and the composition's hfield stores also streamed,
    // the memory-side cache keeps this pass's spectrum for the y pass (256^3
    // greedy iteration: y pass 89 -> 78 us, whole iteration +4 %, measured).
    if (a.mode != XLOAD_PLAIN && ny % RTC == 0) {
      done = true;
      const int x = threadIdx.x;
      const int z = r0 / ny, y0 = r0 - z * ny;
      const size_t sxy = (size_t)NC * ny;
      const float *I = a.idef;
      const float *col0 = I + (size_t)z * sxy + (size_t)y0 * NC + x;
      float col[RTC + 2], zm[RTC], zp[RTC], i1v[RTC], lft[RTC], rgt[RTC];
#pragma unroll
      for (int k = 0; k < RTC + 2; ++k) {
        const int yy = y0 - 1 + k;
        col[k] = (yy >= 0 && yy < ny) ? __builtin_nontemporal_load(col0 + (ptrdiff_t)(k - 1) * NC) : 0.f;
      }
      const int lane = threadIdx.x & 63;
#pragma unroll
      for (int q = 0; q < RTC; ++q) {
        zm[q] = (z > 0) ? __builtin_nontemporal_load(col0 + (ptrdiff_t)q * NC - (ptrdiff_t)sxy) : 0.f;
        zp[q] = (z < nz - 1) ? __builtin_nontemporal_load(col0 + (ptrdiff_t)q * NC + (ptrdiff_t)sxy) : 0.f;
        i1v[q] = load_mul(a, (size_t)z * sxy + (size_t)(y0 + q) * NC + x);
        lft[q] = (lane == 0 && x > 0) ? col0[(ptrdiff_t)q * NC - 1] : 0.f;
        rgt[q] = (lane == 63 && x < NC - 1) ? col0[(ptrdiff_t)q * NC + 1] : 0.f;
      }
#pragma unroll
      for (int q = 0; q < RTC; ++q) {
        const float c = col[q + 1];
        const float up = __shfl_up(c, 1, 64), dn = __shfl_down(c, 1, 64);
        const float cl = (lane == 0) ? lft[q] : up;   // value at x - 1
        const float cr = (lane == 63) ? rgt[q] : dn;  // value at x + 1
        const int y = y0 + q;
        // Array3DUtils::computeGradient (Array3DUtils.txx:783-901), wrap = false
        float gx, gy, gz;
        if (x > 0 && x < NC - 1) gx = div_const((cr - cl), 2.f * a.sp[0], a.r2sp[0]);
        else if (x == 0) gx = div_const((cr - c), a.sp[0], a.rsp[0]);
        else gx = div_const((c - cl), a.sp[0], a.rsp[0]);
        if (ny == 1) gy = 0.f;
        else if (y > 0 && y < ny - 1) gy = div_const((col[q + 2] - col[q]), 2.f * a.sp[1], a.r2sp[1]);
        else if (y == 0) gy = div_const((col[q + 2] - c), a.sp[1], a.rsp[1]);
        else gy = div_const((c - col[q]), a.sp[1], a.rsp[1]);
        if (nz == 1) gz = 0.f;
        else if (z > 0 && z < nz - 1) gz = div_const((zp[q] - zm[q]), 2.f * a.sp[2], a.r2sp[2]);
        else if (z == 0) gz = div_const((zp[q] - c), a.sp[2], a.rsp[2]);
        else gz = div_const((c - zm[q]), a.sp[2], a.rsp[2]);
        const float r = (a.mode == XLOAD_FORCE) ? i1v[q] - c : i1v[q];
        ssd += (double)r * (double)r;
        put(q, x, gx * r, gy * r, gz * r);
      }
    }
  }
  if (!done) {
#pragma unroll
  for (int w = threadIdx.x; w < RT * nx; w += bdim) {
    const int q = w / nx, x = w - q * nx;
    const int row = r0 + q;
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    if (row < NR) {
      const size_t base = (size_t)row * nx + x;
      if (a.mode == XLOAD_PLAIN) {
        f0 = a.in[0][base];
        f1 = a.in[1][base];
        f2 = a.in[2][base];
      } else {
        // gradient of the deformed image (Array3DUtils::computeGradient, wrap=false,
        // Array3DUtils.txx:783-901), residual r = I1 - Idef and force r*grad
        // (GreedyIteratorCPU.cxx:108-117).
        const int y = row % ny, z = row / ny;
        const float *I = a.idef;
        const float c = I[base];
        float gx = 0.f, gy = 0.f, gz = 0.f;
        if (nx > 1) {
          if (x > 0 && x < nx - 1) gx = div_const((I[base + 1] - I[base - 1]), 2.f * a.sp[0], a.r2sp[0]);
          else if (x == 0) gx = div_const((I[base + 1] - c), a.sp[0], a.rsp[0]);
          else gx = div_const((c - I[base - 1]), a.sp[0], a.rsp[0]);
        }
        if (ny > 1) {
          if (y > 0 && y < ny - 1) gy = div_const((I[base + nx] - I[base - nx]), 2.f * a.sp[1], a.r2sp[1]);
          else if (y == 0) gy = div_const((I[base + nx] - c), a.sp[1], a.rsp[1]);
          else gy = div_const((c - I[base - nx]), a.sp[1], a.rsp[1]);
        }
        if (nz > 1) {
          const size_t s = (size_t)nx * ny;
          if (z > 0 && z < nz - 1) gz = div_const((I[base + s] - I[base - s]), 2.f * a.sp[2], a.r2sp[2]);
          else if (z == 0) gz = div_const((I[base + s] - c), a.sp[2], a.rsp[2]);
          else gz = div_const((c - I[base - s]), a.sp[2], a.rsp[2]);
        }
        const float m = load_mul(a, base);
        const float r = (a.mode == XLOAD_FORCE) ? m - c : m;
        ssd += (double)r * (double)r;
        f0 = gx * r;
        f1 = gy * r;
        f2 = gz * r;
      }
    }
    put(q, x, f0, f1, f2);
  }
  }
  __syncthreads();
  if constexpr (NC != 0) ct_fft<NC, 3 * RTC / 2, 3 * RTC / 2 + 1, BDC>(sm, ax.tw, false);
  else lds_fft(sm, ax, NL, LS, false, sm + (size_t)nx * LS);
#pragma unroll
  for (int w = threadIdx.x; w < NL * nxh; w += bdim) {
    const int line = w / nxh, k = w - line * nxh;
    const float2 Z = sm[k * LS + line];
    const float2 Zc = cconj(sm[((nx - k) % nx) * LS + line]);
    const float2 A = cscale(cadd(Z, Zc), 0.5f);
    const float2 d = cscale(csub(Z, Zc), 0.5f);
    const float2 B = make_float2(d.y, -d.x);
    int rowA, cA, rowB, cB;
    line_targets(line, r0, rowA, cA, rowB, cB);
    if (rowA < NR) spec[cA * plane + (size_t)rowA * nxh_pad + k] = A;
    if (rowB < NR) spec[cB * plane + (size_t)rowB * nxh_pad + k] = B;
  }
  if (a.mode == XLOAD_FORCE && a.partials) {
    // deterministic block reduction of the FP64 SSD partial
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ssd += __shfl_down(ssd, off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = ssd;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      const int nw = (bdim + 63) >> 6;
      for (int i = 0; i < nw; ++i) s += red[i];
      a.partials[g] = s;
    }
  }
}

// X inverse: c2r of RT rows (3 components) with an epilogue.  APPLY (COMPOSE
// only): the epilogue also samples I0 at the new field (a.i0 -> a.idef), the next
// iteration's deformed image.  STG (COMPOSE only; NC, RTC, BDC set, ny % RTC == 0):
// the composition reads the current field from the block's rows and their
// neighbours staged in LDS.
template <int NC, int RTC, int BDC, int XU, bool PAIRS = false, bool APPLY = false, bool STG = false>
__global__ void __launch_bounds__(BDC ? BDC : 1024)
k_xinv(XInvArgs a, Axis ax, int nx_rt, int ny, int nz, int nxh_rt, int nxh_pad, int RT_rt,
       const float2 *__restrict__ spec, size_t plane, Batch bt) {
  extern __shared__ float2 sm[];
  __shared__ float redf[16];
  {  // subject blockIdx.y of a batched launch
    const long long sb = blockIdx.y;
    spec += sb * bt.spec;
    for (int c = 0; c < 3; ++c) {
      if (a.h[c]) a.h[c] += sb * bt.field;
      if (a.hnew[c]) a.hnew[c] += sb * bt.field;
      if (a.out[c]) a.out[c] += sb * bt.field;
    }
    if (bt.scalar) a.step += sb;
    if (a.maxpart) a.maxpart += sb * (long long)gridDim.x;
  }
  const int nx = NC ? NC : nx_rt, nxh = NC ? NC / 2 + 1 : nxh_rt, RT = RTC ? RTC : RT_rt;
  const int bdim = BDC ? BDC : (int)blockDim.x;
  const int NL = 3 * RT / 2, LS = NL + 1;
  const int NR = ny * nz;
  const int lb = (int)xcd_block();
  const int g = band_block(lb, a.band, ny / RT);  // the row group
  const int r0 = g * RT;
  if (r0 >= NR) {
    // a padding block of the grid: its max partial is still read
    if ((a.mode == XSTORE_MAXNORM || a.mode == XSTORE_MAXAXPY) && threadIdx.x == 0) a.maxpart[lb] = 0.f;
    return;
  }
#pragma unroll
  for (int w = threadIdx.x; w < NL * nxh; w += bdim) {
    const int line = w / nxh, k = w - line * nxh;
    int rowA, cA, rowB, cB;
    line_targets(line, r0, rowA, cA, rowB, cB);
    float2 A = rowA < NR ? spec[cA * plane + (size_t)rowA * nxh_pad + k] : make_float2(0.f, 0.f);
    float2 B = rowB < NR ? spec[cB * plane + (size_t)rowB * nxh_pad + k] : make_float2(0.f, 0.f);
    if (k == 0 || 2 * k == nx) {  // the imaginary part c2r ignores
      A.y = 0.f;
      B.y = 0.f;
    }
    sm[k * LS + line] = make_float2(A.x - B.y, A.y + B.x);               // A + iB
    if (k > 0 && nx - k >= nxh) sm[(nx - k) * LS + line] = make_float2(A.x + B.y, B.x - A.y);  // conj(A) + i conj(B)
  }
  __syncthreads();
  if constexpr (NC != 0) ct_fft<NC, 3 * RTC / 2, 3 * RTC / 2 + 1, BDC>(sm, ax.tw, true);
  else lds_fft(sm, ax, NL, LS, true, sm + (size_t)nx * LS);
  float mx = 0.f;
  if (a.mode == XSTORE_COMPOSE) {
    // updateDeformation (GreedyIteratorCPU.cxx:81-95): v *= step; h <- composeHV(h, v, sp, PARTIAL_ID).
    // XU voxels per step: their 24*XU gathers are in flight together.
    const float st = *a.step;
    const Dims d{nx, ny, nz};
    // v * step / sp, with the division by the spacing taken as a reciprocal
    // multiply (exact for unit spacing)
    const float isx = 1.f / a.sp[0], isy = 1.f / a.sp[1], isz = 1.f / a.sp[2];
    const int ntot = nx * ny * nz;
    const BufLoad hx = make_buf(a.h[0], ntot), hy = make_buf(a.h[1], ntot), hz = make_buf(a.h[2], ntot);
    if constexpr (STG) {
      // The block's RT rows are y = yb .. yb + RT - 1 of plane zr.  A point whose 8
      // corners lie inside the volume, in rows yb - 1 .. yb + RT of planes zr - 1 ..
      // zr + 1 (the common case: the step moves points by less than a voxel), reads
      // them from those rows staged in LDS, one component at a time; any other
      // point gathers from memory.  The same corners and the same lerp8 as the
      // gathers, so the same results bit for bit.
      constexpr int VP = RTC * NC / BDC;  // voxels per thread
      constexpr int WR = RTC + 2;         // staged rows per plane
      const int zr = r0 / ny, yb = r0 - zr * ny;
      float px[VP], py[VP], pz[VP];
      int win[VP];  // LDS offset of corner c0 in the staged rows (-1: gather from memory)
#pragma unroll
      for (int u = 0; u < VP; ++u) {
        const int w = threadIdx.x + u * BDC;
        const int q = w / NC, x = w - q * NC;
        const int p = q >> 1;
        float v0, v1, v2;
        if ((q & 1) == 0) {
          const float2 t = sm[x * LS + 3 * p];
          v0 = t.x; v1 = t.y; v2 = sm[x * LS + 3 * p + 1].x;
        } else {
          const float2 t = sm[x * LS + 3 * p + 2];
          v0 = t.x; v1 = t.y; v2 = sm[x * LS + 3 * p + 1].y;
        }
        px[u] = fmaf(v0 * st, isx, (float)x);
        py[u] = fmaf(v1 * st, isy, (float)(yb + q));
        pz[u] = fmaf(v2 * st, isz, (float)zr);
        const int x0 = (int)floorf(px[u]), y0 = (int)floorf(py[u]), z0 = (int)floorf(pz[u]);
        const bool ok = x0 >= 0 && x0 < nx - 1 && y0 >= 0 && y0 < ny - 1 && z0 >= 0 && z0 < nz - 1 &&
                        y0 >= yb - 1 && y0 <= yb + RTC - 1 && z0 >= zr - 1 && z0 <= zr;
        win[u] = ok ? ((z0 - zr + 1) * WR + (y0 - yb + 1)) * NC + x0 : -1;
      }
      __syncthreads();  // every read of the transformed rows is done: sm is reused
      float hc[VP][3];
#pragma unroll
      for (int u = 0; u < VP; ++u)
        if (win[u] < 0) {
          const float3 g = trilerp3g<AWK_BG_PARTIAL_ID, BufLoad, false>(hx, hy, hz, d, px[u], py[u], pz[u]);
          hc[u][0] = g.x; hc[u][1] = g.y; hc[u][2] = g.z;
        }
      float *stg = reinterpret_cast<float *>(sm);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (c) __syncthreads();  // the previous component's reads are done
        // rows outside the volume are not staged: no point in the window reads them
        const float *hs = a.h[c];
        for (int e = threadIdx.x; e < 3 * WR * (NC / 4); e += BDC) {
          const int row = e / (NC / 4), x4 = e - row * (NC / 4);
          const int zz = zr - 1 + row / WR, yy = yb - 1 + row % WR;
          if (zz >= 0 && zz < nz && yy >= 0 && yy < ny)
            *reinterpret_cast<float4 *>(stg + row * NC + 4 * x4) =
                *reinterpret_cast<const float4 *>(hs + ((size_t)zz * ny + yy) * NC + 4 * x4);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < VP; ++u)
          if (win[u] >= 0) {
            const float *b = stg + win[u];
            const float t = px[u] - floorf(px[u]), uu = py[u] - floorf(py[u]), vv = pz[u] - floorf(pz[u]);
            hc[u][c] = lerp8(b[0], b[1], b[NC + 1], b[NC], b[WR * NC + NC], b[WR * NC + NC + 1],
                             b[WR * NC + 1], b[WR * NC], t, uu, vv);
          }
      }
#pragma unroll
      for (int u = 0; u < VP; ++u) {
        // rows r0 .. r0 + RT - 1 are contiguous in memory
        const size_t base = (size_t)r0 * NC + threadIdx.x + u * BDC;
        __builtin_nontemporal_store(hc[u][0], a.hnew[0] + base);
        __builtin_nontemporal_store(hc[u][1], a.hnew[1] + base);
        __builtin_nontemporal_store(hc[u][2], a.hnew[2] + base);
      }
      if constexpr (APPLY) {
        const BufLoad ib = make_buf(a.i0, ntot);
#pragma unroll
        for (int u = 0; u < VP; ++u) {
          const size_t base = (size_t)r0 * NC + threadIdx.x + u * BDC;
          a.idef[base] = trilerp1g<AWK_BG_VAL>(ib, d, hc[u][0], hc[u][1], hc[u][2], a.i0_bg);
        }
      }
    } else {
#pragma unroll
    for (int w0 = threadIdx.x; w0 < RT * nx; w0 += bdim * XU) {
      float px[XU], py[XU], pz[XU];
      int row_[XU], x_[XU];
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        const int w = w0 + u * bdim;
        const int q = w / nx, x = w - q * nx;
        row_[u] = (w < RT * nx) ? r0 + q : NR;
        x_[u] = x;
        if (row_[u] < NR) {
          const int p = q >> 1;
          float v0, v1, v2;
          if ((q & 1) == 0) {
            const float2 t = sm[x * LS + 3 * p];
            v0 = t.x; v1 = t.y; v2 = sm[x * LS + 3 * p + 1].x;
          } else {
            const float2 t = sm[x * LS + 3 * p + 2];
            v0 = t.x; v1 = t.y; v2 = sm[x * LS + 3 * p + 1].y;
          }
          const int y = row_[u] % ny, z = row_[u] / ny;
          // explicit fma: the same rounding wherever the position is formed
          px[u] = fmaf(v0 * st, isx, (float)x);
          py[u] = fmaf(v1 * st, isy, (float)y);
          pz[u] = fmaf(v2 * st, isz, (float)z);
        }
      }
      float3 hn[XU];
#pragma unroll
      for (int u = 0; u < XU; ++u)
        if (row_[u] < NR) hn[u] = trilerp3g<AWK_BG_PARTIAL_ID, decltype(hx), PAIRS>(hx, hy, hz, d, px[u], py[u], pz[u]);
#pragma unroll
      for (int u = 0; u < XU; ++u)
        if (row_[u] < NR) {
          const size_t base = (size_t)row_[u] * nx + x_[u];
          // streamed: the next reader is the next iteration's apply (or, with
          // APPLY, the next iteration's composition gathers)
          __builtin_nontemporal_store(hn[u].x, a.hnew[0] + base);
          __builtin_nontemporal_store(hn[u].y, a.hnew[1] + base);
          __builtin_nontemporal_store(hn[u].z, a.hnew[2] + base);
        }
      if constexpr (APPLY) {
        // the next iteration's deformed image I0(hnew), from the new field still in
        // registers: the same sampling as the apply kernel on the stored field
        const BufLoad ib = make_buf(a.i0, ntot);
        float r[XU];
#pragma unroll
        for (int u = 0; u < XU; ++u)
          if (row_[u] < NR) r[u] = trilerp1g<AWK_BG_VAL>(ib, d, hn[u].x, hn[u].y, hn[u].z, a.i0_bg);
#pragma unroll
        for (int u = 0; u < XU; ++u)
          if (row_[u] < NR) a.idef[(size_t)row_[u] * nx + x_[u]] = r[u];
      }
    }
    }
  } else {
    // LDMM coefficients, in float as the reference forms them
    float cva = 0.f, cua = 0.f;
    if (a.mode == XSTORE_AXPBY) {
      const float st = *a.step;
      cva = 1.f - 2.f * st;
      cua = 2.f * st / a.sigma2;
    } else if (a.mode == XSTORE_MAXAXPY) {
      cua = -2.f / a.sigma2;
    }
#pragma unroll
    for (int w = threadIdx.x; w < RT * nx; w += bdim) {
      const int q = w / nx, x = w - q * nx;
      const int row = r0 + q;
      if (row >= NR) continue;
      const int p = q >> 1;
      float v0, v1, v2;
      if ((q & 1) == 0) {
        const float2 t = sm[x * LS + 3 * p];
        v0 = t.x; v1 = t.y; v2 = sm[x * LS + 3 * p + 1].x;
      } else {
        const float2 t = sm[x * LS + 3 * p + 2];
        v0 = t.x; v1 = t.y; v2 = sm[x * LS + 3 * p + 1].y;
      }
      const size_t base = (size_t)row * nx + x;
      if (a.mode == XSTORE_AXPBY) {
        // updateVelocity (LDMMIteratorCPU.cxx:175-190): v = vFac v + uFac u
        a.out[0][base] = cva * a.out[0][base] + cua * v0;
        a.out[1][base] = cva * a.out[1][base] + cua * v1;
        a.out[2][base] = cva * a.out[2][base] + cua * v2;
      } else if (a.mode == XSTORE_MAXAXPY) {
        // calcMaxDisplacement (LDMMIteratorCPU.cxx:192-212): |2 v - 2/sigma^2 u|
        const float q0 = 2.f * a.out[0][base] + cua * v0;
        const float q1 = 2.f * a.out[1][base] + cua * v1;
        const float q2 = 2.f * a.out[2][base] + cua * v2;
        mx = fmaxf(mx, sqrtf(q0 * q0 + q1 * q1 + q2 * q2));
      } else {
        a.out[0][base] = v0;
        a.out[1][base] = v1;
        a.out[2][base] = v2;
        if (a.mode == XSTORE_MAXNORM) mx = fmaxf(mx, sqrtf(v0 * v0 + v1 * v1 + v2 * v2));
      }
    }
  }
  if (a.mode == XSTORE_MAXNORM || a.mode == XSTORE_MAXAXPY) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_down(mx, off, 64));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) redf[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = 0.f;
      const int nw = (bdim + 63) >> 6;
      for (int i = 0; i < nw; ++i) m = fmaxf(m, redf[i]);
      a.maxpart[lb] = m;
    }
  }
}

// Y pass: complex FFT along y for TX kx columns of one (component, z) slab.
__global__ void __launch_bounds__(1024)
k_ypass(Axis ax, int ny, int nz, int nxh, int nxh_pad, int TX, float2 *__restrict__ spec,
        size_t plane, int inv, ReduceArgs ra, Batch bt) {
  extern __shared__ float2 sm[];
  spec += (long long)blockIdx.y * bt.spec;
  const int ntile = (nxh + TX - 1) / TX;
  const int lb = (int)xcd_block();
  if (lb >= 3 * nz * ntile) return;
  const int tile = lb % ntile, rest = lb / ntile;
  const int z = rest % nz, c = rest / nz;
  const int kx0 = tile * TX;
  float2 *base = spec + c * plane + (size_t)z * ny * nxh_pad + kx0;
  for (int w = threadIdx.x; w < TX * ny; w += blockDim.x) {
    const int t = w % TX, y = w / TX;
    sm[y * TX + t] = (kx0 + t < nxh) ? base[(size_t)y * nxh_pad + t] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft(sm, ax, TX, TX, inv != 0, sm + (size_t)ny * TX);
  for (int w = threadIdx.x; w < TX * ny; w += blockDim.x) {
    const int t = w % TX, y = w / TX;
    if (kx0 + t < nxh) base[(size_t)y * nxh_pad + t] = sm[y * TX + t];
  }
  if (ra.partials && lb == 0) {
    __syncthreads();
    batch_reduce(ra, bt);
    energy_block(ra, reinterpret_cast<double *>(sm), blockDim.x);
  }
}



// ------------------------------------------------------------------
// Register-resident two-stage passes for N = RA * RB (RA, RB in {8, 16}).
// Thread (t, j): t = kx column within the tile (lanes, 128-B rows for TX = 16),
// j = butterfly index.  Stage 1 reads its RA inputs straight from HBM, stage 2
// gets its inputs through one LDS transpose; the outputs of the forward stage 2
// stay in registers (Z pass: for the solve and the inverse stage 1).

template <int N, int RA, int RB, int TX>
__global__ void __launch_bounds__(TX *(N / RA > N / RB ? N / RA : N / RB))
k_ypass_reg(const float2 *__restrict__ tw, int nz, int nxh, int nxh_pad, float2 *__restrict__ spec,
            size_t plane, int inv, ReduceArgs ra, Batch bt) {
  constexpr int PA = N / RA, PB = N / RB;
  __shared__ float2 sm[N * TX];
  spec += (long long)blockIdx.y * bt.spec;
  const int t = threadIdx.x % TX, j = threadIdx.x / TX;
  const int ntile = (nxh + TX - 1) / TX;
  const int lb = (int)xcd_block();
  if (lb >= 3 * nz * ntile) return;
  const int tile = lb % ntile, rest = lb / ntile;
  // inv = 0: the forward pass visits (c, z) in reverse, which pays when the x pass
  // wrote the spectrum plane by plane (its last-written planes, the likeliest still
  // held by the memory-side cache, are read first; 256^3: 78.5 -> 75.0 us,
  // measured); inv = 2: in plane order, faster after the x pass's band order
  // (79.9 -> 78.0 us).  Every block transforms its own columns, so the order
  // changes no result.
  const bool rev = inv == 0;
  const int z = rev ? nz - 1 - rest % nz : rest % nz, c = rev ? 2 - rest / nz : rest / nz;
  const int kx = tile * TX + t;
  const bool valid = kx < nxh;
  float2 *base = spec + c * plane + (size_t)z * N * nxh_pad + kx;
  const bool iv = inv == 1;
  if (j < PA) {  // stage 1: radix RA, Ns = 1, inputs y = j + r * PA
    float2 v[RA];
#pragma unroll
    for (int r = 0; r < RA; ++r) v[r] = valid ? base[(size_t)(j + r * PA) * nxh_pad] : make_float2(0.f, 0.f);
    dft<RA>(v, iv);
#pragma unroll
    for (int r = 0; r < RA; ++r) sm[(j * RA + r) * TX + t] = v[r];
  }
  __syncthreads();
  if (j < PB) {  // stage 2: radix RB, Ns = RA, inputs j + r * PB, twiddle W^(r j), outputs j + r * RA
    float2 v[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) v[r] = sm[(j + r * PB) * TX + t];
#pragma unroll
    for (int r = 1; r < RB; ++r) v[r] = iv ? cmulc(v[r], tw[r * j]) : cmul(v[r], tw[r * j]);
    dft<RB>(v, iv);
    if (valid) {
#pragma unroll
      for (int r = 0; r < RB; ++r) base[(size_t)(j + r * RA) * nxh_pad] = v[r];
    }
  }
  if (ra.partials && lb == 0) {
    __syncthreads();
    batch_reduce(ra, bt);
    energy_block(ra, reinterpret_cast<double *>(sm), TX * (PA > PB ? PA : PB));
  }
}
struct SolveArgs {
  const float *cwx, *swx, *cwy, *swy, *cwz, *swz;
  const float *eig;  // 6 per frequency, or null
  float alpha, beta, gamma, scale;
  int inverse_op;  // 1: L^-1 (Cholesky), 0: L
  int divfree;
};


// Per-frequency operator: Cholesky solve of L v = b (DiffOper.cxx:1002-1083), or
// v = L b (1085-1116), or v = M b with the eigen table M = L^(+-p); then the
// divergence-free projection (948-1000) and the 1/N scale.
__device__ __forceinline__ void solve_one(const SolveArgs &s, int kx, int ky, int kz, int ny,
                                          int nxh, float2 &b0, float2 &b1, float2 &b2) {
  float2 v0, v1, v2;
  const float sx = s.swx[kx], sy = s.swy[ky], sz = s.swz[kz];
  if (s.eig) {
    const float *M = s.eig + ((size_t)(kz * ny + ky) * nxh + kx) * 6;
    const float M00 = M[0], M10 = M[1], M11 = M[2], M20 = M[3], M21 = M[4], M22 = M[5];
    v0 = make_float2(M00 * b0.x + M10 * b1.x + M20 * b2.x, M00 * b0.y + M10 * b1.y + M20 * b2.y);
    v1 = make_float2(M10 * b0.x + M11 * b1.x + M21 * b2.x, M10 * b0.y + M11 * b1.y + M21 * b2.y);
    v2 = make_float2(M20 * b0.x + M21 * b1.x + M22 * b2.x, M20 * b0.y + M21 * b1.y + M22 * b2.y);
  } else {
    const float cx = s.cwx[kx], cy = s.cwy[ky], cz = s.cwz[kz];
    const float lam = -s.alpha * (cx + cy + cz) + s.gamma;
    const float L00 = lam - s.beta * cx, L11 = lam - s.beta * cy, L22 = lam - s.beta * cz;
    const float L10 = s.beta * sx * sy, L20 = s.beta * sx * sz, L21 = s.beta * sy * sz;
    if (s.inverse_op) {
      // Same Cholesky factorisation, carried with the reciprocals g_ii = 1/G_ii
      // (rsqrt): every division of the reference becomes a multiplication.
      const float g00 = rsqrtf(L00);
      const float G10 = L10 * g00, G20 = L20 * g00;
      const float g11 = rsqrtf(L11 - G10 * G10);
      const float G21 = (L21 - G20 * G10) * g11;
      const float g22 = rsqrtf(L22 - (G20 * G20 + G21 * G21));
      float y0 = b0.x * g00, y1 = (b1.x - G10 * y0) * g11, y2 = (b2.x - G20 * y0 - G21 * y1) * g22;
      v2.x = y2 * g22; v1.x = (y1 - G21 * v2.x) * g11; v0.x = (y0 - G10 * v1.x - G20 * v2.x) * g00;
      y0 = b0.y * g00; y1 = (b1.y - G10 * y0) * g11; y2 = (b2.y - G20 * y0 - G21 * y1) * g22;
      v2.y = y2 * g22; v1.y = (y1 - G21 * v2.y) * g11; v0.y = (y0 - G10 * v1.y - G20 * v2.y) * g00;
    } else {
      v0 = make_float2(L00 * b0.x + L10 * b1.x + L20 * b2.x, L00 * b0.y + L10 * b1.y + L20 * b2.y);
      v1 = make_float2(L10 * b0.x + L11 * b1.x + L21 * b2.x, L10 * b0.y + L11 * b1.y + L21 * b2.y);
      v2 = make_float2(L20 * b0.x + L21 * b1.x + L22 * b2.x, L20 * b0.y + L21 * b1.y + L22 * b2.y);
    }
  }
  if (s.divfree && (kx | ky | kz)) {
    const float nsq = sx * sx + sy * sy + sz * sz;
    if (nsq > 0.f) {
      const float re = (v0.x * sx + v1.x * sy + v2.x * sz) / nsq;
      const float im = (v0.y * sx + v1.y * sy + v2.y * sz) / nsq;
      v0.x -= re * sx; v1.x -= re * sy; v2.x -= re * sz;
      v0.y -= im * sx; v1.y -= im * sy; v2.y -= im * sz;
    }
  }
  b0 = cscale(v0, s.scale);
  b1 = cscale(v1, s.scale);
  b2 = cscale(v2, s.scale);
}

// The three components' z-lines stay in registers through the solve; the
// occupancy floor (3 blocks of 256 at N = 256, 4 below) makes the compiler fit
// them in 168 / 128 registers without spilling, one more resident wave per SIMD
// than it picks unasked (256^3 greedy: 111 -> 96 us).
template <int N, int RA, int RB, int TX>
__global__ void __launch_bounds__(TX *(N / RA > N / RB ? N / RA : N / RB), N >= 256 ? 3 : 4)
k_zsolve_reg(const float2 *__restrict__ tw, SolveArgs sa, int ny, int nxh, int nxh_pad,
             float2 *__restrict__ spec, size_t plane, Batch bt) {
  constexpr int PA = N / RA, PB = N / RB;  // PB = RA: threads holding the spectrum
  __shared__ float2 sm[N * TX];
  spec += (long long)blockIdx.y * bt.spec;
  const int t = threadIdx.x % TX, j = threadIdx.x / TX;
  const int ntile = (nxh + TX - 1) / TX;
  const int lb = (int)xcd_block();
  if (lb >= ny * ntile) return;
  const int tile = lb % ntile, ky = lb / ntile;
  const int kx = tile * TX + t;
  const bool valid = kx < nxh;
  const size_t zst = (size_t)ny * nxh_pad;
  float2 *base = spec + (size_t)ky * nxh_pad + kx;
  float2 X[3][RB];
  // forward: radix RA then RB; the spectrum at kz = j + r * RA lands in X[c][r]
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (j < PA) {
      float2 v[RA];
#pragma unroll
      for (int r = 0; r < RA; ++r)
        v[r] = valid ? base[c * plane + (size_t)(j + r * PA) * zst] : make_float2(0.f, 0.f);
      dft<RA>(v, false);
#pragma unroll
      for (int r = 0; r < RA; ++r) sm[(j * RA + r) * TX + t] = v[r];
    }
    __syncthreads();
    if (j < PB) {
#pragma unroll
      for (int r = 0; r < RB; ++r) X[c][r] = sm[(j + r * PB) * TX + t];
#pragma unroll
      for (int r = 1; r < RB; ++r) X[c][r] = cmul(X[c][r], tw[r * j]);
      dft<RB>(X[c], false);
    }
    __syncthreads();
  }
  if (j < PB && valid) {
#pragma unroll
    for (int r = 0; r < RB; ++r) solve_one(sa, kx, ky, j + r * RA, ny, nxh, X[0][r], X[1][r], X[2][r]);
  }
  // inverse: radix RB (inputs j + r * RA = the registers) then RA; outputs z = j + r * RB
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (j < PB) {
      dft<RB>(X[c], true);
#pragma unroll
      for (int r = 0; r < RB; ++r) sm[(j * RB + r) * TX + t] = X[c][r];
    }
    __syncthreads();
    if (j < PA) {
      float2 v[RA];
#pragma unroll
      for (int r = 0; r < RA; ++r) v[r] = sm[(j + r * PA) * TX + t];
#pragma unroll
      for (int r = 1; r < RA; ++r) v[r] = cmulc(v[r], tw[r * j]);
      dft<RA>(v, true);
      if (valid) {
#pragma unroll
        for (int r = 0; r < RA; ++r) base[c * plane + (size_t)(j + r * PA) * zst] = v[r];
      }
    }
    __syncthreads();
  }
}

// Z pass: forward FFT along z of 3 components x TX kx columns at one ky, the
// per-frequency solve, inverse FFT along z.
__global__ void __launch_bounds__(1024)
k_zsolve(Axis az, SolveArgs s, int ny, int nz, int nxh, int nxh_pad, int TX,
         float2 *__restrict__ spec, size_t plane, Batch bt) {
  extern __shared__ float2 sm[];
  spec += (long long)blockIdx.y * bt.spec;
  const int NL = 3 * TX;
  const int ntile = (nxh + TX - 1) / TX;
  const int lb = (int)xcd_block();
  if (lb >= ny * ntile) return;
  const int tile = lb % ntile, ky = lb / ntile;
  const int kx0 = tile * TX;
  for (int w = threadIdx.x; w < NL * nz; w += blockDim.x) {
    const int t = w % TX, rest = w / TX;
    const int c = rest % 3, z = rest / 3;
    sm[z * NL + c * TX + t] = (kx0 + t < nxh)
                                  ? spec[c * plane + ((size_t)z * ny + ky) * nxh_pad + kx0 + t]
                                  : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft(sm, az, NL, NL, false, sm + (size_t)nz * NL);
  for (int w = threadIdx.x; w < TX * nz; w += blockDim.x) {
    const int t = w % TX, kz = w / TX;
    const int kx = kx0 + t;
    if (kx >= nxh) continue;
    float2 b0 = sm[kz * NL + t], b1 = sm[kz * NL + TX + t], b2 = sm[kz * NL + 2 * TX + t];
    solve_one(s, kx, ky, kz, ny, nxh, b0, b1, b2);
    sm[kz * NL + t] = b0;
    sm[kz * NL + TX + t] = b1;
    sm[kz * NL + 2 * TX + t] = b2;
  }
  __syncthreads();
  lds_fft(sm, az, NL, NL, true, sm + (size_t)nz * NL);
  for (int w = threadIdx.x; w < NL * nz; w += blockDim.x) {
    const int t = w % TX, rest = w / TX;
    const int c = rest % 3, z = rest / 3;
    if (kx0 + t < nxh) spec[c * plane + ((size_t)z * ny + ky) * nxh_pad + kx0 + t] = sm[z * NL + c * TX + t];
  }
}

// FP64 table of L^p (symmetric, 6 entries) per frequency for the eigen path
// (DiffOper::InitializeEigenLookup + PowL, DiffOper.cxx:641-745, 1176-1218).
// Jacobi rotations on the 3x3 symmetric matrix in double.
__global__ void k_eigen_table(float *__restrict__ out, int nx, int ny, int nz, int nxh, double spx,
                              double spy, double spz, double alpha, double beta, double gamma,
                              double p) {
  const size_t total = (size_t)nxh * ny * nz;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(i % nxh);
    const int ky = (int)((i / nxh) % ny);
    const int kz = (int)(i / ((size_t)nxh * ny));
    const double PI2 = 6.283185307179586476925286766559;
    const double cx = (2.0 * cos(PI2 * kx / nx) - 2.0) / (spx * spx);
    const double cy = (2.0 * cos(PI2 * ky / ny) - 2.0) / (spy * spy);
    const double cz = (2.0 * cos(PI2 * kz / nz) - 2.0) / (spz * spz);
    const double sx = sin(PI2 * kx / nx) / spx, sy = sin(PI2 * ky / ny) / spy,
                 sz = sin(PI2 * kz / nz) / spz;
    const double lam = -alpha * (cx + cy + cz) + gamma;
    double A[3][3] = {{lam - beta * cx, beta * sx * sy, beta * sx * sz},
                      {beta * sx * sy, lam - beta * cy, beta * sy * sz},
                      {beta * sx * sz, beta * sy * sz, lam - beta * cz}};
    double Q[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 32; ++sweep) {
      const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
      const double dg = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
      if (off <= 1e-34 * dg) break;
      for (int pp = 0; pp < 2; ++pp)
        for (int q = pp + 1; q < 3; ++q) {
          if (A[pp][q] == 0.0) continue;
          const double th = (A[q][q] - A[pp][pp]) / (2.0 * A[pp][q]);
          const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
          const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
          for (int k = 0; k < 3; ++k) {
            const double akp = A[k][pp], akq = A[k][q];
            A[k][pp] = c * akp - sn * akq;
            A[k][q] = sn * akp + c * akq;
          }
          for (int k = 0; k < 3; ++k) {
            const double apk = A[pp][k], aqk = A[q][k];
            A[pp][k] = c * apk - sn * aqk;
            A[q][k] = sn * apk + c * aqk;
          }
          for (int k = 0; k < 3; ++k) {
            const double qkp = Q[k][pp], qkq = Q[k][q];
            Q[k][pp] = c * qkp - sn * qkq;
            Q[k][q] = sn * qkp + c * qkq;
          }
        }
    }
    double e[3];
    for (int j = 0; j < 3; ++j) e[j] = pow(A[j][j], p);
    double M[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        M[r][c] = Q[r][0] * Q[c][0] * e[0] + Q[r][1] * Q[c][1] * e[1] + Q[r][2] * Q[c][2] * e[2];
    float *o = out + i * 6;
    o[0] = (float)M[0][0];
    o[1] = (float)M[1][0];
    o[2] = (float)M[1][1];
    o[3] = (float)M[2][0];
    o[4] = (float)M[2][1];
    o[5] = (float)M[2][2];
  }
}

// ------------------------------------------------------------------ host

// LDS of the x passes: the row buffer (LS = NL + 1) plus the Bluestein work area.
static size_t x_lds(const AxisPlan &a, int rt) {
  return ((size_t)a.n * (3 * rt / 2 + 1) + (size_t)a.m * (3 * rt / 2)) * sizeof(float2);
}
// LDS of the y / z passes: n lines-per-column * tx lines, plus the work area.
static size_t yz_lds(const AxisPlan &a, int lines) {
  return ((size_t)a.n + (size_t)a.m) * lines * sizeof(float2);
}

// Dynamic LDS a pass block may take: gfx950's 160 KiB per workgroup, less 1 KiB
// for the kernels' static reduction arrays (the generic pass kernels are set up
// for it once; above 64 KiB a block has the CU's LDS to itself)
constexpr size_t kPassLdsMax = 160 * 1024 - 1024;

static int pick_rt(const AxisPlan &a) {
  // the tuned RT if it fits, else the largest even RT <= 8 whose LDS fits 64 KiB
  const int want = tuning().x_rt;
  if (want > 0 && x_lds(a, want) <= 65536) return want;
  for (int rt = 8; rt >= 2; rt -= 2)  // within 64 KiB where possible (several blocks per CU)
    if (x_lds(a, rt) <= 65536) return rt;
  for (int rt = 8; rt >= 2; rt -= 2)
    if (x_lds(a, rt) <= kPassLdsMax) return rt;
  return 0;
}

static int pick_tx(const AxisPlan &a, int lines_per_col, int max_tx, int want) {
  if (want > 0 && want <= max_tx && yz_lds(a, lines_per_col * want) <= 65536) return want;
  for (int tx = max_tx; tx >= 1; tx /= 2)
    if (yz_lds(a, lines_per_col * tx) <= 65536) return tx;
  for (int tx = max_tx; tx >= 1; tx /= 2)
    if (yz_lds(a, lines_per_col * tx) <= kPassLdsMax) return tx;
  return 0;
}

static int pick_bd(int n, int nl) {
  // n * NL <= 16 * bd, bd multiple of 64, 64..1024
  int bd = (int)(((long long)n * nl + kMaxElemsPerThread - 1) / kMaxElemsPerThread);
  bd = (bd + 63) / 64 * 64;
  if (bd < 256) bd = 256;
  return bd;
}

awk_status fk_init(FluidKernel &k, int nx, int ny, int nz, const float sp[3]) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(AWK_ERR_INVALID, "fluid kernel: sizes must be >= 1");
  if ((long long)nx * ny * nz > kBufMaxElems) return fail(AWK_ERR_INVALID, "fluid kernel: volume too large");
  k.nx = nx; k.ny = ny; k.nz = nz;
  k.nxh = nx / 2 + 1;
  k.nxh_pad = (k.nxh + 7) / 8 * 8;
  for (int i = 0; i < 3; ++i) k.sp[i] = sp[i];
  const int ns[3] = {nx, ny, nz};
  for (int i = 0; i < 3; ++i) {
    k.ax[i].n = ns[i];
    k.ax[i].rad = factor(ns[i]);
    k.ax[i].m = 0;
    k.ax[i].mrad.clear();
    bool big = false;
    for (int r : k.ax[i].rad) big = big || r > 64;
    if (big) {
      // Bluestein: a power-of-two convolution of length m >= 2n - 1
      int m = 1;
      while (m < 2 * ns[i] - 1) m *= 2;
      k.ax[i].m = m;
      k.ax[i].mrad = factor(m);
    }
    if (tuning().fft_order) {  // the stages in the reverse order (a tuned plan)
      std::reverse(k.ax[i].rad.begin(), k.ax[i].rad.end());
      std::reverse(k.ax[i].mrad.begin(), k.ax[i].mrad.end());
    }
    if ((int)k.ax[i].rad.size() > kMaxStages || (int)k.ax[i].mrad.size() > 8)
      return fail(AWK_ERR_INVALID, "fluid kernel: too many FFT stages");
  }
  k.rt = pick_rt(k.ax[0]);
  k.tx_y = pick_tx(k.ax[1], 1, 16, tuning().y_tx);
  k.tx_z = pick_tx(k.ax[2], 3, 8, tuning().z_tx);
  if (!k.rt || !k.tx_y || !k.tx_z) return fail(AWK_ERR_UNSUPPORTED, "fluid kernel: axis too long for LDS staging");
  k.bd_x = pick_bd(k.ax[0].stage_len(), 3 * k.rt / 2);
  k.bd_y = pick_bd(k.ax[1].stage_len(), k.tx_y);
  k.bd_z = pick_bd(k.ax[2].stage_len(), 3 * k.tx_z);
  if (k.bd_x > 1024 || k.bd_y > 1024 || k.bd_z > 1024)
    return fail(AWK_ERR_UNSUPPORTED, "fluid kernel: block too large");
  {
    // the generic pass kernels may take more than the default 64 KiB of dynamic LDS
    static bool lds_attr = false;
    if (!lds_attr) {
      const void *fns[] = {reinterpret_cast<const void *>(&k_xfwd<0, 0, 0>),
                           reinterpret_cast<const void *>(&k_xinv<0, 0, 0, 1>),
                           reinterpret_cast<const void *>(&k_xinv<0, 0, 0, 1, false, true>),
                           reinterpret_cast<const void *>(&k_ypass),
                           reinterpret_cast<const void *>(&k_zsolve)};
      for (const void *f : fns)
        AWK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPassLdsMax));
      lds_attr = true;
    }
  }
  // twiddles and LUT (double -> float); the Bluestein tables follow the three
  // axes' twiddles in the same allocation
  std::vector<float2> tw((size_t)nx + ny + nz);
  size_t off = 0;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < ns[i]; ++j) {
      const double a = -2.0 * M_PI * (double)j / (double)ns[i];
      tw[off + j] = make_float2((float)cos(a), (float)sin(a));
    }
    off += ns[i];
  }
  size_t boff[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < 3; ++i)
    if (k.ax[i].m) bluestein_tables(ns[i], k.ax[i].m, tw, boff[i]);
  const size_t ntw = tw.size();
  const size_t nlut = 2 * (size_t)k.nxh + 2 * (size_t)ny + 2 * (size_t)nz;
  std::vector<float> lut(nlut);
  {
    // DiffOper.cxx:602-625: cosW = (2cos(2 pi k/n) - 2)/sp^2, sinW = sin(2 pi k/n)/sp
    float *p = lut.data();
    const int lens[3] = {k.nxh, ny, nz};
    for (int i = 0; i < 3; ++i) {
      const double s = 2.0 * M_PI / ns[i];
      for (int j = 0; j < lens[i]; ++j) {
        p[j] = (float)((2.0 * cos(s * j) - 2.0) / ((double)sp[i] * sp[i]));
        p[lens[i] + j] = (float)(sin(s * j) / (double)sp[i]);
      }
      p += 2 * lens[i];
    }
  }
  const size_t spec_bytes = 3 * k.spec_plane() * sizeof(float2);
  AWK_HIP(hipMalloc(&k.spec, spec_bytes));
  float2 *twd = nullptr;
  AWK_HIP(hipMalloc(&twd, ntw * sizeof(float2)));
  AWK_HIP(hipMemcpy(twd, tw.data(), ntw * sizeof(float2), hipMemcpyHostToDevice));
  k.ax[0].tw = twd;
  k.ax[1].tw = twd + nx;
  k.ax[2].tw = twd + nx + ny;
  for (int i = 0; i < 3; ++i)
    if (k.ax[i].m) {
      k.ax[i].chirp = twd + boff[i][0];
      k.ax[i].bhat = twd + boff[i][1];
      k.ax[i].mtw = twd + boff[i][2];
    }
  AWK_HIP(hipMalloc(&k.lut, nlut * sizeof(float)));
  AWK_HIP(hipMemcpy(k.lut, lut.data(), nlut * sizeof(float), hipMemcpyHostToDevice));
  k.scratch_bytes = spec_bytes + ntw * sizeof(float2) + nlut * sizeof(float);
  return AWK_OK;
}

void fk_release(FluidKernel &k) {
  if (k.spec) (void)hipFree(k.spec);
  if (k.ax[0].tw) (void)hipFree(k.ax[0].tw);
  if (k.lut) (void)hipFree(k.lut);
  if (k.eig) (void)hipFree(k.eig);
  k.spec = nullptr;
  k.ax[0].tw = k.ax[1].tw = k.ax[2].tw = nullptr;
  k.lut = nullptr;
  k.eig = nullptr;
}

awk_status fk_prepare_eigen(FluidKernel &k, int dir, hipStream_t s) {
  if (k.eig && k.eig_dir == dir) return AWK_OK;
  const size_t nf = (size_t)k.nxh * k.ny * k.nz;
  if (!k.eig) {
    AWK_HIP(hipMalloc(&k.eig, nf * 6 * sizeof(float)));
    k.scratch_bytes += nf * 6 * sizeof(float);
  }
  const double p = dir > 0 ? (double)k.lpow : -(double)k.lpow;
  unsigned grid = (unsigned)std::min<size_t>((nf + 255) / 256, 65536);
  AWK_LAUNCH(k_eigen_table, dim3(grid), dim3(256), 0, s, k.eig, k.nx, k.ny, k.nz, k.nxh,
                     (double)k.sp[0], (double)k.sp[1], (double)k.sp[2], (double)k.alpha,
                     (double)k.beta, (double)k.gamma, p);
  AWK_LAUNCH_CHECK("k_eigen_table");
  k.eig_dir = dir;
  return AWK_OK;
}

int fk_xblocks(const FluidKernel &k) {
  return (int)grid_for((long long)k.ny * k.nz, k.rt);
}

awk_status fk_xfwd(FluidKernel &k, const XFwdArgs &a_in, hipStream_t s, int *nblocks, const Batch *b) {
  const Batch bt = b ? *b : Batch();
  const unsigned grid = grid_for((long long)k.ny * k.nz, k.rt);
  XFwdArgs a = a_in;
  a.band = 0;
  {
    const int bd = tuning().xfwd_band;
    const long long G = k.ny / k.rt, per = grid / kNumXcd;
    if (bd > 0 && k.ny % k.rt == 0 && G * k.nz == (long long)grid && per % G == 0 && G % bd == 0)
      a.band = bd;
  }
  const size_t lds = x_lds(k.ax[0], k.rt);
  const Axis ax = k.ax[0].dev();
#define AWK_XF(NC, RTC, BDC)                                                                     \
  AWK_LAUNCH((k_xfwd<NC, RTC, BDC>), dim3(grid, bt.n), dim3(BDC ? BDC : k.bd_x), lds, s, a, ax, \
                     k.nx, k.ny, k.nz, k.nxh, k.nxh_pad, k.rt, k.spec, k.spec_plane(), bt)
  // block size = row length selects column mode (128^3 LDMM: 17.2 -> 16.3 us
  // against the generic 256-thread form; the spectrum is the same bit for bit)
  if (k.nx == 256 && k.rt == 8) AWK_XF(256, 8, 256);
  else if (k.nx == 128 && k.rt == 8) AWK_XF(128, 8, 128);
  else if (k.nx == 64 && k.rt == 8) AWK_XF(64, 8, 256);
  else AWK_XF(0, 0, 0);
#undef AWK_XF
  AWK_LAUNCH_CHECK("k_xfwd");
  if (nblocks) *nblocks = (int)grid;
  return AWK_OK;
}

awk_status fk_ypass(FluidKernel &k, bool inv, hipStream_t s, const ReduceArgs *red, const Batch *b) {
  const Batch bt = b ? *b : Batch();
  ReduceArgs ra;
  std::memset(&ra, 0, sizeof(ra));
  if (red) ra = *red;
  const float2 *tw = k.ax[1].tw;
  const int ntile16 = (k.nxh + 15) / 16;
  const unsigned g16 = grid_for((long long)3 * k.nz * ntile16, 1);
#define AWK_YR(N, RA, RB)                                                                     \
  AWK_LAUNCH((k_ypass_reg<N, RA, RB, 16>), dim3(g16, bt.n), dim3(16 * (N / RA > N / RB ? N / RA : N / RB)), \
                     0, s, tw, k.nz, k.nxh, k.nxh_pad, k.spec, k.spec_plane(),                    \
                     inv ? 1 : tuning().yfwd_rev ? 0 : 2, ra, bt)
  if (k.ny == 256) AWK_YR(256, 16, 16);
  else if (k.ny == 128) AWK_YR(128, 16, 8);
  else if (k.ny == 64) AWK_YR(64, 8, 8);
  else if (k.ny == 32) AWK_YR(32, 8, 4);
  else {
    const int ntile = (k.nxh + k.tx_y - 1) / k.tx_y;
    const unsigned grid = grid_for((long long)3 * k.nz * ntile, 1);
    const size_t lds = std::max(yz_lds(k.ax[1], k.tx_y), (size_t)k.bd_y * sizeof(double));
    AWK_LAUNCH(k_ypass, dim3(grid, bt.n), dim3(k.bd_y), lds, s, k.ax[1].dev(), k.ny, k.nz, k.nxh,
                       k.nxh_pad, k.tx_y, k.spec, k.spec_plane(), inv ? 1 : 0, ra, bt);
  }
#undef AWK_YR
  AWK_LAUNCH_CHECK("k_ypass");
  return AWK_OK;
}

awk_status fk_zsolve(FluidKernel &k, bool inverse_op, hipStream_t s, const Batch *b) {
  const Batch bt = b ? *b : Batch();
  const bool eigen = (k.flags & AWK_FK_EIGEN) || k.lpow != 1.0f;
  if (eigen) {
    awk_status st = fk_prepare_eigen(k, inverse_op ? -1 : 1, s);
    if (st) return st;
  }
  SolveArgs sa;
  sa.cwx = k.lut;
  sa.swx = k.lut + k.nxh;
  sa.cwy = k.lut + 2 * k.nxh;
  sa.swy = sa.cwy + k.ny;
  sa.cwz = sa.swy + k.ny;
  sa.swz = sa.cwz + k.nz;
  sa.eig = eigen ? k.eig : nullptr;
  sa.alpha = k.alpha;
  sa.beta = k.beta;
  sa.gamma = k.gamma;
  sa.scale = (float)(1.0 / ((double)k.nx * k.ny * k.nz));
  sa.inverse_op = inverse_op ? 1 : 0;
  sa.divfree = (k.flags & AWK_FK_DIVFREE) ? 1 : 0;
  const float2 *tw = k.ax[2].tw;
  const int ntile16 = (k.nxh + 15) / 16;
  const unsigned g16 = grid_for((long long)k.ny * ntile16, 1);
#define AWK_ZR(N, RA, RB)                                                                     \
  AWK_LAUNCH((k_zsolve_reg<N, RA, RB, 16>), dim3(g16, bt.n), dim3(16 * (N / RA > N / RB ? N / RA : N / RB)), \
                     0, s, tw, sa, k.ny, k.nxh, k.nxh_pad, k.spec, k.spec_plane(), bt)
  if (k.nz == 256) AWK_ZR(256, 16, 16);
  else if (k.nz == 128) AWK_ZR(128, 16, 8);
  else if (k.nz == 64) AWK_ZR(64, 8, 8);
  else if (k.nz == 32) AWK_ZR(32, 8, 4);
  else {
    const int ntile = (k.nxh + k.tx_z - 1) / k.tx_z;
    const unsigned grid = grid_for((long long)k.ny * ntile, 1);
    const size_t lds = yz_lds(k.ax[2], 3 * k.tx_z);
    AWK_LAUNCH(k_zsolve, dim3(grid, bt.n), dim3(k.bd_z), lds, s, k.ax[2].dev(), sa, k.ny, k.nz,
                       k.nxh, k.nxh_pad, k.tx_z, k.spec, k.spec_plane(), bt);
  }
#undef AWK_ZR
  AWK_LAUNCH_CHECK("k_zsolve");
  return AWK_OK;
}

awk_status fk_xinv(FluidKernel &k, const XInvArgs &a_in, hipStream_t s, int *nblocks, const Batch *b) {
  const Batch bt = b ? *b : Batch();
  const unsigned grid = grid_for((long long)k.ny * k.nz, k.rt);
  XInvArgs a = a_in;
  a.band = 0;
  {
    // the band block order of the composition, where whole planes of row groups
    // fill each XCD's share of the grid and the band divides a plane
    const int bd = tuning().xinv_band;
    const long long G = k.ny / k.rt, per = grid / kNumXcd;
    if (a.mode == XSTORE_COMPOSE && bd > 0 && k.ny % k.rt == 0 && G * k.nz == (long long)grid &&
        per % G == 0 && G % bd == 0)
      a.band = bd;
  }
  size_t lds = x_lds(k.ax[0], k.rt);
  const Axis ax = k.ax[0].dev();
#define AWK_XI3(NC, RTC, BDC, XU, PR, AP)                                                        \
  AWK_LAUNCH((k_xinv<NC, RTC, BDC, XU, PR, AP>), dim3(grid, bt.n), dim3(BDC ? BDC : k.bd_x), lds, s, a, \
                     ax, k.nx, k.ny, k.nz, k.nxh, k.nxh_pad, k.rt, k.spec, k.spec_plane(), bt)
#define AWK_XIS(NC, BD, AP)                                                                      \
  AWK_LAUNCH((k_xinv<NC, 8, BD, 1, false, AP, true>), dim3(grid, bt.n), dim3(BD), lds, s, a, ax, \
                     k.nx, k.ny, k.nz, k.nxh, k.nxh_pad, k.rt, k.spec, k.spec_plane(), bt)
#define AWK_XI(NC, RTC, BDC, XU)                                                                 \
  AWK_LAUNCH((k_xinv<NC, RTC, BDC, XU>), dim3(grid, bt.n), dim3(BDC ? BDC : k.bd_x), lds, s, a, ax, \
                     k.nx, k.ny, k.nz, k.nxh, k.nxh_pad, k.rt, k.spec, k.spec_plane(), bt)
  const int xu = tuning().xinv_unroll;
  // xinv_pairs: waves whose points all lie inside the volume gather the
  // x-neighbour corners as 8-byte pairs (12 loads per point instead of 24; the
  // same values, so the same result)
  const bool pr = a.mode == XSTORE_COMPOSE && tuning().xinv_pairs;
  // the composition with the next iteration's deformed image (one subject per launch)
  const bool ap = a.mode == XSTORE_COMPOSE && a.i0;
  if (ap && (bt.n != 1 || !a.idef))
    return fail(AWK_ERR_INVALID, "inverse x pass: the fused apply takes one subject and an output image");
  // xinv_stage: the composition reads the field from the block's rows and their
  // neighbours staged in LDS (specialised row lengths, whole blocks in one plane).
  // Measured on the greedy step: 128^3 atlas iteration 1.52 -> 1.35 ms, but at
  // 256-long rows 0.26 -> 0.36 ms per composition (three staged rows per row read
  // through L2, with 4 blocks per CU; 0.30 ms with 512 threads of 4 points each,
  // whose FFT rounds differently), so by default rows up to 128 only.
  const int sk = tuning().xinv_stage;
  const bool stg = a.mode == XSTORE_COMPOSE && (sk == 1 || (sk == 2 && k.nx <= 128)) && k.rt == 8 &&
                   k.ny % 8 == 0 && (k.nx == 256 || k.nx == 128 || k.nx == 64);
  if (stg) {
    lds = std::max(lds, (size_t)3 * (k.rt + 2) * k.nx * sizeof(float));
    if (k.nx == 256) {
      if (ap) AWK_XIS(256, 256, true);
      else AWK_XIS(256, 256, false);
    } else if (k.nx == 128) {
      if (ap) AWK_XIS(128, 256, true);
      else AWK_XIS(128, 256, false);
    } else if (ap) AWK_XIS(64, 256, true);
    else AWK_XIS(64, 256, false);
  } else if (ap) {
    if (k.nx == 256 && k.rt == 8) {
      if (pr) {
        if (xu == 4) AWK_XI3(256, 8, 256, 4, true, true);
        else if (xu == 2) AWK_XI3(256, 8, 256, 2, true, true);
        else AWK_XI3(256, 8, 256, 1, true, true);
      } else if (xu == 4) AWK_XI3(256, 8, 256, 4, false, true);
      else if (xu == 2) AWK_XI3(256, 8, 256, 2, false, true);
      else AWK_XI3(256, 8, 256, 1, false, true);
    } else if (k.nx == 128 && k.rt == 8) AWK_XI3(128, 8, 256, 2, false, true);
    else if (k.nx == 64 && k.rt == 8) AWK_XI3(64, 8, 256, 2, false, true);
    else AWK_XI3(0, 0, 0, 1, false, true);
  } else if (k.nx == 256 && k.rt == 8) {
    if (pr) {
      if (xu == 4) AWK_XI3(256, 8, 256, 4, true, false);
      else if (xu == 2) AWK_XI3(256, 8, 256, 2, true, false);
      else AWK_XI3(256, 8, 256, 1, true, false);
    } else if (xu == 4) AWK_XI(256, 8, 256, 4);
    else if (xu == 2) AWK_XI(256, 8, 256, 2);
    else AWK_XI(256, 8, 256, 1);
  } else if (k.nx == 128 && k.rt == 8) AWK_XI(128, 8, 256, 2);
  else if (k.nx == 64 && k.rt == 8) AWK_XI(64, 8, 256, 2);
  else AWK_XI(0, 0, 0, 1);
#undef AWK_XI
#undef AWK_XI3
#undef AWK_XIS
  AWK_LAUNCH_CHECK("k_xinv");
  if (nblocks) *nblocks = (int)grid;
  return AWK_OK;
}

}  // namespace awk

using namespace awk;

extern "C" {

awk_status awk_fluid_kernel_create(awk_fk_t *out, int nx, int ny, int nz, float spx, float spy,
                                   float spz, float alpha, float beta, float gamma, float lpow,
                                   int flags) {
  if (!out) return fail(AWK_ERR_INVALID, "awk_fluid_kernel_create: null handle pointer");
  *out = nullptr;
  if (!(spx > 0 && spy > 0 && spz > 0)) return fail(AWK_ERR_INVALID, "awk_fluid_kernel_create: spacing must be > 0");
  awk_fluid_kernel *h = new (std::nothrow) awk_fluid_kernel();
  if (!h) return fail(AWK_ERR_ALLOC, "awk_fluid_kernel_create: host allocation failed");
  const float sp[3] = {spx, spy, spz};
  awk_status st = fk_init(h->k, nx, ny, nz, sp);
  if (st) {
    fk_release(h->k);
    delete h;
    return st;
  }
  h->k.alpha = alpha;
  h->k.beta = beta;
  h->k.gamma = gamma;
  h->k.lpow = lpow;
  h->k.flags = flags;
  *out = h;
  return AWK_OK;
}

awk_status awk_fluid_kernel_set_params(awk_fk_t h, float alpha, float beta, float gamma, float lpow,
                                       int flags) {
  if (!h) return fail(AWK_ERR_INVALID, "awk_fluid_kernel_set_params: null handle");
  FluidKernel &k = h->k;
  if (alpha != k.alpha || beta != k.beta || gamma != k.gamma || lpow != k.lpow) k.eig_dir = 0;
  k.alpha = alpha;
  k.beta = beta;
  k.gamma = gamma;
  k.lpow = lpow;
  k.flags = flags;
  return AWK_OK;
}

awk_status awk_fluid_kernel_destroy(awk_fk_t h) {
  if (!h) return AWK_OK;
  fk_release(h->k);
  delete h;
  return AWK_OK;
}

awk_status awk_fluid_kernel_scratch_bytes(awk_fk_t h, unsigned long long *bytes) {
  if (!h || !bytes) return fail(AWK_ERR_INVALID, "awk_fluid_kernel_scratch_bytes: null argument");
  *bytes = h->k.scratch_bytes;
  return AWK_OK;
}

awk_status awk_fluid_kernel_apply(awk_fk_t h, float *x, float *y, float *z, int inverse,
                                  awk_stream_t stream) {
  if (!h || !x || !y || !z) return fail(AWK_ERR_INVALID, "awk_fluid_kernel_apply: null argument");
  FluidKernel &k = h->k;
  hipStream_t s = as_stream(stream);
  XFwdArgs fa;
  std::memset(&fa, 0, sizeof(fa));
  fa.mode = XLOAD_PLAIN;
  fa.in[0] = x;
  fa.in[1] = y;
  fa.in[2] = z;
  awk_status st;
  if ((st = fk_xfwd(k, fa, s, nullptr))) return st;
  if ((st = fk_ypass(k, false, s, nullptr))) return st;
  if ((st = fk_zsolve(k, inverse != 0, s))) return st;
  if ((st = fk_ypass(k, true, s, nullptr))) return st;
  XInvArgs ia;
  std::memset(&ia, 0, sizeof(ia));
  ia.mode = XSTORE_PLAIN;
  ia.out[0] = x;
  ia.out[1] = y;
  ia.out[2] = z;
  return fk_xinv(k, ia, s, nullptr);
}

}  // extern "C"
