// awk_fft.h -- LDS-staged mixed-radix FFT engine for gfx950 (device code).
//
// A batch of NL complex lines of length n lives in LDS in the interleaved layout
// buf[i * LS + line] (LS >= NL) (element i of every line adjacent).  Stockham autosort
// stages run in place: every thread first reads all of its butterflies of the
// stage into registers (at most 16 complex values; the host sizes the block so
// that n * NL <= 16 * blockDim), the block synchronises, then writes them back.
// With lanes mapped to consecutive lines, every LDS read and write of every
// stage touches consecutive addresses: no bank conflicts.
//
// Radices 2, 3, 4, 5, 8 and 16 have register butterflies; any other prime
// factor up to 64 uses a table-driven direct DFT (one thread per output).  Twiddles
// come from a per-axis table tw[k] = exp(-2 pi i k / n) computed in double on the
// host.  An axis with a larger prime factor is transformed by Bluestein's chirp-z
// method (L. Bluestein, "A linear filtering approach to the computation of the
// discrete Fourier transform", 1970): with c_j = exp(-i pi j^2 / n),
//   X_k = c_k sum_j (x_j c_j) conj(c_{k-j}),
// a cyclic convolution of length m (a power of two >= 2n - 1) done as two
// power-of-two FFTs in an LDS work area and a product with the precomputed
// spectrum of conj(c) (scaled by 1/m); the inverse transform is
// conj(DFT(conj x)).
#pragma once

#include "awk_common.h"

namespace awk {

constexpr int kMaxStages = 24;
constexpr int kMaxElemsPerThread = 16;

struct Axis {
  const float2 *tw;  // n twiddles (device)
  int n;
  int nst;
  int rad[kMaxStages];
  // Bluestein (m > 0): chirp c_j (n), the scaled spectrum of conj(c) (m), and
  // the power-of-two sub-transform of length m (twiddles, radices)
  int m;
  const float2 *chirp, *bhat, *mtw;
  int mst;
  int mrad[8];
};

// multiply by -i (forward) or +i (inverse)
__device__ __forceinline__ float2 rot(float2 z, bool inv) {
  return inv ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

template <int R>
__device__ __forceinline__ void dft(float2 *v, bool inv);

template <>
__device__ __forceinline__ void dft<2>(float2 *v, bool) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<3>(float2 *v, bool inv) {
  const float c1 = -0.5f, s1 = 0.86602540378443865f;  // cos, sin(2 pi / 3)
  float2 a = cadd(v[1], v[2]);
  float2 b = cscale(csub(v[1], v[2]), s1);
  float2 m = make_float2(v[0].x + c1 * a.x, v[0].y + c1 * a.y);
  float2 rb = rot(b, inv);  // -i*b (forward)
  v[0] = cadd(v[0], a);
  v[1] = cadd(m, rb);
  v[2] = csub(m, rb);
}

template <>
__device__ __forceinline__ void dft<4>(float2 *v, bool inv) {
  float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  float2 s13 = cadd(v[1], v[3]), d13 = rot(csub(v[1], v[3]), inv);
  v[0] = cadd(s02, s13);
  v[1] = cadd(d02, d13);
  v[2] = csub(s02, s13);
  v[3] = csub(d02, d13);
}

template <>
__device__ __forceinline__ void dft<5>(float2 *v, bool inv) {
  const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
  const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
  float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
  float2 m2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
  float2 n1 = rot(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), inv);
  float2 n2 = rot(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), inv);
  v[0] = cadd(v[0], cadd(a1, a2));
  v[1] = cadd(m1, n1);
  v[4] = csub(m1, n1);
  v[2] = cadd(m2, n2);
  v[3] = csub(m2, n2);
}

template <>
__device__ __forceinline__ void dft<8>(float2 *v, bool inv) {
  const float h = 0.70710678118654752f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e, inv);
  dft<4>(o, inv);
  // o[k] *= W8^k (conjugated for the inverse)
  const float sg = inv ? 1.f : -1.f;
  o[1] = cmul(o[1], make_float2(h, sg * h));
  o[2] = rot(o[2], inv);
  o[3] = cmul(o[3], make_float2(-h, sg * h));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

template <>
__device__ __forceinline__ void dft<16>(float2 *v, bool inv) {
  float2 e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft<8>(e, inv);
  dft<8>(o, inv);
  // W16^k = (cos(pi k / 8), -sin(pi k / 8))
  const float cs[8] = {1.f, 0.92387953251128676f, 0.70710678118654752f, 0.38268343236508977f,
                       0.f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128676f};
  const float sn[8] = {0.f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128676f,
                       1.f, 0.92387953251128676f, 0.70710678118654752f, 0.38268343236508977f};
  const float sg = inv ? 1.f : -1.f;
#pragma unroll
  for (int k = 1; k < 8; ++k) o[k] = cmul(o[k], make_float2(cs[k], sg * sn[k]));
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 8] = csub(e[k], o[k]);
  }
}

// One in-place Stockham stage of radix R over NL interleaved lines of length n.
template <int R>
__device__ __forceinline__ void stage_reg(float2 *buf, int n, int NL, int LS, int Ns,
                                          const float2 *__restrict__ tw, bool inv) {
  constexpr int K = (kMaxElemsPerThread + R - 1) / R;
  const int m = n / R;
  const int total = m * NL;
  const int twstep = n / (Ns * R);
  const int bd = blockDim.x;
  float2 v[K][R];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int w = threadIdx.x + i * bd;
    if (w < total) {
      const int line = w % NL, j = w / NL, k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * m) * LS + line];
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const float2 t = tw[r * k * twstep];
          v[i][r] = inv ? cmulc(v[i][r], t) : cmul(v[i][r], t);
        }
      }
      dft<R>(v[i], inv);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int w = threadIdx.x + i * bd;
    if (w < total) {
      const int line = w % NL, j = w / NL, k = j % Ns;
      const int idxD = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(idxD + r * Ns) * LS + line] = v[i][r];
    }
  }
  __syncthreads();
}

// Direct DFT stage for any other prime radix R: one thread per output element.
__device__ __forceinline__ void stage_generic(float2 *buf, int n, int R, int NL, int LS, int Ns,
                                              const float2 *__restrict__ tw, bool inv) {
  const int m = n / R;
  const int total = n * NL;
  const int twstep = n / (Ns * R);
  const int bd = blockDim.x;
  float2 acc[kMaxElemsPerThread];
#pragma unroll
  for (int i = 0; i < kMaxElemsPerThread; ++i) {
    const int w = threadIdx.x + i * bd;
    acc[i] = make_float2(0.f, 0.f);
    if (w < total) {
      const int line = w % NL, rest = w / NL, q = rest % R, j = rest / R, k = j % Ns;
      float2 a = make_float2(0.f, 0.f);
      for (int r = 0; r < R; ++r) {
        const float2 x = buf[(j + r * m) * LS + line];
        const int e = (r * k * twstep + ((r * q) % R) * m) % n;
        const float2 t = tw[e];
        const float2 p = inv ? cmulc(x, t) : cmul(x, t);
        a.x += p.x;
        a.y += p.y;
      }
      acc[i] = a;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kMaxElemsPerThread; ++i) {
    const int w = threadIdx.x + i * bd;
    if (w < total) {
      const int line = w % NL, rest = w / NL, q = rest % R, j = rest / R, k = j % Ns;
      const int idxD = (j - k) * R + k;
      buf[(idxD + q * Ns) * LS + line] = acc[i];
    }
  }
  __syncthreads();
}

// The Stockham stages of one transform of length n with radices rad[0..nst).
__device__ __forceinline__ void lds_fft_stages(float2 *buf, const float2 *__restrict__ tw, int n,
                                               int nst, const int *rad, int NL, int LS, bool inv) {
  int Ns = 1;
  for (int s = 0; s < nst; ++s) {
    const int R = rad[s];
    switch (R) {
      case 2: stage_reg<2>(buf, n, NL, LS, Ns, tw, inv); break;
      case 3: stage_reg<3>(buf, n, NL, LS, Ns, tw, inv); break;
      case 4: stage_reg<4>(buf, n, NL, LS, Ns, tw, inv); break;
      case 5: stage_reg<5>(buf, n, NL, LS, Ns, tw, inv); break;
      case 8: stage_reg<8>(buf, n, NL, LS, Ns, tw, inv); break;
      case 16: stage_reg<16>(buf, n, NL, LS, Ns, tw, inv); break;
      default: stage_generic(buf, n, R, NL, LS, Ns, tw, inv); break;
    }
    Ns *= R;
  }
}

// Bluestein's transform of the NL lines of buf through the work area work
// (m * NL complex, interleaved work[i * NL + line]).
__device__ __forceinline__ void lds_fft_bluestein(float2 *buf, float2 *work, const Axis &ax,
                                                  int NL, int LS, bool inv) {
  const int n = ax.n, m = ax.m, bd = (int)blockDim.x;
  for (int w = (int)threadIdx.x; w < m * NL; w += bd) {
    const int line = w % NL, j = w / NL;
    float2 v = make_float2(0.f, 0.f);
    if (j < n) {
      const float2 x = buf[j * LS + line];
      v = cmul(inv ? cconj(x) : x, ax.chirp[j]);
    }
    work[w] = v;
  }
  __syncthreads();
  lds_fft_stages(work, ax.mtw, m, ax.mst, ax.mrad, NL, NL, false);
  for (int w = (int)threadIdx.x; w < m * NL; w += bd) work[w] = cmul(work[w], ax.bhat[w / NL]);
  __syncthreads();
  lds_fft_stages(work, ax.mtw, m, ax.mst, ax.mrad, NL, NL, true);
  for (int w = (int)threadIdx.x; w < n * NL; w += bd) {
    const int line = w % NL, k = w / NL;
    const float2 y = cmul(work[w], ax.chirp[k]);
    buf[k * LS + line] = inv ? cconj(y) : y;
  }
  __syncthreads();
}

// Full unnormalised transform of NL interleaved lines held in LDS `buf`.
// Callers must __syncthreads() after filling buf; the result is in buf and
// visible to all threads on return.  work: the Bluestein work area (m * NL
// complex) for axes that need it.
__device__ __forceinline__ void lds_fft(float2 *buf, const Axis &ax, int NL, int LS, bool inv,
                                        float2 *work = nullptr) {
  if (ax.m) lds_fft_bluestein(buf, work, ax, NL, LS, inv);
  else lds_fft_stages(buf, ax.tw, ax.n, ax.nst, ax.rad, NL, LS, inv);
}

}  // namespace awk

namespace awk {

// ---------------------------------------------------------------------------
// Compile-time variants for power-of-two lengths: the same Stockham stages with
// N, the line count NL, the line stride LS and the block size BD known to the
// compiler (index division becomes shifts / multiply-shifts, loops unroll).
constexpr int first_radix(int rem) {
  return rem % 16 == 0 ? 16 : rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : 2;
}

template <int N, int NL, int LS, int BD, int R, int NS>
__device__ __forceinline__ void ct_stage(float2 *buf, const float2 *__restrict__ tw, bool inv) {
  constexpr int M = N / R;
  constexpr int TOTAL = M * NL;
  constexpr int K = (TOTAL + BD - 1) / BD;
  constexpr int TWSTEP = N / (NS * R);
  float2 v[K][R];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int w = threadIdx.x + i * BD;
    if (TOTAL % BD == 0 || w < TOTAL) {
      const int line = w % NL, j = w / NL, k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * M) * LS + line];
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const float2 t = tw[r * k * TWSTEP];
          v[i][r] = inv ? cmulc(v[i][r], t) : cmul(v[i][r], t);
        }
      }
      dft<R>(v[i], inv);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int w = threadIdx.x + i * BD;
    if (TOTAL % BD == 0 || w < TOTAL) {
      const int line = w % NL, j = w / NL, k = j % NS;
      const int idxD = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(idxD + r * NS) * LS + line] = v[i][r];
    }
  }
  __syncthreads();
}

template <int N, int NL, int LS, int BD, int NS = 1>
__device__ __forceinline__ void ct_fft(float2 *buf, const float2 *__restrict__ tw, bool inv) {
  if constexpr (NS < N) {
    constexpr int R = first_radix(N / NS);
    ct_stage<N, NL, LS, BD, R, NS>(buf, tw, inv);
    ct_fft<N, NL, LS, BD, NS * R>(buf, tw, inv);
  }
}

}  // namespace awk
