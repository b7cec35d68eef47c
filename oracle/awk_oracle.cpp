/*
 * awk_oracle.cpp -- CPU restatement of the AtlasWerks greedy-fluid hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See awk_oracle.h.
 * Citations are relative to the AtlasWerks source root.
 *
 * The reference CPU path links FFTW3 (Code/Libraries/Algorithms/Common/DiffOper.cxx:30-99),
 * which is not available here, so the FFT below is this file's own mixed-radix
 * Stockham transform (any length; radix 4/2 butterflies plus a table-driven DFT
 * for every other prime factor).  Its twiddles are computed in double.
 */
#include "awk_oracle.h"

#include <algorithm>
#include <chrono>
#include <thread>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

static int g_mut = 0;
extern "C" void orc_set_mutation(int m) { g_mut = m; }

namespace {

/* ====================================================================== */
/* FFT                                                                    */
/* ====================================================================== */

template <class T>
struct Fft {
  typedef std::complex<T> C;
  int n;
  std::vector<int> rad;
  std::vector<C> tw;   // tw[k] = exp(-2 pi i k / n), computed in double
  std::vector<C> buf;

  explicit Fft(int n_) : n(n_), tw(n_), buf(n_) {
    int m = n;
    while (m % 4 == 0) { rad.push_back(4); m /= 4; }
    while (m % 2 == 0) { rad.push_back(2); m /= 2; }
    for (int p = 3; m > 1; p += 2) {
      while (m % p == 0) { rad.push_back(p); m /= p; }
    }
    for (int k = 0; k < n; ++k) {
      double a = -2.0 * M_PI * (double)k / (double)n;
      tw[k] = C((T)std::cos(a), (T)std::sin(a));
    }
  }

  inline C w(long idx, bool inv) const {
    C t = tw[idx % n];
    return inv ? std::conj(t) : t;
  }

  /* Stockham autosort, out-of-place ping-pong; result left in x. */
  void run(C *x, bool inv) {
    if (n == 1) return;
    C *a = x, *b = buf.data();
    int Ns = 1;
    // butterfly scratch sized to the largest radix (any prime factor, e.g. 181)
    const int rmax = *std::max_element(rad.begin(), rad.end());
    std::vector<C> vv(rmax), oo(rmax);
    C *v = vv.data(), *o = oo.data();
    for (int R : rad) {
      const int m = n / R;
      const int twstep = n / (Ns * R);
      for (int j = 0; j < m; ++j) {
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) {
          C t = a[j + r * m];
          if (k) t *= w((long)r * k * twstep, inv);
          v[r] = t;
        }
        if (R == 2) {
          o[0] = v[0] + v[1];
          o[1] = v[0] - v[1];
        } else if (R == 4) {
          C s02 = v[0] + v[2], d02 = v[0] - v[2];
          C s13 = v[1] + v[3], d13 = v[1] - v[3];
          // forward: multiply d13 by -i ; inverse by +i
          C jd13 = inv ? C(-d13.imag(), d13.real()) : C(d13.imag(), -d13.real());
          o[0] = s02 + s13;
          o[1] = d02 + jd13;
          o[2] = s02 - s13;
          o[3] = d02 - jd13;
        } else {
          const int step = n / R;
          for (int q = 0; q < R; ++q) {
            C acc(0, 0);
            for (int r = 0; r < R; ++r) acc += v[r] * w((long)((r * q) % R) * step, inv);
            o[q] = acc;
          }
        }
        const int idxD = (j / Ns) * Ns * R + k;
        for (int r = 0; r < R; ++r) b[idxD + r * Ns] = o[r];
      }
      std::swap(a, b);
      Ns *= R;
    }
    if (a != x) std::memcpy(x, a, sizeof(C) * n);
  }
};

/* 3-D real <-> half-complex transforms of one scalar plane.  Spectrum layout is
 * [z][y][kx] with kx < nx/2+1, the layout FFTW's r2c produces for rank 3,
 * n = {z, y, x} (DiffOper.cxx:38-66). */
template <class T>
struct Fft3 {
  typedef std::complex<T> C;
  int nx, ny, nz, nxh;
  Fft<T> fx, fy, fz;
  std::vector<C> line, lb;
  Fft3(int x, int y, int z)
      : nx(x), ny(y), nz(z), nxh(x / 2 + 1), fx(x), fy(y), fz(z),
        line(std::max(std::max(x, y), z)), lb(std::max(std::max(x, y), z)) {}

  size_t spec_size() const { return (size_t)nxh * ny * nz; }

  void r2c(const T *in, C *S) {
    const int nl = ny * nz;
    for (int l = 0; l < nl; l += 2) {
      const T *a = in + (size_t)l * nx;
      const T *b = (l + 1 < nl) ? in + (size_t)(l + 1) * nx : nullptr;
      for (int i = 0; i < nx; ++i) line[i] = C(a[i], b ? b[i] : (T)0);
      fx.run(line.data(), false);
      C *A = S + (size_t)l * nxh;
      C *B = (l + 1 < nl) ? S + (size_t)(l + 1) * nxh : nullptr;
      for (int k = 0; k < nxh; ++k) {
        C zk = line[k];
        C zc = std::conj(line[(nx - k) % nx]);
        A[k] = (zk + zc) * (T)0.5;
        if (B) {
          C d = (zk - zc) * (T)0.5;        // = i * B_k
          B[k] = C(d.imag(), -d.real());   // B_k = d / i
        }
      }
    }
    axis_pass(S, false);
  }

  void axis_pass(C *S, bool inv) {
    // y axis
    for (int z = 0; z < nz; ++z)
      for (int k = 0; k < nxh; ++k) {
        C *base = S + (size_t)z * ny * nxh + k;
        for (int y = 0; y < ny; ++y) line[y] = base[(size_t)y * nxh];
        fy.run(line.data(), inv);
        for (int y = 0; y < ny; ++y) base[(size_t)y * nxh] = line[y];
      }
    // z axis
    const size_t pz = (size_t)ny * nxh;
    for (int y = 0; y < ny; ++y)
      for (int k = 0; k < nxh; ++k) {
        C *base = S + (size_t)y * nxh + k;
        for (int z = 0; z < nz; ++z) line[z] = base[z * pz];
        fz.run(line.data(), inv);
        for (int z = 0; z < nz; ++z) base[z * pz] = line[z];
      }
  }

  /* Unnormalised inverse; the DC and Nyquist bins of every x-line are taken as
   * real (the imaginary part FFTW's c2r ignores). Destroys S. */
  void c2r(C *S, T *out) {
    axis_pass(S, true);
    const int nl = ny * nz;
    for (int l = 0; l < nl; l += 2) {
      C *A = S + (size_t)l * nxh;
      C *B = (l + 1 < nl) ? S + (size_t)(l + 1) * nxh : nullptr;
      for (int k = 0; k < nx; ++k) {
        C ak, bk;
        if (k < nxh) {
          ak = A[k];
          bk = B ? B[k] : C(0, 0);
        } else {
          ak = std::conj(A[nx - k]);
          bk = B ? std::conj(B[nx - k]) : C(0, 0);
        }
        if (k == 0 || 2 * k == nx) {
          ak = C(ak.real(), 0);
          bk = C(bk.real(), 0);
        }
        line[k] = ak + C(-bk.imag(), bk.real());  // ak + i*bk
      }
      fx.run(line.data(), true);
      T *a = out + (size_t)l * nx;
      for (int i = 0; i < nx; ++i) a[i] = line[i].real();
      if (B) {
        T *b = out + (size_t)(l + 1) * nx;
        for (int i = 0; i < nx; ++i) b[i] = line[i].imag();
      }
    }
  }
};

/* ====================================================================== */
/* Navier operator                                                         */
/* ====================================================================== */

/* DiffOperT::InitializeOperatorLookupTable (DiffOper.cxx:587-639): computed in
 * double, stored as T. */
template <class T>
void make_lut(int n, double sp, std::vector<T> &cw, std::vector<T> &sw) {
  cw.resize(n);
  sw.resize(n);
  const double s = 2.0 * M_PI / n;
  for (int k = 0; k < n; ++k) {
    cw[k] = (T)((2.0 * std::cos(s * (double)k) - 2.0) / (sp * sp));
    sw[k] = (T)(std::sin(s * (double)k) / sp);
  }
}

/* Jacobi eigen-decomposition of a symmetric 3x3 (double).  Stands in for
 * Matrix3D::factorEVsym / LAPACK dsyev (DataTypes/Matrix3D.txx:273-303): any
 * orthonormal eigenbasis gives the same L^p = Q diag(e^p) Q^T. */
void jacobi3(double A[3][3], double e[3], double Q[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Q[i][j] = (i == j);
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    double diag = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off <= 1e-34 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (A[p][q] == 0.0) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A = J^T A J
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double qkp = Q[k][p], qkq = Q[k][q];
          Q[k][p] = c * qkp - s * qkq;
          Q[k][q] = s * qkp + c * qkq;
        }
      }
  }
  for (int i = 0; i < 3; ++i) e[i] = A[i][i];
}

template <class T>
int diffoper(const int *size, const double *sp, double alpha, double beta, double gamma,
             double lpow, int flags, T *vf, bool inverse) {
  typedef std::complex<T> C;
  const int nx = size[0], ny = size[1], nz = size[2];
  if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
  const bool eigen = (flags & ORC_DO_EIGEN) || lpow != 1.0;
  const bool divfree = flags & ORC_DO_DIVFREE;
  const size_t N = (size_t)nx * ny * nz;
  Fft3<T> F(nx, ny, nz);
  const int nxh = F.nxh;
  std::vector<T> cwx, swx, cwy, swy, cwz, swz;
  make_lut<T>(nx, sp[0], cwx, swx);
  make_lut<T>(ny, sp[1], cwy, swy);
  make_lut<T>(nz, sp[2], cwz, swz);
  if (g_mut == 6) std::swap(cwy, swy);  // self-test: swapped LUT tables

  // scale(1/N) then forward transform of each component
  // (DiffOper.cxx:834-836: mFFTWVectorField->scale(1.0/N); ExecuteForward)
  std::vector<T> plane(N);
  std::vector<C> S[3];
  const T scale = (T)(1.0 / (double)N);
  for (int c = 0; c < 3; ++c) {
    for (size_t i = 0; i < N; ++i) plane[i] = (T)(vf[3 * i + c] * scale);
    S[c].resize(F.spec_size());
    F.r2c(plane.data(), S[c].data());
  }

  const T a = (T)alpha, b = (T)beta, g = (T)gamma;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nxh; ++x) {
        const size_t idx = ((size_t)z * ny + y) * nxh + x;
        C bx = S[0][idx], by = S[1][idx], bz = S[2][idx];
        // DiffOper.cxx:861-879 (lambda held in double as the reference does)
        double lambda = -(double)a * ((double)cwx[x] + (double)cwy[y] + (double)cwz[z]) + (double)g;
        T L00 = (T)(lambda - (double)b * cwx[x]);
        T L11 = (T)(lambda - (double)b * cwy[y]);
        T L22 = (T)(lambda - (double)b * cwz[z]);
        T L10 = b * swx[x] * swy[y];
        T L20 = b * swx[x] * swz[z];
        T L21 = b * swy[y] * swz[z];
        if (g_mut == 1) { L10 = L20 = L21 = 0; }           // dropped grad-div coupling
        if (g_mut == 2) { L10 = -L10; }                      // wrong sign
        if (g_mut == 3) { std::swap(L20, L21); }             // transposed operand
        C vx, vy, vz;
        if (eigen) {
          // InitializeEigenLookup + PowL (DiffOper.cxx:641-745, 1176-1218)
          double A[3][3] = {{(double)L00, (double)L10, (double)L20},
                            {(double)L10, (double)L11, (double)L21},
                            {(double)L20, (double)L21, (double)L22}};
          double e[3], Q[3][3];
          jacobi3(A, e, Q);
          double p = inverse ? -lpow : lpow;
          for (int i = 0; i < 3; ++i) e[i] = std::pow(e[i], p);
          double M[3][3];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              M[i][j] = Q[i][0] * Q[j][0] * e[0] + Q[i][1] * Q[j][1] * e[1] + Q[i][2] * Q[j][2] * e[2];
          T M00 = (T)M[0][0], M10 = (T)M[1][0], M20 = (T)M[2][0];
          T M11 = (T)M[1][1], M21 = (T)M[2][1], M22 = (T)M[2][2];
          vx = M00 * bx + M10 * by + M20 * bz;
          vy = M10 * bx + M11 * by + M21 * bz;
          vz = M20 * bx + M21 * by + M22 * bz;
        } else if (inverse) {
          // InverseOperatorMultiply: Cholesky GG^T = L, two triangular solves
          // on real and imaginary parts (DiffOper.cxx:1002-1083)
          T G00 = std::sqrt(L00);
          T G10 = L10 / G00;
          T G20 = L20 / G00;
          T G11 = L11 - G10 * G10;
          T G21 = L21 - G20 * G10;
          G11 = std::sqrt(G11);
          G21 = G21 / G11;
          T G22 = L22 - (G20 * G20 + G21 * G21);
          G22 = std::sqrt(G22);
          T br[3] = {bx.real(), by.real(), bz.real()};
          T bi[3] = {bx.imag(), by.imag(), bz.imag()};
          T xr[3], xi[3];
          for (int pass = 0; pass < 2; ++pass) {
            T *bb = pass ? bi : br;
            T *xx = pass ? xi : xr;
            T y0 = bb[0] / G00;
            T y1 = (bb[1] - G10 * y0) / G11;
            T y2 = (bb[2] - G20 * y0 - G21 * y1) / G22;
            xx[2] = y2 / G22;
            xx[1] = (y1 - G21 * xx[2]) / G11;
            xx[0] = (y0 - G10 * xx[1] - G20 * xx[2]) / G00;
          }
          vx = C(xr[0], xi[0]);
          vy = C(xr[1], xi[1]);
          vz = C(xr[2], xi[2]);
        } else {
          // OperatorMultiply (DiffOper.cxx:1085-1116)
          vx = L00 * bx + L10 * by + L20 * bz;
          vy = L10 * bx + L11 * by + L21 * bz;
          vz = L20 * bx + L21 * by + L22 * bz;
        }
        if (divfree && (x | y | z)) {
          // ProjectIncomp (DiffOper.cxx:948-1000)
          double nsq = (double)swx[x] * swx[x] + (double)swy[y] * swy[y] + (double)swz[z] * swz[z];
          double re = vx.real() * (double)swx[x] + vy.real() * (double)swy[y] + vz.real() * (double)swz[z];
          double im = vx.imag() * (double)swx[x] + vy.imag() * (double)swy[y] + vz.imag() * (double)swz[z];
          if (nsq > 0) {
            vx = C((T)(vx.real() - re * swx[x] / nsq), (T)(vx.imag() - im * swx[x] / nsq));
            vy = C((T)(vy.real() - re * swy[y] / nsq), (T)(vy.imag() - im * swy[y] / nsq));
            vz = C((T)(vz.real() - re * swz[z] / nsq), (T)(vz.imag() - im * swz[z] / nsq));
          }
        }
        S[0][idx] = vx;
        S[1][idx] = vy;
        S[2][idx] = vz;
      }

  for (int c = 0; c < 3; ++c) {
    F.c2r(S[c].data(), plane.data());
    if (g_mut == 7) for (size_t i = 0; i < N; ++i) plane[i] *= (T)N;  // self-test: double scaling
    for (size_t i = 0; i < N; ++i) vf[3 * i + c] = plane[i];
  }
  return 0;
}

/* ====================================================================== */
/* Interpolation                                                           */
/* ====================================================================== */

inline int safe_mod(int r, int b) {  // Numerics/Arithmetic.h:22-33
  int m = r % b;
  return m < 0 ? m + b : m;
}

/* floor with the reference's negative fix (HField3DUtils.txx:1977-1983) */
inline int ref_floor(double x) {
  int f = (int)x;
  if (x < 0 && x != (double)(int)x) --f;
  return f;
}

/* HField3DUtils::trilerp (HField3DUtils.txx:1969-2300).  h is AoS float. */
template <class TF>
void trilerp_vec(const TF *h, const int *size, double x, double y, double z, double out[3],
                 int bg) {
  int fX = ref_floor(x), fY = ref_floor(y), fZ = ref_floor(z);
  int cX = fX + 1, cY = fY + 1, cZ = fZ + 1;
  const double t = x - fX, u = y - fY, v = z - fZ;
  const double omt = 1.0 - t, omu = 1.0 - u, omv = 1.0 - v;
  const int sX = size[0], sY = size[1], sZ = size[2];
  double c[8][3];
  for (int i = 0; i < 8; ++i) c[i][0] = c[i][1] = c[i][2] = 0;
  bool inside = true;
  if (bg == ORC_BG_WRAP) {
    fX = safe_mod(fX, sX); fY = safe_mod(fY, sY); fZ = safe_mod(fZ, sZ);
    cX = safe_mod(cX, sX); cY = safe_mod(cY, sY); cZ = safe_mod(cZ, sZ);
  } else if (bg == ORC_BG_CLAMP) {
    if (fX < 0) { fX = 0; if (cX < 0) cX = 0; }
    if (cX >= sX) { cX = sX - 1; if (fX >= sX) fX = sX - 1; }
    if (fY < 0) { fY = 0; if (cY < 0) cY = 0; }
    if (cY >= sY) { cY = sY - 1; if (fY >= sY) fY = sY - 1; }
    if (fZ < 0) { fZ = 0; if (cZ < 0) cZ = 0; }
    if (cZ >= sZ) { cZ = sZ - 1; if (fZ >= sZ) fZ = sZ - 1; }
  } else {
    inside = (fX >= 0 && cX < sX && fY >= 0 && cY < sY && fZ >= 0 && cZ < sZ);
  }
  // corner order v0..v7 (HField3DUtils.txx:1998-2003)
  const int cx[8] = {fX, cX, cX, fX, fX, cX, cX, fX};
  const int cy[8] = {fY, fY, cY, cY, cY, cY, fY, fY};
  const int cz[8] = {fZ, fZ, fZ, fZ, cZ, cZ, cZ, cZ};
  auto at = [&](int i, int j, int k, int comp) -> double {
    return (double)h[3 * (((size_t)k * sY + j) * sX + i) + comp];
  };
  if (inside) {
    for (int i = 0; i < 8; ++i)
      for (int q = 0; q < 3; ++q) c[i][q] = at(cx[i], cy[i], cz[i], q);
  } else if (bg == ORC_BG_ID) {
    out[0] = x; out[1] = y; out[2] = z;
    return;
  } else if (bg == ORC_BG_ZERO) {
    out[0] = out[1] = out[2] = 0;
    return;
  } else {  // PARTIAL_ID / PARTIAL_ZERO
    for (int i = 0; i < 8; ++i) {
      bool in = cx[i] >= 0 && cx[i] < sX && cy[i] >= 0 && cy[i] < sY && cz[i] >= 0 && cz[i] < sZ;
      if (in) {
        for (int q = 0; q < 3; ++q) c[i][q] = at(cx[i], cy[i], cz[i], q);
      } else if (bg == ORC_BG_PARTIAL_ID) {
        c[i][0] = cx[i]; c[i][1] = cy[i]; c[i][2] = cz[i];
      }
    }
  }
  if (g_mut == 4) {  // self-test: swapped corners v3/v4
    for (int q = 0; q < 3; ++q) std::swap(c[3][q], c[4][q]);
  }
  for (int q = 0; q < 3; ++q)
    out[q] = omt * (omu * (c[0][q] * omv + c[7][q] * v) + u * (c[3][q] * omv + c[4][q] * v)) +
             t * (omu * (c[1][q] * omv + c[6][q] * v) + u * (c[2][q] * omv + c[5][q] * v));
}

/* Array3DUtils::trilerp<T,bg> (Array3DUtils.txx:59-232).  Scalar strategies:
 * VAL (background value), CLAMP, WRAP; ZERO maps to VAL with 0.  The scalar
 * CLAMP in the reference tests floor > size (it can read one past the end);
 * this restatement clamps with >= as the vector version does. */
template <class TF>
double trilerp_scalar(const TF *a, const int *size, double x, double y, double z, int bg,
                      double background) {
  int fX = ref_floor(x), fY = ref_floor(y), fZ = ref_floor(z);
  int cX = fX + 1, cY = fY + 1, cZ = fZ + 1;
  const double t = x - fX, u = y - fY, v = z - fZ;
  const double omt = 1.0 - t, omu = 1.0 - u, omv = 1.0 - v;
  const int sX = size[0], sY = size[1], sZ = size[2];
  bool inside = true;
  if (bg == ORC_BG_WRAP) {
    fX = safe_mod(fX, sX); fY = safe_mod(fY, sY); fZ = safe_mod(fZ, sZ);
    cX = safe_mod(cX, sX); cY = safe_mod(cY, sY); cZ = safe_mod(cZ, sZ);
  } else if (bg == ORC_BG_CLAMP) {
    if (fX < 0) { fX = 0; if (cX < 0) cX = 0; }
    if (cX >= sX) { cX = sX - 1; if (fX >= sX) fX = sX - 1; }
    if (fY < 0) { fY = 0; if (cY < 0) cY = 0; }
    if (cY >= sY) { cY = sY - 1; if (fY >= sY) fY = sY - 1; }
    if (fZ < 0) { fZ = 0; if (cZ < 0) cZ = 0; }
    if (cZ >= sZ) { cZ = sZ - 1; if (fZ >= sZ) fZ = sZ - 1; }
  } else {
    if (bg == ORC_BG_ZERO) background = 0;
    inside = (fX >= 0 && cX < sX && fY >= 0 && cY < sY && fZ >= 0 && cZ < sZ);
  }
  const int cx[8] = {fX, cX, cX, fX, fX, cX, cX, fX};
  const int cy[8] = {fY, fY, cY, cY, cY, cY, fY, fY};
  const int cz[8] = {fZ, fZ, fZ, fZ, cZ, cZ, cZ, cZ};
  double c[8];
  for (int i = 0; i < 8; ++i) {
    bool in = inside || (cx[i] >= 0 && cx[i] < sX && cy[i] >= 0 && cy[i] < sY && cz[i] >= 0 &&
                         cz[i] < sZ);
    c[i] = in ? (double)a[((size_t)cz[i] * sY + cy[i]) * sX + cx[i]] : background;
  }
  if (g_mut == 4) std::swap(c[3], c[4]);
  return omt * (omu * (c[0] * omv + c[7] * v) + u * (c[3] * omv + c[4] * v)) +
         t * (omu * (c[1] * omv + c[6] * v) + u * (c[2] * omv + c[5] * v));
}

template <class TO>
void apply_t(const float *img, const float *h, const int *size, TO *out, int bg,
             double background) {
  const size_t N = (size_t)size[0] * size[1] * size[2];
  for (size_t i = 0; i < N; ++i)
    out[i] = (TO)trilerp_scalar(img, size, (double)h[3 * i], (double)h[3 * i + 1],
                                (double)h[3 * i + 2], bg, background);
}

/* Array3DUtils::computeGradient (Array3DUtils.txx:783-901) -> AoS */
template <class TO, class TI>
void gradient_t(const TI *a, const int *size, const double *sp, int wrap, TO *g) {
  const int sx = size[0], sy = size[1], sz = size[2];
  auto A = [&](int x, int y, int z) -> double { return (double)a[((size_t)z * sy + y) * sx + x]; };
  auto G = [&](int x, int y, int z, int c) -> TO & { return g[3 * (((size_t)z * sy + y) * sx + x) + c]; };
  const int nd[3] = {sx, sy, sz};
  for (int z = 0; z < sz; ++z)
    for (int y = 0; y < sy; ++y)
      for (int x = 0; x < sx; ++x) {
        const int p[3] = {x, y, z};
        for (int c = 0; c < 3; ++c) {
          const int n = nd[c];
          int pm[3] = {x, y, z}, pp[3] = {x, y, z};
          double val;
          if (n == 1) {
            val = 0;  // degenerate axis: the reference reads out of range here
          } else if (p[c] > 0 && p[c] < n - 1) {
            pp[c] += 1; pm[c] -= 1;
            val = (A(pp[0], pp[1], pp[2]) - A(pm[0], pm[1], pm[2])) / (2.0 * sp[c]);
          } else if (wrap) {
            pp[c] = safe_mod(p[c] + 1, n); pm[c] = safe_mod(p[c] - 1, n);
            val = (A(pp[0], pp[1], pp[2]) - A(pm[0], pm[1], pm[2])) / (2.0 * sp[c]);
          } else if (p[c] == 0) {
            pp[c] = 1;
            val = (A(pp[0], pp[1], pp[2]) - A(x, y, z)) / (g_mut == 5 ? 2.0 * sp[c] : sp[c]);
          } else {
            pm[c] = n - 2;
            val = (A(x, y, z) - A(pm[0], pm[1], pm[2])) / (g_mut == 5 ? 2.0 * sp[c] : sp[c]);
          }
          G(x, y, z, c) = (TO)val;
        }
      }
}

/* ====================================================================== */
/* Greedy loop                                                             */
/* ====================================================================== */

/* One GreedyIteratorCPU::Iterate (GreedyIteratorCPU.cxx:97-138) in precision T.
 * h / hinv: AoS T, I0/I1 float.  vf, scratchI, scratchV: caller scratch. */
template <class T>
struct GreedyState {
  int size[3];
  double sp[3];
  size_t N;
  std::vector<T> Idef, vf, hnew;
  std::vector<float> hf;  // float view of h for trilerp input when T=double
  GreedyState(const int *s, const double *p) {
    for (int i = 0; i < 3; ++i) { size[i] = s[i]; sp[i] = p[i]; }
    N = (size_t)s[0] * s[1] * s[2];
    Idef.resize(N);
    vf.resize(3 * N);
    hnew.resize(3 * N);
  }
};

template <class T>
void trilerp_vec_T(const T *h, const int *size, double x, double y, double z, double out[3], int bg) {
  trilerp_vec<T>(h, size, x, y, z, out, bg);
}

template <class T, class TI1 = float>
double greedy_iterate(GreedyState<T> &st, const orc_greedy_params *p, const float *I0,
                      const TI1 *I1, T *h, T *hinv, bool update_step, double &step,
                      double *energy_float_sum, const int *isize = nullptr,
                      const double *prev_energy = nullptr) {
  const int *size = st.size;
  const size_t N = st.N;
  const int sx = size[0], sy = size[1];
  // 1. deformed image (HField3DUtils::apply, VAL background 0); I0 may live on
  //    its own grid isize (HField3DUtils.txx:728-752 interpolates on the image's size)
  const int *img_size = isize ? isize : size;
  for (size_t i = 0; i < N; ++i)
    st.Idef[i] = (T)trilerp_scalar(I0, img_size, (double)h[3 * i], (double)h[3 * i + 1],
                                   (double)h[3 * i + 2], ORC_BG_VAL, 0.0);
  // 2. gradient into the operator's field
  {
    std::vector<float> idf;
    const float *src;
    if (sizeof(T) == sizeof(float)) {
      src = reinterpret_cast<const float *>(st.Idef.data());
    } else {
      // FP64 oracle: gradient of the double image
      std::vector<T> &I = st.Idef;
      const int nd[3] = {size[0], size[1], size[2]};
      for (int z = 0; z < size[2]; ++z)
        for (int y = 0; y < sy; ++y)
          for (int x = 0; x < sx; ++x) {
            const int pp[3] = {x, y, z};
            for (int c = 0; c < 3; ++c) {
              const int n = nd[c];
              int a[3] = {x, y, z}, b[3] = {x, y, z};
              double den;
              if (n == 1) { st.vf[3 * (((size_t)z * sy + y) * sx + x) + c] = 0; continue; }
              if (pp[c] > 0 && pp[c] < n - 1) { a[c] += 1; b[c] -= 1; den = 2.0 * p->sp[c]; }
              else if (pp[c] == 0) { a[c] = 1; den = p->sp[c]; }
              else { b[c] = n - 2; den = p->sp[c]; }
              double va = I[((size_t)a[2] * sy + a[1]) * sx + a[0]];
              double vb = I[((size_t)b[2] * sy + b[1]) * sx + b[0]];
              st.vf[3 * (((size_t)z * sy + y) * sx + x) + c] = (T)((va - vb) / den);
            }
          }
      src = nullptr;
    }
    if (src) gradient_t<T>(src, size, p->sp, 0, st.vf.data());
  }
  // 3-4. residual r = -(Idef - I1); force = grad * r
  //      (GreedyIteratorCPU.cxx:112-117, DiffOper.cxx:512-525)
  T fsum = 0;
  double dsum = 0;
  for (size_t i = 0; i < N; ++i) {
    T r = (T)((T)(st.Idef[i] - (T)I1[i]) * (T)-1.0);
    st.Idef[i] = r;
    st.vf[3 * i] *= r;
    st.vf[3 * i + 1] *= r;
    st.vf[3 * i + 2] *= r;
    fsum += r * r;  // Array3DUtils.txx:1624-1638 (accumulated in T)
    dsum += (double)r * (double)r;
  }
  const double spprod = p->sp[0] * p->sp[1] * p->sp[2];
  if (energy_float_sum) *energy_float_sum = spprod * (double)fsum;  // ImageUtils.txx:55-58
  // GreedyDeformationData::AddEnergy (GreedyDeformationData.cxx:162-184) runs here,
  // before the update (GreedyIteratorCPU.cxx:120-136): under AutoStepReduce an
  // energy above the previous iteration's halves the step this iteration uses.
  if (p->auto_step_reduce && prev_energy) {
    const double e = (sizeof(T) == sizeof(float)) ? spprod * (double)fsum : spprod * dsum;
    if (e - *prev_energy > 0) step = step / 2.0;
  }
  // 5. v = L^-1 f
  diffoper<T>(size, p->sp, p->alpha, p->beta, p->gamma, 1.0, 0, st.vf.data(), true);
  // 6. step size at the first iteration of a scale level (GreedyIteratorCPU.cxx:65-79,129-134)
  if (update_step) {
    T mx = 0;
    for (size_t i = 0; i < N; ++i) {
      T l = std::sqrt(st.vf[3 * i] * st.vf[3 * i] + st.vf[3 * i + 1] * st.vf[3 * i + 1] +
                      st.vf[3 * i + 2] * st.vf[3 * i + 2]);
      if (l > mx) mx = l;
    }
    step = (double)((T)p->max_pert / mx);
  }
  // 7. updateDeformation (GreedyIteratorCPU.cxx:81-95)
  const T stepT = (T)step;
  for (size_t i = 0; i < 3 * N; ++i) st.vf[i] = (T)(st.vf[i] * stepT);
  const T spx = (T)p->sp[0], spy = (T)p->sp[1], spz = (T)p->sp[2];
  size_t i = 0;
  for (int z = 0; z < size[2]; ++z)
    for (int y = 0; y < sy; ++y)
      for (int x = 0; x < sx; ++x, ++i) {
        double o[3];
        T qx = (T)x + st.vf[3 * i] / spx;
        T qy = (T)y + st.vf[3 * i + 1] / spy;
        T qz = (T)z + st.vf[3 * i + 2] / spz;
        if (g_mut == 8) { qx = (T)x - st.vf[3 * i] / spx; }  // self-test: wrong direction
        trilerp_vec<T>(h, size, (double)qx, (double)qy, (double)qz, o, ORC_BG_PARTIAL_ID);
        st.hnew[3 * i] = (T)o[0];
        st.hnew[3 * i + 1] = (T)o[1];
        st.hnew[3 * i + 2] = (T)o[2];
      }
  std::memcpy(h, st.hnew.data(), sizeof(T) * 3 * N);
  if (p->compute_inverse && hinv) {
    // composeVHInv(v, hinv, PARTIAL_ZERO): hinv(x) - v(hinv(x))/sp
    for (size_t j = 0; j < N; ++j) {
      double o[3];
      trilerp_vec<T>(st.vf.data(), size, (double)hinv[3 * j], (double)hinv[3 * j + 1],
                     (double)hinv[3 * j + 2], o, ORC_BG_PARTIAL_ZERO);
      st.hnew[3 * j] = (T)((T)o[0] / -spx + hinv[3 * j]);
      st.hnew[3 * j + 1] = (T)((T)o[1] / -spy + hinv[3 * j + 1]);
      st.hnew[3 * j + 2] = (T)((T)o[2] / -spz + hinv[3 * j + 2]);
    }
    std::memcpy(hinv, st.hnew.data(), sizeof(T) * 3 * N);
  }
  return spprod * dsum;
}

template <class T>
int greedy_run(const orc_greedy_params *p, const float *I0, const float *I1, T *h, T *hinv,
               double *energy_f, double *energy_d, double *step_out) {
  GreedyState<T> st(p->size, p->sp);
  double step = 0;
  double prev = 0;
  for (int it = 0; it < p->n_iter; ++it) {
    double ef = 0;
    double ed = greedy_iterate<T>(st, p, I0, I1, h, hinv, it == 0, step, &ef, nullptr,
                                  it > 0 ? &prev : nullptr);
    if (energy_f) energy_f[it] = ef;
    if (energy_d) energy_d[it] = ed;
    prev = (sizeof(T) == sizeof(float)) ? ef : ed;
  }
  if (step_out) *step_out = step;
  return 0;
}

/* ====================================================================== */
/* Multiscale: Gaussian downsample, hfield resample                        */
/* ====================================================================== */

/* filter_generate_gauss_full (Utilities/DownsampleFilter/GaussianFilter3D.cxx:
 * 92-140): row 0 the normalised half kernel, row j the kernel renormalised for
 * an element j taps short of a full window (float arithmetic as there). */
std::vector<std::vector<float>> gauss_table(float sigma, int k) {
  std::vector<std::vector<float>> f(k + 1, std::vector<float>(k + 1));
  const float es = (float)(1.0 / (2.0 * sigma * sigma));
  float sc = 1.0f;
  f[0][0] = 1.0f;
  for (int i = 1; i <= k; ++i) {
    f[0][i] = std::exp(-(float)(i * i) * es);
    sc += 2.0f * f[0][i];
  }
  sc = 1.0f / sc;
  for (int i = 0; i <= k; ++i) f[0][i] *= sc;
  float miss = 0.0f;
  for (int j = 1; j <= k; ++j) {
    miss += f[0][k + 1 - j];
    for (int i = 0; i <= k; ++i) f[j][i] = (float)(f[0][i] / (1. - miss));
  }
  return f;
}

/* One axis of filter_symmetric_full (GaussianFilter3D.cxx:159-330): interior
 * elements take the full window; the k elements at either edge take the
 * one-sided window with the renormalised kernel. */
void gauss_axis(const float *in, float *out, const int *size, int axis, float sigma, int k) {
  const int n = size[axis];
  const auto f = gauss_table(sigma, k);
  const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)size[0] : (size_t)size[0] * size[1];
  const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
  for (int b = 0; b < size[o2]; ++b)
    for (int a = 0; a < size[o1]; ++a) {
      size_t base;
      if (axis == 0) base = ((size_t)b * size[1] + a) * size[0];
      else if (axis == 1) base = (size_t)b * size[0] * size[1] + a;
      else base = (size_t)b * size[0] + a;
      auto I = [&](int i) { return in[base + (size_t)i * stride]; };
      for (int x = k; x < n - k; ++x) {
        float acc = I(x) * f[0][0];
        for (int i = 1; i <= k; ++i) acc += (I(x - i) + I(x + i)) * f[0][i];
        out[base + (size_t)x * stride] = acc;
      }
      for (int x = k - 1; x >= 0; --x) {
        const auto &g = f[k - x];
        float acc = I(x) * g[0];
        for (int i = 1; i <= x; ++i) acc += (I(x - i) + I(x + i)) * g[i];
        for (int i = x + 1; i <= k; ++i) acc += I(x + i) * g[i];
        out[base + (size_t)x * stride] = acc;
        const int xr = n - 1 - x;
        acc = I(xr) * g[0];
        for (int i = 1; i <= x; ++i) acc += (I(xr + i) + I(xr - i)) * g[i];
        for (int i = x + 1; i <= k; ++i) acc += I(xr - i) * g[i];
        out[base + (size_t)xr * stride] = acc;
      }
    }
}

/* ====================================================================== */
/* Greedy atlas                                                            */
/* ====================================================================== */

/* GreedyAtlasBuilder / AtlasThread (AtlasThread.txx:95-181, GreedyAtlasBuilder.txx:
 * 236-301, GreedyAtlasThreadGPU.cxx:161-206) at one scale level, FP64: the mean
 * sum_i w_i I0_i(h_i) is formed at the start and after every iteration; each
 * iteration runs one greedy Iterate per subject against the current mean (the
 * step size set at iteration 0); the iteration energy is sum_i w_i E_i. */
int atlas_run_d(const orc_atlas_params *p, const float *const *I0, const int *isizes,
                const double *w, double *const *h, double *const *hinv, double *mean,
                double *energy) {
  const int *size = p->size;
  const size_t N = (size_t)size[0] * size[1] * size[2];
  const int ns = p->n_subjects;
  orc_greedy_params gp;
  for (int i = 0; i < 3; ++i) { gp.size[i] = size[i]; gp.sp[i] = p->sp[i]; }
  gp.alpha = p->alpha; gp.beta = p->beta; gp.gamma = p->gamma; gp.max_pert = p->max_pert;
  gp.n_iter = 1; gp.compute_inverse = hinv ? 1 : 0; gp.auto_step_reduce = 0;
  GreedyState<double> st(size, p->sp);
  std::vector<double> steps(ns, 0.0), m(N);
  auto compute_mean = [&]() {
    std::fill(m.begin(), m.end(), 0.0);
    for (int s = 0; s < ns; ++s)
      for (size_t i = 0; i < N; ++i)
        m[i] += w[s] * trilerp_scalar(I0[s], isizes + 3 * s, h[s][3 * i], h[s][3 * i + 1],
                                      h[s][3 * i + 2], ORC_BG_VAL, 0.0);
  };
  compute_mean();
  for (int it = 0; it < p->n_iter; ++it) {
    double e = 0;
    for (int s = 0; s < ns; ++s)
      e += w[s] * greedy_iterate<double, double>(st, &gp, I0[s], m.data(), h[s],
                                                 hinv ? hinv[s] : nullptr, it == 0, steps[s],
                                                 nullptr, isizes + 3 * s);
    if (energy) energy[it] = e;
    compute_mean();
  }
  if (mean) std::memcpy(mean, m.data(), N * sizeof(double));
  return 0;
}

/* ====================================================================== */
/* LDMM                                                                    */
/* ====================================================================== */

/* cpuSplat3D semantics (cudaSplatAtomic.cu:602-663) for positions and masses of
 * type T, accumulated in double. */
template <class T>
void splat_t(const T *m, const T *h, const int *size, double *out) {
  const int w = size[0], hh = size[1], l = size[2];
  const size_t N = (size_t)w * hh * l;
  for (size_t i = 0; i < N; ++i) out[i] = 0.0;
  for (size_t i = 0; i < N; ++i) {
    const double x = h[3 * i], y = h[3 * i + 1], z = h[3 * i + 2];
    const int xi = ref_floor(x), yi = ref_floor(y), zi = ref_floor(z);
    const double dx = 1.0 - (x - xi), dy = 1.0 - (y - yi), dz = 1.0 - (z - zi);
    const double wx[2] = {dx, 1.0 - dx}, wy[2] = {dy, 1.0 - dy}, wz[2] = {dz, 1.0 - dz};
    for (int c = 0; c < 8; ++c) {
      const int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
      const int X = xi + ox, Y = yi + oy, Z = zi + oz;
      if (X >= 0 && X < w && Y >= 0 && Y < hh && Z >= 0 && Z < l)
        out[((size_t)Z * hh + Y) * w + X] += (double)m[i] * wx[ox] * wy[oy] * wz[oz];
    }
  }
}

/* LDMMIteratorCPU::IterateNew (LDMMIteratorCPU.cxx:529-631) with its helpers
 * updateVelocity (:175-190) and calcMaxDisplacement (:192-212); energies as
 * LDMMEnergy (LDMMEnergy.cxx:48-64, LDMMEnergy.h:42): image energy
 * Prod(sp) * sum(diff^2) / sigma^2, vector energy mean over t of
 * Prod(sp) * <L v_t, v_t> (LDMMIteratorGPU.cxx:250-256).  v: T fields, AoS. */
template <class T>
void ldmm_iterate(const orc_ldmm_params *p, const float *I0, const float *I1, T *v,
                  bool update_step, double &step, double *e_img, double *e_vec, float *alpha) {
  const int *size = p->size;
  const int sx = size[0], sy = size[1], sz = size[2];
  const size_t N = (size_t)sx * sy * sz;
  const int nt = p->n_time_steps;
  const double sig2 = p->sigma * p->sigma;
  const double vol = p->sp[0] * p->sp[1] * p->sp[2];
  std::vector<T> h(3 * N), hs(3 * N), J((size_t)nt * N), diff(N), f(3 * N), lv;
  std::vector<double> si(N);
  auto identity = [&](std::vector<T> &g) {
    size_t i = 0;
    for (int z = 0; z < sz; ++z)
      for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x, ++i) {
          g[3 * i] = (T)x; g[3 * i + 1] = (T)y; g[3 * i + 2] = (T)z;
        }
  };
  // forward: h <- composeHVInv(h, v_t) (PARTIAL_ID); J_t = I0(h)
  identity(h);
  for (int t = 0; t < nt; ++t) {
    const T *vt = v + (size_t)t * 3 * N;
    size_t i = 0;
    for (int z = 0; z < sz; ++z)
      for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x, ++i) {
          double o[3];
          trilerp_vec<T>(h.data(), size, x - (double)vt[3 * i] / p->sp[0],
                         y - (double)vt[3 * i + 1] / p->sp[1], z - (double)vt[3 * i + 2] / p->sp[2],
                         o, ORC_BG_PARTIAL_ID);
          hs[3 * i] = (T)o[0]; hs[3 * i + 1] = (T)o[1]; hs[3 * i + 2] = (T)o[2];
        }
    std::swap(h, hs);
    T *Jt = J.data() + (size_t)t * N;
    for (size_t j = 0; j < N; ++j)
      Jt[j] = (T)trilerp_scalar<float>(I0, size, (double)h[3 * j], (double)h[3 * j + 1],
                                       (double)h[3 * j + 2], ORC_BG_VAL, 0.0);
  }
  double ssd = 0;
  const T *JT = J.data() + (size_t)(nt - 1) * N;
  for (size_t j = 0; j < N; ++j) {
    diff[j] = (T)(JT[j] - (T)I1[j]);
    ssd += (double)diff[j] * (double)diff[j];
  }
  if (e_img) *e_img = vol * ssd / sig2;
  // backward: h <- composeVHInv(v_t, h) (PARTIAL_ZERO); alpha_t = splat(diff, h)
  identity(h);
  double maxdisp = 0, vec = 0;
  for (int t = nt - 1; t >= 0; --t) {
    T *vt = v + (size_t)t * 3 * N;
    for (size_t j = 0; j < N; ++j) {
      double o[3];
      trilerp_vec<T>(vt, size, (double)h[3 * j], (double)h[3 * j + 1], (double)h[3 * j + 2], o,
                     ORC_BG_PARTIAL_ZERO);
      for (int c = 0; c < 3; ++c) hs[3 * j + c] = (T)(o[c] / -p->sp[c] + (double)h[3 * j + c]);
    }
    std::swap(h, hs);
    splat_t<T>(diff.data(), h.data(), size, si.data());
    if (alpha) {
      for (size_t j = 0; j < N; ++j) alpha[(size_t)t * N + j] = (float)(si[j] / sig2);
      continue;
    }
    // force = grad(J_{t-1}) * alpha_t (grad(I0) at t = 0)
    if (t > 0) gradient_t<T>(J.data() + (size_t)(t - 1) * N, size, p->sp, 0, f.data());
    else gradient_t<T>(I0, size, p->sp, 0, f.data());
    for (size_t j = 0; j < N; ++j)
      for (int c = 0; c < 3; ++c) f[3 * j + c] = (T)(f[3 * j + c] * (T)si[j]);
    diffoper<T>(size, p->sp, p->alpha, p->beta, p->gamma, 1.0, 0, f.data(), true);
    if (update_step) {
      for (size_t j = 0; j < N; ++j) {
        double l2 = 0;
        for (int c = 0; c < 3; ++c) {
          const double q = 2.0 * (double)vt[3 * j + c] - (2.0 / sig2) * (double)f[3 * j + c];
          l2 += q * q;
        }
        maxdisp = std::max(maxdisp, std::sqrt(l2));
      }
    } else {
      const T a = (T)(1.0 - 2.0 * step), b = (T)(2.0 * step / sig2);
      for (size_t j = 0; j < 3 * N; ++j) vt[j] = (T)(a * vt[j] + b * f[j]);
      if (e_vec) {
        lv.assign(vt, vt + 3 * N);
        diffoper<T>(size, p->sp, p->alpha, p->beta, p->gamma, 1.0, 0, lv.data(), false);
        double d = 0;
        for (size_t j = 0; j < 3 * N; ++j) d += (double)lv[j] * (double)vt[j];
        vec += vol * d;
      }
    }
  }
  if (update_step && !alpha) step = p->max_pert / maxdisp;
  if (e_vec) *e_vec = update_step ? 0.0 : vec / nt;
}

/* HField3DUtils::jacobian (HField3DUtils.txx:2370-2420): det of the gradients
 * (Array3DUtils::computeGradient, wrap = false) of the three components of an AoS
 * field given in world units, with spacing sp. */
template <class T>
void jacobian_det(const T *hw, const int *size, const double *sp, T *det) {
  const size_t N = (size_t)size[0] * size[1] * size[2];
  std::vector<T> c(N), g1(3 * N), g2(3 * N), g3(3 * N);
  for (int k = 0; k < 3; ++k) {
    for (size_t i = 0; i < N; ++i) c[i] = hw[3 * i + k];
    gradient_t<T>(c.data(), size, sp, 0, k == 0 ? g1.data() : k == 1 ? g2.data() : g3.data());
  }
  for (size_t i = 0; i < N; ++i) {
    const T *a = &g1[3 * i], *b = &g2[3 * i], *e = &g3[3 * i];
    const T t1 = a[0] * (b[1] * e[2] - b[2] * e[1]);
    const T t2 = a[1] * (b[0] * e[2] - b[2] * e[0]);
    const T t3 = a[2] * (b[0] * e[1] - b[1] * e[0]);
    det[i] = t1 - t2 + t3;
  }
}

/* Backward sweep of LDMMIteratorCPU::IterateOrig (LDMMIteratorCPU.cxx:391-428),
 * t = T-1..0: h <- composeHV(h, v_t) (PARTIAL_ID), JT_t = I1(h) (VAL 0),
 * |D(x sp + v_t)| and JacDet_t = applyU(JacDet_{t+1}, v_t) |D(x sp + v_t)|
 * (HField3DUtils::applyU, HField3DUtils.txx:1029-1053: the image at x + v/sp,
 * VAL 0), JacDet_{T-1} = the determinant. */
template <class T>
void ldmm_orig_backward(const orc_ldmm_params *p, const float *I1, const T *v, T *JT, T *JD) {
  const int *size = p->size;
  const int sx = size[0], sy = size[1], sz = size[2];
  const size_t N = (size_t)sx * sy * sz;
  const int nt = p->n_time_steps;
  std::vector<T> h(3 * N), hs(3 * N), det(N), wv(3 * N);
  size_t i = 0;
  for (int z = 0; z < sz; ++z)
    for (int y = 0; y < sy; ++y)
      for (int x = 0; x < sx; ++x, ++i) {
        h[3 * i] = (T)x; h[3 * i + 1] = (T)y; h[3 * i + 2] = (T)z;
      }
  for (int t = nt - 1; t >= 0; --t) {
    const T *vt = v + (size_t)t * 3 * N;
    i = 0;
    for (int z = 0; z < sz; ++z)
      for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x, ++i) {
          double o[3];
          trilerp_vec<T>(h.data(), size, x + (double)vt[3 * i] / p->sp[0],
                         y + (double)vt[3 * i + 1] / p->sp[1], z + (double)vt[3 * i + 2] / p->sp[2],
                         o, ORC_BG_PARTIAL_ID);
          hs[3 * i] = (T)o[0]; hs[3 * i + 1] = (T)o[1]; hs[3 * i + 2] = (T)o[2];
          wv[3 * i] = (T)(x * p->sp[0]) + vt[3 * i];
          wv[3 * i + 1] = (T)(y * p->sp[1]) + vt[3 * i + 1];
          wv[3 * i + 2] = (T)(z * p->sp[2]) + vt[3 * i + 2];
        }
    std::swap(h, hs);
    T *JTt = JT + (size_t)t * N;
    for (size_t j = 0; j < N; ++j)
      JTt[j] = (T)trilerp_scalar<float>(I1, size, (double)h[3 * j], (double)h[3 * j + 1],
                                        (double)h[3 * j + 2], ORC_BG_VAL, 0.0);
    jacobian_det<T>(wv.data(), size, p->sp, det.data());
    T *JDt = JD + (size_t)t * N;
    if (t == nt - 1) {
      for (size_t j = 0; j < N; ++j) JDt[j] = det[j];
    } else {
      const T *JDn = JD + (size_t)(t + 1) * N;
      size_t q = 0;
      for (int z = 0; z < sz; ++z)
        for (int y = 0; y < sy; ++y)
          for (int x = 0; x < sx; ++x, ++q)
            JDt[q] = (T)(trilerp_scalar<T>(JDn, size, x + (double)vt[3 * q] / p->sp[0],
                                           y + (double)vt[3 * q + 1] / p->sp[1],
                                           z + (double)vt[3 * q + 2] / p->sp[2], ORC_BG_VAL, 0.0) *
                         det[q]);
    }
  }
}

/* LDMMIteratorCPU::IterateOrig (LDMMIteratorCPU.cxx:379-527), the
 * Jacobian-determinant gradient (the reference's default: mUseOrigGradientCalc is
 * set true in both iterators' constructors).  After the backward sweep, the
 * forward sweep t = 0..T-1: J = I0(h) (I0 at t = 0), m = (J - JT_t) JacDet_t;
 * alphas-only: alpha_t = m / sigma^2; else u = L^-1 (grad(J) m) and the update of
 * v_t (or the step-size maximum); then h <- composeHVInv(h, v_t) with the
 * updated v_t.  Image energy after the sweep: Prod(sp) sum(I0(h) - I1)^2 /
 * sigma^2 (LDMMIteratorGPU.cxx:430-441, LDMMEnergy.h). */
template <class T>
void ldmm_iterate_orig(const orc_ldmm_params *p, const float *I0, const float *I1, T *v,
                       bool update_step, double &step, double *e_img, double *e_vec,
                       float *alpha) {
  const int *size = p->size;
  const int sx = size[0], sy = size[1], sz = size[2];
  const size_t N = (size_t)sx * sy * sz;
  const int nt = p->n_time_steps;
  const double sig2 = p->sigma * p->sigma;
  const double vol = p->sp[0] * p->sp[1] * p->sp[2];
  std::vector<T> h(3 * N), hs(3 * N), JT((size_t)nt * N), JD((size_t)nt * N), f(3 * N), J(N), lv;
  ldmm_orig_backward<T>(p, I1, v, JT.data(), JD.data());
  size_t i = 0;
  for (int z = 0; z < sz; ++z)
    for (int y = 0; y < sy; ++y)
      for (int x = 0; x < sx; ++x, ++i) {
        h[3 * i] = (T)x; h[3 * i + 1] = (T)y; h[3 * i + 2] = (T)z;
      }
  double maxdisp = 0, vec = 0;
  for (int t = 0; t < nt; ++t) {
    T *vt = v + (size_t)t * 3 * N;
    if (t == 0) {
      for (size_t j = 0; j < N; ++j) J[j] = (T)I0[j];
    } else {
      for (size_t j = 0; j < N; ++j)
        J[j] = (T)trilerp_scalar<float>(I0, size, (double)h[3 * j], (double)h[3 * j + 1],
                                        (double)h[3 * j + 2], ORC_BG_VAL, 0.0);
    }
    const T *JTt = JT.data() + (size_t)t * N, *JDt = JD.data() + (size_t)t * N;
    if (alpha) {
      for (size_t j = 0; j < N; ++j)
        alpha[(size_t)t * N + j] = (float)((J[j] - JTt[j]) * JDt[j] / sig2);
    } else {
      gradient_t<T>(J.data(), size, p->sp, 0, f.data());
      for (size_t j = 0; j < N; ++j) {
        const T a = (T)((J[j] - JTt[j]) * JDt[j]);
        for (int c = 0; c < 3; ++c) f[3 * j + c] = (T)(f[3 * j + c] * a);
      }
      diffoper<T>(size, p->sp, p->alpha, p->beta, p->gamma, 1.0, 0, f.data(), true);
      if (update_step) {
        for (size_t j = 0; j < N; ++j) {
          double l2 = 0;
          for (int c = 0; c < 3; ++c) {
            const double q = 2.0 * (double)vt[3 * j + c] - (2.0 / sig2) * (double)f[3 * j + c];
            l2 += q * q;
          }
          maxdisp = std::max(maxdisp, std::sqrt(l2));
        }
      } else {
        const T a = (T)(1.0 - 2.0 * step), b = (T)(2.0 * step / sig2);
        for (size_t j = 0; j < 3 * N; ++j) vt[j] = (T)(a * vt[j] + b * f[j]);
        if (e_vec) {
          lv.assign(vt, vt + 3 * N);
          diffoper<T>(size, p->sp, p->alpha, p->beta, p->gamma, 1.0, 0, lv.data(), false);
          double d = 0;
          for (size_t j = 0; j < 3 * N; ++j) d += (double)lv[j] * (double)vt[j];
          vec += vol * d;
        }
      }
    }
    i = 0;
    for (int z = 0; z < sz; ++z)
      for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x, ++i) {
          double o[3];
          trilerp_vec<T>(h.data(), size, x - (double)vt[3 * i] / p->sp[0],
                         y - (double)vt[3 * i + 1] / p->sp[1], z - (double)vt[3 * i + 2] / p->sp[2],
                         o, ORC_BG_PARTIAL_ID);
          hs[3 * i] = (T)o[0]; hs[3 * i + 1] = (T)o[1]; hs[3 * i + 2] = (T)o[2];
        }
    std::swap(h, hs);
  }
  if (alpha) return;
  if (update_step) step = p->max_pert / maxdisp;
  if (e_img) {
    double ssd = 0;
    for (size_t j = 0; j < N; ++j) {
      const double d = trilerp_scalar<float>(I0, size, (double)h[3 * j], (double)h[3 * j + 1],
                                             (double)h[3 * j + 2], ORC_BG_VAL, 0.0) - (double)I1[j];
      ssd += d * d;
    }
    *e_img = vol * ssd / sig2;
  }
  if (e_vec) *e_vec = update_step ? 0.0 : vec / nt;
}

template <class T>
int ldmm_run_orig(const orc_ldmm_params *p, const float *I0, const float *I1, T *v, double *e_img,
                  double *e_vec, double *step_out) {
  double step = step_out ? *step_out : 0.0;
  if (step <= 0.0) ldmm_iterate_orig<T>(p, I0, I1, v, true, step, nullptr, nullptr, nullptr);
  for (int it = 0; it < p->n_iter; ++it)
    ldmm_iterate_orig<T>(p, I0, I1, v, false, step, e_img ? e_img + it : nullptr,
                         e_vec ? e_vec + it : nullptr, nullptr);
  if (step_out) *step_out = step;
  return 0;
}

/* LDMM atlas (LDMMAtlasBuilder.txx:270-340, LDMMAtlasThreadGPU.cxx ComputeThreadMean,
 * AtlasThread.txx AtlasThreadMain), single scale, one thread: the mean of the
 * subjects' I_s(h_0) (GetDefTTo1(0): composeHV of v_{T-1}..v_0) weighted by w_s
 * and, with JacobianScale, by JacDet_0 and divided by sum_s w_s JacDet_0; then per
 * iteration each subject's IterateOrig against the mean (I0 = mean, I1 = I_s; a
 * step-size call first at iteration 0 when adaptive), the new mean, and the
 * energy sum_s w_s E_s. */
void ldmm_atlas_mean(const orc_ldmm_params *p, int S, const float *I, const double *w, int jscale,
                     const double *v, double *mean) {
  const size_t N = (size_t)p->size[0] * p->size[1] * p->size[2];
  const int nt = p->n_time_steps;
  std::vector<double> JT((size_t)nt * N), JD((size_t)nt * N), den(N, 0.0);
  std::fill(mean, mean + N, 0.0);
  for (int s = 0; s < S; ++s) {
    ldmm_orig_backward<double>(p, I + (size_t)s * N, v + (size_t)s * nt * 3 * N, JT.data(), JD.data());
    for (size_t i = 0; i < N; ++i) {
      if (jscale) {
        mean[i] += JT[i] * JD[i] * w[s];
        den[i] += JD[i] * w[s];
      } else {
        mean[i] += JT[i] * w[s];
      }
    }
  }
  if (jscale)
    for (size_t i = 0; i < N; ++i) mean[i] /= den[i];
}

int ldmm_atlas_run(const orc_ldmm_params *p, int S, const float *I, const double *w, int jscale,
                   int use_orig, double *v, double *steps, double *mean_out, double *energy) {
  auto iterate = use_orig ? ldmm_iterate_orig<double> : ldmm_iterate<double>;
  const size_t N = (size_t)p->size[0] * p->size[1] * p->size[2];
  const int nt = p->n_time_steps;
  std::vector<double> mean(N);
  std::vector<float> meanf(N);
  ldmm_atlas_mean(p, S, I, w, jscale, v, mean.data());
  for (int it = 0; it < p->n_iter; ++it) {
    for (size_t i = 0; i < N; ++i) meanf[i] = (float)mean[i];
    double e[3] = {0, 0, 0};
    for (int s = 0; s < S; ++s) {
      double *vs = v + (size_t)s * nt * 3 * N;
      if (it == 0 && steps[s] <= 0.0)
        iterate(p, meanf.data(), I + (size_t)s * N, vs, true, steps[s], nullptr, nullptr, nullptr);
      double ei = 0, ev = 0;
      iterate(p, meanf.data(), I + (size_t)s * N, vs, false, steps[s], &ei, &ev, nullptr);
      e[0] += w[s] * ei;
      e[1] += w[s] * ev;
    }
    if (energy) {
      energy[3 * it] = e[0];
      energy[3 * it + 1] = e[1];
      energy[3 * it + 2] = e[0] + e[1];
    }
    ldmm_atlas_mean(p, S, I, w, jscale, v, mean.data());
  }
  std::copy(mean.begin(), mean.end(), mean_out);
  return 0;
}

template <class T>
int ldmm_run(const orc_ldmm_params *p, const float *I0, const float *I1, T *v, double *e_img,
             double *e_vec, double *step_out) {
  // LDMMWarp::RunWarp (LDMMWarp.txx:97-110): with an adaptive step, one extra
  // IterateNew call before the first iteration only sets the step size.
  double step = step_out ? *step_out : 0.0;
  if (step <= 0.0) ldmm_iterate<T>(p, I0, I1, v, true, step, nullptr, nullptr, nullptr);
  for (int it = 0; it < p->n_iter; ++it)
    ldmm_iterate<T>(p, I0, I1, v, false, step, e_img ? e_img + it : nullptr,
                    e_vec ? e_vec + it : nullptr, nullptr);
  if (step_out) *step_out = step;
  return 0;
}

}  // namespace

/* ====================================================================== */
/* C entry points                                                          */
/* ====================================================================== */

extern "C" {

int orc_fft_d(int n, double *data, int inverse) {
  if (n <= 0) return -1;
  Fft<double> f(n);
  f.run(reinterpret_cast<std::complex<double> *>(data), inverse != 0);
  return 0;
}

int orc_fft_f(int n, float *data, int inverse) {
  if (n <= 0) return -1;
  Fft<float> f(n);
  f.run(reinterpret_cast<std::complex<float> *>(data), inverse != 0);
  return 0;
}

void orc_lut(int n, double sp, double *cosW, double *sinW) {
  std::vector<double> c, s;
  make_lut<double>(n, sp, c, s);
  std::copy(c.begin(), c.end(), cosW);
  std::copy(s.begin(), s.end(), sinW);
}

int orc_diffoper_d(const int *size, const double *sp, double alpha, double beta, double gamma,
                   double lpow, int flags, double *vf, int inverse) {
  return diffoper<double>(size, sp, alpha, beta, gamma, lpow, flags, vf, inverse != 0);
}

int orc_diffoper_f(const int *size, const double *sp, double alpha, double beta, double gamma,
                   double lpow, int flags, float *vf, int inverse) {
  return diffoper<float>(size, sp, alpha, beta, gamma, lpow, flags, vf, inverse != 0);
}

void orc_trilerp_vec_d(const float *h, const int *size, double x, double y, double z,
                       double *out3, int bg) {
  trilerp_vec<float>(h, size, x, y, z, out3, bg);
}

double orc_trilerp_scalar_d(const float *img, const int *size, double x, double y, double z,
                            int bg, double background) {
  return trilerp_scalar<float>(img, size, x, y, z, bg, background);
}

void orc_apply_d(const float *img, const float *h, const int *size, double *out, int bg,
                 double background) {
  apply_t<double>(img, h, size, out, bg, background);
}

void orc_apply_f(const float *img, const float *h, const int *size, float *out, int bg,
                 double background) {
  apply_t<float>(img, h, size, out, bg, background);
}

void orc_compose_d(const float *f, const float *g, const int *size, double *h, int bg) {
  const size_t N = (size_t)size[0] * size[1] * size[2];
  for (size_t i = 0; i < N; ++i)
    trilerp_vec<float>(f, size, g[3 * i], g[3 * i + 1], g[3 * i + 2], h + 3 * i, bg);
}

void orc_compose_hv_d(const float *g, const float *v, const int *size, const double *sp,
                      double *h, int bg, int inv) {
  const int sx = size[0], sy = size[1], sz = size[2];
  const double s = inv ? -1.0 : 1.0;
  size_t i = 0;
  for (int z = 0; z < sz; ++z)
    for (int y = 0; y < sy; ++y)
      for (int x = 0; x < sx; ++x, ++i) {
        double qx = x + s * (double)v[3 * i] / sp[0];
        double qy = y + s * (double)v[3 * i + 1] / sp[1];
        double qz = z + s * (double)v[3 * i + 2] / sp[2];
        if (g_mut == 8) qx = x - s * (double)v[3 * i] / sp[0];
        trilerp_vec<float>(g, size, qx, qy, qz, h + 3 * i, bg);
      }
}

void orc_compose_vh_d(const float *v, const float *g, const int *size, const double *sp,
                      double *h, int bg, int inv) {
  const size_t N = (size_t)size[0] * size[1] * size[2];
  const double s = inv ? -1.0 : 1.0;
  for (size_t i = 0; i < N; ++i) {
    double o[3];
    trilerp_vec<float>(v, size, g[3 * i], g[3 * i + 1], g[3 * i + 2], o, bg);
    for (int c = 0; c < 3; ++c) h[3 * i + c] = (double)g[3 * i + c] + s * o[c] / sp[c];
  }
}

void orc_forward_apply_f(const float *img, const float *h, const int *size, float *out,
                         double background, int normalize) {
  // HField3DUtils::forwardApply (HField3DUtils.txx:1281-1413), float accumulation
  const int sx = size[0], sy = size[1], sz = size[2];
  const size_t N = (size_t)sx * sy * sz;
  std::vector<float> count(N, 0.0f);
  for (size_t i = 0; i < N; ++i) out[i] = 0.0f;
  auto O = [&](int x, int y, int z) -> size_t { return ((size_t)z * sy + y) * sx + x; };
  for (size_t i = 0; i < N; ++i) {
    float ix = img[i];
    float hx = h[3 * i], hy = h[3 * i + 1], hz = h[3 * i + 2];
    int fx = (int)hx, fy = (int)hy, fz = (int)hz;
    if (hx < 0 && hx != (float)(int)hx) --fx;
    if (hy < 0 && hy != (float)(int)hy) --fy;
    if (hz < 0 && hz != (float)(int)hz) --fz;
    if (hx > -1 && hx < sx && hy > -1 && hy < sy && hz > -1 && hz < sz) {
      float rx = hx - fx, ry = hy - fy, rz = hz - fz;
      if (g_mut == 9) rx = 1.0f - rx;  // self-test: swapped x weights
      float w000 = (float)((1.0 - rx) * (1.0 - ry) * (1.0 - rz));
      float w001 = (float)((1.0 - rx) * (1.0 - ry) * (rz));
      float w010 = (float)((1.0 - rx) * (ry) * (1.0 - rz));
      float w011 = (float)((1.0 - rx) * (ry) * (rz));
      float w100 = (float)((rx) * (1.0 - ry) * (1.0 - rz));
      float w101 = (float)((rx) * (1.0 - ry) * (rz));
      float w110 = (float)((rx) * (ry) * (1.0 - rz));
      float w111 = (float)((rx) * (ry) * (rz));
      bool fxi = fx >= 0, cxi = fx < sx - 1, fyi = fy >= 0, cyi = fy < sy - 1, fzi = fz >= 0,
           czi = fz < sz - 1;
      if (fxi && fyi && fzi) { out[O(fx, fy, fz)] += w000 * ix; count[O(fx, fy, fz)] += w000; }
      if (fxi && fyi && czi) { out[O(fx, fy, fz + 1)] += w001 * ix; count[O(fx, fy, fz + 1)] += w001; }
      if (fxi && cyi && fzi) { out[O(fx, fy + 1, fz)] += w010 * ix; count[O(fx, fy + 1, fz)] += w010; }
      if (fxi && cyi && czi) { out[O(fx, fy + 1, fz + 1)] += w011 * ix; count[O(fx, fy + 1, fz + 1)] += w011; }
      if (cxi && fyi && fzi) { out[O(fx + 1, fy, fz)] += w100 * ix; count[O(fx + 1, fy, fz)] += w100; }
      if (cxi && fyi && czi) { out[O(fx + 1, fy, fz + 1)] += w101 * ix; count[O(fx + 1, fy, fz + 1)] += w101; }
      if (cxi && cyi && fzi) { out[O(fx + 1, fy + 1, fz)] += w110 * ix; count[O(fx + 1, fy + 1, fz)] += w110; }
      if (cxi && cyi && czi) { out[O(fx + 1, fy + 1, fz + 1)] += w111 * ix; count[O(fx + 1, fy + 1, fz + 1)] += w111; }
    }
  }
  if (normalize)
    for (size_t e = 0; e < N; ++e) out[e] = count[e] > 0 ? out[e] / count[e] : (float)background;
}

void orc_splat_d(const float *src, const float *px, const float *py, const float *pz,
                 const int *size, double *out) {
  // cpuSplat3D (cudaSplatAtomic.cu:602-663) with ADJUST_INT (cudaSplatAtomic.cu:27)
  const int w = size[0], hh = size[1], l = size[2];
  const size_t N = (size_t)w * hh * l;
  for (size_t i = 0; i < N; ++i) out[i] = 0.0;
  auto inside = [&](int x, int y, int z) { return x >= 0 && x < w && y >= 0 && y < hh && z >= 0 && z < l; };
  for (size_t i = 0; i < N; ++i) {
    float x = px[i], y = py[i], z = pz[i];
    int xi = (int)x, yi = (int)y, zi = (int)z;
    if (x < 0 && x != (float)xi) --xi;
    if (y < 0 && y != (float)yi) --yi;
    if (z < 0 && z != (float)zi) --zi;
    float dx = 1.f - (x - xi), dy = 1.f - (y - yi), dz = 1.f - (z - zi);
    if (g_mut == 9) dx = 1.f - dx;
    double m = src[i];
    const double wx[2] = {dx, 1.0 - (double)dx}, wy[2] = {dy, 1.0 - (double)dy},
                 wz[2] = {dz, 1.0 - (double)dz};
    for (int c = 0; c < 8; ++c) {
      int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
      if (inside(xi + ox, yi + oy, zi + oz))
        out[((size_t)(zi + oz) * hh + (yi + oy)) * w + (xi + ox)] += m * wx[ox] * wy[oy] * wz[oz];
    }
  }
}

void orc_splat_index(const float *src, const float *vx, const float *vy, const float *vz,
                     const int *size, int floor_fix, unsigned int *keys, float *vals) {
  // createIndex_kernel (cudaSplat.cu:31-117).  Flattened order: key[4*id + c]
  // for the z-floor corners c = 0..3, key[4*(N + id) + c] for the z-ceil corners;
  // corner c: (x + (c&1), y + (c>>1)).  Sentinel = N.
  const int w = size[0], h = size[1], l = size[2];
  const unsigned int N = (unsigned int)((size_t)w * h * l);
  auto inside = [&](int x, int y, int z) { return x >= 0 && x < w && y >= 0 && y < h && z >= 0 && z < l; };
  unsigned int id = 0;
  for (int k = 0; k < l; ++k)
    for (int j = 0; j < h; ++j)
      for (int i = 0; i < w; ++i, ++id) {
        float mass = src[id];
        float x = i + vx[id], y = j + vy[id], z = k + vz[id];
        int xi = (int)x, yi = (int)y, zi = (int)z;
        if (floor_fix) {
          if (x < 0 && x != (float)xi) --xi;
          if (y < 0 && y != (float)yi) --yi;
          if (z < 0 && z != (float)zi) --zi;
        }
        float dx = 1.f - (x - xi), dy = 1.f - (y - yi), dz = 1.f - (z - zi);
        for (int half = 0; half < 2; ++half) {
          size_t base = 4 * ((size_t)half * N + id);
          float wz = half ? (1.f - dz) : dz;
          for (int c = 0; c < 4; ++c) {
            int ox = c & 1, oy = c >> 1;
            float wx = ox ? (1.f - dx) : dx, wy = oy ? (1.f - dy) : dy;
            if (inside(xi + ox, yi + oy, zi + half)) {
              keys[base + c] = (unsigned int)(((size_t)(zi + half) * h + (yi + oy)) * w + (xi + ox));
              vals[base + c] = mass * wx * wy * wz;
            } else {
              keys[base + c] = N;
              vals[base + c] = 0.f;
            }
          }
        }
      }
}

void orc_splat_sorted(const float *src, const float *vx, const float *vy, const float *vz,
                      const int *size, int floor_fix, float *dst, unsigned int *sorted_keys,
                      float *sorted_vals) {
  const size_t N = (size_t)size[0] * size[1] * size[2];
  std::vector<unsigned int> keys(8 * N);
  std::vector<float> vals(8 * N);
  orc_splat_index(src, vx, vy, vz, size, floor_fix, keys.data(), vals.data());
  std::vector<unsigned int> order(8 * N);
  for (size_t i = 0; i < 8 * N; ++i) order[i] = (unsigned int)i;
  // stable sort by key: equal keys keep the flattened createIndex order
  std::stable_sort(order.begin(), order.end(),
                   [&](unsigned int a, unsigned int b) { return keys[a] < keys[b]; });
  for (size_t i = 0; i < N; ++i) dst[i] = 0.f;
  size_t i = 0;
  while (i < 8 * N) {
    const unsigned int k = keys[order[i]];
    float acc = 0.f;
    size_t j = i;
    for (; j < 8 * N && keys[order[j]] == k; ++j) acc += vals[order[j]];
    if (k < N) dst[k] = acc;  // the sentinel segment (k == N) is dropped
    i = j;
  }
  if (sorted_keys)
    for (size_t t = 0; t < 8 * N; ++t) sorted_keys[t] = keys[order[t]];
  if (sorted_vals)
    for (size_t t = 0; t < 8 * N; ++t) sorted_vals[t] = vals[order[t]];
}

void orc_gradient_d(const float *img, const int *size, const double *sp, int wrap, double *grad) {
  gradient_t<double>(img, size, sp, wrap, grad);
}

void orc_gradient_f(const float *img, const int *size, const double *sp, int wrap, float *grad) {
  gradient_t<float>(img, size, sp, wrap, grad);
}

int orc_greedy_run_f(const orc_greedy_params *p, const float *I0, const float *I1, float *h,
                     float *hinv, double *energy_f, double *energy_d, double *step_out) {
  return greedy_run<float>(p, I0, I1, h, hinv, energy_f, energy_d, step_out);
}

int orc_greedy_run_d(const orc_greedy_params *p, const float *I0, const float *I1, double *h,
                     double *hinv, double *energy_d, double *step_out) {
  return greedy_run<double>(p, I0, I1, h, hinv, nullptr, energy_d, step_out);
}

void orc_gaussian_filter(const float *in, const int *size, const double *sigma, const int *ksize,
                         float *out) {
  // GaussianFilter3D::filter_symmetric_full: x, then y, then z; kernel half sizes
  // clipped to size/2 - 1 (GaussianFilter3D.cxx:178-184)
  const size_t N = (size_t)size[0] * size[1] * size[2];
  std::vector<float> a(in, in + N), b(N, 0.f);
  for (int ax = 0; ax < 3; ++ax) {
    // below 2 voxels (where the reference's unsigned arithmetic skips the clip) no filtering
    const int kmax = std::max(size[ax] / 2 - 1, 0);
    const int k = std::min(std::max(ksize[ax], 0), kmax);
    gauss_axis(a.data(), b.data(), size, ax, (float)sigma[ax], k);
    std::swap(a, b);
  }
  std::memcpy(out, a.data(), N * sizeof(float));
}

void orc_gaussian_downsample(const float *in, const int *size, const int *new_size, float *out) {
  // ImageUtils::gaussianDownsample (ImageUtils.txx:255-290): sigma = sqrt(f/2),
  // kernel 2 ceil(sigma), then resampleNew (ImageUtils.txx:320-347) to the new
  // grid: source index = f x + (f - 1)/2, trilinear, background 0.
  double f[3], sg[3];
  int k[3];
  for (int i = 0; i < 3; ++i) {
    f[i] = (double)size[i] / new_size[i];
    sg[i] = std::sqrt(f[i] / 2.0);
    k[i] = 2 * (int)std::ceil(sg[i]);
  }
  const size_t N = (size_t)size[0] * size[1] * size[2];
  std::vector<float> g(N);
  orc_gaussian_filter(in, size, sg, k, g.data());
  size_t o = 0;
  for (int z = 0; z < new_size[2]; ++z)
    for (int y = 0; y < new_size[1]; ++y)
      for (int x = 0; x < new_size[0]; ++x, ++o)
        out[o] = (float)trilerp_scalar(g.data(), size, f[0] * x + (f[0] - 1) / 2,
                                       f[1] * y + (f[1] - 1) / 2, f[2] * z + (f[2] - 1) / 2,
                                       ORC_BG_VAL, 0.0);
}

void orc_resample_hfield_d(const float *in, const int *in_size, const int *out_size, int bg,
                           int rescale, double *out) {
  // HField3DUtils::resampleNew (HField3DUtils.txx:1814-1885): output index x samples
  // the input at (r - 1)/2 + x r with r = in/out; vectors optionally scaled by 1/r
  double r[3], o[3];
  for (int i = 0; i < 3; ++i) {
    r[i] = (double)in_size[i] / out_size[i];
    o[i] = (r[i] - 1.0) / 2.0;
  }
  size_t q = 0;
  for (int z = 0; z < out_size[2]; ++z)
    for (int y = 0; y < out_size[1]; ++y)
      for (int x = 0; x < out_size[0]; ++x, ++q) {
        double v[3];
        trilerp_vec<float>(in, in_size, (float)(o[0] + x * r[0]), (float)(o[1] + y * r[1]),
                           (float)(o[2] + z * r[2]), v, bg);
        for (int c = 0; c < 3; ++c) out[3 * q + c] = rescale ? v[c] * (1.0 / r[c]) : v[c];
      }
}

int orc_atlas_run_d(const orc_atlas_params *p, const float *const *I0, const int *isizes,
                    const double *weights, double *const *h, double *const *hinv, double *mean,
                    double *energy) {
  return atlas_run_d(p, I0, isizes, weights, h, hinv, mean, energy);
}

int orc_ldmm_run_d(const orc_ldmm_params *p, const float *I0, const float *I1, double *v,
                   double *e_img, double *e_vec, double *step) {
  return ldmm_run<double>(p, I0, I1, v, e_img, e_vec, step);
}

int orc_ldmm_run_orig_d(const orc_ldmm_params *p, const float *I0, const float *I1, double *v,
                        double *e_img, double *e_vec, double *step) {
  return ldmm_run_orig<double>(p, I0, I1, v, e_img, e_vec, step);
}

void orc_ldmm_orig_terms_d(const orc_ldmm_params *p, const float *I1, const double *v, double *JT,
                           double *JD) {
  ldmm_orig_backward<double>(p, I1, v, JT, JD);
}

void orc_ldmm_alphas_orig_d(const orc_ldmm_params *p, const float *I0, const float *I1,
                            const double *v, float *alpha) {
  const size_t n = (size_t)p->n_time_steps * 3 * p->size[0] * p->size[1] * p->size[2];
  std::vector<double> vc(v, v + n);
  double step = 0.0;
  ldmm_iterate_orig<double>(p, I0, I1, vc.data(), false, step, nullptr, nullptr, alpha);
}

int orc_ldmm_atlas_run_d(const orc_ldmm_params *p, int n_subjects, const float *images,
                         const double *weights, int jacobian_scale, int use_orig, double *v,
                         double *steps, double *mean, double *energy) {
  return ldmm_atlas_run(p, n_subjects, images, weights, jacobian_scale, use_orig, v, steps, mean,
                        energy);
}

void orc_jacobian_det_d(const double *hw, const int *size, const double *sp, double *det) {
  jacobian_det<double>(hw, size, sp, det);
}

int orc_ldmm_run_f(const orc_ldmm_params *p, const float *I0, const float *I1, float *v,
                   double *e_img, double *e_vec, double *step) {
  return ldmm_run<float>(p, I0, I1, v, e_img, e_vec, step);
}

void orc_ldmm_alphas_d(const orc_ldmm_params *p, const float *I0, const float *I1,
                       const double *v, float *alpha) {
  std::vector<double> vc(v, v + (size_t)3 * p->n_time_steps * p->size[0] * p->size[1] * p->size[2]);
  double step = 0;
  ldmm_iterate<double>(p, I0, I1, vc.data(), false, step, nullptr, nullptr, alpha);
}

static void greedy_time_pair(const int *size, int n_iter, int seed, double *seconds);

double orc_greedy_time_f(const int *size, int n_iter, int seed) {
  double s = 0.0;
  greedy_time_pair(size, n_iter, seed, &s);
  return s / n_iter;
}

// Subject-parallel rate, as GreedyAtlasThreadCPU runs one thread per image
// (GreedyAtlasThreadCPU.cxx:19-37): n_threads registrations of their own pairs,
// n_iter FP32 iterations each, all threads started together.  Returns the wall
// seconds from the first start to the last finish.
double orc_greedy_time_threads_f(const int *size, int n_iter, int seed, int n_threads) {
  std::vector<std::thread> th;
  std::vector<double> secs(n_threads, 0.0);
  auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < n_threads; ++k)
    th.emplace_back(greedy_time_pair, size, n_iter, seed + k, &secs[k]);
  for (auto &t : th) t.join();
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

static void greedy_time_pair(const int *size, int n_iter, int seed, double *seconds) {
  // Synthetic blob pair, identity start; times n_iter FP32 iterations (1 thread).
  orc_greedy_params p;
  for (int i = 0; i < 3; ++i) { p.size[i] = size[i]; p.sp[i] = 1.0; }
  p.alpha = p.beta = 0.01; p.gamma = 0.001; p.max_pert = 0.1;
  p.n_iter = n_iter; p.compute_inverse = 0; p.auto_step_reduce = 0;
  const size_t N = (size_t)size[0] * size[1] * size[2];
  std::vector<float> I0(N), I1(N), h(3 * N);
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  float cx = size[0] * 0.5f, cy = size[1] * 0.5f, cz = size[2] * 0.5f;
  float s0 = size[0] * 0.15f;
  size_t i = 0;
  for (int z = 0; z < size[2]; ++z)
    for (int y = 0; y < size[1]; ++y)
      for (int x = 0; x < size[0]; ++x, ++i) {
        float d0 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
        float d1 = (x - cx - 2) * (x - cx - 2) + (y - cy) * (y - cy) + (z - cz + 1) * (z - cz + 1);
        I0[i] = std::exp(-d0 / (2 * s0 * s0));
        I1[i] = std::exp(-d1 / (2 * s0 * s0)) + 0.001f * U(rng);
        h[3 * i] = (float)x; h[3 * i + 1] = (float)y; h[3 * i + 2] = (float)z;
      }
  auto t0 = std::chrono::steady_clock::now();
  greedy_run<float>(&p, I0.data(), I1.data(), h.data(), nullptr, nullptr, nullptr, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  *seconds = std::chrono::duration<double>(t1 - t0).count();
}

int orc_fluidwarp_asym_f(const float *fixed, const float *moving, const int *size, double alpha,
                         double beta, double gamma, double max_pert, int n_iter, float *h,
                         float *hinv) {
  // FluidWarp::_computeHFieldAsymmetric (FluidWarp.cxx:925-1175), no Jacobian
  // scaling; L^-1 via the eigen table path (FluidWarp.cxx:973).
  const int sx = size[0], sy = size[1], sz = size[2];
  const size_t N = (size_t)sx * sy * sz;
  const double sp[3] = {1, 1, 1};
  std::vector<float> def(N), vf(3 * N), hinc(3 * N), old(3 * N), tmp(3 * N);
  float delta = 0;
  bool have_delta = false;
  for (int it = 0; it < n_iter; ++it) {
    apply_t<float>(moving, h, size, def.data(), ORC_BG_VAL, 0.0);
    gradient_t<float>(def.data(), size, sp, 0, vf.data());
    // _generateBodyForce (FluidWarp.cxx:2076-2098): di in double
    for (size_t i = 0; i < N; ++i) {
      double di = (double)fixed[i] - (double)def[i];
      for (int c = 0; c < 3; ++c) vf[3 * i + c] = (float)(vf[3 * i + c] * di);
    }
    diffoper<float>(size, sp, alpha, beta, gamma, 1.0, ORC_DO_EIGEN, vf.data(), true);
    if (!have_delta) {
      // _computeDelta (FluidWarp.cxx:2126-2155)
      double mx = 0;
      for (size_t i = 0; i < N; ++i) {
        double l = (double)vf[3 * i] * vf[3 * i] + (double)vf[3 * i + 1] * vf[3 * i + 1] +
                   (double)vf[3 * i + 2] * vf[3 * i + 2];
        if (mx < l) mx = l;
      }
      delta = (float)(max_pert / std::sqrt(mx));
      have_delta = true;
      if (delta > 100) break;
    }
    // _updateHField(h, hinv, v, delta) (FluidWarp.cxx:2236-2275)
    size_t i = 0;
    for (int z = 0; z < sz; ++z)
      for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x, ++i) {
          hinc[3 * i] = x + vf[3 * i] * delta;
          hinc[3 * i + 1] = y + vf[3 * i + 1] * delta;
          hinc[3 * i + 2] = z + vf[3 * i + 2] * delta;
        }
    std::memcpy(old.data(), h, sizeof(float) * 3 * N);
    for (size_t j = 0; j < N; ++j) {
      double o[3];
      trilerp_vec<float>(old.data(), size, hinc[3 * j], hinc[3 * j + 1], hinc[3 * j + 2], o,
                         ORC_BG_PARTIAL_ID);
      for (int c = 0; c < 3; ++c) h[3 * j + c] = (float)o[c];
    }
    if (hinv) {
      i = 0;
      for (int z = 0; z < sz; ++z)
        for (int y = 0; y < sy; ++y)
          for (int x = 0; x < sx; ++x, ++i) {
            // computeInverseZerothOrder (HField3DUtils.txx:585-604)
            old[3 * i] = x + x - hinc[3 * i];
            old[3 * i + 1] = y + y - hinc[3 * i + 1];
            old[3 * i + 2] = z + z - hinc[3 * i + 2];
          }
      for (size_t j = 0; j < N; ++j) {
        double o[3];
        trilerp_vec<float>(old.data(), size, hinv[3 * j], hinv[3 * j + 1], hinv[3 * j + 2], o,
                           ORC_BG_PARTIAL_ID);
        for (int c = 0; c < 3; ++c) tmp[3 * j + c] = (float)o[c];
      }
      std::memcpy(hinv, tmp.data(), sizeof(float) * 3 * N);
    }
  }
  return 0;
}

}  // extern "C"
