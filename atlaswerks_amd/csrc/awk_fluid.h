// awk_fluid.h -- host-side structures of the FluidKernelFFT operator and the
// pass launchers shared by the standalone operator and the fused greedy step.
#pragma once

#include <vector>

#include "awk_fft.h"

namespace awk {

// Per-axis host plan: radix list and the device twiddle table.
struct AxisPlan {
  int n = 0;
  std::vector<int> rad;
  Axis dev() const;
  float2 *tw = nullptr;  // device
  // Bluestein's method (axes with a prime factor above 64): m = the power-of-two
  // convolution length (0: not used), its radices, and device tables inside the
  // operator's twiddle allocation
  int m = 0;
  std::vector<int> mrad;
  const float2 *chirp = nullptr, *bhat = nullptr, *mtw = nullptr;
  // the line length the LDS stages hold (the work area of m when used)
  int stage_len() const { return m > n ? m : n; }
};

// Epilogue of the inverse x pass.
enum XStore {
  XSTORE_PLAIN = 0,    // write v to the output planes
  XSTORE_MAXNORM = 1,  // write v and a per-block max |v|
  XSTORE_COMPOSE = 2,  // h_new(x) = h(x + step * v(x) / sp) (PARTIAL_ID); v not written
  XSTORE_AXPBY = 3,    // out = (1 - 2 step) out + (2 step / sigma^2) v   (LDMM velocity update)
  XSTORE_MAXAXPY = 4   // per-block max |2 out - (2 / sigma^2) v|; out not written
};

// Prologue of the forward x pass.
enum XLoad {
  XLOAD_PLAIN = 0,  // read f from the input planes
  XLOAD_FORCE = 1,     // f = (I1 - Idef) * grad(Idef); per-block FP64 partial sums of (I1-Idef)^2
  XLOAD_GRADMUL = 2,   // f = m * grad(Idef), m = the image i1
  XLOAD_GRADMUL_FX = 3 // f = m * grad(Idef), m = acc * 2^-32 (int64 splat accumulator); acc zeroed
};

struct XFwdArgs {
  int mode;
  const float *in[3];   // PLAIN: input planes
  const float *idef;    // FORCE: deformed image
  const float *i1;      // FORCE: target image
  float sp[3];          // FORCE: spacing for the gradient
  float rsp[3], r2sp[3];  // 1 / sp and 1 / (2 sp), rounded (set by xfwd_spacing)
  double *partials;     // FORCE: one double per block (may be null)
  long long *acc;       // GRADMUL_FX: fixed-point multiplier, cleared after reading
  int band;             // row groups per band of the block order (0: plane order); set by
                        //   fk_xfwd from the xfwd_band knob
};

// The gradient's spacing and the reciprocals of its two divisors.
inline void xfwd_spacing(XFwdArgs &a, const float sp[3]) {
  for (int c = 0; c < 3; ++c) {
    a.sp[c] = sp[c];
    a.rsp[c] = 1.f / sp[c];
    a.r2sp[c] = 1.f / (2.f * sp[c]);
  }
}

// Energy reduction of the greedy step (done by block 0 of the forward y pass).
struct ReduceArgs {
  const double *partials;  // one FP64 partial per x-pass block (null: no reduction)
  int n;
  double scale;            // product of the spacings (ImageUtils::l2NormSqr)
  double *energy_out;      // may be null
  double *prev;            // previous energy
  int *have_prev;
  float *step;             // halved on an energy increase when auto_reduce
  int auto_reduce;
  unsigned *flags;         // sticky AWK_FLAG_* word (may be null)
};

struct XInvArgs {
  int mode;
  float *out[3];            // PLAIN / MAXNORM: output planes
  float *maxpart;           // MAXNORM: one float per block
  const float *h[3];        // COMPOSE: current hfield
  float *hnew[3];           // COMPOSE: updated hfield
  const float *step;        // COMPOSE / AXPBY: step size (device scalar)
  float sp[3];              // COMPOSE: spacing
  float sigma2;             // AXPBY / MAXAXPY: sigma^2
  const float *i0;          // COMPOSE: when set, also idef = I0(hnew) (the next
  float *idef;              //   iteration's deformed image; VAL background i0_bg,
  float i0_bg;              //   I0 on the field's grid; one subject per launch)
  int band;                 // COMPOSE: row groups per band of the block order (0: plane
                            //   order); set by fk_xinv from the xinv_band knob
};

// Subjects of one batched launch: blockIdx.y selects the subject, and each of its
// buffers lies at a fixed stride from subject 0's (every stride 0 and n = 1: one
// subject).  The operator's tables are shared; only the spectrum, the images,
// fields, partial sums and per-subject scalars move.
struct Batch {
  int n = 1;             // subjects in the launch (gridDim.y)
  long long spec = 0;    // float2 elements between spectra
  long long img = 0;     // floats between deformed images (forward x pass, FORCE)
  long long field = 0;   // floats between hfields (inverse x pass, COMPOSE: h and hnew)
  long long part = 0;    // doubles between the x-pass partial arrays
  int scalar = 0;        // step, energy, prev, have_prev: one per subject (stride 1)
};

struct FluidKernel {
  int nx = 0, ny = 0, nz = 0;
  int nxh = 0, nxh_pad = 0;
  float sp[3] = {1, 1, 1};
  float alpha = 0, beta = 0, gamma = 0, lpow = 1;
  int flags = 0;
  AxisPlan ax[3];
  // launch shapes
  int rt = 8;          // spatial rows per x-pass block (even)
  int tx_y = 16;       // kx columns per y-pass block
  int tx_z = 8;        // kx columns per z-pass block
  int bd_x = 256, bd_y = 256, bd_z = 256;
  // device buffers
  float2 *spec = nullptr;  // 3 planes of nz*ny*nxh_pad complex
  float *lut = nullptr;    // cwx[nxh] swx[nxh] cwy[ny] swy[ny] cwz[nz] swz[nz]
  float *eig = nullptr;    // 6 floats per frequency: symmetric L^p (eigen path)
  int eig_dir = 0;         // which power eig holds: +1 (L^lpow) or -1 (L^-lpow), 0 none
  size_t spec_plane() const { return (size_t)nz * ny * nxh_pad; }
  size_t scratch_bytes = 0;
};

awk_status fk_init(FluidKernel &k, int nx, int ny, int nz, const float sp[3]);
void fk_release(FluidKernel &k);
awk_status fk_prepare_eigen(FluidKernel &k, int dir, hipStream_t s);
// Pass launchers; b (may be null: one subject) batches subjects over blockIdx.y.
awk_status fk_xfwd(FluidKernel &k, const XFwdArgs &a, hipStream_t s, int *nblocks,
                   const Batch *b = nullptr);
awk_status fk_ypass(FluidKernel &k, bool inv, hipStream_t s, const ReduceArgs *red,
                    const Batch *b = nullptr);
// step = MaxPert / max over n per-block maxima (GreedyIteratorCPU.cxx:65-79,
// 129-134; LDMMIteratorGPU.cxx:286-289); a NaN, infinite or zero maximum raises
// AWK_FLAG_BAD_STEP in flags (may be null).  The step is stored as computed.
awk_status launch_step_from_max(const float *maxpart, int n, float max_pert, float *step,
                                unsigned *flags, hipStream_t s);
awk_status fk_zsolve(FluidKernel &k, bool inverse_op, hipStream_t s, const Batch *b = nullptr);
awk_status fk_xinv(FluidKernel &k, const XInvArgs &a, hipStream_t s, int *nblocks,
                   const Batch *b = nullptr);
int fk_xblocks(const FluidKernel &k);

}  // namespace awk

struct awk_fluid_kernel {
  awk::FluidKernel k;
};
