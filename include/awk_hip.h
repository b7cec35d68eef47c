/*
 * awk_hip.h -- C-ABI of the MI355X greedy-fluid / LDMM hot path (libawk_hip.so).
 *
 * Plain C: pointers, sizes, status codes.  No HIP or torch types cross this
 * boundary; streams are passed as opaque `awk_stream_t` (a hipStream_t, NULL =
 * the null stream).  All device pointers are float32 device memory owned by the
 * caller.  Every call is asynchronous on its stream unless it returns a host
 * scalar (documented per call).  Errors never throw: a non-zero awk_status is
 * returned and awk_last_error() gives the message (thread-local).
 *
 * Layouts (device):
 *   image         float[nz][ny][nx], x fastest (reference Array3D order,
 *                 DataTypes/Array3D.txx:426-447).
 *   vector field  three planes (x, y, z), each float[nz][ny][nx] (SoA, the
 *                 reference cplVector3DArray, CUDACommon/src/cudaVector3DArray.cu:46-57).
 *   hfield        vector field of voxel-index coordinates (h(x) = x + u(x)).
 *
 * Reference citations are relative to the AtlasWerks source root.
 */
#ifndef AWK_HIP_H
#define AWK_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int awk_status;
enum {
  AWK_OK = 0,
  AWK_ERR_INVALID = 1,      /* bad argument / size / null pointer */
  AWK_ERR_HIP = 2,          /* HIP runtime error */
  AWK_ERR_ALLOC = 3,        /* device allocation failed */
  AWK_ERR_UNSUPPORTED = 4,  /* parameter combination not implemented */
  AWK_ERR_NUMERIC = 5       /* a NaN energy or an invalid step size was detected */
};

/* Numerical events of an iterator or atlas: a sticky device word that the
 * energy and step-size kernels set, read (and cleared) by awk_*_get_flags.
 * Those calls wait for `stream` and return AWK_ERR_NUMERIC when a NaN energy or
 * an invalid step is among the flags (AWK_OK otherwise, the flags still given);
 * an energy increase alone is informational. */
enum {
  AWK_FLAG_NAN_ENERGY = 1,      /* an energy (or its change) was NaN
                                   (GreedyDeformationData.cxx:169-174; the LDMM
                                   iterator throws, LDMMIteratorGPU.cxx:283-285, 439-441) */
  AWK_FLAG_ENERGY_INCREASE = 2, /* an energy rose over the previous iteration's
                                   (EnergyIncreaseEvent, GreedyDeformationData.cxx:175-183) */
  AWK_FLAG_BAD_STEP = 4         /* max|v| was NaN, infinite or zero where a step size was
                                   computed (LDMMIteratorGPU.cxx:238-241) */
};

typedef void *awk_stream_t;

/* Background strategies: HField3DUtils::VectorBackgroundStrategy
 * (UtilitiesDataTypes/HField3DUtils.h) plus the scalar VAL of
 * Array3DUtils::ScalarBackgroundStrategyT (Array3DUtils.h:59-63). */
enum {
  AWK_BG_ID = 0,
  AWK_BG_ZERO = 1,
  AWK_BG_PARTIAL_ID = 2,
  AWK_BG_PARTIAL_ZERO = 3,
  AWK_BG_CLAMP = 4,
  AWK_BG_WRAP = 5,
  AWK_BG_VAL = 6
};

/* Launch-shape knobs chosen by the autotuner (atlaswerks_amd/autotune.py):
 * "apply_vpt", "compose_vpt" (voxels per thread: 1, 2, 4), "xinv_unroll"
 * (voxels per step of the fused inverse-x epilogue: 1, 2, 4), "xinv_pairs" (0/1),
 * "xinv_band" (row groups per band of the composing inverse x pass's block order:
 * 0 plane by plane, 2, 4, 8, 16, 32), "xfwd_band" (the same for the forward x
 * pass), "yfwd_rev" (0/1: the forward y pass visits its planes in reverse), "xinv_stage" (that pass reads the field
 * from the block's rows and their neighbours staged in LDS: 0 never, 1 always,
 * 2 rows of at most 128),
 * "splat_tile" (0 or 1: the fixed-point splat through per-block LDS windows),
 * U-Net convolutions "unet_lds", "unet_pipe", "unet_cc", the greedy atlas's
 * deformed images and mean "atlas_mean" (0: one launch per subject; 1, 2, 4:
 * subjects per group of one fused launch; read when an atlas iteration is
 * launched or captured); results do not depend on them.  FFT plan knobs read when an operator is created:
 * "fft_order", "x_rt", "y_tx", "z_tx" (another butterfly order or tile, results
 * within FP32 rounding).  Process-wide.  No counterpart in the reference. */
awk_status awk_tune_set(const char *name, int value);
awk_status awk_tune_get(const char *name, int *value);

/* Last error message of the calling host thread ("" if none). */
const char *awk_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int awk_version(void);

/* ====================================================================== */
/* FluidKernelFFT: the Navier operator L = -alpha*Lap - beta*grad(div) + gamma */
/* ====================================================================== */

typedef struct awk_fluid_kernel *awk_fk_t;

enum {
  AWK_FK_EIGEN = 1,    /* use precomputed L^(+-LPow) tables (DiffOperParam UseEigenLUT) */
  AWK_FK_DIVFREE = 2   /* project onto divergence-free fields (DiffOperParam DivergenceFree) */
};

/* Replaces DiffOperGPU::SetSize + SetParams (Algorithms/Common/DiffOperGPU.cxx:32-125)
 * and FFTSolverPlan3D::setSize/setParams (CUDACommon/inc/cudaFFTSolver.h:99-129).
 * The operator owns its spectrum scratch and tables; any nx, ny, nz >= 1. */
awk_status awk_fluid_kernel_create(awk_fk_t *out, int nx, int ny, int nz, float spx, float spy,
                                   float spz, float alpha, float beta, float gamma, float lpow,
                                   int flags);
/* DiffOperGPU::SetAlpha/SetBeta/SetGamma/SetLPow/SetUseEigenLUT/SetDivergenceFree
 * (Algorithms/Common/DiffOperGPU.h:40-115). */
awk_status awk_fluid_kernel_set_params(awk_fk_t k, float alpha, float beta, float gamma,
                                       float lpow, int flags);
awk_status awk_fluid_kernel_destroy(awk_fk_t k);
/* Bytes of device scratch the operator holds. */
awk_status awk_fluid_kernel_scratch_bytes(awk_fk_t k, unsigned long long *bytes);

/* In place on the SoA planes x, y, z.  inverse != 0: v = L^-1 f
 * (DiffOperGPU::ApplyInverseOperator(float*,float*,float*), DiffOperGPU.cxx:175-183;
 * KernelInterfaceGPU::ApplyInverseOperator, KernelInterfaceGPU.h:36-41);
 * inverse == 0: f = L v (ApplyOperator, DiffOperGPU.cxx:165-173).
 * Matches DiffOper::ApplyInverseOperatorOnTheFly / ApplyOperatorOnTheFly
 * (DiffOper.cxx:747-905): the 1/N scaling is part of the operator. */
awk_status awk_fluid_kernel_apply(awk_fk_t k, float *x, float *y, float *z, int inverse,
                                  awk_stream_t stream);

/* ====================================================================== */
/* HField3DUtils (cudaHField3DUtils namespace, CUDACommon/inc/cudaHField3DUtils.h) */
/* ====================================================================== */

/* Vector-field layout change on the device, for host AoS fields copied as one
 * block (Vector3D<float> per voxel, the reference VectorField) to and from the
 * device SoA planes: CUDAUtilities::CopyVectorFieldToDevice / FromDevice
 * (Base/CUDAUtilities.cxx:22-72), which transpose on the host one component at
 * a time.  aos: 3n floats. */
awk_status awk_vf_aos_to_soa(const float *aos, float *x, float *y, float *z, long long n,
                             awk_stream_t stream);
awk_status awk_vf_soa_to_aos(const float *x, const float *y, const float *z, float *aos, long long n,
                             awk_stream_t stream);

/* h(x) = x (HField3DUtils::setToIdentity, HField3DUtils.txx:19-35). */
awk_status awk_hfield_set_to_identity(float *hx, float *hy, float *hz, int nx, int ny, int nz,
                                      awk_stream_t stream);
/* v -> h: h = x + v/sp (velocityToH, HField3DUtils.txx:74-90); in place. */
awk_status awk_hfield_velocity_to_h(float *x, float *y, float *z, int nx, int ny, int nz,
                                    float spx, float spy, float spz, awk_stream_t stream);
/* h -> v: v = (h - x)*sp (hToVelocity, HField3DUtils.txx:92-108); in place. */
awk_status awk_hfield_h_to_velocity(float *x, float *y, float *z, int nx, int ny, int nz,
                                    float spx, float spy, float spz, awk_stream_t stream);
/* hinv(x) = 2x - h(x) (computeInverseZerothOrder, HField3DUtils.txx:585-604). */
awk_status awk_hfield_inverse_zeroth_order(float *ox, float *oy, float *oz, const float *hx,
                                           const float *hy, const float *hz, int nx, int ny,
                                           int nz, awk_stream_t stream);

/* def(x) = image(h(x)): HField3DUtils::apply (HField3DUtils.txx:728-752),
 * cudaHField3DUtils::apply (CUDACommon/src/cudaHField3DUtils.cu:580-640).
 * bg: AWK_BG_VAL (with `background`), AWK_BG_ZERO, AWK_BG_CLAMP, AWK_BG_WRAP. */
awk_status awk_hfield_apply(float *out, const float *img, const float *hx, const float *hy,
                            const float *hz, int nx, int ny, int nz, int bg, float background,
                            awk_stream_t stream);

/* The same with an image on its own nx_img*ny_img*nz_img grid: the output and
 * the hfield have nx*ny*nz voxels and h holds coordinates in the image's voxels.
 * HField3DUtils::apply takes the output size from defImage and interpolates the
 * image on its own size (HField3DUtils.txx:728-752). */
awk_status awk_hfield_apply_grid(float *out, const float *img, int nx_img, int ny_img, int nz_img,
                                 const float *hx, const float *hy, const float *hz, int nx, int ny,
                                 int nz, int bg, float background, awk_stream_t stream);

/* h(x) = M [x y z 1]^T with M a row-major 3x4 voxel-to-voxel affine map:
 * HField3DUtils::initializeFromAffine (HField3DUtils.txx:249-287), with the
 * world <-> voxel conversions of both grids folded into M by the caller, so the
 * target grid may differ from the hfield's own. */
awk_status awk_hfield_init_affine(float *hx, float *hy, float *hz, int nx, int ny, int nz,
                                  const float m[12], awk_stream_t stream);

/* h(x) = f(g(x)): HField3DUtils::compose (HField3DUtils.txx:290-310),
 * cudaHField3DUtils::compose (cudaHField3DUtils.cu:1163-1226).  out may not
 * alias f. */
awk_status awk_hfield_compose(float *ox, float *oy, float *oz, const float *fx, const float *fy,
                              const float *fz, const float *gx, const float *gy, const float *gz,
                              int nx, int ny, int nz, int bg, awk_stream_t stream);

/* h(x) = g(x + s*delta*v(x)/sp), s = +1 (composeHV, HField3DUtils.txx:432-455) or
 * -1 if inverse (composeHVInv, HField3DUtils.txx:457-480); the spacing variant of
 * cudaHField3DUtilsSp.cu:454-470.  `delta` scales v (1 for the plain call; the
 * greedy update passes its step size).  out may not alias g. */
awk_status awk_hfield_compose_hv(float *ox, float *oy, float *oz, const float *gx,
                                 const float *gy, const float *gz, const float *vx,
                                 const float *vy, const float *vz, int nx, int ny, int nz,
                                 float spx, float spy, float spz, float delta, int inverse,
                                 int bg, awk_stream_t stream);

/* h(x) = g(x) + s*v(g(x))/sp, s = +1 (composeVH) or -1 (composeVHInv)
 * (HField3DUtils.txx:370-430; cudaHField3DUtilsSp.cu:360-390).  out may not alias v. */
awk_status awk_hfield_compose_vh(float *ox, float *oy, float *oz, const float *vx,
                                 const float *vy, const float *vz, const float *gx,
                                 const float *gy, const float *gz, int nx, int ny, int nz,
                                 float spx, float spy, float spz, int inverse, int bg,
                                 awk_stream_t stream);

/* Gradient (central differences inside, one-sided at the edges, or periodic if
 * wrap): Array3DUtils::computeGradient (Array3DUtils.txx:783-901),
 * cplComputeGradient (CUDACommon/src/cudaImage3D.cu:620-656). */
awk_status awk_gradient(float *gx, float *gy, float *gz, const float *img, int nx, int ny,
                        int nz, float spx, float spy, float spz, int wrap,
                        awk_stream_t stream);

/* ====================================================================== */
/* Splatting (forward apply)                                                */
/* ====================================================================== */

enum {
  AWK_SPLAT_FIXED64 = 0,   /* deterministic: int64 fixed point (2^-32), any order */
  AWK_SPLAT_FIXED20 = 1,   /* the reference's int32 S2p20 fixed point (cplMacro.h:17,67-91) */
  AWK_SPLAT_FLOAT = 2      /* float atomics (fast, order-dependent rounding) */
};

/* dst(y) = sum over x with h(x) near y of w(h(x), y) * src(x): positions are an
 * hfield (voxel coordinates).  cplSplat3DH (CUDACommon/inc/cudaSplat.h:127-128,
 * cudaSplatAtomic.cu:569-599) / HField3DUtils::forwardApply with normalize=false
 * (HField3DUtils.txx:1281-1413).  Corners outside the volume are dropped. */
awk_status awk_splat(float *dst, const float *src, const float *hx, const float *hy,
                     const float *hz, int nx, int ny, int nz, int mode, awk_stream_t stream);

/* awk_splat without allocations: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under hipGraph capture, serving loops).
 * FIXED64 accumulates in a file-scope device array of 2^22 int64 cells (32 MiB per
 * GPU, held until the process ends; its address is fetched once per device with
 * hipGetSymbolAddress); FIXED20 and FLOAT accumulate in dst as awk_splat does.
 * Grids above 2^22 voxels return AWK_ERR_INVALID.  Results equal awk_splat's bit
 * for bit (FIXED64, FIXED20).  Calls that overlap on one GPU share its array and
 * race; each call leaves the cells it used cleared. */
awk_status awk_splat_noalloc(float *dst, const float *src, const float *hx, const float *hy,
                             const float *hz, int nx, int ny, int nz, int mode,
                             awk_stream_t stream);
/* Sort/scan splatting, cplSplating (CUDACommon/inc/cudaSplat.h:32-34,
 * cudaSplat.cu:175-238): positions are voxel + (vx, vy, vz) (a velocity in
 * voxels).  Every voxel emits 8 (cell, weight * src) pairs in the createIndex
 * order (cudaSplat.cu:31-117; sentinel cell nx*ny*nz for corners outside the
 * volume), a stable radix sort orders them by cell, and each cell's pairs are
 * summed in sorted order.  Deterministic.  reference_truncation != 0 keeps the
 * reference's int() truncation of negative positions (cudaSplat.cu:50-52,
 * bit-exact keys); 0 floors them (the trilinear weights the other splats use).
 * keys_out / vals_out (8*nx*ny*nz each, device, may be NULL) receive the sorted
 * pairs.  At most 2^28 voxels. */
awk_status awk_splat_sorted(float *dst, const float *src, const float *vx, const float *vy,
                            const float *vz, int nx, int ny, int nz, int reference_truncation,
                            unsigned int *keys_out, float *vals_out, awk_stream_t stream);

/* ====================================================================== */
/* Greedy-fluid iteration                                                   */
/* ====================================================================== */

typedef struct awk_greedy *awk_greedy_t;

/* One registration's device state: the caller's I0, I1 (float images) and
 * Def1To0 hfield (SoA, updated in place), optional Def0To1; the iterator owns
 * its operator and scratch.  GreedyIteratorGPU (Algorithms/Greedy/GreedyIteratorGPU.cxx:24-70)
 * + GreedyDeformationDataGPU.  Parameters: GreedyIteratorParam::MaxPert
 * (GreedyIteratorParam.h:29-34), DiffOperParam alpha/beta/gamma. */
awk_status awk_greedy_create(awk_greedy_t *out, int nx, int ny, int nz, float spx, float spy,
                             float spz, float alpha, float beta, float gamma, float max_pert);
awk_status awk_greedy_destroy(awk_greedy_t g);

/* Request a new step size (MaxPert / max|v|) at the next iteration
 * (DeformationIteratorInterface::UpdateStepSizeNextIteration; GreedyWarp.txx:99). */
awk_status awk_greedy_update_step_next(awk_greedy_t g);
/* The step is a device scalar: set_step writes it in the order of `stream`, after
 * the work already queued there; get_step reads it after that work (it waits for
 * `stream`). */
awk_status awk_greedy_set_step(awk_greedy_t g, float step, awk_stream_t stream);
awk_status awk_greedy_get_step(awk_greedy_t g, float *step, awk_stream_t stream);

/* One GreedyIteratorCPU::Iterate / GreedyIteratorGPU::Iterate
 * (GreedyIteratorCPU.cxx:97-138; GreedyIteratorGPU.cxx:92-135):
 *   Idef = I0(h); f = (I1 - Idef) * grad(Idef); E = prod(sp) * sum (I1-Idef)^2;
 *   v = L^-1 f; [step = MaxPert/max|v| if requested]; h' = h(x + step*v/sp)
 *   (PARTIAL_ID); if hinv given: hinv' = hinv - step*v(hinv)/sp (PARTIAL_ZERO).
 * h_in -> h_out (SoA planes): passing distinct buffers and swapping them between
 * iterations avoids the copy the reference makes (GreedyIteratorGPU.cxx:78-80);
 * h_out == h_in is allowed (the update then goes through iterator scratch).
 * hinv_in/hinv_out may be NULL (no Def0To1).  The energy is summed in FP64 on
 * the device in a fixed order (deterministic) and, if energy_dev is non-NULL,
 * stored there (device double) without a host sync. */
awk_status awk_greedy_iterate(awk_greedy_t g, const float *I0, const float *I1,
                              const float *const h_in[3], float *const h_out[3],
                              const float *const hinv_in[3], float *const hinv_out[3],
                              double *energy_dev, awk_stream_t stream);

/* n iterations back to back (GreedyWarp::RunWarp's inner loop,
 * GreedyWarp.txx:94-111, without the host between iterations): equal bit for
 * bit to n calls of awk_greedy_iterate that alternate the two fields.
 * Iteration i reads h_a and writes h_b when i is even, the other way round when
 * it is odd, so the result is in h_b for odd n and in h_a for even n; the fields
 * must not overlap.  energies_dev (n device doubles, may be NULL) receives each
 * iteration's energy.  Once the step is known, each iteration's composition
 * kernel also samples I0 at the new field, which is the next iteration's
 * deformed image, so the apply kernel (and its read of the field) runs only in
 * the first iteration.  No inverse field (awk_greedy_iterate computes Def0To1). */
awk_status awk_greedy_iterate_n(awk_greedy_t g, const float *I0, const float *I1, float *const h_a[3],
                                float *const h_b[3], int n, double *energies_dev,
                                awk_stream_t stream);

/* Sticky numerical flags since the last clear (AWK_FLAG_*); waits for stream;
 * clear != 0 resets them in stream order.  Returns AWK_ERR_NUMERIC on a NaN
 * energy or an invalid step.  Replaces the NaN and energy-increase checks of
 * GreedyDeformationData::AddEnergy (GreedyDeformationData.cxx:162-184), which
 * the reference makes on the host after every iteration. */
awk_status awk_greedy_get_flags(awk_greedy_t g, unsigned *flags, int clear, awk_stream_t stream);
/* AutoStepReduce (GreedyParam.h:61-65, GreedyDeformationData.cxx:176-183): when
 * on, an energy increase over the previous iteration halves the step before
 * this iteration's update (done on the device). */
awk_status awk_greedy_set_auto_step_reduce(awk_greedy_t g, int on);

/* Per-kernel timing of awk_greedy_iterate / awk_greedy_iterate_n, for
 * benchmarks: while on, each stage's kernel is launched with start/stop events
 * attached to its dispatch (kernel duration only).  Stages of the steady-state
 * iteration:
 *   0 apply, 1 x pass + force + SSD partials, 2 y pass (+ energy reduction),
 *   3 z pass + solve, 4 inverse y pass, 5 inverse x pass + hfield update,
 *   6 inverse x pass + hfield update + the next iteration's deformed image
 *     (awk_greedy_iterate_n, in place of 5 and of the next iteration's 0).
 * No counterpart in the reference (its GreedyIteratorGPU has no timers). */
#define AWK_GREEDY_NSTAGES 7
awk_status awk_greedy_set_profiling(awk_greedy_t g, int on);
/* Sums the recorded stage times (ms) over all iterations since profiling was
 * switched on or last read (synchronises), and resets the record.
 * stage_launches (may be NULL) receives how many launches of each stage were
 * timed; iterations the number of iterations recorded. */
awk_status awk_greedy_get_profile(awk_greedy_t g, double *stage_ms, int *stage_launches,
                                  int *iterations);

/* ====================================================================== */
/* Scale levels                                                             */
/* ====================================================================== */

/* GaussianFilter3D (Utilities/DownsampleFilter/GaussianFilter3D.cxx:159-330):
 * separable Gaussian with half kernel sizes half_* (clipped to n/2 - 1, at most
 * 15); within half_* of an edge the one-sided kernel is renormalised. */
awk_status awk_gaussian_filter(float *out, const float *in, int nx, int ny, int nz, float sigma_x,
                               float sigma_y, float sigma_z, int half_x, int half_y, int half_z,
                               awk_stream_t stream);
/* ImageUtils::gaussianDownsample (ImageUtils.txx:255-290): sigma = sqrt(f/2),
 * half size 2 ceil(sigma) per axis (f = n / n_out), then resampleNew at the
 * centred points f x + (f - 1)/2 (trilinear, background 0). */
awk_status awk_gaussian_downsample(float *out, int nx_out, int ny_out, int nz_out, const float *in,
                                   int nx, int ny, int nz, awk_stream_t stream);
/* HField3DUtils::resampleNew (HField3DUtils.txx:1814-1885): output voxel x samples
 * the input at (r - 1)/2 + x r (r = n_in / n_out) with background strategy bg;
 * rescale != 0 scales the vectors by 1/r (MultiscaleManager::UpsampleToLevel with
 * ScaleVectorFields). */
awk_status awk_hfield_resample(float *ox, float *oy, float *oz, int nx_out, int ny_out, int nz_out,
                               const float *ix, const float *iy, const float *iz, int nx, int ny,
                               int nz, int bg, int rescale, awk_stream_t stream);
/* h += sign * M [x y z 1]^T: hToVelocity / velocityToH (HField3DUtils.txx:19-128)
 * for M = identity, generalised to an affine base map. */
awk_status awk_hfield_add_affine(float *hx, float *hy, float *hz, int nx, int ny, int nz,
                                 const float m[12], float sign, awk_stream_t stream);
/* h(x) = M [x + u(x); 1]: velocityToH with unit spacing (HField3DUtils.txx:19-57)
 * followed by an affine voxel-to-voxel map, e.g. the grid map of a subject whose
 * displacement u is given in template voxels (the warm start's initial hfield).
 * h may alias u. */
awk_status awk_hfield_affine_displace(float *hx, float *hy, float *hz, const float *ux,
                                      const float *uy, const float *uz, int nx, int ny, int nz,
                                      const float m[12], awk_stream_t stream);

/* ====================================================================== */
/* Greedy atlas                                                             */
/* ====================================================================== */

typedef struct awk_atlas *awk_atlas_t;

/* GreedyAtlasBuilder + AtlasThread + GreedyAtlasThreadGPU (Algorithms/Common/
 * AtlasThread.txx:95-181, Algorithms/Greedy/GreedyAtlasBuilder.txx:236-301,
 * GreedyAtlasThreadGPU.cxx:85-206) on one device for n_subjects subjects.  The
 * mean and every hfield live on the nx*ny*nz template grid (spacing sp*); each
 * hfield holds coordinates in its subject's own image grid, so images of
 * different sizes can share one atlas (the reference requires equal sizes,
 * WeightedImageSet.cxx:316-318). */
awk_status awk_atlas_create(awk_atlas_t *out, int nx, int ny, int nz, float spx, float spy,
                            float spz, float alpha, float beta, float gamma, float max_pert,
                            int n_subjects);
awk_status awk_atlas_destroy(awk_atlas_t a);
/* Subject s: image I0 (nx_img*ny_img*nz_img, device, kept by pointer), weight,
 * initial hfield h0 (3 template-grid planes, copied).  Resets its step size. */
awk_status awk_atlas_set_subject(awk_atlas_t a, int s, const float *I0, int nx_img, int ny_img,
                                 int nz_img, float weight, const float *const h0[3],
                                 awk_stream_t stream);
awk_status awk_atlas_get_hfield(awk_atlas_t a, int s, float *const h_out[3], awk_stream_t stream);
/* Track subject s's inverse field Def0To1 from hinv0 (device, copied; identity, or
 * the inverse of the initial map), updated every iteration by composeVHInv of the
 * step, PARTIAL_ZERO, as GreedyIteratorCPU::updateDeformation does for
 * ComputeInverseHField (GreedyIteratorCPU.cxx:82-95; the reference's GPU atlas
 * rejects inverse fields, GreedyAtlasBuilderGPU.cxx:451-458).  The subject's image
 * must be on the template grid; call after awk_atlas_set_subject (which drops a
 * tracked inverse).  Such subjects take the unfused update (v, then composeHV and
 * composeVHInv). */
awk_status awk_atlas_set_subject_inverse(awk_atlas_t a, int s, const float *const hinv0[3],
                                         awk_stream_t stream);
awk_status awk_atlas_get_hinv(awk_atlas_t a, int s, float *const out[3], awk_stream_t stream);
/* Every subject's next iteration sets its step size (iteration 0 of a scale
 * level, GreedyAtlasThreadGPU::BeginImageIteration). */
awk_status awk_atlas_update_step_next(awk_atlas_t a);
/* Numerical flags of all subjects' iterations (AWK_FLAG_*), as awk_greedy_get_flags. */
awk_status awk_atlas_get_flags(awk_atlas_t a, unsigned *flags, int clear, awk_stream_t stream);
/* Steady-state iterations are captured into hipGraphs and replayed (default on). */
awk_status awk_atlas_set_graph(awk_atlas_t a, int on);
/* Subjects per batched launch in steady-state iterations (those that set no step
 * size and track no inverse): each pass of the greedy step is one launch over up
 * to n subjects (at most as many as 8 GiB of spectra hold; 0: one launch per
 * subject and pass, on 3 concurrent stream lanes).  Default: as many as keep the
 * batch's spectra within 128 MiB.  Results are bit-identical for every n. */
awk_status awk_atlas_set_batch(awk_atlas_t a, int n);
/* n_iter atlas iterations: each subject's greedy step against the current
 * weighted mean, then the new mean (AtlasThreadMain's loop).  energy_dev (n_iter
 * device doubles, may be NULL): sum_s w_s E_s per iteration (FP64, subject
 * order).  No host synchronisation. */
awk_status awk_atlas_iterate(awk_atlas_t a, int n_iter, double *energy_dev, awk_stream_t stream);
/* The weighted mean sum_s w_s I0_s(h_s) of the current hfields (ComputeThreadMean). */
awk_status awk_atlas_mean(awk_atlas_t a, float *mean_out, awk_stream_t stream);
/* I0_s(h_s) on the template grid (DefToMean). */
awk_status awk_atlas_deformed_image(awk_atlas_t a, int s, float *out, awk_stream_t stream);

/* ====================================================================== */
/* LDMM (IterateOrig, the reference's default, and IterateNew)              */
/* ====================================================================== */

typedef struct awk_ldmm *awk_ldmm_t;

/* LDMMIteratorGPU (Algorithms/LDMM/LDMMIteratorGPU.h, LDMMIteratorGPU.cxx:26-60)
 * for one nx*ny*nz grid with n_time_steps velocity fields, the Navier operator
 * (alpha, beta, gamma), sigma and MaxPert.  The first iterate computes the step
 * size (UseAdaptiveStepSize, LDMMWarp.txx:97-103) unless awk_ldmm_set_step is
 * called first.  The gradient is IterateOrig's, as the reference's constructor
 * sets mUseOrigGradientCalc (LDMMIteratorGPU.cxx:37); awk_ldmm_set_gradient
 * switches.  At most 2^28 voxels. */
awk_status awk_ldmm_create(awk_ldmm_t *out, int nx, int ny, int nz, float spx, float spy,
                           float spz, float alpha, float beta, float gamma, int n_time_steps,
                           float sigma, float max_pert);
awk_status awk_ldmm_destroy(awk_ldmm_t l);
/* mUseOrigGradientCalc (LDMMIteratorGPU.cxx:37, :151-155): 1 = IterateOrig
 * (Jacobian-determinant gradient, the default), 0 = IterateNew (adjoint splat). */
awk_status awk_ldmm_set_gradient(awk_ldmm_t l, int use_orig);
/* IterateNew's splat of the image difference to the backward-composed positions:
 * 0 = int64 fixed point through LDS tiles (cplSplat3DH's role, cudaSplatAtomic.cu:
 * 185-247, made order-independent; the default), 1 = the sort/scan splat
 * (cplSplating, cudaSplat.cu:175-238: createIndex keys, stable radix sort,
 * sequential segment sums; positions floored as the atomic splat does).  Both are
 * deterministic.  Mode 1 allocates its scratch here (about 130 bytes per voxel). */
awk_status awk_ldmm_set_splat(awk_ldmm_t l, int mode);
/* LDMMIteratorGPU::UpdateStepSizeNextIteration: the next iterate only sets the step. */
awk_status awk_ldmm_update_step_next(awk_ldmm_t l);
/* Numerical flags (AWK_FLAG_*), as awk_greedy_get_flags: the reference's LDMM
 * iterator throws on a NaN image energy or max displacement
 * (LDMMIteratorGPU.cxx:238-241, 283-285, 439-441); here the iterate calls stay
 * asynchronous and AWK_ERR_NUMERIC comes from this call. */
awk_status awk_ldmm_get_flags(awk_ldmm_t l, unsigned *flags, int clear, awk_stream_t stream);
/* Stream-ordered like awk_greedy_set_step / awk_greedy_get_step. */
awk_status awk_ldmm_set_step(awk_ldmm_t l, float step, awk_stream_t stream);
awk_status awk_ldmm_get_step(awk_ldmm_t l, float *step, awk_stream_t stream); /* waits for stream */

/* One Iterate(computeAlphasOnly = false) (LDMMIteratorGPU.cxx:147-156).
 * v: 3 * n_time_steps device planes, v[3t + c] = component c of v_t, updated in
 * place: v_t = (1 - 2 eps) v_t + (2 eps / sigma^2) L^-1(f_t), with
 *   IterateOrig (LDMMIteratorGPU.cxx:295-441): f_t = grad(J_t) (J_t - JT_t) JacDet_t,
 *     J_t = I0 o h_t through the already updated v_0..v_{t-1};
 *   IterateNew (LDMMIteratorGPU.cxx:159-293): f_t = grad(J_{t-1}) alpha_t.
 * energy_dev (3 device doubles, may be NULL): image energy Prod(sp) sum(diff^2) /
 * sigma^2 (IterateOrig: after the update; IterateNew: before it), vector energy
 * mean_t Prod(sp) <L v_t, v_t> after the update (only when vec_energy != 0: one
 * extra operator application per time step), and their sum (LDMMEnergy.cxx:48-64).
 * FP64, fixed summation order.  No host synchronisation. */
awk_status awk_ldmm_iterate(awk_ldmm_t l, const float *I0, const float *I1, float *const v[],
                            double *energy_dev, int vec_energy, awk_stream_t stream);
/* Iterate(computeAlphasOnly = true); v is not changed.  IterateOrig:
 * alpha[t] = (J_t - JT_t) JacDet_t / sigma^2 (LDMMIteratorGPU.cxx:366-371);
 * IterateNew: alpha[t] = splat(J_{T-1} - I1, h_t) / sigma^2 (LDMMIteratorGPU.cxx:197-204). */
awk_status awk_ldmm_alphas(awk_ldmm_t l, const float *I0, const float *I1, float *const v[],
                           float *const alpha[], awk_stream_t stream);
/* Copies J_t = I0 o h_t of the last iterate, h_t composed through v_0..v_t (the
 * reference's mdJ0t[t]). */
awk_status awk_ldmm_deformed_image(awk_ldmm_t l, int t, float *out, awk_stream_t stream);

/* ====================================================================== */
/* LDMM atlas                                                               */
/* ====================================================================== */

typedef struct awk_ldmm_atlas *awk_ldmm_atlas_t;

/* LDMMAtlasBuilder (Algorithms/LDMM/LDMMAtlasBuilder.txx) with LDMMAtlasThreadGPU
 * (LDMMAtlasThreadGPU.cxx) on one nx*ny*nz grid, single scale: n_subjects LDMM
 * registrations of the mean (I0) to the subjects (I1), each with n_time_steps
 * velocity fields; jacobian_scale = LDMMAtlasParam JacobianScale (default true in
 * the reference, LDMMAtlasParam.h:43-46). */
awk_status awk_ldmm_atlas_create(awk_ldmm_atlas_t *out, int nx, int ny, int nz, float spx,
                                 float spy, float spz, float alpha, float beta, float gamma,
                                 int n_time_steps, float sigma, float max_pert, int n_subjects,
                                 int jacobian_scale);
awk_status awk_ldmm_atlas_destroy(awk_ldmm_atlas_t a);
/* Subject s: its image (device, copied), weight (WeightedImageSet::GetWeight) and
 * 3 * n_time_steps device velocity planes v[3t + c], updated in place by
 * awk_ldmm_atlas_iterate; the planes stay the caller's and must outlive the atlas
 * (or the next set_subject of s).  Synchronises the stream. */
awk_status awk_ldmm_atlas_set_subject(awk_ldmm_atlas_t a, int s, const float *image, float weight,
                                      float *const v[], awk_stream_t stream);
/* The subjects' gradient (awk_ldmm_set_gradient; IterateOrig by default) and
 * IterateNew's splat (awk_ldmm_set_splat). */
awk_status awk_ldmm_atlas_set_gradient(awk_ldmm_atlas_t a, int use_orig);
awk_status awk_ldmm_atlas_set_splat(awk_ldmm_atlas_t a, int mode);
/* Every subject computes its step size at its next iteration
 * (LDMMAtlasThreadGPU::BeginImageIteration with UseAdaptiveStepSize). */
awk_status awk_ldmm_atlas_update_step_next(awk_ldmm_atlas_t a);
/* Numerical flags of all subjects' iterations (AWK_FLAG_*), as awk_ldmm_get_flags. */
awk_status awk_ldmm_atlas_get_flags(awk_ldmm_atlas_t a, unsigned *flags, int clear,
                                    awk_stream_t stream);
awk_status awk_ldmm_atlas_set_step(awk_ldmm_atlas_t a, float step, awk_stream_t stream);
awk_status awk_ldmm_atlas_get_step(awk_ldmm_atlas_t a, int s, float *step,
                                   awk_stream_t stream); /* waits for stream */
/* n_iter iterations of AtlasThread::AtlasThreadMain's loop (AtlasThread.txx): each
 * subject's LDMM iteration against the current mean, then the new mean
 * sum_s w_s I_s(h_0) JacDet_0 / sum_s w_s JacDet_0 (ComputeThreadMean and
 * LDMMAtlasBuilder::ComputeMean, LDMMAtlasBuilder.txx:270-340).  energy_dev
 * (3 * n_iter device doubles, may be NULL): sum_s w_s (image, vector, total)
 * energy per iteration, FP64 in subject order.  No host synchronisation. */
awk_status awk_ldmm_atlas_iterate(awk_ldmm_atlas_t a, int n_iter, double *energy_dev,
                                  int vec_energy, awk_stream_t stream);
/* The mean of the current fields (the initial mean before any iteration). */
awk_status awk_ldmm_atlas_mean(awk_ldmm_atlas_t a, float *mean_out, awk_stream_t stream);
/* Subject s's image at time t, I_s(h_t) with h_t = GetDefTTo1(t) of the current
 * fields (LDMMDeformationData::GetI1AtT, the DefToMeanTime images of
 * LDMMAtlasBuilder::WriteIntermediateImages, LDMMAtlasBuilder.txx:730-750). */
awk_status awk_ldmm_atlas_deformed_image(awk_ldmm_atlas_t a, int s, int t, float *out,
                                         awk_stream_t stream);

/* ====================================================================== */
/* U-Net warm start (no reference counterpart)                              */
/* ====================================================================== */

typedef struct awk_unet *awk_unet_t;

/* Number of fp32 weights the network takes.  Blob order, per 3x3x3 convolution
 * l = 0..13 (encoder 2->16, 16->16, 16->32, 32->32, 32->64, 64->64, 64->128,
 * 128->128; decoder 192->64, 64->64, 96->32, 32->32, 48->16, 16->16, where the
 * decoder inputs are [upsampled coarser level, skip] concatenated): weights in
 * PyTorch OIDHW order, then the bias; finally the 1x1x1 head W[3][16], b[3]. */
size_t awk_unet_weight_count(void);
/* Network for an nx*ny*nz grid (each a multiple of 8, at most 2^24 voxels).
 * Weights are rounded to bf16 on upload. */
awk_status awk_unet_create(awk_unet_t *out, int nx, int ny, int nz, const float *weights,
                           size_t n_weights);
/* As awk_unet_create, with activations for up to max_batch volumes (1..64, at most
 * 2^26 voxels over the batch) for awk_unet_forward_batch. */
awk_status awk_unet_create_batch(awk_unet_t *out, int nx, int ny, int nz, int max_batch,
                                 const float *weights, size_t n_weights);
awk_status awk_unet_destroy(awk_unet_t u);
/* disp = U-Net(template, subject): bf16 activations, fp32 accumulation (MFMA),
 * LeakyReLU(0.2), 2x2x2 average pooling, nearest x2 upsampling; fp32 output. */
awk_status awk_unet_forward(awk_unet_t u, const float *templ, const float *subject,
                            float *const disp[3], awk_stream_t stream);
/* batch volumes in one pass (batch <= max_batch): disp[3b + c] receives component c
 * of pair b's displacement.  Each output equals, bit for bit, the single-pair
 * forward of its pair: the volumes are stacked as one grid whose convolution taps
 * stay inside their own volume, and the summation order of every output is chosen
 * from one volume's size. */
awk_status awk_unet_forward_batch(awk_unet_t u, int batch, const float *const templ[],
                                  const float *const subject[], float *const disp[],
                                  awk_stream_t stream);
/* Training on the device (no reference counterpart): MSE between the predicted
 * and a target displacement, backward on MFMA, Adam (beta 0.9 / 0.999, eps 1e-8)
 * on fp32 master weights initialised from `weights`; every step refreshes the
 * bf16 weights of `net`.  Gradient sums use float atomics: steps are not
 * bit-reproducible. */
typedef struct awk_unet_trainer *awk_unet_trainer_t;
awk_status awk_unet_trainer_create(awk_unet_trainer_t *out, awk_unet_t net, const float *weights,
                                   size_t n_weights);
awk_status awk_unet_trainer_destroy(awk_unet_trainer_t t);
/* One step on one (template, subject, target displacement) triple; *loss_dev
 * (device double) is incremented by the step's loss. */
awk_status awk_unet_trainer_step(awk_unet_trainer_t t, const float *templ, const float *subject,
                                 const float *const target[3], float lr, double *loss_dev,
                                 awk_stream_t stream);
/* Gradient of the loss with respect to the weight blob (host copy; synchronises),
 * without an update. */
awk_status awk_unet_trainer_gradients(awk_unet_trainer_t t, const float *templ, const float *subject,
                                      const float *const target[3], float *grad_host,
                                      double *loss_host);
/* The fp32 master weights, in blob order (synchronises). */
awk_status awk_unet_trainer_get_weights(awk_unet_trainer_t t, float *host_out, size_t n);
/* Diagnostic: 64 x 4 floats, lane l's accumulator registers of one
 * v_mfma_f32_16x16x32_bf16 on fixed small-integer operands. */
awk_status awk_unet_mfma_probe(float *out, awk_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* AWK_HIP_H */
