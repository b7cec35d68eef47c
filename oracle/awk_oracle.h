/*
 * awk_oracle.h -- CPU restatement of the AtlasWerks greedy-fluid / LDMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (atlaswerks_amd/libawk_hip.so) never links it.
 *
 * Each function names the reference file:line it restates (paths relative to
 * the AtlasWerks source root).  Layouts follow the reference CPU containers:
 * images are x-fastest float arrays (Array3D<float>, DataTypes/Array3D.txx:426-447),
 * vector fields are AoS (x,y,z) triples (VectorField = Array3D<Vector3D<float>>,
 * Base/AtlasWerksTypes.h:29-32).
 *
 * Template precision: every "_f" entry point mirrors the reference's float
 * (T=float) arithmetic, every "_d" entry point runs the same algorithm in
 * double and is the FP64 oracle the GPU kernels are compared against.
 */
#ifndef AWK_ORACLE_H
#define AWK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Background strategies: numbering matches the C-ABI (include/awk_hip.h). */
enum {
  ORC_BG_ID = 0,
  ORC_BG_ZERO = 1,
  ORC_BG_PARTIAL_ID = 2,
  ORC_BG_PARTIAL_ZERO = 3,
  ORC_BG_CLAMP = 4,
  ORC_BG_WRAP = 5,
  ORC_BG_VAL = 6 /* scalar only: background value */
};

/* DiffOper flags */
enum { ORC_DO_EIGEN = 1, ORC_DO_DIVFREE = 2 };

/* Deliberate-mistake switch used only by the oracle's own self-tests to show
 * that the pinning checks catch plausible errors (0 = correct oracle). */
void orc_set_mutation(int m);

/* ---- FFT (unnormalised, forward = exp(-2 pi i jk/n)) ---- */
int orc_fft_d(int n, double *data, int inverse);
int orc_fft_f(int n, float *data, int inverse);

/* ---- Navier operator LUT (DiffOper.cxx:587-639) ---- */
void orc_lut(int n, double sp, double *cosW, double *sinW);

/* ---- DiffOper L / L^-1 on an AoS vector field (DiffOper.cxx:747-1218) ---- */
int orc_diffoper_d(const int *size, const double *sp, double alpha, double beta,
                   double gamma, double lpow, int flags, double *vf, int inverse);
int orc_diffoper_f(const int *size, const double *sp, double alpha, double beta,
                   double gamma, double lpow, int flags, float *vf, int inverse);

/* ---- interpolation (HField3DUtils.txx:1969-2300, Array3DUtils.txx:59-232) ---- */
void orc_trilerp_vec_d(const float *h, const int *size, double x, double y, double z,
                       double *out3, int bg);
double orc_trilerp_scalar_d(const float *img, const int *size, double x, double y,
                            double z, int bg, double background);

/* ---- hfield utilities (HField3DUtils.txx) ---- */
/* apply: def(x) = image(h(x)) (HField3DUtils.txx:728-752) */
void orc_apply_d(const float *img, const float *h, const int *size, double *out,
                 int bg, double background);
void orc_apply_f(const float *img, const float *h, const int *size, float *out,
                 int bg, double background);
/* compose: h(x) = f(g(x)) (HField3DUtils.txx:290-310) */
void orc_compose_d(const float *f, const float *g, const int *size, double *h, int bg);
/* composeHV: h(x) = g(x + v(x)/sp)   (HField3DUtils.txx:432-455)
 * composeHVInv: h(x) = g(x - v(x)/sp) (HField3DUtils.txx:457-480) */
void orc_compose_hv_d(const float *g, const float *v, const int *size, const double *sp,
                      double *h, int bg, int inv);
/* composeVH: h(x) = g(x) + v(g(x))/sp ; composeVHInv: g(x) - v(g(x))/sp
 * (HField3DUtils.txx:370-430) */
void orc_compose_vh_d(const float *v, const float *g, const int *size, const double *sp,
                      double *h, int bg, int inv);
/* forwardApply (splat, CPU float accumulation) (HField3DUtils.txx:1281-1413) */
void orc_forward_apply_f(const float *img, const float *h, const int *size, float *out,
                         double background, int normalize);
/* cpuSplat3D (double accumulation) (cudaSplatAtomic.cu:602-663); SoA positions */
void orc_splat_d(const float *src, const float *px, const float *py, const float *pz,
                 const int *size, double *out);
/* sort/scan splat keys (cudaSplat.cu:31-117): positions = voxel + v (SoA v).
 * keys[8N], vals[8N] in the reference's flattened uint4/float4 order; sentinel N. */
void orc_splat_index(const float *src, const float *vx, const float *vy, const float *vz,
                     const int *size, int floor_fix, unsigned int *keys, float *vals);

/* cplSplating (cudaSplat.cu:175-238) restated: createIndex keys/values, a stable
 * sort by key (std::stable_sort), then per-cell sums taken sequentially in the
 * sorted order (float).  sorted_keys/sorted_vals (8N each) may be NULL. */
void orc_splat_sorted(const float *src, const float *vx, const float *vy, const float *vz,
                      const int *size, int floor_fix, float *dst, unsigned int *sorted_keys,
                      float *sorted_vals);

/* ---- gradient / residual / energy ---- */
/* computeGradient (Array3DUtils.txx:783-901), output AoS */
void orc_gradient_d(const float *img, const int *size, const double *sp, int wrap,
                    double *grad);
void orc_gradient_f(const float *img, const int *size, const double *sp, int wrap,
                    float *grad);

/* ---- greedy loop (GreedyIteratorCPU.cxx:65-138, GreedyWarp.txx:69-125) ---- */
typedef struct {
  int size[3];
  double sp[3];
  double alpha, beta, gamma;
  double max_pert;
  int n_iter;
  int compute_inverse;   /* Def0To1 via composeVHInv, PARTIAL_ZERO */
  int auto_step_reduce;  /* GreedyParam.h:61-65 */
} orc_greedy_params;

/* Runs n_iter greedy iterations at one scale level starting from h (AoS, in/out).
 * energy_f: energies as the reference records them (float sum x spacing product);
 * energy_d: the same residual summed in double.  step_out: final step size. */
int orc_greedy_run_f(const orc_greedy_params *p, const float *I0, const float *I1,
                     float *h, float *hinv, double *energy_f, double *energy_d,
                     double *step_out);
int orc_greedy_run_d(const orc_greedy_params *p, const float *I0, const float *I1,
                     double *h, double *hinv, double *energy_d, double *step_out);

/* One FP32 iteration timed for the CPU baseline (same as orc_greedy_run_f with
 * n_iter=1 but with preallocated state; returns seconds per iteration). */
double orc_greedy_time_f(const int *size, int n_iter, int seed);
/* Wall seconds of n_threads concurrent registrations (one thread each), n_iter each. */
double orc_greedy_time_threads_f(const int *size, int n_iter, int seed, int n_threads);

/* ---- Multiscale (ImageUtils.txx:255-347, GaussianFilter3D.cxx, HField3DUtils.txx:1814-1885) ---- */
/* GaussianFilter3D with sigma[3] and half kernel sizes ksize[3] (float). */
void orc_gaussian_filter(const float *in, const int *size, const double *sigma, const int *ksize,
                         float *out);
/* ImageUtils::gaussianDownsample to new_size (VAL background 0, trilinear). */
void orc_gaussian_downsample(const float *in, const int *size, const int *new_size, float *out);
/* HField3DUtils::resampleNew of an AoS vector field (bg: vector strategy). */
void orc_resample_hfield_d(const float *in, const int *in_size, const int *out_size, int bg,
                           int rescale, double *out);

/* ---- Greedy atlas (AtlasThread.txx:95-181, GreedyAtlasBuilder.txx:236-301) ---- */
typedef struct {
  int size[3];       /* the template (mean) grid; every hfield lives on it */
  double sp[3];
  double alpha, beta, gamma;
  double max_pert;
  int n_iter;
  int n_subjects;
} orc_atlas_params;

/* One scale level of GreedyAtlasBuilder in FP64.  I0[s]: subject image on its own
 * grid isizes[3s..3s+2]; h[s]: AoS hfield on the template grid holding voxel
 * coordinates of subject s (in/out); weights w_s.  mean (N, may be NULL): the
 * final weighted mean; energy (n_iter, may be NULL): sum_s w_s E_s per iteration. */
int orc_atlas_run_d(const orc_atlas_params *p, const float *const *I0, const int *isizes,
                    const double *weights, double *const *h, double *const *hinv, double *mean,
                    double *energy);

/* ---- LDMM (LDMMIteratorCPU.cxx:175-212, 529-631; LDMMEnergy.cxx:48-64) ---- */
typedef struct {
  int size[3];
  double sp[3];
  double alpha, beta, gamma;
  double sigma;
  double max_pert;
  int n_time_steps;
  int n_iter;
} orc_ldmm_params;

/* n_iter IterateNew updates of velocity fields v (n_time_steps AoS fields, in/out).
 * *step <= 0 on entry: one extra call first computes the step size (MaxPert /
 * max over t of |2 v_t - 2/sigma^2 L^-1 f_t|) and leaves v unchanged, as
 * mUpdateStepSizeNextIter does.  e_img / e_vec (n_iter each, may be NULL): image
 * energy (before the update) and vector energy (after it).  *step: in/out. */
int orc_ldmm_run_d(const orc_ldmm_params *p, const float *I0, const float *I1, double *v,
                   double *e_img, double *e_vec, double *step);
int orc_ldmm_run_f(const orc_ldmm_params *p, const float *I0, const float *I1, float *v,
                   double *e_img, double *e_vec, double *step);
/* The same loop with LDMMIteratorCPU::IterateOrig (LDMMIteratorCPU.cxx:379-527),
 * the Jacobian-determinant gradient; the image energy is measured after the
 * update (IterateOrig's order). */
int orc_ldmm_run_orig_d(const orc_ldmm_params *p, const float *I0, const float *I1, double *v,
                        double *e_img, double *e_vec, double *step);
/* IterateOrig's backward sweep: JT_t = I1(h_t) and JacDet_t, [T][nz][ny][nx] each. */
void orc_ldmm_orig_terms_d(const orc_ldmm_params *p, const float *I1, const double *v, double *JT,
                           double *JD);
/* IterateOrig(computeAlphasOnly = true): alpha_t = (I0(h_t) - JT_t) JacDet_t / sigma^2. */
void orc_ldmm_alphas_orig_d(const orc_ldmm_params *p, const float *I0, const float *I1,
                            const double *v, float *alpha);
/* LDMM atlas (LDMMAtlasBuilder.txx:270-340, LDMMAtlasThreadGPU.cxx, AtlasThread.txx),
 * single scale: images [S][N] (the subjects, I1 of each registration), weights
 * [S], v [S][T][N][3] in/out, steps [S] in/out (<= 0: computed by a step-size call
 * at the first iteration), mean [N] out (after the last iteration), energy
 * [n_iter][3] (image, vector, total; sum_s w_s E_s).  p->n_iter iterations; each
 * subject's iteration is IterateOrig (use_orig) or IterateNew. */
int orc_ldmm_atlas_run_d(const orc_ldmm_params *p, int n_subjects, const float *images,
                         const double *weights, int jacobian_scale, int use_orig, double *v,
                         double *steps, double *mean, double *energy);
/* HField3DUtils::jacobian (HField3DUtils.txx:2370-2420) of an AoS field in world
 * units with spacing sp. */
void orc_jacobian_det_d(const double *hw, const int *size, const double *sp, double *det);
/* IterateNew(computeAlphasOnly = true): alpha_t = splat(J0t_T - I1, h_t) / sigma^2
 * for every t (n_time_steps images, float). */
void orc_ldmm_alphas_d(const orc_ldmm_params *p, const float *I0, const float *I1,
                       const double *v, float *alpha);

/* ---- FluidWarp::_computeHFieldAsymmetric (FluidWarp.cxx:925-1175) ---- */
int orc_fluidwarp_asym_f(const float *fixed, const float *moving, const int *size,
                         double alpha, double beta, double gamma, double max_pert,
                         int n_iter, float *h, float *hinv);

#ifdef __cplusplus
}
#endif
#endif
