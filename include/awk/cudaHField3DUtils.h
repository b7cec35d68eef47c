// cudaHField3DUtils.h -- the reference's GPU hfield utility namespace
// (CUDACommon/inc/cudaHField3DUtils.h) over the C-ABI: same function names,
// argument order (output first, then inputs, then w, h, l, [spacing], background,
// stream) and defaults.
#ifndef AWK_CUDAHFIELD3DUTILS_H
#define AWK_CUDAHFIELD3DUTILS_H

#include "cplVector3DArray.h"

namespace cudaHField3DUtils {

inline void setToIdentity(float *hx, float *hy, float *hz, int w, int h, int l, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_set_to_identity(hx, hy, hz, w, h, l, s));
}
inline void setToIdentity(cplVector3DArray &d_h, const Vector3Di &n, awkStream_t s = nullptr) {
  setToIdentity(d_h.x, d_h.y, d_h.z, n.x, n.y, n.z, s);
}

// in place: h = x + v/sp ; v = (h - x) * sp
inline void velocityToH_I(cplVector3DArray &d_h, const Vector3Di &n, const Vector3Df &sp, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_velocity_to_h(d_h.x, d_h.y, d_h.z, n.x, n.y, n.z, sp.x, sp.y, sp.z, s));
}
inline void velocityToH_I(cplVector3DArray &d_h, const Vector3Di &n, awkStream_t s = nullptr) {
  velocityToH_I(d_h, n, Vector3Df(1, 1, 1), s);
}
inline void hToVelocity_I(cplVector3DArray &d_v, const Vector3Di &n, const Vector3Df &sp, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_h_to_velocity(d_v.x, d_v.y, d_v.z, n.x, n.y, n.z, sp.x, sp.y, sp.z, s));
}
inline void hToVelocity_I(cplVector3DArray &d_v, const Vector3Di &n, awkStream_t s = nullptr) {
  hToVelocity_I(d_v, n, Vector3Df(1, 1, 1), s);
}

// hinv(x) = 2x - h(x)
inline void computeInverseZerothOrder(cplVector3DArray &d_hinv, const cplVector3DArray &d_h,
                                      const Vector3Di &n, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_inverse_zeroth_order(d_hinv.x, d_hinv.y, d_hinv.z, d_h.x, d_h.y, d_h.z,
                                            n.x, n.y, n.z, s));
}

// d_o(x) = d_i(h(x)), background 0 (cudaHField3DUtils::apply)
inline void apply(float *d_o, const float *d_i, const float *hx, const float *hy, const float *hz,
                  int w, int h, int l, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_apply(d_o, d_i, hx, hy, hz, w, h, l, AWK_BG_VAL, 0.f, s));
}
inline void apply(float *d_o, const float *d_i, const cplVector3DArray &d_h, const Vector3Di &n,
                  awkStream_t s = nullptr) {
  apply(d_o, d_i, d_h.x, d_h.y, d_h.z, n.x, n.y, n.z, s);
}

// h(x) = f(g(x))
inline void compose(cplVector3DArray &d_h, const cplVector3DArray &d_f, const cplVector3DArray &d_g,
                    const Vector3Di &n, int bg = BACKGROUND_STRATEGY_ZERO, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_compose(d_h.x, d_h.y, d_h.z, d_f.x, d_f.y, d_f.z, d_g.x, d_g.y, d_g.z, n.x,
                               n.y, n.z, bg, s));
}

// h(x) = g(x + v(x)/sp) (composeHV), g(x - v(x)/sp) (composeHVInv)
inline void composeHV(cplVector3DArray &d_h, const cplVector3DArray &d_g, const cplVector3DArray &d_v,
                      const Vector3Di &n, const Vector3Df &sp, BackgroundStrategy bg = BACKGROUND_STRATEGY_ID,
                      awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_compose_hv(d_h.x, d_h.y, d_h.z, d_g.x, d_g.y, d_g.z, d_v.x, d_v.y, d_v.z,
                                  n.x, n.y, n.z, sp.x, sp.y, sp.z, 1.f, 0, bg, s));
}
inline void composeHV(cplVector3DArray &d_h, const cplVector3DArray &d_g, const cplVector3DArray &d_v,
                      const Vector3Di &n, BackgroundStrategy bg = BACKGROUND_STRATEGY_ID, awkStream_t s = nullptr) {
  composeHV(d_h, d_g, d_v, n, Vector3Df(1, 1, 1), bg, s);
}
// h(x) = g(x + delta * v(x))
inline void composeHV(cplVector3DArray &d_h, const cplVector3DArray &d_g, const cplVector3DArray &d_v,
                      float delta, const Vector3Di &n, BackgroundStrategy bg = BACKGROUND_STRATEGY_ID,
                      awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_compose_hv(d_h.x, d_h.y, d_h.z, d_g.x, d_g.y, d_g.z, d_v.x, d_v.y, d_v.z,
                                  n.x, n.y, n.z, 1.f, 1.f, 1.f, delta, 0, bg, s));
}
inline void composeHVInv(cplVector3DArray &d_h, const cplVector3DArray &d_g, const cplVector3DArray &d_v,
                         const Vector3Di &n, const Vector3Df &sp, BackgroundStrategy bg = BACKGROUND_STRATEGY_ID,
                         awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_compose_hv(d_h.x, d_h.y, d_h.z, d_g.x, d_g.y, d_g.z, d_v.x, d_v.y, d_v.z,
                                  n.x, n.y, n.z, sp.x, sp.y, sp.z, 1.f, 1, bg, s));
}

// h(x) = g(x) + v(g(x))/sp (composeVH), g(x) - v(g(x))/sp (composeVHInv)
inline void composeVH(cplVector3DArray &d_h, const cplVector3DArray &d_v, const cplVector3DArray &d_g,
                      const Vector3Di &n, const Vector3Df &sp, BackgroundStrategy bg, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_compose_vh(d_h.x, d_h.y, d_h.z, d_v.x, d_v.y, d_v.z, d_g.x, d_g.y, d_g.z,
                                  n.x, n.y, n.z, sp.x, sp.y, sp.z, 0, bg, s));
}
inline void composeVHInv(cplVector3DArray &d_h, const cplVector3DArray &d_v, const cplVector3DArray &d_g,
                         const Vector3Di &n, const Vector3Df &sp, BackgroundStrategy bg, awkStream_t s = nullptr) {
  AWK_CHECK(awk_hfield_compose_vh(d_h.x, d_h.y, d_h.z, d_v.x, d_v.y, d_v.z, d_g.x, d_g.y, d_g.z,
                                  n.x, n.y, n.z, sp.x, sp.y, sp.z, 1, bg, s));
}

}  // namespace cudaHField3DUtils

// cplComputeGradient (CUDACommon/src/cudaImage3D.cu:620-656)
inline void cplComputeGradient(cplVector3DArray &d_g, const float *d_i, int w, int h, int l,
                               float spX = 1.f, float spY = 1.f, float spZ = 1.f, bool wrap = false,
                               awkStream_t s = nullptr) {
  AWK_CHECK(awk_gradient(d_g.x, d_g.y, d_g.z, d_i, w, h, l, spX, spY, spZ, wrap ? 1 : 0, s));
}

#endif
