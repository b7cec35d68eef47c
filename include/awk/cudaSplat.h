// cudaSplat.h -- splatting with the reference's entry names (CUDACommon/inc/cudaSplat.h:
// cplSplat3DH at :127-128).  The default accumulation is the reference's int32
// fixed point S2p20 (cudaSplatAtomic.cu:185-247), so drop-in callers get its
// results; AWK_SPLAT_FIXED64 is the deterministic 64-bit fixed point with 2^-32
// resolution and no 2048 range limit.
#ifndef AWK_CUDASPLAT_H
#define AWK_CUDASPLAT_H

#include "cplVector3DArray.h"

// d_dst(h(x)) += d_src(x), positions = hfield d_h (voxel coordinates)
inline void cplSplat3DH(float *d_dst, float *d_src, cplVector3DArray &d_h, const Vector3Di &n,
                        awkStream_t s = nullptr, int mode = AWK_SPLAT_FIXED20) {
  AWK_CHECK(awk_splat(d_dst, d_src, d_h.x, d_h.y, d_h.z, n.x, n.y, n.z, mode, s));
}
inline void cplSplat3DH(float *d_dst, float *d_src, float *d_px, float *d_py, float *d_pz,
                        unsigned int sizeX, unsigned int sizeY, unsigned int sizeZ,
                        awkStream_t s = nullptr, int mode = AWK_SPLAT_FIXED20) {
  AWK_CHECK(awk_splat(d_dst, d_src, d_px, d_py, d_pz, (int)sizeX, (int)sizeY, (int)sizeZ, mode, s));
}

// cplSplating (CUDACommon/inc/cudaSplat.h:32-34): sort/scan splat to voxel + v,
// keeping the reference's truncation of negative positions.
inline void cplSplating(float *d_dst, float *d_src, float *d_vx, float *d_vy, float *d_vz,
                        unsigned int w, unsigned int h, unsigned int l, awkStream_t s = nullptr) {
  AWK_CHECK(awk_splat_sorted(d_dst, d_src, d_vx, d_vy, d_vz, (int)w, (int)h, (int)l, 1, nullptr,
                             nullptr, s));
}

#endif
