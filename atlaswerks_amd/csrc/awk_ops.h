// awk_ops.h -- host-side launchers shared between translation units.
#pragma once

#include "awk_common.h"

namespace awk {

awk_status launch_apply(float *out, const float *img, const float *hx, const float *hy,
                        const float *hz, int nx, int ny, int nz, int bg, float background,
                        hipStream_t s);
awk_status launch_apply_grid(float *out, const float *img, int inx, int iny, int inz,
                             const float *hx, const float *hy, const float *hz, int nx, int ny,
                             int nz, int bg, float background, hipStream_t s);
awk_status launch_compose_hv(float *ox, float *oy, float *oz, const float *gx, const float *gy,
                             const float *gz, const float *vx, const float *vy, const float *vz,
                             int nx, int ny, int nz, const float sp[3], float delta,
                             const float *delta_dev, int inverse, int bg, hipStream_t s);
awk_status launch_compose_vh(float *ox, float *oy, float *oz, const float *vx, const float *vy,
                             const float *vz, const float *gx, const float *gy, const float *gz,
                             int nx, int ny, int nz, const float sp[3], const float *delta_dev,
                             int inverse, int bg, hipStream_t s);
awk_status launch_splat(float *dst, const float *src, const float *hx, const float *hy,
                        const float *hz, int nx, int ny, int nz, int mode, hipStream_t s);

}  // namespace awk
