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
// Sort/scan splat (awk_sortsplat.hip): positions = grid + (vx, vy, vz), or the
// positions themselves when absolute; truncate = the reference's int() cast;
// work = splat_sorted_scratch_bytes() of device memory or null (stream-ordered).
size_t splat_sorted_scratch_bytes(int nx, int ny, int nz);
awk_status launch_splat_sorted(float *dst, const float *src, const float *vx, const float *vy,
                               const float *vz, int nx, int ny, int nz, int truncate,
                               unsigned *keys_out, float *vals_out, hipStream_t s, int absolute,
                               void *work);
awk_status launch_splat(float *dst, const float *src, const float *hx, const float *hy,
                        const float *hz, int nx, int ny, int nz, int mode, hipStream_t s);

}  // namespace awk
