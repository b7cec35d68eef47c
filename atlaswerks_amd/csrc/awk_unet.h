// awk_unet.h -- shared definitions of the U-Net warm start (forward and training).
#pragma once

#include "awk_ops.h"

namespace awk {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kConvWaves = 4;
constexpr int kConvBlock = 64 * kConvWaves;
constexpr int kPosPerWave = 16;

struct ConvArgs {
  const bf16 *x;      // input [N][C]
  const bf16 *w;      // [cout][kpad], k = tap*C + ci
  const float *bias;  // [cout] or null
  const bf16 *dact;   // gradient mode: the activation whose derivative scales the output; else null
  bf16 *out;          // [N][cout]
  int C, cout, kpad, nchunks;
  unsigned c8magic;   // floor(2^32 / (C/8)) + 1 for C > 8: chunk / (C/8) as a multiply-high
  int nx, ny, nz;     // one volume's grid
  int batch;          // volumes stacked along z (positions p = ((b*nz + z)*ny + y)*nx + x)
  float slope;        // LeakyReLU slope (forward: applied; gradient mode: derivative)
  // Concatenated input read in place (LDS form only): channels [0, ca) are the
  // half-resolution tensor xa at (x/2, y/2, z/2) (nearest x2 upsampling), channels
  // [ca, C) are x with C - ca channels.  xa = null: x holds all C channels.
  const bf16 *xa;
  int ca;
};

awk_status launch_conv3(const ConvArgs &g, hipStream_t s);
// whether launch_conv3 takes the LDS-staged form for g (the only form that reads a
// concatenated input in place)
bool conv3_lds_ok(const ConvArgs &g);

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (bf16)0.f;
  return z;
}

// input channels (as stored: the first layer's 2 are padded to 8) and outputs
struct LayerSpec {
  int cin, cout;
};
constexpr LayerSpec kUnetLayers[14] = {
    {8, 16},   {16, 16},   // level 0
    {16, 32},  {32, 32},   // level 1
    {32, 64},  {64, 64},   // level 2
    {64, 128}, {128, 128}, // level 3 (bottom)
    {192, 64}, {64, 64},   // decoder level 2: [up(level 3), level 2]
    {96, 32},  {32, 32},   // decoder level 1
    {48, 16},  {16, 16},   // decoder level 0
};
// input channels the weight blob carries per layer (the first layer has 2)
constexpr int kUnetBlobCin[14] = {2, 16, 16, 32, 32, 64, 64, 128, 192, 64, 96, 32, 48, 16};
constexpr float kUnetSlope = 0.2f;

}  // namespace awk

struct awk_unet {
  int nx = 0, ny = 0, nz = 0;
  int max_batch = 1;  // activations hold this many stacked volumes
  awk::bf16 *w[14] = {};
  float *bias[14] = {};
  int kpad[14] = {};
  float *head_w = nullptr, *head_b = nullptr;
  // activations
  awk::bf16 *x0 = nullptr, *t0 = nullptr, *s0 = nullptr, *p0 = nullptr;
  awk::bf16 *t1 = nullptr, *s1 = nullptr, *p1 = nullptr;
  awk::bf16 *t2 = nullptr, *s2 = nullptr, *p2 = nullptr;
  awk::bf16 *t3 = nullptr, *s3 = nullptr;
  awk::bf16 *c2 = nullptr, *d2a = nullptr, *d2 = nullptr;
  awk::bf16 *c1 = nullptr, *d1a = nullptr, *d1 = nullptr;
  awk::bf16 *c0 = nullptr, *d0a = nullptr, *d0 = nullptr;
};

namespace awk {
// nb volumes (nb <= u->max_batch): templ[b], subject[b] -> disp[3b..3b+2].
// keep_concat: write the decoder's concatenated inputs (c2, c1, c0), which the
// weight gradients read; else the convolutions that can read them in place do.
awk_status unet_forward(awk_unet *u, int nb, const float *const *templ, const float *const *subject,
                        float *const *disp, hipStream_t s, bool keep_concat = false);
awk_status unet_pool(const bf16 *in, bf16 *out, int nx, int ny, int nz, int C, hipStream_t s);
awk_status unet_upcat(bf16 *out, const bf16 *a, int ca, const bf16 *b, int cb, int nx, int ny,
                      int nz, hipStream_t s);
}  // namespace awk
