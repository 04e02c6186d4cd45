// snl_wavefront.h -- wavefront buffers shared by the host runtime and the
// hipRTC kernels: SoA path-field indices, queue counters and kernel arguments.
#pragma once
#include "snl_params.h"

namespace glsl {

// float path fields
enum {
  F_OX, F_OY, F_OZ, F_DX, F_DY, F_DZ, F_HX, F_HY, F_HZ,
  F_TP, F_L = F_TP + 3, F_LS = F_L + 3, F_WL = F_LS + 3, F_XYZ, F_RGB = F_XYZ + 3,
  F_MIS = F_RGB + 3,           // sky, sun, sphere
  F_RNG = F_MIS + 3,           // 4 floats
  F_NTP = F_RNG + 4,           // throughput at the NEE vertex (3)
  F_NEE = F_NTP + 3,           // per light: dir(3) tmax f(3) Li(3) den cos mis  -> 13
  F_NEE_STRIDE = 13,
  F_COUNT = F_NEE + 3 * F_NEE_STRIDE
};
// int path fields
enum { I_FLAGS, I_HMAT, I_NEE, I_COUNT };
// flags bits
#define PF_INDIELE 1
#define PF_SKIP 2
#define PF_VSHIFT 8
#define PF_ASHIFT 16

// counters
// C_TRACE_STEPS (SDF march steps) and C_NORMALS (normal evaluations, 6 SDF
// evaluations each) are 64-bit counters in 8-byte aligned slots
enum { C_TRACE_IN, C_NEE_IN, C_TRACE_OUT, C_NEE_OUT, C_MAT, C_FETCH = C_MAT + 4, C_TRACE_STEPS = 10, C_NORMALS = 12,
       C_COUNT = 16 };

struct SnlKArgs {
  const SnlParams* P;
  const void* U;
  float* pf;          // F_COUNT * cap floats
  int* pi;            // I_COUNT * cap ints
  long long cap;
  int* qtrace_in;
  int* qnee_in;
  int* qtrace_out;
  int* qnee_out;
  int* qmat;          // 4 * cap
  int* cnt;
  float* radiance;    // 4 floats per pixel: X Y Z N
  float* rngstate;    // 4 floats per pixel
  int W, H, npix;
  int vertex, sample, spf;
  int chunk;          // trace refill chunk per wave
  int itemBase, itemEnd;  // persistent kernel: queue range of this tile
  int tileIndex, tileCount;  // persistent kernel: tile k of K takes every K-th 64-item stripe
  int shadeThresh;        // persistent kernel: lanes waiting before a wave shades
  int frames;             // persistent kernel: frames each pixel runs before the lane moves on
  int maxBounces;         // __MAX_BOUNCES__: the last vertex index a path may reach
  int imagePix;           // W * H; a state of several chains holds npix = chains * imagePix
  int burst;              // persistent kernel: march steps per lane between shading checks
  int* tileCost;          // queue kernels: march steps per 8x8 pixel tile are added here (or null)
  const int* tileOrder;   // queue position of a tile -> the tile it takes (or null: in order)
};

#define SNL_PF(k, i) a.pf[(long long)(k) * a.cap + (i)]
#define SNL_PI(k, i) a.pi[(long long)(k) * a.cap + (i)]


}  // namespace glsl
