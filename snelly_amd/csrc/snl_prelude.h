// snl_prelude.h -- definitions visible to scene code, placed before the
// translated scene exactly where the reference splices __SHADER__
// (pathtracer-fragment-shader.glsl:102-140): the numeric tolerances, the
// material ids and rand().
#pragma once
#include "glsl.h"
#include "snl_params.h"

#define DENOM_TOLERANCE 1.0e-7f
#define PDF_EPSILON 1.0e-6f
#define THROUGHPUT_EPSILON 1.0e-5f
#define RADIANCE_EPSILON 1.0e-6f
#define M_PI 3.141592653589793f
#define MAT_INVAL -1
#define MAT_DIELE 0
#define MAT_METAL 1
#define MAT_SURFA 2
#define MAT_VOLUM 3  // the pick program's medium bound (pick-fragment-shader.glsl:24)

namespace glsl {
// Hachisuka's combined multiplicative LCG evaluated with Schrage's method in
// float (pathtracer-fragment-shader.glsl:119-130).  With integer seeds every
// intermediate is an exact integer < 2^24, so the output is exact given a
// correctly rounded division.
SNL_HD float snl_rand(vec4& rnd) {
  const vec4 q(1225.0f, 1585.0f, 2457.0f, 2098.0f);
  const vec4 r(1112.0f, 367.0f, 92.0f, 265.0f);
  const vec4 a(3423.0f, 2646.0f, 1707.0f, 1999.0f);
  const vec4 m(4194287.0f, 4194277.0f, 4194191.0f, 4194167.0f);
  // the divisions are correctly rounded in every build mode (floor(rnd/q) must
  // be exact), hence the explicit round-to-nearest intrinsic
#if defined(__HIP_DEVICE_COMPILE__)
  vec4 quo(__fdiv_rn(rnd.x, q.x), __fdiv_rn(rnd.y, q.y), __fdiv_rn(rnd.z, q.z), __fdiv_rn(rnd.w, q.w));
#else
  vec4 quo = rnd / q;
#endif
  vec4 beta = floor(quo);
  vec4 p = a * (rnd - beta * q) - beta * r;
  beta = (1.0f - sign(p)) * 0.5f * m;
  rnd = p + beta;
#if defined(__HIP_DEVICE_COMPILE__)
  vec4 u(__fdiv_rn(rnd.x, m.x), __fdiv_rn(rnd.y, m.y), __fdiv_rn(rnd.z, m.z), __fdiv_rn(rnd.w, m.w));
#else
  vec4 u = rnd / m;
#endif
  return fract(dot(u, vec4(1.0f, -1.0f, 1.0f, -1.0f)));
}
}  // namespace glsl
