// Accurate single-precision sin, cos, tan, exp, log, pow, atan, acos for
// GLSL ES 3.00, built from +, *, floor and bit casts only, so a GLES
// implementation computes them exactly as written.  make_reference_fixtures.py
// puts this block ahead of the reference's shader code and maps the builtins
// onto it with #defines: GLSL leaves builtin precision to the implementation,
// and the one used for fixtures (SwiftShader 4.1) has ~1e-3 relative error in
// sin/cos/tan inside large programs.  mix and normalize follow their
// definitions, and min, max and clamp theirs, for the same reason.
//
// Algorithms and coefficients: the Cephes single-precision library by
// S. L. Moshier (sinf.c, tanf.c, expf.c, logf.c, atanf.c, asinf.c): Cody-Waite
// range reduction and minimax polynomials, about 1-2 ulp on the reduced ranges.

float snl_sinp(float x) {
  float z = x * x;
  return ((-1.9515295891e-4 * z + 8.3321608736e-3) * z - 1.6666654611e-1) * z * x + x;
}
float snl_cosp(float x) {
  float z = x * x;
  return ((2.443315711809948e-5 * z - 1.388731625493765e-3) * z + 4.166664568298827e-2) * z * z - 0.5 * z + 1.0;
}
// x = j*pi/4 + r with |r| <= pi/4 (three-part pi/4 for the reduction)
float snl_reduce(float x, out int j) {
  float ax = abs(x);
  float y = floor(ax * 1.27323954473516);
  int k = int(y);
  if ((k & 1) == 1) { k += 1; y += 1.0; }
  j = k & 7;
  return ((ax - y * 0.78515625) - y * 2.4187564849853515625e-4) - y * 3.77489497744594108e-8;
}
float snl_sin(float x) {
  int j;
  float r = snl_reduce(x, j);
  float s = x < 0.0 ? -1.0 : 1.0;
  if (j > 3) { s = -s; j -= 4; }
  float v = (j == 1 || j == 2) ? snl_cosp(r) : snl_sinp(r);
  return s * v;
}
float snl_cos(float x) {
  int j;
  float r = snl_reduce(x, j);
  float s = 1.0;
  if (j > 3) { s = -s; j -= 4; }
  if (j > 1) s = -s;
  float v = (j == 1 || j == 2) ? snl_sinp(r) : snl_cosp(r);
  return s * v;
}
// Cephes tanf: tan(r) = r + r^3 P(r^2) on |r| <= pi/4, and -1/tan on odd
// octant pairs (about 1 ulp; sin/cos would round twice more)
float snl_tan(float x) {
  int j;
  float r = snl_reduce(x, j);
  float z = r * r;
  float y = (((((9.38540185543e-3 * z + 3.11992232697e-3) * z + 2.44301354525e-2) * z + 5.34112807005e-2) * z
               + 1.33387994085e-1) * z + 3.33331568548e-1) * z * r + r;
  if ((j & 2) == 2) y = -1.0 / y;
  return x < 0.0 ? -y : y;
}

float snl_pow2i(int n) { return intBitsToFloat((clamp(n, -126, 127) + 127) << 23); }
float snl_exp(float x) {
  if (x > 88.72283905206835) return intBitsToFloat(0x7f800000);
  if (x < -103.278929903431851103) return 0.0;
  float n = floor(x * 1.44269504088896341 + 0.5);
  float r = (x - n * 0.693359375) - n * -2.12194440e-4;
  float z = r * r;
  float p = (((((1.9875691500e-4 * r + 1.3981999507e-3) * r + 8.3334519073e-3) * r + 4.1665795894e-2) * r
              + 1.6666665459e-1) * r + 5.0000001201e-1) * z + r + 1.0;
  int ni = int(n);
  // 2^n in two factors keeps results down to the denormal range
  return (p * snl_pow2i(ni / 2)) * snl_pow2i(ni - ni / 2);
}
float snl_log(float x) {
  if (x < 0.0 || x != x) return intBitsToFloat(0x7fc00000);
  if (x == 0.0) return intBitsToFloat(0xff800000);
  if (isinf(x)) return x;
  int bits = floatBitsToInt(x);
  int e = ((bits >> 23) & 0xff) - 126;
  float m = intBitsToFloat((bits & 0x007fffff) | 0x3f000000);  // in [0.5, 1)
  if (m < 0.70710678118654752440) { e -= 1; m = m + m - 1.0; } else { m = m - 1.0; }
  float z = m * m;
  float y = ((((((((7.0376836292e-2 * m - 1.1514610310e-1) * m + 1.1676998740e-1) * m - 1.2420140846e-1) * m
                 + 1.4249322787e-1) * m - 1.6668057665e-1) * m + 2.0000714765e-1) * m - 2.4999993993e-1) * m
            + 3.3333331174e-1) * m * z;
  float fe = float(e);
  y += -2.12194440e-4 * fe;
  y += -0.5 * z;
  return m + y + 0.693359375 * fe;
}
float snl_pow(float x, float y) {
  if (y == 0.0) return 1.0;
  if (x == 0.0) return y > 0.0 ? 0.0 : intBitsToFloat(0x7f800000);
  return snl_exp(y * snl_log(x));
}
float snl_atan(float x) {
  float a = abs(x), y = 0.0;
  if (a > 2.414213562373095) { y = 1.5707963267948966; a = -1.0 / a; }
  else if (a > 0.4142135623730950) { y = 0.7853981633974483; a = (a - 1.0) / (a + 1.0); }
  float z = a * a;
  y += (((8.05374449538e-2 * z - 1.38776856032e-1) * z + 1.99777106478e-1) * z - 3.33329491539e-1) * z * a + a;
  return x < 0.0 ? -y : y;
}
float snl_atan(float y, float x) {
  if (x == 0.0) return y > 0.0 ? 1.5707963267948966 : (y < 0.0 ? -1.5707963267948966 : 0.0);
  float t = snl_atan(y / x);
  if (x < 0.0) t += (y < 0.0 ? -3.141592653589793 : 3.141592653589793);
  return t;
}
float snl_asinp(float a) {  // asin on [0, 0.5]
  float z = a * a;
  return ((((4.2163199048e-2 * z + 2.4181311049e-2) * z + 4.5470025998e-2) * z + 7.4953002686e-2) * z
          + 1.6666752422e-1) * z * a + a;
}
float snl_acos(float x) {
  if (x < -0.5) return 3.141592653589793 - 2.0 * snl_asinp(sqrt(0.5 * (1.0 + x)));
  if (x > 0.5) return 2.0 * snl_asinp(sqrt(0.5 * (1.0 - x)));
  return 1.5707963267948966 - (x < 0.0 ? -snl_asinp(-x) : snl_asinp(x));
}
float snl_asin(float x) {
  float a = abs(x), r;
  if (a > 0.5) r = 1.5707963267948966 - 2.0 * snl_asinp(sqrt(0.5 * (1.0 - a)));
  else r = snl_asinp(a);
  return x < 0.0 ? -r : r;
}

#define SNL_VEC1(F)                                                              \
  vec2 F(vec2 v) { return vec2(F(v.x), F(v.y)); }                                \
  vec3 F(vec3 v) { return vec3(F(v.x), F(v.y), F(v.z)); }                        \
  vec4 F(vec4 v) { return vec4(F(v.x), F(v.y), F(v.z), F(v.w)); }
SNL_VEC1(snl_sin)
SNL_VEC1(snl_cos)
SNL_VEC1(snl_tan)
SNL_VEC1(snl_exp)
SNL_VEC1(snl_log)
SNL_VEC1(snl_atan)
SNL_VEC1(snl_acos)
SNL_VEC1(snl_asin)
vec2 snl_pow(vec2 a, vec2 b) { return vec2(snl_pow(a.x, b.x), snl_pow(a.y, b.y)); }
vec3 snl_pow(vec3 a, vec3 b) { return vec3(snl_pow(a.x, b.x), snl_pow(a.y, b.y), snl_pow(a.z, b.z)); }
vec4 snl_pow(vec4 a, vec4 b) { return vec4(snl_pow(a.x, b.x), snl_pow(a.y, b.y), snl_pow(a.z, b.z), snl_pow(a.w, b.w)); }
vec2 snl_atan(vec2 a, vec2 b) { return vec2(snl_atan(a.x, b.x), snl_atan(a.y, b.y)); }
vec3 snl_atan(vec3 a, vec3 b) { return vec3(snl_atan(a.x, b.x), snl_atan(a.y, b.y), snl_atan(a.z, b.z)); }
#undef SNL_VEC1

// min, max and clamp as GLSL ES 3.00 defines them (section 8.3): min(x, y) is
// "y if y < x, otherwise x", max(x, y) "y if x < y, otherwise x", clamp(x, a, b)
// is min(max(x, a), b).  SwiftShader 4.1's builtins return the second operand
// when either is NaN (min(3.0, NaN) = NaN), the definition returns x.
#define SNL_MINMAX(T, B, S)                                                      \
  T snl_min(T x, T y) { return mix(x, y, lessThan(y, x)); }                      \
  T snl_max(T x, T y) { return mix(x, y, lessThan(x, y)); }                      \
  T snl_min(T x, S y) { return snl_min(x, T(y)); }                               \
  T snl_max(T x, S y) { return snl_max(x, T(y)); }                               \
  T snl_clamp(T x, T a, T b) { return snl_min(snl_max(x, a), b); }               \
  T snl_clamp(T x, S a, S b) { return snl_min(snl_max(x, T(a)), T(b)); }
float snl_min(float x, float y) { return y < x ? y : x; }
float snl_max(float x, float y) { return x < y ? y : x; }
float snl_clamp(float x, float a, float b) { return snl_min(snl_max(x, a), b); }
int snl_min(int x, int y) { return y < x ? y : x; }
int snl_max(int x, int y) { return x < y ? y : x; }
int snl_clamp(int x, int a, int b) { return snl_min(snl_max(x, a), b); }
SNL_MINMAX(vec2, bvec2, float)
SNL_MINMAX(vec3, bvec3, float)
SNL_MINMAX(vec4, bvec4, float)
#undef SNL_MINMAX

// mix as GLSL ES 3.00 defines it (section 8.3), x * (1 - a) + y * a, and
// normalize as x / length(x) divided per component.  SwiftShader 4.1 computes
// mix as x + a * (y - x), which differs in the last bit for about half of all
// operands, and normalize by a multiply with an inverse square root (1-2 ulp):
// those last bits move camera-ray hit points, and the normals pass's central
// differences (eps = 2e-4) magnify them to ~1e-4.  The boolean-select forms
// of mix are exact and keep the builtin.
float snl_mix(float x, float y, float a) { return x * (1.0 - a) + y * a; }
vec2 snl_mix(vec2 x, vec2 y, float a) { return x * (1.0 - a) + y * a; }
vec3 snl_mix(vec3 x, vec3 y, float a) { return x * (1.0 - a) + y * a; }
vec4 snl_mix(vec4 x, vec4 y, float a) { return x * (1.0 - a) + y * a; }
vec2 snl_mix(vec2 x, vec2 y, vec2 a) { return x * (1.0 - a) + y * a; }
vec3 snl_mix(vec3 x, vec3 y, vec3 a) { return x * (1.0 - a) + y * a; }
vec4 snl_mix(vec4 x, vec4 y, vec4 a) { return x * (1.0 - a) + y * a; }
float snl_mix(float x, float y, bool a) { return mix(x, y, a); }
vec2 snl_mix(vec2 x, vec2 y, bvec2 a) { return mix(x, y, a); }
vec3 snl_mix(vec3 x, vec3 y, bvec3 a) { return mix(x, y, a); }
vec4 snl_mix(vec4 x, vec4 y, bvec4 a) { return mix(x, y, a); }
float snl_normalize(float x) { return x / abs(x); }
vec2 snl_normalize(vec2 x) { float l = sqrt(x.x * x.x + x.y * x.y); return vec2(x.x / l, x.y / l); }
vec3 snl_normalize(vec3 x) {
  float l = sqrt(x.x * x.x + x.y * x.y + x.z * x.z);
  return vec3(x.x / l, x.y / l, x.z / l);
}
vec4 snl_normalize(vec4 x) {
  float l = sqrt(x.x * x.x + x.y * x.y + x.z * x.z + x.w * x.w);
  return vec4(x.x / l, x.y / l, x.z / l, x.w / l);
}

#define mix snl_mix
#define normalize snl_normalize
#define min snl_min
#define max snl_max
#define clamp snl_clamp
#define sin snl_sin
#define cos snl_cos
#define tan snl_tan
#define exp snl_exp
#define log snl_log
#define pow snl_pow
#define atan snl_atan
#define acos snl_acos
#define asin snl_asin
