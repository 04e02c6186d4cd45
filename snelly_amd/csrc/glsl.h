// glsl.h -- GLSL ES 3.00 vector/matrix types and builtins for HIP device code
// (gfx950, compiled at run time with hipRTC) and for host C++.
//
// Scene snippets written for Snelly's WebGL2 shaders (SDF_SURFACE, SDF_METAL,
// SDF_DIELECTRIC and the material hooks; see js/sceneAPI.js:120-198 in the
// reference) are translated token-wise to C++ and compiled against this header.
// Semantics follow the GLSL ES 3.00 specification:
//   * vectors of float / int / bool with read swizzles of up to 4 components
//     (xyzw and rgba sets) and write swizzles when no component repeats;
//   * matrices are column-major, m[i] is column i, M*v and v*M as in GLSL;
//   * mod(x,y) = x - y*floor(x/y), fract(x) = x - floor(x), mix(x,y,a) =
//     x*(1-a) + y*a, smoothstep uses t*t*(3-2t), sign(0) = 0;
//   * texture() on 1..4-channel float samplers with GL_LINEAR / GL_NEAREST and
//     CLAMP_TO_EDGE / REPEAT addressing (texel centres at (i+0.5)/N), the way
//     GLU.Texture configures them (js/gl.js:485-516).
// Every function is inline; nothing here allocates or touches global state.
#pragma once

#if defined(__HIPCC_RTC__) || defined(__HIPCC__) || defined(__HIP__)
#define SNL_HD __host__ __device__ inline
#define SNL_DEVICE_CODE 1
#else
#include <cmath>
#define SNL_HD inline
#endif

namespace glsl {

// ---------------------------------------------------------------- scalars
#if defined(__HIP_DEVICE_COMPILE__)
SNL_HD float snl_sqrtf(float x) { return __builtin_sqrtf(x); }
SNL_HD float snl_floorf(float x) { return __builtin_floorf(x); }
SNL_HD float snl_ceilf(float x) { return __builtin_ceilf(x); }
SNL_HD float snl_truncf(float x) { return __builtin_truncf(x); }
SNL_HD float snl_fabsf(float x) { return __builtin_fabsf(x); }
#else
SNL_HD float snl_sqrtf(float x) { return ::sqrtf(x); }
SNL_HD float snl_floorf(float x) { return ::floorf(x); }
SNL_HD float snl_ceilf(float x) { return ::ceilf(x); }
SNL_HD float snl_truncf(float x) { return ::truncf(x); }
SNL_HD float snl_fabsf(float x) { return ::fabsf(x); }
#endif

SNL_HD float abs(float x) { return snl_fabsf(x); }
SNL_HD int abs(int x) { return x < 0 ? -x : x; }
SNL_HD float sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
SNL_HD int sign(int x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }
SNL_HD float floor(float x) { return snl_floorf(x); }
SNL_HD float ceil(float x) { return snl_ceilf(x); }
SNL_HD float trunc(float x) { return snl_truncf(x); }
SNL_HD float round(float x) { return snl_floorf(x + 0.5f); }
SNL_HD float roundEven(float x) { return ::rintf(x); }
SNL_HD float fract(float x) { return x - snl_floorf(x); }
SNL_HD float mod(float x, float y) { return x - y * snl_floorf(x / y); }
// GLSL leaves min/max of NaN undefined; fminf/fmaxf map to one v_min/v_max_f32
#if defined(__HIP_DEVICE_COMPILE__)
SNL_HD float min(float a, float b) { return __builtin_fminf(a, b); }
SNL_HD float max(float a, float b) { return __builtin_fmaxf(a, b); }
#else
SNL_HD float min(float a, float b) { return b < a ? b : a; }
SNL_HD float max(float a, float b) { return a < b ? b : a; }
#endif
SNL_HD int min(int a, int b) { return b < a ? b : a; }
SNL_HD int max(int a, int b) { return a < b ? b : a; }
SNL_HD unsigned min(unsigned a, unsigned b) { return b < a ? b : a; }
SNL_HD unsigned max(unsigned a, unsigned b) { return a < b ? b : a; }
SNL_HD float clamp(float x, float lo, float hi) { return min(max(x, lo), hi); }
SNL_HD int clamp(int x, int lo, int hi) { return min(max(x, lo), hi); }
SNL_HD float mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
SNL_HD float mix(float x, float y, bool a) { return a ? y : x; }
SNL_HD float step(float edge, float x) { return x < edge ? 0.0f : 1.0f; }
SNL_HD float smoothstep(float e0, float e1, float x) {
  float t = clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
SNL_HD float sqrt(float x) { return snl_sqrtf(x); }
SNL_HD float inversesqrt(float x) { return 1.0f / snl_sqrtf(x); }
SNL_HD float pow(float x, float y) { return ::powf(x, y); }
SNL_HD float exp(float x) { return ::expf(x); }
SNL_HD float exp2(float x) { return ::exp2f(x); }
SNL_HD float log(float x) { return ::logf(x); }
SNL_HD float log2(float x) { return ::log2f(x); }
SNL_HD float sin(float x) { return ::sinf(x); }
SNL_HD float cos(float x) { return ::cosf(x); }
SNL_HD float tan(float x) { return ::tanf(x); }
SNL_HD float asin(float x) { return ::asinf(x); }
SNL_HD float acos(float x) { return ::acosf(x); }
SNL_HD float atan(float y, float x) { return ::atan2f(y, x); }
SNL_HD float atan(float x) { return ::atanf(x); }
SNL_HD float sinh(float x) { return ::sinhf(x); }
SNL_HD float cosh(float x) { return ::coshf(x); }
SNL_HD float tanh(float x) { return ::tanhf(x); }
SNL_HD float asinh(float x) { return ::asinhf(x); }
SNL_HD float acosh(float x) { return ::acoshf(x); }
SNL_HD float atanh(float x) { return ::atanhf(x); }
SNL_HD float radians(float d) { return d * 0.017453292519943295f; }
SNL_HD float degrees(float r) { return r * 57.29577951308232f; }
SNL_HD float length(float x) { return abs(x); }
SNL_HD float distance(float a, float b) { return abs(a - b); }
SNL_HD float dot(float a, float b) { return a * b; }
SNL_HD float normalize(float x) { return sign(x); }
SNL_HD bool isnan(float x) { return x != x; }
SNL_HD bool isinf(float x) { return snl_fabsf(x) == __builtin_huge_valf(); }

struct vec2; struct vec3; struct vec4;
struct ivec2; struct ivec3; struct ivec4;
struct bvec2; struct bvec3; struct bvec4;

// ------------------------------------------------------------ swizzle proxies
// A proxy has the same storage as its parent vector (N floats) and names the
// components it selects.  Reading converts to the result vector V; writing
// assigns component-wise (only meaningful when no index repeats, as in GLSL).
template <int N, class V, int A, int B>
struct sw2 {
  float e[N];
  SNL_HD operator V() const;
  SNL_HD sw2& operator=(const V& v);
  SNL_HD sw2& operator=(const sw2& o) { return *this = V(o); }
  template <class T> SNL_HD sw2& operator+=(const T& t) { return *this = V(*this) + t; }
  template <class T> SNL_HD sw2& operator-=(const T& t) { return *this = V(*this) - t; }
  template <class T> SNL_HD sw2& operator*=(const T& t) { return *this = V(*this) * t; }
  template <class T> SNL_HD sw2& operator/=(const T& t) { return *this = V(*this) / t; }
  SNL_HD V operator-() const { return -V(*this); }
  SNL_HD float operator[](int i) const { return i == 0 ? e[A] : e[B]; }
};
template <int N, class V, int A, int B, int C>
struct sw3 {
  float e[N];
  SNL_HD operator V() const;
  SNL_HD sw3& operator=(const V& v);
  SNL_HD sw3& operator=(const sw3& o) { return *this = V(o); }
  template <class T> SNL_HD sw3& operator+=(const T& t) { return *this = V(*this) + t; }
  template <class T> SNL_HD sw3& operator-=(const T& t) { return *this = V(*this) - t; }
  template <class T> SNL_HD sw3& operator*=(const T& t) { return *this = V(*this) * t; }
  template <class T> SNL_HD sw3& operator/=(const T& t) { return *this = V(*this) / t; }
  SNL_HD V operator-() const { return -V(*this); }
  SNL_HD float operator[](int i) const { return i == 0 ? e[A] : (i == 1 ? e[B] : e[C]); }
};
template <int N, class V, int A, int B, int C, int D>
struct sw4 {
  float e[N];
  SNL_HD operator V() const;
  SNL_HD sw4& operator=(const V& v);
  SNL_HD sw4& operator=(const sw4& o) { return *this = V(o); }
  template <class T> SNL_HD sw4& operator+=(const T& t) { return *this = V(*this) + t; }
  template <class T> SNL_HD sw4& operator-=(const T& t) { return *this = V(*this) - t; }
  template <class T> SNL_HD sw4& operator*=(const T& t) { return *this = V(*this) * t; }
  template <class T> SNL_HD sw4& operator/=(const T& t) { return *this = V(*this) / t; }
  SNL_HD V operator-() const { return -V(*this); }
  SNL_HD float operator[](int i) const {
    return i == 0 ? e[A] : (i == 1 ? e[B] : (i == 2 ? e[C] : e[D]));
  }
};

struct mat2; struct mat3; struct mat4;

// GLSL constructors may be given more components than the vector holds as
// long as the last argument is (partly) used -- vec3(a, b, v2) takes v2.x.
// Such calls land on a variadic constructor that flattens its arguments.
template <class T> struct snl_ncomp { static constexpr int n = 1; };
template <> struct snl_ncomp<vec2> { static constexpr int n = 2; };
template <> struct snl_ncomp<vec3> { static constexpr int n = 3; };
template <> struct snl_ncomp<vec4> { static constexpr int n = 4; };
template <> struct snl_ncomp<ivec2> { static constexpr int n = 2; };
template <> struct snl_ncomp<ivec3> { static constexpr int n = 3; };
template <> struct snl_ncomp<ivec4> { static constexpr int n = 4; };
template <int N, class V, int A, int B> struct snl_ncomp<sw2<N, V, A, B>> { static constexpr int n = 2; };
template <int N, class V, int A, int B, int C> struct snl_ncomp<sw3<N, V, A, B, C>> { static constexpr int n = 3; };
template <int N, class V, int A, int B, int C, int D> struct snl_ncomp<sw4<N, V, A, B, C, D>> { static constexpr int n = 4; };
template <class... A> struct snl_sumc { static constexpr int n = 0; };
template <class H, class... R> struct snl_sumc<H, R...> { static constexpr int n = snl_ncomp<H>::n + snl_sumc<R...>::n; };
template <bool B> struct snl_if {};
template <> struct snl_if<true> { typedef int type; };
template <class T> SNL_HD void snl_put(float* t, int& k, const T& v) { t[k++] = float(v); }
// an argument that may stand for an out/inout parameter of vector type V:
// V itself or a swizzle of a vector that converts to V (GLSL copies in and out)
template <class A, class V> struct snl_lv { static constexpr bool ok = false; };
template <class V> struct snl_lv<V, V> { static constexpr bool ok = true; };
template <int N, class V, int A, int B> struct snl_lv<sw2<N, V, A, B>, V> { static constexpr bool ok = true; };
template <int N, class V, int A, int B, int C> struct snl_lv<sw3<N, V, A, B, C>, V> { static constexpr bool ok = true; };
template <int N, class V, int A, int B, int C, int D> struct snl_lv<sw4<N, V, A, B, C, D>, V> { static constexpr bool ok = true; };
#define SNL_VARIADIC_CTOR(T, K)                                                              \
  template <class A0, class A1, class... R, typename snl_if<(snl_sumc<A0, A1, R...>::n >= K)>::type = 0> \
  SNL_HD T(const A0& a0, const A1& a1, const R&... r) {                                      \
    float t[snl_sumc<A0, A1, R...>::n];                                                      \
    int k = 0;                                                                               \
    snl_put(t, k, a0);                                                                       \
    snl_put(t, k, a1);                                                                       \
    (snl_put(t, k, r), ...);                                                                 \
    for (int i = 0; i < K; ++i) e[i] = t[i];                                                 \
  }

// Component-wise arithmetic for a float vector type T with K components,
// written once and stamped out for vec2/vec3/vec4.  Declared as friends so
// that argument-dependent lookup finds them for swizzle proxies too.
#define SNL_VEC_OPS(T, K)                                                                 \
  SNL_HD T& operator+=(const T& o) { for (int i = 0; i < K; ++i) e[i] += o.e[i]; return *this; } \
  SNL_HD T& operator-=(const T& o) { for (int i = 0; i < K; ++i) e[i] -= o.e[i]; return *this; } \
  SNL_HD T& operator*=(const T& o) { for (int i = 0; i < K; ++i) e[i] *= o.e[i]; return *this; } \
  SNL_HD T& operator/=(const T& o) { for (int i = 0; i < K; ++i) e[i] /= o.e[i]; return *this; } \
  SNL_HD T& operator+=(float s) { for (int i = 0; i < K; ++i) e[i] += s; return *this; }      \
  SNL_HD T& operator-=(float s) { for (int i = 0; i < K; ++i) e[i] -= s; return *this; }      \
  SNL_HD T& operator*=(float s) { for (int i = 0; i < K; ++i) e[i] *= s; return *this; }      \
  SNL_HD T& operator/=(float s) { for (int i = 0; i < K; ++i) e[i] /= s; return *this; }      \
  SNL_HD float& operator[](int i) { return e[i]; }                                          \
  SNL_HD float operator[](int i) const { return e[i]; }                                     \
  SNL_HD T operator-() const { T r; for (int i = 0; i < K; ++i) r.e[i] = -e[i]; return r; } \
  SNL_HD T operator+() const { return *this; }                                              \
  friend SNL_HD T operator+(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] + b.e[i]; return r; } \
  friend SNL_HD T operator-(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] - b.e[i]; return r; } \
  friend SNL_HD T operator*(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] * b.e[i]; return r; } \
  friend SNL_HD T operator/(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] / b.e[i]; return r; } \
  friend SNL_HD T operator+(const T& a, float s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] + s; return r; } \
  friend SNL_HD T operator-(const T& a, float s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] - s; return r; } \
  friend SNL_HD T operator*(const T& a, float s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] * s; return r; } \
  friend SNL_HD T operator/(const T& a, float s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] / s; return r; } \
  friend SNL_HD T operator+(float s, const T& a) { T r; for (int i = 0; i < K; ++i) r.e[i] = s + a.e[i]; return r; } \
  friend SNL_HD T operator-(float s, const T& a) { T r; for (int i = 0; i < K; ++i) r.e[i] = s - a.e[i]; return r; } \
  friend SNL_HD T operator*(float s, const T& a) { T r; for (int i = 0; i < K; ++i) r.e[i] = s * a.e[i]; return r; } \
  friend SNL_HD T operator/(float s, const T& a) { T r; for (int i = 0; i < K; ++i) r.e[i] = s / a.e[i]; return r; } \
  friend SNL_HD bool operator==(const T& a, const T& b) { for (int i = 0; i < K; ++i) if (!(a.e[i] == b.e[i])) return false; return true; } \
  friend SNL_HD bool operator!=(const T& a, const T& b) { return !(a == b); }

struct vec2 {
  union {
    struct { float x, y; };
    struct { float r, g; };
    struct { float s, t; };
    float e[2];
#include "glsl_swz2.h"
  };
  SNL_HD vec2() : x(0.0f), y(0.0f) {}
  SNL_HD explicit vec2(float v) : x(v), y(v) {}
  SNL_HD vec2(float a, float b) : x(a), y(b) {}
  SNL_HD vec2(const vec2& o) : x(o.x), y(o.y) {}
  SNL_VARIADIC_CTOR(vec2, 2)
  SNL_HD explicit vec2(const vec3& o);
  SNL_HD explicit vec2(const vec4& o);
  SNL_HD explicit vec2(const ivec2& o);
  SNL_HD explicit vec2(const bvec2& o);
  SNL_HD vec2& operator=(const vec2& o) { x = o.x; y = o.y; return *this; }
  SNL_VEC_OPS(vec2, 2)
  SNL_HD vec2& operator*=(const mat2& m);
};

struct vec3 {
  union {
    struct { float x, y, z; };
    struct { float r, g, b; };
    float e[3];
#include "glsl_swz3.h"
  };
  SNL_HD vec3() : x(0.0f), y(0.0f), z(0.0f) {}
  SNL_HD explicit vec3(float v) : x(v), y(v), z(v) {}
  SNL_HD vec3(float a, float b, float c) : x(a), y(b), z(c) {}
  SNL_HD vec3(const vec2& a, float c) : x(a.x), y(a.y), z(c) {}
  SNL_HD vec3(float a, const vec2& b) : x(a), y(b.x), z(b.y) {}
  SNL_HD vec3(const vec3& o) : x(o.x), y(o.y), z(o.z) {}
  SNL_VARIADIC_CTOR(vec3, 3)
  SNL_HD explicit vec3(const vec4& o);
  SNL_HD explicit vec3(const ivec3& o);
  SNL_HD explicit vec3(const bvec3& o);
  SNL_HD vec3& operator=(const vec3& o) { x = o.x; y = o.y; z = o.z; return *this; }
  SNL_VEC_OPS(vec3, 3)
  SNL_HD vec3& operator*=(const mat3& m);
};

struct vec4 {
  union {
    struct { float x, y, z, w; };
    struct { float r, g, b, a; };
    float e[4];
#include "glsl_swz4.h"
  };
  SNL_HD vec4() : x(0.0f), y(0.0f), z(0.0f), w(0.0f) {}
  SNL_HD explicit vec4(float v) : x(v), y(v), z(v), w(v) {}
  SNL_HD vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
  SNL_HD vec4(const vec3& v, float d) : x(v.x), y(v.y), z(v.z), w(d) {}
  SNL_HD vec4(float a, const vec3& v) : x(a), y(v.x), z(v.y), w(v.z) {}
  SNL_HD vec4(const vec2& v, float c, float d) : x(v.x), y(v.y), z(c), w(d) {}
  SNL_HD vec4(float a, const vec2& v, float d) : x(a), y(v.x), z(v.y), w(d) {}
  SNL_HD vec4(float a, float b, const vec2& v) : x(a), y(b), z(v.x), w(v.y) {}
  SNL_HD vec4(const vec2& u, const vec2& v) : x(u.x), y(u.y), z(v.x), w(v.y) {}
  SNL_HD vec4(const vec4& o) : x(o.x), y(o.y), z(o.z), w(o.w) {}
  SNL_VARIADIC_CTOR(vec4, 4)
  SNL_HD explicit vec4(const ivec4& o);
  SNL_HD explicit vec4(const bvec4& o);
  SNL_HD vec4& operator=(const vec4& o) { x = o.x; y = o.y; z = o.z; w = o.w; return *this; }
  SNL_VEC_OPS(vec4, 4)
  SNL_HD vec4& operator*=(const mat4& m);
};

SNL_HD vec2::vec2(const vec3& o) : x(o.x), y(o.y) {}
SNL_HD vec2::vec2(const vec4& o) : x(o.x), y(o.y) {}
SNL_HD vec3::vec3(const vec4& o) : x(o.x), y(o.y), z(o.z) {}

SNL_HD void snl_put(float* t, int& k, const vec2& v) { t[k++] = v.x; t[k++] = v.y; }
SNL_HD void snl_put(float* t, int& k, const vec3& v) { t[k++] = v.x; t[k++] = v.y; t[k++] = v.z; }
SNL_HD void snl_put(float* t, int& k, const vec4& v) { t[k++] = v.x; t[k++] = v.y; t[k++] = v.z; t[k++] = v.w; }
template <int N, class V, int A, int B>
SNL_HD void snl_put(float* t, int& k, const sw2<N, V, A, B>& v) { t[k++] = v.e[A]; t[k++] = v.e[B]; }
template <int N, class V, int A, int B, int C>
SNL_HD void snl_put(float* t, int& k, const sw3<N, V, A, B, C>& v) { t[k++] = v.e[A]; t[k++] = v.e[B]; t[k++] = v.e[C]; }
template <int N, class V, int A, int B, int C, int D>
SNL_HD void snl_put(float* t, int& k, const sw4<N, V, A, B, C, D>& v) {
  t[k++] = v.e[A]; t[k++] = v.e[B]; t[k++] = v.e[C]; t[k++] = v.e[D];
}

// GLSL float(x): a scalar conversion, or the first component of a vector
template <class T> SNL_HD float snl_float(const T& v) { return float(v); }
SNL_HD float snl_float(const vec2& v) { return v.x; }
SNL_HD float snl_float(const vec3& v) { return v.x; }
SNL_HD float snl_float(const vec4& v) { return v.x; }
template <int N, class V, int A, int B> SNL_HD float snl_float(const sw2<N, V, A, B>& v) { return v.e[A]; }
template <int N, class V, int A, int B, int C> SNL_HD float snl_float(const sw3<N, V, A, B, C>& v) { return v.e[A]; }
template <int N, class V, int A, int B, int C, int D> SNL_HD float snl_float(const sw4<N, V, A, B, C, D>& v) { return v.e[A]; }

// swizzle proxy conversions (need the complete vector types)
template <class V> struct vmake;
template <> struct vmake<vec2> { SNL_HD static vec2 m(float a, float b) { return vec2(a, b); } };
template <> struct vmake<vec3> { SNL_HD static vec3 m(float a, float b, float c) { return vec3(a, b, c); } };
template <> struct vmake<vec4> { SNL_HD static vec4 m(float a, float b, float c, float d) { return vec4(a, b, c, d); } };

template <int N, class V, int A, int B>
SNL_HD sw2<N, V, A, B>::operator V() const { return vec2(e[A], e[B]); }
template <int N, class V, int A, int B>
SNL_HD sw2<N, V, A, B>& sw2<N, V, A, B>::operator=(const V& v) { float a = v.x, b = v.y; e[A] = a; e[B] = b; return *this; }
template <int N, class V, int A, int B, int C>
SNL_HD sw3<N, V, A, B, C>::operator V() const { return vec3(e[A], e[B], e[C]); }
template <int N, class V, int A, int B, int C>
SNL_HD sw3<N, V, A, B, C>& sw3<N, V, A, B, C>::operator=(const V& v) {
  float a = v.x, b = v.y, c = v.z; e[A] = a; e[B] = b; e[C] = c; return *this;
}
template <int N, class V, int A, int B, int C, int D>
SNL_HD sw4<N, V, A, B, C, D>::operator V() const { return vec4(e[A], e[B], e[C], e[D]); }
template <int N, class V, int A, int B, int C, int D>
SNL_HD sw4<N, V, A, B, C, D>& sw4<N, V, A, B, C, D>::operator=(const V& v) {
  float a = v.x, b = v.y, c = v.z, d = v.w; e[A] = a; e[B] = b; e[C] = c; e[D] = d; return *this;
}

// ------------------------------------------------------------ integer / bool vectors
#define SNL_IVEC_OPS(T, K)                                                                \
  SNL_HD int& operator[](int i) { return e[i]; }                                            \
  SNL_HD int operator[](int i) const { return e[i]; }                                       \
  SNL_HD T operator-() const { T r; for (int i = 0; i < K; ++i) r.e[i] = -e[i]; return r; } \
  SNL_HD T& operator+=(const T& o) { for (int i = 0; i < K; ++i) e[i] += o.e[i]; return *this; } \
  SNL_HD T& operator-=(const T& o) { for (int i = 0; i < K; ++i) e[i] -= o.e[i]; return *this; } \
  SNL_HD T& operator*=(const T& o) { for (int i = 0; i < K; ++i) e[i] *= o.e[i]; return *this; } \
  SNL_HD T& operator*=(int s) { for (int i = 0; i < K; ++i) e[i] *= s; return *this; }        \
  friend SNL_HD T operator+(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] + b.e[i]; return r; } \
  friend SNL_HD T operator-(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] - b.e[i]; return r; } \
  friend SNL_HD T operator*(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] * b.e[i]; return r; } \
  friend SNL_HD T operator/(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] / b.e[i]; return r; } \
  friend SNL_HD T operator%(const T& a, const T& b) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] % b.e[i]; return r; } \
  friend SNL_HD T operator+(const T& a, int s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] + s; return r; } \
  friend SNL_HD T operator-(const T& a, int s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] - s; return r; } \
  friend SNL_HD T operator*(const T& a, int s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] * s; return r; } \
  friend SNL_HD T operator/(const T& a, int s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] / s; return r; } \
  friend SNL_HD T operator%(const T& a, int s) { T r; for (int i = 0; i < K; ++i) r.e[i] = a.e[i] % s; return r; } \
  friend SNL_HD T operator*(int s, const T& a) { T r; for (int i = 0; i < K; ++i) r.e[i] = s * a.e[i]; return r; } \
  friend SNL_HD bool operator==(const T& a, const T& b) { for (int i = 0; i < K; ++i) if (a.e[i] != b.e[i]) return false; return true; } \
  friend SNL_HD bool operator!=(const T& a, const T& b) { return !(a == b); }

struct ivec2 {
  union { struct { int x, y; }; struct { int r, g; }; int e[2]; };
  SNL_HD ivec2() : x(0), y(0) {}
  SNL_HD explicit ivec2(int v) : x(v), y(v) {}
  SNL_HD ivec2(int a, int b) : x(a), y(b) {}
  SNL_HD explicit ivec2(const vec2& v) : x(int(v.x)), y(int(v.y)) {}
  SNL_IVEC_OPS(ivec2, 2)
};
struct ivec3 {
  union { struct { int x, y, z; }; struct { int r, g, b; }; int e[3]; };
  SNL_HD ivec3() : x(0), y(0), z(0) {}
  SNL_HD explicit ivec3(int v) : x(v), y(v), z(v) {}
  SNL_HD ivec3(int a, int b, int c) : x(a), y(b), z(c) {}
  SNL_HD explicit ivec3(const vec3& v) : x(int(v.x)), y(int(v.y)), z(int(v.z)) {}
  SNL_IVEC_OPS(ivec3, 3)
};
struct ivec4 {
  union { struct { int x, y, z, w; }; struct { int r, g, b, a; }; int e[4]; };
  SNL_HD ivec4() : x(0), y(0), z(0), w(0) {}
  SNL_HD explicit ivec4(int v) : x(v), y(v), z(v), w(v) {}
  SNL_HD ivec4(int a, int b, int c, int d) : x(a), y(b), z(c), w(d) {}
  SNL_HD explicit ivec4(const vec4& v) : x(int(v.x)), y(int(v.y)), z(int(v.z)), w(int(v.w)) {}
  SNL_IVEC_OPS(ivec4, 4)
};
SNL_HD vec2::vec2(const ivec2& o) : x(float(o.x)), y(float(o.y)) {}
SNL_HD vec3::vec3(const ivec3& o) : x(float(o.x)), y(float(o.y)), z(float(o.z)) {}
SNL_HD vec4::vec4(const ivec4& o) : x(float(o.x)), y(float(o.y)), z(float(o.z)), w(float(o.w)) {}

struct bvec2 { bool x, y; SNL_HD bvec2() : x(false), y(false) {} SNL_HD bvec2(bool a, bool b) : x(a), y(b) {} SNL_HD explicit bvec2(bool a) : x(a), y(a) {} };
struct bvec3 { bool x, y, z; SNL_HD bvec3() : x(false), y(false), z(false) {} SNL_HD bvec3(bool a, bool b, bool c) : x(a), y(b), z(c) {} SNL_HD explicit bvec3(bool a) : x(a), y(a), z(a) {} };
struct bvec4 { bool x, y, z, w; SNL_HD bvec4() : x(false), y(false), z(false), w(false) {} SNL_HD bvec4(bool a, bool b, bool c, bool d) : x(a), y(b), z(c), w(d) {} SNL_HD explicit bvec4(bool a) : x(a), y(a), z(a), w(a) {} };
SNL_HD vec2::vec2(const bvec2& o) : x(o.x ? 1.0f : 0.0f), y(o.y ? 1.0f : 0.0f) {}
SNL_HD vec3::vec3(const bvec3& o) : x(o.x ? 1.0f : 0.0f), y(o.y ? 1.0f : 0.0f), z(o.z ? 1.0f : 0.0f) {}
SNL_HD vec4::vec4(const bvec4& o) : x(o.x ? 1.0f : 0.0f), y(o.y ? 1.0f : 0.0f), z(o.z ? 1.0f : 0.0f), w(o.w ? 1.0f : 0.0f) {}
SNL_HD void snl_put(float* t, int& k, const ivec2& v) { t[k++] = float(v.x); t[k++] = float(v.y); }
SNL_HD void snl_put(float* t, int& k, const ivec3& v) { t[k++] = float(v.x); t[k++] = float(v.y); t[k++] = float(v.z); }
SNL_HD void snl_put(float* t, int& k, const ivec4& v) {
  t[k++] = float(v.x); t[k++] = float(v.y); t[k++] = float(v.z); t[k++] = float(v.w);
}
SNL_HD bool any(bvec2 b) { return b.x || b.y; }
SNL_HD bool any(bvec3 b) { return b.x || b.y || b.z; }
SNL_HD bool any(bvec4 b) { return b.x || b.y || b.z || b.w; }
SNL_HD bool all(bvec2 b) { return b.x && b.y; }
SNL_HD bool all(bvec3 b) { return b.x && b.y && b.z; }
SNL_HD bool all(bvec4 b) { return b.x && b.y && b.z && b.w; }
SNL_HD bvec2 not_(bvec2 b) { return bvec2(!b.x, !b.y); }
SNL_HD bvec3 not_(bvec3 b) { return bvec3(!b.x, !b.y, !b.z); }
SNL_HD bvec4 not_(bvec4 b) { return bvec4(!b.x, !b.y, !b.z, !b.w); }

// ------------------------------------------------------------ component-wise builtins
#define SNL_MAP1(FN)                                                                        \
  SNL_HD vec2 FN(const vec2& v) { return vec2(FN(v.x), FN(v.y)); }                          \
  SNL_HD vec3 FN(const vec3& v) { return vec3(FN(v.x), FN(v.y), FN(v.z)); }                 \
  SNL_HD vec4 FN(const vec4& v) { return vec4(FN(v.x), FN(v.y), FN(v.z), FN(v.w)); }
SNL_MAP1(abs) SNL_MAP1(sign) SNL_MAP1(floor) SNL_MAP1(ceil) SNL_MAP1(trunc) SNL_MAP1(round)
SNL_MAP1(roundEven) SNL_MAP1(fract) SNL_MAP1(sqrt) SNL_MAP1(inversesqrt) SNL_MAP1(exp) SNL_MAP1(exp2)
SNL_MAP1(log) SNL_MAP1(log2) SNL_MAP1(sin) SNL_MAP1(cos) SNL_MAP1(tan) SNL_MAP1(asin)
SNL_MAP1(acos) SNL_MAP1(atan) SNL_MAP1(sinh) SNL_MAP1(cosh) SNL_MAP1(tanh) SNL_MAP1(radians) SNL_MAP1(degrees)

#define SNL_MAP2(FN)                                                                        \
  SNL_HD vec2 FN(const vec2& a, const vec2& b) { return vec2(FN(a.x, b.x), FN(a.y, b.y)); } \
  SNL_HD vec3 FN(const vec3& a, const vec3& b) { return vec3(FN(a.x, b.x), FN(a.y, b.y), FN(a.z, b.z)); } \
  SNL_HD vec4 FN(const vec4& a, const vec4& b) { return vec4(FN(a.x, b.x), FN(a.y, b.y), FN(a.z, b.z), FN(a.w, b.w)); } \
  SNL_HD vec2 FN(const vec2& a, float b) { return vec2(FN(a.x, b), FN(a.y, b)); }           \
  SNL_HD vec3 FN(const vec3& a, float b) { return vec3(FN(a.x, b), FN(a.y, b), FN(a.z, b)); } \
  SNL_HD vec4 FN(const vec4& a, float b) { return vec4(FN(a.x, b), FN(a.y, b), FN(a.z, b), FN(a.w, b)); }
SNL_MAP2(mod) SNL_MAP2(min) SNL_MAP2(max) SNL_MAP2(pow) SNL_MAP2(atan)

SNL_HD vec2 step(const vec2& e, const vec2& x) { return vec2(step(e.x, x.x), step(e.y, x.y)); }
SNL_HD vec3 step(const vec3& e, const vec3& x) { return vec3(step(e.x, x.x), step(e.y, x.y), step(e.z, x.z)); }
SNL_HD vec4 step(const vec4& e, const vec4& x) { return vec4(step(e.x, x.x), step(e.y, x.y), step(e.z, x.z), step(e.w, x.w)); }
SNL_HD vec2 step(float e, const vec2& x) { return vec2(step(e, x.x), step(e, x.y)); }
SNL_HD vec3 step(float e, const vec3& x) { return vec3(step(e, x.x), step(e, x.y), step(e, x.z)); }
SNL_HD vec4 step(float e, const vec4& x) { return vec4(step(e, x.x), step(e, x.y), step(e, x.z), step(e, x.w)); }

#define SNL_MAP3(T)                                                                          \
  SNL_HD T clamp(const T& x, float lo, float hi) { return min(max(x, lo), hi); }              \
  SNL_HD T clamp(const T& x, const T& lo, const T& hi) { return min(max(x, lo), hi); }        \
  SNL_HD T mix(const T& x, const T& y, float a) { return x * (1.0f - a) + y * a; }            \
  SNL_HD T mix(const T& x, const T& y, const T& a) { return x * (1.0f - a) + y * a; }         \
  SNL_HD T smoothstep(float e0, float e1, const T& x) { T r; for (int i = 0; i < int(sizeof(T) / 4); ++i) r.e[i] = smoothstep(e0, e1, x.e[i]); return r; } \
  SNL_HD T smoothstep(const T& e0, const T& e1, const T& x) { T r; for (int i = 0; i < int(sizeof(T) / 4); ++i) r.e[i] = smoothstep(e0.e[i], e1.e[i], x.e[i]); return r; }
SNL_MAP3(vec2) SNL_MAP3(vec3) SNL_MAP3(vec4)

SNL_HD float dot(const vec2& a, const vec2& b) { return a.x * b.x + a.y * b.y; }
SNL_HD float dot(const vec3& a, const vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SNL_HD float dot(const vec4& a, const vec4& b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
SNL_HD float length(const vec2& v) { return sqrt(dot(v, v)); }
SNL_HD float length(const vec3& v) { return sqrt(dot(v, v)); }
SNL_HD float length(const vec4& v) { return sqrt(dot(v, v)); }
SNL_HD float distance(const vec2& a, const vec2& b) { return length(a - b); }
SNL_HD float distance(const vec3& a, const vec3& b) { return length(a - b); }
SNL_HD float distance(const vec4& a, const vec4& b) { return length(a - b); }
SNL_HD vec2 normalize(const vec2& v) { return v / length(v); }
SNL_HD vec3 normalize(const vec3& v) { return v / length(v); }
SNL_HD vec4 normalize(const vec4& v) { return v / length(v); }
SNL_HD vec3 cross(const vec3& a, const vec3& b) {
  return vec3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
SNL_HD vec2 reflect(const vec2& i, const vec2& n) { return i - 2.0f * dot(n, i) * n; }
SNL_HD vec3 reflect(const vec3& i, const vec3& n) { return i - 2.0f * dot(n, i) * n; }
SNL_HD vec3 refract(const vec3& i, const vec3& n, float eta) {
  float d = dot(n, i);
  float k = 1.0f - eta * eta * (1.0f - d * d);
  if (k < 0.0f) return vec3(0.0f);
  return eta * i - (eta * d + sqrt(k)) * n;
}
SNL_HD vec3 faceforward(const vec3& n, const vec3& i, const vec3& nref) { return dot(nref, i) < 0.0f ? n : -n; }

#define SNL_CMP(FN, OP)                                                                          \
  SNL_HD bvec2 FN(const vec2& a, const vec2& b) { return bvec2(a.x OP b.x, a.y OP b.y); }        \
  SNL_HD bvec3 FN(const vec3& a, const vec3& b) { return bvec3(a.x OP b.x, a.y OP b.y, a.z OP b.z); } \
  SNL_HD bvec4 FN(const vec4& a, const vec4& b) { return bvec4(a.x OP b.x, a.y OP b.y, a.z OP b.z, a.w OP b.w); }
SNL_CMP(lessThan, <) SNL_CMP(lessThanEqual, <=) SNL_CMP(greaterThan, >) SNL_CMP(greaterThanEqual, >=)
SNL_CMP(equal, ==) SNL_CMP(notEqual, !=)

SNL_HD ivec2 abs(const ivec2& v) { return ivec2(abs(v.x), abs(v.y)); }
SNL_HD ivec3 abs(const ivec3& v) { return ivec3(abs(v.x), abs(v.y), abs(v.z)); }
SNL_HD ivec4 abs(const ivec4& v) { return ivec4(abs(v.x), abs(v.y), abs(v.z), abs(v.w)); }

// ------------------------------------------------------------ matrices (column-major)
struct mat2 {
  vec2 c[2];
  SNL_HD mat2() { c[0] = vec2(1.0f, 0.0f); c[1] = vec2(0.0f, 1.0f); }
  SNL_HD explicit mat2(float d) { c[0] = vec2(d, 0.0f); c[1] = vec2(0.0f, d); }
  SNL_HD mat2(float a, float b, float cc, float d) { c[0] = vec2(a, b); c[1] = vec2(cc, d); }
  SNL_HD mat2(const vec2& a, const vec2& b) { c[0] = a; c[1] = b; }
  SNL_HD explicit mat2(const vec4& v) { c[0] = vec2(v.x, v.y); c[1] = vec2(v.z, v.w); }
  SNL_HD explicit mat2(const mat3& m);
  SNL_HD vec2& operator[](int i) { return c[i]; }
  SNL_HD const vec2& operator[](int i) const { return c[i]; }
};
struct mat3 {
  vec3 c[3];
  SNL_HD mat3() { c[0] = vec3(1.0f, 0.0f, 0.0f); c[1] = vec3(0.0f, 1.0f, 0.0f); c[2] = vec3(0.0f, 0.0f, 1.0f); }
  SNL_HD explicit mat3(float d) { c[0] = vec3(d, 0.0f, 0.0f); c[1] = vec3(0.0f, d, 0.0f); c[2] = vec3(0.0f, 0.0f, d); }
  SNL_HD mat3(float a0, float a1, float a2, float b0, float b1, float b2, float c0, float c1, float c2) {
    c[0] = vec3(a0, a1, a2); c[1] = vec3(b0, b1, b2); c[2] = vec3(c0, c1, c2);
  }
  SNL_HD mat3(const vec3& a, const vec3& b, const vec3& d) { c[0] = a; c[1] = b; c[2] = d; }
  SNL_HD explicit mat3(const mat4& m);
  SNL_HD vec3& operator[](int i) { return c[i]; }
  SNL_HD const vec3& operator[](int i) const { return c[i]; }
};
struct mat4 {
  vec4 c[4];
  SNL_HD mat4() { c[0] = vec4(1, 0, 0, 0); c[1] = vec4(0, 1, 0, 0); c[2] = vec4(0, 0, 1, 0); c[3] = vec4(0, 0, 0, 1); }
  SNL_HD explicit mat4(float d) { c[0] = vec4(d, 0, 0, 0); c[1] = vec4(0, d, 0, 0); c[2] = vec4(0, 0, d, 0); c[3] = vec4(0, 0, 0, d); }
  SNL_HD mat4(float a0, float a1, float a2, float a3, float b0, float b1, float b2, float b3,
              float c0, float c1, float c2, float c3, float d0, float d1, float d2, float d3) {
    c[0] = vec4(a0, a1, a2, a3); c[1] = vec4(b0, b1, b2, b3); c[2] = vec4(c0, c1, c2, c3); c[3] = vec4(d0, d1, d2, d3);
  }
  SNL_HD mat4(const vec4& a, const vec4& b, const vec4& d, const vec4& f) { c[0] = a; c[1] = b; c[2] = d; c[3] = f; }
  SNL_HD vec4& operator[](int i) { return c[i]; }
  SNL_HD const vec4& operator[](int i) const { return c[i]; }
};
SNL_HD mat2::mat2(const mat3& m) { c[0] = vec2(m.c[0]); c[1] = vec2(m.c[1]); }
SNL_HD mat3::mat3(const mat4& m) { c[0] = vec3(m.c[0]); c[1] = vec3(m.c[1]); c[2] = vec3(m.c[2]); }

#define SNL_MAT_OPS(M, V, K)                                                                \
  SNL_HD V operator*(const M& m, const V& v) { V r(0.0f); for (int j = 0; j < K; ++j) r += m.c[j] * v.e[j]; return r; } \
  SNL_HD V operator*(const V& v, const M& m) { V r; for (int j = 0; j < K; ++j) r.e[j] = dot(v, m.c[j]); return r; } \
  SNL_HD M operator*(const M& a, const M& b) { M r; for (int j = 0; j < K; ++j) r.c[j] = a * b.c[j]; return r; } \
  SNL_HD M operator*(const M& a, float s) { M r; for (int j = 0; j < K; ++j) r.c[j] = a.c[j] * s; return r; } \
  SNL_HD M operator*(float s, const M& a) { M r; for (int j = 0; j < K; ++j) r.c[j] = s * a.c[j]; return r; } \
  SNL_HD M operator+(const M& a, const M& b) { M r; for (int j = 0; j < K; ++j) r.c[j] = a.c[j] + b.c[j]; return r; } \
  SNL_HD M operator-(const M& a, const M& b) { M r; for (int j = 0; j < K; ++j) r.c[j] = a.c[j] - b.c[j]; return r; } \
  SNL_HD M operator-(const M& a) { M r; for (int j = 0; j < K; ++j) r.c[j] = -a.c[j]; return r; } \
  SNL_HD M operator/(const M& a, float s) { M r; for (int j = 0; j < K; ++j) r.c[j] = a.c[j] / s; return r; } \
  SNL_HD M operator+(const M& a, float s) { M r; for (int j = 0; j < K; ++j) r.c[j] = a.c[j] + s; return r; } \
  SNL_HD M operator-(const M& a, float s) { M r; for (int j = 0; j < K; ++j) r.c[j] = a.c[j] - s; return r; } \
  SNL_HD M& operator*=(M& a, float s) { for (int j = 0; j < K; ++j) a.c[j] *= s; return a; } \
  SNL_HD M& operator/=(M& a, float s) { for (int j = 0; j < K; ++j) a.c[j] /= s; return a; } \
  SNL_HD M& operator+=(M& a, const M& b) { for (int j = 0; j < K; ++j) a.c[j] += b.c[j]; return a; } \
  SNL_HD M& operator-=(M& a, const M& b) { for (int j = 0; j < K; ++j) a.c[j] -= b.c[j]; return a; } \
  SNL_HD M transpose(const M& a) { M r; for (int i = 0; i < K; ++i) for (int j = 0; j < K; ++j) r.c[i].e[j] = a.c[j].e[i]; return r; } \
  SNL_HD M matrixCompMult(const M& a, const M& b) { M r; for (int j = 0; j < K; ++j) r.c[j] = a.c[j] * b.c[j]; return r; }
SNL_MAT_OPS(mat2, vec2, 2) SNL_MAT_OPS(mat3, vec3, 3) SNL_MAT_OPS(mat4, vec4, 4)

SNL_HD vec2& vec2::operator*=(const mat2& m) { return *this = *this * m; }
SNL_HD vec3& vec3::operator*=(const mat3& m) { return *this = *this * m; }
SNL_HD vec4& vec4::operator*=(const mat4& m) { return *this = *this * m; }
SNL_HD mat2& operator*=(mat2& a, const mat2& b) { return a = a * b; }
SNL_HD mat3& operator*=(mat3& a, const mat3& b) { return a = a * b; }
SNL_HD mat4& operator*=(mat4& a, const mat4& b) { return a = a * b; }

SNL_HD float determinant(const mat2& m) { return m.c[0].x * m.c[1].y - m.c[1].x * m.c[0].y; }
SNL_HD float determinant(const mat3& m) { return dot(m.c[0], cross(m.c[1], m.c[2])); }
SNL_HD mat2 inverse(const mat2& m) {
  float d = determinant(m);
  return mat2(m.c[1].y / d, -m.c[0].y / d, -m.c[1].x / d, m.c[0].x / d);
}
SNL_HD mat3 inverse(const mat3& m) {
  vec3 r0 = cross(m.c[1], m.c[2]), r1 = cross(m.c[2], m.c[0]), r2 = cross(m.c[0], m.c[1]);
  float d = dot(m.c[0], r0);
  return transpose(mat3(r0 / d, r1 / d, r2 / d));
}

// ------------------------------------------------------------ samplers
// A sampler is a plain descriptor of a float texture resident in device memory:
// row-major texels of `ch` floats.  flags bit 0 = GL_LINEAR (else NEAREST),
// bit 1 = CLAMP_TO_EDGE (else REPEAT).
struct sampler2D { const float* data; int w, h, ch, flags; };
struct sampler3D { const float* data; int w, h, d, ch, flags; };

SNL_HD int snl_wrap(int i, int n, bool clampEdge) {
  if (clampEdge) return i < 0 ? 0 : (i >= n ? n - 1 : i);
  int m = i % n;
  return m < 0 ? m + n : m;
}
// channels a texture lacks read as (0, 0, 0, 1); the component writes are
// spelled out so that no vector is indexed by a runtime value (that would put
// it in scratch memory)
SNL_HD vec4 snl_texel(const float* p, int ch) {
  vec4 r(p[0], 0.0f, 0.0f, 1.0f);
  if (ch > 1) r.y = p[1];
  if (ch > 2) r.z = p[2];
  if (ch > 3) r.w = p[3];
  return r;
}
SNL_HD vec4 texture(const sampler2D& s, const vec2& uv) {
  if (!s.data || s.w <= 0 || s.h <= 0) return vec4(0.0f, 0.0f, 0.0f, 1.0f);
  bool ce = (s.flags & 2) != 0;
  if (!(s.flags & 1)) {
    int i = snl_wrap(int(floor(uv.x * float(s.w))), s.w, ce);
    int j = snl_wrap(int(floor(uv.y * float(s.h))), s.h, ce);
    return snl_texel(s.data + (size_t(j) * s.w + i) * s.ch, s.ch);
  }
  float fu = uv.x * float(s.w) - 0.5f, fv = uv.y * float(s.h) - 0.5f;
  float iu = floor(fu), iv = floor(fv);
  float au = fu - iu, av = fv - iv;
  int i0 = snl_wrap(int(iu), s.w, ce), i1 = snl_wrap(int(iu) + 1, s.w, ce);
  int j0 = snl_wrap(int(iv), s.h, ce), j1 = snl_wrap(int(iv) + 1, s.h, ce);
  vec4 t00 = snl_texel(s.data + (size_t(j0) * s.w + i0) * s.ch, s.ch);
  vec4 t10 = snl_texel(s.data + (size_t(j0) * s.w + i1) * s.ch, s.ch);
  vec4 t01 = snl_texel(s.data + (size_t(j1) * s.w + i0) * s.ch, s.ch);
  vec4 t11 = snl_texel(s.data + (size_t(j1) * s.w + i1) * s.ch, s.ch);
  return mix(mix(t00, t10, au), mix(t01, t11, au), av);
}
SNL_HD vec4 texture(const sampler3D& s, const vec3& uvw) {
  if (!s.data || s.w <= 0 || s.h <= 0 || s.d <= 0) return vec4(0.0f, 0.0f, 0.0f, 1.0f);
  bool ce = (s.flags & 2) != 0;
  if (!(s.flags & 1)) {
    int i = snl_wrap(int(floor(uvw.x * float(s.w))), s.w, ce);
    int j = snl_wrap(int(floor(uvw.y * float(s.h))), s.h, ce);
    int k = snl_wrap(int(floor(uvw.z * float(s.d))), s.d, ce);
    return snl_texel(s.data + ((size_t(k) * s.h + j) * s.w + i) * s.ch, s.ch);
  }
  float f[3] = {uvw.x * float(s.w) - 0.5f, uvw.y * float(s.h) - 0.5f, uvw.z * float(s.d) - 0.5f};
  int n[3] = {s.w, s.h, s.d};
  int i0[3], i1[3];
  float a[3];
  for (int q = 0; q < 3; ++q) {
    float fl = floor(f[q]);
    a[q] = f[q] - fl;
    i0[q] = snl_wrap(int(fl), n[q], ce);
    i1[q] = snl_wrap(int(fl) + 1, n[q], ce);
  }
  vec4 acc(0.0f);
  for (int c = 0; c < 8; ++c) {
    int ii = (c & 1) ? i1[0] : i0[0], jj = (c & 2) ? i1[1] : i0[1], kk = (c & 4) ? i1[2] : i0[2];
    float wgt = ((c & 1) ? a[0] : 1.0f - a[0]) * ((c & 2) ? a[1] : 1.0f - a[1]) * ((c & 4) ? a[2] : 1.0f - a[2]);
    acc += wgt * snl_texel(s.data + ((size_t(kk) * s.h + jj) * s.w + ii) * s.ch, s.ch);
  }
  return acc;
}
SNL_HD vec4 texture(const sampler2D& s, const vec2& uv, float /*bias*/) { return texture(s, uv); }
SNL_HD vec4 textureLod(const sampler2D& s, const vec2& uv, float /*lod*/) { return texture(s, uv); }
SNL_HD vec4 textureLod(const sampler3D& s, const vec3& uvw, float /*lod*/) { return texture(s, uvw); }
SNL_HD ivec2 textureSize(const sampler2D& s, int) { return ivec2(s.w, s.h); }
SNL_HD ivec3 textureSize(const sampler3D& s, int) { return ivec3(s.w, s.h, s.d); }

}  // namespace glsl
