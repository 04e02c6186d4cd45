// vec.h -- minimal float vectors for the CPU oracle (test infrastructure only).
// Independent of the product's glsl.h: plain structs, explicit component math,
// every operation rounded to float (x86-64 SSE, no FMA contraction).
#pragma once
#include <cmath>

namespace orc {

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

inline V2 v2(float x, float y) { return {x, y}; }
inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 v3(float s) { return {s, s, s}; }
inline V4 v4(float x, float y, float z, float w) { return {x, y, z, w}; }

inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator+(V3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
inline V3 operator-(V3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
inline V3 operator-(float s, V3 a) { return {s - a.x, s - a.y, s - a.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator/(V2 a, V2 b) { return {a.x / b.x, a.y / b.y}; }
inline V2 operator*(float s, V2 a) { return {s * a.x, s * a.y}; }
inline V2 operator+(float s, V2 a) { return {s + a.x, s + a.y}; }

inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return a < b ? b : a; }
inline float fclamp(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
inline V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline V3 vmax(V3 a, float s) { return {fmax_(a.x, s), fmax_(a.y, s), fmax_(a.z, s)}; }
inline V3 vmin(V3 a, V3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline V3 vclamp(V3 a, float lo, float hi) { return {fclamp(a.x, lo, hi), fclamp(a.y, lo, hi), fclamp(a.z, lo, hi)}; }
inline float fmod_glsl(float x, float y) { return x - y * std::floor(x / y); }
inline V3 vmod(V3 a, float y) { return {fmod_glsl(a.x, y), fmod_glsl(a.y, y), fmod_glsl(a.z, y)}; }
inline V3 vexp(V3 a) { return {std::exp(a.x), std::exp(a.y), std::exp(a.z)}; }

}  // namespace orc
