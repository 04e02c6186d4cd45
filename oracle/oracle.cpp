// oracle.cpp -- CPU restatement of Snelly's shader programs, used only as the
// checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
// It is test infrastructure, never the product path.
//
// Structure follows the reference megakernel line by line (one pixel per
// loop iteration, cameraPath with inline traceRay and shadow rays):
//   shaders/pathtracer-fragment-shader.glsl  (rand 119-130, BSDFs 379-852,
//     traceRay 881-920, normal 922-940, lights 1051-1212,
//     transmittance 1003-1050, directSurfaceLighting 1221-1295,
//     atmosphere 1299-1457, randomwalk_SSS 1468-1593,
//     constructPrimaryRay 1602-1622, cameraPath 1625-1926, pathtrace 1934-1989)
//   shaders/normals-fragment-shader.glsl     (traceDistance 70-113, main 293-345)
//   shaders/ao-fragment-shader.glsl          (Transmittance 154-159, main 395-488)
//   shaders/simplepathtracer-fragment-shader.glsl (traceRay 508-534, fog 570-604,
//     directSurfaceLighting 719-770, SSS 783-882, cameraPath 913-1083, main 1091-1129)
//   shaders/pick-fragment-shader.glsl        (main 133-152)
//   shaders/tonemapper-fragment-shader.glsl  (main 36-97)
// Scenes are hand-written C++ SDFs (no GLSL translation), so the oracle is
// independent of the product's scene compiler.  Scene hooks keep their
// defaults (constant subsurface albedo, ...) except where a scene selects one
// of the hand-written variants (surface albedo, volume emission field).
//
// Parity status of each part is recorded in the tests that pin it; see
// tests/test_oracle.py (closed forms, invariants, brute force).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_COUNT
// the counting build: every float below is a counting value (counting.h)
#include "counting.h"
#define COUNT_SCOPE(c) orccount::Scope count_scope_(orccount::c)
#define COUNT_ADD(field) (++orccount::tc.field)
#else
#define COUNT_SCOPE(c)
#define COUNT_ADD(field)
#endif

#include "vec.h"

using namespace orc;

#ifdef ORACLE_COUNT
namespace orccount {
thread_local Counters tc{};
thread_local int cat = CAT_SHADE;
Counters total{};
// adds the calling thread's counts to the totals (inside each thread of a
// parallel render, after its share of the pixels)
void flush() {
#ifdef _OPENMP
#pragma omp critical(orccount_flush)
#endif
  {
    for (int c = 0; c < CAT_COUNT; ++c) {
      total.flops[c] += tc.flops[c];
      total.trans[c] += tc.trans[c];
      total.ops[c] += tc.ops[c];
    }
    total.steps += tc.steps;
    total.normals += tc.normals;
    total.paths += tc.paths;
    tc = Counters{};
  }
}
}  // namespace orccount
#define COUNT_FLUSH() orccount::flush()
#else
#define COUNT_FLUSH()
#endif

namespace {

const float DENOM_TOLERANCE = 1.0e-7f, PDF_EPSILON = 1.0e-6f, THROUGHPUT_EPSILON = 1.0e-5f,
            RADIANCE_EPSILON = 1.0e-6f;
const float PI_F = 3.141592653589793f;
enum { MAT_INVAL = -1, MAT_DIELE = 0, MAT_METAL = 1, MAT_SURFA = 2 };

// ---------------------------------------------------------------- parameters (the uniforms)
struct Params {
  float resolution[2];
  float camPos[3], camDir[3], camX[3], camY[3];
  float camFovy, camAspect, camAperture, camFocalDistance;
  float filterRadius, radianceClamp, skipProbability, shadowStrength;
  float maxStepsIsMiss, wavelengthSamples;
  float lengthScale, minLengthScale, maxLengthScale;
  float haveEnvMap, envMapVisible, envMapPhiRotation, envMapThetaRotation, envMapTransitionAngle, envMapRotation;
  float skyPower, skyTintUp[3], skyTintDown[3];
  float sunPower, sunAngularSize, sunLatitude, sunLongitude, sunColor[3], sunDir[3], sunVisibleDirectly;
  float sphereLightPosition[3], sphereLightRadius, sphereLightPower, sphereLightColor[3];
  float metalRoughness, metalSpecAlbedoRGB[3];
  float dieleRoughness, dieleAbsorptionRGB[3], dieleSpecAlbedoRGB[3];
  float surfaceDiffuseAlbedoRGB[3], surfaceSpecAlbedoRGB[3], surfaceRoughness, surfaceIor, subsurface;
  float subsurfaceAlbedoRGB[3], subsurfaceMFP, subsurfaceAnisotropy, subsurfaceDiffuseWeight, maxSSSSteps;
  float atmosphereBoundsMin[3], atmosphereBoundsMax[3], atmosphereExtinctionCoeffRGB[3];
  float atmosphereScatteringCoeffRGB[3], atmosphereMFP, atmosphereAnisotropy;
  float fogEnable, fogMaxOpticalDepth, fogMFP, fogEmission[3], fogTint[3];
  float volumeEmission, volumeEmissionColorRGB[3], volumeExtinction;
  float exposure, invGamma, contrast, saturation, hueShift;
};
Params P;

struct Named { const char* name; float* p; int n; };
std::vector<Named> table() {
  return {
      {"resolution", P.resolution, 2}, {"camPos", P.camPos, 3}, {"camDir", P.camDir, 3}, {"camX", P.camX, 3},
      {"camY", P.camY, 3}, {"camFovy", &P.camFovy, 1}, {"camAspect", &P.camAspect, 1},
      {"camAperture", &P.camAperture, 1}, {"camFocalDistance", &P.camFocalDistance, 1},
      {"filterRadius", &P.filterRadius, 1}, {"radianceClamp", &P.radianceClamp, 1},
      {"skipProbability", &P.skipProbability, 1}, {"shadowStrength", &P.shadowStrength, 1},
      {"maxStepsIsMiss", &P.maxStepsIsMiss, 1}, {"wavelengthSamples", &P.wavelengthSamples, 1},
      {"lengthScale", &P.lengthScale, 1}, {"minLengthScale", &P.minLengthScale, 1},
      {"maxLengthScale", &P.maxLengthScale, 1}, {"haveEnvMap", &P.haveEnvMap, 1},
      {"envMapVisible", &P.envMapVisible, 1}, {"envMapPhiRotation", &P.envMapPhiRotation, 1},
      {"envMapThetaRotation", &P.envMapThetaRotation, 1}, {"envMapTransitionAngle", &P.envMapTransitionAngle, 1},
      {"envMapRotation", &P.envMapRotation, 1}, {"skyPower", &P.skyPower, 1}, {"skyTintUp", P.skyTintUp, 3},
      {"skyTintDown", P.skyTintDown, 3}, {"sunPower", &P.sunPower, 1}, {"sunAngularSize", &P.sunAngularSize, 1},
      {"sunLatitude", &P.sunLatitude, 1}, {"sunLongitude", &P.sunLongitude, 1}, {"sunColor", P.sunColor, 3},
      {"sunDir", P.sunDir, 3}, {"sunVisibleDirectly", &P.sunVisibleDirectly, 1},
      {"sphereLightPosition", P.sphereLightPosition, 3}, {"sphereLightRadius", &P.sphereLightRadius, 1},
      {"sphereLightPower", &P.sphereLightPower, 1}, {"sphereLightColor", P.sphereLightColor, 3},
      {"metalRoughness", &P.metalRoughness, 1}, {"metalSpecAlbedoRGB", P.metalSpecAlbedoRGB, 3},
      {"dieleRoughness", &P.dieleRoughness, 1}, {"dieleAbsorptionRGB", P.dieleAbsorptionRGB, 3},
      {"dieleSpecAlbedoRGB", P.dieleSpecAlbedoRGB, 3}, {"surfaceDiffuseAlbedoRGB", P.surfaceDiffuseAlbedoRGB, 3},
      {"surfaceSpecAlbedoRGB", P.surfaceSpecAlbedoRGB, 3}, {"surfaceRoughness", &P.surfaceRoughness, 1},
      {"surfaceIor", &P.surfaceIor, 1}, {"subsurface", &P.subsurface, 1},
      {"subsurfaceAlbedoRGB", P.subsurfaceAlbedoRGB, 3}, {"subsurfaceMFP", &P.subsurfaceMFP, 1},
      {"subsurfaceAnisotropy", &P.subsurfaceAnisotropy, 1},
      {"subsurfaceDiffuseWeight", &P.subsurfaceDiffuseWeight, 1}, {"maxSSSSteps", &P.maxSSSSteps, 1},
      {"exposure", &P.exposure, 1}, {"invGamma", &P.invGamma, 1}, {"contrast", &P.contrast, 1},
      {"saturation", &P.saturation, 1}, {"hueShift", &P.hueShift, 1},
      {"atmosphereBoundsMin", P.atmosphereBoundsMin, 3}, {"atmosphereBoundsMax", P.atmosphereBoundsMax, 3},
      {"atmosphereExtinctionCoeffRGB", P.atmosphereExtinctionCoeffRGB, 3},
      {"atmosphereScatteringCoeffRGB", P.atmosphereScatteringCoeffRGB, 3}, {"atmosphereMFP", &P.atmosphereMFP, 1},
      {"atmosphereAnisotropy", &P.atmosphereAnisotropy, 1}, {"fogEnable", &P.fogEnable, 1},
      {"fogMaxOpticalDepth", &P.fogMaxOpticalDepth, 1}, {"fogMFP", &P.fogMFP, 1},
      {"fogEmission", P.fogEmission, 3}, {"fogTint", P.fogTint, 3},
      {"volumeEmission", &P.volumeEmission, 1}, {"volumeEmissionColorRGB", P.volumeEmissionColorRGB, 3},
      {"volumeExtinction", &P.volumeExtinction, 1}};
}
inline V3 A3(const float* a) { return v3(a[0], a[1], a[2]); }

// ---------------------------------------------------------------- build-time switches (the #defines)
struct Config {
  int dispersive = 0, interactive = 0, maxBounces = 3, maxMarch = 256, spf = 1, sphereLight = 0;
  int atmosphere = 0, fog = 0, maxAtmScatters = 1;  // HAS_ATMOSPHERE, HAS_FOG, __MAX_ATMOSPHERE_SCATTERS__
  int mode = 0;  // Renderer.renderMode: 0 'pt', 1 'ao', 2 'ptsimple' 
  int medium = 0;  // HAS_VOLUME: the scene's SDF_VOLUME medium (kind below)
} C;

// ---------------------------------------------------------------- scene (hand-written SDFs)
enum { SDF_NONE, SDF_SPHERE, SDF_SPHERE_PLANE, SDF_MENGER, SDF_MANDELBOX, SDF_TORUS, SDF_PLANE };
struct SdfSpec { int kind = SDF_NONE; float a = 0, b = 0, c = 0, d = 0; };
SdfSpec sdfs[3];        // indexed by material id (DIELE, METAL, SURFA)
int albedoHook = 0;     // 1: sphere/plane albedo (0.8 above y=-1+1e-3, else 0.5)
int emissionHook = 0;   // HAS_VOLUME_EMISSION; 1: VOLUME_EMISSION = emission where X.y > 0, else emission/4

// SDFs are templates on the scalar type: float for rendering, Counted for the
// algorithmic FLOP count of one evaluation (+ - * / and sqrt count 1 each;
// min/max/abs/floor/clamp are counted separately as non-FLOP VALU operations).
struct Counted {
  float v;
  static long long flops, ops;
};
long long Counted::flops = 0, Counted::ops = 0;
inline Counted cf(float v) { return Counted{v}; }
inline Counted operator+(Counted a, Counted b) { ++Counted::flops; return {a.v + b.v}; }
inline Counted operator-(Counted a, Counted b) { ++Counted::flops; return {a.v - b.v}; }
inline Counted operator*(Counted a, Counted b) { ++Counted::flops; return {a.v * b.v}; }
inline Counted operator/(Counted a, Counted b) { ++Counted::flops; return {a.v / b.v}; }
inline Counted operator-(Counted a) { return {-a.v}; }
inline bool operator>(Counted a, Counted b) { ++Counted::ops; return a.v > b.v; }
inline Counted sqrt_(Counted a) { ++Counted::flops; return {std::sqrt(a.v)}; }
inline Counted floor_(Counted a) { ++Counted::ops; return {std::floor(a.v)}; }
inline Counted abs_(Counted a) { return {std::fabs(a.v)}; }  // a source modifier on gfx9
inline Counted min_(Counted a, Counted b) { ++Counted::ops; return {fmin_(a.v, b.v)}; }
inline Counted max_(Counted a, Counted b) { ++Counted::ops; return {fmax_(a.v, b.v)}; }
inline Counted pow_(Counted a, Counted b) { ++Counted::flops; return {std::pow(a.v, b.v)}; }
inline float sqrt_(float a) { return std::sqrt(a); }
inline float floor_(float a) { return std::floor(a); }
inline float abs_(float a) { return std::fabs(a); }
inline float min_(float a, float b) { return fmin_(a, b); }
inline float max_(float a, float b) { return fmax_(a, b); }
inline float pow_(float a, float b) { return std::pow(a, b); }
template <class T> inline T lit(float x) { return T{x}; }
template <> inline float lit<float>(float x) { return x; }
template <class T> inline T clamp_(T x, float lo, float hi) { return min_(max_(x, lit<T>(lo)), lit<T>(hi)); }

template <class T>
T sdBoxT(const T X[3], float size) {
  T d[3];
  for (int k = 0; k < 3; ++k) d[k] = abs_(X[k]) - lit<T>(size);
  T m = min_(max_(d[0], max_(d[1], d[2])), lit<T>(0.0f));
  T q[3];
  for (int k = 0; k < 3; ++k) q[k] = max_(d[k], lit<T>(0.0f));
  return m + sqrt_(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
}
// exampleScenes/fractals/menger.html:116-134 (without the rotation), size a, scale b, iterations c
template <class T>
T mengerT(const T X[3], float size, float scale, int iter) {
  T sd = sdBoxT(X, size);
  T s = lit<T>(scale);
  for (int i = 0; i < iter; i++) {
    T a[3], r[3];
    for (int k = 0; k < 3; ++k) {
      T xs = X[k] * s;
      a[k] = (xs - lit<T>(2.0f) * floor_(xs / lit<T>(2.0f))) - lit<T>(1.0f);
    }
    s = s * lit<T>(3.0f);
    for (int k = 0; k < 3; ++k) r[k] = abs_(lit<T>(1.0f) - lit<T>(3.0f) * abs_(a[k]));
    T da = max_(r[0], r[1]), db = max_(r[1], r[2]), dc = max_(r[2], r[0]);
    T c = (min_(da, min_(db, dc)) - lit<T>(1.0f)) / s;
    sd = max_(sd, c);
  }
  return sd;
}
// exampleScenes/fractals/mandelbox.html:122-135 (without the rotation); the
// per-scene constants are INIT() work, outside the per-evaluation count
template <class T>
T mandelboxT(const T pos[3], float Scale, int iters, float bailout, float MinRad2) {
  float sc[4], absScalem1, AbsScaleRaisedTo1mIters;
  {
    COUNT_SCOPE(CAT_INIT);
    for (int k = 0; k < 3; ++k) sc[k] = Scale / MinRad2;
    sc[3] = std::fabs(Scale) / MinRad2;
    absScalem1 = std::fabs(Scale - 1.0f);
    AbsScaleRaisedTo1mIters = std::pow(std::fabs(Scale), float(1 - iters));
  }
  T p[4] = {pos[0], pos[1], pos[2], lit<T>(1.0f)}, p0[4] = {pos[0], pos[1], pos[2], lit<T>(1.0f)};
  for (int i = 0; i < iters; i++) {
    for (int k = 0; k < 3; ++k) p[k] = clamp_(p[k], -1.0f, 1.0f) * lit<T>(2.0f) - p[k];
    T r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    T f = clamp_(max_(lit<T>(MinRad2) / r2, lit<T>(MinRad2)), 0.0f, 1.0f);
    for (int k = 0; k < 4; ++k) p[k] = p[k] * f;
    for (int k = 0; k < 4; ++k) p[k] = p[k] * lit<T>(sc[k]) + p0[k];
    if (r2 > lit<T>(bailout)) break;
  }
  return (sqrt_(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - lit<T>(absScalem1)) / p[3] - lit<T>(AbsScaleRaisedTo1mIters);
}
template <class T>
T evalSdfT(const SdfSpec& s, const T X[3]) {
  switch (s.kind) {
    case SDF_SPHERE: return sqrt_(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]) - lit<T>(s.a);
    case SDF_SPHERE_PLANE: return min_(sqrt_(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]) - lit<T>(1.0f), X[1] + lit<T>(1.0f));
    case SDF_MENGER: return mengerT(X, s.a, s.b, (int)s.c);
    case SDF_MANDELBOX: return mandelboxT(X, s.a, (int)s.b, s.c, s.d);
    case SDF_TORUS: {
      T qx = sqrt_(X[0] * X[0] + X[2] * X[2]) - lit<T>(s.a);
      return sqrt_(qx * qx + X[1] * X[1]) - lit<T>(s.b);
    }
    case SDF_PLANE: return X[1] + lit<T>(s.a);  // the plane y = -a
  }
  return lit<T>(1.0e20f);
}
float evalSdf(const SdfSpec& s, V3 X) {
  float x[3] = {X.x, X.y, X.z};
  return evalSdfT<float>(s, x);
}
bool hasMat(int m) { return sdfs[m].kind != SDF_NONE; }
V3 surfaceDiffuseHook(V3 Cc, V3 X) {
  if (albedoHook == 1) return X.y > -1.0f + 1.0e-3f ? v3(0.8f) : v3(0.5f);
  return Cc;
}

// ---------------------------------------------------------------- the medium's scene hooks (hand-written)
// The heterogeneous medium of a scene that defines SDF_VOLUME: its bound
// SDF_VOLUME, VOLUME_EXTINCTION and VOLUME_EXTINCTION_MAX, restated from the
// scene GLSL of the workloads (snelly_amd/scenes.py) with the same float
// operations.  Kinds:
//   1  the config-3 cloud: density max(0, 1 - |p| + 0.3 fbm3(p)) at
//      p = b X, bound |X| < a (a = 5.2, b = 1/4), majorant 1.3
//   2  a linear ball: density max(0, 1 - |X|), bound |X| < 1, majorant 1
//   3  a box of half-extents (a, b, c) of density 1
enum { MEDIUM_NONE, MEDIUM_CLOUD, MEDIUM_LINEAR_BALL, MEDIUM_BOX };
struct MediumSpec { int kind = MEDIUM_NONE; float a = 0, b = 0, c = 0, d = 0; } medium;

// value noise on the integer lattice: corner values from an integer hash
// (nested quadratic maps mod the prime 32749, every product < 2^31), scaled
// by 2^-15 into [0, 1), blended with smoothstep weights
int latticeHash(int x) {
  int h = x % 32749;
  return (h * h + 7919) % 32749;
}
float latticeValue(int x) { return float(latticeHash(x)) * (1.0f / 32768.0f); }
float mixGL(float x, float y, float a) { return x * (1.0f - a) + y * a; }
float valueNoise(V3 p) {
  V3 i = v3(std::floor(p.x), std::floor(p.y), std::floor(p.z));
  V3 f = p - i;
  V3 u = f * f * (3.0f - 2.0f * f);
  int cx = (int)i.x, cy = (int)i.y, cz = (int)i.z;
  int hx0 = latticeHash(cx & 4095), hx1 = latticeHash((cx + 1) & 4095);
  int y0 = cy & 4095, y1 = (cy + 1) & 4095, z0 = cz & 4095, z1 = (cz + 1) & 4095;
  int h00 = latticeHash(hx0 + y0), h10 = latticeHash(hx1 + y0);
  int h01 = latticeHash(hx0 + y1), h11 = latticeHash(hx1 + y1);
  float n000 = latticeValue(h00 + z0), n100 = latticeValue(h10 + z0);
  float n010 = latticeValue(h01 + z0), n110 = latticeValue(h11 + z0);
  float n001 = latticeValue(h00 + z1), n101 = latticeValue(h10 + z1);
  float n011 = latticeValue(h01 + z1), n111 = latticeValue(h11 + z1);
  float x00 = mixGL(n000, n100, u.x), x10 = mixGL(n010, n110, u.x);
  float x01 = mixGL(n001, n101, u.x), x11 = mixGL(n011, n111, u.x);
  return mixGL(mixGL(x00, x10, u.y), mixGL(x01, x11, u.y), u.z);
}
// five octaves of (2 noise - 1), weights 1/2 ... 1/32, each octave at twice
// the frequency of the last plus an offset: |fbm3| < 31/32
float fbm3(V3 p) {
  float sum = 0.0f, amp = 0.5f;
  V3 q = 2.0f * p;
  for (int i = 0; i < 5; i++) {
    sum = sum + amp * (2.0f * valueNoise(q) - 1.0f);
    amp = amp * 0.5f;
    q = 2.0f * q + v3(1.7f, 9.2f, 3.4f);
  }
  return sum;
}
float sdBoxV(V3 X, V3 h) {
  V3 d = vabs(X) - h;
  return fmin_(fmax_(d.x, fmax_(d.y, d.z)), 0.0f) + length(vmax(d, 0.0f));
}
float mediumBoundHook(V3 X) {  // SDF_VOLUME
  switch (medium.kind) {
    case MEDIUM_CLOUD: return length(X) - medium.a;
    case MEDIUM_LINEAR_BALL: return length(X) - 1.0f;
    case MEDIUM_BOX: return sdBoxV(X, v3(medium.a, medium.b, medium.c));
  }
  return 1.0e20f;
}
float mediumExtinctionHook(float e, V3 X) {  // VOLUME_EXTINCTION
  switch (medium.kind) {
    case MEDIUM_CLOUD: {
      V3 p = medium.b * X;
      return e * fmax_(0.0f, 1.0f - length(p) + 0.3f * fbm3(p));
    }
    case MEDIUM_LINEAR_BALL: return e * fmax_(0.0f, 1.0f - length(X));
    case MEDIUM_BOX: return sdBoxV(X, v3(medium.a, medium.b, medium.c)) <= 0.0f ? e : 0.0f;
  }
  return 0.0f;
}
float mediumExtinctionMaxHook(float e) {  // VOLUME_EXTINCTION_MAX
  return medium.kind == MEDIUM_CLOUD ? 1.3f * e : e;
}

// ---------------------------------------------------------------- materials' IOR code
int dieleModel = 1;  // 0 Abbe, 1 Sellmeier (squared poles), 2 Sellmeier2, 3 Polyanskiy, 4 Cauchy
std::vector<float> dieleC;
int metalModel = 0;  // 0 tabulated, 1 analytic n = 0.2 + 0.8t, k = 3.5 - 1.5t
std::vector<float> metalN(64, 1.0f), metalK(64, 0.0f);

float IOR_DIELE(float wl) {
  const std::vector<float>& c = dieleC;
  auto Cn = [&](int i) { return i - 1 < (int)c.size() ? c[i - 1] : 0.0f; };
  float um = 1.0e-3f * wl;
  switch (dieleModel) {
    case 0: {  // js/materials.js:846-861
      float iorVal = Cn(1), abbe = Cn(2);
      if (abbe <= 0.0f) return iorVal;
      float nD2 = 0.5893f; nD2 *= nD2;
      float nF2 = 0.4861f; nF2 *= nF2;
      float nC2 = 0.6563f; nC2 *= nC2;
      float cauchyC = (iorVal - 1.0f) / abbe * (nC2 * nF2) / (nC2 - nF2);
      float cauchyB = iorVal - cauchyC / nD2;
      return cauchyB + cauchyC / (um * um);
    }
    case 1: case 2: {  // js/materials.js:905-930, 956-981
      float l2 = um * um;
      float n2 = 1.0f + Cn(1);
      int terms = ((int)c.size() - 1) / 2;
      for (int t = 1; t <= terms; ++t) {
        float pole = dieleModel == 1 ? Cn(2 * t + 1) * Cn(2 * t + 1) : Cn(2 * t + 1);
        n2 = n2 + Cn(2 * t) * l2 / (l2 - pole);
      }
      return fmax_(std::sqrt(std::fabs(n2)), 1.0e-3f);
    }
    case 3: {  // js/materials.js:1014-1034
      float n2 = Cn(1) + Cn(2) * std::pow(um, Cn(3)) / (um * um - std::pow(Cn(4), Cn(5)));
      return fmax_(std::sqrt(std::fabs(n2)), 1.0e-3f);
    }
    case 4: {  // js/materials.js:1064-1089
      float n = Cn(1);
      int terms = ((int)c.size() - 1) / 2;
      for (int t = 1; t <= terms; ++t) n = n + Cn(2 * t) * std::pow(um, Cn(2 * t + 1));
      return fmax_(n, 1.0e-3f);
    }
  }
  return 1.5f;
}
// GL_LINEAR, CLAMP_TO_EDGE lookup of a 1-D float table at texel centres (i+0.5)/N
float lerpTable(const float* t, int n, float u) {
  float f = u * float(n) - 0.5f;
  float fl = std::floor(f);
  float a = f - fl;
  int i0 = (int)fl, i1 = i0 + 1;
  i0 = i0 < 0 ? 0 : (i0 >= n ? n - 1 : i0);
  i1 = i1 < 0 ? 0 : (i1 >= n ? n - 1 : i1);
  return t[i0] * (1.0f - a) + t[i1] * a;
}
float IOR_METAL(float wl) {
  if (metalModel == 1) { float t = (wl - 390.0f) / 360.0f; return 0.2f + 0.8f * t; }
  return lerpTable(metalN.data(), 64, (wl - 390.0f) * float(1.0 / (750.0 - 390.0)));
}
float K_METAL_table(float wl) {
  if (metalModel == 1) { float t = (wl - 390.0f) / 360.0f; return 3.5f - 1.5f * t; }
  return lerpTable(metalK.data(), 64, (wl - 390.0f) * float(1.0 / (750.0 - 390.0)));
}
// TabulatedMetal's K_METAL samples the sampler kTex, which the reference never
// assigns a unit (it binds the k texture as "k_tex", js/materials.js:664-665),
// so it reads unit 0: the radiance texture being read this frame
// (js/renderer.js:590, 706-713), NEAREST and CLAMP_TO_EDGE, red channel.
std::vector<float> unit0Radiance;
int unit0W = 0, unit0H = 0;
float K_METAL(float wl) {
  if (metalModel == 1) return K_METAL_table(wl);
  float u = (wl - 390.0f) * float(1.0 / (750.0 - 390.0));
  if (unit0W <= 0) return 0.0f;
  int i = (int)std::floor(u * float(unit0W)), j = (int)std::floor(0.5f * float(unit0H));
  i = i < 0 ? 0 : (i >= unit0W ? unit0W - 1 : i);
  j = j < 0 ? 0 : (j >= unit0H ? unit0H - 1 : j);
  return unit0Radiance[4 * ((size_t)j * unit0W + i)];
}

std::vector<float> lutXYZ(4096, 0.0f), lutICDF(4096, 0.0f);
// the environment map: linear RGB texels, row 0 first, sampled with GL LINEAR
// filtering and CLAMP_TO_EDGE (the texture GLU.loadImageAndCreateTextureInfo makes)
std::vector<float> envTex;
int envW = 0, envH = 0;
V3 envTexel(int i, int j) {
  i = i < 0 ? 0 : (i >= envW ? envW - 1 : i);
  j = j < 0 ? 0 : (j >= envH ? envH - 1 : j);
  const float* t = &envTex[3 * ((size_t)j * envW + i)];
  return v3(t[0], t[1], t[2]);
}
V3 mixV(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
V3 sampleEnv(float u, float v) {
  float fu = u * float(envW) - 0.5f, fv = v * float(envH) - 0.5f;
  float iu = std::floor(fu), iv = std::floor(fv);
  float au = fu - iu, av = fv - iv;
  int i0 = (int)iu, j0 = (int)iv;
  return mixV(mixV(envTexel(i0, j0), envTexel(i0 + 1, j0), au), mixV(envTexel(i0, j0 + 1), envTexel(i0 + 1, j0 + 1), au), av);
}
V3 lookupXYZ(float w) {
  float f = w * 1024.0f - 0.5f, fl = std::floor(f), a = f - fl;
  int i0 = (int)fl, i1 = i0 + 1;
  i0 = i0 < 0 ? 0 : (i0 > 1023 ? 1023 : i0);
  i1 = i1 < 0 ? 0 : (i1 > 1023 ? 1023 : i1);
  const float* t0 = &lutXYZ[4 * i0];
  const float* t1 = &lutXYZ[4 * i1];
  return v3(t0[0] * (1.0f - a) + t1[0] * a, t0[1] * (1.0f - a) + t1[1] * a, t0[2] * (1.0f - a) + t1[2] * a);
}

// ---------------------------------------------------------------- RNG (pathtracer-fragment-shader.glsl:119-130)
inline float randf(float r[4]) {
  static const float q[4] = {1225.0f, 1585.0f, 2457.0f, 2098.0f};
  static const float rr[4] = {1112.0f, 367.0f, 92.0f, 265.0f};
  static const float a[4] = {3423.0f, 2646.0f, 1707.0f, 1999.0f};
  static const float m[4] = {4194287.0f, 4194277.0f, 4194191.0f, 4194167.0f};
  float s[4];
  for (int k = 0; k < 4; ++k) {
    float beta = std::floor(r[k] / q[k]);
    float p = a[k] * (r[k] - beta * q[k]) - beta * rr[k];
    beta = (1.0f - fsign(p)) * 0.5f * m[k];
    r[k] = p + beta;
    s[k] = r[k] / m[k];
  }
  float d = s[0] * 1.0f + s[1] * -1.0f + s[2] * 1.0f + s[3] * -1.0f;
  return d - std::floor(d);
}

// ---------------------------------------------------------------- radiance type helpers
inline float rt(float x, float) { return x; }
inline V3 rt(float x, V3) { return v3(x); }
inline float maxC(float f) { return f; }
inline float maxC(V3 v) { return fmax_(fmax_(v.x, v.y), v.z); }
inline float avgC(float f) { return f; }
inline float avgC(V3 v) { return (v.x + v.y + v.z) / 3.0f; }
inline float minR(float a, float b) { return fmin_(a, b); }
inline V3 minR(V3 a, V3 b) { return vmin(a, b); }
inline float expR(float a) { return std::exp(a); }
inline V3 expR(V3 a) { return vexp(a); }
inline float lenR(float a) { return std::fabs(a); }
inline float lenR(V3 a) { return length(a); }
inline float absR(float a) { return std::fabs(a); }
inline V3 absR(V3 a) { return vabs(a); }
inline float mixR(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline V3 mixR(V3 x, V3 y, float a) { return x * (1.0f - a) + y * a; }
inline float albedo(V3 RGB, V3 rgb, float) { return dot(RGB, rgb); }  // dispersive rgbToAlbedo
inline V3 albedo(V3 RGB, V3, V3) { return RGB; }

// ---------------------------------------------------------------- basis, colour (152-293)
struct Basis { V3 nW, tW, bW; };
V3 safe_normalize(V3 N) { float l = length(N); return N / fmax_(l, DENOM_TOLERANCE); }
Basis makeBasis(V3 nW, bool safe = true) {
  Basis b;
  b.nW = nW;
  if (std::fabs(nW.z) < std::fabs(nW.x)) b.tW = v3(nW.z, 0.0f, -nW.x);
  else b.tW = v3(0.0f, nW.z, -nW.y);
  b.tW = safe ? safe_normalize(b.tW) : normalize(b.tW);
  b.bW = cross(nW, b.tW);
  return b;
}
V3 worldToLocal(V3 v, const Basis& b) { return v3(dot(v, b.tW), dot(v, b.bW), dot(v, b.nW)); }
V3 localToWorld(V3 v, const Basis& b) { return b.tW * v.x + b.bW * v.y + b.nW * v.z; }
float cosTheta2(V3 n) { return n.z * n.z; }
float tanTheta2(V3 n) { float ct2 = cosTheta2(n); return fmax_(0.0f, 1.0f - ct2) / fmax_(ct2, 1.0e-7f); }
float tanTheta(V3 n) { return std::sqrt(fmax_(0.0f, tanTheta2(n))); }
V3 xyzToRgb(V3 XYZ) {
  V3 RGB;
  RGB.x = 3.2404542f * XYZ.x - 1.5371385f * XYZ.y - 0.4985314f * XYZ.z;
  RGB.y = -0.9692660f * XYZ.x + 1.8760108f * XYZ.y + 0.0415560f * XYZ.z;
  RGB.z = 0.0556434f * XYZ.x - 0.2040259f * XYZ.y + 1.0572252f * XYZ.z;
  float delta = -fmin_(0.0f, fmin_(fmin_(RGB.x, RGB.y), RGB.z));
  RGB.x += delta; RGB.y += delta; RGB.z += delta;
  float sum = RGB.x + RGB.y + RGB.z;
  return RGB / sum;
}
V3 rgbToXyz(V3 RGB) {
  return v3(0.4124564f * RGB.x + 0.3575761f * RGB.y + 0.1804375f * RGB.z,
            0.2126729f * RGB.x + 0.7151522f * RGB.y + 0.0721750f * RGB.z,
            0.0193339f * RGB.x + 0.1191920f * RGB.y + 0.9503041f * RGB.z);
}

// ---------------------------------------------------------------- sampling (300-372)
V3 sampleSphereUniformly(float* rnd) {
  float z = 1.0f - 2.0f * randf(rnd);
  float phi = 2.0f * PI_F * randf(rnd);
  return v3(std::cos(phi), std::sin(phi), z);  // as the reference: x, y not scaled by sqrt(1 - z^2)
}
float phaseFunction(float mu, float g) {
  float gSqr = g * g;
  return (1.0f / (4.0f * PI_F)) * (1.0f - gSqr) / std::pow(1.0f - 2.0f * g * mu + gSqr, 1.5f);
}
V3 sampleHemisphereUniformly(float* rnd) {
  float xi = randf(rnd);
  float r = std::sqrt(1.0f - xi * xi);
  float phi = 2.0f * PI_F * randf(rnd);
  return v3(r * std::cos(phi), r * std::sin(phi), xi);
}
V3 sampleHemisphereCosineWeighted(float* rnd, float& pdf) {
  float r = std::sqrt(randf(rnd));
  float theta = 2.0f * PI_F * randf(rnd);
  float x = r * std::cos(theta), y = r * std::sin(theta);
  float z = std::sqrt(fmax_(0.0f, 1.0f - x * x - y * y));
  pdf = std::fabs(z) / PI_F;
  return v3(x, y, z);
}
float pdfHemisphereCosineWeighted(V3 w) { return w.z < 0.0f ? 0.0f : w.z / PI_F; }
float powerHeuristic(float a, float b) { return a / (a + b); }
V3 samplePhaseFunction(V3 rayDir, float g, float* rnd) {
  float costheta;
  if (std::fabs(g) < 1.0e-3f) costheta = 1.0f - 2.0f * randf(rnd);
  else costheta = 1.0f / (2.0f * g) * (1.0f + g * g - ((1.0f - g * g) * (1.0f - g + 2.0f * g * randf(rnd))));
  float sintheta = std::sqrt(fmax_(0.0f, 1.0f - costheta * costheta));
  float phi = 2.0f * PI_F * randf(rnd);
  Basis b = makeBasis(rayDir);
  return costheta * rayDir + sintheta * (std::cos(phi) * b.tW + std::sin(phi) * b.bW);
}
// Henyey-Greenstein sampling by inversion of its CDF in cos(theta) (Pharr,
// Jakob & Humphreys, Physically Based Rendering, 3rd ed., 15.2.3): the
// heterogeneous medium's phase sampling; draws cos(theta), then the azimuth
V3 sampleHenyeyGreenstein(V3 dir, float g, float* rnd, float& pdf) {
  float xi = randf(rnd);
  float cosTheta;
  if (std::fabs(g) < 1.0e-3f) cosTheta = 1.0f - 2.0f * xi;
  else {
    float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * xi);
    cosTheta = (1.0f + g * g - sq * sq) / (2.0f * g);
  }
  cosTheta = fclamp(cosTheta, -1.0f, 1.0f);
  float sinTheta = std::sqrt(fmax_(0.0f, 1.0f - cosTheta * cosTheta));
  float phi = 2.0f * PI_F * randf(rnd);
  Basis b = makeBasis(dir);
  pdf = phaseFunction(cosTheta, g);
  return localToWorld(v3(sinTheta * std::cos(phi), sinTheta * std::sin(phi), cosTheta), b);
}
// a uniformly distributed unit direction (the medium's sky sample)
V3 sampleSphereUnit(float* rnd) {
  float z = 1.0f - 2.0f * randf(rnd);
  float r = std::sqrt(fmax_(0.0f, 1.0f - z * z));
  float phi = 2.0f * PI_F * randf(rnd);
  return v3(r * std::cos(phi), r * std::sin(phi), z);
}

// ---------------------------------------------------------------- microfacet + Fresnel (379-441)
float microfacetEval(V3 m, float roughness) {
  float t2 = tanTheta2(m), c2 = cosTheta2(m);
  float rs = roughness * roughness;
  float exponent = t2 / fmax_(rs, 1.0e-9f);
  return std::exp(-exponent) / fmax_(PI_F * rs * c2 * c2, 1.0e-9f);
}
V3 microfacetSample(float* rnd, float roughness) {
  float phiM = (2.0f * PI_F) * randf(rnd);
  float cosPhiM = std::cos(phiM), sinPhiM = std::sin(phiM);
  float tanThetaMSqr = -roughness * roughness * std::log(fmax_(1.0e-9f, randf(rnd)));
  float cosThetaM = 1.0f / std::sqrt(1.0f + tanThetaMSqr);
  float sinThetaM = std::sqrt(fmax_(0.0f, 1.0f - cosThetaM * cosThetaM));
  return safe_normalize(v3(sinThetaM * cosPhiM, sinThetaM * sinPhiM, cosThetaM));
}
float microfacetPDF(V3 m, float roughness) { return microfacetEval(m, roughness) * std::fabs(m.z); }
float smithG1(V3 v, V3 m, float roughness) {
  float tanThetaAbs = std::fabs(tanTheta(v));
  if (tanThetaAbs < 1.0e-6f) return 1.0f;
  if (dot(v, m) * v.z <= 0.0f) return 0.0f;
  float a = 1.0f / (fmax_(roughness, 1.0e-6f) * tanThetaAbs);
  if (a >= 1.6f) return 1.0f;
  float aSqr = a * a;
  return (3.535f * a + 2.181f * aSqr) / (1.0f + 2.276f * a + 2.577f * aSqr);
}
float smithG2(V3 wo, V3 wi, V3 m, float r) { return smithG1(wo, m, r) * smithG1(wi, m, r); }
float fresnelDielectricReflectance(float c, float eta_ti) {
  float g = eta_ti * eta_ti - 1.0f + c * c;
  if (g > 0.0f) {
    g = std::sqrt(g);
    float A = (g - c) / (g + c);
    float B = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * A * A * (1.0f + B * B);
  }
  return 1.0f;
}
float fresnelMetalReflectance(float cosi, float ior, float k) {
  float cosip = std::fabs(cosi), cosi2 = cosip * cosip;
  float iikk = ior * ior + k * k, iikkc = iikk * cosi2, twoEtaCosi = 2.0f * ior * cosip;
  float Rparl2 = (iikkc - twoEtaCosi + 1.0f) / (iikkc + twoEtaCosi + 1.0f);
  float Rperp2 = (iikk - twoEtaCosi + cosi2) / (iikk + twoEtaCosi + cosi2);
  return 0.5f * (Rparl2 + Rperp2);
}
bool refractTransmitted(V3 n, float eta, V3 wt, V3& wi) {
  float wtn = dot(wt, n);
  float disc = 1.0f - eta * eta * (1.0f - wtn * wtn);
  if (disc < 0.0f) return false;
  wi = eta * wt - n * fsign(wtn) * (eta * std::fabs(wtn) - std::sqrt(disc));
  return true;
}

// ---------------------------------------------------------------- the integrator, templated on RadianceType
template <class R>
struct Tracer {
  V3 rgb;
  float wl;

  R A(V3 RGB) const { return albedo(RGB, rgb, R()); }
  R one() const { return rt(1.0f, R()); }
  R zero() const { return rt(0.0f, R()); }

  bool traceRay(V3 start, V3 dir, V3& hit, int& material, float maxMarchDist) const {
    COUNT_SCOPE(CAT_MARCH);
    material = MAT_INVAL;
    float minMarch = P.minLengthScale;
    float sdf = 1.0e20f;
    COUNT_ADD(steps);
    for (int m : {MAT_SURFA, MAT_METAL, MAT_DIELE}) if (hasMat(m)) sdf = fmin_(sdf, std::fabs(evalSdf(sdfs[m], start)));
    float InitialSign = fsign(sdf);
    float t = InitialSign * sdf;
    if (std::fabs(t) >= maxMarchDist) return false;
    for (int n = 0; n < C.maxMarch; n++) {
      COUNT_ADD(steps);
      V3 pW = start + t * dir;
      sdf = 1.0e20f;
      for (int m : {MAT_SURFA, MAT_METAL, MAT_DIELE}) {
        if (!hasMat(m)) continue;
        float s = std::fabs(evalSdf(sdfs[m], pW));
        if (s < minMarch) { material = m; hit = start + t * dir; return true; }
        sdf = fmin_(sdf, s);
      }
      t += InitialSign * sdf;
      if (std::fabs(t) >= maxMarchDist) return false;
    }
    hit = start + t * dir;
    return P.maxStepsIsMiss == 0.0f;
  }
  V3 normal(V3 pW, int m, bool safe = true) const {
    COUNT_SCOPE(CAT_NORMAL);
    COUNT_ADD(normals);
    float e = 2.0f * P.minLengthScale;
    const SdfSpec& s = sdfs[m];
    V3 N = v3(evalSdf(s, pW + v3(e, 0, 0)) - evalSdf(s, pW - v3(e, 0, 0)),
              evalSdf(s, pW + v3(0, e, 0)) - evalSdf(s, pW - v3(0, e, 0)),
              evalSdf(s, pW + v3(0, 0, e)) - evalSdf(s, pW - v3(0, 0, e)));
    return safe ? safe_normalize(N) : normalize(N);
  }

  // Dielectric (477-633)
  R evaluateDielectric(const Basis& basis, V3 wi, V3 wo) const {
    float ior = IOR_DIELE(wl);
    bool reflected = wo.z * wi.z > 0.0f;
    R alb = A(A3(P.dieleSpecAlbedoRGB));
    V3 bI = -wo, bO = wi;  // fromCamera
    bool entering = bI.z < 0.0f;
    V3 h;
    if (reflected) h = safe_normalize(wo + wi);
    else if (entering) h = safe_normalize(bI - ior * bO);
    else h = safe_normalize(ior * bI - bO);
    if (h.z < 0.0f) h = h * -1.0f;
    float roughness = P.dieleRoughness;
    float D = microfacetEval(h, roughness);
    float G = smithG2(-bI, bO, h, roughness);
    float eta = entering ? ior : 1.0f / ior;
    R Fr = alb * fresnelDielectricReflectance(std::fabs(wi.z), eta);
    if (reflected) return alb * Fr * D * G / fmax_(4.0f * std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
    float im = dot(-bI, h), om = dot(bO, h);
    float sd = im + eta * om;
    float dwh = eta * eta * std::fabs(om) / fmax_(sd * sd, DENOM_TOLERANCE);
    return (one() - Fr) * G * D * std::fabs(im) * dwh / fmax_(std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
  }
  float pdfDielectric(const Basis&, V3 wi, V3 wo) const {
    float ior = IOR_DIELE(wl);
    bool reflected = wo.z * wi.z > 0.0f;
    R alb = A(A3(P.dieleSpecAlbedoRGB));
    V3 bI = -wo, bO = wi;
    bool entering = bI.z < 0.0f;
    float eta = entering ? ior : 1.0f / ior;
    R Fr = alb * fresnelDielectricReflectance(std::fabs(wi.z), eta);
    V3 h;
    float dwh, pdf;
    if (reflected) {
      h = safe_normalize(wo + wi);
      dwh = 1.0f / fmax_(4.0f * std::fabs(dot(wi, h)), DENOM_TOLERANCE);
      pdf = avgC(Fr);
    } else {
      if (entering) h = safe_normalize(bI - ior * bO);
      else h = safe_normalize(ior * bI - bO);
      if (h.z < 0.0f) h = h * -1.0f;
      float im = dot(-bI, h), om = dot(bO, h);
      float sd = im + eta * om;
      dwh = eta * eta * std::fabs(om) / fmax_(sd * sd, DENOM_TOLERANCE);
      pdf = 1.0f - avgC(Fr);
    }
    pdf *= microfacetPDF(h, P.dieleRoughness);
    return std::fabs(pdf * dwh);
  }
  R sampleDielectric(const Basis&, V3 wi, V3& wo, float& pdfOut, float* rnd) const {
    float ior = IOR_DIELE(wl);
    R alb = A(A3(P.dieleSpecAlbedoRGB));
    float etaRefl = (wi.z >= 0.0f) ? ior : 1.0f / ior;
    R Fr = alb * fresnelDielectricReflectance(std::fabs(wi.z), etaRefl);
    float roughness = P.dieleRoughness;
    V3 m = microfacetSample(rnd, roughness);
    float microPDF = microfacetPDF(m, roughness);
    float reflectProb = avgC(Fr);
    if (randf(rnd) < reflectProb) {
      wo = -wi + 2.0f * dot(wi, m) * m;
      if (wo.z * wi.z < 0.0f) wo = wo * -1.0f;
      float D = microfacetEval(m, roughness), G = smithG2(wi, wo, m, roughness);
      R f = Fr * D * G / fmax_(4.0f * std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
      float dwh = 1.0f / fmax_(std::fabs(4.0f * dot(wi, m)), DENOM_TOLERANCE);
      pdfOut = microPDF * reflectProb * dwh;
      return f;
    }
    V3 bI = -wo, bO = wi;  // fromCamera: beamIncident from (uninitialised) woutputL, then overwritten
    bool entering = bI.z < 0.0f;
    float eta = entering ? ior : 1.0f / ior;
    if (!refractTransmitted(m, eta, bO, bI)) return zero();
    wo = -bI;
    float im = dot(-bI, m);
    R Frm = rt(fresnelDielectricReflectance(std::fabs(im), eta), R());
    R Trm = one() - alb * Frm;
    float D = microfacetEval(m, roughness), G = smithG2(-bI, bO, m, roughness);
    float om = dot(bO, m);
    float sd = im + eta * om;
    float dwh = eta * eta * std::fabs(om) / fmax_(sd * sd, DENOM_TOLERANCE);
    R f = (std::fabs(im) * dwh * Trm * G * D) / fmax_(std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
    pdfOut = (1.0f - reflectProb) * microPDF * std::fabs(dwh);
    return f;
  }

  // Metal (643-708)
  R evaluateMetal(V3 wi, V3 wo) const {
    if (wi.z < 0.0f) return zero();
    float Fr = fresnelMetalReflectance(wi.z, IOR_METAL(wl), K_METAL(wl));
    V3 h = normalize(wo + wi);
    float r = P.metalRoughness;
    float D = microfacetEval(h, r), G = smithG2(wi, wo, h, r);
    R spec = A(A3(P.metalSpecAlbedoRGB));
    return spec * Fr * D * G / fmax_(4.0f * std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
  }
  float pdfMetal(V3 wi, V3 wo) const {
    if (wi.z < 0.0f) return PDF_EPSILON;
    V3 h = safe_normalize(wo + wi);
    float dwh = 1.0f / fmax_(std::fabs(4.0f * dot(wi, h)), DENOM_TOLERANCE);
    return microfacetPDF(h, P.metalRoughness) * dwh;
  }
  R sampleMetal(V3 wi, V3& wo, float& pdfOut, float* rnd) const {
    if (wi.z < 0.0f) return zero();
    float Fr = fresnelMetalReflectance(wi.z, IOR_METAL(wl), K_METAL(wl));
    float r = P.metalRoughness;
    V3 m = microfacetSample(rnd, r);
    wo = -wi + 2.0f * dot(wi, m) * m;
    if (wo.z < DENOM_TOLERANCE) wo.z *= -1.0f;
    float D = microfacetEval(m, r), G = smithG2(wi, wo, m, r);
    R spec = A(A3(P.metalSpecAlbedoRGB));
    R f = spec * Fr * D * G / fmax_(4.0f * std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
    float dwh = 1.0f / fmax_(std::fabs(4.0f * dot(wi, m)), DENOM_TOLERANCE);
    pdfOut = microfacetPDF(m, r) * dwh;
    return f;
  }

  // Surface (716-807)
  R diffuseRefl(V3 X) const { return A(surfaceDiffuseHook(A3(P.surfaceDiffuseAlbedoRGB), X)); }
  R evaluateSurface(V3 X, V3 wi, V3 wo) const {
    if (wi.z < 0.0f) return zero();
    R diffuseAlbedo = (1.0f - P.subsurface) * diffuseRefl(X);
    R specAlbedo = A(A3(P.surfaceSpecAlbedoRGB));
    float ior = P.surfaceIor, r = P.surfaceRoughness;
    float Fr = fresnelDielectricReflectance(wo.z, ior);
    V3 h = normalize(wo + wi);
    float D = microfacetEval(h, r), G = smithG2(wi, wo, h, r);
    R f = Fr * specAlbedo * D * G / fmax_(4.0f * std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
    float E = fresnelDielectricReflectance(wi.z, ior);
    f = f + (1.0f - E) * diffuseAlbedo / PI_F;
    return f;
  }
  float pdfSurface(V3 X, V3 wi, V3 wo) const {
    if (wi.z < 0.0f) return PDF_EPSILON;
    R diffuseAlbedo = (1.0f - P.subsurface) * diffuseRefl(X);
    R specAlbedo = A(A3(P.surfaceSpecAlbedoRGB));
    float E = fresnelDielectricReflectance(std::fabs(wi.z), P.surfaceIor);
    float specW = E * avgC(specAlbedo), diffW = (1.0f - E) * avgC(diffuseAlbedo);
    float specProb = specW / fmax_(specW + diffW, DENOM_TOLERANCE);
    float diffusePdf = pdfHemisphereCosineWeighted(wo);
    V3 h = safe_normalize(wo + wi);
    float dwh = 1.0f / fmax_(std::fabs(4.0f * dot(wi, h)), DENOM_TOLERANCE);
    float specularPdf = microfacetPDF(h, P.surfaceRoughness) * dwh;
    return specProb * specularPdf + (1.0f - specProb) * diffusePdf;
  }
  R sampleSurface(V3 X, V3 wi, V3& wo, float& pdfOut, float* rnd) const {
    if (wi.z < 0.0f) return zero();
    R diffuseAlbedo = (1.0f - P.subsurface) * diffuseRefl(X);
    R specAlbedo = A(A3(P.surfaceSpecAlbedoRGB));
    float ior = P.surfaceIor, r = P.surfaceRoughness;
    float E = fresnelDielectricReflectance(std::fabs(wi.z), ior);
    float specW = E * avgC(specAlbedo), diffW = (1.0f - E) * avgC(diffuseAlbedo);
    float specProb = specW / fmax_(specW + diffW, DENOM_TOLERANCE);
    if (randf(rnd) >= specProb) {
      wo = sampleHemisphereCosineWeighted(rnd, pdfOut);
      pdfOut *= (1.0f - specProb);
      return (1.0f - E) * diffuseAlbedo / PI_F;
    }
    V3 m = microfacetSample(rnd, r);
    wo = -wi + 2.0f * dot(wi, m) * m;
    pdfOut = specProb;
    if (wo.z < DENOM_TOLERANCE) return zero();
    float D = microfacetEval(m, r), G = smithG2(wi, wo, m, r);
    float Fr = fresnelDielectricReflectance(std::fabs(wi.z), ior);
    R f = Fr * specAlbedo * D * G / fmax_(4.0f * std::fabs(wo.z) * std::fabs(wi.z), DENOM_TOLERANCE);
    float dwh = 1.0f / fmax_(std::fabs(4.0f * dot(wi, m)), DENOM_TOLERANCE);
    pdfOut *= microfacetPDF(m, r) * dwh;
    return f;
  }

  R evaluateBsdf(V3 X, const Basis& b, V3 wi, V3 wo, int m) const {
    if (m == MAT_SURFA) return evaluateSurface(X, wi, wo);
    if (m == MAT_METAL) return evaluateMetal(wi, wo);
    if (m == MAT_DIELE) return evaluateDielectric(b, wi, wo);
    return zero();
  }
  float pdfBsdf(V3 X, const Basis& b, V3 wi, V3 wo, int m) const {
    if (m == MAT_SURFA) return pdfSurface(X, wi, wo);
    if (m == MAT_METAL) return pdfMetal(wi, wo);
    if (m == MAT_DIELE) return pdfDielectric(b, wi, wo);
    return PDF_EPSILON;
  }
  R sampleBsdf(V3 X, const Basis& b, V3 wi, int m, V3& wo, float& pdf, float* rnd) const {
    if (m == MAT_SURFA) return sampleSurface(X, wi, wo, pdf, rnd);
    if (m == MAT_METAL) return sampleMetal(wi, wo, pdf, rnd);
    if (m == MAT_DIELE) return sampleDielectric(b, wi, wo, pdf, rnd);
    pdf = 1.0f;
    return zero();
  }

  // lights (1051-1212)
  R environmentRadiance(V3 dirW) const {  // environmentRadianceRGB (1051-1070)
    float rot_phi = PI_F * P.envMapPhiRotation / 180.0f;
    float rot_theta = PI_F * P.envMapThetaRotation / 180.0f;
    V3 pole = v3(std::sin(rot_theta), std::cos(rot_theta), 0.0f);
    V3 RGB = v3(1.0f);
    if (P.haveEnvMap != 0.0f) {
      Basis sb = makeBasis(pole);
      V3 dirL = worldToLocal(dirW, sb);
      float phi = std::atan2(dirL.y, dirL.x) + PI_F + rot_phi;
      phi -= 2.0f * PI_F * std::floor(phi / (2.0f * PI_F));
      float theta = std::acos(dirL.z);
      RGB = sampleEnv(phi / (2.0f * PI_F), theta / PI_F);
    }
    float t = dot(dirW, pole);
    float tt = P.envMapTransitionAngle / 180.0f;
    float e0 = -tt, e1 = tt;
    float s = fclamp((t - e0) / (e1 - e0), 0.0f, 1.0f);
    s = s * s * (3.0f - 2.0f * s);
    V3 tint = A3(P.skyTintDown) * (1.0f - s) + A3(P.skyTintUp) * s;
    RGB = RGB * (P.skyPower * tint);
    return A(RGB);
  }
  V3 sampleSunDir(float* rnd, float& pdfDir) const {
    float theta_max = P.sunAngularSize * PI_F / 180.0f;
    float theta = theta_max * std::sqrt(randf(rnd));
    float ct = std::cos(theta), st = std::sqrt(fmax_(0.0f, 1.0f - ct * ct));
    float phi = 2.0f * PI_F * randf(rnd);
    float solid_angle = 2.0f * PI_F * (1.0f - std::cos(theta_max));
    pdfDir = 1.0f / solid_angle;
    return localToWorld(v3(st * std::cos(phi), st * std::sin(phi), ct), makeBasis(A3(P.sunDir)));
  }
  R sunRadiance(V3 dir) const {
    float theta_max = P.sunAngularSize * PI_F / 180.0f;
    if (dot(dir, A3(P.sunDir)) < std::cos(theta_max)) return zero();
    return A(P.sunPower * A3(P.sunColor));
  }
  R sphereLightRadiance(V3 pW, V3 dir, V3& pHit) const {
    V3 m = pW - A3(P.sphereLightPosition);
    float b = dot(m, dir);
    float c = dot(m, m) - P.sphereLightRadius * P.sphereLightRadius;
    if (c > 0.0f && b > 0.0f) return zero();
    float discr = b * b - c;
    if (discr < 0.0f) return zero();
    float t = fmax_(-b - std::sqrt(discr), 0.0f);
    pHit = pW + t * dir;
    return A(P.sphereLightPower * A3(P.sphereLightColor));
  }
  R sampleSphereLight(float* rnd, V3 pW, V3& pHit, float& pdfDir) const {
    if (P.sphereLightPower < RADIANCE_EPSILON) return zero();
    V3 cp = normalize(pW - A3(P.sphereLightPosition));
    Basis hb = makeBasis(cp);
    V3 vL = sampleHemisphereUniformly(rnd);
    V3 vW = localToWorld(vL, hb);
    float pdfArea = 1.0f / (2.0f * PI_F * P.sphereLightRadius * P.sphereLightRadius);
    pHit = A3(P.sphereLightPosition) + P.sphereLightRadius * vW;
    V3 dHit = pHit - pW;
    float dHit2 = dot(dHit, dHit);
    dHit = normalize(dHit);
    V3 nHit = normalize(pHit - A3(P.sphereLightPosition));
    float costheta = std::fabs(dot(nHit, dHit));
    pdfDir = pdfArea * dHit2 / fmax_(DENOM_TOLERANCE, costheta);
    return A(P.sphereLightPower * A3(P.sphereLightColor));
  }

  // homogeneous atmosphere and fog (856-868, 975-1050)
  R VOLUME_SCATTERING_EVAL() const { return A(A3(P.atmosphereScatteringCoeffRGB) / (P.atmosphereMFP * P.lengthScale)); }
  R VOLUME_EXTINCTION_EVAL() const { return A(A3(P.atmosphereExtinctionCoeffRGB) / (P.atmosphereMFP * P.lengthScale)); }
  R VOLUME_EMISSION_EVAL(V3 X) const {
    V3 e = A3(P.volumeEmissionColorRGB) * P.volumeEmission;
    return A(X.y > 0.0f ? e : e * 0.25f);
  }
  bool atmosphereHit(V3 pos, V3 dir, float& t0, float& t1) const {
    V3 dL = v3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    V3 lo = (A3(P.atmosphereBoundsMin) * P.lengthScale - pos) * dL;
    V3 hi = (A3(P.atmosphereBoundsMax) * P.lengthScale - pos) * dL;
    V3 tmp = vmin(lo, hi);
    hi = lo + hi - tmp;
    lo = tmp;
    bool hit = !(lo.x > hi.y || lo.y > hi.x || lo.x > hi.z || lo.z > hi.x || lo.y > hi.z || lo.z > hi.y);
    t0 = fmax_(fmax_(lo.x, lo.y), lo.z);
    t1 = fmin_(fmin_(hi.x, hi.y), hi.z);
    return hit;
  }
  bool atmosphereSegment(V3 pW, V3 dir, float len, float& t0, float& t1) const {
    if (!atmosphereHit(pW, dir, t0, t1)) return false;
    t0 = fmax_(0.0f, t0);
    if (t1 < 0.0f || len < t0) return false;
    t1 = fmin_(len, t1);
    return true;
  }
  R transmittanceOverFreeSegment(V3 pW, V3 dir, float len) const {
    R Tr = one();
    if (C.atmosphere) {
      R ext = VOLUME_EXTINCTION_EVAL();
      if (lenR(ext) == 0.0f) return one();
      float t0 = 0.0f, t1 = 0.0f;
      if (!atmosphereSegment(pW, dir, len, t0, t1)) return one();
      Tr = expR(-((t1 - t0) * ext));
    }
    if (C.fog && P.fogEnable != 0.0f) {
      V3 fogAbsorption = v3(1.0f) - A3(P.fogTint);
      V3 depth = P.fogMaxOpticalDepth * (v3(1.0f) - vexp(-(len * fogAbsorption / P.fogMFP)));
      Tr = Tr * expR(-A(depth));
    }
    return Tr;
  }
  // transmittanceOverSegment (1042-1050); with the heterogeneous medium also
  // its ratio-tracking estimate (random draws)
  R transmittanceOverSegment(V3 pW, V3 dir, float len, float* rnd) const {
    V3 h; int m;
    if (traceRay(pW, dir, h, m, len)) return zero();
    R Tr = transmittanceOverFreeSegment(pW, dir, len);
    if (C.medium) Tr = Tr * mediumTransmittance(pW, dir, len, rnd);
    return Tr;
  }

  // ---------------- heterogeneous medium (HAS_VOLUME; the build's extension:
  // the reference declares the hooks, js/renderer.js:258-263, but its path
  // tracer never calls them).  Medium where SDF_VOLUME <= 0, grey extinction
  // VOLUME_EXTINCTION(volumeExtinction, X) / lengthScale under the majorant
  // VOLUME_EXTINCTION_MAX(volumeExtinction) / lengthScale, albedo s / (s + a)
  // of the Volume colours, Henyey-Greenstein phase of the anisotropy.  Delta
  // tracking for free flights, ratio tracking for shadow rays; where the
  // bound exceeds 1/16 of the majorant's mean free path the walk jumps it.
  static constexpr int kVolumeSteps = 1024;
  static constexpr float kVolumeSkip = 0.0625f;
  float mediumSigmaMax() const { return mediumExtinctionMaxHook(P.volumeExtinction) / P.lengthScale; }
  float mediumSigma(V3 X) const { return mediumExtinctionHook(P.volumeExtinction, X) / P.lengthScale; }
  R mediumAlbedo() const {
    V3 sc = A3(P.atmosphereScatteringCoeffRGB);
    V3 ab = A3(P.atmosphereExtinctionCoeffRGB) - A3(P.atmosphereScatteringCoeffRGB);
    return A(sc / vmax(sc + ab, DENOM_TOLERANCE));
  }
  float mediumFreeFlight(V3 o, V3 d, float tmax, float* rnd) const {
    COUNT_SCOPE(CAT_MEDIUM);
    float sigmaMax = mediumSigmaMax();
    if (!(sigmaMax > 0.0f)) return tmax;
    float skip = kVolumeSkip / sigmaMax;
    float t = 0.0f;
    for (int n = 0; n < kVolumeSteps; ++n) {
      float dv = mediumBoundHook(o + t * d);
      if (dv > skip) {
        t += dv;
        if (t >= tmax) return tmax;
        continue;
      }
      t -= std::log(1.0f - randf(rnd)) / sigmaMax;
      if (t >= tmax) return tmax;
      if (randf(rnd) * sigmaMax < mediumSigma(o + t * d)) return t;
    }
    return tmax;
  }
  float mediumTransmittance(V3 o, V3 d, float tmax, float* rnd) const {
    COUNT_SCOPE(CAT_MEDIUM);
    float sigmaMax = mediumSigmaMax();
    if (!(sigmaMax > 0.0f)) return 1.0f;
    float skip = kVolumeSkip / sigmaMax;
    float T = 1.0f, t = 0.0f;
    for (int n = 0; n < kVolumeSteps; ++n) {
      float dv = mediumBoundHook(o + t * d);
      if (dv > skip) {
        t += dv;
        if (t >= tmax) break;
        continue;
      }
      t -= std::log(1.0f - randf(rnd)) / sigmaMax;
      if (t >= tmax) break;
      T *= 1.0f - mediumSigma(o + t * d) / sigmaMax;
    }
    return T;
  }
  float atmosphericScatteringProbability(R Tr) const {
    if (!C.atmosphere) return 0.0f;
    R ext = VOLUME_EXTINCTION_EVAL(), sca = VOLUME_SCATTERING_EVAL();
    return avgC(sca / ext) * (1.0f - avgC(Tr));
  }
  R shadowFactor(V3 o, V3 d, float len, float* rnd) const {
    R Tr = transmittanceOverSegment(o, d, len, rnd);
    return absR(one() - P.shadowStrength * (one() - Tr));
  }

  // directSurfaceLighting (1221-1295)
  R directSurfaceLighting(V3 pW, const Basis& b, V3 winW, int m, float* rnd, float& skyPdf, float& sunPdf, float& sphPdf) const {
    V3 wi = worldToLocal(winW, b);
    R Ld = zero();
    if (P.skyPower > RADIANCE_EPSILON) {
      R Li = zero();
      V3 wL, wW;
      if (!(P.skyPower < RADIANCE_EPSILON)) {
        wL = sampleHemisphereCosineWeighted(rnd, skyPdf);
        wW = localToWorld(wL, b);
        Li = environmentRadiance(wW);
      }
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float bp = fmax_(PDF_EPSILON, pdfBsdf(pW, b, wi, wL, m));
          R f = evaluateBsdf(pW, b, wi, wL, m);
          float w = powerHeuristic(skyPdf, bp);
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, skyPdf) * std::fabs(dot(wW, b.nW)) * w;
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      R Li = zero();
      V3 wL, wW;
      wW = sampleSunDir(rnd, sunPdf);
      wL = worldToLocal(wW, b);
      if (!(wL.z < 0.0f)) Li = sunRadiance(wW);
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float bp = fmax_(PDF_EPSILON, pdfBsdf(pW, b, wi, wL, m));
          R f = evaluateBsdf(pW, b, wi, wL, m);
          float w = powerHeuristic(sunPdf, bp);
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, sunPdf) * std::fabs(dot(wW, b.nW)) * w;
        }
      }
    }
    if (C.sphereLight && (P.sphereLightPower > RADIANCE_EPSILON) &&
        (dot(pW - A3(P.sphereLightPosition), b.nW) <= P.sphereLightRadius)) {
      V3 pHit;
      R Li = sampleSphereLight(rnd, pW, pHit, sphPdf);
      if (avgC(Li) > RADIANCE_EPSILON) {
        float hitDist = fmax_(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
        V3 wW = normalize(pHit - pW);
        V3 wL = worldToLocal(wW, b);
        Li = Li * shadowFactor(pW, wW, hitDist, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float bp = fmax_(PDF_EPSILON, pdfBsdf(pW, b, wi, wL, m));
          R f = evaluateBsdf(pW, b, wi, wL, m);
          float w = powerHeuristic(sphPdf, bp);
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, sphPdf) * std::fabs(dot(wW, b.nW)) * w;
        }
      }
    }
    return minR(rt(P.radianceClamp, R()), Ld);
  }

  // directVolumeLighting (1299-1363)
  R directVolumeLighting(V3 pW, V3 rayDir, float* rnd) const {
    R Ld = zero();
    if (P.skyPower > RADIANCE_EPSILON) {
      V3 wW = v3(0.0f);
      float skyPdf = 0.0f;
      R Li = zero();
      if (!(P.skyPower < RADIANCE_EPSILON)) {
        wW = sampleSphereUniformly(rnd);
        skyPdf = 1.0f / (4.0f * PI_F);
        Li = environmentRadiance(wW);
      }
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON)
          Ld = Ld + phaseFunction(dot(-rayDir, -wW), P.atmosphereAnisotropy) * Li / fmax_(PDF_EPSILON, skyPdf);
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      float sunPdf = 0.0f;
      V3 wW = sampleSunDir(rnd, sunPdf);
      R Li = sunRadiance(wW);
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON)
          Ld = Ld + phaseFunction(dot(-rayDir, -wW), P.atmosphereAnisotropy) * Li / fmax_(PDF_EPSILON, sunPdf);
      }
    }
    if (C.sphereLight && P.sphereLightPower > RADIANCE_EPSILON) {
      float sphPdf = 0.0f;
      V3 pHit;
      R Li = sampleSphereLight(rnd, pW, pHit, sphPdf);
      if (avgC(Li) > RADIANCE_EPSILON) {
        float hitDist = fmax_(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
        V3 wW = normalize(pHit - pW);
        Li = Li * shadowFactor(pW, wW, hitDist, rnd);
        if (avgC(Li) > RADIANCE_EPSILON)
          Ld = Ld + phaseFunction(dot(-rayDir, -wW), P.atmosphereAnisotropy) * Li / fmax_(PDF_EPSILON, sphPdf);
      }
    }
    return minR(rt(P.radianceClamp, R()), Ld);
  }

  // direct light at a scattering point of the heterogeneous medium: one sample
  // per light (sky uniform over the sphere, sun cone, sphere light), shadow
  // rays with ratio tracking, balance heuristic against the phase sample
  R directMediumLighting(V3 pW, V3 rayDir, float g, float* rnd, float& skyPdf, float& sunPdf, float& sphPdf) const {
    R Ld = zero();
    if (P.skyPower > RADIANCE_EPSILON) {
      V3 wW = sampleSphereUnit(rnd);
      skyPdf = 1.0f / (4.0f * PI_F);
      R Li = environmentRadiance(wW);
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float PF = phaseFunction(dot(rayDir, wW), g);
          Ld = Ld + PF * Li / skyPdf * powerHeuristic(skyPdf, PF);
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      V3 wW = sampleSunDir(rnd, sunPdf);
      R Li = sunRadiance(wW);
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float PF = phaseFunction(dot(rayDir, wW), g);
          Ld = Ld + PF * Li / fmax_(PDF_EPSILON, sunPdf) * powerHeuristic(sunPdf, PF);
        }
      }
    }
    if (C.sphereLight && P.sphereLightPower > RADIANCE_EPSILON) {
      V3 pHit;
      R Li = sampleSphereLight(rnd, pW, pHit, sphPdf);
      if (avgC(Li) > RADIANCE_EPSILON) {
        float hitDist = fmax_(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
        V3 wW = normalize(pHit - pW);
        Li = Li * shadowFactor(pW, wW, hitDist, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float PF = phaseFunction(dot(rayDir, wW), g);
          Ld = Ld + PF * Li / fmax_(PDF_EPSILON, sphPdf) * powerHeuristic(sphPdf, PF);
        }
      }
    }
    return minR(rt(P.radianceClamp, R()), Ld);
  }

  // a scatter distance in the atmosphere part [t0, t1] of the segment, drawn
  // with density proportional to the transmittance (1392-1398, 1430-1436)
  bool scatterPoint(V3 pW, V3 rayDir, float len, float* rnd, V3& pS, R& TrS, float& invPdf) const {
    R ext = VOLUME_EXTINCTION_EVAL();
    float en = avgC(ext);
    float t0 = 0.0f, t1 = 0.0f;
    if (!atmosphereSegment(pW, rayDir, len, t0, t1)) return false;
    float T01 = std::exp(-(t1 - t0) * en);
    float ts = t0 - std::log(1.0f - randf(rnd) * (1.0f - T01)) / en;
    pS = pW + ts * rayDir;
    R od = (ts - t0) * ext;
    TrS = expR(-od);
    invPdf = (1.0f - T01) * std::exp(avgC(od)) / en;
    return true;
  }

  // atmosphericInscatteringRadiance (1365-1417)
  R atmosphericInscatteringRadiance(V3 pW, V3 rayDir, float len, float* rnd) const {
    R Ls = zero();
    if (C.fog && P.fogEnable != 0.0f) Ls = (one() - transmittanceOverFreeSegment(pW, rayDir, len)) * A(A3(P.fogEmission));
    R ext = VOLUME_EXTINCTION_EVAL(), sca = VOLUME_SCATTERING_EVAL();
    if (avgC(ext) < RADIANCE_EPSILON || avgC(sca) < RADIANCE_EPSILON) return Ls;
    V3 pS;
    R TrS;
    float invPdf;
    if (!scatterPoint(pW, rayDir, len, rnd, pS, TrS, invPdf)) return Ls;
    R Li = directVolumeLighting(pS, rayDir, rnd);
    Ls = Ls + TrS * sca * Li * invPdf;
    if (emissionHook) {
      V3 wo = samplePhaseFunction(rayDir, P.atmosphereAnisotropy, rnd);
      V3 h;
      int m;
      if (traceRay(pS, wo, h, m, P.maxLengthScale)) Ls = Ls + TrS * sca * VOLUME_EMISSION_EVAL(h) * invPdf;
    }
    return Ls;
  }

  // atmosphericScatterSample (1419-1442)
  R atmosphericScatterSample(V3 pW, V3 rayDir, float len, float* rnd, V3& pS, V3& wo) const {
    R TrS;
    float invPdf;
    if (!scatterPoint(pW, rayDir, len, rnd, pS, TrS, invPdf)) return zero();
    wo = samplePhaseFunction(rayDir, P.atmosphereAnisotropy, rnd);
    return VOLUME_SCATTERING_EVAL() * TrS * invPdf;
  }

  // randomwalk_SSS (1468-1518)
  bool randomwalkSSS(V3 pW, const Basis& b, float* rnd, R albedoSSS, R& walkT, V3& pExit) const {
    V3 pWalk = pW - 3.0f * P.minLengthScale * b.nW;
    float pdfDir;
    V3 dL = sampleHemisphereCosineWeighted(rnd, pdfDir);
    V3 dW = -localToWorld(dL, b);
    walkT = one();
    for (int n = 0; n < 2048; ++n) {
      if ((float)n > P.maxSSSSteps) break;
      float step = -std::log(randf(rnd)) * P.subsurfaceMFP * P.lengthScale;
      V3 h;
      int m;
      if (traceRay(pWalk, dW, h, m, step)) {
        pExit = h;
        return true;
      }
      if (n > 4) {
        float cont = fclamp(10.0f * maxC(walkT), 0.0f, 1.0f);
        if (randf(rnd) < 1.0f - cont) break;
        walkT = walkT / cont;
      }
      dW = samplePhaseFunction(dW, P.subsurfaceAnisotropy, rnd);
      pWalk = pWalk + step * dW;
      walkT = walkT * albedoSSS;
    }
    return false;
  }

  // SSS_exit_radiance (1521-1593): a Lambertian lobe at the exit point
  R exitRadianceSSS(V3 pW, const Basis& b, float* rnd, R albedoExit, float& skyPdf, float& sunPdf, float& sphPdf) const {
    V3 dPw = 3.0f * P.minLengthScale * b.nW;
    R Ld = zero();
    R f = mixR(one(), albedoExit, P.subsurfaceDiffuseWeight) / PI_F;
    if (P.skyPower > RADIANCE_EPSILON) {
      V3 wL = v3(0.0f), wW = v3(0.0f);
      R Li = zero();
      if (!(P.skyPower < RADIANCE_EPSILON)) {
        wL = sampleHemisphereCosineWeighted(rnd, skyPdf);
        wW = localToWorld(wL, b);
        Li = environmentRadiance(wW);
      }
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW + dPw, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float w = powerHeuristic(skyPdf, pdfHemisphereCosineWeighted(wL));
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, skyPdf) * std::fabs(dot(wW, b.nW)) * w;
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      V3 wW = sampleSunDir(rnd, sunPdf);
      V3 wL = worldToLocal(wW, b);
      R Li = wL.z < 0.0f ? zero() : sunRadiance(wW);
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * shadowFactor(pW + dPw, wW, P.maxLengthScale, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float w = powerHeuristic(sunPdf, pdfHemisphereCosineWeighted(wL));
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, sunPdf) * std::fabs(dot(wW, b.nW)) * w;
        }
      }
    }
    if (C.sphereLight && (P.sphereLightPower > RADIANCE_EPSILON) &&
        (dot(pW - A3(P.sphereLightPosition), b.nW) <= P.sphereLightRadius)) {
      V3 pHit;
      R Li = sampleSphereLight(rnd, pW, pHit, sphPdf);
      if (avgC(Li) > RADIANCE_EPSILON) {
        float hitDist = fmax_(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
        V3 wW = normalize(pHit - pW);
        V3 wL = worldToLocal(wW, b);
        Li = Li * shadowFactor(pW + dPw, wW, hitDist, rnd);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float w = powerHeuristic(sphPdf, pdfHemisphereCosineWeighted(wL));
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, sphPdf) * std::fabs(dot(wW, b.nW)) * w;
        }
      }
    }
    return minR(rt(P.radianceClamp, R()), Ld);
  }

  // cameraPath (1625-1926)
  R cameraPath(V3 start, V3 dir, float* rnd) const {
    R L = zero();
    float misSky = 1.0f, misSun = 1.0f, misSph = 1.0f;
    V3 pW = start, rayDir = dir;
    bool inDielectric = hasMat(MAT_DIELE) && evalSdf(sdfs[MAT_DIELE], start) < 0.0f;
    R throughput = one();
    int scatters = 0;
    for (int vertex = 0; vertex <= C.maxBounces; ++vertex) {
      if (maxC(throughput) < THROUGHPUT_EPSILON) break;
      V3 pNext;
      int hm;
      bool hit = traceRay(pW, rayDir, pNext, hm, P.maxLengthScale);
      float rayLength = P.maxLengthScale;
      if (hit) rayLength = length(pNext - pW);
      if (C.medium) {
        // delta tracking up to the surface hit (maxLengthScale on a miss); a
        // real collision makes this vertex a scattering point
        float tc = mediumFreeFlight(pW, rayDir, rayLength, rnd);
        if (tc < rayLength) {
          if (C.sphereLight) {
            V3 pl;
            R LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? zero() : sphereLightRadiance(pW, rayDir, pl);
            if (avgC(LiSph) > RADIANCE_EPSILON && length(pl - pW) < tc) L = L + throughput * misSph * LiSph;
          }
          if (hasMat(MAT_DIELE) && inDielectric) throughput = throughput * expR(-tc * A(A3(P.dieleAbsorptionRGB)));
          pW = pW + tc * rayDir;
          throughput = throughput * mediumAlbedo();
          float g = P.atmosphereAnisotropy;
          float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
          if (!(hasMat(MAT_DIELE) && inDielectric))
            L = L + throughput * directMediumLighting(pW, rayDir, g, rnd, skyPdf, sunPdf, sphPdf);
          float phasePdf = 0.0f;
          rayDir = sampleHenyeyGreenstein(rayDir, g, rnd, phasePdf);
          misSky = powerHeuristic(phasePdf, skyPdf);
          misSun = powerHeuristic(phasePdf, sunPdf);
          misSph = powerHeuristic(phasePdf, sphPdf);
          if (maxC(throughput) < THROUGHPUT_EPSILON) break;
          continue;
        }
      }
      if (!hit) {
        float segLen = P.maxLengthScale;
        R Tr = one();
        bool lit = false;
        if (C.sphereLight) {
          V3 pl;
          R LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? zero() : sphereLightRadiance(pW, rayDir, pl);
          if (avgC(LiSph) > RADIANCE_EPSILON) {
            lit = true;
            segLen = length(pl - pW);
            Tr = transmittanceOverFreeSegment(pW, rayDir, segLen);
            float sp = atmosphericScatteringProbability(Tr);
            L = L + throughput * Tr * misSph * LiSph / (1.0f - sp);
          }
        }
        if (!lit) {
          segLen = P.maxLengthScale;
          Tr = transmittanceOverFreeSegment(pW, rayDir, segLen);
          float sp = atmosphericScatteringProbability(Tr);
          if (avgC(Tr) > RADIANCE_EPSILON) {
            if (!(vertex == 0 && P.envMapVisible == 0.0f) && misSky > 0.0f)
              L = L + throughput * Tr * misSky * environmentRadiance(rayDir) / (1.0f - sp);
            if (!(vertex == 0 && P.sunVisibleDirectly == 0.0f) && misSun > 0.0f)
              L = L + throughput * Tr * misSun * sunRadiance(rayDir) / (1.0f - sp);
          }
        }
        if (C.atmosphere) {
          if (vertex < 2) L = L + throughput * atmosphericInscatteringRadiance(pW, rayDir, segLen, rnd);
          if (scatters + 1 < C.maxAtmScatters) {
            float sp = atmosphericScatteringProbability(Tr);
            if (sp > randf(rnd)) {
              V3 pS = v3(0.0f), wo = v3(0.0f);
              throughput = throughput * atmosphericScatterSample(pW, rayDir, segLen, rnd, pS, wo) / fmax_(DENOM_TOLERANCE, sp);
              pW = pS;
              rayDir = wo;
              scatters++;
              continue;
            }
          }
        }
        break;
      }
      if (hm == MAT_INVAL) break;  // max-step "hit" without a material: undefined in the reference; ends the path
      R Tr = one();
      bool hitSphereLight = false;
      if (C.sphereLight) {
        V3 pl;
        R LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? zero() : sphereLightRadiance(pW, rayDir, pl);
        if (avgC(LiSph) > RADIANCE_EPSILON) {
          float lsl = length(pl - pW);
          if (lsl < rayLength) {
            hitSphereLight = true;
            Tr = transmittanceOverFreeSegment(pW, rayDir, lsl);
            L = L + throughput * Tr * misSph * LiSph;
            if (C.atmosphere && vertex < 2) L = L + throughput * atmosphericInscatteringRadiance(pW, rayDir, lsl, rnd);
          }
        }
      }
      if (hasMat(MAT_DIELE) && inDielectric) {
        R absorption = A(A3(P.dieleAbsorptionRGB));
        throughput = throughput * expR(-rayLength * absorption);
      }
      if (C.atmosphere) {
        if (vertex < 2) L = L + throughput * atmosphericInscatteringRadiance(pW, rayDir, rayLength, rnd);
        if (!hitSphereLight) Tr = transmittanceOverFreeSegment(pW, rayDir, rayLength);
        if (scatters + 1 < C.maxAtmScatters) {
          float sp = atmosphericScatteringProbability(Tr);
          if (sp > randf(rnd)) {
            V3 pS = v3(0.0f), wo = v3(0.0f);
            throughput = throughput * atmosphericScatterSample(pW, rayDir, rayLength, rnd, pS, wo) / fmax_(DENOM_TOLERANCE, sp);
            pW = pS;
            rayDir = wo;
            scatters++;
            continue;
          }
          throughput = throughput / fmax_(DENOM_TOLERANCE, 1.0f - sp);
        }
        if (hitSphereLight) break;
        throughput = throughput * transmittanceOverFreeSegment(pW, rayDir, rayLength);
      }
      pW = pNext;
      V3 nW = normal(pW, hm);
      V3 ngW = nW;
      Basis b = makeBasis(nW);
      bool sss = false;
      float probSSS = 0.0f;
      R albedoSSS = zero();
      if (hasMat(MAT_SURFA) && P.subsurfaceMFP > 0.0f) {
        albedoSSS = A(A3(P.subsurfaceAlbedoRGB));
        float diffW = (1.0f - P.subsurface) * avgC(diffuseRefl(pW));
        float specW = avgC(A(A3(P.surfaceSpecAlbedoRGB)));
        float sssW = P.subsurface * avgC(albedoSSS) * 3.0f;
        float total = diffW + specW + sssW;
        probSSS = sssW / (total + DENOM_TOLERANCE);
        probSSS = fclamp(probSSS, 0.0f, 1.0f - PDF_EPSILON);
        sss = randf(rnd) < probSSS;
      }
      if (!sss) {
        V3 winW = -rayDir;
        V3 wi = worldToLocal(winW, b);
        // WebGL zero-initialises uninitialised locals (woutputL, bsdfPdf)
        V3 wo = v3(0.0f);
        float bsdfPdf = 0.0f;
        R f = sampleBsdf(pW, b, wi, hm, wo, bsdfPdf, rnd);
        V3 woW = localToWorld(wo, b);
        if (hm == MAT_DIELE && dot(winW, ngW) * dot(woW, ngW) < 0.0f) inDielectric = !inDielectric;
        if (emissionHook) L = L + throughput * VOLUME_EMISSION_EVAL(pW);
        rayDir = woW;
        pW = pW + ngW * fsign(dot(rayDir, ngW)) * 3.0f * P.minLengthScale;
        float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
        if (!(hasMat(MAT_DIELE) && inDielectric))
          L = L + throughput * directSurfaceLighting(pW, b, winW, hm, rnd, skyPdf, sunPdf, sphPdf);
        R fOverPdf = minR(rt(P.radianceClamp, R()), f / fmax_(PDF_EPSILON, bsdfPdf));
        R st = fOverPdf * std::fabs(dot(woW, nW)) / fmax_(PDF_EPSILON, 1.0f - probSSS);
        throughput = throughput * st;
        misSky = powerHeuristic(bsdfPdf, skyPdf);
        misSun = powerHeuristic(bsdfPdf, sunPdf);
        misSph = powerHeuristic(bsdfPdf, sphPdf);
      } else {
        R walkT = one();
        V3 pExit = v3(0.0f);
        b.nW = ngW;  // the entry uses the geometric normal
        if (!randomwalkSSS(pW, b, rnd, albedoSSS, walkT, pExit)) break;
        pW = pExit;
        nW = normal(pExit, MAT_SURFA);
        Basis be = makeBasis(nW);
        float bsdfPdf;
        V3 dL = sampleHemisphereCosineWeighted(rnd, bsdfPdf);
        V3 woW = localToWorld(dL, be);
        rayDir = woW;
        throughput = throughput * walkT;
        float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
        R albedoExit = diffuseRefl(pW);
        if (!(hasMat(MAT_DIELE) && inDielectric))
          L = L + throughput * exitRadianceSSS(pExit, be, rnd, albedoExit, skyPdf, sunPdf, sphPdf);
        misSky = powerHeuristic(bsdfPdf, skyPdf);
        misSun = powerHeuristic(bsdfPdf, sunPdf);
        misSph = powerHeuristic(bsdfPdf, sphPdf);
        R f = mixR(one(), albedoExit, P.subsurfaceDiffuseWeight) / PI_F;
        R fOverPdf = minR(rt(P.radianceClamp, R()), f / fmax_(PDF_EPSILON, bsdfPdf));
        throughput = throughput * (fOverPdf * std::fabs(dot(woW, nW)) / fmax_(PDF_EPSILON, probSSS));
        pW = pW + nW * fsign(dot(rayDir, nW)) * 3.0f * P.minLengthScale;
      }
      if (maxC(throughput) < THROUGHPUT_EPSILON) break;
    }
    return L;
  }
};

inline V3 toXYZ(float L, V3 xyz) { return xyz * L; }  // dispersive: xyz * L
inline V3 toXYZ(V3 L, V3) { return rgbToXyz(L); }

// GLSL's tan, correctly rounded (through double), as the HIP path computes
// it: the camera's frustum height scales every primary ray
inline float tanCR(float x) {
#ifdef ORACLE_COUNT
  orccount::trans();
#endif
  return (float)std::tan((double)x);
}

float sample_jitter(float xi) { return xi < 0.5f ? std::sqrt(2.0f * xi) - 1.0f : 1.0f - std::sqrt(2.0f - 2.0f * xi); }

void constructPrimaryRay(float px, float py, float* rnd, V3& start, V3& dir, bool safe) {
  float ndcx = -1.0f + 2.0f * (px / P.resolution[0]);
  float ndcy = -1.0f + 2.0f * (py / P.resolution[1]);
  float fh = tanCR(0.5f * (P.camFovy * 0.017453292519943295f));
  float fw = P.camAspect * fh;
  V3 s = -fw * ndcx * A3(P.camX) + fh * ndcy * A3(P.camY);
  dir = safe ? safe_normalize(A3(P.camDir) + s) : normalize(A3(P.camDir) + s);
  if (P.camAperture <= 0.0f) { start = A3(P.camPos); return; }
  V3 focal = A3(P.camPos) + P.camFocalDistance * dir / dot(dir, A3(P.camDir));
  float lensRadial = P.camAperture * std::sqrt(randf(rnd));
  float theta = 2.0f * PI_F * randf(rnd);
  V3 lensPos = A3(P.camPos) + lensRadial * (-A3(P.camX) * std::cos(theta) + A3(P.camY) * std::sin(theta));
  start = lensPos;
  dir = safe ? safe_normalize(focal - lensPos) : normalize(focal - lensPos);
}

// pathtrace + main (1934-2009) for one pixel
template <class R>
void pathtracePixel(int ix, int iy, float* rad, float* rnd) {
  if (C.interactive) {
    if (randf(rnd) < P.skipProbability) return;  // radiance unchanged, rng advanced
  }
  Tracer<R> T;
  V3 xyz = v3(0.0f);
  if (C.dispersive) {
    float ws = P.wavelengthSamples;
    float xi = 0.1f + 0.8f * (0.5f + std::floor(ws * randf(rnd))) / ws;
    float w = lerpTable(lutICDF.data(), 4096, xi);
    T.wl = 390.0f + (750.0f - 390.0f) * w;
    xyz = lookupXYZ(w);
    T.rgb = xyzToRgb(xyz);
  } else {
    T.wl = 390.0f + (750.0f - 390.0f) * 0.5f;
    T.rgb = v3(0.0f);
  }
  R L = T.zero();
  float px = ix + 0.5f, py = iy + 0.5f;
  for (int n = 0; n < C.spf; ++n) {
    float jx = 0.5f * P.filterRadius * sample_jitter(randf(rnd));
    float jy = 0.5f * P.filterRadius * sample_jitter(randf(rnd));
    V3 s, d;
    constructPrimaryRay(px + jx, py + jy, rnd, s, d, true);
    COUNT_ADD(paths);
    L = L + T.cameraPath(s, d, rnd);
  }
  L = L / float(C.spf);
  V3 c = toXYZ(L, xyz);
  float oldN = rad[3], newN = oldN + 1.0f;
  V3 newL = (oldN * v3(rad[0], rad[1], rad[2]) + c) / newN;
  rad[0] = newL.x; rad[1] = newL.y; rad[2] = newL.z; rad[3] = newN;
}

// traceDistance (normals-fragment-shader.glsl:70-113, ao-fragment-shader.glsl:100-144):
// t starts at 0 and the step precedes the evaluation
bool traceDistance(V3 s, V3 d, float maxDist, V3& hit, int& material) {
  float sdf = 1.0e20f;
  for (int m : {MAT_SURFA, MAT_METAL, MAT_DIELE}) if (hasMat(m)) sdf = fmin_(sdf, std::fabs(evalSdf(sdfs[m], s)));
  float InitialSign = fsign(sdf), t = 0.0f;
  int iters = 0;
  material = MAT_INVAL;
  for (int k = 0; k < C.maxMarch; k++) {
    t += InitialSign * sdf;
    if (t >= maxDist) break;
    V3 pW = s + t * d;
    sdf = 1.0e20f;
    bool h = false;
    for (int m : {MAT_SURFA, MAT_METAL, MAT_DIELE}) {
      if (!hasMat(m)) continue;
      float v = std::fabs(evalSdf(sdfs[m], pW));
      if (v < P.minLengthScale) { material = m; h = true; break; }
      sdf = fmin_(sdf, v);
    }
    if (h) break;
    iters++;
  }
  hit = s + t * d;
  return !(t >= maxDist) && !(P.maxStepsIsMiss != 0.0f && iters >= C.maxMarch);
}

inline void accumulate(float* rad, V3 c) {
  float oldN = rad[3], newN = oldN + 1.0f;
  V3 newL = (oldN * v3(rad[0], rad[1], rad[2]) + c) / newN;
  rad[0] = newL.x; rad[1] = newL.y; rad[2] = newL.z; rad[3] = newN;
}

// ambient occlusion main (ao-fragment-shader.glsl:395-488)
void aoPixel(int ix, int iy, float* rad, float* rnd) {
  const float RAD_EPS = 1.0e-5f;  // this program's RADIANCE_EPSILON (62)
  if (C.interactive && randf(rnd) < P.skipProbability) return;
  Tracer<V3> T;
  T.wl = 570.0f;
  T.rgb = v3(0.0f);
  V3 RGB = v3(0.0f);
  for (int n = 0; n < C.spf; ++n) {
    float r1 = randf(rnd);
    float r2 = randf(rnd);
    float px = ix + 0.5f + (-0.5f + r1), py = iy + 0.5f + (-0.5f + r2);
    V3 s, d;
    constructPrimaryRay(px, py, rnd, s, d, false);
    V3 pW;
    int m;
    if (traceDistance(s, d, P.maxLengthScale, pW, m)) {
      if (m == MAT_INVAL) continue;  // max-step "hit": no normal defined; the sample adds nothing
      V3 nW = T.normal(pW, m, false);
      Basis b = makeBasis(nW, false);
      V3 f = surfaceDiffuseHook(A3(P.surfaceDiffuseAlbedoRGB), pW) / PI_F;
      V3 Ld = v3(0.0f);
      V3 dPw = 3.0f * P.minLengthScale * nW;
      V3 h = v3(0.0f);
      int hm;
      if (P.skyPower > RAD_EPS) {
        float skyPdf = 0.0f;
        V3 wL = sampleHemisphereCosineWeighted(rnd, skyPdf);
        V3 wW = localToWorld(wL, b);
        V3 Li = T.environmentRadiance(wW);
        Li = Li * (traceDistance(pW + dPw, wW, P.maxLengthScale, h, hm) ? 0.0f : 1.0f);
        Ld = Ld + f * Li / fmax_(PDF_EPSILON, skyPdf) * std::fabs(dot(wW, nW));
      }
      if (P.sunPower > RAD_EPS) {
        float theta_max = P.sunAngularSize * PI_F / 180.0f;
        float theta = theta_max * std::sqrt(randf(rnd));
        float ct = std::cos(theta), st = std::sqrt(fmax_(0.0f, 1.0f - ct * ct));
        float phi = 2.0f * PI_F * randf(rnd);
        float sunPdf = 1.0f / (2.0f * PI_F * (1.0f - std::cos(theta_max)));
        V3 wW = localToWorld(v3(st * std::cos(phi), st * std::sin(phi), ct), makeBasis(A3(P.sunDir), false));
        V3 wL = worldToLocal(wW, b);
        V3 Li = wL.z < 0.0f ? v3(0.0f) : T.sunRadiance(wW);
        if (avgC(Li) > RAD_EPS) {
          Li = Li * (traceDistance(pW + dPw, wW, P.maxLengthScale, h, hm) ? 0.0f : 1.0f);
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, sunPdf) * std::fabs(dot(wW, nW));
        }
      }
      RGB = RGB + Ld;
    } else if (P.envMapVisible != 0.0f) {
      RGB = RGB + T.environmentRadiance(d);
    }
  }
  accumulate(rad, rgbToXyz(RGB / float(C.spf)));
}

// the simplified path tracer (simplepathtracer-fragment-shader.glsl): RGB,
// the Surface material only, fog only, no sphere light, no shadowStrength
struct Simple : Tracer<V3> {
  bool traceSurface(V3 start, V3 dir, V3& hit, int& material, float maxMarchDist) const {  // 508-534
    material = MAT_INVAL;
    if (!hasMat(MAT_SURFA)) return false;
    float sdf = fmin_(1.0e20f, std::fabs(evalSdf(sdfs[MAT_SURFA], start)));
    float InitialSign = fsign(sdf);
    float t = InitialSign * sdf;
    if (std::fabs(t) >= maxMarchDist) return false;
    for (int n = 0; n < C.maxMarch; n++) {
      V3 pW = start + t * dir;
      float v = std::fabs(evalSdf(sdfs[MAT_SURFA], pW));
      if (v < P.minLengthScale) { material = MAT_SURFA; hit = start + t * dir; return true; }
      sdf = fmin_(1.0e20f, v);
      t += InitialSign * sdf;
      if (std::fabs(t) >= maxMarchDist) return false;
    }
    return P.maxStepsIsMiss == 0.0f;
  }
  V3 trFree(float len) const {  // 570-582
    V3 Tr = v3(1.0f);
    if (C.fog && P.fogEnable != 0.0f) {
      V3 fogAbsorption = v3(1.0f) - A3(P.fogTint);
      Tr = Tr * vexp(-(P.fogMaxOpticalDepth * (v3(1.0f) - vexp(-(len * fogAbsorption / P.fogMFP)))));
    }
    return Tr;
  }
  V3 trSeg(V3 pW, V3 dir, float len) const {  // 584-597
    V3 h = v3(0.0f);
    int m;
    if (traceSurface(pW, dir, h, m, len)) return v3(0.0f);
    return C.fog ? trFree(len) : v3(1.0f);
  }
  V3 inscatter(V3 Tr) const { return P.fogEnable == 0.0f ? v3(0.0f) : (v3(1.0f) - Tr) * A3(P.fogEmission); }
  V3 direct(V3 pW, const Basis& b, V3 winW, float* rnd, float& skyPdf, float& sunPdf) const {  // 719-770
    V3 wi = worldToLocal(winW, b);
    V3 Ld = v3(0.0f);
    if (P.skyPower > RADIANCE_EPSILON) {
      V3 wL = sampleHemisphereCosineWeighted(rnd, skyPdf);
      V3 wW = localToWorld(wL, b);
      V3 Li = environmentRadiance(wW);
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * trSeg(pW, wW, P.maxLengthScale);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float bp = fmax_(PDF_EPSILON, pdfBsdf(pW, b, wi, wL, MAT_SURFA));
          V3 f = evaluateBsdf(pW, b, wi, wL, MAT_SURFA);
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, skyPdf) * std::fabs(dot(wW, b.nW)) * powerHeuristic(skyPdf, bp);
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      V3 wW = sampleSunDir(rnd, sunPdf);
      V3 wL = worldToLocal(wW, b);
      V3 Li = wL.z < 0.0f ? v3(0.0f) : sunRadiance(wW);
      if (avgC(Li) > RADIANCE_EPSILON) {
        Li = Li * trSeg(pW, wW, P.maxLengthScale);
        if (avgC(Li) > RADIANCE_EPSILON) {
          float bp = fmax_(PDF_EPSILON, pdfBsdf(pW, b, wi, wL, MAT_SURFA));
          V3 f = evaluateBsdf(pW, b, wi, wL, MAT_SURFA);
          Ld = Ld + f * Li / fmax_(PDF_EPSILON, sunPdf) * std::fabs(dot(wW, b.nW)) * powerHeuristic(sunPdf, bp);
        }
      }
    }
    return vmin(v3(P.radianceClamp), Ld);
  }
  bool walk(V3 pW, const Basis& b, float* rnd, V3 albedoSSS, V3& walkT, V3& pExit) const {  // 783-834
    V3 pWalk = pW - 3.0f * P.minLengthScale * b.nW;
    float pdfDir;
    V3 dW = -localToWorld(sampleHemisphereCosineWeighted(rnd, pdfDir), b);
    walkT = v3(1.0f);
    for (int n = 0; n < 256; ++n) {
      if ((float)n > P.maxSSSSteps) break;
      float step = -std::log(randf(rnd)) * P.subsurfaceMFP * P.lengthScale;
      V3 h = v3(0.0f);
      int m;
      if (traceSurface(pWalk, dW, h, m, step)) { pExit = h; return true; }
      if (n > 4) {
        float cont = fclamp(10.0f * maxC(walkT), 0.0f, 1.0f);
        if (randf(rnd) < 1.0f - cont) break;
        walkT = walkT / cont;
      }
      dW = samplePhaseFunction(dW, P.subsurfaceAnisotropy, rnd);
      pWalk = pWalk + step * dW;
      walkT = walkT * albedoSSS;
    }
    return false;
  }
  V3 exitRad(V3 pW, const Basis& b, float* rnd, V3 albedoExit, float& skyPdf, float& sunPdf) const {  // 837-882
    V3 Ld = v3(0.0f);
    V3 f = mixR(v3(1.0f), albedoExit, P.subsurfaceDiffuseWeight) / PI_F;
    if (P.skyPower > RADIANCE_EPSILON) {
      V3 wL = sampleHemisphereCosineWeighted(rnd, skyPdf);
      V3 wW = localToWorld(wL, b);
      V3 Li = environmentRadiance(wW) * trSeg(pW, wW, P.maxLengthScale);
      if (avgC(Li) > RADIANCE_EPSILON)
        Ld = Ld + f * Li / fmax_(PDF_EPSILON, skyPdf) * std::fabs(dot(wW, b.nW)) *
                      powerHeuristic(skyPdf, pdfHemisphereCosineWeighted(wL));
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      V3 wW = sampleSunDir(rnd, sunPdf);
      V3 wL = worldToLocal(wW, b);
      V3 Li = (wL.z < 0.0f ? v3(0.0f) : sunRadiance(wW)) * trSeg(pW, wW, P.maxLengthScale);
      if (avgC(Li) > RADIANCE_EPSILON)
        Ld = Ld + f * Li / fmax_(PDF_EPSILON, sunPdf) * std::fabs(dot(wW, b.nW)) *
                      powerHeuristic(sunPdf, pdfHemisphereCosineWeighted(wL));
    }
    return vmin(v3(P.radianceClamp), Ld);
  }
  V3 path(V3 start, V3 dir, float* rnd) const {  // cameraPath 913-1083
    V3 L = v3(0.0f);
    float misSky = 1.0f, misSun = 1.0f;
    V3 pW = start, rayDir = dir;
    V3 throughput = v3(1.0f);
    for (int vertex = 0; vertex <= C.maxBounces; ++vertex) {
      if (maxC(throughput) < THROUGHPUT_EPSILON) break;
      V3 pNext;
      int hm;
      bool hit = traceSurface(pW, rayDir, pNext, hm, P.maxLengthScale);
      float rayLength = P.maxLengthScale;
      V3 Tr;
      if (hit) {
        if (hm == MAT_INVAL) break;  // max-step "hit" without a material: ends the path
        rayLength = length(pNext - pW);
        Tr = trFree(rayLength);
      } else {
        Tr = trFree(P.maxLengthScale);
        if (!C.fog || avgC(Tr) > RADIANCE_EPSILON) {
          if (!(vertex == 0 && P.envMapVisible == 0.0f) && misSky > 0.0f)
            L = L + throughput * Tr * misSky * environmentRadiance(rayDir);
          if (!(vertex == 0 && P.sunVisibleDirectly == 0.0f) && misSun > 0.0f)
            L = L + throughput * Tr * misSun * sunRadiance(rayDir);
        }
        if (C.fog && vertex < 2) L = L + throughput * inscatter(Tr);
        break;
      }
      if (C.fog) {
        if (vertex < 2) L = L + throughput * inscatter(Tr);
        throughput = throughput * Tr;
      }
      pW = pNext;
      V3 nW = normal(pW, MAT_SURFA);
      V3 ngW = nW;
      Basis b = makeBasis(nW);
      bool sss = false;
      float probSSS = 0.0f;
      V3 albedoSSS = v3(0.0f);
      if (hasMat(MAT_SURFA) && P.subsurfaceMFP > 0.0f) {
        albedoSSS = A3(P.subsurfaceAlbedoRGB);
        float diffW = (1.0f - P.subsurface) * avgC(diffuseRefl(pW));
        float specW = avgC(A3(P.surfaceSpecAlbedoRGB));
        float sssW = P.subsurface * avgC(albedoSSS) * 3.0f;
        probSSS = fclamp(sssW / (diffW + specW + sssW + DENOM_TOLERANCE), 0.0f, 1.0f - PDF_EPSILON);
        sss = randf(rnd) < probSSS;
      }
      if (!sss) {
        V3 winW = -rayDir;
        V3 wi = worldToLocal(winW, b);
        V3 wo = v3(0.0f);
        float bsdfPdf = 0.0f;
        V3 f = sampleBsdf(pW, b, wi, MAT_SURFA, wo, bsdfPdf, rnd);
        V3 woW = localToWorld(wo, b);
        rayDir = woW;
        pW = pW + ngW * fsign(dot(rayDir, ngW)) * 3.0f * P.minLengthScale;
        float skyPdf = 0.0f, sunPdf = 0.0f;
        L = L + throughput * direct(pW, b, winW, rnd, skyPdf, sunPdf);
        V3 fOverPdf = vmin(v3(P.radianceClamp), f / fmax_(PDF_EPSILON, bsdfPdf));
        throughput = throughput * (fOverPdf * std::fabs(dot(woW, nW)) / fmax_(PDF_EPSILON, 1.0f - probSSS));
        misSky = powerHeuristic(bsdfPdf, skyPdf);
        misSun = powerHeuristic(bsdfPdf, sunPdf);
      } else {
        V3 walkT = v3(1.0f), pExit = v3(0.0f);
        b.nW = ngW;
        if (!walk(pW, b, rnd, albedoSSS, walkT, pExit)) break;
        pW = pExit;
        nW = normal(pExit, MAT_SURFA);
        Basis be = makeBasis(nW);
        float bsdfPdf;
        V3 woW = localToWorld(sampleHemisphereCosineWeighted(rnd, bsdfPdf), be);
        rayDir = woW;
        throughput = throughput * walkT;
        float skyPdf = 0.0f, sunPdf = 0.0f;
        V3 albedoExit = diffuseRefl(pW);
        L = L + throughput * exitRad(pExit, be, rnd, albedoExit, skyPdf, sunPdf);
        misSky = powerHeuristic(bsdfPdf, skyPdf);
        misSun = powerHeuristic(bsdfPdf, sunPdf);
        V3 f = mixR(v3(1.0f), albedoExit, P.subsurfaceDiffuseWeight) / PI_F;
        V3 fOverPdf = vmin(v3(P.radianceClamp), f / fmax_(PDF_EPSILON, bsdfPdf));
        throughput = throughput * (fOverPdf * std::fabs(dot(woW, nW)) / fmax_(PDF_EPSILON, probSSS));
        pW = pW + nW * fsign(dot(rayDir, nW)) * 3.0f * P.minLengthScale;
      }
    }
    return L;
  }
};

// pathtrace + main of the simplified program (1091-1129): no interactive skip
void simplePixel(int ix, int iy, float* rad, float* rnd) {
  Simple T;
  T.wl = 390.0f + (750.0f - 390.0f) * 0.5f;
  T.rgb = v3(0.0f);
  V3 L = v3(0.0f);
  float px = ix + 0.5f, py = iy + 0.5f;
  for (int n = 0; n < C.spf; ++n) {
    float jx = 0.5f * P.filterRadius * sample_jitter(randf(rnd));
    float jy = 0.5f * P.filterRadius * sample_jitter(randf(rnd));
    V3 s, d;
    constructPrimaryRay(px + jx, py + jy, rnd, s, d, true);
    L = L + T.path(s, d, rnd);
  }
  accumulate(rad, rgbToXyz(L / float(C.spf)));
}

// normals program main (normals-fragment-shader.glsl:293-345)
void normalsPixel(int ix, int iy, float* rad, float* rnd) {
  Tracer<V3> T;
  V3 colorXYZ = v3(0.0f);
  for (int n = 0; n < C.spf; ++n) {
    float r1 = randf(rnd);
    float r2 = randf(rnd);
    float px = ix + 0.5f + (-0.5f + r1), py = iy + 0.5f + (-0.5f + r2);
    V3 s, d;
    constructPrimaryRay(px, py, rnd, s, d, false);
    V3 hit;
    int material;
    bool isHit = traceDistance(s, d, P.maxLengthScale, hit, material);
    if (isHit) {
      V3 nW = material == MAT_INVAL ? v3(0.0f) : T.normal(hit, material, false);
      colorXYZ = colorXYZ + rgbToXyz(0.5f * (nW + v3(1.0f)));
    } else if (P.envMapVisible != 0.0f) {
      // environmentRadianceXYZ (normals-fragment-shader.glsl:234-253)
      if (P.haveEnvMap != 0.0f) {
        float phi = std::atan2(d.x, d.z) + PI_F + PI_F * P.envMapRotation / 180.0f;
        phi -= 2.0f * PI_F * std::floor(phi / (2.0f * PI_F));
        float theta = std::acos(d.y);
        colorXYZ = colorXYZ + rgbToXyz(sampleEnv(phi / (2.0f * PI_F), theta / PI_F) * P.skyPower);
      } else {
        colorXYZ = colorXYZ + P.skyPower * v3(1.0f);
      }
    }
  }
  colorXYZ = colorXYZ / float(C.spf);
  float oldN = rad[3], newN = oldN + 1.0f;
  V3 newL = (oldN * v3(rad[0], rad[1], rad[2]) + colorXYZ) / newN;
  rad[0] = newL.x; rad[1] = newL.y; rad[2] = newL.z; rad[3] = newN;
}

// one frame of a pixel with the program of Renderer.renderMode
void renderPixel(int ix, int iy, float* rad, float* rnd) {
  if (C.mode == 1) aoPixel(ix, iy, rad, rnd);
  else if (C.mode == 2) simplePixel(ix, iy, rad, rnd);
  else if (C.dispersive) pathtracePixel<float>(ix, iy, rad, rnd);
  else pathtracePixel<V3>(ix, iy, rad, rnd);
}

uint32_t pcg_next(uint64_t& state, uint64_t inc) {
  uint64_t old = state;
  state = old * 6364136223846793005ull + inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((-rot) & 31u));
}

}  // namespace

extern "C" {

void oracle_reset() {
  memset(&P, 0, sizeof P);
  C = Config();
  for (auto& s : sdfs) s = SdfSpec();
  albedoHook = 0;
  emissionHook = 0;
  envW = envH = 0;
  envTex.clear();
  dieleModel = 1;
  dieleC.clear();
  metalModel = 0;
  medium = MediumSpec();
}

int oracle_set(const char* name, const float* v, int n) {
  for (auto& e : table()) {
    if (!strcmp(e.name, name)) {
      for (int i = 0; i < e.n && i < n; ++i) e.p[i] = v[i];
      return 0;
    }
  }
  return -1;
}

void oracle_config(int dispersive, int interactive, int maxBounces, int maxMarch, int spf, int sphereLight) {
  C.dispersive = dispersive; C.interactive = interactive; C.maxBounces = maxBounces;
  C.maxMarch = maxMarch; C.spf = spf; C.sphereLight = sphereLight;
}

// material id: 0 dielectric, 1 metal, 2 surface
void oracle_sdf(int material, int kind, float a, float b, float c, float d) {
  sdfs[material].kind = kind; sdfs[material].a = a; sdfs[material].b = b; sdfs[material].c = c; sdfs[material].d = d;
}
void oracle_albedo_hook(int kind) { albedoHook = kind; }
void oracle_emission_hook(int kind) { emissionHook = kind; }
// Renderer.renderMode: 0 'pt', 1 'ao', 2 'ptsimple'
void oracle_mode(int mode) { C.mode = mode; }
// the environment map: w x h linear RGB texels, row 0 first (w = 0: none)
void oracle_envmap(int w, int h, const float* rgb) {
  envW = w;
  envH = h;
  envTex.assign(rgb, rgb + (size_t)3 * w * h);
}
// HAS_ATMOSPHERE, HAS_FOG and __MAX_ATMOSPHERE_SCATTERS__
void oracle_volume(int atmosphere, int fog, int maxAtmScatters) {
  C.atmosphere = atmosphere; C.fog = fog; C.maxAtmScatters = maxAtmScatters;
}
// the heterogeneous medium of a scene with SDF_VOLUME (HAS_VOLUME): kind 0
// none, MEDIUM_CLOUD / MEDIUM_LINEAR_BALL / MEDIUM_BOX with parameters a..d
void oracle_medium(int kind, float a, float b, float c, float d) {
  medium.kind = kind;
  medium.a = a; medium.b = b; medium.c = c; medium.d = d;
  C.medium = kind != MEDIUM_NONE;
}
float oracle_value_noise(const float* p) { return valueNoise(v3(p[0], p[1], p[2])); }
float oracle_lattice_value(int x) { return latticeValue(x); }
float oracle_fbm3(const float* p) { return fbm3(v3(p[0], p[1], p[2])); }
// sigma(X) in units of the extinction uniform (VOLUME_EXTINCTION(1, X)) and the bound SDF_VOLUME(X)
float oracle_medium_density(const float* p) { return mediumExtinctionHook(1.0f, v3(p[0], p[1], p[2])); }
float oracle_medium_bound(const float* p) { return mediumBoundHook(v3(p[0], p[1], p[2])); }
// delta tracking (free-flight distance, tmax if none) and ratio tracking
// (transmittance estimate) along [0, tmax) of the ray o + t d; rnd4 advances
float oracle_medium_free_flight(const float* o, const float* d, float tmax, float* rnd4) {
  Tracer<V3> T;
  return T.mediumFreeFlight(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), tmax, rnd4);
}
float oracle_medium_transmittance(const float* o, const float* d, float tmax, float* rnd4) {
  Tracer<V3> T;
  return T.mediumTransmittance(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), tmax, rnd4);
}
void oracle_sample_hg(const float* dir, float g, float* rnd4, float* out3, float* pdf) {
  V3 w = sampleHenyeyGreenstein(v3(dir[0], dir[1], dir[2]), g, rnd4, *pdf);
  out3[0] = w.x; out3[1] = w.y; out3[2] = w.z;
}
float oracle_phase_function(float mu, float g) { return phaseFunction(mu, g); }
#ifdef ORACLE_COUNT
// the counting build's totals since the last reset: FLOPs, transcendentals and
// other operations per category (shade, march, normal, medium, init), then SDF
// evaluations of the sphere tracer, normals and camera paths: 18 values
void oracle_count_reset() { orccount::total = orccount::Counters{}; orccount::tc = orccount::Counters{}; }
void oracle_count_get(long long* out) {
  const orccount::Counters& t = orccount::total;
  const int n = orccount::CAT_COUNT;
  for (int c = 0; c < n; ++c) {
    out[c] = t.flops[c];
    out[n + c] = t.trans[c];
    out[2 * n + c] = t.ops[c];
  }
  out[3 * n] = t.steps;
  out[3 * n + 1] = t.normals;
  out[3 * n + 2] = t.paths;
}
#endif
void oracle_dielectric(int model, const float* coeffs, int n) { dieleModel = model; dieleC.assign(coeffs, coeffs + n); }
void oracle_metal(int model, const float* n64, const float* k64) {
  metalModel = model;
  if (n64) metalN.assign(n64, n64 + 64);
  if (k64) metalK.assign(k64, k64 + 64);
}
void oracle_luts(const float* xyz4096, const float* icdf4096) {
  lutXYZ.assign(xyz4096, xyz4096 + 4096);
  lutICDF.assign(icdf4096, icdf4096 + 4096);
}

float oracle_rand(float* state4) { return randf(state4); }
float oracle_ior_diele(float wl) { return IOR_DIELE(wl); }
float oracle_ior_metal(float wl) { return IOR_METAL(wl); }
float oracle_k_metal(float wl) { return K_METAL_table(wl); }
float oracle_fresnel_dielectric(float c, float eta) { return fresnelDielectricReflectance(c, eta); }
float oracle_fresnel_metal(float c, float n, float k) { return fresnelMetalReflectance(c, n, k); }
float oracle_microfacet_eval(const float* m, float r) { return microfacetEval(v3(m[0], m[1], m[2]), r); }
float oracle_smith_g1(const float* v, const float* m, float r) { return smithG1(v3(v[0], v[1], v[2]), v3(m[0], m[1], m[2]), r); }
float oracle_sdf_eval(int material, const float* p) { return evalSdf(sdfs[material], v3(p[0], p[1], p[2])); }
// algorithmic FLOPs (and other VALU ops) of one SDF evaluation at p
void oracle_sdf_count(int material, const float* p, long long* flops, long long* ops) {
  Counted::flops = 0;
  Counted::ops = 0;
  Counted x[3] = {cf(p[0]), cf(p[1]), cf(p[2])};
  evalSdfT<Counted>(sdfs[material], x);
  *flops = Counted::flops;
  *ops = Counted::ops;
}

// bsdf evaluate / pdf / sample in the local frame at wavelength wl (dispersive rgb from the LUT w)
void oracle_bsdf(int material, float wl, const float* rgb, const float* wi, const float* wo, float* f3, float* pdf) {
  Tracer<V3> T;
  T.wl = wl;
  T.rgb = v3(rgb[0], rgb[1], rgb[2]);
  Basis b = makeBasis(v3(0, 0, 1));
  V3 f = T.evaluateBsdf(v3(0, 0, 0), b, v3(wi[0], wi[1], wi[2]), v3(wo[0], wo[1], wo[2]), material);
  f3[0] = f.x; f3[1] = f.y; f3[2] = f.z;
  *pdf = T.pdfBsdf(v3(0, 0, 0), b, v3(wi[0], wi[1], wi[2]), v3(wo[0], wo[1], wo[2]), material);
}
void oracle_bsdf_sample(int material, float wl, const float* wi, float* rnd4, float* wo, float* f3, float* pdf) {
  Tracer<V3> T;
  T.wl = wl;
  T.rgb = v3(0.0f);
  Basis b = makeBasis(v3(0, 0, 1));
  V3 w = v3(0.0f);
  float p = 1.0f;
  V3 f = T.sampleBsdf(v3(0, 0, 0), b, v3(wi[0], wi[1], wi[2]), material, w, p, rnd4);
  wo[0] = w.x; wo[1] = w.y; wo[2] = w.z;
  f3[0] = f.x; f3[1] = f.y; f3[2] = f.z;
  *pdf = p;
}

void oracle_seed(float* rng, int npix, unsigned seed) {
  const uint32_t m[4] = {4194287u, 4194277u, 4194191u, 4194167u};
  for (int i = 0; i < npix; ++i) {
    uint64_t state = 0, inc = (54ull << 1u) | 1u;
    pcg_next(state, inc);
    state += (uint64_t)((uint32_t)i ^ seed);
    pcg_next(state, inc);
    for (int k = 0; k < 4; ++k) rng[4 * i + k] = float(1u + pcg_next(state, inc) % (m[k] - 1u));
  }
}

// pt frames for the pixels i with i % stride == offset (stride 1 = all)
int oracle_render(float* radiance, float* rng, int W, int H, int frames, int stride, int offset, int threads) {
  long long npix = (long long)W * H;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  if (metalModel == 0 && hasMat(MAT_METAL)) {
    // K_METAL reads the radiance as of the frame start: frames in order
    unit0W = W;
    unit0H = H;
    for (int f = 0; f < frames; ++f) {
      unit0Radiance.assign(radiance, radiance + 4 * npix);
#ifdef _OPENMP
#pragma omp parallel
#endif
      {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (long long i = offset; i < npix; i += stride) {
          renderPixel((int)(i % W), (int)(i / W), radiance + 4 * i, rng + 4 * i);
        }
        COUNT_FLUSH();
      }
    }
    return 0;
  }
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (long long i = offset; i < npix; i += stride) {
      int ix = (int)(i % W), iy = (int)(i / W);
      for (int f = 0; f < frames; ++f) renderPixel(ix, iy, radiance + 4 * i, rng + 4 * i);
    }
    COUNT_FLUSH();
  }
  return 0;
}

int oracle_normals(float* radiance, float* rng, int W, int H, int frames, int threads) {
  long long npix = (long long)W * H;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (long long i = 0; i < npix; ++i)
    for (int f = 0; f < frames; ++f) normalsPixel((int)(i % W), (int)(i / W), radiance + 4 * i, rng + 4 * i);
  return 0;
}

// pick program (pick-fragment-shader.glsl:116-152)
int oracle_pick(const float* xy, int n, float* out) {
  for (int k = 0; k < n; ++k) {
    float dummy[4] = {1, 1, 1, 1};
    V3 s, d;
    float ap = P.camAperture;
    P.camAperture = 0.0f;
    constructPrimaryRay(xy[2 * k], xy[2 * k + 1], dummy, s, d, false);
    P.camAperture = ap;
    // pick-fragment-shader.glsl:58-113 with its HAS_VOLUME branch (74, 98-99):
    // the medium's bound is pickable as material 3 (MAT_VOLUM)
    float sdf = 1.0e20f;
    for (int m : {MAT_SURFA, MAT_METAL, MAT_DIELE}) if (hasMat(m)) sdf = fmin_(sdf, std::fabs(evalSdf(sdfs[m], s)));
    if (C.medium) sdf = fmin_(sdf, std::fabs(mediumBoundHook(s)));
    float InitialSign = fsign(sdf), t = 0.0f;
    int iters = 0, material = MAT_INVAL;
    for (int q = 0; q < C.maxMarch; q++) {
      t += InitialSign * sdf;
      if (t >= P.maxLengthScale) break;
      V3 pW = s + t * d;
      sdf = 1.0e20f;
      bool h = false;
      for (int m : {MAT_SURFA, MAT_METAL, MAT_DIELE}) {
        if (!hasMat(m)) continue;
        float v = std::fabs(evalSdf(sdfs[m], pW));
        if (v < P.minLengthScale) { material = m; h = true; break; }
        sdf = fmin_(sdf, v);
      }
      if (h) break;
      if (C.medium) {
        float v = std::fabs(mediumBoundHook(pW));
        if (v < P.minLengthScale) { material = 3; break; }
        sdf = fmin_(sdf, v);
      }
      iters++;
    }
    bool isHit = !(t >= P.maxLengthScale) && !(P.maxStepsIsMiss != 0.0f && iters >= C.maxMarch);
    float dist = P.maxLengthScale;
    if (isHit) dist = length((s + t * d) - s);
    out[2 * k] = dist;
    out[2 * k + 1] = float(material);
  }
  return 0;
}

// GLSL ES 3.00 builtins as defined: step(edge, x) = x < edge ? 0 : 1,
// mix(x, y, a) = x (1 - a) + y a, fract(x) = x - floor(x), min(x, y) = y < x ? y : x
static float glStep(float edge, float x) { return x < edge ? 0.0f : 1.0f; }
static float glMix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static float glFract(float x) { return x - std::floor(x); }
static float glMin(float x, float y) { return y < x ? y : x; }

// rgb2hsv (tonemapper-fragment-shader.glsl:19-27)
static void rgb2hsv(const float c[3], float hsv[3]) {
  const float K[4] = {0.0f, -1.0f / 3.0f, 2.0f / 3.0f, -1.0f};
  float t = glStep(c[2], c[1]);
  const float a[4] = {c[2], c[1], K[3], K[2]}, b[4] = {c[1], c[2], K[0], K[1]};
  float p[4];
  for (int k = 0; k < 4; ++k) p[k] = glMix(a[k], b[k], t);
  float u = glStep(p[0], c[0]);
  const float a2[4] = {p[0], p[1], p[3], c[0]}, b2[4] = {c[0], p[1], p[2], p[0]};
  float q[4];
  for (int k = 0; k < 4; ++k) q[k] = glMix(a2[k], b2[k], u);
  float d = q[0] - glMin(q[3], q[1]);
  float e = 1.0e-10f;
  hsv[0] = std::fabs(q[2] + (q[3] - q[1]) / (6.0f * d + e));
  hsv[1] = d / (q[0] + e);
  hsv[2] = q[0];
}

// hsv2rgb (tonemapper-fragment-shader.glsl:29-34)
static void hsv2rgb(const float c[3], float rgb[3]) {
  const float K[4] = {1.0f, 2.0f / 3.0f, 1.0f / 3.0f, 3.0f};
  for (int k = 0; k < 3; ++k) {
    float p = std::fabs(glFract(c[0] + K[k]) * 6.0f - K[3]);
    float cl = std::fmin(std::fmax(p - K[0], 0.0f), 1.0f);
    rgb[k] = c[2] * glMix(K[0], cl, c[1]);
  }
}

// tonemapper (tonemapper-fragment-shader.glsl:36-97)
int oracle_tonemap(const float* radiance, long long npix, float* rgb) {
  auto tm = [](float L) { return L / (1.0f + L); };
  for (long long i = 0; i < npix; ++i) {
    float X = radiance[4 * i], Y = radiance[4 * i + 1], Z = radiance[4 * i + 2];
    float c[3] = {3.2406f * X - 1.5372f * Y - 0.4986f * Z, -0.9689f * X + 1.8758f * Y + 0.0415f * Z,
                  0.0557f * X - 0.2040f * Y + 1.0570f * Z};
    float delta = -fmin_(0.0f, fmin_(fmin_(c[0], c[1]), c[2]));
    for (float& v : c) v += delta;
    for (float& v : c) v = std::pow(v, P.invGamma);
    float e = std::pow(2.0f, P.exposure);
    for (float& v : c) v = tm(v * e);
    float mean = (c[0] + c[1] + c[2]) / 3.0f;
    for (float& v : c) { float dv = v - mean; v = mean + fsign(dv) * std::pow(std::fabs(dv), 1.0f / P.saturation); }
    for (float& v : c) { float dv = v - 0.5f; v = 0.5f + fsign(dv) * std::pow(std::fabs(dv), 1.0f / P.contrast); }
    if (P.hueShift > 0.0f) {  // (85-91)
      float hsv[3];
      rgb2hsv(c, hsv);
      float h = hsv[0] + P.hueShift;
      hsv[0] = h - 1.0f * std::floor(h / 1.0f);  // mod(h, 1.0)
      hsv2rgb(hsv, c);
    }
    rgb[3 * i] = c[0]; rgb[3 * i + 1] = c[1]; rgb[3 * i + 2] = c[2];
  }
  return 0;
}

}  // extern "C"
