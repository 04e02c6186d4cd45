// snl_pt.h -- wavefront spectral SDF path tracer for gfx950 (wave64), compiled
// at run time with hipRTC together with one translated scene (struct SnlScene).
//
// Reference algorithm: shaders/pathtracer-fragment-shader.glsl (cameraPath
// 1625-1926, pathtrace 1934-1989, main 1991-2009) re-organised as a wavefront:
//
//   snl_camera      pathtrace() up to the first traceRay: interactive skip,
//                   hero wavelength, pixel jitter, primary ray      (1934-1970)
//   snl_trace       persistent sphere tracer with per-wave lane refill; runs
//                   the continuation ray of each live path (traceRay 881-920)
//                   and the deferred shadow rays of each NEE job
//                   (directSurfaceLighting 1221-1295), then bins paths by the
//                   material they hit
//   snl_shade_*     the body of the vertex loop after traceRay, one kernel per
//                   material (miss / surface / metal / dielectric)
//   snl_accumulate  running-mean XYZ update + RNG write-back   (1972-1988)
//
// Random numbers are drawn per path in exactly the reference's call order, and
// shadow-ray results are folded into L before the next vertex's contributions,
// so each pixel's estimator (including its float summation order) is the
// reference's.  Shadow rays can be deferred because no random number drawn in
// directSurfaceLighting depends on a shadow-ray outcome.
//
// All path state is SoA in HBM: field k of path i lives at pf[k*cap + i].
#pragma once

namespace glsl {

#ifdef DISPERSION_ENABLED
typedef float RT;
SNL_HD float rt1(float x) { return x; }
#define SNL_RTN 1
#else
typedef vec3 RT;
SNL_HD vec3 rt1(float x) { return vec3(x); }
#define SNL_RTN 3
#endif

SNL_HD float maxComponent(const vec3& v) { return max(max(v.r, v.g), v.b); }
SNL_HD float maxComponent(float f) { return f; }
SNL_HD float averageComponent(const vec3& v) { return (v.r + v.g + v.b) / 3.0f; }
SNL_HD float averageComponent(float f) { return f; }
// whether a radiance value has a nonzero component (NaN counts as nonzero)
SNL_HD bool anyNonzero(const vec3& v) { return v.r != 0.0f || v.g != 0.0f || v.b != 0.0f; }
SNL_HD bool anyNonzero(float f) { return f != 0.0f; }

struct Basis { vec3 nW, tW, bW; };

SNL_HD vec3 safe_normalize(const vec3& N) { float l = length(N); return N / max(l, DENOM_TOLERANCE); }
SNL_HD float cosTheta2(const vec3& n) { return n.z * n.z; }
SNL_HD float cosTheta(const vec3& n) { return n.z; }
SNL_HD float sinTheta2(const vec3& n) { return 1.0f - cosTheta2(n); }
SNL_HD float tanTheta2(const vec3& n) { float ct2 = cosTheta2(n); return max(0.0f, 1.0f - ct2) / max(ct2, 1.0e-7f); }
SNL_HD float tanTheta(const vec3& n) { return sqrt(max(0.0f, tanTheta2(n))); }

// pathtracer-fragment-shader.glsl:175-194 (branch on |n.z| < |n.x|)
SNL_HD Basis makeBasis(const vec3& nW) {
  Basis b;
  b.nW = nW;
  if (abs(nW.z) < abs(nW.x)) b.tW = vec3(nW.z, 0.0f, -nW.x);
  else b.tW = vec3(0.0f, nW.z, -nW.y);
  b.tW = safe_normalize(b.tW);
  b.bW = cross(nW, b.tW);
  return b;
}
// normals-fragment-shader.glsl:160-180 uses plain normalize in the same frame
SNL_HD Basis makeBasisPlain(const vec3& nW) {
  Basis b;
  b.nW = nW;
  if (abs(nW.z) < abs(nW.x)) b.tW = vec3(nW.z, 0.0f, -nW.x);
  else b.tW = vec3(0.0f, nW.z, -nW.y);
  b.tW = normalize(b.tW);
  b.bW = cross(nW, b.tW);
  return b;
}
SNL_HD vec3 worldToLocal(const vec3& v, const Basis& b) { return vec3(dot(v, b.tW), dot(v, b.bW), dot(v, b.nW)); }
SNL_HD vec3 localToWorld(const vec3& v, const Basis& b) { return b.tW * v.x + b.bW * v.y + b.nW * v.z; }

// linear sRGB <-> XYZ (pathtracer-fragment-shader.glsl:209-236; 7-digit matrices)
SNL_HD vec3 xyzToRgb(const vec3& XYZ) {
  vec3 RGB;
  RGB.r = 3.2404542f * XYZ.x - 1.5371385f * XYZ.y - 0.4985314f * XYZ.z;
  RGB.g = -0.9692660f * XYZ.x + 1.8760108f * XYZ.y + 0.0415560f * XYZ.z;
  RGB.b = 0.0556434f * XYZ.x - 0.2040259f * XYZ.y + 1.0572252f * XYZ.z;
  float delta = -min(0.0f, min(min(RGB.r, RGB.g), RGB.b));
  RGB.r += delta; RGB.g += delta; RGB.b += delta;
  float sum = RGB.r + RGB.g + RGB.b;
  RGB /= sum;
  return RGB;
}
SNL_HD vec3 rgbToXyz(const vec3& RGB) {
  vec3 XYZ;
  XYZ.x = 0.4124564f * RGB.r + 0.3575761f * RGB.g + 0.1804375f * RGB.b;
  XYZ.y = 0.2126729f * RGB.r + 0.7151522f * RGB.g + 0.0721750f * RGB.b;
  XYZ.z = 0.0193339f * RGB.r + 0.1191920f * RGB.g + 0.9503041f * RGB.b;
  return XYZ;
}
#ifdef DISPERSION_ENABLED
SNL_HD float rgbToAlbedo(const vec3& RGB, const vec3& rgb) { return dot(RGB, rgb); }
#else
SNL_HD vec3 rgbToAlbedo(const vec3& RGB, const vec3& /*rgb*/) { return RGB; }
#endif

// ------------------------------------------------------------ sampling (300-372)
SNL_HD vec3 sampleSphereUniformly(vec4& rnd) {
  float z = 1.0f - 2.0f * snl_rand(rnd);
  float r = sqrt(max(0.0f, 1.0f - z * z));
  float phi = 2.0f * float(M_PI) * snl_rand(rnd);
  (void)r;  // the reference leaves x and y unscaled by r: the direction is not unit length
  return vec3(cos(phi), sin(phi), z);
}
SNL_HD vec3 sampleHemisphereUniformly(vec4& rnd) {
  float xi = snl_rand(rnd);
  float r = sqrt(1.0f - xi * xi);
  float phi = 2.0f * float(M_PI) * snl_rand(rnd);
  return vec3(r * cos(phi), r * sin(phi), xi);
}
SNL_HD vec3 sampleHemisphereCosineWeighted(vec4& rnd, float& pdf) {
  float r = sqrt(snl_rand(rnd));
  float theta = 2.0f * float(M_PI) * snl_rand(rnd);
  float x = r * cos(theta);
  float y = r * sin(theta);
  float z = sqrt(max(0.0f, 1.0f - x * x - y * y));
  pdf = abs(z) / float(M_PI);
  return vec3(x, y, z);
}
SNL_HD float pdfHemisphereCosineWeighted(const vec3& wiL) {
  if (wiL.z < 0.0f) return 0.0f;
  return wiL.z / float(M_PI);
}
SNL_HD float powerHeuristic(float a, float b) { return a / (a + b); }

SNL_HD float phaseFunction(float mu, float g) {
  float gSqr = g * g;
  return (1.0f / (4.0f * float(M_PI))) * (1.0f - gSqr) / pow(1.0f - 2.0f * g * mu + gSqr, 1.5f);
}
SNL_HD vec3 samplePhaseFunction(const vec3& rayDir, float g, vec4& rnd) {
  float costheta;
  if (abs(g) < 1.0e-3f) costheta = 1.0f - 2.0f * snl_rand(rnd);
  else costheta = 1.0f / (2.0f * g) * (1.0f + g * g - ((1.0f - g * g) * (1.0f - g + 2.0f * g * snl_rand(rnd))));
  float sintheta = sqrt(max(0.0f, 1.0f - costheta * costheta));
  float phi = 2.0f * float(M_PI) * snl_rand(rnd);
  Basis basis = makeBasis(rayDir);
  return costheta * rayDir + sintheta * (cos(phi) * basis.tW + sin(phi) * basis.bW);
}
// Henyey-Greenstein sampling by inversion of its CDF in cos(theta) (the
// published form, e.g. Pharr, Jakob & Humphreys, Physically Based Rendering,
// 3rd ed., 15.2.3): the heterogeneous medium's phase sampling.  Returns wo
// with cos(theta) = dot(dir, wo) distributed as phaseFunction(., g), and that
// density in pdf.  Draws: cos(theta), then the azimuth.
SNL_HD vec3 sampleHenyeyGreenstein(const vec3& dir, float g, vec4& rnd, float& pdf) {
  float xi = snl_rand(rnd);
  float cosTheta;
  if (abs(g) < 1.0e-3f) cosTheta = 1.0f - 2.0f * xi;
  else {
    float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * xi);
    cosTheta = (1.0f + g * g - sq * sq) / (2.0f * g);
  }
  cosTheta = clamp(cosTheta, -1.0f, 1.0f);
  float sinTheta = sqrt(max(0.0f, 1.0f - cosTheta * cosTheta));
  float phi = 2.0f * float(M_PI) * snl_rand(rnd);
  Basis b = makeBasis(dir);
  pdf = phaseFunction(cosTheta, g);
  return localToWorld(vec3(sinTheta * cos(phi), sinTheta * sin(phi), cosTheta), b);
}
// a uniformly distributed unit direction (density 1 / 4 pi): the sky sample of
// the heterogeneous medium's direct lighting
SNL_HD vec3 sampleSphereUnit(vec4& rnd) {
  float z = 1.0f - 2.0f * snl_rand(rnd);
  float r = sqrt(max(0.0f, 1.0f - z * z));
  float phi = 2.0f * float(M_PI) * snl_rand(rnd);
  return vec3(r * cos(phi), r * sin(phi), z);
}

// ------------------------------------------------------------ Beckmann microfacet (379-441)
SNL_HD float microfacetEval(const vec3& m, float roughness) {
  float t2 = tanTheta2(m);
  float c2 = cosTheta2(m);
  float roughnessSqr = roughness * roughness;
  float epsilon = 1.0e-9f;
  float exponent = t2 / max(roughnessSqr, epsilon);
  return exp(-exponent) / max(float(M_PI) * roughnessSqr * c2 * c2, epsilon);
}
SNL_HD vec3 microfacetSample(vec4& rnd, float roughness) {
  float phiM = (2.0f * float(M_PI)) * snl_rand(rnd);
  float cosPhiM = cos(phiM);
  float sinPhiM = sin(phiM);
  float epsilon = 1.0e-9f;
  float tanThetaMSqr = -roughness * roughness * log(max(epsilon, snl_rand(rnd)));
  float cosThetaM = 1.0f / sqrt(1.0f + tanThetaMSqr);
  float sinThetaM = sqrt(max(0.0f, 1.0f - cosThetaM * cosThetaM));
  return safe_normalize(vec3(sinThetaM * cosPhiM, sinThetaM * sinPhiM, cosThetaM));
}
SNL_HD float microfacetPDF(const vec3& m, float roughness) { return microfacetEval(m, roughness) * abs(cosTheta(m)); }
// Walter et al. 2007 rational approximation of the Beckmann Smith G1
SNL_HD float smithG1(const vec3& vLocal, const vec3& mLocal, float roughness) {
  float tanThetaAbs = abs(tanTheta(vLocal));
  float epsilon = 1.0e-6f;
  if (tanThetaAbs < epsilon) return 1.0f;
  if (dot(vLocal, mLocal) * vLocal.z <= 0.0f) return 0.0f;
  float a = 1.0f / (max(roughness, epsilon) * tanThetaAbs);
  if (a >= 1.6f) return 1.0f;
  float aSqr = a * a;
  return (3.535f * a + 2.181f * aSqr) / (1.0f + 2.276f * a + 2.577f * aSqr);
}
SNL_HD float smithG2(const vec3& woL, const vec3& wiL, const vec3& mLocal, float roughness) {
  return smithG1(woL, mLocal, roughness) * smithG1(wiL, mLocal, roughness);
}
SNL_HD float fresnelDielectricReflectance(float cosi, float eta_ti) {
  float c = cosi;
  float g = eta_ti * eta_ti - 1.0f + c * c;
  if (g > 0.0f) {
    g = sqrt(g);
    float A = (g - c) / (g + c);
    float B = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * A * A * (1.0f + B * B);
  }
  return 1.0f;
}
SNL_HD float fresnelMetalReflectance(float cosi, float ior, float k) {
  float cosip = abs(cosi);
  float cosi2 = cosip * cosip;
  float iikk = ior * ior + k * k;
  float iikkc = iikk * cosi2;
  float twoEtaCosi = 2.0f * ior * cosip;
  float Rparl2 = (iikkc - twoEtaCosi + 1.0f) / (iikkc + twoEtaCosi + 1.0f);
  float Rperp2 = (iikk - twoEtaCosi + cosi2) / (iikk + twoEtaCosi + cosi2);
  return 0.5f * (Rparl2 + Rperp2);
}
SNL_HD bool refraction_given_transmitted_direction(const vec3& n, float eta_ti, const vec3& wt, vec3& wi) {
  float wtn = dot(wt, n);
  float disc = 1.0f - eta_ti * eta_ti * (1.0f - wtn * wtn);
  if (disc < 0.0f) return false;
  wi = eta_ti * wt - n * sign(wtn) * (eta_ti * abs(wtn) - sqrt(disc));
  return true;
}
SNL_HD bool refraction_given_incident_direction(const vec3& n, float eta_ti, const vec3& wi, vec3& wt) {
  float win = dot(wi, n);
  float disc = 1.0f - (1.0f - win * win) / (eta_ti * eta_ti);
  if (disc < 0.0f) return false;
  wt = wi / eta_ti - n * sign(win) * (abs(win) / eta_ti - sqrt(disc));
  return true;
}

// ------------------------------------------------------------ path tracer bound to a scene
template <class SC>
struct PT {
  SC& S;
  const SnlParams& P;
  SNL_HD PT(SC& s, const SnlParams& p) : S(s), P(p) {}
  // SDF evaluations of traceRay, for the step statistics
  unsigned long long steps = 0;

  // ---------------- SDF marching (traceRay, 881-920)
  // Unified single-step form used by the persistent tracer: evaluates the
  // enabled SDFs at p; returns the material hit (|sdf| < minMarch, priority
  // surface -> metal -> dielectric) or MAT_INVAL, and the min |sdf|.
  SNL_HD int sdfStep(const vec3& p, bool checkHit, float& sdf) {
    float minMarch = P.minLengthScale;
    sdf = 1.0e20f;
    if constexpr (SC::kHasSurface) {
      float s = abs(S.SDF_SURFACE(p));
      if (checkHit && s < minMarch) return MAT_SURFA;
      sdf = min(sdf, s);
    }
    if constexpr (SC::kHasMetal) {
      float s = abs(S.SDF_METAL(p));
      if (checkHit && s < minMarch) return MAT_METAL;
      sdf = min(sdf, s);
    }
    if constexpr (SC::kHasDielectric) {
      float s = abs(S.SDF_DIELECTRIC(p));
      if (checkHit && s < minMarch) return MAT_DIELE;
      sdf = min(sdf, s);
    }
    return MAT_INVAL;
  }

  // Whole traceRay, used where a single trace is inline (SSS walk, volume emission).
  SNL_HD bool traceRay(const vec3& start, const vec3& dir, vec3& hit, int& material, float maxMarchDist) {
    material = MAT_INVAL;
    float sdf;
    sdfStep(start, false, sdf);
    ++steps;
    float InitialSign = sign(sdf);
    float t = InitialSign * sdf;
    if (abs(t) >= maxMarchDist) return false;
    for (int n = 0; n < SC::kMaxMarchSteps; n++) {
      vec3 pW = start + t * dir;
      int m = sdfStep(pW, true, sdf);
      ++steps;
      if (m != MAT_INVAL) { material = m; hit = start + t * dir; return true; }
      t += InitialSign * sdf;
      if (abs(t) >= maxMarchDist) return false;
    }
    hit = start + t * dir;
    return !P.maxStepsIsMiss;
  }

  // First-hit marcher of the normals and pick programs (traceDistance,
  // normals-fragment-shader.glsl:70-113): t starts at 0 and the step precedes
  // the evaluation; `iters` counts non-hit evaluations.  `volume`: the pick
  // program's HAS_VOLUME branch (pick-fragment-shader.glsl:74-100), where the
  // medium's bound SDF_VOLUME is hit last in priority as MAT_VOLUM.
  SNL_HD bool traceDistance(const vec3& start, const vec3& dir, float maxDist, vec3& hit, int& material,
                            bool volume = false) {
    material = MAT_INVAL;
    float sdf;
    sdfStep(start, false, sdf);
    if constexpr (SC::kHasVolume) if (volume) sdf = min(sdf, abs(S.SDF_VOLUME(start)));
    float InitialSign = sign(sdf);
    float t = 0.0f;
    int iters = 0;
    for (int n = 0; n < SC::kMaxMarchSteps; n++) {
      t += InitialSign * sdf;
      if (t >= maxDist) break;
      vec3 pW = start + t * dir;
      int m = sdfStep(pW, true, sdf);
      if (m != MAT_INVAL) { material = m; break; }
      if constexpr (SC::kHasVolume) {
        if (volume) {
          float sv = abs(S.SDF_VOLUME(pW));
          if (sv < P.minLengthScale) { material = MAT_VOLUM; break; }
          sdf = min(sdf, sv);
        }
      }
      iters++;
    }
    hit = start + t * dir;
    if (t >= maxDist) return false;
    if (P.maxStepsIsMiss && iters >= SC::kMaxMarchSteps) return false;
    return true;
  }

  SNL_HD float sdfOf(const vec3& p, int material) {
    if constexpr (SC::kHasSurface) if (material == MAT_SURFA) return S.SDF_SURFACE(p);
    if constexpr (SC::kHasMetal) if (material == MAT_METAL) return S.SDF_METAL(p);
    if constexpr (SC::kHasDielectric) if (material == MAT_DIELE) return S.SDF_DIELECTRIC(p);
    return 0.0f;
  }
  // central-difference gradient (922-940); `safe` selects safe_normalize (pathtracer)
  // or normalize (normals program, 133-151)
  SNL_HD vec3 normal(const vec3& pW, int material, bool safe) {
    float normalEpsilon = 2.0f * P.minLengthScale;
    vec3 e(normalEpsilon, 0.0f, 0.0f);
    vec3 Xp = pW + vec3(e.x, e.y, e.y), Xn = pW - vec3(e.x, e.y, e.y);
    vec3 Yp = pW + vec3(e.y, e.x, e.y), Yn = pW - vec3(e.y, e.x, e.y);
    vec3 Zp = pW + vec3(e.y, e.y, e.x), Zn = pW - vec3(e.y, e.y, e.x);
    vec3 N(sdfOf(Xp, material) - sdfOf(Xn, material), sdfOf(Yp, material) - sdfOf(Yn, material),
           sdfOf(Zp, material) - sdfOf(Zn, material));
    return safe ? safe_normalize(N) : normalize(N);
  }

  SNL_HD vec3 perturbNormal(const vec3& X, const Basis& basis, int material) {
    if constexpr (SC::kHasSurfaceNM) if (material == MAT_SURFA) return localToWorld(normalize(vec3(S.SURFACE_NORMAL_MAP(X, basis.nW))), basis);
    if constexpr (SC::kHasMetalNM) if (material == MAT_METAL) return localToWorld(normalize(vec3(S.METAL_NORMAL_MAP(X, basis.nW))), basis);
    if constexpr (SC::kHasDielectricNM) if (material == MAT_DIELE) return localToWorld(normalize(vec3(S.DIELECTRIC_NORMAL_MAP(X, basis.nW))), basis);
    return basis.nW;
  }

  // ---------------- Dielectric (477-633)
  SNL_HD RT DIELECTRIC_SPEC_REFL_EVAL(const vec3& X, const vec3& winputL, const Basis& basis, const vec3& rgb) {
    vec3 winputW = localToWorld(winputL, basis);
    vec3 reflRGB = S.DIELECTRIC_SPECULAR_REFLECTANCE(P.dieleSpecAlbedoRGB, X, basis.nW, winputW);
    return rgbToAlbedo(reflRGB, rgb);
  }
  SNL_HD RT DIELECTRIC_ABSORPTION_EVAL(const vec3& X, const vec3& rgb) {
    vec3 absorptionRGB = S.DIELECTRIC_ABSORPTION(P.dieleAbsorptionRGB, X);
    return rgbToAlbedo(absorptionRGB, rgb);
  }
  SNL_HD RT evaluateDielectric(const vec3& X, const Basis& basis, const vec3& winputL, const vec3& woutputL,
                               float wavelength_nm, const vec3& rgb, bool fromCamera) {
    float ior = S.IOR_DIELE(wavelength_nm);
    bool reflected = cosTheta(woutputL) * cosTheta(winputL) > 0.0f;
    RT dielectricAlbedo = DIELECTRIC_SPEC_REFL_EVAL(X, winputL, basis, rgb);
    vec3 beamIncidentL = fromCamera ? -woutputL : -winputL;
    vec3 beamOutgoingL = fromCamera ? winputL : woutputL;
    bool entering = (beamIncidentL.z < 0.0f);
    vec3 h;
    if (reflected) h = safe_normalize(woutputL + winputL);
    else if (entering) h = safe_normalize(beamIncidentL - ior * beamOutgoingL);
    else h = safe_normalize(ior * beamIncidentL - beamOutgoingL);
    if (cosTheta(h) < 0.0f) h *= -1.0f;
    float roughness = S.DIELECTRIC_ROUGHNESS(P.dieleRoughness, X, basis.nW);
    float D = microfacetEval(h, roughness);
    float G = smithG2(-beamIncidentL, beamOutgoingL, h, roughness);
    RT f;
    float eta_ti = entering ? ior : 1.0f / ior;
    RT Fr = dielectricAlbedo * fresnelDielectricReflectance(abs(winputL.z), eta_ti);
    if (reflected) {
      f = dielectricAlbedo * Fr * D * G / max(4.0f * abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
    } else {
      float im = dot(-beamIncidentL, h);
      float om = dot(beamOutgoingL, h);
      float sqrtDenom = im + eta_ti * om;
      float dwh_dwo = eta_ti * eta_ti * abs(om) / max(sqrtDenom * sqrtDenom, DENOM_TOLERANCE);
      f = (rt1(1.0f) - Fr) * G * D * abs(im) * dwh_dwo /
          max(abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
    }
    return f;
  }
  SNL_HD float pdfDielectric(const vec3& X, const Basis& basis, const vec3& winputL, const vec3& woutputL,
                             float wavelength_nm, const vec3& rgb, bool fromCamera) {
    float ior = S.IOR_DIELE(wavelength_nm);
    bool reflected = cosTheta(woutputL) * cosTheta(winputL) > 0.0f;
    RT dielectricAlbedo = DIELECTRIC_SPEC_REFL_EVAL(X, winputL, basis, rgb);
    vec3 beamIncidentL = fromCamera ? -woutputL : -winputL;
    vec3 beamOutgoingL = fromCamera ? winputL : woutputL;
    bool entering = (beamIncidentL.z < 0.0f);
    float eta_ti = entering ? ior : 1.0f / ior;
    RT Fr = dielectricAlbedo * fresnelDielectricReflectance(abs(winputL.z), eta_ti);
    vec3 h;
    float dwh_dwo;
    float pdf;
    if (reflected) {
      h = safe_normalize(woutputL + winputL);
      dwh_dwo = 1.0f / max(4.0f * abs(dot(winputL, h)), DENOM_TOLERANCE);
      pdf = averageComponent(Fr);
    } else {
      if (entering) h = safe_normalize(beamIncidentL - ior * beamOutgoingL);
      else h = safe_normalize(ior * beamIncidentL - beamOutgoingL);
      if (cosTheta(h) < 0.0f) h *= -1.0f;
      float im = dot(-beamIncidentL, h);
      float om = dot(beamOutgoingL, h);
      float sqrtDenom = im + eta_ti * om;
      dwh_dwo = eta_ti * eta_ti * abs(om) / max(sqrtDenom * sqrtDenom, DENOM_TOLERANCE);
      pdf = 1.0f - averageComponent(Fr);
    }
    float roughness = S.DIELECTRIC_ROUGHNESS(P.dieleRoughness, X, basis.nW);
    pdf *= microfacetPDF(h, roughness);
    return abs(pdf * dwh_dwo);
  }
  SNL_HD RT sampleDielectric(const vec3& X, const Basis& basis, const vec3& winputL, float wavelength_nm,
                             const vec3& rgb, bool fromCamera, vec3& woutputL, float& pdfOut, vec4& rnd) {
    float ior = S.IOR_DIELE(wavelength_nm);
    RT dielectricAlbedo = DIELECTRIC_SPEC_REFL_EVAL(X, winputL, basis, rgb);
    float eta_ti_refl = (winputL.z >= 0.0f) ? ior : 1.0f / ior;
    RT Fr = dielectricAlbedo * fresnelDielectricReflectance(abs(winputL.z), eta_ti_refl);
    float roughness = S.DIELECTRIC_ROUGHNESS(P.dieleRoughness, X, basis.nW);
    vec3 m = microfacetSample(rnd, roughness);
    float microPDF = microfacetPDF(m, roughness);
    float reflectProb = averageComponent(Fr);
    if (snl_rand(rnd) < reflectProb) {
      woutputL = -winputL + 2.0f * dot(winputL, m) * m;
      if (woutputL.z * winputL.z < 0.0f) woutputL *= -1.0f;
      float D = microfacetEval(m, roughness);
      float G = smithG2(winputL, woutputL, m, roughness);
      RT f = Fr * D * G / max(4.0f * abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
      float dwh_dwo = 1.0f / max(abs(4.0f * dot(winputL, m)), DENOM_TOLERANCE);
      pdfOut = microPDF * reflectProb * dwh_dwo;
      return f;
    } else {
      vec3 beamIncidentL = fromCamera ? -woutputL : -winputL;
      vec3 beamOutgoingL = fromCamera ? winputL : woutputL;
      bool entering = (beamIncidentL.z < 0.0f);
      float eta_ti = entering ? ior : 1.0f / ior;
      if (fromCamera) {
        if (!refraction_given_transmitted_direction(m, eta_ti, beamOutgoingL, beamIncidentL)) return rt1(0.0f);
        woutputL = -beamIncidentL;
      } else {
        beamIncidentL = -winputL;
        if (!refraction_given_incident_direction(m, eta_ti, beamIncidentL, beamOutgoingL)) return rt1(0.0f);
        woutputL = beamOutgoingL;
      }
      float im = dot(-beamIncidentL, m);
      RT Frm = rt1(fresnelDielectricReflectance(abs(im), eta_ti));
      RT Trm = rt1(1.0f) - dielectricAlbedo * Frm;
      vec3 wh = m;
      float D = microfacetEval(wh, roughness);
      float G = smithG2(-beamIncidentL, beamOutgoingL, wh, roughness);
      float om = dot(beamOutgoingL, m);
      float sqrtDenom = im + eta_ti * om;
      float dwh_dwo = eta_ti * eta_ti * abs(om) / max(sqrtDenom * sqrtDenom, DENOM_TOLERANCE);
      RT f = RT(abs(im) * dwh_dwo * Trm * G * D) / max(abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
      pdfOut = (1.0f - reflectProb) * microPDF * abs(dwh_dwo);
      return f;
    }
  }

  // ---------------- Metal (643-708)
  SNL_HD RT METAL_SPEC_REFL_EVAL(const vec3& X, const vec3& winputL, const Basis& basis, const vec3& rgb) {
    vec3 winputW = localToWorld(winputL, basis);
    vec3 reflRGB = S.METAL_SPECULAR_REFLECTANCE(P.metalSpecAlbedoRGB, X, basis.nW, winputW);
    return rgbToAlbedo(reflRGB, rgb);
  }
  SNL_HD RT evaluateMetal(const vec3& X, const Basis& basis, const vec3& winputL, const vec3& woutputL,
                          float wavelength_nm, const vec3& rgb) {
    if (winputL.z < 0.0f) return rt1(0.0f);
    float ior = S.IOR_METAL(wavelength_nm);
    float k = S.K_METAL(wavelength_nm);
    float Fr = fresnelMetalReflectance(winputL.z, ior, k);
    vec3 h = normalize(woutputL + winputL);
    float roughness = S.METAL_ROUGHNESS(P.metalRoughness, X, basis.nW);
    float D = microfacetEval(h, roughness);
    float G = smithG2(winputL, woutputL, h, roughness);
    RT specAlbedo = METAL_SPEC_REFL_EVAL(X, winputL, basis, rgb);
    return specAlbedo * Fr * D * G / max(4.0f * abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
  }
  SNL_HD float pdfMetal(const vec3& X, const Basis& basis, const vec3& winputL, const vec3& woutputL,
                        float /*wavelength_nm*/, const vec3& /*rgb*/) {
    // (the reference also evaluates IOR_DIELE and K_METAL here; the values are unused)
    if (winputL.z < 0.0f) return PDF_EPSILON;
    vec3 h = safe_normalize(woutputL + winputL);
    float dwh_dwo = 1.0f / max(abs(4.0f * dot(winputL, h)), DENOM_TOLERANCE);
    float roughness = S.METAL_ROUGHNESS(P.metalRoughness, X, basis.nW);
    return microfacetPDF(h, roughness) * dwh_dwo;
  }
  SNL_HD RT sampleMetal(const vec3& X, const Basis& basis, const vec3& winputL, float wavelength_nm,
                        const vec3& rgb, vec3& woutputL, float& pdfOut, vec4& rnd) {
    if (winputL.z < 0.0f) return rt1(0.0f);
    float ior = S.IOR_METAL(wavelength_nm);
    float k = S.K_METAL(wavelength_nm);
    float Fr = fresnelMetalReflectance(winputL.z, ior, k);
    float roughness = S.METAL_ROUGHNESS(P.metalRoughness, X, basis.nW);
    vec3 m = microfacetSample(rnd, roughness);
    woutputL = -winputL + 2.0f * dot(winputL, m) * m;
    if (woutputL.z < DENOM_TOLERANCE) woutputL.z *= -1.0f;
    float D = microfacetEval(m, roughness);
    float G = smithG2(winputL, woutputL, m, roughness);
    RT specAlbedo = METAL_SPEC_REFL_EVAL(X, winputL, basis, rgb);
    RT f = specAlbedo * Fr * D * G / max(4.0f * abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
    float dwh_dwo = 1.0f / max(abs(4.0f * dot(winputL, m)), DENOM_TOLERANCE);
    pdfOut = microfacetPDF(m, roughness) * dwh_dwo;
    return f;
  }

  // ---------------- Surface / uber (716-807)
  SNL_HD RT SURFACE_DIFFUSE_REFL_EVAL(const vec3& X, const vec3& nW, const vec3& winputW, const vec3& rgb) {
    vec3 reflRGB = S.SURFACE_DIFFUSE_REFLECTANCE(P.surfaceDiffuseAlbedoRGB, X, nW, winputW);
    return rgbToAlbedo(reflRGB, rgb);
  }
  SNL_HD RT SUBSURFACE_ALBEDO_EVAL(const vec3& X, const vec3& nW, const vec3& rgb) {
    vec3 reflRGB = S.SUBSURFACE_ALBEDO(P.subsurfaceAlbedoRGB, X, nW);
    return rgbToAlbedo(reflRGB, rgb);
  }
  SNL_HD RT SURFACE_SPEC_REFL_EVAL(const vec3& X, const vec3& nW, const vec3& winputW, const vec3& rgb) {
    vec3 reflRGB = S.SURFACE_SPECULAR_REFLECTANCE(P.surfaceSpecAlbedoRGB, X, nW, winputW);
    return rgbToAlbedo(reflRGB, rgb);
  }
  SNL_HD RT evaluateSurface(const vec3& X, const Basis& basis, const vec3& winputL, const vec3& woutputL,
                            float /*wavelength_nm*/, const vec3& rgb) {
    if (winputL.z < 0.0f) return rt1(0.0f);
    vec3 winputW = localToWorld(winputL, basis);
    RT diffuseAlbedo = (1.0f - P.subsurface) * SURFACE_DIFFUSE_REFL_EVAL(X, basis.nW, winputW, rgb);
    RT specAlbedo = SURFACE_SPEC_REFL_EVAL(X, basis.nW, winputW, rgb);
    float ior = P.surfaceIor;
    float roughness = S.SURFACE_ROUGHNESS(P.surfaceRoughness, X, basis.nW);
    float Fr = fresnelDielectricReflectance(woutputL.z, ior);
    vec3 h = normalize(woutputL + winputL);
    float D = microfacetEval(h, roughness);
    float G = smithG2(winputL, woutputL, h, roughness);
    RT f = Fr * specAlbedo * D * G / max(4.0f * abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
    float E = fresnelDielectricReflectance(winputL.z, ior);
    f += (1.0f - E) * diffuseAlbedo / float(M_PI);
    return f;
  }
  SNL_HD float pdfSurface(const vec3& X, const Basis& basis, const vec3& winputL, const vec3& woutputL,
                          float /*wavelength_nm*/, const vec3& rgb) {
    if (winputL.z < 0.0f) return PDF_EPSILON;
    vec3 winputW = localToWorld(winputL, basis);
    RT diffuseAlbedo = (1.0f - P.subsurface) * SURFACE_DIFFUSE_REFL_EVAL(X, basis.nW, winputW, rgb);
    RT specAlbedo = SURFACE_SPEC_REFL_EVAL(X, basis.nW, winputW, rgb);
    float ior = P.surfaceIor;
    float E = fresnelDielectricReflectance(abs(winputL.z), ior);
    float specWeight = (E)*averageComponent(specAlbedo);
    float diffuseWeight = (1.0f - E) * averageComponent(diffuseAlbedo);
    float weightSum = max(specWeight + diffuseWeight, DENOM_TOLERANCE);
    float specProb = specWeight / weightSum;
    float roughness = S.SURFACE_ROUGHNESS(P.surfaceRoughness, X, basis.nW);
    float diffusePdf = pdfHemisphereCosineWeighted(woutputL);
    vec3 h = safe_normalize(woutputL + winputL);
    float dwh_dwo = 1.0f / max(abs(4.0f * dot(winputL, h)), DENOM_TOLERANCE);
    float specularPdf = microfacetPDF(h, roughness) * dwh_dwo;
    return specProb * specularPdf + (1.0f - specProb) * diffusePdf;
  }
  SNL_HD RT sampleSurface(const vec3& X, const Basis& basis, const vec3& winputL, float /*wavelength_nm*/,
                          const vec3& rgb, vec3& woutputL, float& pdfOut, vec4& rnd) {
    if (winputL.z < 0.0f) return rt1(0.0f);
    vec3 winputW = localToWorld(winputL, basis);
    RT diffuseAlbedo = (1.0f - P.subsurface) * SURFACE_DIFFUSE_REFL_EVAL(X, basis.nW, winputW, rgb);
    RT specAlbedo = SURFACE_SPEC_REFL_EVAL(X, basis.nW, winputW, rgb);
    float ior = P.surfaceIor;
    float roughness = S.SURFACE_ROUGHNESS(P.surfaceRoughness, X, basis.nW);
    float E = fresnelDielectricReflectance(abs(winputL.z), ior);
    float specWeight = (E)*averageComponent(specAlbedo);
    float diffuseWeight = (1.0f - E) * averageComponent(diffuseAlbedo);
    float weightSum = max(specWeight + diffuseWeight, DENOM_TOLERANCE);
    float specProb = specWeight / weightSum;
    if (snl_rand(rnd) >= specProb) {
      woutputL = sampleHemisphereCosineWeighted(rnd, pdfOut);
      pdfOut *= (1.0f - specProb);
      return (1.0f - E) * diffuseAlbedo / float(M_PI);
    } else {
      vec3 m = microfacetSample(rnd, roughness);
      woutputL = -winputL + 2.0f * dot(winputL, m) * m;
      pdfOut = specProb;
      if (woutputL.z < DENOM_TOLERANCE) return rt1(0.0f);
      float D = microfacetEval(m, roughness);
      float G = smithG2(winputL, woutputL, m, roughness);
      float Fr = fresnelDielectricReflectance(abs(winputL.z), ior);
      RT f = Fr * specAlbedo * D * G / max(4.0f * abs(cosTheta(woutputL)) * abs(cosTheta(winputL)), DENOM_TOLERANCE);
      float dwh_dwo = 1.0f / max(abs(4.0f * dot(winputL, m)), DENOM_TOLERANCE);
      pdfOut *= microfacetPDF(m, roughness) * dwh_dwo;
      return f;
    }
  }

  // ---------------- BSDF dispatch (813-852)
  SNL_HD RT evaluateBsdf(const vec3& X, const Basis& basis, const vec3& wi, const vec3& wo, int material,
                         float wl, const vec3& rgb, bool fromCamera) {
    if constexpr (SC::kHasSurface) if (material == MAT_SURFA) return evaluateSurface(X, basis, wi, wo, wl, rgb);
    if constexpr (SC::kHasMetal) if (material == MAT_METAL) return evaluateMetal(X, basis, wi, wo, wl, rgb);
    if constexpr (SC::kHasDielectric) if (material == MAT_DIELE) return evaluateDielectric(X, basis, wi, wo, wl, rgb, fromCamera);
    return rt1(0.0f);
  }
  SNL_HD RT sampleBsdf(const vec3& X, const Basis& basis, const vec3& wi, int material, float wl,
                       const vec3& rgb, bool fromCamera, vec3& wo, float& pdfOut, vec4& rnd) {
    if constexpr (SC::kHasSurface) if (material == MAT_SURFA) return sampleSurface(X, basis, wi, wl, rgb, wo, pdfOut, rnd);
    if constexpr (SC::kHasMetal) if (material == MAT_METAL) return sampleMetal(X, basis, wi, wl, rgb, wo, pdfOut, rnd);
    if constexpr (SC::kHasDielectric) if (material == MAT_DIELE) return sampleDielectric(X, basis, wi, wl, rgb, fromCamera, wo, pdfOut, rnd);
    pdfOut = 1.0f;
    return rt1(0.0f);
  }
  SNL_HD float pdfBsdf(const vec3& X, const Basis& basis, const vec3& wi, const vec3& wo, int material,
                       float wl, const vec3& rgb, bool fromCamera) {
    if constexpr (SC::kHasSurface) if (material == MAT_SURFA) return pdfSurface(X, basis, wi, wo, wl, rgb);
    if constexpr (SC::kHasMetal) if (material == MAT_METAL) return pdfMetal(X, basis, wi, wo, wl, rgb);
    if constexpr (SC::kHasDielectric) if (material == MAT_DIELE) return pdfDielectric(X, basis, wi, wo, wl, rgb, fromCamera);
    return PDF_EPSILON;
  }

  // ---------------- atmosphere (971-1039)
  SNL_HD RT VOLUME_SCATTERING_EVAL(const vec3& rgb) {
    return rgbToAlbedo(P.atmosphereScatteringCoeffRGB / (P.atmosphereMFP * P.lengthScale), rgb);
  }
  SNL_HD RT VOLUME_EXTINCTION_EVAL(const vec3& rgb) {
    return rgbToAlbedo(P.atmosphereExtinctionCoeffRGB / (P.atmosphereMFP * P.lengthScale), rgb);
  }
  SNL_HD RT VOLUME_EMISSION_EVAL(const vec3& X, const vec3& rgb) {
    vec3 emission = S.VOLUME_EMISSION(P.volumeEmissionColorRGB * P.volumeEmission, X);
    return rgbToAlbedo(emission, rgb);
  }
  SNL_HD bool atmosphereHit(const vec3& rayPos, const vec3& rayDir, float& t0, float& t1) {
    vec3 dL = 1.0f / rayDir;
    vec3 lo = (P.atmosphereBoundsMin * P.lengthScale - rayPos) * dL;
    vec3 hi = (P.atmosphereBoundsMax * P.lengthScale - rayPos) * dL;
    vec3 tmp = min(lo, hi); hi = lo + hi - tmp; lo = tmp;
    bool hit = !(lo.x > hi.y || lo.y > hi.x || lo.x > hi.z || lo.z > hi.x || lo.y > hi.z || lo.z > hi.y);
    t0 = max(max(lo.x, lo.y), lo.z);
    t1 = min(min(hi.x, hi.y), hi.z);
    return hit;
  }
  SNL_HD bool atmosphereSegment(const vec3& pW, const vec3& rayDir, float segmentLength, float& t0, float& t1) {
    if (!atmosphereHit(pW, rayDir, t0, t1)) return false;
    t0 = max(0.0f, t0);
    if (t1 < 0.0f || segmentLength < t0) return false;
    t1 = min(segmentLength, t1);
    return true;
  }
  SNL_HD RT transmittanceOverFreeSegment(const vec3& pW, const vec3& rayDir, float segmentLength, const vec3& rgb) {
    RT Tr = rt1(1.0f);
    if constexpr (SC::kHasAtmosphere) {
      RT extinction = VOLUME_EXTINCTION_EVAL(rgb);
      if (length(extinction) == 0.0f) return rt1(1.0f);
      float t0 = 0.0f, t1 = 0.0f;
      if (!atmosphereSegment(pW, rayDir, segmentLength, t0, t1)) return rt1(1.0f);
      RT opticalDepth = (t1 - t0) * extinction;
      Tr = exp(-opticalDepth);
    }
    if constexpr (SC::kHasFog) {
      if (P.fogEnable) {
        vec3 fogAbsorption = vec3(1.0f) - P.fogTint;
        vec3 opticalDepthFog = P.fogMaxOpticalDepth * (vec3(1.0f) - exp(-segmentLength * fogAbsorption / P.fogMFP));
        Tr *= exp(-rgbToAlbedo(opticalDepthFog, rgb));
      }
    }
    return Tr;
  }
  SNL_HD float atmosphericScatteringProbability(const RT& Tr, const vec3& rgb) {
    if constexpr (SC::kHasAtmosphere) {
      RT extinction = VOLUME_EXTINCTION_EVAL(rgb);
      RT scattering = VOLUME_SCATTERING_EVAL(rgb);
      float albedo_norm = averageComponent(scattering / extinction);
      return albedo_norm * (1.0f - averageComponent(Tr));
    }
    return 0.0f;
  }

  // ---------------- heterogeneous medium (HAS_VOLUME)
  // The reference's scene API declares a medium's hooks -- SDF_VOLUME, and
  // VOLUME_EXTINCTION(_MAX), VOLUME_SCATTERING/ABSORPTION_COLOR and
  // VOLUME_ANISOTROPY with their injected defaults (js/sceneAPI.js:124-131,
  // js/renderer.js:258-263) -- but its path tracer never calls them.  Here a
  // scene that defines SDF_VOLUME gets that medium, in place of the
  // homogeneous atmosphere:
  //   * it lies where SDF_VOLUME(X) <= 0, and SDF_VOLUME(X) = d > 0 means no
  //     medium within distance d of X;
  //   * grey extinction sigma(X) = VOLUME_EXTINCTION(volumeExtinction, X) /
  //     lengthScale, zero outside, bounded by the majorant sigmaMax =
  //     VOLUME_EXTINCTION_MAX(volumeExtinction) / lengthScale;
  //   * single-scattering albedo s / (s + a) per channel, with s and a the
  //     VOLUME_SCATTERING_COLOR / VOLUME_ABSORPTION_COLOR of the Volume
  //     material's scatteringColor / absorptionColor;
  //   * Henyey-Greenstein phase function of g = VOLUME_ANISOTROPY(anisotropy, X).
  // Free flights are sampled by delta tracking and shadow rays weighted by
  // ratio tracking, both against sigmaMax.  Where SDF_VOLUME exceeds
  // kVolumeSkip / sigmaMax the walk jumps that distance: nothing there can
  // collide, and the exponential's memorylessness lets tracking restart after
  // the jump.  At most kVolumeSteps steps per segment.
  static constexpr int kVolumeSteps = 1024;
  static constexpr float kVolumeSkip = 0.0625f;
  SNL_HD float volumeSigmaMax() {
    if constexpr (SC::kHasVolume) return S.VOLUME_EXTINCTION_MAX(P.volumeExtinction) / P.lengthScale;
    return 0.0f;
  }
  SNL_HD float volumeSigma(const vec3& X) {
    if constexpr (SC::kHasVolume) return S.VOLUME_EXTINCTION(P.volumeExtinction, X) / P.lengthScale;
    return 0.0f;
  }
  SNL_HD float volumeBound(const vec3& X) {
    if constexpr (SC::kHasVolume) return S.SDF_VOLUME(X);
    return 1.0e20f;
  }
  SNL_HD RT volumeAlbedo(const vec3& X, const vec3& rgb) {
    if constexpr (SC::kHasVolume) {
      vec3 sc = S.VOLUME_SCATTERING_COLOR(P.atmosphereScatteringCoeffRGB, X);
      vec3 ab = S.VOLUME_ABSORPTION_COLOR(P.atmosphereExtinctionCoeffRGB - P.atmosphereScatteringCoeffRGB, X);
      return rgbToAlbedo(sc / max(sc + ab, vec3(DENOM_TOLERANCE)), rgb);
    }
    return rt1(0.0f);
  }
  SNL_HD float volumeAnisotropy(const vec3& X) {
    if constexpr (SC::kHasVolume) return S.VOLUME_ANISOTROPY(P.atmosphereAnisotropy, X);
    return 0.0f;
  }
  // delta tracking over [0, tmax) of the ray (o, d): the distance of the first
  // real collision, or tmax if none
  SNL_HD float volumeFreeFlight(const vec3& o, const vec3& d, float tmax, vec4& rnd) {
    float sigmaMax = volumeSigmaMax();
    if (!(sigmaMax > 0.0f)) return tmax;
    float skip = kVolumeSkip / sigmaMax;
    float t = 0.0f;
    for (int n = 0; n < kVolumeSteps; ++n) {
      float dv = volumeBound(o + t * d);
      if (dv > skip) {
        t += dv;
        if (t >= tmax) return tmax;
        continue;
      }
      t -= log(1.0f - snl_rand(rnd)) / sigmaMax;
      if (t >= tmax) return tmax;
      if (snl_rand(rnd) * sigmaMax < volumeSigma(o + t * d)) return t;
    }
    return tmax;
  }
  // ratio tracking over [0, tmax): an unbiased estimate of the medium's
  // transmittance along the segment
  SNL_HD float volumeTransmittance(const vec3& o, const vec3& d, float tmax, vec4& rnd) {
    float sigmaMax = volumeSigmaMax();
    if (!(sigmaMax > 0.0f)) return 1.0f;
    float skip = kVolumeSkip / sigmaMax;
    float T = 1.0f, t = 0.0f;
    for (int n = 0; n < kVolumeSteps; ++n) {
      float dv = volumeBound(o + t * d);
      if (dv > skip) {
        t += dv;
        if (t >= tmax) break;
        continue;
      }
      t -= log(1.0f - snl_rand(rnd)) / sigmaMax;
      if (t >= tmax) break;
      T *= 1.0f - volumeSigma(o + t * d) / sigmaMax;
    }
    return T;
  }

  // ---------------- lights (1051-1212)
  SNL_HD vec3 environmentRadianceRGB(const vec3& dirW) {
    float rot_phi = float(M_PI) * P.envMapPhiRotation / 180.0f;
    float rot_theta = float(M_PI) * P.envMapThetaRotation / 180.0f;
    vec3 sky_pole(sin(rot_theta), cos(rot_theta), 0.0f);
    vec3 RGB(1.0f);
    if (P.haveEnvMap) {
      Basis sky_basis = makeBasis(sky_pole);
      vec3 dirL = worldToLocal(dirW, sky_basis);
      float phi = atan(dirL.y, dirL.x) + float(M_PI) + rot_phi;
      phi -= 2.0f * float(M_PI) * floor(phi / (2.0f * float(M_PI)));
      float theta = acos(dirL.z);
      float u = phi / (2.0f * float(M_PI));
      float v = theta / float(M_PI);
      RGB = vec3(texture(P.envMap, vec2(u, v)));
    }
    float t = dot(dirW, sky_pole);
    float tt = P.envMapTransitionAngle / 180.0f;
    RGB *= P.skyPower * mix(P.skyTintDown, P.skyTintUp, smoothstep(-tt, tt, t));
    return RGB;
  }
  SNL_HD RT environmentRadiance(const vec3& dir, const vec3& rgb) { return rgbToAlbedo(environmentRadianceRGB(dir), rgb); }
  SNL_HD RT sampleSkyAtSurface(const Basis& basis, const vec3& rgb, vec4& rnd, vec3& woutputL, vec3& woutputW, float& pdfDir) {
    if (P.skyPower < RADIANCE_EPSILON) return rt1(0.0f);
    woutputL = sampleHemisphereCosineWeighted(rnd, pdfDir);
    woutputW = localToWorld(woutputL, basis);
    return environmentRadiance(woutputW, rgb);
  }
  SNL_HD RT sampleSkyInVolume(const vec3& rgb, vec4& rnd, vec3& woutputW, float& pdfDir) {
    if (P.skyPower < RADIANCE_EPSILON) return rt1(0.0f);
    woutputW = sampleSphereUniformly(rnd);
    pdfDir = 1.0f / (4.0f * float(M_PI));
    return environmentRadiance(woutputW, rgb);
  }
  SNL_HD vec3 sampleSunDir(vec4& rnd, float& pdfDir) {
    float theta_max = P.sunAngularSize * float(M_PI) / 180.0f;
    float theta = theta_max * sqrt(snl_rand(rnd));
    float costheta = cos(theta);
    float sintheta = sqrt(max(0.0f, 1.0f - costheta * costheta));
    float phi = 2.0f * float(M_PI) * snl_rand(rnd);
    float cosphi = cos(phi);
    float sinphi = sin(phi);
    float solid_angle = 2.0f * float(M_PI) * (1.0f - cos(theta_max));
    pdfDir = 1.0f / solid_angle;
    return localToWorld(vec3(sintheta * cosphi, sintheta * sinphi, costheta), makeBasis(P.sunDir));
  }
  SNL_HD RT sunRadiance(const vec3& dir, const vec3& rgb) {
    float theta_max = P.sunAngularSize * float(M_PI) / 180.0f;
    if (dot(dir, P.sunDir) < cos(theta_max)) return rt1(0.0f);
    return rgbToAlbedo(P.sunPower * P.sunColor, rgb);
  }
  SNL_HD RT sampleSunAtSurface(const Basis& basis, const vec3& rgb, vec4& rnd, vec3& woutputL, vec3& woutputW, float& pdfDir) {
    if (P.sunPower < RADIANCE_EPSILON) return rt1(0.0f);
    woutputW = sampleSunDir(rnd, pdfDir);
    woutputL = worldToLocal(woutputW, basis);
    if (woutputL.z < 0.0f) return rt1(0.0f);
    return sunRadiance(woutputW, rgb);
  }
  SNL_HD RT sampleSunInVolume(const vec3& rgb, vec4& rnd, vec3& woutputW, float& pdfDir) {
    if (P.sunPower < RADIANCE_EPSILON) return rt1(0.0f);
    woutputW = sampleSunDir(rnd, pdfDir);
    return sunRadiance(woutputW, rgb);
  }
  SNL_HD vec3 sampleSphereLightPosition(vec4& rnd, const vec3& pW, float& pdfArea) {
    vec3 cp = normalize(pW - P.sphereLightPosition);
    Basis hb = makeBasis(cp);
    vec3 vL = sampleHemisphereUniformly(rnd);
    vec3 vW = localToWorld(vL, hb);
    pdfArea = 1.0f / (2.0f * float(M_PI) * P.sphereLightRadius * P.sphereLightRadius);
    return P.sphereLightPosition + P.sphereLightRadius * vW;
  }
  SNL_HD RT sphereLightRadiance(const vec3& pW, const vec3& dir, const vec3& rgb, vec3& pHit) {
    vec3 m = pW - P.sphereLightPosition;
    float b = dot(m, dir);
    float c = dot(m, m) - P.sphereLightRadius * P.sphereLightRadius;
    if (c > 0.0f && b > 0.0f) return rt1(0.0f);
    float discr = b * b - c;
    if (discr < 0.0f) return rt1(0.0f);
    float t = -b - sqrt(discr);
    t = max(t, 0.0f);
    pHit = pW + t * dir;
    return rgbToAlbedo(P.sphereLightPower * P.sphereLightColor, rgb);
  }
  SNL_HD RT sampleSphereLight(const vec3& rgb, vec4& rnd, const vec3& pW, vec3& pHit, float& pdfDir) {
    if (P.sphereLightPower < RADIANCE_EPSILON) return rt1(0.0f);
    float pdfArea;
    pHit = sampleSphereLightPosition(rnd, pW, pdfArea);
    vec3 dHit = pHit - pW;
    float dHit2 = dot(dHit, dHit);
    dHit = normalize(dHit);
    vec3 nHit = normalize(pHit - P.sphereLightPosition);
    float costheta = abs(dot(nHit, dHit));
    pdfDir = pdfArea * dHit2 / max(DENOM_TOLERANCE, costheta);
    return rgbToAlbedo(P.sphereLightPower * P.sphereLightColor, rgb);
  }

  // ---------------- camera (1602-1622)
  SNL_HD void constructPrimaryRay(const vec2& pixel, vec4& rnd, vec3& primaryStart, vec3& primaryDir, bool safe) {
    vec2 ndc = -1.0f + 2.0f * (pixel / P.resolution);
    float fh = P.camFrustumHeight;  // tan(0.5 * radians(camFovy)), formed once per upload
    float fw = P.camAspect * fh;
    vec3 s = -fw * ndc.x * P.camX + fh * ndc.y * P.camY;
    primaryDir = safe ? safe_normalize(P.camDir + s) : normalize(P.camDir + s);
    if (P.camAperture <= 0.0f) { primaryStart = P.camPos; return; }
    vec3 focalPlaneHit = P.camPos + P.camFocalDistance * primaryDir / dot(primaryDir, P.camDir);
    float lensRadial = P.camAperture * sqrt(snl_rand(rnd));
    float theta = 2.0f * float(M_PI) * snl_rand(rnd);
    vec3 lensPos = P.camPos + lensRadial * (-P.camX * cos(theta) + P.camY * sin(theta));
    primaryStart = lensPos;
    primaryDir = safe ? safe_normalize(focalPlaneHit - lensPos) : normalize(focalPlaneHit - lensPos);
  }
};

SNL_HD RT ldRT(const SnlKArgs& a, int k, int i) {
#ifdef DISPERSION_ENABLED
  return SNL_PF(k, i);
#else
  return vec3(SNL_PF(k, i), SNL_PF(k + 1, i), SNL_PF(k + 2, i));
#endif
}
SNL_HD void stRT(const SnlKArgs& a, int k, int i, const RT& v) {
#ifdef DISPERSION_ENABLED
  SNL_PF(k, i) = v;
#else
  SNL_PF(k, i) = v.x; SNL_PF(k + 1, i) = v.y; SNL_PF(k + 2, i) = v.z;
#endif
}
SNL_HD vec3 ld3(const SnlKArgs& a, int k, int i) { return vec3(SNL_PF(k, i), SNL_PF(k + 1, i), SNL_PF(k + 2, i)); }
SNL_HD void st3(const SnlKArgs& a, int k, int i, const vec3& v) { SNL_PF(k, i) = v.x; SNL_PF(k + 1, i) = v.y; SNL_PF(k + 2, i) = v.z; }

#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline int snl_lane() { return __lane_id(); }
// Workgroup-staged queue append: the entries of the (at most 256-lane)
// workgroup are gathered in LDS in (wave, lane) order and written to `queue`
// as one contiguous run after a single atomic on `counter`: coalesced stores
// and one atomic per workgroup.  Every lane of the workgroup must call it (it
// synchronises the workgroup); queue order within a run is the lanes' order.
__device__ inline void snl_append_staged(int* counter, int* queue, bool pred, int value) {
  __shared__ int stage[256];
  __shared__ int count[5];  // per wave, then the run's base
  const int lane = snl_lane(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long m = __ballot(pred);
  if (lane == 0) count[wave] = __popcll(m);
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    int c = count[w];
    off += w < wave ? c : 0;
    tot += c;
  }
  if (pred) stage[off + __popcll(m & ((1ull << lane) - 1ull))] = value;
  if (threadIdx.x == 0) count[4] = tot ? atomicAdd(counter, tot) : 0;
  __syncthreads();
  const int base = count[4];
  for (int t = threadIdx.x; t < tot; t += blockDim.x) queue[base + t] = stage[t];
  __syncthreads();  // the buffers serve the next call
}
#endif

}  // namespace glsl
