// snl_modes.h -- the other integrators of Renderer.renderMode (js/renderer.js:
// 334-352, 508-515): ambient occlusion ('ao', shaders/ao-fragment-shader.glsl)
// and the simplified path tracer ('ptsimple',
// shaders/simplepathtracer-fragment-shader.glsl).  A program built with
// SNL_BUILD_AO or SNL_BUILD_PTSIMPLE compiles one of them as snl_mode_kernel.
// Both are RGB integrators (no wavelength sampling): such programs are
// compiled without DISPERSION_ENABLED, so RT is vec3.
//
// Launch shape: a wave takes 64 queue items (pixels) at a time and each lane
// runs all frames of its pixel, the RNG state and running mean in registers.
#pragma once

namespace glsl {

#if defined(SNL_MODE_AO) || defined(SNL_MODE_PTSIMPLE)

// the running mean of main (ao 476-487, simplepathtracer 1108-1120)
SNL_HD void snl_accumulate(vec4& accum, const vec3& XYZ) {
  float oldN = accum.w, newN = oldN + 1.0f;
  accum = vec4((oldN * vec3(accum) + XYZ) / newN, newN);
}

template <class SC, class PixelFrames>
__device__ inline void mode_queue(const SnlKArgs& a, PixelFrames&& frames) {
  const int lane = snl_lane();
  const int nItems = a.itemEnd - a.itemBase;
  while (true) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&a.cnt[C_FETCH], 64);
    base = __shfl(base, 0);
    if (base >= nItems) break;
    int item = base + lane;
    if (item >= nItems) continue;
    int pix = snl_item_pixel(a, snl_tile_item(a, item));
    vec4 rnd = ld4(a.rngstate, pix);
    vec4 accum = ld4(a.radiance, pix);
    vec2 pixel(float(pix % a.W) + 0.5f, float((pix % a.imagePix) / a.W) + 0.5f);
    frames(pixel, rnd, accum);
    st4(a.radiance, pix, accum);
    st4(a.rngstate, pix, rnd);
  }
}

#endif

#ifdef SNL_MODE_AO
// ---------------------------------------------------------------- ambient occlusion
template <class SC>
struct AO {
  SC& S;
  PT<SC>& pt;
  const SnlParams& P;
  static constexpr float kRadianceEps = 1.0e-5f;  // the AO program's RADIANCE_EPSILON (62)
  SNL_HD AO(SC& s, PT<SC>& t, const SnlParams& p) : S(s), pt(t), P(p) {}

  // Transmittance (154-159): whether the ray is blocked within maxLengthScale
  SNL_HD float transmittance(const vec3& start, const vec3& dir) {
    vec3 hit;
    int m;
    return pt.traceDistance(start, dir, P.maxLengthScale, hit, m) ? 0.0f : 1.0f;
  }
  // sampleSunDir (327-342) in the frame of sunBasis = makeBasis(sunDir) (409)
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
    return localToWorld(vec3(sintheta * cosphi, sintheta * sinphi, costheta), makeBasisPlain(P.sunDir));
  }

  // one sample of main's loop (411-473)
  SNL_HD vec3 sample(const vec2& pixel, vec4& rnd) {
    float r1 = snl_rand(rnd);
    float r2 = snl_rand(rnd);
    vec2 pixelj = pixel + (-0.5f + vec2(r1, r2));
    vec3 primaryStart, primaryDir;
    if constexpr (SC::kHasCustomCamera) S.CONSTRUCT_PRIMARY_RAY(pixelj, rnd, primaryStart, primaryDir);
    else pt.constructPrimaryRay(pixelj, rnd, primaryStart, primaryDir, false);
    vec3 pW;
    vec3 woW = -primaryDir;
    int hitMaterial;
    bool hit = pt.traceDistance(primaryStart, primaryDir, P.maxLengthScale, pW, hitMaterial);
    vec3 RGB(0.0f);
    if (hit) {
      // a max-step "hit" has no material, and the program's normal() no value
      // for it: the sample adds nothing
      if (hitMaterial == MAT_INVAL) return RGB;
      vec3 nW = pt.normal(pW, hitMaterial, false);
      Basis basis = makeBasisPlain(nW);
      if constexpr (SC::kHasSurfaceNM) {
        if (hitMaterial == MAT_SURFA) nW = localToWorld(normalize(vec3(S.SURFACE_NORMAL_MAP(pW, basis.nW))), basis);
        else nW = basis.nW;
        basis = makeBasisPlain(nW);
      }
      // every material is shaded with the Surface diffuse colour (446)
      vec3 f = S.SURFACE_DIFFUSE_REFLECTANCE(P.surfaceDiffuseAlbedoRGB, pW, nW, woW) / float(M_PI);
      vec3 Ldirect(0.0f);
      vec3 dPw = 3.0f * P.minLengthScale * nW;
      if (P.skyPower > kRadianceEps) {
        // sampleSkyAtSurface (313-319)
        float skyPdf = 0.0f;
        vec3 woutputL = sampleHemisphereCosineWeighted(rnd, skyPdf);
        vec3 woutputW = localToWorld(woutputL, basis);
        vec3 Li = pt.environmentRadianceRGB(woutputW);
        Li *= transmittance(pW + dPw, woutputW);
        Ldirect += f * Li / max(PDF_EPSILON, skyPdf) * abs(dot(woutputW, nW));
      }
      if (P.sunPower > kRadianceEps) {
        // sampleSunAtSurface (360-368)
        float sunPdf = 0.0f;
        vec3 woutputW = sampleSunDir(rnd, sunPdf);
        vec3 woutputL = worldToLocal(woutputW, basis);
        vec3 Li = woutputL.z < 0.0f ? vec3(0.0f) : vec3(pt.sunRadiance(woutputW, vec3(0.0f)));
        if (averageComponent(Li) > kRadianceEps) {
          Li *= transmittance(pW + dPw, woutputW);
          Ldirect += f * Li / max(PDF_EPSILON, sunPdf) * abs(dot(woutputW, nW));
        }
      }
      RGB += Ldirect;
    } else if (P.envMapVisible) {
      RGB += pt.environmentRadianceRGB(primaryDir);
    }
    return RGB;
  }
};

template <class SC>
__device__ inline void ao_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  AO<SC> ao(S, pt, P);
  mode_queue<SC>(a, [&](const vec2& pixel, vec4& rnd, vec4& accum) {
    for (int frame = 0; frame < a.frames; ++frame) {
#ifdef INTERACTIVE_MODE
      if (snl_rand(rnd) < P.skipProbability) continue;  // main (398-408)
#endif
      vec3 RGB(0.0f);
      for (int n = 0; n < a.spf; ++n) RGB += ao.sample(pixel, rnd);
      RGB /= float(a.spf);
      snl_accumulate(accum, rgbToXyz(RGB));
    }
  });
}

extern "C" __global__ void __launch_bounds__(256) snl_mode_kernel(SnlKArgs a) { ao_body<SnlScene>(a); }
#endif  // SNL_MODE_AO

#ifdef SNL_MODE_PTSIMPLE
// ---------------------------------------------------------------- simplified path tracer
// The simplified program knows only the Surface material, fog (no
// atmosphere, no sphere light) and applies no shadowStrength.
template <class SC>
struct Simple {
  SC& S;
  PT<SC>& pt;
  const SnlParams& P;
  SNL_HD Simple(SC& s, PT<SC>& t, const SnlParams& p) : S(s), pt(t), P(p) {}

  // traceRay (508-534): SDF_SURFACE only
  SNL_HD bool traceRay(const vec3& start, const vec3& dir, vec3& hit, int& material, float maxMarchDist) {
    material = MAT_INVAL;
    if constexpr (!SC::kHasSurface) {
      return false;
    } else {
      float minMarch = P.minLengthScale;
      float sdf = abs(S.SDF_SURFACE(start));
      float InitialSign = sign(sdf);
      float t = InitialSign * sdf;
      if (abs(t) >= maxMarchDist) return false;
      for (int n = 0; n < SNL_MAX_MARCH_STEPS; n++) {
        vec3 pW = start + t * dir;
        float s = abs(S.SDF_SURFACE(pW));
        if (s < minMarch) {
          material = MAT_SURFA;
          hit = start + t * dir;
          return true;
        }
        sdf = min(1.0e20f, s);
        t += InitialSign * sdf;
        if (abs(t) >= maxMarchDist) return false;
      }
      return !P.maxStepsIsMiss;
    }
  }
  // transmittanceOverFreeSegment (570-582): fog only
  SNL_HD vec3 transmittanceOverFreeSegment(float segmentLength) {
    vec3 Tr(1.0f);
    if constexpr (SC::kHasFog) {
      if (P.fogEnable) {
        vec3 fogAbsorption = vec3(1.0f) - P.fogTint;
        vec3 opticalDepthFog = P.fogMaxOpticalDepth * (vec3(1.0f) - exp(-segmentLength * fogAbsorption / P.fogMFP));
        Tr *= exp(-opticalDepthFog);
      }
    }
    return Tr;
  }
  // transmittanceOverSegment (584-597)
  SNL_HD vec3 transmittanceOverSegment(const vec3& pW, const vec3& rayDir, float segmentLength) {
    vec3 pW_surface;
    int hitMaterial;
    if (traceRay(pW, rayDir, pW_surface, hitMaterial, segmentLength)) return vec3(0.0f);
    if constexpr (SC::kHasFog) return transmittanceOverFreeSegment(segmentLength);
    return vec3(1.0f);
  }
  // atmosphericInscatteringRadiance (599-604)
  SNL_HD vec3 atmosphericInscatteringRadiance(const vec3& Tr) {
    if (!P.fogEnable) return vec3(0.0f);
    return (1.0f - Tr) * P.fogEmission;
  }

  // directSurfaceLighting (719-770): sky and sun.  A BSDF value of exactly
  // zero makes a light's term exactly zero, so its shadow ray is not traced.
  SNL_HD vec3 directSurfaceLighting(const vec3& pW, const Basis& basis, const vec3& winputW, vec4& rnd,
                                    float& skyPdf, float& sunPdf) {
    vec3 winputL = worldToLocal(winputW, basis);
    vec3 Ldirect(0.0f);
    vec3 woutputL(0.0f), woutputW(0.0f);
    const float wl = 390.0f + (750.0f - 390.0f) * 0.5f;
    const vec3 rgb(0.0f);
    if (P.skyPower > RADIANCE_EPSILON) {
      vec3 Li = pt.sampleSkyAtSurface(basis, rgb, rnd, woutputL, woutputW, skyPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        vec3 f = pt.evaluateBsdf(pW, basis, winputL, woutputL, MAT_SURFA, wl, rgb, true);
        if (anyNonzero(f)) {
          Li *= transmittanceOverSegment(pW, woutputW, P.maxLengthScale);
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            float bsdfPdf = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, woutputL, MAT_SURFA, wl, rgb, true));
            float misWeight = powerHeuristic(skyPdf, bsdfPdf);
            Ldirect += f * Li / max(PDF_EPSILON, skyPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
          }
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      vec3 Li = pt.sampleSunAtSurface(basis, rgb, rnd, woutputL, woutputW, sunPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        vec3 f = pt.evaluateBsdf(pW, basis, winputL, woutputL, MAT_SURFA, wl, rgb, true);
        if (anyNonzero(f)) {
          Li *= transmittanceOverSegment(pW, woutputW, P.maxLengthScale);
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            float bsdfPdf = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, woutputL, MAT_SURFA, wl, rgb, true));
            float misWeight = powerHeuristic(sunPdf, bsdfPdf);
            Ldirect += f * Li / max(PDF_EPSILON, sunPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
          }
        }
      }
    }
    return min(vec3(P.radianceClamp), Ldirect);
  }

  // randomwalk_SSS (783-834), at most 256 steps
  SNL_HD bool randomwalk_SSS(const vec3& pW, const Basis& basis, vec4& rnd, const vec3& subsurfaceAlbedo,
                             vec3& walk_throughput, vec3& pExit) {
    vec3 pWalk = pW - 3.0f * P.minLengthScale * basis.nW;
    float pdfDir;
    vec3 dirwalkL = sampleHemisphereCosineWeighted(rnd, pdfDir);
    vec3 dirwalkW = -localToWorld(dirwalkL, basis);
    walk_throughput = vec3(1.0f);
    for (int n = 0; n < 256; ++n) {
      if (n > P.maxSSSSteps) break;
      float walk_step = -log(snl_rand(rnd)) * P.subsurfaceMFP * P.lengthScale;
      vec3 pHit;
      int hm;
      if (traceRay(pWalk, dirwalkW, pHit, hm, walk_step)) {
        pExit = pHit;
        return true;
      }
      if (n > 4) {
        float continuation_prob = clamp(10.0f * maxComponent(walk_throughput), 0.0f, 1.0f);
        if (snl_rand(rnd) < 1.0f - continuation_prob) break;
        walk_throughput /= continuation_prob;
      }
      dirwalkW = samplePhaseFunction(dirwalkW, P.subsurfaceAnisotropy, rnd);
      pWalk += walk_step * dirwalkW;
      walk_throughput *= subsurfaceAlbedo;
    }
    return false;
  }
  // SSS_exit_radiance (837-882): the shadow rays start at the exit point itself
  SNL_HD vec3 SSS_exit_radiance(const vec3& pW, const Basis& basis, vec4& rnd, const vec3& diffuseAlbedoExit,
                                float& skyPdf, float& sunPdf) {
    vec3 Ldirect(0.0f);
    vec3 woutputL(0.0f), woutputW(0.0f);
    const vec3 rgb(0.0f);
    vec3 f = mix(vec3(1.0f), diffuseAlbedoExit, P.subsurfaceDiffuseWeight) / float(M_PI);
    if (P.skyPower > RADIANCE_EPSILON) {
      vec3 Li = pt.sampleSkyAtSurface(basis, rgb, rnd, woutputL, woutputW, skyPdf);
      Li *= transmittanceOverSegment(pW, woutputW, P.maxLengthScale);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        float misWeight = powerHeuristic(skyPdf, pdfHemisphereCosineWeighted(woutputL));
        Ldirect += f * Li / max(PDF_EPSILON, skyPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      vec3 Li = pt.sampleSunAtSurface(basis, rgb, rnd, woutputL, woutputW, sunPdf);
      Li *= transmittanceOverSegment(pW, woutputW, P.maxLengthScale);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        float misWeight = powerHeuristic(sunPdf, pdfHemisphereCosineWeighted(woutputL));
        Ldirect += f * Li / max(PDF_EPSILON, sunPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
      }
    }
    return min(vec3(P.radianceClamp), Ldirect);
  }

  // cameraPath (913-1083)
  SNL_HD vec3 cameraPath(const vec3& primaryStart, const vec3& primaryDir, vec4& rnd, int maxBounces) {
    const float wl = 390.0f + (750.0f - 390.0f) * 0.5f;
    const vec3 rgb(0.0f);
    vec3 L(0.0f);
    float misWeightSky = 1.0f, misWeightSun = 1.0f;
    vec3 pW = primaryStart, rayDir = primaryDir;
    vec3 throughput(1.0f);
    for (int vertex = 0; vertex <= maxBounces; ++vertex) {
      if (maxComponent(throughput) < THROUGHPUT_EPSILON) break;
      vec3 pW_next;
      int hitMaterial;
      bool hit = traceRay(pW, rayDir, pW_next, hitMaterial, P.maxLengthScale);
      float rayLength = P.maxLengthScale;
      vec3 Tr;
      if (hit) {
        // a max-step "hit" has no material: the path ends
        if (hitMaterial == MAT_INVAL) break;
        rayLength = length(pW_next - pW);
        Tr = transmittanceOverFreeSegment(rayLength);
      } else {
        Tr = transmittanceOverFreeSegment(P.maxLengthScale);
        bool lit = true;
        if constexpr (SC::kHasFog) lit = averageComponent(Tr) > RADIANCE_EPSILON;
        if (lit) {
          if (!(vertex == 0 && !P.envMapVisible) && misWeightSky > 0.0f)
            L += throughput * Tr * misWeightSky * vec3(pt.environmentRadiance(rayDir, rgb));
          if (!(vertex == 0 && !P.sunVisibleDirectly) && misWeightSun > 0.0f)
            L += throughput * Tr * misWeightSun * vec3(pt.sunRadiance(rayDir, rgb));
        }
        if constexpr (SC::kHasFog) {
          if (vertex < 2) L += throughput * atmosphericInscatteringRadiance(Tr);
        }
        break;
      }
      if constexpr (SC::kHasFog) {
        if (vertex < 2) L += throughput * atmosphericInscatteringRadiance(Tr);
        throughput *= Tr;
      }
      pW = pW_next;
      vec3 nW = pt.normal(pW, MAT_SURFA, true);
      vec3 ngW = nW;
      Basis basis = makeBasis(nW);
      if constexpr (SC::kHasNormalMap) {
        nW = basis.nW;
        if constexpr (SC::kHasSurfaceNM) nW = localToWorld(normalize(vec3(S.SURFACE_NORMAL_MAP(pW, basis.nW))), basis);
        if (dot(nW, rayDir) > 0.0f) nW = 2.0f * ngW * dot(ngW, nW) - nW;
        basis = makeBasis(nW);
      }
      bool do_subsurface_walk = false;
      float prob_sss = 0.0f;
      vec3 subsurfaceAlbedo(0.0f);
      if constexpr (SC::kHasSurface) {
        if (P.subsurfaceMFP > 0.0f) {
          subsurfaceAlbedo = vec3(pt.SUBSURFACE_ALBEDO_EVAL(pW, basis.nW, rgb));
          float diffuse_weight = (1.0f - P.subsurface) * averageComponent(pt.SURFACE_DIFFUSE_REFL_EVAL(pW, basis.nW, -rayDir, rgb));
          float spec_weight = averageComponent(pt.SURFACE_SPEC_REFL_EVAL(pW, basis.nW, -rayDir, rgb));
          float subsurface_weight = P.subsurface * averageComponent(subsurfaceAlbedo) * 3.0f;
          float total_weight = diffuse_weight + spec_weight + subsurface_weight;
          prob_sss = subsurface_weight / (total_weight + DENOM_TOLERANCE);
          prob_sss = clamp(prob_sss, 0.0f, 1.0f - PDF_EPSILON);
          do_subsurface_walk = (snl_rand(rnd) < prob_sss);
        }
      }
      if (!do_subsurface_walk) {
        vec3 winputW = -rayDir;
        vec3 winputL = worldToLocal(winputW, basis);
        vec3 woutputL(0.0f);  // WebGL zero-initialises these locals
        float bsdfPdf = 0.0f;
        vec3 f = pt.sampleBsdf(pW, basis, winputL, MAT_SURFA, wl, rgb, true, woutputL, bsdfPdf, rnd);
        vec3 woutputW = localToWorld(woutputL, basis);
        if constexpr (SC::kHasVolumeEmission) L += throughput * vec3(pt.VOLUME_EMISSION_EVAL(pW, rgb));
        rayDir = woutputW;
        pW += ngW * sign(dot(rayDir, ngW)) * 3.0f * P.minLengthScale;
        float skyPdf = 0.0f, sunPdf = 0.0f;
        L += throughput * directSurfaceLighting(pW, basis, winputW, rnd, skyPdf, sunPdf);
        vec3 fOverPdf = min(vec3(P.radianceClamp), f / max(PDF_EPSILON, bsdfPdf));
        throughput *= fOverPdf * abs(dot(woutputW, nW)) / max(PDF_EPSILON, 1.0f - prob_sss);
        misWeightSky = powerHeuristic(bsdfPdf, skyPdf);
        misWeightSun = powerHeuristic(bsdfPdf, sunPdf);
      } else {
        if constexpr (SC::kHasSurface) {
          vec3 walk_throughput(1.0f);
          vec3 pExit;
          basis.nW = ngW;
          if (!randomwalk_SSS(pW, basis, rnd, subsurfaceAlbedo, walk_throughput, pExit)) break;
          pW = pExit;
          nW = pt.normal(pExit, MAT_SURFA, true);
          Basis basis_exit = makeBasis(nW);
          float bsdfPdf;
          vec3 dirExitL = sampleHemisphereCosineWeighted(rnd, bsdfPdf);
          vec3 woutputW = localToWorld(dirExitL, basis_exit);
          rayDir = woutputW;
          throughput *= walk_throughput;
          float skyPdf = 0.0f, sunPdf = 0.0f;
          vec3 diffuseAlbedoExit = vec3(pt.SURFACE_DIFFUSE_REFL_EVAL(pW, nW, -rayDir, rgb));
          L += throughput * SSS_exit_radiance(pExit, basis_exit, rnd, diffuseAlbedoExit, skyPdf, sunPdf);
          misWeightSky = powerHeuristic(bsdfPdf, skyPdf);
          misWeightSun = powerHeuristic(bsdfPdf, sunPdf);
          vec3 f = mix(vec3(1.0f), diffuseAlbedoExit, P.subsurfaceDiffuseWeight) / float(M_PI);
          vec3 fOverPdf = min(vec3(P.radianceClamp), f / max(PDF_EPSILON, bsdfPdf));
          throughput *= fOverPdf * abs(dot(woutputW, nW)) / max(PDF_EPSILON, prob_sss);
          pW += nW * sign(dot(rayDir, nW)) * 3.0f * P.minLengthScale;
        }
      }
    }
    return L;
  }
};

template <class SC>
__device__ inline void ptsimple_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  Simple<SC> sp(S, pt, P);
  mode_queue<SC>(a, [&](const vec2& pixel, vec4& rnd, vec4& accum) {
    // pathtrace + main (1091-1129): no interactive skip in this program
    for (int frame = 0; frame < a.frames; ++frame) {
      vec3 L(0.0f);
      for (int n = 0; n < a.spf; ++n) {
        float jx = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
        float jy = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
        vec2 pixelj = pixel + vec2(jx, jy);
        vec3 primaryStart, primaryDir;
        if constexpr (SC::kHasCustomCamera) S.CONSTRUCT_PRIMARY_RAY(pixelj, rnd, primaryStart, primaryDir);
        else pt.constructPrimaryRay(pixelj, rnd, primaryStart, primaryDir, true);
        L += sp.cameraPath(primaryStart, primaryDir, rnd, a.maxBounces);
      }
      L /= float(a.spf);
      snl_accumulate(accum, rgbToXyz(L));
    }
  });
}

extern "C" __global__ void __launch_bounds__(256) snl_mode_kernel(SnlKArgs a) { ptsimple_body<SnlScene>(a); }
#endif  // SNL_MODE_PTSIMPLE

}  // namespace glsl
