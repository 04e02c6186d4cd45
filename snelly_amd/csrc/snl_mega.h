// snl_mega.h -- the path tracer as one loop per lane, shadow and probe rays
// traced where the reference traces them.
//
// This engine renders programs with the homogeneous atmosphere
// (HAS_ATMOSPHERE): there a vertex may trace up to five extra rays -- the
// in-scattering estimate's light samples and its emission probe -- between
// random draws, which the queue-based engines do not stage.  It follows
// cameraPath (pathtracer-fragment-shader.glsl:1625-1926) with
// directSurfaceLighting (1221-1295), the atmosphere routines (1297-1457) and
// the sub-surface walk (1468-1593).  Each lane runs all frames of a pixel of
// the launch's queue, one path vertex per loop iteration: a lane whose path
// ends starts the next at once, and one whose pixel is done takes another.
#pragma once

namespace glsl {

template <class SC>
struct Mega {
  PT<SC>& pt;
  const SnlParams& P;
  SNL_HD Mega(PT<SC>& t, const SnlParams& p) : pt(t), P(p) {}

  // transmittanceOverSegment (1042-1050); with the heterogeneous medium also
  // its ratio-tracking estimate over the segment (random draws)
  SNL_HD RT transmittanceOverSegment(const vec3& pW, const vec3& rayDir, float segmentLength, const vec3& rgb,
                                     vec4& rnd) {
    vec3 pHit;
    int hm;
    if (pt.traceRay(pW, rayDir, pHit, hm, segmentLength)) return rt1(0.0f);
    RT Tr = pt.transmittanceOverFreeSegment(pW, rayDir, segmentLength, rgb);
    if constexpr (SC::kHasVolume) Tr *= pt.volumeTransmittance(pW, rayDir, segmentLength, rnd);
    (void)rnd;
    return Tr;
  }

  // directSurfaceLighting (1221-1295).  A BSDF value of exactly zero makes a
  // light's term exactly zero whatever its shadow ray finds, so that ray is
  // not traced (the BSDF evaluation draws no random numbers).
  SNL_HD RT directSurfaceLighting(const vec3& pW, const Basis& basis, const vec3& winputW, int material, float wl,
                                  const vec3& rgb, vec4& rnd, float& skyPdf, float& sunPdf, float& sphPdf) {
    vec3 winputL = worldToLocal(winputW, basis);
    RT Ldirect = rt1(0.0f);
    vec3 woutputL(0.0f), woutputW(0.0f);
    if (P.skyPower > RADIANCE_EPSILON) {
      RT Li = pt.sampleSkyAtSurface(basis, rgb, rnd, woutputL, woutputW, skyPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT f = pt.evaluateBsdf(pW, basis, winputL, woutputL, material, wl, rgb, true);
        if (anyNonzero(f)) {
          RT Tr = transmittanceOverSegment(pW, woutputW, P.maxLengthScale, rgb, rnd);
          Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            float bsdfPdf = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, woutputL, material, wl, rgb, true));
            float misWeight = powerHeuristic(skyPdf, bsdfPdf);
            Ldirect += f * Li / max(PDF_EPSILON, skyPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
          }
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      RT Li = pt.sampleSunAtSurface(basis, rgb, rnd, woutputL, woutputW, sunPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT f = pt.evaluateBsdf(pW, basis, winputL, woutputL, material, wl, rgb, true);
        if (anyNonzero(f)) {
          RT Tr = transmittanceOverSegment(pW, woutputW, P.maxLengthScale, rgb, rnd);
          Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            float bsdfPdf = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, woutputL, material, wl, rgb, true));
            float misWeight = powerHeuristic(sunPdf, bsdfPdf);
            Ldirect += f * Li / max(PDF_EPSILON, sunPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
          }
        }
      }
    }
    if constexpr (SC::kHasSphereLight) {
      if ((P.sphereLightPower > RADIANCE_EPSILON) && (dot(pW - P.sphereLightPosition, basis.nW) <= P.sphereLightRadius)) {
        vec3 pHit;
        RT Li = pt.sampleSphereLight(rgb, rnd, pW, pHit, sphPdf);
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float hitDist = max(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
          woutputW = normalize(pHit - pW);
          woutputL = worldToLocal(woutputW, basis);
          RT f = pt.evaluateBsdf(pW, basis, winputL, woutputL, material, wl, rgb, true);
          if (anyNonzero(f)) {
            RT Tr = transmittanceOverSegment(pW, woutputW, hitDist, rgb, rnd);
            Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
            if (averageComponent(Li) > RADIANCE_EPSILON) {
              float bsdfPdf = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, woutputL, material, wl, rgb, true));
              float misWeight = powerHeuristic(sphPdf, bsdfPdf);
              Ldirect += f * Li / max(PDF_EPSILON, sphPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
            }
          }
        }
    }
    }
    return min(rt1(P.radianceClamp), Ldirect);
  }

  // directVolumeLighting (1299-1363)
  SNL_HD RT directVolumeLighting(const vec3& pW, const vec3& rayDir, const vec3& rgb, vec4& rnd) {
    RT Ldirect = rt1(0.0f);
    if (P.skyPower > RADIANCE_EPSILON) {
      vec3 woutputW(0.0f);
      float skyPdf = 0.0f;
      RT Li = pt.sampleSkyInVolume(rgb, rnd, woutputW, skyPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT Tr = transmittanceOverSegment(pW, woutputW, P.maxLengthScale, rgb, rnd);
        Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float PF = phaseFunction(dot(-rayDir, -woutputW), P.atmosphereAnisotropy);
          Ldirect += PF * Li / max(PDF_EPSILON, skyPdf);
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      vec3 woutputW(0.0f);
      float sunPdf = 0.0f;
      RT Li = pt.sampleSunInVolume(rgb, rnd, woutputW, sunPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT Tr = transmittanceOverSegment(pW, woutputW, P.maxLengthScale, rgb, rnd);
        Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float PF = phaseFunction(dot(-rayDir, -woutputW), P.atmosphereAnisotropy);
          Ldirect += PF * Li / max(PDF_EPSILON, sunPdf);
        }
      }
    }
    if constexpr (SC::kHasSphereLight) {
      if (P.sphereLightPower > RADIANCE_EPSILON) {
        float sphPdf = 0.0f;
        vec3 pHit;
        RT Li = pt.sampleSphereLight(rgb, rnd, pW, pHit, sphPdf);
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float hitDist = max(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
          vec3 woutputW = normalize(pHit - pW);
          RT Tr = transmittanceOverSegment(pW, woutputW, hitDist, rgb, rnd);
          Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            float PF = phaseFunction(dot(-rayDir, -woutputW), P.atmosphereAnisotropy);
            Ldirect += PF * Li / max(PDF_EPSILON, sphPdf);
          }
        }
    }
    }
    return min(rt1(P.radianceClamp), Ldirect);
  }

  // Direct light at a scattering point of the heterogeneous medium: one sample
  // of each light (the sky uniformly over the sphere, the sun's cone, the
  // sphere light's visible cap), shadowed by geometry and weighted by the
  // medium's ratio-tracking transmittance, combined with the phase-function
  // sample of the continuation by the balance heuristic.  The light pdfs go
  // back in skyPdf / sunPdf / sphPdf for the continuation's MIS weights.
  SNL_HD RT directMediumLighting(const vec3& pW, const vec3& rayDir, float g, const vec3& rgb, vec4& rnd,
                                 float& skyPdf, float& sunPdf, float& sphPdf) {
    RT Ldirect = rt1(0.0f);
    if (P.skyPower > RADIANCE_EPSILON) {
      vec3 woutputW = sampleSphereUnit(rnd);
      skyPdf = 1.0f / (4.0f * float(M_PI));
      RT Li = pt.environmentRadiance(woutputW, rgb);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT Tr = transmittanceOverSegment(pW, woutputW, P.maxLengthScale, rgb, rnd);
        Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float PF = phaseFunction(dot(rayDir, woutputW), g);
          Ldirect += PF * Li / skyPdf * powerHeuristic(skyPdf, PF);
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      vec3 woutputW(0.0f);
      RT Li = pt.sampleSunInVolume(rgb, rnd, woutputW, sunPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT Tr = transmittanceOverSegment(pW, woutputW, P.maxLengthScale, rgb, rnd);
        Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float PF = phaseFunction(dot(rayDir, woutputW), g);
          Ldirect += PF * Li / max(PDF_EPSILON, sunPdf) * powerHeuristic(sunPdf, PF);
        }
      }
    }
    if constexpr (SC::kHasSphereLight) {
      if (P.sphereLightPower > RADIANCE_EPSILON) {
        vec3 pHit;
        RT Li = pt.sampleSphereLight(rgb, rnd, pW, pHit, sphPdf);
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float hitDist = max(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
          vec3 woutputW = normalize(pHit - pW);
          RT Tr = transmittanceOverSegment(pW, woutputW, hitDist, rgb, rnd);
          Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            float PF = phaseFunction(dot(rayDir, woutputW), g);
            Ldirect += PF * Li / max(PDF_EPSILON, sphPdf) * powerHeuristic(sphPdf, PF);
          }
        }
    }
    }
    return min(rt1(P.radianceClamp), Ldirect);
  }

  // atmosphericInscatteringRadiance (1365-1417)
  SNL_HD RT atmosphericInscatteringRadiance(const vec3& pW, const vec3& rayDir, float segmentLength, const vec3& rgb,
                                            vec4& rnd) {
    RT Ls = rt1(0.0f);
    if constexpr (SC::kHasFog) {
      if (P.fogEnable) {
        RT TrFog = pt.transmittanceOverFreeSegment(pW, rayDir, segmentLength, rgb);
        Ls = (rt1(1.0f) - TrFog) * rgbToAlbedo(P.fogEmission, rgb);
      }
    }
    RT extinction = pt.VOLUME_EXTINCTION_EVAL(rgb);
    RT scattering = pt.VOLUME_SCATTERING_EVAL(rgb);
    float extinction_norm = averageComponent(extinction);
    float scattering_norm = averageComponent(scattering);
    if (extinction_norm < RADIANCE_EPSILON || scattering_norm < RADIANCE_EPSILON) return Ls;
    float t0 = 0.0f, t1 = 0.0f;
    if (!pt.atmosphereSegment(pW, rayDir, segmentLength, t0, t1)) return Ls;
    float T01 = exp(-(t1 - t0) * extinction_norm);
    float t_scatter = t0 - log(1.0f - snl_rand(rnd) * (1.0f - T01)) / extinction_norm;
    vec3 pW_scatter = pW + t_scatter * rayDir;
    RT opticalDepth_scatter = (t_scatter - t0) * extinction;
    RT Tr_pW = exp(-opticalDepth_scatter);
    float invDistancePdf = (1.0f - T01) * exp(averageComponent(opticalDepth_scatter)) / extinction_norm;
    RT Li = directVolumeLighting(pW_scatter, rayDir, rgb, rnd);
    Ls += Tr_pW * scattering * Li * invDistancePdf;
    if constexpr (SC::kHasVolumeEmission) {
      vec3 woW = samplePhaseFunction(rayDir, P.atmosphereAnisotropy, rnd);
      vec3 pW_hit;
      int hm;
      if (pt.traceRay(pW_scatter, woW, pW_hit, hm, P.maxLengthScale))
        Ls += Tr_pW * scattering * pt.VOLUME_EMISSION_EVAL(pW_hit, rgb) * invDistancePdf;
    }
    return Ls;
  }

  // atmosphericScatterSample (1419-1442)
  SNL_HD RT atmosphericScatterSample(const vec3& pW, const vec3& rayDir, float segmentLength, const vec3& rgb,
                                     vec4& rnd, vec3& pW_scatter, vec3& woutputW) {
    RT extinction = pt.VOLUME_EXTINCTION_EVAL(rgb);
    RT scattering = pt.VOLUME_SCATTERING_EVAL(rgb);
    float extinction_norm = averageComponent(extinction);
    float t0 = 0.0f, t1 = 0.0f;
    if (!pt.atmosphereSegment(pW, rayDir, segmentLength, t0, t1)) return rt1(0.0f);
    float T01 = exp(-(t1 - t0) * extinction_norm);
    float t_scatter = t0 - log(1.0f - snl_rand(rnd) * (1.0f - T01)) / extinction_norm;
    pW_scatter = pW + t_scatter * rayDir;
    RT opticalDepth_scatter = (t_scatter - t0) * extinction;
    RT Tr_pW = exp(-opticalDepth_scatter);
    float invDistancePdf = (1.0f - T01) * exp(averageComponent(opticalDepth_scatter)) / extinction_norm;
    woutputW = samplePhaseFunction(rayDir, P.atmosphereAnisotropy, rnd);
    return scattering * Tr_pW * invDistancePdf;
  }

  // randomwalk_SSS (1468-1518)
  SNL_HD bool randomwalk_SSS(const vec3& pW, const Basis& basis, vec4& rnd, float MFP, const RT& subsurfaceAlbedo,
                             RT& walk_throughput, vec3& pExit) {
    vec3 pWalk = pW - 3.0f * P.minLengthScale * basis.nW;
    float pdfDir;
    vec3 dirwalkL = sampleHemisphereCosineWeighted(rnd, pdfDir);
    vec3 dirwalkW = -localToWorld(dirwalkL, basis);
    walk_throughput = rt1(1.0f);
    for (int n = 0; n < 2048; ++n) {
      if (n > P.maxSSSSteps) break;
      float walk_step = -log(snl_rand(rnd)) * MFP * P.lengthScale;
      vec3 pHit;
      int hm;
      if (pt.traceRay(pWalk, dirwalkW, pHit, hm, walk_step)) {
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

  // SSS_exit_radiance (1521-1593)
  SNL_HD RT SSS_exit_radiance(const vec3& pW, const Basis& basis, const vec3& rgb, vec4& rnd,
                              const RT& diffuseAlbedoExit, float& skyPdf, float& sunPdf, float& sphPdf) {
    vec3 dPw = 3.0f * P.minLengthScale * basis.nW;
    RT Ldirect = rt1(0.0f);
    vec3 woutputL(0.0f), woutputW(0.0f);
    RT f = mix(rt1(1.0f), diffuseAlbedoExit, P.subsurfaceDiffuseWeight) / float(M_PI);
    if (P.skyPower > RADIANCE_EPSILON) {
      RT Li = pt.sampleSkyAtSurface(basis, rgb, rnd, woutputL, woutputW, skyPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT Tr = transmittanceOverSegment(pW + dPw, woutputW, P.maxLengthScale, rgb, rnd);
        Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float misWeight = powerHeuristic(skyPdf, pdfHemisphereCosineWeighted(woutputL));
          Ldirect += f * Li / max(PDF_EPSILON, skyPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
        }
      }
    }
    if (P.sunPower > RADIANCE_EPSILON) {
      RT Li = pt.sampleSunAtSurface(basis, rgb, rnd, woutputL, woutputW, sunPdf);
      if (averageComponent(Li) > RADIANCE_EPSILON) {
        RT Tr = transmittanceOverSegment(pW + dPw, woutputW, P.maxLengthScale, rgb, rnd);
        Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float misWeight = powerHeuristic(sunPdf, pdfHemisphereCosineWeighted(woutputL));
          Ldirect += f * Li / max(PDF_EPSILON, sunPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
        }
      }
    }
    if constexpr (SC::kHasSphereLight) {
      if ((P.sphereLightPower > RADIANCE_EPSILON) && (dot(pW - P.sphereLightPosition, basis.nW) <= P.sphereLightRadius)) {
        vec3 pHit;
        RT Li = pt.sampleSphereLight(rgb, rnd, pW, pHit, sphPdf);
        if (averageComponent(Li) > RADIANCE_EPSILON) {
          float hitDist = max(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
          woutputW = normalize(pHit - pW);
          woutputL = worldToLocal(woutputW, basis);
          RT Tr = transmittanceOverSegment(pW + dPw, woutputW, hitDist, rgb, rnd);
          Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            float misWeight = powerHeuristic(sphPdf, pdfHemisphereCosineWeighted(woutputL));
            Ldirect += f * Li / max(PDF_EPSILON, sphPdf) * abs(dot(woutputW, basis.nW)) * misWeight;
          }
        }
    }
    }
    return min(rt1(P.radianceClamp), Ldirect);
  }

  // one camera path (cameraPath, 1625-1926) as the state of a lane, so that a
  // lane whose path ends starts its next one at once instead of waiting for
  // the longest path of its wave
  struct Path {
    vec3 pW, rayDir;
    RT throughput, L;
    float misWeightSky, misWeightSun, misWeightSph;
    int vertex, atmosphere_scatters;
    bool inDielectric;
  };
  SNL_HD void startPath(Path& ps, const vec3& primaryStart, const vec3& primaryDir) {
    ps.L = rt1(0.0f);
    ps.misWeightSky = ps.misWeightSun = ps.misWeightSph = 1.0f;
    ps.pW = primaryStart;
    ps.rayDir = primaryDir;
    ps.inDielectric = false;
    if constexpr (SC::kHasDielectric) ps.inDielectric = pt.S.SDF_DIELECTRIC(primaryStart) < 0.0f;
    ps.throughput = rt1(1.0f);
    ps.atmosphere_scatters = 0;
    ps.vertex = 0;
  }
  // the loop's increment and condition: whether the path goes on
  SNL_HD static bool advance(Path& ps, int maxBounces) { return ++ps.vertex <= maxBounces; }

  // A real collision of the ray (ps.pW, ps.rayDir) with the heterogeneous
  // medium at distance tc, before the surface it would hit: the vertex
  // scatters there.  The sphere light, if the ray passes it first, adds its
  // emission as on the way to a surface hit (1721-1735); Beer absorption inside
  // a dielectric applies over tc (1739-1745); the throughput takes the
  // albedo; direct light as in a surface vertex (not inside a dielectric,
  // 1846-1850); the continuation follows the phase function.
  SNL_HD bool mediumVertex(Path& ps, float tc, const vec3& rgb, vec4& rnd, int maxBounces) {
    if constexpr (SC::kHasSphereLight) {
      {
        vec3 pLightHit;
        RT LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? rt1(0.0f) : pt.sphereLightRadiance(ps.pW, ps.rayDir, rgb, pLightHit);
        if (averageComponent(LiSph) > RADIANCE_EPSILON && length(pLightHit - ps.pW) < tc)
          ps.L += ps.throughput * ps.misWeightSph * LiSph;
    }
    }
    if constexpr (SC::kHasDielectric) {
      if (ps.inDielectric) ps.throughput *= exp(-tc * pt.DIELECTRIC_ABSORPTION_EVAL(ps.pW, rgb));
    }
    ps.pW += tc * ps.rayDir;
    ps.throughput *= pt.volumeAlbedo(ps.pW, rgb);
    float g = pt.volumeAnisotropy(ps.pW);
    float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
    if (!ps.inDielectric) ps.L += ps.throughput * directMediumLighting(ps.pW, ps.rayDir, g, rgb, rnd, skyPdf, sunPdf, sphPdf);
    float phasePdf = 0.0f;
    ps.rayDir = sampleHenyeyGreenstein(ps.rayDir, g, rnd, phasePdf);
    ps.misWeightSky = powerHeuristic(phasePdf, skyPdf);
    ps.misWeightSun = powerHeuristic(phasePdf, sunPdf);
    ps.misWeightSph = powerHeuristic(phasePdf, sphPdf);
    if (maxComponent(ps.throughput) < THROUGHPUT_EPSILON) return false;
    return advance(ps, maxBounces);
  }

  // the body of cameraPath's vertex loop for vertex ps.vertex: `break` ends the
  // path (false), `continue` and the end of the body go to the next vertex
  SNL_HD bool vertexStep(Path& ps, float wl, const vec3& rgb, vec4& rnd, int maxBounces, long long& normals) {
    RT& L = ps.L;
    RT& throughput = ps.throughput;
    vec3& pW = ps.pW;
    vec3& rayDir = ps.rayDir;
    float& misWeightSky = ps.misWeightSky;
    float& misWeightSun = ps.misWeightSun;
    float& misWeightSph = ps.misWeightSph;
    bool& inDielectric = ps.inDielectric;
    int& atmosphere_scatters = ps.atmosphere_scatters;
    const int vertex = ps.vertex;
    (void)misWeightSph;
    (void)atmosphere_scatters;
    (void)vertex;
    if (maxComponent(throughput) < THROUGHPUT_EPSILON) return false;
    vec3 pW_next;
    int hitMaterial;
    bool hit = pt.traceRay(pW, rayDir, pW_next, hitMaterial, P.maxLengthScale);
    float rayLength = P.maxLengthScale;
    if (hit) rayLength = length(pW_next - pW);
    if constexpr (SC::kHasVolume) {
      // delta tracking up to the surface hit (or maxLengthScale on a miss)
      float tc = pt.volumeFreeFlight(pW, rayDir, rayLength, rnd);
      if (tc < rayLength) return mediumVertex(ps, tc, rgb, rnd, maxBounces);
    }
    if (!hit) {
      float lightSegmentLength = P.maxLengthScale;
      RT Tr = rt1(1.0f);
      bool litSph = false;
      if constexpr (SC::kHasSphereLight) {
        {
          vec3 pLightHit;
          RT LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? rt1(0.0f) : pt.sphereLightRadiance(pW, rayDir, rgb, pLightHit);
          if (averageComponent(LiSph) > RADIANCE_EPSILON) {
            litSph = true;
            lightSegmentLength = length(pLightHit - pW);
            Tr = pt.transmittanceOverFreeSegment(pW, rayDir, lightSegmentLength, rgb);
            float scatter_prob = pt.atmosphericScatteringProbability(Tr, rgb);
            L += throughput * Tr * misWeightSph * LiSph / (1.0f - scatter_prob);
          }
      }
      }
      if (!litSph) {
        lightSegmentLength = P.maxLengthScale;
        Tr = pt.transmittanceOverFreeSegment(pW, rayDir, lightSegmentLength, rgb);
        float scatter_prob = pt.atmosphericScatteringProbability(Tr, rgb);
        if (averageComponent(Tr) > RADIANCE_EPSILON) {
          if (!(vertex == 0 && !P.envMapVisible) && misWeightSky > 0.0f)
            L += throughput * Tr * misWeightSky * pt.environmentRadiance(rayDir, rgb) / (1.0f - scatter_prob);
          if (!(vertex == 0 && !P.sunVisibleDirectly) && misWeightSun > 0.0f)
            L += throughput * Tr * misWeightSun * pt.sunRadiance(rayDir, rgb) / (1.0f - scatter_prob);
        }
      }
      if constexpr (SC::kHasAtmosphere) {
        if (vertex < 2) L += throughput * atmosphericInscatteringRadiance(pW, rayDir, lightSegmentLength, rgb, rnd);
        if (atmosphere_scatters + 1 < SC::kMaxAtmosphereScatters) {
          float scatter_prob = pt.atmosphericScatteringProbability(Tr, rgb);
          if (scatter_prob > snl_rand(rnd)) {
            vec3 pW_scatter(0.0f), woutputW(0.0f);
            throughput *= atmosphericScatterSample(pW, rayDir, lightSegmentLength, rgb, rnd, pW_scatter, woutputW) /
                          max(DENOM_TOLERANCE, scatter_prob);
            pW = pW_scatter;
            rayDir = woutputW;
            atmosphere_scatters++;
            return advance(ps, maxBounces);
          }
        }
      }
      return false;
    }
    if (hitMaterial < 0) return false;  // a max-step "hit" has no material: the path ends

    RT Tr = rt1(1.0f);
    bool hitSphereLight = false;
    if constexpr (SC::kHasSphereLight) {
      {
        vec3 pLightHit;
        RT LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? rt1(0.0f) : pt.sphereLightRadiance(pW, rayDir, rgb, pLightHit);
        if (averageComponent(LiSph) > RADIANCE_EPSILON) {
          float lightSegmentLength = length(pLightHit - pW);
          if (lightSegmentLength < rayLength) {
            hitSphereLight = true;
            Tr = pt.transmittanceOverFreeSegment(pW, rayDir, lightSegmentLength, rgb);
            L += throughput * Tr * misWeightSph * LiSph;
            if constexpr (SC::kHasAtmosphere) {
              if (vertex < 2) L += throughput * atmosphericInscatteringRadiance(pW, rayDir, lightSegmentLength, rgb, rnd);
            }
          }
        }
    }
    }
    if constexpr (SC::kHasDielectric) {
      if (inDielectric) throughput *= exp(-rayLength * pt.DIELECTRIC_ABSORPTION_EVAL(pW, rgb));
    }
    if constexpr (SC::kHasAtmosphere) {
      if (vertex < 2) L += throughput * atmosphericInscatteringRadiance(pW, rayDir, rayLength, rgb, rnd);
      if (!hitSphereLight) Tr = pt.transmittanceOverFreeSegment(pW, rayDir, rayLength, rgb);
      if (atmosphere_scatters + 1 < SC::kMaxAtmosphereScatters) {
        float scatter_prob = pt.atmosphericScatteringProbability(Tr, rgb);
        if (scatter_prob > snl_rand(rnd)) {
          vec3 pW_scatter(0.0f), woutputW(0.0f);
          throughput *= atmosphericScatterSample(pW, rayDir, rayLength, rgb, rnd, pW_scatter, woutputW) /
                        max(DENOM_TOLERANCE, scatter_prob);
          pW = pW_scatter;
          rayDir = woutputW;
          atmosphere_scatters++;
          return advance(ps, maxBounces);
        } else {
          throughput /= max(DENOM_TOLERANCE, 1.0f - scatter_prob);
        }
      }
      if (hitSphereLight) return false;
      throughput *= pt.transmittanceOverFreeSegment(pW, rayDir, rayLength, rgb);
    }
    (void)Tr;
    (void)hitSphereLight;

    // surface interaction
    pW = pW_next;
    vec3 nW = pt.normal(pW, hitMaterial, true);
    ++normals;
    vec3 ngW = nW;
    Basis basis = makeBasis(nW);
    if constexpr (SC::kHasNormalMap) {
      nW = pt.perturbNormal(pW, basis, hitMaterial);
      if ((dot(nW, rayDir) > 0.0f) && (hitMaterial != MAT_DIELE)) nW = 2.0f * ngW * dot(ngW, nW) - nW;
      basis = makeBasis(nW);
    }
    bool do_subsurface_walk = false;
    float prob_sss = 0.0f;
    RT subsurfaceAlbedo = rt1(0.0f);
    if constexpr (SC::kHasSurface) {
      if (P.subsurfaceMFP > 0.0f) {
        subsurfaceAlbedo = pt.SUBSURFACE_ALBEDO_EVAL(pW, basis.nW, rgb);
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
      RT f = pt.sampleBsdf(pW, basis, winputL, hitMaterial, wl, rgb, true, woutputL, bsdfPdf, rnd);
      vec3 woutputW = localToWorld(woutputL, basis);
      if constexpr (SC::kHasDielectric) {
        if (hitMaterial == MAT_DIELE && dot(winputW, ngW) * dot(woutputW, ngW) < 0.0f) inDielectric = !inDielectric;
      }
      if constexpr (SC::kHasVolumeEmission) L += throughput * pt.VOLUME_EMISSION_EVAL(pW, rgb);
      rayDir = woutputW;
      pW += ngW * sign(dot(rayDir, ngW)) * 3.0f * P.minLengthScale;
      float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
      if (!inDielectric)
        L += throughput * directSurfaceLighting(pW, basis, winputW, hitMaterial, wl, rgb, rnd, skyPdf, sunPdf, sphPdf);
      RT fOverPdf = min(rt1(P.radianceClamp), f / max(PDF_EPSILON, bsdfPdf));
      RT surface_throughput = fOverPdf * abs(dot(woutputW, nW)) / max(PDF_EPSILON, 1.0f - prob_sss);
      throughput *= surface_throughput;
      misWeightSky = powerHeuristic(bsdfPdf, skyPdf);
      misWeightSun = powerHeuristic(bsdfPdf, sunPdf);
      misWeightSph = powerHeuristic(bsdfPdf, sphPdf);
    } else {
      if constexpr (SC::kHasSurface) {
        RT diffuseAlbedoEntry = pt.SURFACE_DIFFUSE_REFL_EVAL(pW, basis.nW, -rayDir, rgb);
        (void)diffuseAlbedoEntry;  // weighs the entry lobe, which the walk leaves at 1 (1486-1489)
        RT walk_throughput = rt1(1.0f);
        vec3 pExit;
        basis.nW = ngW;
        if (!randomwalk_SSS(pW, basis, rnd, P.subsurfaceMFP, subsurfaceAlbedo, walk_throughput, pExit)) return false;
        pW = pExit;
        nW = pt.normal(pExit, MAT_SURFA, true);
        ++normals;
        Basis basis_exit = makeBasis(nW);
        float bsdfPdf;
        vec3 dirExitL = sampleHemisphereCosineWeighted(rnd, bsdfPdf);
        vec3 woutputW = localToWorld(dirExitL, basis_exit);
        rayDir = woutputW;
        throughput *= walk_throughput;
        float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
        RT diffuseAlbedoExit = pt.SURFACE_DIFFUSE_REFL_EVAL(pW, nW, -rayDir, rgb);
        if (!inDielectric)
          L += throughput * SSS_exit_radiance(pExit, basis_exit, rgb, rnd, diffuseAlbedoExit, skyPdf, sunPdf, sphPdf);
        misWeightSky = powerHeuristic(bsdfPdf, skyPdf);
        misWeightSun = powerHeuristic(bsdfPdf, sunPdf);
        misWeightSph = powerHeuristic(bsdfPdf, sphPdf);
        RT f = mix(rt1(1.0f), diffuseAlbedoExit, P.subsurfaceDiffuseWeight) / float(M_PI);
        RT fOverPdf = min(rt1(P.radianceClamp), f / max(PDF_EPSILON, bsdfPdf));
        RT surface_exit_throughput = fOverPdf * abs(dot(woutputW, nW));
        throughput *= surface_exit_throughput / max(PDF_EPSILON, prob_sss);
        pW += nW * sign(dot(rayDir, nW)) * 3.0f * P.minLengthScale;
      }
    }
    if (maxComponent(throughput) < THROUGHPUT_EPSILON) return false;
    return advance(ps, maxBounces);
  }
};

template <class SC>
__device__ inline void mega_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  Mega<SC> mg(pt, P);
  const int lane = snl_lane();
  const unsigned long long below = (1ull << lane) - 1ull;
  const int nItems = a.itemEnd - a.itemBase;
  int* fetch = &a.cnt[C_FETCH];
  int rb = 0, re = 0;
  bool exhausted = false;
  // the lane's pixel: its frames run back to back, RNG and running mean in registers
  bool havePix = false, frameOpen = false, inPath = false;
  int pix = 0, frame = 0, sample = 0;
  vec4 rnd(1.0f), accum(0.0f);
  vec2 pixel(0.0f);
  // the frame being rendered (pathtrace, 1934-1989)
  float wl = 0.0f;
  vec3 rgb(0.0f), xyz(0.0f);
  RT Lframe = rt1(0.0f);
  typename Mega<SC>::Path ps;
  long long normals = 0;
  unsigned long long pixSteps = 0;  // pt.steps when the lane took its pixel
  while (true) {
    // ---- lanes without a pixel take the next ones of the range
    unsigned long long need = __ballot(!havePix);
    if (need != 0ull && rb >= re && !exhausted) {
      int base = 0;
      if (lane == 0) base = atomicAdd(fetch, a.chunk);
      base = __shfl(base, 0);
      if (base >= nItems) exhausted = true;
      else { rb = base; re = min(base + a.chunk, nItems); }
    }
    if (need != 0ull && rb < re) {
      int avail = re - rb;
      int rank = __popcll(need & below);
      if (!havePix && rank < avail) {
        pix = snl_item_pixel(a, snl_tile_item(a, rb + rank));
        havePix = true;
        frame = 0;
        sample = 0;
        frameOpen = false;
        pixSteps = pt.steps;
        rnd = ld4(a.rngstate, pix);
        accum = ld4(a.radiance, pix);
        pixel = vec2(float(pix % a.W) + 0.5f, float((pix % a.imagePix) / a.W) + 0.5f);
      }
      rb += min(avail, (int)__popcll(need));
    }
    if (__ballot(havePix) == 0ull) {
      if (exhausted) break;
      continue;
    }
    // ---- a lane whose path ended starts the next sample, opening frames and
    // handing back its pixel after the last one
    if (havePix && !inPath) {
      while (!frameOpen) {
        if (frame >= a.frames) {
          st4(a.radiance, pix, accum);
          st4(a.rngstate, pix, rnd);
          havePix = false;
          if (a.tileCost) {
            // the pixel's march steps to its 8x8 tile; the 16 pixels with even
            // coordinates stand for the tile (a quarter of the atomics)
            int i = pix % a.imagePix, y = i / a.W, x = i - y * a.W;
            if (((x | y) & 1) == 0) {
              unsigned long long d = pt.steps - pixSteps;
              atomicAdd(&a.tileCost[(y >> 3) * (a.W >> 3) + (x >> 3)], (int)(d < 0x7fffffffull ? d : 0x7fffffffull));
            }
          }
          break;
        }
        if constexpr (SC::kInteractive) {
          // main (1991-2007): a skipped frame leaves the pixel as it is
          if (snl_rand(rnd) < P.skipProbability) { ++frame; continue; }
        }
#ifdef DISPERSION_ENABLED
        float xi = 0.1f + 0.8f * (0.5f + floor(float(P.wavelengthSamples) * snl_rand(rnd))) / float(P.wavelengthSamples);
        float w = texture(P.ICDF, vec2(xi, 0.5f)).r;
        wl = 390.0f + (750.0f - 390.0f) * w;
        xyz = vec3(texture(P.WavelengthToXYZ, vec2(w, 0.5f)));
        rgb = xyzToRgb(xyz);
#else
        wl = 390.0f + (750.0f - 390.0f) * 0.5f;
        rgb = vec3(0.0f);
#endif
        Lframe = rt1(0.0f);
        sample = 0;
        frameOpen = true;
      }
      if (havePix) {
        float jx = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
        float jy = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
        vec2 pixelj = pixel + vec2(jx, jy);
        vec3 primaryStart, primaryDir;
        if constexpr (SC::kHasCustomCamera) S.CONSTRUCT_PRIMARY_RAY(pixelj, rnd, primaryStart, primaryDir);
        else pt.constructPrimaryRay(pixelj, rnd, primaryStart, primaryDir, true);
        mg.startPath(ps, primaryStart, primaryDir);
        inPath = true;
      }
    }
    // ---- one vertex of every running path
    if (inPath) {
      inPath = mg.vertexStep(ps, wl, rgb, rnd, a.maxBounces, normals);
      if (!inPath) {
        Lframe += ps.L;
        if (++sample == a.spf) {
          RT L = Lframe / float(a.spf);
#ifdef DISPERSION_ENABLED
          vec3 colorXYZ = xyz * L;
#else
          vec3 colorXYZ = rgbToXyz(L);
#endif
          float oldN = accum.w, newN = oldN + 1.0f;
          accum = vec4((oldN * vec3(accum) + colorXYZ) / newN, newN);
          frameOpen = false;
          ++frame;
        }
      }
    }
  }
  unsigned long long tot = pt.steps, ntot = normals;
  for (int off = 32; off > 0; off >>= 1) {
    tot += __shfl_xor(tot, off);
    ntot += __shfl_xor(ntot, off);
  }
  if (lane == 0 && tot) atomicAdd(reinterpret_cast<unsigned long long*>(&a.cnt[C_TRACE_STEPS]), tot);
  if (lane == 0 && ntot) atomicAdd(reinterpret_cast<unsigned long long*>(&a.cnt[C_NORMALS]), ntot);
}

// four waves per SIMD (at most 128 VGPRs) unless the runtime asks for fewer:
// it builds the program again with SNL_MEGA_WAVES=1 when four make this kernel
// spill to scratch
#ifndef SNL_MEGA_WAVES
#define SNL_MEGA_WAVES 4
#endif
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256, SNL_MEGA_WAVES) snl_mega(SnlKArgs a) { mega_body<SnlScene>(a); }
#endif

}  // namespace glsl
