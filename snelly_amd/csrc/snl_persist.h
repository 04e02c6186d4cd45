// snl_persist.h -- persistent path kernel: one launch renders `frames` frames
// of a range of pixels with no per-bounce or per-frame barrier.
//
// Every lane runs whole camera paths (cameraPath, pathtracer-fragment-shader.glsl
// 1625-1926) as a small state machine held in registers:
//   CAMERA -> TRACE (continuation ray) -> SHADE -> TRACE (shadow rays of the
//   vertex's NEE, one after another) -> TRACE (next continuation) ... -> DONE
// and takes a new pixel from the tile's queue once its pixel has run all its
// frames.  A pixel's frames run back to back in one lane, its RNG state and
// running mean held in registers: the per-pixel sequence of operations is the
// same as one launch per frame (pixels never read each other's state), so the
// result is too.  A wave advances all its tracing lanes by one SDF step per
// iteration; lanes that need shading wait until at least `shadeThresh` of them
// do (or nothing is tracing), so the expensive shading code runs with most
// lanes active while marching keeps the rest busy.  Random numbers are drawn
// in the reference's per-path order (light samples are drawn at the vertex,
// their shadow rays traced afterwards), so results match the wavefront kernels
// and the reference bit for bit where the float arithmetic does.
#pragma once

namespace glsl {

enum { PS_IDLE = 0, PS_CAMERA, PS_TRACE, PS_SHADE, PS_DONE };

template <class SC>
__device__ inline void persist_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  const int lane = snl_lane();
  const unsigned long long below = (1ull << lane) - 1ull;
  const int nItems = a.itemEnd - a.itemBase;
  int* fetch = &a.cnt[C_FETCH];

  int rb = 0, re = 0;
  bool exhausted = false;
  int st = PS_IDLE;
  int pix = 0, sample = 0, frame = 0;
  vec4 accum(0.0f);  // the pixel's (X, Y, Z, N) over its frames
  // path state
  vec3 o(0.0f), d(0.0f), hp(0.0f);
  RT tp = rt1(1.0f), L = rt1(0.0f), Ls = rt1(0.0f);
  float wl = 0.0f;
  vec3 xyz(0.0f), rgb(0.0f);
  float misSky = 1.0f, misSun = 1.0f, misSph = 1.0f;
  int vertex = 0;
  bool inDiele = false, skip = false;
  vec4 rnd(1.0f);
  // tracer state
  float t = 0.0f, sgn = 0.0f, tmax = 0.0f;
  int steps = 0, hm = MAT_INVAL;
  bool init = false, shadow = false, contPending = false;
  vec3 contDir(0.0f);
  // NEE of the current vertex: per-light record (direction, length, f, Li, pdf denominator, cosine, MIS weight)
  int neeMask = 0, light = 0;
  RT acc = rt1(0.0f), ntp = rt1(0.0f);
  vec3 nd0(0.0f), nd1(0.0f), nd2(0.0f);
  float nt0 = 0.0f, nt1 = 0.0f, nt2 = 0.0f;
  RT nf0 = rt1(0.0f), nf1 = rt1(0.0f), nf2 = rt1(0.0f), nl0 = rt1(0.0f), nl1 = rt1(0.0f), nl2 = rt1(0.0f);
  float nden0 = 1.0f, nden1 = 1.0f, nden2 = 1.0f, ncos0 = 0.0f, ncos1 = 0.0f, ncos2 = 0.0f;
  float nmis0 = 0.0f, nmis1 = 0.0f, nmis2 = 0.0f;
  long long stepTotal = 0, normalTotal = 0;
  long long pixSteps = 0;  // stepTotal when the lane took its pixel
  // the pixel's march steps to its 8x8 tile (the 16 pixels with even
  // coordinates stand for the tile), for a render's tile order
  auto addTileCost = [&]() {
    if (!a.tileCost) return;
    int i = pix % a.imagePix, y = i / a.W, x = i - y * a.W;
    if (((x | y) & 1) == 0) {
      long long d = stepTotal - pixSteps;
      atomicAdd(&a.tileCost[(y >> 3) * (a.W >> 3) + (x >> 3)], (int)(d < 0x7fffffffll ? d : 0x7fffffffll));
    }
  };

  auto startRay = [&](const vec3& dir, float len, bool isShadow) {
    d = dir;
    tmax = len;
    shadow = isShadow;
    init = true;
    steps = 0;
    st = PS_TRACE;
  };
  auto startLight = [&](int l) {
    light = l;
    if (l == 0) startRay(nd0, nt0, true);
    else if (l == 1) startRay(nd1, nt1, true);
    else startRay(nd2, nt2, true);
  };
  // all shadow rays of the vertex are resolved: add the clamped direct light
  // (directSurfaceLighting's return, 1294) and continue the path
  auto finishNee = [&]() {
    L += ntp * min(rt1(P.radianceClamp), acc);
    neeMask = 0;
    if (contPending) { contPending = false; startRay(contDir, P.maxLengthScale, false); }
    else st = PS_DONE;
  };

  while (true) {
    // ---- refill idle lanes with pixels of this tile
    unsigned long long idleMask = __ballot(st == PS_IDLE);
    if (idleMask != 0ull && rb >= re && !exhausted) {
      int base = 0;
      if (lane == 0) base = atomicAdd(fetch, a.chunk);
      base = __shfl(base, 0);
      if (base >= nItems) exhausted = true;
      else { rb = base; re = min(base + a.chunk, nItems); }
    }
    if (idleMask != 0ull && rb < re) {
      int avail = re - rb;
      int rank = __popcll(idleMask & below);
      if (st == PS_IDLE && rank < avail) {
        pix = snl_item_pixel(a, snl_tile_item(a, rb + rank));
        sample = 0;
        frame = 0;
        pixSteps = stepTotal;
        st = PS_CAMERA;
      }
      rb += min(avail, (int)__popcll(idleMask));
    }
    unsigned long long busy = __ballot(st != PS_IDLE);
    if (busy == 0ull) {
      if (exhausted) break;
      continue;
    }

    // ---- up to a.burst march steps for every tracing lane (traceRay, 881-920);
    // the shading gate is checked once per burst
    for (int burst = 0; burst < a.burst; ++burst) {
      if (st == PS_TRACE) {
        vec3 p = init ? o : o + t * d;
        float sdf;
        int m = pt.sdfStep(p, !init, sdf);
        ++stepTotal;
        bool done = false, hit = false;
        if (init) {
          init = false;
          sgn = sign(sdf);
          t = sgn * sdf;
          if (abs(t) >= tmax) done = true;
        } else if (m != MAT_INVAL) {
          done = true;
          hit = true;
        } else {
          t += sgn * sdf;
          ++steps;
          if (abs(t) >= tmax) done = true;
          else if (steps >= SC::kMaxMarchSteps) { done = true; hit = !P.maxStepsIsMiss; m = -2; p = o + t * d; }
        }
        if (done) {
          if (!shadow) {
            hm = hit ? m : MAT_INVAL;
            hp = p;
            st = PS_SHADE;
          } else {
            // fold one shadow ray into the vertex's direct light (1236-1245)
            RT Tr = rt1(0.0f);
            if (!hit) Tr = pt.transmittanceOverFreeSegment(o, d, tmax, rgb);
            RT f = light == 0 ? nf0 : (light == 1 ? nf1 : nf2);
            RT Li = light == 0 ? nl0 : (light == 1 ? nl1 : nl2);
            float den = light == 0 ? nden0 : (light == 1 ? nden1 : nden2);
            float cs = light == 0 ? ncos0 : (light == 1 ? ncos1 : ncos2);
            float w = light == 0 ? nmis0 : (light == 1 ? nmis1 : nmis2);
            Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
            if (averageComponent(Li) > RADIANCE_EPSILON) acc += f * Li / den * cs * w;
            neeMask &= ~(1 << light);
            if (neeMask) startLight(__ffs(neeMask) - 1);
            else finishNee();
          }
        }
      }
      if (__ballot(st == PS_TRACE) == 0ull) break;
      if ((int)__popcll(__ballot(st == PS_CAMERA || st == PS_SHADE || st == PS_DONE)) >= a.shadeThresh) break;
    }

    // ---- shading, when enough lanes wait for it
    unsigned long long waiting = __ballot(st == PS_CAMERA || st == PS_SHADE || st == PS_DONE);
    unsigned long long tracing = __ballot(st == PS_TRACE);
    if (waiting == 0ull) continue;
    if ((int)__popcll(waiting) < a.shadeThresh && tracing != 0ull) continue;

    if (st == PS_SHADE) {
      if (hm == MAT_INVAL) {
        // miss (1657-1712)
        bool lit = false;
        if constexpr (SC::kHasSphereLight) {
          vec3 pLightHit;
          RT LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? rt1(0.0f) : pt.sphereLightRadiance(o, d, rgb, pLightHit);
          if (averageComponent(LiSph) > RADIANCE_EPSILON) {
            lit = true;
            RT Tr = pt.transmittanceOverFreeSegment(o, d, length(pLightHit - o), rgb);
            float sp = pt.atmosphericScatteringProbability(Tr, rgb);
            L += tp * Tr * misSph * LiSph / (1.0f - sp);
          }
        }
        if (!lit) {
          RT Tr = pt.transmittanceOverFreeSegment(o, d, P.maxLengthScale, rgb);
          float sp = pt.atmosphericScatteringProbability(Tr, rgb);
          if (averageComponent(Tr) > RADIANCE_EPSILON) {
            if (!(vertex == 0 && !P.envMapVisible) && misSky > 0.0f)
              L += tp * Tr * misSky * pt.environmentRadiance(d, rgb) / (1.0f - sp);
            if (!(vertex == 0 && !P.sunVisibleDirectly) && misSun > 0.0f)
              L += tp * Tr * misSun * pt.sunRadiance(d, rgb) / (1.0f - sp);
          }
        }
        st = PS_DONE;
      } else if (hm < 0) {
        st = PS_DONE;  // a max-step "hit" without a material ends the path
      } else {
        // surface interaction (1714-1923)
        const int hitMaterial = hm;
        vec3 pW = o, rayDir = d, pW_next = hp;
        float rayLength = length(pW_next - pW);
        if constexpr (SC::kHasSphereLight) {
          vec3 pLightHit;
          RT LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? rt1(0.0f) : pt.sphereLightRadiance(pW, rayDir, rgb, pLightHit);
          if (averageComponent(LiSph) > RADIANCE_EPSILON) {
            float lsl = length(pLightHit - pW);
            if (lsl < rayLength) {
              RT Tr = pt.transmittanceOverFreeSegment(pW, rayDir, lsl, rgb);
              L += tp * Tr * misSph * LiSph;
            }
          }
        }
        if constexpr (SC::kHasDielectric) {
          if (inDiele) tp *= exp(-rayLength * pt.DIELECTRIC_ABSORPTION_EVAL(pW, rgb));
        }
        pW = pW_next;
        vec3 nW = pt.normal(pW, hitMaterial, true);
        ++normalTotal;
        vec3 ngW = nW;
        Basis basis = makeBasis(nW);
        if constexpr (SC::kHasNormalMap) {
          nW = pt.perturbNormal(pW, basis, hitMaterial);
          if ((dot(nW, rayDir) > 0.0f) && (hitMaterial != MAT_DIELE)) nW = 2.0f * ngW * dot(ngW, nW) - nW;
          basis = makeBasis(nW);
        }
        float prob_sss = 0.0f;
        bool sss = false;
        RT subsurfaceAlbedo = rt1(0.0f);
        if constexpr (SC::kHasSurface) {
          if (hitMaterial == MAT_SURFA && P.subsurfaceMFP > 0.0f) {
            subsurfaceAlbedo = pt.SUBSURFACE_ALBEDO_EVAL(pW, basis.nW, rgb);
            float dw = (1.0f - P.subsurface) * averageComponent(pt.SURFACE_DIFFUSE_REFL_EVAL(pW, basis.nW, -rayDir, rgb));
            float sw = averageComponent(pt.SURFACE_SPEC_REFL_EVAL(pW, basis.nW, -rayDir, rgb));
            float ssw = P.subsurface * averageComponent(subsurfaceAlbedo) * 3.0f;
            prob_sss = clamp(ssw / (dw + sw + ssw + DENOM_TOLERANCE), 0.0f, 1.0f - PDF_EPSILON);
            sss = (snl_rand(rnd) < prob_sss);
          }
        }
        if (sss) {
          // the sub-surface walk (1866-1918) is not run by this kernel; the
          // runtime routes programs with sub-surface scattering to the
          // wavefront kernels, so this branch is unreachable
          st = PS_DONE;
        } else {
          vec3 winputW = -rayDir;
          vec3 winputL = worldToLocal(winputW, basis);
          vec3 woutputL(0.0f);  // WebGL zero-initialises these locals
          float bsdfPdf = 0.0f;
          RT f = pt.sampleBsdf(pW, basis, winputL, hitMaterial, wl, rgb, true, woutputL, bsdfPdf, rnd);
          vec3 woutputW = localToWorld(woutputL, basis);
          if (hitMaterial == MAT_DIELE && dot(winputW, ngW) * dot(woutputW, ngW) < 0.0f) inDiele = !inDiele;
          if constexpr (SC::kHasVolumeEmission) L += tp * pt.VOLUME_EMISSION_EVAL(pW, rgb);
          rayDir = woutputW;
          pW += ngW * sign(dot(rayDir, ngW)) * 3.0f * P.minLengthScale;
          float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
          // a light whose BSDF value is exactly zero adds exactly zero whatever
          // its shadow ray finds: that ray is not traced
          neeMask = 0;
          if (!inDiele) {
            vec3 wL, wW;
            if (P.skyPower > RADIANCE_EPSILON) {
              RT Li = pt.sampleSkyAtSurface(basis, rgb, rnd, wL, wW, skyPdf);
              if (averageComponent(Li) > RADIANCE_EPSILON) {
                float lp = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, wL, hitMaterial, wl, rgb, true));
                nf0 = pt.evaluateBsdf(pW, basis, winputL, wL, hitMaterial, wl, rgb, true);
                nl0 = Li; nd0 = wW; nt0 = P.maxLengthScale;
                nden0 = max(PDF_EPSILON, skyPdf); ncos0 = abs(dot(wW, basis.nW)); nmis0 = powerHeuristic(skyPdf, lp);
                if (anyNonzero(nf0)) neeMask |= 1;
              }
            }
            if (P.sunPower > RADIANCE_EPSILON) {
              RT Li = pt.sampleSunAtSurface(basis, rgb, rnd, wL, wW, sunPdf);
              if (averageComponent(Li) > RADIANCE_EPSILON) {
                float lp = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, wL, hitMaterial, wl, rgb, true));
                nf1 = pt.evaluateBsdf(pW, basis, winputL, wL, hitMaterial, wl, rgb, true);
                nl1 = Li; nd1 = wW; nt1 = P.maxLengthScale;
                nden1 = max(PDF_EPSILON, sunPdf); ncos1 = abs(dot(wW, basis.nW)); nmis1 = powerHeuristic(sunPdf, lp);
                if (anyNonzero(nf1)) neeMask |= 2;
              }
            }
            if constexpr (SC::kHasSphereLight) {
              if ((P.sphereLightPower > RADIANCE_EPSILON) && (dot(pW - P.sphereLightPosition, basis.nW) <= P.sphereLightRadius)) {
                vec3 pHit;
                RT Li = pt.sampleSphereLight(rgb, rnd, pW, pHit, sphPdf);
                if (averageComponent(Li) > RADIANCE_EPSILON) {
                  wW = normalize(pHit - pW);
                  wL = worldToLocal(wW, basis);
                  float lp = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, wL, hitMaterial, wl, rgb, true));
                  nf2 = pt.evaluateBsdf(pW, basis, winputL, wL, hitMaterial, wl, rgb, true);
                  nl2 = Li; nd2 = wW; nt2 = max(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
                  nden2 = max(PDF_EPSILON, sphPdf); ncos2 = abs(dot(wW, basis.nW)); nmis2 = powerHeuristic(sphPdf, lp);
                  if (anyNonzero(nf2)) neeMask |= 4;
                }
              }
            }
          }
          ntp = tp;
          acc = rt1(0.0f);
          RT fOverPdf = min(rt1(P.radianceClamp), f / max(PDF_EPSILON, bsdfPdf));
          tp *= fOverPdf * abs(dot(woutputW, nW)) / max(PDF_EPSILON, 1.0f - prob_sss);
          misSky = powerHeuristic(bsdfPdf, skyPdf);
          misSun = powerHeuristic(bsdfPdf, sunPdf);
          misSph = powerHeuristic(bsdfPdf, sphPdf);
          o = pW;
          contDir = rayDir;
          contPending = !(maxComponent(tp) < THROUGHPUT_EPSILON) && vertex + 1 <= a.maxBounces;
          if (contPending) ++vertex;
          if (neeMask) startLight(__ffs(neeMask) - 1);
          else if (contPending) { contPending = false; startRay(contDir, P.maxLengthScale, false); }
          else st = PS_DONE;
        }
      }
    }

    if (st == PS_DONE) {
      // end of one camera path: sum the frame's samples, accumulate at the last (1954-1988)
      if (!skip) Ls += L;
      ++sample;
      if (sample < a.spf && !skip) {
        st = PS_CAMERA;
      } else {
        if (!skip) {
          RT Lf = Ls / float(a.spf);
#ifdef DISPERSION_ENABLED
          vec3 colorXYZ = xyz * Lf;
#else
          vec3 colorXYZ = rgbToXyz(Lf);
#endif
          float oldN = accum.w, newN = oldN + 1.0f;
          accum = vec4((oldN * vec3(accum) + colorXYZ) / newN, newN);
        }
        if (++frame < a.frames) {
          sample = 0;
          st = PS_CAMERA;
        } else {
          st4(a.radiance, pix, accum);
          st4(a.rngstate, pix, rnd);
          addTileCost();
          st = PS_IDLE;
        }
      }
    }

    if (st == PS_CAMERA) {
      // pathtrace() up to the first traceRay (1934-1970)
      if (sample == 0) {
        if (frame == 0) {
          rnd = ld4(a.rngstate, pix);
          accum = ld4(a.radiance, pix);
        }
        skip = false;
        if constexpr (SC::kInteractive) {
          if (snl_rand(rnd) < P.skipProbability) skip = true;
        }
        Ls = rt1(0.0f);
        if (!skip) {
#ifdef DISPERSION_ENABLED
          float xi = 0.1f + 0.8f * (0.5f + floor(float(P.wavelengthSamples) * snl_rand(rnd))) / float(P.wavelengthSamples);
          float w = texture(P.ICDF, vec2(xi, 0.5f)).r;
          wl = 390.0f + (750.0f - 390.0f) * w;
          xyz = vec3(texture(P.WavelengthToXYZ, vec2(w, 0.5f)));
          rgb = xyzToRgb(xyz);
#else
          wl = 390.0f + (750.0f - 390.0f) * 0.5f;
          xyz = vec3(0.0f);
          rgb = vec3(0.0f);
#endif
        }
      }
      if (skip) {
        // a skipped frame leaves the pixel's mean as it is (1994-2004)
        if (++frame < a.frames) {
          sample = 0;
        } else {
          st4(a.radiance, pix, accum);
          st4(a.rngstate, pix, rnd);
          addTileCost();
          st = PS_IDLE;
        }
      } else {
        vec2 pixel(float(pix % a.W) + 0.5f, float((pix % a.imagePix) / a.W) + 0.5f);
        float jx = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
        float jy = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
        vec2 pixelj = pixel + vec2(jx, jy);
        vec3 primaryStart, primaryDir;
        if constexpr (SC::kHasCustomCamera) S.CONSTRUCT_PRIMARY_RAY(pixelj, rnd, primaryStart, primaryDir);
        else pt.constructPrimaryRay(pixelj, rnd, primaryStart, primaryDir, true);
        inDiele = false;
        if constexpr (SC::kHasDielectric) inDiele = S.SDF_DIELECTRIC(primaryStart) < 0.0f;
        o = primaryStart;
        tp = rt1(1.0f);
        L = rt1(0.0f);
        misSky = misSun = misSph = 1.0f;
        vertex = 0;
        neeMask = 0;
        contPending = false;
        startRay(primaryDir, P.maxLengthScale, false);
      }
    }
  }
  unsigned long long tot = stepTotal, ntot = normalTotal;
  for (int off = 32; off > 0; off >>= 1) {
    tot += __shfl_xor(tot, off);
    ntot += __shfl_xor(ntot, off);
  }
  if (lane == 0 && tot) atomicAdd(reinterpret_cast<unsigned long long*>(&a.cnt[C_TRACE_STEPS]), tot);
  if (lane == 0 && ntot) atomicAdd(reinterpret_cast<unsigned long long*>(&a.cnt[C_NORMALS]), ntot);
}

#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_persist(SnlKArgs a) { persist_body<SnlScene>(a); }
#endif

// zero the tile's queue counter before a persistent launch
extern "C" __global__ void snl_persist_reset(SnlKArgs a) {
  if (threadIdx.x == 0) a.cnt[C_FETCH] = 0;
}

}  // namespace glsl
