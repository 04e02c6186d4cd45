// snl_kernels.h -- the gfx950 kernels of one compiled Snelly program.
// Included by the hipRTC source after the translated scene (struct SnlScene).
// Launch shapes and their order per frame are driven by runtime.cpp.
#pragma once

namespace glsl {

// Kernel bodies are templates on the scene type so that `if constexpr` on the
// scene's feature flags discards the calls into hooks a scene does not define.
#define SNL_SCENE_SETUP                                               \
  const SnlParams& P = *a.P;                                          \
  SC S(a.P, (const typename SC::User*)a.U);                           \
  S.INIT();                                                           \
  PT<SC> pt(S, P);                                                    \
  (void)pt;

SNL_HD float sample_jitter(float xi) {  // triangle filter (1928-1932)
  return xi < 0.5f ? sqrt(2.0f * xi) - 1.0f : 1.0f - sqrt(2.0f - 2.0f * xi);
}

// The queue item of entry j of a launch's range: with K > 1 tiles launched
// side by side, tile k takes the 64-item stripes k, k + K, k + 2K, ..., so the
// tiles hold alike mixes of cheap and expensive pixels.
__device__ inline int snl_tile_item(const SnlKArgs& a, int j) {
  if (a.tileCount <= 1) return a.itemBase + j;
  return ((j >> 6) * a.tileCount + a.tileIndex) * 64 + (j & 63);
}

// The pixel a queue item stands for.  Where the image divides into 8x8 tiles,
// 64 consecutive items cover one tile, so a wave's lanes start neighbouring
// rays; otherwise items are pixels in row order.  Chains of a multi-chain
// state follow one another.  With a tile order, queue position t takes tile
// tileOrder[t] (a render's later launches take the costliest tiles first).
__device__ inline int snl_item_pixel(const SnlKArgs& a, int item) {
  if ((a.W & 7) || (a.H & 7)) return item;
  int chain = item / a.imagePix, i = item - chain * a.imagePix;
  int tile = i >> 6, j = i & 63, tw = a.W >> 3;
  if (a.tileOrder) tile = a.tileOrder[tile];
  int tr = tile / tw, tc = tile - tr * tw;
  return chain * a.imagePix + (tr * 8 + (j >> 3)) * a.W + tc * 8 + (j & 7);
}

__device__ inline vec4 ld4(const float* p, long long i) {
  const float4 v = reinterpret_cast<const float4*>(p)[i];
  return vec4(v.x, v.y, v.z, v.w);
}
__device__ inline void st4(float* p, long long i, const vec4& v) {
  reinterpret_cast<float4*>(p)[i] = make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline vec4 ldRng(const SnlKArgs& a, int i) { return vec4(SNL_PF(F_RNG, i), SNL_PF(F_RNG + 1, i), SNL_PF(F_RNG + 2, i), SNL_PF(F_RNG + 3, i)); }
__device__ inline void stRng(const SnlKArgs& a, int i, const vec4& r) {
  SNL_PF(F_RNG, i) = r.x; SNL_PF(F_RNG + 1, i) = r.y; SNL_PF(F_RNG + 2, i) = r.z; SNL_PF(F_RNG + 3, i) = r.w;
}

// ---------------------------------------------------------------- counters
// vertex == -2 also clears the 64-bit step counter kept across a render call
extern "C" __global__ void snl_counters_reset(SnlKArgs a) {
  int k = threadIdx.x;
  bool stat = k == C_TRACE_STEPS || k == C_TRACE_STEPS + 1 || k == C_NORMALS || k == C_NORMALS + 1;
  if (k < C_COUNT && (a.vertex == -2 || !stat)) a.cnt[k] = 0;
}
// the output queues of one stage become the input queues of the next
extern "C" __global__ void snl_counters_rotate(SnlKArgs a) {
  if (threadIdx.x == 0) {
    // vertex < 0: the loop is over, continuation rays are dropped
    a.cnt[C_TRACE_IN] = a.vertex < 0 ? 0 : a.cnt[C_TRACE_OUT];
    a.cnt[C_NEE_IN] = a.cnt[C_NEE_OUT];
    a.cnt[C_TRACE_OUT] = 0;
    a.cnt[C_NEE_OUT] = 0;
    for (int k = 0; k < 4; ++k) a.cnt[C_MAT + k] = 0;
    a.cnt[C_FETCH] = 0;
  }
}

// ---------------------------------------------------------------- camera
// pathtrace() up to cameraPath's first traceRay (1934-1970, main 1991-2007).
template <class SC>
__device__ inline void camera_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool alive = false;
  if (i < a.npix) {
    vec4 rnd;
    int flags = 0;
    if (a.sample == 0) {
      rnd = ld4(a.rngstate, i);
#ifdef INTERACTIVE_MODE
      if (snl_rand(rnd) < P.skipProbability) flags |= PF_SKIP;
#endif
      if (!(flags & PF_SKIP)) {
#ifdef DISPERSION_ENABLED
        float xi = 0.1f + 0.8f * (0.5f + floor(float(P.wavelengthSamples) * snl_rand(rnd))) / float(P.wavelengthSamples);
        float w = texture(P.ICDF, vec2(xi, 0.5f)).r;
        float wavelength_nm = 390.0f + (750.0f - 390.0f) * w;
        vec3 xyz = vec3(texture(P.WavelengthToXYZ, vec2(w, 0.5f)));
        vec3 rgb = xyzToRgb(xyz);
#else
        float wavelength_nm = 390.0f + (750.0f - 390.0f) * 0.5f;
        vec3 xyz(0.0f), rgb(0.0f);
#endif
        SNL_PF(F_WL, i) = wavelength_nm;
        st3(a, F_XYZ, i, xyz);
        st3(a, F_RGB, i, rgb);
        stRT(a, F_LS, i, rt1(0.0f));
      }
    } else {
      rnd = ldRng(a, i);
      flags = SNL_PI(I_FLAGS, i) & PF_SKIP;
    }
    if (!(flags & PF_SKIP)) {
      vec2 pixel(float(i % a.W) + 0.5f, float((i % a.imagePix) / a.W) + 0.5f);
      float jx = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
      float jy = 0.5f * P.filterRadius * sample_jitter(snl_rand(rnd));
      vec2 pixelj = pixel + vec2(jx, jy);
      vec3 primaryStart, primaryDir;
      if constexpr (SC::kHasCustomCamera) S.CONSTRUCT_PRIMARY_RAY(pixelj, rnd, primaryStart, primaryDir);
      else pt.constructPrimaryRay(pixelj, rnd, primaryStart, primaryDir, true);
      bool inDiele = false;
      if constexpr (SC::kHasDielectric) inDiele = S.SDF_DIELECTRIC(primaryStart) < 0.0f;
      st3(a, F_OX, i, primaryStart);
      st3(a, F_DX, i, primaryDir);
      stRT(a, F_TP, i, rt1(1.0f));
      stRT(a, F_L, i, rt1(0.0f));
      SNL_PF(F_MIS, i) = 1.0f; SNL_PF(F_MIS + 1, i) = 1.0f; SNL_PF(F_MIS + 2, i) = 1.0f;
      flags |= inDiele ? PF_INDIELE : 0;   // vertex 0, no atmosphere scatters
      alive = true;
    }
    stRng(a, i, rnd);
    SNL_PI(I_FLAGS, i) = flags;
    SNL_PI(I_NEE, i) = 0;
  }
  snl_append_staged(&a.cnt[C_TRACE_OUT], a.qtrace_out, alive, i);
}
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_camera(SnlKArgs a) { camera_body<SnlScene>(a); }
#endif

// ---------------------------------------------------------------- trace
// Persistent sphere tracer.  Jobs [0, nT) are continuation rays (the path's
// current o, d with tmax = maxLengthScale); jobs [nT, nT+nN) are NEE jobs of
// up to three shadow rays each.  A wave owns a contiguous range of job indices
// grabbed with one atomic per `chunk` jobs; a lane whose ray terminates takes
// the next job of the range at once, so lanes stay busy until the queue drains
// (sphere-trace step counts vary by 10-100x between neighbouring pixels).
template <class SC>
__device__ inline void trace_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  const int nT = a.cnt[C_TRACE_IN];
  const int nN = a.cnt[C_NEE_IN];
  const int total = nT + nN;
  const int lane = snl_lane();
  const unsigned long long below = (1ull << lane) - 1ull;
  const float minMarch = P.minLengthScale;
  (void)minMarch;

  int rb = 0, re = 0;          // wave-uniform job range
  bool exhausted = false;      // wave-uniform
  int path = -1;               // path of the current job, -1 = idle
  bool nee = false;
  int mask = 0, light = 0;
  vec3 o, d;
  float tmax = 0.0f, t = 0.0f, sgn = 0.0f;
  int steps = 0;
  bool init = false;
  RT acc = rt1(0.0f);
  long long stepTotal = 0;

  while (true) {
    unsigned long long idleMask = __ballot(path < 0);
    if (idleMask != 0ull && rb >= re && !exhausted) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&a.cnt[C_FETCH], a.chunk);
      base = __shfl(base, 0);
      if (base >= total) exhausted = true;
      else { rb = base; re = min(base + a.chunk, total); }
    }
    if (idleMask != 0ull && rb < re) {
      int avail = re - rb;
      int rank = __popcll(idleMask & below);
      if (path < 0 && rank < avail) {
        int job = rb + rank;
        if (job < nT) {
          path = a.qtrace_in[job];
          nee = false;
          o = ld3(a, F_OX, path);
          d = ld3(a, F_DX, path);
          tmax = P.maxLengthScale;
        } else {
          path = a.qnee_in[job - nT];
          nee = true;
          mask = SNL_PI(I_NEE, path);
          light = __ffs(mask) - 1;
          o = ld3(a, F_OX, path);
          d = ld3(a, F_NEE + light * F_NEE_STRIDE, path);
          tmax = SNL_PF(F_NEE + light * F_NEE_STRIDE + 3, path);
          acc = rt1(0.0f);
        }
        init = true;
        steps = 0;
      }
      rb += min(avail, (int)__popcll(idleMask));
    }
    if (__ballot(path >= 0) == 0ull) {
      if (exhausted) break;
      continue;
    }
    if (path >= 0) {
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
        done = true; hit = true;
      } else {
        t += sgn * sdf;
        ++steps;
        if (abs(t) >= tmax) done = true;
        else if (steps >= SNL_MAX_MARCH_STEPS) { done = true; hit = !P.maxStepsIsMiss; m = -2; p = o + t * d; }
      }
      if (done) {
        if (!nee) {
          // continuation ray: record the hit for the shade stage
          SNL_PI(I_HMAT, path) = hit ? m : MAT_INVAL;
          if (hit) st3(a, F_HX, path, p);
          path = -1;
        } else {
          // one shadow ray of directSurfaceLighting (1236-1245): fold it in
          const int b = F_NEE + light * F_NEE_STRIDE;
          RT Tr = rt1(0.0f);
          if (!hit) {
            vec3 rgb = ld3(a, F_RGB, path);
            Tr = pt.transmittanceOverFreeSegment(o, d, tmax, rgb);
          }
          RT Li = ldRT(a, b + 7, path);
          Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
          if (averageComponent(Li) > RADIANCE_EPSILON) {
            RT f = ldRT(a, b + 4, path);
            acc += f * Li / SNL_PF(b + 10, path) * SNL_PF(b + 11, path) * SNL_PF(b + 12, path);
          }
          mask &= ~(1 << light);
          if (mask) {
            light = __ffs(mask) - 1;
            d = ld3(a, F_NEE + light * F_NEE_STRIDE, path);
            tmax = SNL_PF(F_NEE + light * F_NEE_STRIDE + 3, path);
            init = true;
            steps = 0;
          } else {
            RT Ldirect = min(rt1(P.radianceClamp), acc);
            RT L = ldRT(a, F_L, path);
            L += ldRT(a, F_NTP, path) * Ldirect;
            stRT(a, F_L, path, L);
            SNL_PI(I_NEE, path) = 0;
            path = -1;
          }
        }
      }
    }
  }
  // per-wave step count for the roofline accounting
  unsigned long long tot = stepTotal;
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
  if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&a.cnt[C_TRACE_STEPS]), tot);
}
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_trace(SnlKArgs a) { trace_body<SnlScene>(a); }
#endif

// ---------------------------------------------------------------- bin by material
// queue 0 = miss (or an invalid max-step hit), 1 surface, 2 metal, 3 dielectric
extern "C" __global__ void __launch_bounds__(256) snl_bin(SnlKArgs a) {
  const int n = a.cnt[C_TRACE_IN];
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    int j = base + threadIdx.x;
    bool act = j < n;
    int path = act ? a.qtrace_in[j] : 0;
    int m = act ? SNL_PI(I_HMAT, path) : -3;
    int q = m == MAT_SURFA ? 1 : (m == MAT_METAL ? 2 : (m == MAT_DIELE ? 3 : 0));
    for (int k = 0; k < 4; ++k) {
      bool mine = act && q == k;
      snl_append_staged(&a.cnt[C_MAT + k], a.qmat + (long long)k * a.cap, mine, path);
    }
  }
}

// ---------------------------------------------------------------- shade: miss
// cameraPath's miss branch (1657-1712): sphere light, sky, sun; the path ends.
template <class SC>
__device__ inline void shade_miss_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  const int n = a.cnt[C_MAT + 0];
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    int j = base + threadIdx.x;
    if (j >= n) continue;
    int i = a.qmat[j];
    if (SNL_PI(I_HMAT, i) != MAT_INVAL) continue;  // a max-step "hit" with no material: the path ends
    vec3 pW = ld3(a, F_OX, i), rayDir = ld3(a, F_DX, i), rgb = ld3(a, F_RGB, i);
    RT throughput = ldRT(a, F_TP, i), L = ldRT(a, F_L, i);
    int flags = SNL_PI(I_FLAGS, i);
    int vertex = (flags >> PF_VSHIFT) & 0xff;
    bool lightHit = false;
#ifdef HAS_SPHERE_LIGHT
    {
      vec3 pLightHit;
      RT LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? rt1(0.0f) : pt.sphereLightRadiance(pW, rayDir, rgb, pLightHit);
      if (averageComponent(LiSph) > RADIANCE_EPSILON) {
        lightHit = true;
        float lightSegmentLength = length(pLightHit - pW);
        RT Tr = pt.transmittanceOverFreeSegment(pW, rayDir, lightSegmentLength, rgb);
        float scatter_prob = pt.atmosphericScatteringProbability(Tr, rgb);
        L += throughput * Tr * SNL_PF(F_MIS + 2, i) * LiSph / (1.0f - scatter_prob);
      }
    }
#endif
    if (!lightHit) {
      RT Tr = pt.transmittanceOverFreeSegment(pW, rayDir, P.maxLengthScale, rgb);
      float scatter_prob = pt.atmosphericScatteringProbability(Tr, rgb);
      if (averageComponent(Tr) > RADIANCE_EPSILON) {
        float misSky = SNL_PF(F_MIS, i), misSun = SNL_PF(F_MIS + 1, i);
        if (!(vertex == 0 && !P.envMapVisible) && misSky > 0.0f)
          L += throughput * Tr * misSky * pt.environmentRadiance(rayDir, rgb) / (1.0f - scatter_prob);
        if (!(vertex == 0 && !P.sunVisibleDirectly) && misSun > 0.0f)
          L += throughput * Tr * misSun * pt.sunRadiance(rayDir, rgb) / (1.0f - scatter_prob);
      }
    }
    stRT(a, F_L, i, L);
  }
}
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_shade_miss(SnlKArgs a) { shade_miss_body<SnlScene>(a); }
#endif

// ---------------------------------------------------------------- shade: surface hit
// cameraPath's hit branch (1714-1923) for the paths of one material.  Shadow
// rays of directSurfaceLighting are written as an NEE job for the next trace.
template <class SC, int MAT>
__device__ inline void shade_hit(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  const int n = a.cnt[C_MAT + (MAT == MAT_SURFA ? 1 : (MAT == MAT_METAL ? 2 : 3))];
  const int qoff = (MAT == MAT_SURFA ? 1 : (MAT == MAT_METAL ? 2 : 3));
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    int j = base + threadIdx.x;
    bool act = j < n;
    bool cont = false, hasNee = false;
    int i = 0;
    if (act) {
      i = a.qmat[(long long)qoff * a.cap + j];
      vec3 pW = ld3(a, F_OX, i), rayDir = ld3(a, F_DX, i), pW_next = ld3(a, F_HX, i);
      vec3 rgb = ld3(a, F_RGB, i);
      float wavelength_nm = SNL_PF(F_WL, i);
      RT throughput = ldRT(a, F_TP, i), L = ldRT(a, F_L, i);
      float misWeightSky = SNL_PF(F_MIS, i), misWeightSun = SNL_PF(F_MIS + 1, i), misWeightSph = SNL_PF(F_MIS + 2, i);
      int flags = SNL_PI(I_FLAGS, i);
      int vertex = (flags >> PF_VSHIFT) & 0xff;
      bool inDielectric = (flags & PF_INDIELE) != 0;
      vec4 rnd = ldRng(a, i);
      const int hitMaterial = MAT;

      float rayLength = length(pW_next - pW);
#ifdef HAS_SPHERE_LIGHT
      {
        vec3 pLightHit;
        RT LiSph = (P.sphereLightPower < RADIANCE_EPSILON) ? rt1(0.0f) : pt.sphereLightRadiance(pW, rayDir, rgb, pLightHit);
        if (averageComponent(LiSph) > RADIANCE_EPSILON) {
          float lightSegmentLength = length(pLightHit - pW);
          if (lightSegmentLength < rayLength) {
            RT Tr = pt.transmittanceOverFreeSegment(pW, rayDir, lightSegmentLength, rgb);
            L += throughput * Tr * misWeightSph * LiSph;
          }
        }
      }
#endif
      if constexpr (SC::kHasDielectric) {
        if (inDielectric) {
          RT absorption = pt.DIELECTRIC_ABSORPTION_EVAL(pW, rgb);
          throughput *= exp(-rayLength * absorption);
        }
      }
      pW = pW_next;
      vec3 nW = pt.normal(pW, hitMaterial, true);
      {
        unsigned long long nb = __ballot(1);
        if (snl_lane() == __ffsll((long long)nb) - 1)
          atomicAdd(reinterpret_cast<unsigned long long*>(&a.cnt[C_NORMALS]), (unsigned long long)__popcll(nb));
      }
      vec3 ngW = nW;
      Basis basis = makeBasis(nW);
      if constexpr (SC::kHasNormalMap) {
        nW = pt.perturbNormal(pW, basis, hitMaterial);
        if ((dot(nW, rayDir) > 0.0f) && (hitMaterial != MAT_DIELE)) nW = 2.0f * ngW * dot(ngW, nW) - nW;
        basis = makeBasis(nW);
      }
      float prob_sss = 0.0f;
      bool do_subsurface_walk = false;
      RT subsurfaceAlbedo = rt1(0.0f);
      if constexpr (MAT == MAT_SURFA) {
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
      bool alive = true;
      if (!do_subsurface_walk) {
        vec3 winputW = -rayDir;
        vec3 winputL = worldToLocal(winputW, basis);
        // WebGL zero-initialises the uninitialised locals woutputL and bsdfPdf
        // (sampleDielectric reads woutputL before writing it)
        vec3 woutputL(0.0f);
        float bsdfPdf = 0.0f;
        RT f = pt.sampleBsdf(pW, basis, winputL, hitMaterial, wavelength_nm, rgb, true, woutputL, bsdfPdf, rnd);
        vec3 woutputW = localToWorld(woutputL, basis);
        if constexpr (MAT == MAT_DIELE) {
          if (dot(winputW, ngW) * dot(woutputW, ngW) < 0.0f) inDielectric = !inDielectric;
        }
        if constexpr (SC::kHasVolumeEmission) L += throughput * pt.VOLUME_EMISSION_EVAL(pW, rgb);
        rayDir = woutputW;
        pW += ngW * sign(dot(rayDir, ngW)) * 3.0f * P.minLengthScale;

        float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
        if (!inDielectric) {
          // directSurfaceLighting (1221-1295) with the shadow rays deferred; a
          // light whose BSDF value is exactly zero adds exactly zero, its
          // shadow ray is not traced
          int nmask = 0;
          vec3 wL, wW;
          if (P.skyPower > RADIANCE_EPSILON) {
            RT Li = pt.sampleSkyAtSurface(basis, rgb, rnd, wL, wW, skyPdf);
            if (averageComponent(Li) > RADIANCE_EPSILON) {
              float lp = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, wL, hitMaterial, wavelength_nm, rgb, true));
              RT fl = pt.evaluateBsdf(pW, basis, winputL, wL, hitMaterial, wavelength_nm, rgb, true);
              const int b = F_NEE + 0 * F_NEE_STRIDE;
              st3(a, b, i, wW); SNL_PF(b + 3, i) = P.maxLengthScale;
              stRT(a, b + 4, i, fl); stRT(a, b + 7, i, Li);
              SNL_PF(b + 10, i) = max(PDF_EPSILON, skyPdf); SNL_PF(b + 11, i) = abs(dot(wW, basis.nW));
              SNL_PF(b + 12, i) = powerHeuristic(skyPdf, lp);
              if (anyNonzero(fl)) nmask |= 1;
            }
          }
          if (P.sunPower > RADIANCE_EPSILON) {
            RT Li = pt.sampleSunAtSurface(basis, rgb, rnd, wL, wW, sunPdf);
            if (averageComponent(Li) > RADIANCE_EPSILON) {
              float lp = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, wL, hitMaterial, wavelength_nm, rgb, true));
              RT fl = pt.evaluateBsdf(pW, basis, winputL, wL, hitMaterial, wavelength_nm, rgb, true);
              const int b = F_NEE + 1 * F_NEE_STRIDE;
              st3(a, b, i, wW); SNL_PF(b + 3, i) = P.maxLengthScale;
              stRT(a, b + 4, i, fl); stRT(a, b + 7, i, Li);
              SNL_PF(b + 10, i) = max(PDF_EPSILON, sunPdf); SNL_PF(b + 11, i) = abs(dot(wW, basis.nW));
              SNL_PF(b + 12, i) = powerHeuristic(sunPdf, lp);
              if (anyNonzero(fl)) nmask |= 2;
            }
          }
#ifdef HAS_SPHERE_LIGHT
          if ((P.sphereLightPower > RADIANCE_EPSILON) && (dot(pW - P.sphereLightPosition, basis.nW) <= P.sphereLightRadius)) {
            vec3 pHit;
            RT Li = pt.sampleSphereLight(rgb, rnd, pW, pHit, sphPdf);
            if (averageComponent(Li) > RADIANCE_EPSILON) {
              float hitDist = max(0.0f, length(pHit - pW) - 3.0f * P.minLengthScale);
              wW = normalize(pHit - pW);
              wL = worldToLocal(wW, basis);
              float lp = max(PDF_EPSILON, pt.pdfBsdf(pW, basis, winputL, wL, hitMaterial, wavelength_nm, rgb, true));
              RT fl = pt.evaluateBsdf(pW, basis, winputL, wL, hitMaterial, wavelength_nm, rgb, true);
              const int b = F_NEE + 2 * F_NEE_STRIDE;
              st3(a, b, i, wW); SNL_PF(b + 3, i) = hitDist;
              stRT(a, b + 4, i, fl); stRT(a, b + 7, i, Li);
              SNL_PF(b + 10, i) = max(PDF_EPSILON, sphPdf); SNL_PF(b + 11, i) = abs(dot(wW, basis.nW));
              SNL_PF(b + 12, i) = powerHeuristic(sphPdf, lp);
              if (anyNonzero(fl)) nmask |= 4;
            }
          }
#endif
          if (nmask) {
            SNL_PI(I_NEE, i) = nmask;
            stRT(a, F_NTP, i, throughput);
            hasNee = true;
          }
        }
        RT fOverPdf = min(rt1(P.radianceClamp), f / max(PDF_EPSILON, bsdfPdf));
        RT surface_throughput = fOverPdf * abs(dot(woutputW, nW)) / max(PDF_EPSILON, 1.0f - prob_sss);
        throughput *= surface_throughput;
        misWeightSky = powerHeuristic(bsdfPdf, skyPdf);
        misWeightSun = powerHeuristic(bsdfPdf, sunPdf);
        misWeightSph = powerHeuristic(bsdfPdf, sphPdf);
      } else {
        // sub-surface random walk (1866-1918) -- traced inline, rare
        if constexpr (MAT == MAT_SURFA) {
          alive = false;  // filled in by the walk below when it exits
          RT diffuseAlbedoEntry = pt.SURFACE_DIFFUSE_REFL_EVAL(pW, basis.nW, -rayDir, rgb);
          basis.nW = ngW;
          // randomwalk_SSS (1468-1518)
          vec3 pWalk = pW - 3.0f * P.minLengthScale * basis.nW;
          float pdfDir;
          vec3 dirwalkL = sampleHemisphereCosineWeighted(rnd, pdfDir);
          vec3 dirwalkW = -localToWorld(dirwalkL, basis);
          RT walk_throughput = rt1(1.0f);
          bool success = false;
          vec3 pExit;
          for (int k = 0; k < 2048; ++k) {
            if (k > P.maxSSSSteps) break;
            float walk_step = -log(snl_rand(rnd)) * P.subsurfaceMFP * P.lengthScale;
            vec3 pHit; int hm;
            if (pt.traceRay(pWalk, dirwalkW, pHit, hm, walk_step)) { pExit = pHit; success = true; break; }
            if (k > 4) {
              float continuation_prob = clamp(10.0f * maxComponent(walk_throughput), 0.0f, 1.0f);
              if (snl_rand(rnd) < 1.0f - continuation_prob) break;
              walk_throughput /= continuation_prob;
            }
            dirwalkW = samplePhaseFunction(dirwalkW, P.subsurfaceAnisotropy, rnd);
            pWalk += walk_step * dirwalkW;
            walk_throughput *= subsurfaceAlbedo;
          }
          if (success) {
            alive = true;
            pW = pExit;
            nW = pt.normal(pExit, MAT_SURFA, true);
            Basis basis_exit = makeBasis(nW);
            float bsdfPdf;
            vec3 dirExitL = sampleHemisphereCosineWeighted(rnd, bsdfPdf);
            vec3 woutputW = localToWorld(dirExitL, basis_exit);
            rayDir = woutputW;
            throughput *= walk_throughput;
            float skyPdf = 0.0f, sunPdf = 0.0f, sphPdf = 0.0f;
            RT diffuseAlbedoExit = pt.SURFACE_DIFFUSE_REFL_EVAL(pW, nW, -rayDir, rgb);
            RT fx = mix(rt1(1.0f), diffuseAlbedoExit, P.subsurfaceDiffuseWeight) / float(M_PI);
            if (!inDielectric) {
              // SSS_exit_radiance (1521-1593), shadow rays traced inline
              vec3 dPw = 3.0f * P.minLengthScale * basis_exit.nW;
              RT Ld = rt1(0.0f);
              vec3 wL, wW;
              if (P.skyPower > RADIANCE_EPSILON) {
                RT Li = pt.sampleSkyAtSurface(basis_exit, rgb, rnd, wL, wW, skyPdf);
                if (averageComponent(Li) > RADIANCE_EPSILON) {
                  vec3 ph; int hm;
                  RT Tr = pt.traceRay(pExit + dPw, wW, ph, hm, P.maxLengthScale) ? rt1(0.0f) : pt.transmittanceOverFreeSegment(pExit + dPw, wW, P.maxLengthScale, rgb);
                  Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
                  if (averageComponent(Li) > RADIANCE_EPSILON)
                    Ld += fx * Li / max(PDF_EPSILON, skyPdf) * abs(dot(wW, basis_exit.nW)) * powerHeuristic(skyPdf, pdfHemisphereCosineWeighted(wL));
                }
              }
              if (P.sunPower > RADIANCE_EPSILON) {
                RT Li = pt.sampleSunAtSurface(basis_exit, rgb, rnd, wL, wW, sunPdf);
                if (averageComponent(Li) > RADIANCE_EPSILON) {
                  vec3 ph; int hm;
                  RT Tr = pt.traceRay(pExit + dPw, wW, ph, hm, P.maxLengthScale) ? rt1(0.0f) : pt.transmittanceOverFreeSegment(pExit + dPw, wW, P.maxLengthScale, rgb);
                  Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
                  if (averageComponent(Li) > RADIANCE_EPSILON)
                    Ld += fx * Li / max(PDF_EPSILON, sunPdf) * abs(dot(wW, basis_exit.nW)) * powerHeuristic(sunPdf, pdfHemisphereCosineWeighted(wL));
                }
              }
#ifdef HAS_SPHERE_LIGHT
              if ((P.sphereLightPower > RADIANCE_EPSILON) && (dot(pExit - P.sphereLightPosition, basis_exit.nW) <= P.sphereLightRadius)) {
                vec3 pHit;
                RT Li = pt.sampleSphereLight(rgb, rnd, pExit, pHit, sphPdf);
                if (averageComponent(Li) > RADIANCE_EPSILON) {
                  float hitDist = max(0.0f, length(pHit - pExit) - 3.0f * P.minLengthScale);
                  wW = normalize(pHit - pExit);
                  wL = worldToLocal(wW, basis_exit);
                  vec3 ph; int hm;
                  RT Tr = pt.traceRay(pExit + dPw, wW, ph, hm, hitDist) ? rt1(0.0f) : pt.transmittanceOverFreeSegment(pExit + dPw, wW, hitDist, rgb);
                  Li *= abs(rt1(1.0f) - P.shadowStrength * (rt1(1.0f) - Tr));
                  if (averageComponent(Li) > RADIANCE_EPSILON)
                    Ld += fx * Li / max(PDF_EPSILON, sphPdf) * abs(dot(wW, basis_exit.nW)) * powerHeuristic(sphPdf, pdfHemisphereCosineWeighted(wL));
                }
              }
#endif
              L += throughput * min(rt1(P.radianceClamp), Ld);
            }
            misWeightSky = powerHeuristic(bsdfPdf, skyPdf);
            misWeightSun = powerHeuristic(bsdfPdf, sunPdf);
            misWeightSph = powerHeuristic(bsdfPdf, sphPdf);
            RT fOverPdf = min(rt1(P.radianceClamp), fx / max(PDF_EPSILON, bsdfPdf));
            throughput *= fOverPdf * abs(dot(woutputW, nW)) / max(PDF_EPSILON, prob_sss);
            pW += nW * sign(dot(rayDir, nW)) * 3.0f * P.minLengthScale;
          }
        }
      }
      if (alive && !(maxComponent(throughput) < THROUGHPUT_EPSILON) && vertex + 1 <= a.maxBounces) {
        cont = true;
        flags = (flags & ~(0xff << PF_VSHIFT)) | ((vertex + 1) << PF_VSHIFT);
      }
      flags = inDielectric ? (flags | PF_INDIELE) : (flags & ~PF_INDIELE);
      st3(a, F_OX, i, pW);
      st3(a, F_DX, i, rayDir);
      stRT(a, F_TP, i, throughput);
      stRT(a, F_L, i, L);
      SNL_PF(F_MIS, i) = misWeightSky; SNL_PF(F_MIS + 1, i) = misWeightSun; SNL_PF(F_MIS + 2, i) = misWeightSph;
      SNL_PI(I_FLAGS, i) = flags;
      stRng(a, i, rnd);
    }
    snl_append_staged(&a.cnt[C_TRACE_OUT], a.qtrace_out, cont, i);
    snl_append_staged(&a.cnt[C_NEE_OUT], a.qnee_out, hasNee, i);
  }
}

#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_shade_surface(SnlKArgs a) { if constexpr (SnlScene::kHasSurface) shade_hit<SnlScene, MAT_SURFA>(a); }
#endif
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_shade_metal(SnlKArgs a) { if constexpr (SnlScene::kHasMetal) shade_hit<SnlScene, MAT_METAL>(a); }
#endif
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_shade_dielectric(SnlKArgs a) { if constexpr (SnlScene::kHasDielectric) shade_hit<SnlScene, MAT_DIELE>(a); }
#endif

// ---------------------------------------------------------------- accumulate
// running mean of XYZ and RNG write-back (1972-1988); also sums the frame's
// samples when maxSamplesPerFrame > 1 (1954-1972).
extern "C" __global__ void __launch_bounds__(256) snl_accumulate(SnlKArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.npix) return;
  int flags = SNL_PI(I_FLAGS, i);
  bool skip = (flags & PF_SKIP) != 0;
  if (!skip) stRT(a, F_LS, i, ldRT(a, F_LS, i) + ldRT(a, F_L, i));
  if (a.sample != a.spf - 1) return;
  if (!skip) {
    RT L = ldRT(a, F_LS, i) / float(a.spf);
#ifdef DISPERSION_ENABLED
    vec3 colorXYZ = ld3(a, F_XYZ, i) * L;
#else
    vec3 colorXYZ = rgbToXyz(L);
#endif
    vec4 old = ld4(a.radiance, i);
    float oldN = old.w;
    float newN = oldN + 1.0f;
    vec3 newL = (oldN * vec3(old) + colorXYZ) / newN;
    st4(a.radiance, i, vec4(newL, newN));
  }
  st4(a.rngstate, i, ldRng(a, i));
}

// ---------------------------------------------------------------- first-hit normals program
// normals-fragment-shader.glsl:293-345
SNL_HD vec3 environmentRadianceXYZ_normals(const SnlParams& P, const vec3& dir) {
  float phi = atan(dir.x, dir.z) + float(M_PI) + float(M_PI) * P.envMapRotation / 180.0f;
  phi -= 2.0f * float(M_PI) * floor(phi / (2.0f * float(M_PI)));
  float theta = acos(dir.y);
  float u = phi / (2.0f * float(M_PI));
  float v = theta / float(M_PI);
  if (P.haveEnvMap) {
    vec3 RGB = vec3(texture(P.envMap, vec2(u, v)));
    RGB *= P.skyPower;
    return rgbToXyz(RGB);
  }
  return P.skyPower * vec3(1.0f);
}
template <class SC>
__device__ inline void normals_body(const SnlKArgs& a) {
  SNL_SCENE_SETUP
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.npix) return;
  vec4 rnd = ld4(a.rngstate, i);
  vec2 pixel(float(i % a.W) + 0.5f, float((i % a.imagePix) / a.W) + 0.5f);
  vec3 colorXYZ(0.0f);
  for (int n = 0; n < a.spf; ++n) {
    float r1 = snl_rand(rnd);
    float r2 = snl_rand(rnd);
    vec2 pixelj = pixel + (-0.5f + vec2(r1, r2));
    vec3 primaryStart, primaryDir;
    pt.constructPrimaryRay(pixelj, rnd, primaryStart, primaryDir, false);
    vec3 pW;
    int hitMaterial;
    bool hit = pt.traceDistance(primaryStart, primaryDir, P.maxLengthScale, pW, hitMaterial);
    if (hit) {
      vec3 nW = (hitMaterial == MAT_INVAL) ? vec3(0.0f) : pt.normal(pW, hitMaterial, false);
      if constexpr (SC::kHasSurfaceNM) {
        vec3 ngW = nW;
        Basis basis = makeBasisPlain(nW);
        if (hitMaterial == MAT_SURFA) nW = localToWorld(normalize(vec3(S.SURFACE_NORMAL_MAP(pW, basis.nW))), basis);
        if ((dot(nW, primaryDir) > 0.0f) && (hitMaterial != MAT_DIELE)) nW = 2.0f * ngW * dot(ngW, nW) - nW;
      }
      colorXYZ += rgbToXyz(0.5f * (nW + vec3(1.0f)));
    } else {
      if (P.envMapVisible) colorXYZ += environmentRadianceXYZ_normals(P, primaryDir);
    }
  }
  colorXYZ /= float(a.spf);
  vec4 old = ld4(a.radiance, i);
  float oldN = old.w;
  float newN = oldN + 1.0f;
  vec3 newL = (oldN * vec3(old) + colorXYZ) / newN;
  st4(a.radiance, i, vec4(newL, newN));
  st4(a.rngstate, i, rnd);
}
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_normals(SnlKArgs a) { normals_body<SnlScene>(a); }
#endif

// ---------------------------------------------------------------- pick / first-hit depth
// pick-fragment-shader.glsl:116-152 evaluated at a list of pixel positions
// (xy[2k], xy[2k+1]) in gl_FragCoord units; out[2k] = distance, out[2k+1] = material.
template <class SC>
__device__ inline void pick_body(const SnlKArgs& a, const float* xy, float* out, int n) {
  SNL_SCENE_SETUP
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  vec2 pixel(xy[2 * k], xy[2 * k + 1]);
  vec2 ndc = -1.0f + 2.0f * (pixel / P.resolution);
  float fh = P.camFrustumHeight;  // tan(0.5 * radians(camFovy)), formed once per upload
  float fw = P.camAspect * fh;
  vec3 s = -fw * ndc.x * P.camX + fh * ndc.y * P.camY;
  vec3 primaryDir = normalize(P.camDir + s);
  vec3 primaryStart = P.camPos;
  vec3 pW;
  int hitMaterial;
  bool hit = pt.traceDistance(primaryStart, primaryDir, P.maxLengthScale, pW, hitMaterial, true);
  float d = P.maxLengthScale;
  if (hit) d = length(pW - primaryStart);
  out[2 * k] = d;
  out[2 * k + 1] = float(hitMaterial);
}
#ifndef SNL_FUSED
extern "C" __global__ void __launch_bounds__(256) snl_pick(SnlKArgs a, const float* xy, float* out, int n) { pick_body<SnlScene>(a, xy, out, n); }
#endif

// ---------------------------------------------------------------- tonemap
// tonemapper-fragment-shader.glsl:36-97 (4-digit XYZ->RGB, gamma before
// exposure, L/(1+L), saturation, contrast, optional hue shift).
SNL_HD float toneMap(float L) { return L / (1.0f + L); }
SNL_HD vec3 rgb2hsv(const vec3& c) {
  vec4 K(0.0f, -1.0f / 3.0f, 2.0f / 3.0f, -1.0f);
  vec4 p = mix(vec4(c.b, c.g, K.w, K.z), vec4(c.g, c.b, K.x, K.y), step(c.b, c.g));
  vec4 q = mix(vec4(p.x, p.y, p.w, c.r), vec4(c.r, p.y, p.z, p.x), step(p.x, c.r));
  float d = q.x - min(q.w, q.y);
  float e = 1.0e-10f;
  return vec3(abs(q.z + (q.w - q.y) / (6.0f * d + e)), d / (q.x + e), q.x);
}
SNL_HD vec3 hsv2rgb(const vec3& c) {
  vec4 K(1.0f, 2.0f / 3.0f, 1.0f / 3.0f, 3.0f);
  vec3 p = abs(fract(vec3(c.x) + vec3(K.x, K.y, K.z)) * 6.0f - vec3(K.w));
  return c.z * mix(vec3(K.x), clamp(p - vec3(K.x), 0.0f, 1.0f), c.y);
}
extern "C" __global__ void __launch_bounds__(256) snl_tonemap(const SnlParams* Pp, const float* radiance, float* rgbOut,
                                                              unsigned char* rgba8, int npix) {
  const SnlParams& P = *Pp;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  float X = radiance[4 * i], Y = radiance[4 * i + 1], Z = radiance[4 * i + 2];
  vec3 RGB;
  RGB.r = 3.2406f * X - 1.5372f * Y - 0.4986f * Z;
  RGB.g = -0.9689f * X + 1.8758f * Y + 0.0415f * Z;
  RGB.b = 0.0557f * X - 0.2040f * Y + 1.0570f * Z;
  float delta = -min(0.0f, min(min(RGB.r, RGB.g), RGB.b));
  RGB.r += delta; RGB.g += delta; RGB.b += delta;
  RGB = pow(RGB, vec3(P.invGamma));
  RGB *= pow(2.0f, P.exposure);
  float R = toneMap(RGB.r), G = toneMap(RGB.g), B = toneMap(RGB.b);
  float mean = (R + G + B) / 3.0f;
  float dR = R - mean, dG = G - mean, dB = B - mean;
  R = mean + sign(dR) * pow(abs(dR), 1.0f / P.saturation);
  G = mean + sign(dG) * pow(abs(dG), 1.0f / P.saturation);
  B = mean + sign(dB) * pow(abs(dB), 1.0f / P.saturation);
  dR = R - 0.5f; dG = G - 0.5f; dB = B - 0.5f;
  R = 0.5f + sign(dR) * pow(abs(dR), 1.0f / P.contrast);
  G = 0.5f + sign(dG) * pow(abs(dG), 1.0f / P.contrast);
  B = 0.5f + sign(dB) * pow(abs(dB), 1.0f / P.contrast);
  vec3 C(R, G, B);
  if (P.hueShift > 0.0f) {
    vec3 hsv = rgb2hsv(C);
    hsv.r = mod(hsv.r + P.hueShift, 1.0f);
    C = hsv2rgb(hsv);
  }
  if (rgbOut) { rgbOut[3 * i] = C.r; rgbOut[3 * i + 1] = C.g; rgbOut[3 * i + 2] = C.b; }
  if (rgba8) {
    // GL float -> UNORM8 conversion: round(clamp(c, 0, 1) * 255)
    rgba8[4 * i] = (unsigned char)floor(clamp(C.r, 0.0f, 1.0f) * 255.0f + 0.5f);
    rgba8[4 * i + 1] = (unsigned char)floor(clamp(C.g, 0.0f, 1.0f) * 255.0f + 0.5f);
    rgba8[4 * i + 2] = (unsigned char)floor(clamp(C.b, 0.0f, 1.0f) * 255.0f + 0.5f);
    rgba8[4 * i + 3] = 255;
  }
}

}  // namespace glsl

