// snl_params.h -- the built-in uniform block of Snelly's shader programs.
//
// One struct holds every uniform that the reference's pathtracer, normals,
// pick and tonemapper fragment shaders declare
// (shaders/pathtracer-fragment-shader.glsl:4-99, normals-fragment-shader.glsl:3-44,
//  pick-fragment-shader.glsl:5-18, tonemapper-fragment-shader.glsl:3-12), so a
// program's uniforms can be set by name exactly as GLU.Shader.uniform* does
// (js/gl.js:255-408).  The same header is compiled into the host runtime
// (which uses the X-macro table for name -> offset reflection) and into every
// hipRTC program (which reads the block through a const pointer).
//
// Layout: plain glsl types (4-byte aligned floats/ints, vec3 = 12 bytes),
// samplers (8-byte aligned descriptors) first.  Booleans are stored as int,
// as uniform1i sets them.
#pragma once
#include "glsl.h"

// type codes shared with the C ABI (include/snelly.h: SNL_T_*)
#define SNL_T_FLOAT 1
#define SNL_T_VEC2 2
#define SNL_T_VEC3 3
#define SNL_T_VEC4 4
#define SNL_T_INT 5
#define SNL_T_MAT2 6
#define SNL_T_MAT3 7
#define SNL_T_MAT4 8
#define SNL_T_SAMPLER2D 9
#define SNL_T_SAMPLER3D 10

// X(type, type code, name)
#define SNL_BUILTIN_SAMPLERS(X)                                   \
  X(sampler2D, SNL_T_SAMPLER2D, WavelengthToXYZ)                  \
  X(sampler2D, SNL_T_SAMPLER2D, ICDF)                             \
  X(sampler2D, SNL_T_SAMPLER2D, iorTex)                           \
  X(sampler2D, SNL_T_SAMPLER2D, kTex)                             \
  X(sampler2D, SNL_T_SAMPLER2D, envMap)

#define SNL_BUILTIN_VALUES(X)                                     \
  X(vec2, SNL_T_VEC2, resolution)                                 \
  X(vec3, SNL_T_VEC3, camPos)                                     \
  X(vec3, SNL_T_VEC3, camDir)                                     \
  X(vec3, SNL_T_VEC3, camX)                                       \
  X(vec3, SNL_T_VEC3, camY)                                       \
  X(float, SNL_T_FLOAT, camFovy)                                  \
  X(float, SNL_T_FLOAT, camAspect)                                \
  X(float, SNL_T_FLOAT, camAperture)                              \
  X(float, SNL_T_FLOAT, camFocalDistance)                         \
  X(float, SNL_T_FLOAT, filterRadius)                             \
  X(float, SNL_T_FLOAT, radianceClamp)                            \
  X(float, SNL_T_FLOAT, skipProbability)                          \
  X(float, SNL_T_FLOAT, shadowStrength)                           \
  X(int, SNL_T_INT, maxStepsIsMiss)                               \
  X(int, SNL_T_INT, wavelengthSamples)                            \
  X(float, SNL_T_FLOAT, lengthScale)                              \
  X(float, SNL_T_FLOAT, minLengthScale)                           \
  X(float, SNL_T_FLOAT, maxLengthScale)                           \
  X(int, SNL_T_INT, haveEnvMap)                                   \
  X(int, SNL_T_INT, envMapVisible)                                \
  X(float, SNL_T_FLOAT, envMapPhiRotation)                        \
  X(float, SNL_T_FLOAT, envMapThetaRotation)                      \
  X(float, SNL_T_FLOAT, envMapTransitionAngle)                    \
  X(float, SNL_T_FLOAT, envMapRotation)                           \
  X(float, SNL_T_FLOAT, skyPower)                                 \
  X(vec3, SNL_T_VEC3, skyTintUp)                                  \
  X(vec3, SNL_T_VEC3, skyTintDown)                                \
  X(float, SNL_T_FLOAT, sunPower)                                 \
  X(float, SNL_T_FLOAT, sunAngularSize)                           \
  X(float, SNL_T_FLOAT, sunLatitude)                              \
  X(float, SNL_T_FLOAT, sunLongitude)                             \
  X(vec3, SNL_T_VEC3, sunColor)                                   \
  X(vec3, SNL_T_VEC3, sunDir)                                     \
  X(int, SNL_T_INT, sunVisibleDirectly)                           \
  X(vec3, SNL_T_VEC3, sphereLightPosition)                        \
  X(float, SNL_T_FLOAT, sphereLightRadius)                        \
  X(float, SNL_T_FLOAT, sphereLightPower)                         \
  X(vec3, SNL_T_VEC3, sphereLightColor)                           \
  X(float, SNL_T_FLOAT, metalRoughness)                           \
  X(vec3, SNL_T_VEC3, metalSpecAlbedoRGB)                         \
  X(float, SNL_T_FLOAT, dieleRoughness)                           \
  X(vec3, SNL_T_VEC3, dieleAbsorptionRGB)                         \
  X(vec3, SNL_T_VEC3, dieleSpecAlbedoRGB)                         \
  X(vec3, SNL_T_VEC3, surfaceDiffuseAlbedoRGB)                    \
  X(vec3, SNL_T_VEC3, surfaceSpecAlbedoRGB)                       \
  X(float, SNL_T_FLOAT, surfaceRoughness)                         \
  X(float, SNL_T_FLOAT, surfaceIor)                               \
  X(float, SNL_T_FLOAT, subsurface)                               \
  X(vec3, SNL_T_VEC3, subsurfaceAlbedoRGB)                        \
  X(float, SNL_T_FLOAT, subsurfaceMFP)                            \
  X(float, SNL_T_FLOAT, subsurfaceAnisotropy)                     \
  X(float, SNL_T_FLOAT, subsurfaceDiffuseWeight)                  \
  X(int, SNL_T_INT, maxSSSSteps)                                  \
  X(vec3, SNL_T_VEC3, atmosphereBoundsMin)                        \
  X(vec3, SNL_T_VEC3, atmosphereBoundsMax)                        \
  X(vec3, SNL_T_VEC3, atmosphereExtinctionCoeffRGB)               \
  X(vec3, SNL_T_VEC3, atmosphereScatteringCoeffRGB)               \
  X(float, SNL_T_FLOAT, atmosphereMFP)                            \
  X(float, SNL_T_FLOAT, atmosphereAnisotropy)                     \
  X(int, SNL_T_INT, fogEnable)                                    \
  X(float, SNL_T_FLOAT, fogMaxOpticalDepth)                       \
  X(float, SNL_T_FLOAT, fogMFP)                                   \
  X(vec3, SNL_T_VEC3, fogEmission)                                \
  X(vec3, SNL_T_VEC3, fogTint)                                    \
  X(float, SNL_T_FLOAT, volumeEmission)                           \
  X(vec3, SNL_T_VEC3, volumeEmissionColorRGB)                     \
  X(float, SNL_T_FLOAT, volumeExtinction)                         \
  X(vec2, SNL_T_VEC2, mousePick)                                  \
  X(float, SNL_T_FLOAT, exposure)                                 \
  X(float, SNL_T_FLOAT, invGamma)                                 \
  X(float, SNL_T_FLOAT, contrast)                                 \
  X(float, SNL_T_FLOAT, saturation)                               \
  X(float, SNL_T_FLOAT, hueShift)

namespace glsl {
struct SnlParams {
#define SNL_DECL(T, C, N) T N;
  SNL_BUILTIN_SAMPLERS(SNL_DECL)
  SNL_BUILTIN_VALUES(SNL_DECL)
#undef SNL_DECL
  // Derived by the runtime at upload, not a uniform: the camera's frustum
  // height tan(0.5 radians(camFovy)), correctly rounded (through double).  The
  // reference's shaders compute it per fragment (pathtracer-fragment-shader.glsl:1607,
  // normals-fragment-shader.glsl:276, pick-fragment-shader.glsl:126,
  // ao-fragment-shader.glsl:378, simplepathtracer-fragment-shader.glsl:874);
  // it scales every primary ray, so one ulp of it moves every first hit.
  float camFrustumHeight;
};
}  // namespace glsl
