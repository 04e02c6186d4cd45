// snelly_napi.cpp -- Node-API (N-API v8) addon exposing libsnelly_hip.so to the
// JavaScript host in snelly_amd/js.  Each function is a thin wrapper of one
// C-ABI entry point (include/snelly.h); handles travel as napi externals and
// errors become JS exceptions carrying snl_last_error().
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/snelly.h"

namespace {

#define NAPI_CALL(env, call)                                          \
  do {                                                                \
    if ((call) != napi_ok) {                                          \
      napi_throw_error(env, nullptr, "N-API call failed: " #call);    \
      return nullptr;                                                 \
    }                                                                 \
  } while (0)

napi_value throwSnl(napi_env env, int rc, snl_ctx ctx, const std::string& extra = "") {
  std::string msg = "snelly error " + std::to_string(rc) + ": " + snl_last_error(ctx);
  if (!extra.empty()) msg += "\n" + extra;
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

struct Args {
  size_t argc = 12;
  napi_value argv[12];
  bool ok(napi_env env, napi_callback_info info) {
    return napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) == napi_ok;
  }
};

void* getPtr(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_external) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, v, &p);
  return p;
}
int getInt(napi_env env, napi_value v) {
  int32_t i = 0;
  napi_get_value_int32(env, v, &i);
  return i;
}
double getNum(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}
std::string getStr(napi_env env, napi_value v) {
  size_t n = 0;
  napi_get_value_string_utf8(env, v, nullptr, 0, &n);
  std::string s(n, '\0');
  napi_get_value_string_utf8(env, v, &s[0], n + 1, &n);
  return s;
}
// data pointer + element count of a TypedArray (nullptr for null/undefined)
void* typedData(napi_env env, napi_value v, size_t* count = nullptr) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return nullptr;
  napi_typedarray_type type;
  size_t length;
  void* data;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &off);
  if (count) *count = length;
  return data;
}
// the TypedArray's data when it has `type` elements and at least `need` of
// them; otherwise a JS TypeError / RangeError is pending and nullptr returns
void* checkedArray(napi_env env, napi_value v, napi_typedarray_type want, size_t need, const char* what) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, nullptr, (std::string(what) + ": expected a TypedArray").c_str());
    return nullptr;
  }
  napi_typedarray_type type;
  size_t length;
  void* data;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &off);
  if (type != want) {
    napi_throw_type_error(env, nullptr,
                          (std::string(what) + (want == napi_int32_array ? ": expected an Int32Array"
                                                                          : ": expected a Float32Array")).c_str());
    return nullptr;
  }
  if (length < need) {
    napi_throw_range_error(env, nullptr, (std::string(what) + ": " + std::to_string(length) + " elements, " +
                                          std::to_string(need) + " needed").c_str());
    return nullptr;
  }
  return data;
}
// float count of a state's radiance or RNG buffer (4 per pixel of every chain)
bool stateFloats(snl_state s, size_t* n) {
  int w = 0, h = 0, k = 0;
  if (snl_state_info(s, &w, &h, &k) != SNL_OK) return false;
  *n = (size_t)w * h * k * 4;
  return true;
}
napi_value ext(napi_env env, void* p) {
  napi_value v;
  napi_create_external(env, p, nullptr, nullptr, &v);
  return v;
}
napi_value num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}
napi_value undef(napi_env env) {
  napi_value v;
  napi_get_undefined(env, &v);
  return v;
}
napi_value floatArray(napi_env env, size_t n, float** data) {
  napi_value ab, arr;
  void* raw = nullptr;
  napi_create_arraybuffer(env, n * sizeof(float), &raw, &ab);
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr);
  *data = static_cast<float*>(raw);
  return arr;
}

napi_value CtxCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_ctx c = nullptr;
  int rc = snl_ctx_create(a.argc > 0 ? getInt(env, a.argv[0]) : 0, &c);
  if (rc) return throwSnl(env, rc, nullptr);
  return ext(env, c);
}
napi_value CtxDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_ctx_destroy(static_cast<snl_ctx>(getPtr(env, a.argv[0])));
  return undef(env);
}
napi_value Sync(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_ctx c = static_cast<snl_ctx>(getPtr(env, a.argv[0]));
  int rc = snl_sync(c);
  if (rc) return throwSnl(env, rc, c);
  return undef(env);
}
// programBuild(ctx|null, shader, ior, defines, maxMarch, maxBounces, maxAtm, spf, flags)
// with ctx null: compile and cache only (no GPU)
napi_value ProgramBuild(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info) || a.argc < 9) { napi_throw_error(env, nullptr, "programBuild: 9 arguments"); return nullptr; }
  snl_ctx c = static_cast<snl_ctx>(getPtr(env, a.argv[0]));
  std::string shader = getStr(env, a.argv[1]), ior = getStr(env, a.argv[2]), defines = getStr(env, a.argv[3]);
  int mm = getInt(env, a.argv[4]), mb = getInt(env, a.argv[5]), ma = getInt(env, a.argv[6]), spf = getInt(env, a.argv[7]);
  unsigned flags = (unsigned)getInt(env, a.argv[8]);
  std::vector<char> log(1 << 16, '\0');
  if (!c) {
    int rc = snl_program_precompile(shader.c_str(), ior.c_str(), defines.c_str(), mm, mb, ma, spf, flags, log.data(),
                                    (int)log.size());
    if (rc) return throwSnl(env, rc, nullptr, log.data());
    return undef(env);
  }
  snl_prog p = nullptr;
  int rc = snl_program_build(c, shader.c_str(), ior.c_str(), defines.c_str(), mm, mb, ma, spf, flags, &p, log.data(),
                             (int)log.size());
  if (rc) return throwSnl(env, rc, c, log.data());
  return ext(env, p);
}
napi_value ProgramDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_program_destroy(static_cast<snl_prog>(getPtr(env, a.argv[0])));
  return undef(env);
}
napi_value UniformIndex(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  return num(env, snl_uniform_index(static_cast<snl_prog>(getPtr(env, a.argv[0])), getStr(env, a.argv[1]).c_str()));
}
// uniformSet(prog, idx, type, typedArray, count)
napi_value UniformSet(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_prog p = static_cast<snl_prog>(getPtr(env, a.argv[0]));
  const int type = getInt(env, a.argv[2]), count = getInt(env, a.argv[4]);
  int per = 0;  // values of one element of `type` (include/snelly.h: SNL_T_*)
  switch (type) {
    case SNL_T_FLOAT: case SNL_T_INT: per = 1; break;
    case SNL_T_VEC2: per = 2; break;
    case SNL_T_VEC3: per = 3; break;
    case SNL_T_VEC4: case SNL_T_MAT2: per = 4; break;
    case SNL_T_MAT3: per = 9; break;
    case SNL_T_MAT4: per = 16; break;
    default: napi_throw_type_error(env, nullptr, "uniformSet: bad type"); return nullptr;
  }
  if (count <= 0) { napi_throw_range_error(env, nullptr, "uniformSet: count must be positive"); return nullptr; }
  void* data = checkedArray(env, a.argv[3], type == SNL_T_INT ? napi_int32_array : napi_float32_array,
                            (size_t)per * count, "uniformSet");
  if (!data) return nullptr;
  int rc = snl_uniform_set(p, getInt(env, a.argv[1]), type, data, count);
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
napi_value UniformTexture(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  int rc = snl_uniform_texture(static_cast<snl_prog>(getPtr(env, a.argv[0])), getInt(env, a.argv[1]),
                               static_cast<snl_tex>(getPtr(env, a.argv[2])));
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
// textureCreate(ctx, w, h, d, channels, linear, clamp, Float32Array|null)
napi_value TextureCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_ctx c = static_cast<snl_ctx>(getPtr(env, a.argv[0]));
  snl_tex t = nullptr;
  const int w = getInt(env, a.argv[1]), h = getInt(env, a.argv[2]), d = getInt(env, a.argv[3]), ch = getInt(env, a.argv[4]);
  const float* texels = nullptr;
  napi_valuetype vt;
  napi_typeof(env, a.argv[7], &vt);
  if (vt != napi_null && vt != napi_undefined) {
    if (w <= 0 || h <= 0 || d <= 0 || ch <= 0) { napi_throw_range_error(env, nullptr, "textureCreate: bad shape"); return nullptr; }
    texels = static_cast<const float*>(checkedArray(env, a.argv[7], napi_float32_array, (size_t)w * h * d * ch,
                                                    "textureCreate"));
    if (!texels) return nullptr;
  }
  int rc = snl_texture_create(c, w, h, d, ch, getInt(env, a.argv[5]), getInt(env, a.argv[6]), texels, &t);
  if (rc) return throwSnl(env, rc, c);
  return ext(env, t);
}
napi_value TextureUpload(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_tex t = static_cast<snl_tex>(getPtr(env, a.argv[0]));
  int w = 0, h = 0, d = 0, ch = 0;
  if (snl_texture_info(t, &w, &h, &d, &ch) != SNL_OK) return throwSnl(env, SNL_E_ARG, nullptr);
  const float* texels = static_cast<const float*>(checkedArray(env, a.argv[1], napi_float32_array,
                                                               (size_t)w * h * d * ch, "textureUpload"));
  if (!texels) return nullptr;
  int rc = snl_texture_upload(t, texels);
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
napi_value TextureDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_texture_destroy(static_cast<snl_tex>(getPtr(env, a.argv[0])));
  return undef(env);
}
napi_value StateCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_ctx c = static_cast<snl_ctx>(getPtr(env, a.argv[0]));
  snl_state s = nullptr;
  int rc = snl_state_create(c, getInt(env, a.argv[1]), getInt(env, a.argv[2]), &s);
  if (rc) return throwSnl(env, rc, c);
  return ext(env, s);
}
napi_value StateDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_state_destroy(static_cast<snl_state>(getPtr(env, a.argv[0])));
  return undef(env);
}
napi_value StateSeed(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  uint32_t seed = 0;
  napi_get_value_uint32(env, a.argv[1], &seed);
  int rc = snl_state_seed(static_cast<snl_state>(getPtr(env, a.argv[0])), seed);
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
napi_value StateClear(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  int rc = snl_state_clear(static_cast<snl_state>(getPtr(env, a.argv[0])));
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
// stateRead(state, which) -> Float32Array(4 * width * height * chains), sized
// from the state itself
napi_value StateRead(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_state st = static_cast<snl_state>(getPtr(env, a.argv[0]));
  size_t n = 0;
  if (!stateFloats(st, &n)) return throwSnl(env, SNL_E_ARG, nullptr);
  float* out = nullptr;
  napi_value arr = floatArray(env, n, &out);
  int rc = snl_state_read(st, getInt(env, a.argv[1]), out);
  if (rc) return throwSnl(env, rc, nullptr);
  return arr;
}
napi_value StateWrite(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_state st = static_cast<snl_state>(getPtr(env, a.argv[0]));
  size_t n = 0;
  if (!stateFloats(st, &n)) return throwSnl(env, SNL_E_ARG, nullptr);
  const float* in = static_cast<const float*>(checkedArray(env, a.argv[2], napi_float32_array, n, "stateWrite"));
  if (!in) return nullptr;
  int rc = snl_state_write(st, getInt(env, a.argv[1]), in);
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
napi_value Render(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  int rc = snl_render(static_cast<snl_prog>(getPtr(env, a.argv[0])), static_cast<snl_state>(getPtr(env, a.argv[1])),
                      getInt(env, a.argv[2]));
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
napi_value RenderNormals(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  int rc = snl_render_normals(static_cast<snl_prog>(getPtr(env, a.argv[0])),
                              static_cast<snl_state>(getPtr(env, a.argv[1])), getInt(env, a.argv[2]));
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
// pick(prog, Float32Array xy) -> Float32Array (distance, material) pairs
napi_value Pick(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  size_t n = 0;
  typedData(env, a.argv[1], &n);
  const float* xy = static_cast<const float*>(checkedArray(env, a.argv[1], napi_float32_array, n & ~(size_t)1, "pick"));
  if (!xy) return nullptr;
  n &= ~(size_t)1;
  float* out = nullptr;
  napi_value arr = floatArray(env, n, &out);
  int rc = snl_pick(static_cast<snl_prog>(getPtr(env, a.argv[0])), xy, (int)(n / 2), out);
  if (rc) return throwSnl(env, rc, nullptr);
  return arr;
}
// tonemap(prog, state) -> Float32Array rgb (3 per pixel of the state)
napi_value Tonemap(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  size_t n4 = 0;
  if (!stateFloats(static_cast<snl_state>(getPtr(env, a.argv[1])), &n4)) return throwSnl(env, SNL_E_ARG, nullptr);
  float* out = nullptr;
  napi_value arr = floatArray(env, n4 / 4 * 3, &out);
  int rc = snl_tonemap(static_cast<snl_prog>(getPtr(env, a.argv[0])), static_cast<snl_state>(getPtr(env, a.argv[1])),
                       out, nullptr);
  if (rc) return throwSnl(env, rc, nullptr);
  return arr;
}
// tonemapRGBA8(prog, state) -> Uint8Array rgba (4 per pixel of the state, GL UNORM8)
napi_value TonemapRGBA8(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  size_t n = 0;
  if (!stateFloats(static_cast<snl_state>(getPtr(env, a.argv[1])), &n)) return throwSnl(env, SNL_E_ARG, nullptr);
  napi_value ab, arr;
  void* raw = nullptr;
  napi_create_arraybuffer(env, n, &raw, &ab);
  napi_create_typedarray(env, napi_uint8_array, n, ab, 0, &arr);
  int rc = snl_tonemap(static_cast<snl_prog>(getPtr(env, a.argv[0])), static_cast<snl_state>(getPtr(env, a.argv[1])),
                       nullptr, static_cast<unsigned char*>(raw));
  if (rc) return throwSnl(env, rc, nullptr);
  return arr;
}
napi_value Stats(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  snl_stats s{};
  snl_get_stats(static_cast<snl_prog>(getPtr(env, a.argv[0])), &s);
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "ms_total", num(env, s.ms_total));
  napi_set_named_property(env, o, "trace_steps", num(env, (double)s.trace_steps));
  napi_set_named_property(env, o, "paths", num(env, (double)s.paths));
  napi_set_named_property(env, o, "normal_evals", num(env, (double)s.normal_evals));
  return o;
}
napi_value SetEngine(napi_env env, napi_callback_info info) {
  Args a;
  if (!a.ok(env, info)) return nullptr;
  int rc = snl_set_engine(static_cast<snl_prog>(getPtr(env, a.argv[0])), getInt(env, a.argv[1]),
                          a.argc > 2 ? getInt(env, a.argv[2]) : 0, a.argc > 3 ? getInt(env, a.argv[3]) : 0,
                          a.argc > 4 ? getInt(env, a.argv[4]) : 0, a.argc > 5 ? getInt(env, a.argv[5]) : 0);
  if (rc) return throwSnl(env, rc, nullptr);
  return undef(env);
}
napi_value Version(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_string_utf8(env, snl_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback cb; } fns[] = {
      {"ctxCreate", CtxCreate}, {"ctxDestroy", CtxDestroy}, {"sync", Sync}, {"programBuild", ProgramBuild},
      {"programDestroy", ProgramDestroy}, {"uniformIndex", UniformIndex}, {"uniformSet", UniformSet},
      {"uniformTexture", UniformTexture}, {"textureCreate", TextureCreate}, {"textureUpload", TextureUpload},
      {"textureDestroy", TextureDestroy}, {"stateCreate", StateCreate}, {"stateDestroy", StateDestroy},
      {"stateSeed", StateSeed}, {"stateClear", StateClear}, {"stateRead", StateRead}, {"stateWrite", StateWrite},
      {"render", Render}, {"renderNormals", RenderNormals}, {"pick", Pick}, {"tonemap", Tonemap}, {"tonemapRGBA8", TonemapRGBA8},
      {"stats", Stats}, {"setEngine", SetEngine}, {"version", Version}};
  for (auto& f : fns) {
    napi_value fn;
    NAPI_CALL(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_CALL(env, napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
