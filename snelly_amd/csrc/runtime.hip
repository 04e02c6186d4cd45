// runtime.hip -- libsnelly_hip.so: the C ABI (include/snelly.h) over HIP on gfx950.
//
// A program is one hipRTC module built from a scene snippet (translated by
// translate.cpp) plus the wavefront kernels (snl_pt.h, snl_kernels.h).  A state
// holds one image's accumulation + RNG buffers and the SoA path buffers.  A
// frame is a fixed launch sequence on the context's stream with no host
// synchronisation: queue sizes live in device counters that each kernel reads.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <elf.h>
#include <sys/stat.h>

#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/snelly.h"
#include "embedded_sources.h"
#include "snl_wavefront.h"
#include "translate.h"

using namespace glsl;

// ---------------------------------------------------------------- objects
#define SNL_MAX_TILES 4
struct FusedModule;
// One program as a part of a batch module: its translated scene, its #define
// lines, its __MAX_*__ limits and the build modes a module must share.
struct FusedPart {
  std::string sceneCode, defs;
  unsigned long long key = 0;  // hash of the scene, its defines and limits: the program's kind
  int maxMarch = 256, maxAtm = 1;
  bool fastDiv = false, precise = false, dispersion = false, interactive = false;
  bool fusable = false;  // a path-tracer program (not 'ao' / 'ptsimple')
  bool sameGroup(const FusedPart& o) const {
    return fastDiv == o.fastDiv && precise == o.precise && dispersion == o.dispersion && interactive == o.interactive;
  }
};
struct snl_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t tileStream[SNL_MAX_TILES] = {};  // default-priority streams for pipelined pixel tiles
  hipEvent_t tileEvent[SNL_MAX_TILES] = {};
  hipEvent_t joinEvent = nullptr;
  // batch streams (default priority), created on first use; their count is
  // SNL_BATCH_STREAMS (default 16).  Kernels on them run concurrently up to
  // the process's hardware queue count (GPU_MAX_HW_QUEUES).
  std::vector<hipStream_t> batchStream;
  std::vector<hipEvent_t> batchEvent;
  std::string err;
  int numCU = 256;
  int* scratch = nullptr;  // this device's copy of the allocation-free render's `scratch` array
  std::set<snl_batch_s*> batches;  // submitted batches not yet released
  std::map<unsigned long long, FusedModule*> fused;  // batch modules by source hash
  bool batchFusion = true;  // snl_ctx_set_batch_fusion (effective once a batch module is loaded)
  hipStream_t graphStream = nullptr;  // where batch graphs are captured from (created on first use)
};

struct snl_tex_s {
  snl_ctx ctx;
  float* d = nullptr;
  int w = 0, h = 0, depth = 1, ch = 4, flags = 0;
  std::set<snl_prog_s*> users;  // programs with a sampler uniform bound to this texture
};

struct snl_state_s {
  snl_ctx ctx;
  int w = 0, h = 0, chains = 1;
  long long npix = 0;  // w * h * chains
  float* radiance = nullptr;
  float* rng = nullptr;
  float* pf = nullptr;
  int* pi = nullptr;
  int* q = nullptr;    // 8 * cap ints: trace A/B, nee A/B, mat[4]
  int* cnt = nullptr;  // C_COUNT ints
  int* tileCnt = nullptr;  // SNL_MAX_TILES * C_COUNT ints (persistent engine)
  float* radPrev = nullptr;  // radiance as of the frame start, for samplers left on texture unit 0
  double* errPartial = nullptr;  // snl_state_error's block partials and result (allocated on first use)
  int* tileCost = nullptr;   // per 8x8 tile: march steps of the render's first frame, then the
                             // cost-bucket histogram and the tile order (ordered megakernel renders)
  snl_batch_s* busy = nullptr;  // the submitted batch whose job `busyJob` renders into this state
  int busyJob = -1;
};

struct snl_prog_s {
  snl_ctx ctx;
  hipModule_t mod = nullptr;
  hipFunction_t k_reset, k_rotate, k_camera, k_trace, k_bin, k_miss, k_surf, k_metal, k_diele, k_accum, k_normals,
      k_pick, k_tonemap, k_persist, k_persist_reset, k_mega, k_mode = nullptr;
  unsigned mode = 0;  // SNL_BUILD_AO or SNL_BUILD_PTSIMPLE, else 0 (the path tracer)
  int engine = SNL_ENGINE_MEGAKERNEL;
  bool orderTiles = true;  // queue-kernel renders take tiles costliest first (SNL_TILE_ORDER=0: queue order)
  int tiles = 1, persistBlocks = 0, shadeThresh = 32, persistChunk = 64, marchBurst = 4;
  bool hasSurface = false, hasMetal = false, hasDiele = false, hasAtmosphere = false;
  // programs only the megakernel renders: the homogeneous atmosphere (a vertex
  // traces up to five extra rays between random draws) and the heterogeneous
  // medium (delta and ratio tracking inside the vertex)
  bool megaOnly = false;
  int megaBlocks = 0, modeBlocks = 0;
  // the megakernel's two builds: the one the program renders with by default,
  // and, where four waves per SIMD spill, that four-wave build (its code object
  // kept on the host, loaded on first use by snl_set_mega_variant)
  hipFunction_t k_megaMain = nullptr, k_megaSpill = nullptr;
  std::vector<char> spillCode;
  hipModule_t spillMod = nullptr;
  int megaVariant = 0;
  int maxBounces = 3, maxMarch = 256, spf = 1;
  SnlParams hostP;
  std::vector<char> hostU;
  SnlParams* dP = nullptr;
  void* dU = nullptr;
  bool dirty = true;
  std::vector<snl::UniformInfo> user;
  int traceBlocks = 0, traceChunk = 64, shadeBlocks = 0;
  bool profiling = false;
  snl_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr, et0 = nullptr, et1 = nullptr;
  std::vector<hipEvent_t> traceEvents;  // start/stop pairs of the trace launches (profiling)
  // 2D samplers the scene code reads (uniform indices) and those the host has
  // bound; an unbound one samples texture unit 0, the radiance being read
  std::vector<int> sampled2D;
  std::set<int> bound;
  std::map<int, snl_tex_s*> texOf;  // sampler uniform -> the texture bound to it (null once destroyed)
  snl_batch_s* busy = nullptr;  // the submitted batch whose job `busyJob` runs this program
  int busyJob = -1;
  FusedPart part;  // what a batch module holding this program needs of it
};

// A batch module: the megakernel of several programs in one code object, one
// instantiation per distinct program (its "kind"), run by one persistent
// launch over all jobs (snl_mega_fused); built on first use, kept per context.
struct FusedModule {
  hipModule_t mod = nullptr;
  hipFunction_t k_run = nullptr, k_reset = nullptr;
  int blocks = 0;
  std::map<unsigned long long, int> kindOf;  // program kind (FusedPart::key) -> instantiation
  FusedPart group;                           // the build modes its programs share
};

// A submitted batch (snl_batch_submit): its jobs' launches are queued on the
// context's batch streams; each job's end is an event on its stream.
struct snl_batch_s {
  snl_ctx ctx = nullptr;
  int njobs = 0, pending = 0;
  std::vector<snl_prog> progs;
  std::vector<snl_state> states;
  std::vector<int> frames, stream;
  std::vector<char> unit0, ordered, done;
  std::vector<double> work;            // the job's estimated cost (orders jobs onto streams)
  std::vector<long long> steps, normals;  // the job's counters, read when it is seen complete
  std::vector<float> ms;               // submit to the job's end (hipEvents)
  std::vector<hipEvent_t> ev;          // per job: after its last launch
  hipEvent_t ev0 = nullptr;            // before the first launch
  // the jobs one batch module runs in one persistent launch (may be empty),
  // and that launch's arguments (host copy and device copy)
  FusedModule* fm = nullptr;
  std::vector<int> fusedJobs;
  std::vector<char> hFused;
  void* dFused = nullptr;
};

namespace {

thread_local std::string g_err;  // for calls without a context

int fail(snl_ctx ctx, int code, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  g_err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, SNL_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct BuiltinInfo {
  const char* name;
  int offset, type, size;
};
const std::vector<BuiltinInfo>& builtins() {
  static const std::vector<BuiltinInfo> v = {
#define SNL_X(T, C, N) {#N, (int)offsetof(SnlParams, N), C, (int)sizeof(T)},
      SNL_BUILTIN_SAMPLERS(SNL_X) SNL_BUILTIN_VALUES(SNL_X)
#undef SNL_X
  };
  return v;
}
const std::set<std::string>& builtinNames() {
  static std::set<std::string> s = [] {
    std::set<std::string> r;
    for (auto& b : builtins()) r.insert(b.name);
    return r;
  }();
  return s;
}

unsigned long long fnv1a(const std::string& s, unsigned long long h = 1469598103934665603ull) {
  for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
  return h;
}

std::string libDir() {
  Dl_info info;
  if (dladdr((void*)&fnv1a, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t s = p.rfind('/');
    if (s != std::string::npos) return p.substr(0, s);
  }
  return ".";
}
std::string cacheDir() {
  const char* e = getenv("SNL_KCACHE");
  std::string d = e && *e ? e : libDir() + "/_kcache";
  mkdir(d.c_str(), 0755);
  return d;
}

struct Compiled {
  std::vector<char> code;
  std::vector<char> spillCode;  // the four-wave megakernel build that spilled, when it did
  snl::TranslateResult tr;
  std::set<std::string> defines;
  std::string defs;  // the program's #define lines
  bool fastDiv = false;
  int maxAtm = 1;
};

// Compile `source` with `opts` through hipRTC, or fetch the code object from
// the on-disk cache (keyed by the source, the options and the embedded headers).
static int compileCached(snl_ctx ctx, const std::string& source, const std::vector<std::string>& opts,
                         std::vector<char>& code, char* log, int loglen) {
  std::string key = source;
  for (auto& o : opts) key += "\n" + o;
  for (int i = 0; i < kNumEmbedded; ++i) key += kEmbeddedSources[i];
  char name[64];
  snprintf(name, sizeof name, "%016llx.hsaco", fnv1a(key));
  std::string path = cacheDir() + "/" + name;
  {
    std::ifstream f(path, std::ios::binary);
    if (f) {
      code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
      if (!code.empty()) return SNL_OK;
    }
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source.c_str(), "snl_program.hip", kNumEmbedded, kEmbeddedSources,
                          kEmbeddedNames) != HIPRTC_SUCCESS)
    return fail(ctx, SNL_E_COMPILE, "hiprtcCreateProgram failed");
  std::vector<const char*> copts;
  for (auto& o : opts) copts.push_back(o.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  std::string plog(ls + 1, '\0');
  if (ls) hiprtcGetProgramLog(prog, &plog[0]);
  if (r != HIPRTC_SUCCESS) {
    if (log && loglen > 0) snprintf(log, loglen, "%s", plog.c_str());
    hiprtcDestroyProgram(&prog);
    std::string head = plog.substr(0, 1500);
    return fail(ctx, SNL_E_COMPILE, "hipRTC compile failed:\n%s", head.c_str());
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    f.write(code.data(), (std::streamsize)code.size());
  }
  rename(tmp.c_str(), path.c_str());
  return SNL_OK;
}

// Private-segment (scratch) bytes per lane of a kernel in an AMDGPU code
// object: the second field of its kernel descriptor, the symbol `<kernel>.kd`.
// -1 when the object or the symbol cannot be read.
static long kernelPrivateBytes(const std::vector<char>& code, const char* kdName) {
  if (code.size() < sizeof(Elf64_Ehdr)) return -1;
  const char* base = code.data();
  Elf64_Ehdr eh;
  memcpy(&eh, base, sizeof eh);
  if (memcmp(eh.e_ident, ELFMAG, SELFMAG) != 0 || eh.e_ident[EI_CLASS] != ELFCLASS64) return -1;
  if (eh.e_shoff + (size_t)eh.e_shnum * sizeof(Elf64_Shdr) > code.size()) return -1;
  std::vector<Elf64_Shdr> sh(eh.e_shnum);
  memcpy(sh.data(), base + eh.e_shoff, sh.size() * sizeof(Elf64_Shdr));
  for (const Elf64_Shdr& t : sh) {
    if ((t.sh_type != SHT_SYMTAB && t.sh_type != SHT_DYNSYM) || t.sh_link >= sh.size()) continue;
    const Elf64_Shdr& str = sh[t.sh_link];
    if (t.sh_offset + t.sh_size > code.size() || str.sh_offset + str.sh_size > code.size()) continue;
    size_t n = t.sh_size / sizeof(Elf64_Sym);
    for (size_t i = 0; i < n; ++i) {
      Elf64_Sym sym;
      memcpy(&sym, base + t.sh_offset + i * sizeof(Elf64_Sym), sizeof sym);
      if (sym.st_name >= str.sh_size || sym.st_shndx == SHN_UNDEF || sym.st_shndx >= sh.size()) continue;
      const char* nm = base + str.sh_offset + sym.st_name;
      if (strncmp(nm, kdName, str.sh_size - sym.st_name) != 0) continue;
      const Elf64_Shdr& sec = sh[sym.st_shndx];
      size_t off = sec.sh_offset + (sym.st_value - sec.sh_addr) + 4;
      if (sym.st_value < sec.sh_addr || off + 4 > code.size()) return -1;
      uint32_t priv;
      memcpy(&priv, base + off, 4);
      return (long)priv;
    }
  }
  return -1;
}

// An integer field (e.g. ".vgpr_spill_count") of a kernel's entry in the
// AMDGPU metadata note of a code object: the msgpack map that holds
// ".name" = kernel lists its keys in order, so the field's first occurrence
// after the name is the kernel's.  -1 when not found.
static long kernelMetaInt(const std::vector<char>& code, const std::string& kernel, const std::string& key) {
  auto packed = [](const std::string& t) {  // a msgpack string of fewer than 32 bytes
    std::string m(1, (char)(0xa0 | (int)t.size()));
    return m + t;
  };
  const std::string bytes(code.begin(), code.end());
  const std::string name = packed(".name") + packed(kernel);
  size_t at = bytes.find(name);
  if (at == std::string::npos || key.size() >= 32) return -1;
  size_t k = bytes.find(packed(key), at);
  if (k == std::string::npos) return -1;
  size_t v = k + 1 + key.size();
  if (v >= bytes.size()) return -1;
  const unsigned char c = (unsigned char)bytes[v];
  auto be = [&](int n) {
    long r = 0;
    for (int i = 1; i <= n && v + i < bytes.size(); ++i) r = (r << 8) | (unsigned char)bytes[v + i];
    return r;
  };
  if (c < 0x80) return c;
  if (c == 0xcc) return be(1);
  if (c == 0xcd) return be(2);
  if (c == 0xce) return be(4);
  return -1;
}

// Assemble the hipRTC source, compile (or fetch from the on-disk cache).
int compileProgram(snl_ctx ctx, const char* shader, const char* ior, const char* defines, int maxMarch,
                   int maxBounces, int maxAtm, int spf, unsigned flags, Compiled& out, char* log, int loglen) {
  const unsigned mode = flags & (SNL_BUILD_AO | SNL_BUILD_PTSIMPLE);
  if (mode == (SNL_BUILD_AO | SNL_BUILD_PTSIMPLE))
    return fail(ctx, SNL_E_ARG, "SNL_BUILD_AO and SNL_BUILD_PTSIMPLE exclude each other");
  // the AO and simplified programs are RGB only: DISPERSION_ENABLED is dropped
  std::string defs = defines ? defines : "";
  if (mode) {
    const std::string d = "#define DISPERSION_ENABLED";
    for (size_t at = defs.find(d); at != std::string::npos; at = defs.find(d, at)) {
      size_t e = at + d.size();
      if (e < defs.size() && (isalnum((unsigned char)defs[e]) || defs[e] == '_')) { at = e; continue; }
      defs.erase(at, d.size());
    }
  }
  std::string glsl = std::string(shader ? shader : "") + "\n" + (ior ? ior : "") + "\n";
  // Renderer.compileShaders substitutes these tokens throughout the assembled
  // source, scene code included (js/renderer.js:288-294)
  const std::pair<const char*, int> tokens[] = {{"__MAX_MARCH_STEPS__", maxMarch},
                                                {"__MAX_BOUNCES__", maxBounces},
                                                {"__MAX_ATMOSPHERE_SCATTERS__", maxAtm},
                                                {"__MAX_SAMPLES_PER_FRAME__", spf}};
  for (const auto& tk : tokens) {
    std::string v = std::to_string(tk.second);
    for (size_t at = glsl.find(tk.first); at != std::string::npos; at = glsl.find(tk.first, at + v.size()))
      glsl.replace(at, strlen(tk.first), v);
  }
  out.defines = snl::parseDefines(defs);
  if (!snl::translateScene(glsl, builtinNames(), out.defines, out.tr)) {
    if (log && loglen > 0) snprintf(log, loglen, "%s", out.tr.error.c_str());
    return fail(ctx, SNL_E_COMPILE, "scene translation failed: %s", out.tr.error.c_str());
  }
  // the fast build lets the scene's divisions become products with a 1-ulp
  // reciprocal (within GLSL ES 3.00's 2.5-ulp bound for x / y), so a divisor that
  // is constant along a march is inverted once; the RNG in the prelude stays
  // correctly rounded
  const bool fastDiv = (flags & SNL_BUILD_FAST) && !(flags & SNL_BUILD_PRECISE);
  std::ostringstream src;
  src << "#include \"snl_prelude.h\"\n#include \"snl_wavefront.h\"\n";
  src << defs << "\n";
  if (fastDiv) src << "#pragma clang fp reciprocal(on)\n";
  src << "namespace glsl {\n" << out.tr.code << "\n}  // namespace glsl\n";
  if (fastDiv) src << "#pragma clang fp reciprocal(off)\n";
  src << "#line 1 \"snl_pt.h\"\n#include \"snl_pt.h\"\n#include \"snl_kernels.h\"\n#include \"snl_persist.h\"\n#include \"snl_mega.h\"\n"
         "#include \"snl_modes.h\"\n";
  std::string source = src.str();
  if (const char* dump = getenv("SNL_DUMP_SOURCE")) {  // for offline ISA inspection
    std::ofstream f(dump);
    f << source;
  }

  std::vector<std::string> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                   "-DSNL_MAX_MARCH_STEPS=" + std::to_string(maxMarch),
                                   "-DSNL_MAX_ATMOSPHERE_SCATTERS=" + std::to_string(maxAtm),
                                   "-DSNL_MAX_SAMPLES_PER_FRAME=" + std::to_string(spf)};
  if (mode == SNL_BUILD_AO) opts.push_back("-DSNL_MODE_AO");
  if (mode == SNL_BUILD_PTSIMPLE) opts.push_back("-DSNL_MODE_PTSIMPLE");
  if (flags & SNL_BUILD_PRECISE) opts.push_back("-ffp-contract=off");
  if (fastDiv) {
    opts.push_back("-fno-hip-fp32-correctly-rounded-divide-sqrt");
    // subnormal results flush to zero (GLSL ES 3.00 lets highp flush them): no
    // scaling around sqrt, reciprocals and divisions
    opts.push_back("-fgpu-flush-denormals-to-zero");
  }
  out.defs = defs;
  out.fastDiv = fastDiv;
  out.maxAtm = maxAtm;
  int rc = compileCached(ctx, source, opts, out.code, log, loglen);
  if (rc) return rc;
  // The megakernel asks for four waves per SIMD (at most 128 VGPRs); where that
  // makes it spill more than SNL_MEGA_SPILLS VGPRs (default 0) it is built
  // again without the bound.
  const long spillsOk = getenv("SNL_MEGA_SPILLS") ? atol(getenv("SNL_MEGA_SPILLS")) : 0;
  long spills = kernelMetaInt(out.code, "snl_mega", ".vgpr_spill_count");
  if (spills < 0) spills = kernelPrivateBytes(out.code, "snl_mega.kd") > 0 ? 1 << 20 : 0;
  if (spills > spillsOk) {
    out.spillCode.swap(out.code);
    opts.push_back("-DSNL_MEGA_WAVES=1");
    rc = compileCached(ctx, source, opts, out.code, log, loglen);
  }
  return rc;
}

FusedPart partOf(const Compiled& c, int maxMarch, unsigned flags) {
  FusedPart f;
  f.sceneCode = c.tr.code;
  f.defs = c.defs;
  f.maxMarch = maxMarch;
  f.maxAtm = c.maxAtm;
  f.fastDiv = c.fastDiv;
  f.precise = (flags & SNL_BUILD_PRECISE) != 0;
  f.dispersion = c.defines.count("DISPERSION_ENABLED") > 0;
  f.interactive = c.defines.count("INTERACTIVE_MODE") > 0;
  f.fusable = (flags & (SNL_BUILD_AO | SNL_BUILD_PTSIMPLE)) == 0;
  f.key = fnv1a(f.defs + "\n" + std::to_string(maxMarch) + " " + std::to_string(f.maxAtm) + "\n" + f.sceneCode);
  return f;
}

// The batch module of the programs `kinds` (one build group, sorted by key):
// each program's translated scene in a namespace of its own, between its own
// #defines and __MAX_*__ limits, then the kernel headers once, then a
// persistent kernel that walks the jobs in the order given and runs each with
// its program's megakernel body -- the code and arithmetic of that program's
// snl_mega, so a job renders bit for bit as it does alone.
std::string fusedSource(const std::vector<const FusedPart*>& kinds) {
  const FusedPart& g = *kinds[0];
  std::ostringstream src;
  src << "#define SNL_FUSED 1\n#include \"snl_prelude.h\"\n#include \"snl_wavefront.h\"\n";
  if (g.dispersion) src << "#define DISPERSION_ENABLED\n";
  if (g.interactive) src << "#define INTERACTIVE_MODE\n";
  if (g.fastDiv) src << "#pragma clang fp reciprocal(on)\n";
  for (size_t k = 0; k < kinds.size(); ++k) {
    const FusedPart& p = *kinds[k];
    std::set<std::string> names;
    std::istringstream in(p.defs);
    std::string line;
    src << "#define SNL_MAX_MARCH_STEPS " << p.maxMarch << "\n#define SNL_MAX_ATMOSPHERE_SCATTERS " << p.maxAtm
        << "\n";
    while (std::getline(in, line)) {
      std::istringstream w(line);
      std::string d, name;
      w >> d >> name;
      if (d != "#define" || name.empty() || name == "DISPERSION_ENABLED" || name == "INTERACTIVE_MODE") continue;
      src << line << "\n";
      names.insert(name);
    }
    src << "namespace glsl {\nnamespace snlk" << k << " {\n" << p.sceneCode << "\n}  // namespace snlk" << k
        << "\n}  // namespace glsl\n";
    for (const auto& n : names) src << "#undef " << n << "\n";
    src << "#undef SNL_MAX_MARCH_STEPS\n#undef SNL_MAX_ATMOSPHERE_SCATTERS\n";
  }
  if (g.fastDiv) src << "#pragma clang fp reciprocal(off)\n";
  // for the single-program kernels' templates, which this module does not instantiate
  src << "#define SNL_MAX_MARCH_STEPS 256\n#define SNL_MAX_ATMOSPHERE_SCATTERS 1\n";
  src << "#line 1 \"snl_pt.h\"\n#include \"snl_pt.h\"\n#include \"snl_kernels.h\"\n#include \"snl_persist.h\"\n"
         "#include \"snl_mega.h\"\n";
  src << "namespace glsl {\n"
         "struct SnlFusedArgs {\n"
         "  const SnlKArgs* jobs;  // per job: its program's parameters and its state\n"
         "  const int* kind;       // per job: its program's instantiation\n"
         "  const int* order;      // the jobs in the order the waves take them\n"
         "  int njobs;\n"
         "};\n"
         "extern \"C\" __global__ void snl_fused_reset(SnlFusedArgs f) {\n"
         "  if ((int)blockIdx.x < f.njobs && threadIdx.x < C_COUNT) f.jobs[blockIdx.x].cnt[threadIdx.x] = 0;\n"
         "}\n"
         "// each kind's megakernel body as a function of its own, so that its\n"
         "// registers are allocated as in its own kernel (inlined into one switch,\n"
         "// the kinds' bodies would share one allocation)\n"
         "template <class SC>\n"
         "__device__ __noinline__ void snl_fused_job(const SnlKArgs& a) { mega_body<SC>(a); }\n"
         "extern \"C\" __global__ void __launch_bounds__(256, SNL_MEGA_WAVES) snl_mega_fused(SnlFusedArgs f) {\n"
         "  for (int q = 0; q < f.njobs; ++q) {\n"
         "    const int j = f.order[q];\n"
         "    switch (f.kind[j]) {\n";
  for (size_t k = 0; k < kinds.size(); ++k)
    src << "      case " << k << ": snl_fused_job<snlk" << k << "::SnlScene>(f.jobs[j]); break;\n";
  src << "      default: break;\n    }\n  }\n}\n}  // namespace glsl\n";
  return src.str();
}

// compile (or fetch from the cache) the batch module of `kinds`; no device needed
int compileFused(snl_ctx ctx, std::vector<const FusedPart*> kinds, std::vector<char>& code, std::string& source) {
  if (kinds.empty()) return fail(ctx, SNL_E_ARG, "a batch module needs programs");
  std::sort(kinds.begin(), kinds.end(), [](const FusedPart* a, const FusedPart* b) { return a->key < b->key; });
  kinds.erase(std::unique(kinds.begin(), kinds.end(), [](const FusedPart* a, const FusedPart* b) {
                return a->key == b->key; }), kinds.end());
  for (const FusedPart* k : kinds) {
    if (!k->fusable) return fail(ctx, SNL_E_ARG, "a batch module holds path-tracer programs only");
    if (!k->sameGroup(*kinds[0]))
      return fail(ctx, SNL_E_ARG, "a batch module's programs share dispersion, interactive mode and build mode");
  }
  source = fusedSource(kinds);
  const FusedPart& g = *kinds[0];
  std::vector<std::string> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  if (g.precise) opts.push_back("-ffp-contract=off");
  if (g.fastDiv) {
    opts.push_back("-fno-hip-fp32-correctly-rounded-divide-sqrt");
    opts.push_back("-fgpu-flush-denormals-to-zero");
  }
  // The kinds' bodies are functions of their own, so the module's private
  // segment holds their call stack and does not show spills; the VGPR spill
  // count of the code object's metadata does.  Four waves per SIMD where that
  // spills at most a few registers, else three, else two.
  int rc = SNL_OK;
  const int forced = getenv("SNL_FUSED_WAVES") ? atoi(getenv("SNL_FUSED_WAVES")) : 0;
  for (int waves : {4, 3, 2}) {
    if (forced && waves != forced) continue;
    std::vector<std::string> o = opts;
    o.push_back("-DSNL_MEGA_WAVES=" + std::to_string(waves));
    if ((rc = compileCached(ctx, source, o, code, nullptr, 0))) return rc;
    long spills = kernelMetaInt(code, "snl_mega_fused", ".vgpr_spill_count");
    if (forced || (spills >= 0 && spills <= 8)) return SNL_OK;
  }
  return rc;
}

int buildFused(snl_ctx ctx, const std::vector<const FusedPart*>& kinds, FusedModule*& out) {
  std::vector<char> code;
  std::string source;
  int rc = compileFused(ctx, kinds, code, source);
  if (rc) return rc;
  const unsigned long long key = fnv1a(source);
  auto it = ctx->fused.find(key);
  if (it != ctx->fused.end()) { out = it->second; return SNL_OK; }
  std::unique_ptr<FusedModule> m(new FusedModule);
  HIPCHK(ctx, hipModuleLoadData(&m->mod, code.data()));
  HIPCHK(ctx, hipModuleGetFunction(&m->k_run, m->mod, "snl_mega_fused"));
  HIPCHK(ctx, hipModuleGetFunction(&m->k_reset, m->mod, "snl_fused_reset"));
  int occ = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, m->k_run, 256, 0) != hipSuccess || occ < 1) occ = 2;
  m->blocks = ctx->numCU * occ;
  std::vector<const FusedPart*> sorted(kinds);
  std::sort(sorted.begin(), sorted.end(), [](const FusedPart* a, const FusedPart* b) { return a->key < b->key; });
  sorted.erase(std::unique(sorted.begin(), sorted.end(), [](const FusedPart* a, const FusedPart* b) {
                 return a->key == b->key; }), sorted.end());
  for (size_t k = 0; k < sorted.size(); ++k) m->kindOf[sorted[k]->key] = (int)k;
  m->group = *sorted[0];
  m->group.sceneCode.clear();
  out = m.get();
  ctx->fused[key] = m.release();
  return SNL_OK;
}

void fillArgs(snl_prog p, snl_state st, SnlKArgs& a) {
  memset(&a, 0, sizeof a);
  a.P = p->dP;
  a.U = p->dU;
  long long cap = st->npix;
  a.cap = cap;
  a.pf = st->pf;
  a.pi = st->pi;
  a.qmat = st->q ? st->q + 4 * cap : nullptr;
  a.cnt = st->cnt;
  a.radiance = st->radiance;
  a.rngstate = st->rng;
  a.W = st->w;
  a.H = st->h;
  a.npix = (int)st->npix;
  a.imagePix = st->w * st->h;
  a.spf = p->spf;
  a.chunk = p->traceChunk;
  a.frames = 1;
  a.maxBounces = p->maxBounces;
  a.burst = p->marchBurst;
}

int launch(snl_ctx ctx, hipFunction_t f, int grid, int block, SnlKArgs& a) {
  void* args[] = {&a};
  HIPCHK(ctx, hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, ctx->stream, args, nullptr));
  return SNL_OK;
}

// ---- submitted batches: a program or state that a job of a submitted batch
// renders is busy until that job completes; every call that touches its device
// memory first waits for that job (settle)

// job j of b is complete: read its counters, free its program and state
int jobDone(snl_batch_s* b, int j) {
  if (b->done[j]) return SNL_OK;
  snl_ctx ctx = b->ctx;
  b->done[j] = 1;
  --b->pending;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, b->ev0, b->ev[j]) == hipSuccess) b->ms[j] = ms;
  HIPCHK(ctx, hipMemcpyAsync(&b->steps[j], b->states[j]->tileCnt + C_TRACE_STEPS, 8, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&b->normals[j], b->states[j]->tileCnt + C_NORMALS, 8, hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (b->progs[j]->busy == b && b->progs[j]->busyJob == j) b->progs[j]->busy = nullptr;
  if (b->states[j]->busy == b && b->states[j]->busyJob == j) b->states[j]->busy = nullptr;
  return SNL_OK;
}
int jobWait(snl_batch_s* b, int j) {
  if (b->done[j]) return SNL_OK;
  HIPCHK(b->ctx, hipEventSynchronize(b->ev[j]));
  return jobDone(b, j);
}
// 1 complete, 0 still running, < 0 error
int jobQuery(snl_batch_s* b, int j) {
  if (b->done[j]) return 1;
  hipError_t e = hipEventQuery(b->ev[j]);
  if (e == hipErrorNotReady) return 0;
  if (e != hipSuccess) return fail(b->ctx, SNL_E_HIP, "job %d of a batch: %s", j, hipGetErrorString(e));
  int rc = jobDone(b, j);
  return rc ? rc : 1;
}
template <class T>
int settle(T* o) {
  return (o && o->busy) ? jobWait(o->busy, o->busyJob) : SNL_OK;
}

int upload(snl_prog p) {
  int rc = settle(p);
  if (rc) return rc;
  if (!p->dirty) return SNL_OK;
  // 0.5 * radians(camFovy) in float, as the shaders form the argument
  const float halfFovy = 0.5f * (p->hostP.camFovy * 0.017453292519943295f);
  p->hostP.camFrustumHeight = (float)std::tan((double)halfFovy);
  HIPCHK(p->ctx, hipMemcpyAsync(p->dP, &p->hostP, sizeof(SnlParams), hipMemcpyHostToDevice, p->ctx->stream));
  if (!p->hostU.empty())
    HIPCHK(p->ctx, hipMemcpyAsync(p->dU, p->hostU.data(), p->hostU.size(), hipMemcpyHostToDevice, p->ctx->stream));
  // the host shadow may change right after this call returns
  HIPCHK(p->ctx, hipStreamSynchronize(p->ctx->stream));
  p->dirty = false;
  return SNL_OK;
}

// PCG32-seeded integer RNG states (see snelly.h: snl_state_seed)
__device__ inline unsigned pcg_next(unsigned long long& state, unsigned long long inc) {
  unsigned long long old = state;
  state = old * 6364136223846793005ull + inc;
  unsigned xorshifted = (unsigned)(((old >> 18u) ^ old) >> 27u);
  unsigned rot = (unsigned)(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}
__global__ void seed_kernel(float* radiance, float* rngstate, long long npix, unsigned seed, int mode) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  if (mode & 1) {
    unsigned long long state = 0ull, inc = (54ull << 1u) | 1u;
    pcg_next(state, inc);
    state += (unsigned long long)((unsigned)i ^ seed);
    pcg_next(state, inc);
    const unsigned m[4] = {4194287u, 4194277u, 4194191u, 4194167u};
    for (int k = 0; k < 4; ++k) rngstate[4 * i + k] = float(1u + pcg_next(state, inc) % (m[k] - 1u));
  }
  if (mode & 2) {
    radiance[4 * i] = 0.0f;
    radiance[4 * i + 1] = 0.0f;
    radiance[4 * i + 2] = 0.0f;
    radiance[4 * i + 3] = 1.0f;
  }
}

int blocksFor(long long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// ---------------------------------------------------------------- allocation-free render's scratch
// snl_render_noalloc keeps the render's queue counters and its tile-order
// buffers (costs, cost histogram, order) in this array instead of the state's
// allocations.  Its size bounds the image: at most SNL_NOALLOC_MAX_TILES 8x8
// tiles (a 4096 x 4096 image), 2.1 MB per GPU, held until the process ends.
// On the host the array's name is no device address: set-up resolves each
// device's copy with hipGetSymbolAddress and every host use goes through that.
constexpr int kNoallocCostBuckets = 128;  // = kCostBuckets of the tile order
constexpr size_t kScratchInts = C_COUNT + 2 * (size_t)SNL_NOALLOC_MAX_TILES + kNoallocCostBuckets;
__device__ __align__(256) int scratch[kScratchInts];
std::mutex g_scratchMu;
std::map<int, int*> g_scratchPtr;  // device -> address of its copy of `scratch`

// the address of `scratch` on `device`, resolved once per device with that
// device current (the caller has made it current)
int scratchFor(int device, int** out) {
  std::lock_guard<std::mutex> lk(g_scratchMu);
  auto it = g_scratchPtr.find(device);
  if (it == g_scratchPtr.end()) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(nullptr, SNL_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                  hipGetErrorString(e));
    it = g_scratchPtr.emplace(device, static_cast<int*>(ptr)).first;
  }
  *out = it->second;
  return SNL_OK;
}

}  // namespace

// ---------------------------------------------------------------- API
extern "C" {

const char* snl_version(void) { return "snelly-mi355x 0.1 (gfx950)"; }

const char* snl_last_error(snl_ctx ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int snl_ctx_create(int device, snl_ctx* out) {
  if (!out) return fail(nullptr, SNL_E_ARG, "out is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= device || device < 0)
    return fail(nullptr, SNL_E_HIP, "no HIP device %d (count %d): %s", device, n, hipGetErrorString(e));
  std::unique_ptr<snl_ctx_s> c(new snl_ctx_s);
  c->device = device;
  HIPCHK(nullptr, hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(nullptr, hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, SNL_E_UNSUPPORTED, "device %d is %s, this library is built for gfx950", device, prop.gcnArchName);
  c->numCU = prop.multiProcessorCount;
  HIPCHK(nullptr, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int k = 0; k < SNL_MAX_TILES; ++k) {
    HIPCHK(nullptr, hipStreamCreateWithFlags(&c->tileStream[k], hipStreamNonBlocking));
    HIPCHK(nullptr, hipEventCreateWithFlags(&c->tileEvent[k], hipEventDisableTiming));
  }
  HIPCHK(nullptr, hipEventCreateWithFlags(&c->joinEvent, hipEventDisableTiming));
  if (int rc = scratchFor(device, &c->scratch)) return rc;
  *out = c.release();
  return SNL_OK;
}

int snl_ctx_destroy(snl_ctx ctx) {
  if (!ctx) return SNL_OK;
  hipSetDevice(ctx->device);
  while (!ctx->batches.empty()) (void)snl_batch_release(*ctx->batches.begin());
  if (ctx->graphStream) (void)hipStreamDestroy(ctx->graphStream);
  for (auto& kv : ctx->fused) {
    if (kv.second->mod) (void)hipModuleUnload(kv.second->mod);
    delete kv.second;
  }
  ctx->fused.clear();
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  for (int k = 0; k < SNL_MAX_TILES; ++k) {
    if (ctx->tileStream[k]) hipStreamDestroy(ctx->tileStream[k]);
    if (ctx->tileEvent[k]) hipEventDestroy(ctx->tileEvent[k]);
  }
  if (ctx->joinEvent) hipEventDestroy(ctx->joinEvent);
  for (auto& st : ctx->batchStream) hipStreamDestroy(st);
  for (auto& ev : ctx->batchEvent) hipEventDestroy(ev);
  delete ctx;
  return SNL_OK;
}

int snl_sync(snl_ctx ctx) {
  if (!ctx) return fail(nullptr, SNL_E_ARG, "ctx is NULL");
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return SNL_OK;
}

int snl_program_precompile(const char* shader, const char* ior, const char* defines, int maxMarch, int maxBounces,
                           int maxAtm, int spf, unsigned flags, char* log, int loglen) {
  Compiled c;
  return compileProgram(nullptr, shader, ior, defines, maxMarch, maxBounces, maxAtm, spf, flags, c, log, loglen);
}

int snl_program_build(snl_ctx ctx, const char* shader, const char* ior, const char* defines, int maxMarch,
                      int maxBounces, int maxAtm, int spf, unsigned flags, snl_prog* out, char* log, int loglen) {
  if (!ctx || !out) return fail(ctx, SNL_E_ARG, "NULL argument");
  if (maxMarch < 1 || maxBounces < 0 || spf < 1) return fail(ctx, SNL_E_ARG, "bad loop bounds");
  Compiled c;
  int rc = compileProgram(ctx, shader, ior, defines, maxMarch, maxBounces, maxAtm, spf, flags, c, log, loglen);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<snl_prog_s> p(new snl_prog_s);
  p->ctx = ctx;
  HIPCHK(ctx, hipModuleLoadData(&p->mod, c.code.data()));
  struct { hipFunction_t* f; const char* n; } fns[] = {
      {&p->k_reset, "snl_counters_reset"}, {&p->k_rotate, "snl_counters_rotate"}, {&p->k_camera, "snl_camera"},
      {&p->k_trace, "snl_trace"}, {&p->k_bin, "snl_bin"}, {&p->k_miss, "snl_shade_miss"},
      {&p->k_surf, "snl_shade_surface"}, {&p->k_metal, "snl_shade_metal"}, {&p->k_diele, "snl_shade_dielectric"},
      {&p->k_accum, "snl_accumulate"}, {&p->k_normals, "snl_normals"}, {&p->k_pick, "snl_pick"},
      {&p->k_tonemap, "snl_tonemap"}, {&p->k_persist, "snl_persist"}, {&p->k_persist_reset, "snl_persist_reset"},
      {&p->k_mega, "snl_mega"}};
  for (auto& f : fns) HIPCHK(ctx, hipModuleGetFunction(f.f, p->mod, f.n));
  p->mode = flags & (SNL_BUILD_AO | SNL_BUILD_PTSIMPLE);
  if (p->mode) HIPCHK(ctx, hipModuleGetFunction(&p->k_mode, p->mod, "snl_mode_kernel"));
  p->hasSurface = c.defines.count("HAS_SURFACE") > 0;
  p->hasMetal = c.defines.count("HAS_METAL") > 0;
  p->hasDiele = c.defines.count("HAS_DIELECTRIC") > 0;
  p->hasAtmosphere = c.defines.count("HAS_ATMOSPHERE") > 0;
  p->megaOnly = p->hasAtmosphere || c.defines.count("HAS_VOLUME") > 0;
  p->maxBounces = maxBounces;
  p->maxMarch = maxMarch;
  p->spf = spf;
  p->part = partOf(c, maxMarch, flags);
  memset(&p->hostP, 0, sizeof(SnlParams));
  p->user = c.tr.uniforms;
  p->hostU.assign(c.tr.userBlockSize, 0);
  {
    // samplers the program can read: those the scene code names, and those of
    // the materials' IOR code when something calls it -- the metal's
    // IOR_METAL / K_METAL run only in the metal BSDF (under HAS_METAL) unless
    // the scene code calls them itself, so a program without a metal never
    // reads the default tabulated metal's kTex
    std::string text = shader ? shader : "";
    auto mentionsIn = [](const std::string& t, const std::string& name) {
      for (size_t at = t.find(name); at != std::string::npos; at = t.find(name, at + 1)) {
        bool l = at == 0 || !(isalnum((unsigned char)t[at - 1]) || t[at - 1] == '_');
        size_t e = at + name.size();
        bool r = e >= t.size() || !(isalnum((unsigned char)t[e]) || t[e] == '_');
        if (l && r) return true;
      }
      return false;
    };
    if (ior && (p->hasMetal || mentionsIn(text, "IOR_METAL") || mentionsIn(text, "K_METAL")))
      text += std::string("\n") + ior;
    auto mentions = [&](const std::string& name) { return mentionsIn(text, name); };
    const auto& b = builtins();
    for (int i = 0; i < (int)b.size(); ++i)
      if (b[i].type == SNL_T_SAMPLER2D && mentions(b[i].name)) p->sampled2D.push_back(i);
    for (int j = 0; j < (int)p->user.size(); ++j)
      if (p->user[j].type == SNL_T_SAMPLER2D) p->sampled2D.push_back(1000 + j);
  }
  HIPCHK(ctx, hipMalloc((void**)&p->dP, sizeof(SnlParams)));
  HIPCHK(ctx, hipMalloc(&p->dU, p->hostU.size()));
  if (const char* e = getenv("SNL_TILE_ORDER")) p->orderTiles = atoi(e) != 0;
  p->traceBlocks = ctx->numCU * 8;
  p->traceChunk = 64;
  p->shadeBlocks = ctx->numCU * 8;
  int occ = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, p->k_persist, 256, 0) != hipSuccess || occ < 1) occ = 2;
  p->persistBlocks = ctx->numCU * occ;
  occ = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, p->k_mega, 256, 0) != hipSuccess || occ < 1) occ = 2;
  p->megaBlocks = ctx->numCU * occ;
  p->k_megaMain = p->k_mega;
  p->spillCode.swap(c.spillCode);
  if (p->mode) {
    occ = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, p->k_mode, 256, 0) != hipSuccess || occ < 1) occ = 2;
    p->modeBlocks = ctx->numCU * occ;
  }
  HIPCHK(ctx, hipEventCreate(&p->ev0));
  HIPCHK(ctx, hipEventCreate(&p->ev1));
  HIPCHK(ctx, hipEventCreate(&p->et0));
  HIPCHK(ctx, hipEventCreate(&p->et1));
  *out = p.release();
  return SNL_OK;
}

int snl_program_destroy(snl_prog p) {
  if (!p) return SNL_OK;
  hipSetDevice(p->ctx->device);
  (void)settle(p);
  hipStreamSynchronize(p->ctx->stream);
  if (p->dP) hipFree(p->dP);
  if (p->dU) hipFree(p->dU);
  if (p->mod) hipModuleUnload(p->mod);
  if (p->spillMod) hipModuleUnload(p->spillMod);
  for (auto& kv : p->texOf)
    if (kv.second) kv.second->users.erase(p);
  for (hipEvent_t e : {p->ev0, p->ev1, p->et0, p->et1}) if (e) hipEventDestroy(e);
  for (hipEvent_t e : p->traceEvents) hipEventDestroy(e);
  delete p;
  return SNL_OK;
}

int snl_uniform_index(snl_prog p, const char* name) {
  if (!p || !name) return -1;
  const auto& b = builtins();
  for (size_t i = 0; i < b.size(); ++i) if (!strcmp(b[i].name, name)) return (int)i;
  for (size_t j = 0; j < p->user.size(); ++j) if (p->user[j].name == name) return 1000 + (int)j;
  return -1;
}

static int typeClass(int t) { return (t == SNL_T_INT) ? 1 : (t >= SNL_T_SAMPLER2D ? 2 : 0); }
static int typeFloats(int t) {
  switch (t) {
    case SNL_T_FLOAT: case SNL_T_INT: return 1;
    case SNL_T_VEC2: return 2;
    case SNL_T_VEC3: return 3;
    case SNL_T_VEC4: case SNL_T_MAT2: return 4;
    case SNL_T_MAT3: return 9;
    case SNL_T_MAT4: return 16;
  }
  return 0;
}

int snl_uniform_set(snl_prog p, int idx, int type, const void* data, int count) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  if (idx < 0) return SNL_OK;  // unknown uniform: silently ignored, as WebGL does for a null location
  if (!data || count <= 0) return fail(p->ctx, SNL_E_ARG, "no data");
  char* base;
  int off, utype, bytes;
  if (idx < 1000) {
    const auto& b = builtins();
    if (idx >= (int)b.size()) return fail(p->ctx, SNL_E_ARG, "bad uniform index %d", idx);
    base = reinterpret_cast<char*>(&p->hostP);
    off = b[idx].offset;
    utype = b[idx].type;
    bytes = b[idx].size;
  } else {
    int j = idx - 1000;
    if (j >= (int)p->user.size()) return fail(p->ctx, SNL_E_ARG, "bad uniform index %d", idx);
    base = p->hostU.data();
    off = p->user[j].offset;
    utype = p->user[j].type;
    bytes = p->user[j].size * p->user[j].count;
  }
  // a float setter on an int uniform (or the reverse) is a GL error with no effect
  if (typeClass(type) != typeClass(utype)) return SNL_OK;
  if (typeClass(type) == 2) return fail(p->ctx, SNL_E_ARG, "use snl_uniform_texture for samplers");
  int n = typeFloats(type) * count * 4;
  memcpy(base + off, data, (size_t)(n < bytes ? n : bytes));
  p->dirty = true;
  return SNL_OK;
}

int snl_uniform_texture(snl_prog p, int idx, snl_tex t) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  if (idx < 0) return SNL_OK;
  char* base;
  int off, utype;
  if (idx < 1000) {
    const auto& b = builtins();
    if (idx >= (int)b.size()) return fail(p->ctx, SNL_E_ARG, "bad uniform index");
    base = reinterpret_cast<char*>(&p->hostP);
    off = b[idx].offset;
    utype = b[idx].type;
  } else {
    int j = idx - 1000;
    if (j >= (int)p->user.size()) return fail(p->ctx, SNL_E_ARG, "bad uniform index");
    base = p->hostU.data();
    off = p->user[j].offset;
    utype = p->user[j].type;
  }
  if (utype != SNL_T_SAMPLER2D && utype != SNL_T_SAMPLER3D) return SNL_OK;  // not a sampler: GL error, no effect
  if (t) p->bound.insert(idx);
  else p->bound.erase(idx);
  // keep the texture -> program links that snl_texture_destroy follows
  snl_tex_s* old = p->texOf.count(idx) ? p->texOf[idx] : nullptr;
  if (t) p->texOf[idx] = t;
  else p->texOf.erase(idx);
  if (old && old != t) {
    bool still = false;
    for (auto& kv : p->texOf) still |= kv.second == old;
    if (!still) old->users.erase(p);
  }
  if (t) t->users.insert(p);
  if (utype == SNL_T_SAMPLER2D) {
    sampler2D s{t ? t->d : nullptr, t ? t->w : 0, t ? t->h : 0, t ? t->ch : 4, t ? t->flags : 0};
    memcpy(base + off, &s, sizeof s);
  } else {
    sampler3D s{t ? t->d : nullptr, t ? t->w : 0, t ? t->h : 0, t ? t->depth : 0, t ? t->ch : 4, t ? t->flags : 0};
    memcpy(base + off, &s, sizeof s);
  }
  p->dirty = true;
  return SNL_OK;
}

int snl_texture_create(snl_ctx ctx, int w, int h, int depth, int ch, int linear, int clamp, const float* texels,
                       snl_tex* out) {
  if (!ctx || !out || w <= 0 || h <= 0 || depth <= 0 || ch < 1 || ch > 4) return fail(ctx, SNL_E_ARG, "bad texture shape");
  std::unique_ptr<snl_tex_s> t(new snl_tex_s);
  t->ctx = ctx;
  t->w = w; t->h = h; t->depth = depth; t->ch = ch;
  t->flags = (linear ? 1 : 0) | (clamp ? 2 : 0);
  size_t bytes = (size_t)w * h * depth * ch * sizeof(float);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (hipMalloc((void**)&t->d, bytes) != hipSuccess) return fail(ctx, SNL_E_NOMEM, "texture allocation of %zu bytes failed", bytes);
  if (texels) HIPCHK(ctx, hipMemcpyAsync(t->d, texels, bytes, hipMemcpyHostToDevice, ctx->stream));
  else HIPCHK(ctx, hipMemsetAsync(t->d, 0, bytes, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *out = t.release();
  return SNL_OK;
}

int snl_texture_info(snl_tex t, int* width, int* height, int* depth, int* channels) {
  if (!t) return fail(nullptr, SNL_E_ARG, "texture is NULL");
  if (width) *width = t->w;
  if (height) *height = t->h;
  if (depth) *depth = t->depth;
  if (channels) *channels = t->ch;
  return SNL_OK;
}
int snl_texture_upload(snl_tex t, const float* texels) {
  if (!t || !texels) return fail(t ? t->ctx : nullptr, SNL_E_ARG, "NULL argument");
  size_t bytes = (size_t)t->w * t->h * t->depth * t->ch * sizeof(float);
  // a submitted batch job of a program that samples the texture reads the old texels
  for (snl_prog_s* p : t->users)
    if (int rc = settle(p)) return rc;
  HIPCHK(t->ctx, hipMemcpyAsync(t->d, texels, bytes, hipMemcpyHostToDevice, t->ctx->stream));
  HIPCHK(t->ctx, hipStreamSynchronize(t->ctx->stream));
  return SNL_OK;
}

int snl_texture_destroy(snl_tex t) {
  if (!t) return SNL_OK;
  hipStreamSynchronize(t->ctx->stream);
  // a submitted batch job of a program that samples the texture finishes first
  for (snl_prog_s* p : t->users)
    if (int rc = settle(p)) return rc;
  // a sampler still bound to the texture now reads its unit's missing texture,
  // (0, 0, 0, 1) as in WebGL, instead of freed memory; it stays assigned a unit
  for (snl_prog_s* p : t->users) {
    for (auto& kv : p->texOf) {
      if (kv.second != t) continue;
      int idx = kv.first;
      char* base = idx < 1000 ? reinterpret_cast<char*>(&p->hostP) : p->hostU.data();
      int off = idx < 1000 ? builtins()[idx].offset : p->user[idx - 1000].offset;
      int utype = idx < 1000 ? builtins()[idx].type : p->user[idx - 1000].type;
      if (utype == SNL_T_SAMPLER2D) {
        sampler2D s{nullptr, 0, 0, 4, 0};
        memcpy(base + off, &s, sizeof s);
      } else {
        sampler3D s{nullptr, 0, 0, 0, 4, 0};
        memcpy(base + off, &s, sizeof s);
      }
      kv.second = nullptr;
      p->dirty = true;
    }
  }
  if (t->d) hipFree(t->d);
  delete t;
  return SNL_OK;
}

int snl_state_create(snl_ctx ctx, int w, int h, snl_state* out) { return snl_state_create_chains(ctx, w, h, 1, out); }

int snl_state_create_chains(snl_ctx ctx, int w, int h, int chains, snl_state* out) {
  if (!ctx || !out || w <= 0 || h <= 0 || chains <= 0) return fail(ctx, SNL_E_ARG, "bad state size");
  if ((long long)w * h * chains > 0x7fffffffLL) return fail(ctx, SNL_E_ARG, "state too large");
  std::unique_ptr<snl_state_s> s(new snl_state_s);
  s->ctx = ctx;
  s->w = w; s->h = h; s->chains = chains;
  s->npix = (long long)w * h * chains;
  long long n = s->npix;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // radiance and RNG (32 B per pixel) plus counters; the wavefront engine's
  // path and queue buffers are allocated on its first use
  struct { void** p; size_t bytes; } allocs[] = {
      {(void**)&s->radiance, (size_t)n * 16}, {(void**)&s->rng, (size_t)n * 16},
      {(void**)&s->cnt, (size_t)C_COUNT * 4}, {(void**)&s->tileCnt, (size_t)SNL_MAX_TILES * C_COUNT * 4}};
  for (auto& a : allocs) {
    if (hipMalloc(a.p, a.bytes) != hipSuccess) {
      for (auto& b : allocs) if (*b.p) hipFree(*b.p);
      return fail(ctx, SNL_E_NOMEM, "state allocation of %zu bytes failed", a.bytes);
    }
  }
  HIPCHK(ctx, hipMemsetAsync(s->cnt, 0, C_COUNT * 4, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(s->tileCnt, 0, SNL_MAX_TILES * C_COUNT * 4, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *out = s.release();
  return snl_state_seed(*out, 0);
}

int snl_state_destroy(snl_state s) {
  if (!s) return SNL_OK;
  hipSetDevice(s->ctx->device);
  (void)settle(s);
  hipStreamSynchronize(s->ctx->stream);
  for (void* p : {(void*)s->radiance, (void*)s->rng, (void*)s->pf, (void*)s->pi, (void*)s->q, (void*)s->cnt,
                  (void*)s->tileCnt, (void*)s->radPrev, (void*)s->tileCost, (void*)s->errPartial})
    if (p) hipFree(p);
  delete s;
  return SNL_OK;
}

// Error estimate of a state of K >= 2 chains of one image.  Per pixel and
// channel, with u_k the unbiased mean XYZ * N / (N - 1) of chain k (the count
// starts at 1, js/renderer.js:735-745) and m their mean over chains:
//   out[0] += sum_k (u_k - m)^2 / (K - 1) / K   (variance of m)
//   out[1] += m^2
// One partial per block, summed in a second pass in block order.
__global__ void chain_error_kernel(const float4* rad, long long imagePix, int K, double* partial) {
  __shared__ double s0[256], s1[256];
  double v2 = 0.0, m2 = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < imagePix;
       i += (long long)gridDim.x * blockDim.x) {
    double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
    for (int k = 0; k < K; ++k) {
      float4 x = rad[k * imagePix + i];
      double f = x.w > 1.0f ? (double)x.w / (x.w - 1.0) : 0.0;
      double u[3] = {x.x * f, x.y * f, x.z * f};
      for (int c = 0; c < 3; ++c) { sum[c] += u[c]; sq[c] += u[c] * u[c]; }
    }
    for (int c = 0; c < 3; ++c) {
      double m = sum[c] / K;
      double var = (sq[c] - K * m * m) / (K - 1);
      v2 += (var > 0.0 ? var : 0.0) / K;
      m2 += m * m;
    }
  }
  s0[threadIdx.x] = v2;
  s1[threadIdx.x] = m2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) { s0[threadIdx.x] += s0[threadIdx.x + k]; s1[threadIdx.x] += s1[threadIdx.x + k]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { partial[2 * blockIdx.x] = s0[0]; partial[2 * blockIdx.x + 1] = s1[0]; }
}
__global__ void error_sum_kernel(const double* partial, int n, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double v2 = 0.0, m2 = 0.0;
  for (int k = 0; k < n; ++k) { v2 += partial[2 * k]; m2 += partial[2 * k + 1]; }
  out[0] = v2;
  out[1] = m2;
}

static int seedMode(snl_state s, unsigned seed, int mode) {
  if (!s) return fail(nullptr, SNL_E_ARG, "state is NULL");
  if (int rc = settle(s)) return rc;
  hipLaunchKernelGGL(seed_kernel, dim3(blocksFor(s->npix)), dim3(256), 0, s->ctx->stream, s->radiance, s->rng,
                     s->npix, seed, mode);
  HIPCHK(s->ctx, hipGetLastError());
  HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return SNL_OK;
}
int snl_state_seed(snl_state s, unsigned seed) { return seedMode(s, seed, 3); }
int snl_state_clear(snl_state s) { return seedMode(s, 0, 2); }

int snl_state_info(snl_state s, int* width, int* height, int* chains) {
  if (!s) return fail(nullptr, SNL_E_ARG, "state is NULL");
  if (width) *width = s->w;
  if (height) *height = s->h;
  if (chains) *chains = s->chains;
  return SNL_OK;
}

int snl_state_read(snl_state s, int which, float* outp) {
  if (!s || !outp || which < 0 || which > 1) return fail(s ? s->ctx : nullptr, SNL_E_ARG, "bad argument");
  if (int rc = settle(s)) return rc;
  HIPCHK(s->ctx, hipMemcpyAsync(outp, which ? s->rng : s->radiance, (size_t)s->npix * 16, hipMemcpyDeviceToHost,
                                s->ctx->stream));
  HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return SNL_OK;
}

int snl_state_write(snl_state s, int which, const float* in) {
  if (!s || !in || which < 0 || which > 1) return fail(s ? s->ctx : nullptr, SNL_E_ARG, "bad argument");
  if (int rc = settle(s)) return rc;
  HIPCHK(s->ctx, hipMemcpyAsync(which ? s->rng : s->radiance, in, (size_t)s->npix * 16, hipMemcpyHostToDevice,
                                s->ctx->stream));
  HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return SNL_OK;
}

// A GL sampler uniform that is never assigned a unit reads unit 0, where the
// reference's Renderer has bound the radiance texture being read this frame
// (js/renderer.js:590, 706-713) -- e.g. TabulatedMetal's kTex, whose texture is
// bound under the name "k_tex" (js/materials.js:664-665).  Such samplers get a
// NEAREST/CLAMP view of a per-frame snapshot of the radiance buffer, since the
// kernels accumulate into that buffer while the frame runs.  Returns whether
// snapshots are needed; copyUnit0 takes one.
static int prepareUnit0(snl_prog p, snl_state st, bool& active) {
  active = false;
  for (int idx : p->sampled2D) active |= !p->bound.count(idx);
  if (!active) return SNL_OK;
  snl_ctx ctx = p->ctx;
  if (!st->radPrev) {
    if (hipMalloc((void**)&st->radPrev, (size_t)st->npix * 16) != hipSuccess)
      return fail(ctx, SNL_E_NOMEM, "radiance snapshot allocation failed");
  }
  sampler2D view{st->radPrev, st->w, st->h, 4, 2};
  for (int idx : p->sampled2D) {
    if (p->bound.count(idx)) continue;
    char* base = idx < 1000 ? reinterpret_cast<char*>(&p->hostP) : p->hostU.data();
    int off = idx < 1000 ? builtins()[idx].offset : p->user[idx - 1000].offset;
    memcpy(base + off, &view, sizeof view);
  }
  p->dirty = true;
  return SNL_OK;
}
static int copyUnit0(snl_ctx ctx, snl_state st, hipStream_t s) {
  HIPCHK(ctx, hipMemcpyAsync(st->radPrev, st->radiance, (size_t)st->npix * 16, hipMemcpyDeviceToDevice, s));
  return SNL_OK;
}

static int wavefrontBuffers(snl_ctx ctx, snl_state st) {
  if (st->pf) return SNL_OK;
  long long n = st->npix;
  struct { void** p; size_t bytes; } allocs[] = {
      {(void**)&st->pf, (size_t)n * F_COUNT * 4}, {(void**)&st->pi, (size_t)n * I_COUNT * 4},
      {(void**)&st->q, (size_t)n * 8 * 4}};
  for (auto& a : allocs) {
    if (hipMalloc(a.p, a.bytes) != hipSuccess) {
      for (auto& b : allocs) if (*b.p) { hipFree(*b.p); *b.p = nullptr; }
      return fail(ctx, SNL_E_NOMEM, "wavefront buffers of %zu bytes: allocation failed", a.bytes);
    }
  }
  HIPCHK(ctx, hipMemsetAsync(st->pi, 0, (size_t)n * I_COUNT * 4, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return SNL_OK;
}

static int renderWavefront(snl_prog p, snl_state st, int frames) {
  snl_ctx ctx = p->ctx;
  if (st->ctx != ctx) return fail(ctx, SNL_E_ARG, "program and state belong to different contexts");
  if (int rc0 = wavefrontBuffers(ctx, st)) return rc0;
  bool unit0 = false;
  int rc = prepareUnit0(p, st, unit0);
  if (rc) return rc;
  if ((rc = upload(p))) return rc;
  hipStream_t s = ctx->stream;
  SnlKArgs a;
  fillArgs(p, st, a);
  long long cap = st->npix;
  int* qt[2] = {st->q, st->q + cap};
  int* qn[2] = {st->q + 2 * cap, st->q + 3 * cap};
  int pixBlocks = blocksFor(st->npix);
  int shadeBlocks = std::min(pixBlocks, p->shadeBlocks);
  p->stats = snl_stats{};
  a.vertex = -2;  // zero every counter, the step counter included
  if ((rc = launch(ctx, p->k_reset, 1, 64, a))) return rc;
  HIPCHK(ctx, hipEventRecord(p->ev0, s));
  double traceMs = 0.0;
  int traceLaunches = 0;
  for (int f = 0; f < frames; ++f) {
    if (unit0 && (rc = copyUnit0(ctx, st, s))) return rc;
    for (int smp = 0; smp < p->spf; ++smp) {
      a.sample = smp;
      int cur = 0;
      a.qtrace_in = qt[1]; a.qnee_in = qn[1]; a.qtrace_out = qt[0]; a.qnee_out = qn[0];
      a.vertex = 0;
      if ((rc = launch(ctx, p->k_reset, 1, 64, a))) return rc;
      if ((rc = launch(ctx, p->k_camera, pixBlocks, 256, a))) return rc;
      for (int v = 0; v <= p->maxBounces + 1; ++v) {
        bool last = (v == p->maxBounces + 1);
        a.vertex = last ? -1 : v;
        if ((rc = launch(ctx, p->k_rotate, 1, 64, a))) return rc;
        // the queues written by the previous stage are read now
        a.qtrace_in = qt[cur]; a.qnee_in = qn[cur];
        a.qtrace_out = qt[1 - cur]; a.qnee_out = qn[1 - cur];
        cur = 1 - cur;
        if (p->profiling) {
          while ((int)p->traceEvents.size() < 2 * (traceLaunches + 1)) {
            hipEvent_t e;
            HIPCHK(ctx, hipEventCreate(&e));
            p->traceEvents.push_back(e);
          }
          HIPCHK(ctx, hipEventRecord(p->traceEvents[2 * traceLaunches], s));
        }
        if ((rc = launch(ctx, p->k_trace, p->traceBlocks, 256, a))) return rc;
        if (p->profiling) HIPCHK(ctx, hipEventRecord(p->traceEvents[2 * traceLaunches + 1], s));
        ++traceLaunches;
        if (last) break;
        if ((rc = launch(ctx, p->k_bin, shadeBlocks, 256, a))) return rc;
        if ((rc = launch(ctx, p->k_miss, shadeBlocks, 256, a))) return rc;
        if (p->hasSurface && (rc = launch(ctx, p->k_surf, shadeBlocks, 256, a))) return rc;
        if (p->hasMetal && (rc = launch(ctx, p->k_metal, shadeBlocks, 256, a))) return rc;
        if (p->hasDiele && (rc = launch(ctx, p->k_diele, shadeBlocks, 256, a))) return rc;
      }
      if ((rc = launch(ctx, p->k_accum, pixBlocks, 256, a))) return rc;
    }
  }
  HIPCHK(ctx, hipEventRecord(p->ev1, s));
  HIPCHK(ctx, hipEventSynchronize(p->ev1));
  float ms = 0.0f;
  HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev0, p->ev1));
  if (p->profiling) {
    for (int k = 0; k < traceLaunches; ++k) {
      float t = 0.0f;
      HIPCHK(ctx, hipEventElapsedTime(&t, p->traceEvents[2 * k], p->traceEvents[2 * k + 1]));
      traceMs += t;
    }
  }
  long long steps = 0, normals = 0;
  HIPCHK(ctx, hipMemcpyAsync(&steps, st->cnt + C_TRACE_STEPS, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipMemcpyAsync(&normals, st->cnt + C_NORMALS, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  p->stats.normal_evals = normals;
  p->stats.ms_total = ms;
  p->stats.ms_trace = traceMs;
  p->stats.trace_steps = steps;
  p->stats.paths = (long long)frames * p->spf * st->npix;
  p->stats.trace_launches = traceLaunches;
  return SNL_OK;
}

// Persistent engine: per frame and pixel tile one launch of snl_persist; the
// tiles run on their own streams so one tile's tail overlaps another's work.
// The megakernel (snl_mega) and the AO / simplified integrators
// (snl_mode_kernel) launch the same way.
// Tile order of a queue-kernel render (megakernel or persistent): its first
// frame runs as a launch of its own that adds pixels' march steps to their
// 8x8 tiles; the tiles are then
// sorted by that cost, most first (a counting sort over log-spaced buckets,
// four per doubling), and the remaining frames take them in that order, so the
// longest pixels start first and the launch does not end on a few of them.
// Each pixel still runs its frames in order in one lane: results do not change.
constexpr int kCostBuckets = 128;
static_assert(kCostBuckets == kNoallocCostBuckets, "the allocation-free render's scratch holds the cost histogram");
__device__ inline int costBucket(int c) {
  int b = (int)(log2f((float)c + 1.0f) * 4.0f);
  b = b < 0 ? 0 : (b >= kCostBuckets ? kCostBuckets - 1 : b);
  return kCostBuckets - 1 - b;  // costliest first
}
__global__ void tile_zero_kernel(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void tile_hist_kernel(const int* cost, int n, int* hist) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&hist[costBucket(cost[i])], 1);
}
__global__ void tile_scan_kernel(int* hist) {  // exclusive prefix sum, in place
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int run = 0;
    for (int b = 0; b < kCostBuckets; ++b) {
      int h = hist[b];
      hist[b] = run;
      run += h;
    }
  }
}
__global__ void tile_scatter_kernel(const int* cost, int n, int* cursor, int* order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) order[atomicAdd(&cursor[costBucket(cost[i])], 1)] = i;
}

// whether a render of `frames` frames of `st` with `kern` takes its tiles in
// cost order, and the state's tile buffers for it (allocated on first use)
static bool tileOrdered(snl_prog p, snl_state st, hipFunction_t kern, int frames) {
  return (kern == p->k_mega || kern == p->k_persist) && frames >= 2 && st->w % 8 == 0 && st->h % 8 == 0 &&
         p->orderTiles;
}
static int ensureTileBuffers(snl_ctx ctx, snl_state st) {
  if (!st->tileCost) {
    size_t nt = (size_t)(st->w >> 3) * (st->h >> 3);
    HIPCHK(ctx, hipMalloc((void**)&st->tileCost, (2 * nt + kCostBuckets) * sizeof(int)));
  }
  return SNL_OK;
}
// frame 1 with tile costs, the tile order, then the other frames in that order,
// all on stream ts (a is the launch's argument block, its queue counter reset
// before each launch); tileBuf holds 2 * tiles + kCostBuckets ints
static int launchOrdered(snl_ctx ctx, snl_prog p, snl_state st, hipFunction_t kern, int blocks, hipStream_t ts,
                         SnlKArgs a, int frames, int* tileBuf) {
  const int nt = (st->w >> 3) * (st->h >> 3), g = (nt + 255) / 256;
  int* cost = tileBuf;
  int* hist = cost + nt;
  int* order = hist + kCostBuckets;
  void* args[] = {&a};
  hipLaunchKernelGGL(tile_zero_kernel, dim3(g), dim3(256), 0, ts, cost, nt);
  hipLaunchKernelGGL(tile_zero_kernel, dim3(1), dim3(kCostBuckets), 0, ts, hist, kCostBuckets);
  a.frames = 1;
  a.tileCost = cost;
  a.tileOrder = nullptr;
  HIPCHK(ctx, hipModuleLaunchKernel(p->k_persist_reset, 1, 1, 1, 64, 1, 1, 0, ts, args, nullptr));
  HIPCHK(ctx, hipModuleLaunchKernel(kern, blocks, 1, 1, 256, 1, 1, 0, ts, args, nullptr));
  hipLaunchKernelGGL(tile_hist_kernel, dim3(g), dim3(256), 0, ts, cost, nt, hist);
  hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(1), 0, ts, hist);
  hipLaunchKernelGGL(tile_scatter_kernel, dim3(g), dim3(256), 0, ts, cost, nt, hist, order);
  HIPCHK(ctx, hipGetLastError());
  a.frames = frames - 1;
  a.tileCost = nullptr;
  a.tileOrder = order;
  HIPCHK(ctx, hipModuleLaunchKernel(p->k_persist_reset, 1, 1, 1, 64, 1, 1, 0, ts, args, nullptr));
  HIPCHK(ctx, hipModuleLaunchKernel(kern, blocks, 1, 1, 256, 1, 1, 0, ts, args, nullptr));
  return SNL_OK;
}

// scratch (or null): the allocation-free render's counters (C_COUNT ints) and
// tile buffers, used in place of the state's; the render then runs one tile
// set on one stream
static int renderPersistent(snl_prog p, snl_state st, int frames, hipFunction_t kern, int blocksAll,
                            int* scratch = nullptr) {
  snl_ctx ctx = p->ctx;
  bool unit0 = false;
  int rc = prepareUnit0(p, st, unit0);
  if (rc) return rc;
  if ((rc = upload(p))) return rc;
  int K = std::max(1, std::min(p->tiles, SNL_MAX_TILES));
  // a frame must see the whole previous frame through the unit-0 snapshot,
  // so such programs run their frames in order on one stream
  if ((long long)K > st->npix || unit0 || scratch) K = 1;
  int* counters = scratch ? scratch : st->tileCnt;
  p->stats = snl_stats{};
  SnlKArgs base;
  fillArgs(p, st, base);
  base.chunk = p->persistChunk;
  base.shadeThresh = p->shadeThresh;
  // zero the tile counters (step counters included) on the main stream, then fork
  SnlKArgs z = base;
  z.vertex = -2;
  for (int k = 0; k < K; ++k) {
    z.cnt = counters + k * C_COUNT;
    if ((rc = launch(ctx, p->k_reset, 1, 64, z))) return rc;
  }
  HIPCHK(ctx, hipEventRecord(p->ev0, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->joinEvent, ctx->stream));
  int blocks = std::max(1, blocksAll / K);
  // queue-kernel renders of an 8x8-tiled image (all its chains) on one stream
  // take their tiles costliest first after the first frame (see tile_scatter_kernel)
  const bool ordered = K == 1 && !unit0 && tileOrdered(p, st, kern, frames);
  int* tileBuf = scratch ? scratch + C_COUNT : nullptr;
  if (ordered && !scratch) {
    if ((rc = ensureTileBuffers(ctx, st))) return rc;
    tileBuf = st->tileCost;
  }
  for (int k = 0; k < K; ++k) {
    hipStream_t ts = ctx->tileStream[k];
    HIPCHK(ctx, hipStreamWaitEvent(ts, ctx->joinEvent, 0));
    SnlKArgs a = base;
    a.cnt = counters + k * C_COUNT;
    // tile k: the 64-pixel stripes s with s % K == k (the last one may be short)
    long long stripes = (st->npix + 63) / 64, last = stripes - 1;
    long long mine = stripes / K + (k < stripes % K ? 1 : 0);
    long long items = mine * 64;
    if (mine > 0 && last % K == k) items -= stripes * 64 - st->npix;
    a.itemBase = 0;
    a.itemEnd = (int)items;
    a.tileIndex = k;
    a.tileCount = K;
    void* args[] = {&a};
    if (unit0) {
      // one launch per frame, each after a snapshot of the radiance
      for (int f = 0; f < frames; ++f) {
        if ((rc = copyUnit0(ctx, st, ts))) return rc;
        HIPCHK(ctx, hipModuleLaunchKernel(p->k_persist_reset, 1, 1, 1, 64, 1, 1, 0, ts, args, nullptr));
        HIPCHK(ctx, hipModuleLaunchKernel(kern, blocks, 1, 1, 256, 1, 1, 0, ts, args, nullptr));
      }
    } else if (ordered) {
      // frame 1 records the tiles' costs; the rest take the costliest tiles first
      if ((rc = launchOrdered(ctx, p, st, kern, blocks, ts, a, frames, tileBuf))) return rc;
    } else if (frames > 0) {
      // every pixel runs all its frames in one lane: one launch per tile
      a.frames = frames;
      HIPCHK(ctx, hipModuleLaunchKernel(p->k_persist_reset, 1, 1, 1, 64, 1, 1, 0, ts, args, nullptr));
      HIPCHK(ctx, hipModuleLaunchKernel(kern, blocks, 1, 1, 256, 1, 1, 0, ts, args, nullptr));
    }
    HIPCHK(ctx, hipEventRecord(ctx->tileEvent[k], ts));
    HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->tileEvent[k], 0));
  }
  HIPCHK(ctx, hipEventRecord(p->ev1, ctx->stream));
  HIPCHK(ctx, hipEventSynchronize(p->ev1));
  float ms = 0.0f;
  HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev0, p->ev1));
  long long steps = 0, normals = 0;
  for (int k = 0; k < K; ++k) {
    long long sk = 0, nk = 0;
    HIPCHK(ctx, hipMemcpyAsync(&sk, counters + k * C_COUNT + C_TRACE_STEPS, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(&nk, counters + k * C_COUNT + C_NORMALS, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    steps += sk;
    normals += nk;
  }
  p->stats.normal_evals = normals;
  p->stats.ms_total = ms;
  p->stats.ms_trace = ms;  // the persistent kernel is the whole frame
  p->stats.trace_steps = steps;
  p->stats.paths = (long long)frames * p->spf * st->npix;
  p->stats.trace_launches = unit0 ? frames * K : (ordered ? 2 : K);
  return SNL_OK;
}

int snl_state_error(snl_state st, double* out) {
  if (!st || !out) return fail(st ? st->ctx : nullptr, SNL_E_ARG, "NULL argument");
  if (int rc = settle(st)) return rc;
  if (st->chains < 2) return fail(st->ctx, SNL_E_ARG, "the error estimate needs a state of two or more chains");
  snl_ctx ctx = st->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int blocks = 256;
  // block partials and the result: allocated once per state, on its first estimate
  if (!st->errPartial && hipMalloc((void**)&st->errPartial, (2 * blocks + 2) * sizeof(double)) != hipSuccess)
    return fail(ctx, SNL_E_NOMEM, "error-estimate partials: allocation failed");
  double* d = st->errPartial;
  hipLaunchKernelGGL(chain_error_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float4*>(st->radiance), (long long)st->w * st->h, st->chains, d);
  hipLaunchKernelGGL(error_sum_kernel, dim3(1), dim3(64), 0, ctx->stream, d, blocks, d + 2 * blocks);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(out, d + 2 * blocks, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return SNL_OK;
}

int snl_ctx_mem_info(snl_ctx ctx, unsigned long long* free_bytes, unsigned long long* total_bytes) {
  if (!ctx) return fail(nullptr, SNL_E_ARG, "ctx is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  size_t f = 0, t = 0;
  HIPCHK(ctx, hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return SNL_OK;
}

unsigned long long snl_state_bytes(int width, int height, int chains) {
  if (width <= 0 || height <= 0) return 0ull;
  const unsigned long long npix = (unsigned long long)width * height * (unsigned long long)(chains > 0 ? chains : 1);
  const unsigned long long tiles = (unsigned long long)(width >> 3) * (unsigned long long)(height >> 3);
  return npix * 32ull                                        // radiance + RNG
         + (unsigned long long)(SNL_MAX_TILES + 1) * C_COUNT * 4ull  // queue counters
         + (2ull * tiles + kCostBuckets) * 4ull                 // tile costs, histogram, order
         + (2ull * 256 + 2) * 8ull;                             // error-estimate partials
}

unsigned long long snl_program_state_bytes(snl_prog p, int width, int height, int chains) {
  unsigned long long b = snl_state_bytes(width, height, chains);
  if (!p) return b;
  bool unit0 = false;
  for (int idx : p->sampled2D) unit0 |= !p->bound.count(idx);
  if (unit0)  // the per-frame radiance snapshot for samplers left on texture unit 0
    b += (unsigned long long)width * height * (unsigned long long)(chains > 0 ? chains : 1) * 16ull;
  return b;
}

// batched jobs run the megakernel, or the persistent kernel where that is
// chosen and the program needs neither the sub-surface walk nor the atmosphere
// (the wavefront engine's per-bounce launches are not batched: its jobs run the
// megakernel); AO and simplified programs their own kernel
static hipFunction_t queueKernel(snl_prog p, int& blocks) {
  if (p->mode) { blocks = p->modeBlocks; return p->k_mode; }
  if (p->engine == SNL_ENGINE_PERSISTENT && !p->megaOnly && !(p->hostP.subsurfaceMFP > 0.0f)) {
    blocks = p->persistBlocks;
    return p->k_persist;
  }
  blocks = p->megaBlocks;
  return p->k_mega;
}

static int enqueueBatch(snl_batch_s* b, const std::vector<int>& order, std::vector<double>& load, bool graph);

// the batch module launch's arguments (per fused job its kernel arguments and
// kind, then the jobs heaviest first), in device memory owned by the batch
static int fusedArgs(snl_batch_s* b) {
  snl_ctx ctx = b->ctx;
  const int n = (int)b->fusedJobs.size();
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
    double wx = (double)b->states[b->fusedJobs[x]]->npix * b->frames[b->fusedJobs[x]] * b->progs[b->fusedJobs[x]]->spf;
    double wy = (double)b->states[b->fusedJobs[y]]->npix * b->frames[b->fusedJobs[y]] * b->progs[b->fusedJobs[y]]->spf;
    double lx = (double)b->frames[b->fusedJobs[x]] * (b->progs[b->fusedJobs[x]]->maxBounces + 1) * 4.0e5;
    double ly = (double)b->frames[b->fusedJobs[y]] * (b->progs[b->fusedJobs[y]]->maxBounces + 1) * 4.0e5;
    return std::max(wx, lx) > std::max(wy, ly);
  });
  const size_t argBytes = sizeof(SnlKArgs) * n, intBytes = sizeof(int) * n;
  b->hFused.assign(argBytes + 2 * intBytes, 0);
  SnlKArgs* ha = reinterpret_cast<SnlKArgs*>(b->hFused.data());
  int* hk = reinterpret_cast<int*>(b->hFused.data() + argBytes);
  int* ho = hk + n;
  for (int i = 0; i < n; ++i) {
    const int j = b->fusedJobs[i];
    snl_prog p = b->progs[j];
    snl_state st = b->states[j];
    fillArgs(p, st, ha[i]);
    ha[i].chunk = p->persistChunk;
    ha[i].shadeThresh = p->shadeThresh;
    ha[i].cnt = st->tileCnt;
    ha[i].itemBase = 0;
    ha[i].itemEnd = (int)st->npix;
    ha[i].frames = b->frames[j];
    ha[i].tileCost = nullptr;
    ha[i].tileOrder = nullptr;
    hk[i] = b->fm->kindOf.at(p->part.key);
    ho[i] = ord[i];
  }
  HIPCHK(ctx, hipMalloc(&b->dFused, b->hFused.size()));
  HIPCHK(ctx, hipMemcpyAsync(b->dFused, b->hFused.data(), b->hFused.size(), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return SNL_OK;
}

// Everything a batch needs before its launches: checks, uniform uploads, the
// per-job counters zeroed (graph == false: on the main stream; a captured
// batch zeroes them in its own launch sequence), streams, tile buffers, the
// jobs a batch module runs, events, the jobs' order and the streams' loads.
static int prepareBatch(int njobs, const snl_prog* progs, const snl_state* states, const int* frames, bool graph,
                        std::unique_ptr<snl_batch_s>& b, std::vector<int>& order, std::vector<double>& load) {
  if (njobs <= 0 || !progs || !states || !frames) return fail(nullptr, SNL_E_ARG, "bad batch");
  snl_ctx ctx = progs[0] ? progs[0]->ctx : nullptr;
  if (!ctx) return fail(nullptr, SNL_E_ARG, "NULL program in batch");
  // a program's uniforms (resolution, unit-0 snapshot) belong to one job at a
  // time, as a state's buffers do: each may appear once
  std::set<snl_state> seen;
  std::set<snl_prog> seenProg;
  for (int j = 0; j < njobs; ++j) {
    if (!progs[j] || !states[j] || frames[j] < 0) return fail(ctx, SNL_E_ARG, "bad job %d", j);
    if (progs[j]->ctx != ctx || states[j]->ctx != ctx) return fail(ctx, SNL_E_ARG, "job %d: other context", j);
    if (!seen.insert(states[j]).second) return fail(ctx, SNL_E_ARG, "job %d: state used twice", j);
    if (!seenProg.insert(progs[j]).second)
      return fail(ctx, SNL_E_ARG, "job %d: program used twice (its uniforms are per job; build one per job)", j);
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc;
  // a program or state still rendering in an earlier batch finishes first
  for (int j = 0; j < njobs; ++j)
    if ((rc = settle(progs[j])) || (rc = settle(states[j]))) return rc;
  // parameters up, per-job counters zeroed, all on the main stream
  b.reset(new snl_batch_s);
  b->ctx = ctx;
  b->njobs = njobs;
  b->progs.assign(progs, progs + njobs);
  b->states.assign(states, states + njobs);
  b->frames.assign(frames, frames + njobs);
  b->unit0.assign(njobs, 0);
  b->ordered.assign(njobs, 0);
  b->done.assign(njobs, 0);
  b->stream.assign(njobs, 0);
  b->work.assign(njobs, 0.0);
  b->steps.assign(njobs, 0);
  b->normals.assign(njobs, 0);
  b->ms.assign(njobs, 0.0f);
  for (int j = 0; j < njobs; ++j) {
    bool u = false;
    if ((rc = prepareUnit0(progs[j], states[j], u))) return rc;
    if (u && graph)
      return fail(ctx, SNL_E_UNSUPPORTED, "job %d: a program that samples texture unit 0 needs a snapshot per "
                  "frame, which a captured batch does not take", j);
    b->unit0[j] = u;
    if ((rc = upload(progs[j]))) return rc;
    if (graph) continue;
    SnlKArgs z;
    fillArgs(progs[j], states[j], z);
    z.vertex = -2;
    z.cnt = states[j]->tileCnt;
    if ((rc = launch(ctx, progs[j]->k_reset, 1, 64, z))) return rc;
  }
  if (ctx->batchStream.empty()) {
    int n = 16;
    if (const char* e = getenv("SNL_BATCH_STREAMS")) n = std::max(1, std::min(32, atoi(e)));
    ctx->batchStream.resize(n);
    ctx->batchEvent.resize(n);
    for (int k = 0; k < n; ++k) {
      HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->batchStream[k], hipStreamNonBlocking));
      HIPCHK(ctx, hipEventCreateWithFlags(&ctx->batchEvent[k], hipEventDisableTiming));
    }
  }
  // queue-kernel jobs of an 8x8-tiled image take their tiles in cost order;
  // their tile buffers are allocated before anything is launched
  for (int j = 0; j < njobs; ++j) {
    int resident = 0;
    b->ordered[j] = !b->unit0[j] && tileOrdered(progs[j], states[j], queueKernel(progs[j], resident), frames[j]);
    if (b->ordered[j] && (rc = ensureTileBuffers(ctx, states[j]))) return rc;
  }
  // the jobs that one loaded batch module can run together: megakernel
  // programs of its build group whose kinds it holds (the module covering the
  // most jobs; fewer than two gain nothing)
  if (ctx->batchFusion && !ctx->fused.empty()) {
    std::vector<int> cand;
    for (int j = 0; j < njobs; ++j) {
      int resident = 0;
      if (progs[j]->part.fusable && !b->unit0[j] && frames[j] > 0 &&
          queueKernel(progs[j], resident) == progs[j]->k_mega)
        cand.push_back(j);
    }
    for (auto& kv : ctx->fused) {
      FusedModule* m = kv.second;
      std::vector<int> got;
      for (int j : cand)
        if (progs[j]->part.sameGroup(m->group) && m->kindOf.count(progs[j]->part.key)) got.push_back(j);
      if (got.size() > b->fusedJobs.size()) { b->fm = m; b->fusedJobs = got; }
    }
    if (b->fusedJobs.size() < 2) { b->fm = nullptr; b->fusedJobs.clear(); }
    for (int j : b->fusedJobs) b->ordered[j] = 0;  // one launch in queue order, jobs ordered by cost
  }
  if (b->fm && (rc = fusedArgs(b.get()))) return rc;
  b->ev.assign(njobs, nullptr);
  HIPCHK(ctx, hipEventCreate(&b->ev0));
  for (int j = 0; j < njobs; ++j) HIPCHK(ctx, hipEventCreate(&b->ev[j]));
  // Longest job first onto the least loaded stream.  A job lasts about the
  // larger of its throughput time (all its path samples at a device-wide rate)
  // and its latency (its frames run back to back in one lane: frames times a
  // deep path's time, ~bounces); both scales are coarse and only order jobs.
  // A stream's load counts the unfinished jobs of earlier batches queued on it.
  order.assign(njobs, 0);
  for (int j = 0; j < njobs; ++j) {
    order[j] = j;
    double samples = (double)states[j]->npix * frames[j] * progs[j]->spf;
    double latency = (double)frames[j] * progs[j]->spf * (progs[j]->maxBounces + 1) * 4.0e5;
    b->work[j] = std::max(samples, latency);
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->work[x] > b->work[y]; });
  const int NS = (int)ctx->batchStream.size();
  load.assign(NS, 0.0);
  for (snl_batch_s* o : ctx->batches)
    for (int j = 0; j < o->njobs; ++j) {
      if ((rc = jobQuery(o, j)) < 0) return rc;
      if (rc == 0) load[o->stream[j]] += o->work[j];
    }
  return SNL_OK;
}

int snl_batch_submit(int njobs, const snl_prog* progs, const snl_state* states, const int* frames,
                     snl_batch* out) {
  if (!out) return fail(nullptr, SNL_E_ARG, "out is NULL");
  *out = nullptr;
  std::unique_ptr<snl_batch_s> b;
  std::vector<int> order;
  std::vector<double> load;
  int rc = prepareBatch(njobs, progs, states, frames, false, b, order, load);
  if (rc) return rc;
  snl_ctx ctx = b->ctx;
  const int NS = (int)ctx->batchStream.size();
  // on a failure part way, what was queued finishes before the batch is dropped
  auto drop = [&](int code) {
    for (int k = 0; k < NS; ++k) (void)hipStreamSynchronize(ctx->batchStream[k]);
    for (hipEvent_t e : b->ev)
      if (e) (void)hipEventDestroy(e);
    (void)hipEventDestroy(b->ev0);
    if (b->dFused) (void)hipFree(b->dFused);
    return code;
  };
  if ((rc = enqueueBatch(b.get(), order, load, false))) return drop(rc);
  for (int j = 0; j < njobs; ++j) {
    progs[j]->busy = states[j]->busy = b.get();
    progs[j]->busyJob = states[j]->busyJob = j;
  }
  b->pending = njobs;
  ctx->batches.insert(b.get());
  *out = b.release();
  return SNL_OK;
}

// the launches of a submitted batch, jobs in `order`, each onto the least
// loaded batch stream, with the job's end event after them
// (graph: the launches are being captured from ctx->graphStream -- no events
// per job, and each stream-launched job zeroes its counters first)
static int enqueueBatch(snl_batch_s* b, const std::vector<int>& order, std::vector<double>& load, bool graph) {
  snl_ctx ctx = b->ctx;
  const int NS = (int)ctx->batchStream.size();
  int rc;
  hipStream_t origin = graph ? ctx->graphStream : ctx->stream;
  if (!graph) HIPCHK(ctx, hipEventRecord(b->ev0, origin));
  HIPCHK(ctx, hipEventRecord(ctx->joinEvent, origin));
  for (int k = 0; k < NS; ++k) HIPCHK(ctx, hipStreamWaitEvent(ctx->batchStream[k], ctx->joinEvent, 0));
  const int* frames = b->frames.data();
  std::vector<char> isFused(b->njobs, 0);
  // the persistent launch first: launched after the stream jobs it ran slower
  // (0.585 against 0.451 s on config 4)
  if (b->fm) {
    // one persistent launch over the module's jobs, heaviest first: its waves
    // take pixels from one job until that job's queue runs dry, then move on
    const int n = (int)b->fusedJobs.size();
    double w = 0.0;
    long long blocksWanted = 0;
    for (int i = 0; i < n; ++i) {
      isFused[b->fusedJobs[i]] = 1;
      w += b->work[b->fusedJobs[i]];
      blocksWanted += (b->states[b->fusedJobs[i]]->npix + 255) / 256;
    }
    const size_t argBytes = sizeof(SnlKArgs) * n, intBytes = sizeof(int) * n;
    int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[k] += w;
    hipStream_t ts = ctx->batchStream[k];
    struct {
      const void* jobs;
      const void* kind;
      const void* order;
      int njobs;
    } fa = {b->dFused, (const char*)b->dFused + argBytes, (const char*)b->dFused + argBytes + intBytes, n};
    void* fargs[] = {&fa};
    HIPCHK(ctx, hipModuleLaunchKernel(b->fm->k_reset, n, 1, 1, 64, 1, 1, 0, ts, fargs, nullptr));
    // SNL_FUSED_GRID: percent of the resident grid the launch takes (default
    // 100); less leaves room for the jobs on streams beside it
    long long grid = b->fm->blocks;
    if (const char* e = getenv("SNL_FUSED_GRID")) grid = grid * std::max(1, std::min(100, atoi(e))) / 100;
    const int blocks = (int)std::max(1LL, std::min<long long>(grid, blocksWanted));
    HIPCHK(ctx, hipModuleLaunchKernel(b->fm->k_run, blocks, 1, 1, 256, 1, 1, 0, ts, fargs, nullptr));
    for (int i = 0; i < n; ++i) {
      b->stream[b->fusedJobs[i]] = k;
      if (!graph) HIPCHK(ctx, hipEventRecord(b->ev[b->fusedJobs[i]], ts));
    }
  }
  for (int j : order) {
    if (isFused[j]) continue;
    int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[k] += b->work[j];
    b->stream[j] = k;
    snl_prog p = b->progs[j];
    snl_state st = b->states[j];
    hipStream_t ts = ctx->batchStream[k];
    SnlKArgs a;
    fillArgs(p, st, a);
    a.chunk = p->persistChunk;
    a.shadeThresh = p->shadeThresh;
    a.cnt = st->tileCnt;
    a.itemBase = 0;
    a.itemEnd = (int)st->npix;
    // one lane per pixel up to the program's full persistent grid; the
    // megakernel takes sub-surface and atmosphere programs
    int resident = 0;
    hipFunction_t kern = queueKernel(p, resident);
    long long want = (st->npix + 255) / 256;
    int blocks = (int)std::max(1LL, std::min<long long>(resident, want));
    void* args[] = {&a};
    if (graph) {  // the counters, zeroed in the launch sequence itself
      SnlKArgs z = a;
      z.vertex = -2;
      void* zargs[] = {&z};
      HIPCHK(ctx, hipModuleLaunchKernel(p->k_reset, 1, 1, 1, 64, 1, 1, 0, ts, zargs, nullptr));
    }
    if (b->ordered[j]) {  // the costliest tiles first after frame 1, as in snl_render
      if ((rc = launchOrdered(ctx, p, st, kern, blocks, ts, a, frames[j], st->tileCost))) return rc;
    } else {
      int launches = b->unit0[j] ? frames[j] : (frames[j] > 0 ? 1 : 0);
      a.frames = b->unit0[j] ? 1 : frames[j];
      for (int f = 0; f < launches; ++f) {
        if (b->unit0[j] && (rc = copyUnit0(ctx, st, ts))) return rc;
        HIPCHK(ctx, hipModuleLaunchKernel(p->k_persist_reset, 1, 1, 1, 64, 1, 1, 0, ts, args, nullptr));
        HIPCHK(ctx, hipModuleLaunchKernel(kern, blocks, 1, 1, 256, 1, 1, 0, ts, args, nullptr));
      }
    }
    if (!graph) HIPCHK(ctx, hipEventRecord(b->ev[j], ts));
  }
  return SNL_OK;
}

int snl_job_poll(snl_batch b, int job) {
  if (!b) return fail(nullptr, SNL_E_ARG, "batch is NULL");
  if (job < -1 || job >= b->njobs) return fail(b->ctx, SNL_E_ARG, "job %d of a %d-job batch", job, b->njobs);
  HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
  if (job >= 0) return jobQuery(b, job);
  for (int j = 0; j < b->njobs; ++j) {
    int rc = jobQuery(b, j);
    if (rc < 0) return rc;
  }
  return b->pending == 0 ? 1 : 0;
}

int snl_batch_wait(snl_batch b, snl_stats* stats) {
  if (!b) return fail(nullptr, SNL_E_ARG, "batch is NULL");
  snl_ctx ctx = b->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  for (int j = 0; j < b->njobs; ++j) {
    int rc = jobWait(b, j);
    if (rc) return rc;
  }
  if (stats) {
    *stats = snl_stats{};
    for (int j = 0; j < b->njobs; ++j) {
      stats->ms_total = std::max(stats->ms_total, (double)b->ms[j]);
      stats->trace_steps += b->steps[j];
      stats->normal_evals += b->normals[j];
      stats->paths += (long long)b->frames[j] * b->progs[j]->spf * b->states[j]->npix;
      stats->trace_launches += b->unit0[j] ? b->frames[j] : (b->ordered[j] ? 2 : 1);
    }
    if (b->fm) stats->trace_launches -= (int)b->fusedJobs.size() - 1;  // one launch for all of them
    stats->ms_trace = stats->ms_total;
  }
  return SNL_OK;
}

int snl_batch_job_stats(snl_batch b, int job, snl_stats* out) {
  if (!b || !out) return fail(b ? b->ctx : nullptr, SNL_E_ARG, "NULL argument");
  if (job < 0 || job >= b->njobs) return fail(b->ctx, SNL_E_ARG, "job %d of a %d-job batch", job, b->njobs);
  HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
  int rc = jobWait(b, job);
  if (rc) return rc;
  *out = snl_stats{};
  out->ms_total = out->ms_trace = b->ms[job];
  out->trace_steps = b->steps[job];
  out->normal_evals = b->normals[job];
  out->paths = (long long)b->frames[job] * b->progs[job]->spf * b->states[job]->npix;
  out->trace_launches = b->unit0[job] ? b->frames[job] : (b->ordered[job] ? 2 : 1);
  return SNL_OK;
}

int snl_batch_release(snl_batch b) {
  if (!b) return SNL_OK;
  snl_ctx ctx = b->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = SNL_OK;
  for (int j = 0; j < b->njobs; ++j)
    if (!rc) rc = jobWait(b, j);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->dFused) (void)hipFree(b->dFused);
  ctx->batches.erase(b);
  delete b;
  return rc;
}

int snl_render_batch(int njobs, const snl_prog* progs, const snl_state* states, const int* frames,
                     snl_stats* stats) {
  if (stats) *stats = snl_stats{};
  if (njobs == 0) return SNL_OK;
  snl_batch b = nullptr;
  int rc = snl_batch_submit(njobs, progs, states, frames, &b);
  if (rc) return rc;
  rc = snl_batch_wait(b, stats);
  int rc2 = snl_batch_release(b);
  return rc ? rc : rc2;
}

// A batch captured as a hipGraph: its launch sequence (counters zeroed by
// kernels, tile-order passes, the batch module's persistent launch, the other
// jobs' launches over the batch streams) recorded once and replayed.
struct snl_graph_s {
  snl_ctx ctx = nullptr;
  std::unique_ptr<snl_batch_s> b;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~snl_graph_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (b) {
      for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
      if (b->ev0) (void)hipEventDestroy(b->ev0);
      if (b->dFused) (void)hipFree(b->dFused);
    }
  }
};

int snl_batch_graph_create(int njobs, const snl_prog* progs, const snl_state* states, const int* frames,
                           snl_graph* out) {
  if (!out) return fail(nullptr, SNL_E_ARG, "out is NULL");
  *out = nullptr;
  std::unique_ptr<snl_graph_s> g(new snl_graph_s);
  std::vector<int> order;
  std::vector<double> load;
  int rc = prepareBatch(njobs, progs, states, frames, true, g->b, order, load);
  if (rc) return rc;
  snl_ctx ctx = g->ctx = g->b->ctx;
  if (!ctx->graphStream) HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->graphStream, hipStreamNonBlocking));
  HIPCHK(ctx, hipEventCreate(&g->e0));
  HIPCHK(ctx, hipEventCreate(&g->e1));
  HIPCHK(ctx, hipStreamBeginCapture(ctx->graphStream, hipStreamCaptureModeThreadLocal));
  rc = enqueueBatch(g->b.get(), order, load, true);
  // every batch stream joins back into the capture's origin
  for (size_t k = 0; !rc && k < ctx->batchStream.size(); ++k) {
    if (hipEventRecord(ctx->batchEvent[k], ctx->batchStream[k]) != hipSuccess ||
        hipStreamWaitEvent(ctx->graphStream, ctx->batchEvent[k], 0) != hipSuccess)
      rc = fail(ctx, SNL_E_HIP, "joining batch stream %d into the capture failed", (int)k);
  }
  hipError_t e = hipStreamEndCapture(ctx->graphStream, &g->graph);
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, SNL_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  HIPCHK(ctx, hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  return SNL_OK;
}

int snl_batch_graph_launch(snl_graph g, snl_stats* stats) {
  if (!g) return fail(nullptr, SNL_E_ARG, "graph is NULL");
  snl_ctx ctx = g->ctx;
  snl_batch_s* b = g->b.get();
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc;
  // what the jobs' programs and states may still be doing finishes first; the
  // programs' current uniforms go up (the graph reads them where they live)
  for (int j = 0; j < b->njobs; ++j)
    if ((rc = settle(b->progs[j])) || (rc = settle(b->states[j])) || (rc = upload(b->progs[j]))) return rc;
  HIPCHK(ctx, hipEventRecord(g->e0, ctx->stream));
  HIPCHK(ctx, hipGraphLaunch(g->exec, ctx->stream));
  HIPCHK(ctx, hipEventRecord(g->e1, ctx->stream));
  HIPCHK(ctx, hipEventSynchronize(g->e1));
  if (stats) {
    *stats = snl_stats{};
    float ms = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, g->e0, g->e1));
    stats->ms_total = stats->ms_trace = ms;
    for (int j = 0; j < b->njobs; ++j) {
      long long sk = 0, nk = 0;
      HIPCHK(ctx, hipMemcpyAsync(&sk, b->states[j]->tileCnt + C_TRACE_STEPS, 8, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(ctx, hipMemcpyAsync(&nk, b->states[j]->tileCnt + C_NORMALS, 8, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      stats->trace_steps += sk;
      stats->normal_evals += nk;
      stats->paths += (long long)b->frames[j] * b->progs[j]->spf * b->states[j]->npix;
      stats->trace_launches += b->ordered[j] ? 2 : 1;
    }
    if (b->fm) stats->trace_launches -= (int)b->fusedJobs.size() - 1;
  }
  return SNL_OK;
}

int snl_batch_graph_destroy(snl_graph g) {
  if (!g) return SNL_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
  return SNL_OK;
}

int snl_batch_module(snl_ctx ctx, int nprogs, const snl_prog* progs) {
  if (!ctx || nprogs <= 0 || !progs) return fail(ctx, SNL_E_ARG, "bad argument");
  std::vector<const FusedPart*> kinds;
  for (int i = 0; i < nprogs; ++i) {
    if (!progs[i] || progs[i]->ctx != ctx) return fail(ctx, SNL_E_ARG, "program %d: NULL or another context", i);
    kinds.push_back(&progs[i]->part);
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  FusedModule* m = nullptr;
  return buildFused(ctx, kinds, m);
}

int snl_batch_module_precompile(int nprogs, const char* const* shaders, const char* const* iors,
                                const char* const* defines, const int* maxMarch, const int* maxAtm, unsigned flags) {
  if (nprogs <= 0 || !shaders || !iors || !defines || !maxMarch || !maxAtm)
    return fail(nullptr, SNL_E_ARG, "bad argument");
  std::vector<FusedPart> parts(nprogs);
  for (int i = 0; i < nprogs; ++i) {
    Compiled c;
    int rc = compileProgram(nullptr, shaders[i], iors[i], defines[i], maxMarch[i], 1, maxAtm[i], 1, flags, c,
                            nullptr, 0);
    if (rc) return rc;
    parts[i] = partOf(c, maxMarch[i], flags);
  }
  std::vector<const FusedPart*> kinds;
  for (auto& f : parts) kinds.push_back(&f);
  std::vector<char> code;
  std::string source;
  return compileFused(nullptr, kinds, code, source);
}

int snl_ctx_set_batch_fusion(snl_ctx ctx, int on) {
  if (!ctx) return fail(nullptr, SNL_E_ARG, "ctx is NULL");
  ctx->batchFusion = on != 0;
  return SNL_OK;
}

int snl_render(snl_prog p, snl_state st, int frames) {
  if (!p || !st || frames < 0) return fail(p ? p->ctx : nullptr, SNL_E_ARG, "bad argument");
  if (st->ctx != p->ctx) return fail(p->ctx, SNL_E_ARG, "program and state belong to different contexts");
  if (int rc = settle(st)) return rc;
  // only the megakernel renders the atmosphere; the persistent kernel does not
  // run the sub-surface walk, such programs use the wavefront unless the
  // megakernel is asked for
  if (p->mode) return renderPersistent(p, st, frames, p->k_mode, p->modeBlocks);
  if (p->megaOnly || p->engine == SNL_ENGINE_MEGAKERNEL) return renderPersistent(p, st, frames, p->k_mega, p->megaBlocks);
  bool needWavefront = p->hostP.subsurfaceMFP > 0.0f;
  if (p->engine == SNL_ENGINE_PERSISTENT && !needWavefront)
    return renderPersistent(p, st, frames, p->k_persist, p->persistBlocks);
  return renderWavefront(p, st, frames);
}

int snl_render_noalloc(snl_prog p, snl_state st, int frames) {
  if (!p || !st || frames < 0) return fail(p ? p->ctx : nullptr, SNL_E_ARG, "bad argument");
  snl_ctx ctx = p->ctx;
  if (st->ctx != ctx) return fail(ctx, SNL_E_ARG, "program and state belong to different contexts");
  if (int rc = settle(st)) return rc;
  const long long tiles = (long long)(st->w >> 3) * (st->h >> 3);
  if (tiles > SNL_NOALLOC_MAX_TILES)
    return fail(ctx, SNL_E_ARG,
                "snl_render_noalloc: a %d x %d image has %lld 8x8 tiles, its scratch holds at most %d "
                "(a 4096 x 4096 image); render it with snl_render",
                st->w, st->h, tiles, SNL_NOALLOC_MAX_TILES);
  for (int idx : p->sampled2D)
    if (!p->bound.count(idx))
      return fail(ctx, SNL_E_UNSUPPORTED,
                  "snl_render_noalloc: the program samples texture unit 0 (an unbound sampler2D), whose "
                  "radiance snapshot is an allocation; render it with snl_render");
  if (!ctx->scratch) return fail(ctx, SNL_E_HIP, "snl_render_noalloc: scratch address not resolved");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // the kernel snl_render picks; the wavefront engine's path buffers are allocations
  if (p->mode) return renderPersistent(p, st, frames, p->k_mode, p->modeBlocks, ctx->scratch);
  if (p->megaOnly || p->engine == SNL_ENGINE_MEGAKERNEL)
    return renderPersistent(p, st, frames, p->k_mega, p->megaBlocks, ctx->scratch);
  if (p->engine == SNL_ENGINE_PERSISTENT && !(p->hostP.subsurfaceMFP > 0.0f))
    return renderPersistent(p, st, frames, p->k_persist, p->persistBlocks, ctx->scratch);
  return fail(ctx, SNL_E_UNSUPPORTED,
              "snl_render_noalloc: this program renders with the wavefront engine, whose path buffers are "
              "allocations; choose the megakernel or the persistent engine");
}

int snl_set_engine(snl_prog p, int engine, int tiles, int persist_blocks, int shade_thresh, int chunk) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  if (engine != SNL_ENGINE_WAVEFRONT && engine != SNL_ENGINE_PERSISTENT && engine != SNL_ENGINE_MEGAKERNEL)
    return fail(p->ctx, SNL_E_ARG, "bad engine");
  p->engine = engine;
  if (tiles > 0) p->tiles = std::min(tiles, SNL_MAX_TILES);
  if (persist_blocks > 0) (engine == SNL_ENGINE_MEGAKERNEL ? p->megaBlocks : p->persistBlocks) = persist_blocks;
  if (shade_thresh > 0) p->shadeThresh = std::min(shade_thresh, 64);
  if (chunk > 0) p->persistChunk = chunk;
  return SNL_OK;
}

int snl_mega_variants(snl_prog p) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  return p->spillCode.empty() && !p->spillMod ? 1 : 2;
}

int snl_set_mega_variant(snl_prog p, int variant) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  if (variant != 0 && variant != 1) return fail(p->ctx, SNL_E_ARG, "bad megakernel variant %d", variant);
  if (variant == 1 && !p->spillMod) {
    if (p->spillCode.empty())
      return fail(p->ctx, SNL_E_UNSUPPORTED, "the program's megakernel has one build (four waves per SIMD do not spill)");
    HIPCHK(p->ctx, hipSetDevice(p->ctx->device));
    HIPCHK(p->ctx, hipModuleLoadData(&p->spillMod, p->spillCode.data()));
    HIPCHK(p->ctx, hipModuleGetFunction(&p->k_megaSpill, p->spillMod, "snl_mega"));
    std::vector<char>().swap(p->spillCode);
  }
  p->k_mega = variant ? p->k_megaSpill : p->k_megaMain;
  p->megaVariant = variant;
  return SNL_OK;
}

int snl_render_normals(snl_prog p, snl_state st, int frames) {
  if (!p || !st || frames < 0) return fail(p ? p->ctx : nullptr, SNL_E_ARG, "bad argument");
  int rc = settle(st);
  if (rc || (rc = upload(p))) return rc;
  SnlKArgs a;
  fillArgs(p, st, a);
  for (int f = 0; f < frames; ++f)
    if ((rc = launch(p->ctx, p->k_normals, blocksFor(st->npix), 256, a))) return rc;
  HIPCHK(p->ctx, hipStreamSynchronize(p->ctx->stream));
  return SNL_OK;
}

int snl_pick(snl_prog p, const float* xy, int n, float* outp) {
  if (!p || !xy || !outp || n < 0) return fail(p ? p->ctx : nullptr, SNL_E_ARG, "bad argument");
  if (n == 0) return SNL_OK;
  int rc = upload(p);
  if (rc) return rc;
  snl_ctx ctx = p->ctx;
  float *dxy = nullptr, *dout = nullptr;
  HIPCHK(ctx, hipMalloc((void**)&dxy, (size_t)n * 8));
  if (hipMalloc((void**)&dout, (size_t)n * 8) != hipSuccess) { hipFree(dxy); return fail(ctx, SNL_E_NOMEM, "pick buffers"); }
  SnlKArgs a;
  memset(&a, 0, sizeof a);
  a.P = p->dP;
  a.U = p->dU;
  void* args[] = {&a, &dxy, &dout, &n};
  // every copy goes through the context's (non-blocking) stream
  hipError_t e = hipMemcpyAsync(dxy, xy, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipModuleLaunchKernel(p->k_pick, blocksFor(n), 1, 1, 256, 1, 1, 0, ctx->stream, args, nullptr);
  if (e == hipSuccess) e = hipMemcpyAsync(outp, dout, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(dxy);
  hipFree(dout);
  if (e != hipSuccess) return fail(ctx, SNL_E_HIP, "pick: %s", hipGetErrorString(e));
  return SNL_OK;
}

int snl_tonemap(snl_prog p, snl_state st, float* rgb, unsigned char* rgba8) {
  if (!p || !st) return fail(p ? p->ctx : nullptr, SNL_E_ARG, "bad argument");
  int rc = settle(st);
  if (rc || (rc = upload(p))) return rc;
  snl_ctx ctx = p->ctx;
  long long n = st->npix;
  float* drgb = nullptr;
  unsigned char* d8 = nullptr;
  if (rgb) HIPCHK(ctx, hipMalloc((void**)&drgb, (size_t)n * 12));
  if (rgba8 && hipMalloc((void**)&d8, (size_t)n * 4) != hipSuccess) { if (drgb) hipFree(drgb); return fail(ctx, SNL_E_NOMEM, "tonemap buffer"); }
  int npix = (int)n;
  const SnlParams* dP = p->dP;
  float* rad = st->radiance;
  void* args[] = {&dP, &rad, &drgb, &d8, &npix};
  hipError_t e = hipModuleLaunchKernel(p->k_tonemap, blocksFor(n), 1, 1, 256, 1, 1, 0, ctx->stream, args, nullptr);
  if (e == hipSuccess && rgb) e = hipMemcpyAsync(rgb, drgb, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && rgba8) e = hipMemcpyAsync(rgba8, d8, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (drgb) hipFree(drgb);
  if (d8) hipFree(d8);
  if (e != hipSuccess) return fail(ctx, SNL_E_HIP, "tonemap: %s", hipGetErrorString(e));
  return SNL_OK;
}

int snl_get_stats(snl_prog p, snl_stats* out) {
  if (!p || !out) return fail(nullptr, SNL_E_ARG, "bad argument");
  *out = p->stats;
  return SNL_OK;
}

int snl_set_tuning(snl_prog p, int traceBlocks, int traceChunk, int shadeBlocks) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  if (traceBlocks > 0) p->traceBlocks = traceBlocks;
  if (traceChunk > 0) p->traceChunk = traceChunk;
  if (shadeBlocks > 0) p->shadeBlocks = shadeBlocks;
  return SNL_OK;
}

int snl_set_march_burst(snl_prog p, int steps) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  if (steps < 1 || steps > 64) return fail(p->ctx, SNL_E_ARG, "march burst must be 1..64");
  p->marchBurst = steps;
  return SNL_OK;
}

int snl_set_profiling(snl_prog p, int on) {
  if (!p) return fail(nullptr, SNL_E_ARG, "prog is NULL");
  p->profiling = on != 0;
  return SNL_OK;
}

}  // extern "C"
