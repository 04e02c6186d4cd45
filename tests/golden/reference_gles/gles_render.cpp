// gles_render.cpp -- runs a GLSL ES 3.00 fragment program over a W x H pixel
// grid with a GLES 3 implementation loaded at run time (dlopen), the way the
// reference's Renderer drives WebGL2: a full-screen quad (js/renderer.js:171-184),
// two RGBA32F state textures read at units 0/1 and written through a two-target
// framebuffer, ping-ponged each frame (js/renderer.js:586-616, 692-733).
//
// It is test tooling: make_fixtures.py uses it in the development container to
// turn the reference's shaders into golden vectors.  It never ships.
//
//   gles_render GLES_LIB_DIR SPEC
//
// SPEC is a text file of directives (one per line):
//   size W H | frames N | vertex PATH | fragment PATH
//   state RAD_IN RNG_IN                    initial state (W*H*4 float32 each)
//   texture UNIT NAME W H CH LINEAR CLAMP PATH   bind a float texture at UNIT and
//                                          point sampler NAME at it ('-': no sampler)
//   texture_srgb8 UNIT NAME W H PATH       bind W*H RGB bytes as an SRGB8 texture (LINEAR,
//                                          CLAMP_TO_EDGE), as GLU.loadImageAndCreateTextureInfo
//                                          uploads an image (js/gl.js:532-559)
//   u1i NAME v | u1f NAME v | u2f NAME a b | u3f NAME a b c | u4f NAME a b c d
//   u1fv NAME n v... | u3fv NAME n v...
//   outputs 1|2                            render targets (2: radiance + rng)
//   pick                                   one draw per pixel into a 1x1 viewport
//                                          with uniform mousePick = pixel centre
//                                          (the pick program reads no state)
//   out RAD_OUT [RNG_OUT]
// Uniform values are parsed as doubles and rounded to float, as WebGL does with
// JavaScript numbers.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

// ------------------------------------------------------------------ EGL / GLES
// The handful of entry points and enumerants used here (Khronos EGL 1.4 and
// OpenGL ES 3.0 specifications).
typedef void* EGLDisplay;
typedef void* EGLConfig;
typedef void* EGLContext;
typedef void* EGLSurface;
typedef int EGLint;
typedef unsigned EGLBoolean;
typedef unsigned GLenum;
typedef unsigned GLuint;
typedef int GLint;
typedef int GLsizei;
typedef float GLfloat;
typedef unsigned char GLboolean;
typedef char GLchar;
typedef long GLsizeiptr;

enum : EGLint {
  EGL_NONE_ = 0x3038, EGL_RENDERABLE_TYPE_ = 0x3040, EGL_OPENGL_ES3_BIT_ = 0x40, EGL_SURFACE_TYPE_ = 0x3033,
  EGL_PBUFFER_BIT_ = 0x1, EGL_CONTEXT_CLIENT_VERSION_ = 0x3098, EGL_OPENGL_ES_API_ = 0x30A0,
  EGL_WIDTH_ = 0x3057, EGL_HEIGHT_ = 0x3056
};
enum : GLenum {
  GL_FRAGMENT_SHADER_ = 0x8B30, GL_VERTEX_SHADER_ = 0x8B31, GL_COMPILE_STATUS_ = 0x8B81, GL_LINK_STATUS_ = 0x8B82,
  GL_TEXTURE_2D_ = 0x0DE1, GL_TEXTURE0_ = 0x84C0, GL_RGBA32F_ = 0x8814, GL_R32F_ = 0x822E, GL_RG32F_ = 0x8230,
  GL_RGB32F_ = 0x8815, GL_SRGB8_ = 0x8C41, GL_UNSIGNED_BYTE_ = 0x1401, GL_RGBA_ = 0x1908, GL_RGB_ = 0x1907, GL_RG_ = 0x8227, GL_RED_ = 0x1903, GL_FLOAT_ = 0x1406,
  GL_FRAMEBUFFER_ = 0x8D40, GL_COLOR_ATTACHMENT0_ = 0x8CE0, GL_FRAMEBUFFER_COMPLETE_ = 0x8CD5,
  GL_TEXTURE_MIN_FILTER_ = 0x2801, GL_TEXTURE_MAG_FILTER_ = 0x2800, GL_LINEAR_ = 0x2601, GL_NEAREST_ = 0x2600,
  GL_TEXTURE_WRAP_S_ = 0x2802, GL_TEXTURE_WRAP_T_ = 0x2803, GL_CLAMP_TO_EDGE_ = 0x812F, GL_REPEAT_ = 0x2901,
  GL_ARRAY_BUFFER_ = 0x8892, GL_STATIC_DRAW_ = 0x88E4, GL_TRIANGLE_FAN_ = 0x0006, GL_BLEND_ = 0x0BE2,
  GL_VERSION_ = 0x1F02, GL_RENDERER_ = 0x1F01, GL_EXTENSIONS_ = 0x1F03, GL_READ_FRAMEBUFFER_ = 0x8CA8
};

struct Api {
  void* egl = nullptr;
  void* gles = nullptr;
#define EGLF(ret, name, ...) ret (*name)(__VA_ARGS__) = nullptr;
#define GLF(ret, name, ...) ret (*name)(__VA_ARGS__) = nullptr;
#include "gles_api.inc"
#undef EGLF
#undef GLF
  bool load(const std::string& dir) {
    gles = dlopen((dir + "/libGLESv2.so").c_str(), RTLD_NOW | RTLD_GLOBAL);
    egl = dlopen((dir + "/libEGL.so").c_str(), RTLD_NOW | RTLD_GLOBAL);
    if (!gles || !egl) { fprintf(stderr, "dlopen: %s\n", dlerror()); return false; }
    bool ok = true;
#define EGLF(ret, name, ...) ok &= (*(void**)&name = dlsym(egl, #name)) != nullptr;
#define GLF(ret, name, ...) ok &= (*(void**)&name = dlsym(gles, #name)) != nullptr;
#include "gles_api.inc"
#undef EGLF
#undef GLF
    return ok;
  }
};

static Api gl;

static void die(const char* what) {
  fprintf(stderr, "gles_render: %s (GL error 0x%x)\n", what, gl.glGetError ? gl.glGetError() : 0);
  exit(2);
}

static std::string slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) { fprintf(stderr, "cannot read %s\n", path.c_str()); exit(2); }
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

static std::vector<float> readFloats(const std::string& path, size_t n) {
  std::string s = slurp(path);
  if (s.size() != n * sizeof(float)) {
    fprintf(stderr, "%s: expected %zu floats, found %zu bytes\n", path.c_str(), n, s.size());
    exit(2);
  }
  std::vector<float> v(n);
  memcpy(v.data(), s.data(), s.size());
  return v;
}

static GLuint compile(GLenum type, const std::string& src) {
  GLuint s = gl.glCreateShader(type);
  const char* p = src.c_str();
  GLint len = (GLint)src.size();
  gl.glShaderSource(s, 1, &p, &len);
  gl.glCompileShader(s);
  GLint ok = 0;
  gl.glGetShaderiv(s, GL_COMPILE_STATUS_, &ok);
  if (!ok) {
    std::vector<char> log(1 << 16);
    gl.glGetShaderInfoLog(s, (GLsizei)log.size(), nullptr, log.data());
    fprintf(stderr, "shader compile failed:\n%s\n", log.data());
    exit(3);
  }
  return s;
}

static GLuint makeTexture(int w, int h, int ch, bool linear, bool clamp, const float* data) {
  static const GLenum ifmt[] = {GL_R32F_, GL_RG32F_, GL_RGB32F_, GL_RGBA32F_};
  static const GLenum fmt[] = {GL_RED_, GL_RG_, GL_RGB_, GL_RGBA_};
  GLuint t = 0;
  gl.glGenTextures(1, &t);
  gl.glBindTexture(GL_TEXTURE_2D_, t);
  gl.glTexImage2D(GL_TEXTURE_2D_, 0, (GLint)ifmt[ch - 1], w, h, 0, fmt[ch - 1], GL_FLOAT_, data);
  gl.glTexParameteri(GL_TEXTURE_2D_, GL_TEXTURE_WRAP_S_, clamp ? GL_CLAMP_TO_EDGE_ : GL_REPEAT_);
  gl.glTexParameteri(GL_TEXTURE_2D_, GL_TEXTURE_WRAP_T_, clamp ? GL_CLAMP_TO_EDGE_ : GL_REPEAT_);
  gl.glTexParameteri(GL_TEXTURE_2D_, GL_TEXTURE_MIN_FILTER_, linear ? GL_LINEAR_ : GL_NEAREST_);
  gl.glTexParameteri(GL_TEXTURE_2D_, GL_TEXTURE_MAG_FILTER_, linear ? GL_LINEAR_ : GL_NEAREST_);
  return t;
}

static GLuint makeSrgb8Texture(int w, int h, const unsigned char* rgb) {
  GLuint t = 0;
  gl.glGenTextures(1, &t);
  gl.glBindTexture(GL_TEXTURE_2D_, t);
  gl.glTexImage2D(GL_TEXTURE_2D_, 0, (GLint)GL_SRGB8_, w, h, 0, GL_RGB_, GL_UNSIGNED_BYTE_, rgb);
  gl.glTexParameteri(GL_TEXTURE_2D_, GL_TEXTURE_WRAP_S_, GL_CLAMP_TO_EDGE_);
  gl.glTexParameteri(GL_TEXTURE_2D_, GL_TEXTURE_WRAP_T_, GL_CLAMP_TO_EDGE_);
  gl.glTexParameteri(GL_TEXTURE_2D_, GL_TEXTURE_MIN_FILTER_, GL_LINEAR_);
  return t;
}

struct Tex { int unit; std::string name; GLuint id; };
struct Uni { std::string kind, name; std::vector<double> v; };

int main(int argc, char** argv) {
  if (argc != 3) { fprintf(stderr, "usage: gles_render GLES_LIB_DIR SPEC\n"); return 1; }
  if (!gl.load(argv[1])) die("cannot load the GLES implementation");

  int W = 0, H = 0, frames = 1, outputs = 2;
  bool pickMode = false;
  std::string vsPath, fsPath, radIn, rngIn, radOut, rngOut;
  std::vector<std::pair<int, std::vector<std::string>>> texSpecs;
  std::vector<Uni> unis;
  {
    std::istringstream in(slurp(argv[2]));
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream w(line);
      std::string d;
      if (!(w >> d) || d[0] == '#') continue;
      if (d == "size") w >> W >> H;
      else if (d == "frames") w >> frames;
      else if (d == "outputs") w >> outputs;
      else if (d == "pick") { pickMode = true; outputs = 1; }
      else if (d == "vertex") w >> vsPath;
      else if (d == "fragment") w >> fsPath;
      else if (d == "state") w >> radIn >> rngIn;
      else if (d == "out") { w >> radOut; w >> rngOut; }
      else if (d == "texture_srgb8") {
        std::vector<std::string> f(5);
        for (auto& x : f) w >> x;
        texSpecs.push_back({-1 - atoi(f[0].c_str()), f});
      }
      else if (d == "texture") {
        std::vector<std::string> f(8);
        for (auto& x : f) w >> x;
        texSpecs.push_back({atoi(f[0].c_str()), f});
      } else if (d[0] == 'u') {
        Uni u;
        u.kind = d;
        w >> u.name;
        std::string x;
        while (w >> x) u.v.push_back(strtod(x.c_str(), nullptr));
        unis.push_back(u);
      } else {
        fprintf(stderr, "unknown directive %s\n", d.c_str());
        return 1;
      }
    }
  }
  if (W <= 0 || H <= 0) die("bad size");

  // ---------------------------------------------------------------- context
  EGLDisplay dpy = gl.eglGetDisplay(nullptr);
  EGLint maj, min;
  if (!gl.eglInitialize(dpy, &maj, &min)) die("eglInitialize");
  gl.eglBindAPI(EGL_OPENGL_ES_API_);
  EGLint cattr[] = {EGL_RENDERABLE_TYPE_, EGL_OPENGL_ES3_BIT_, EGL_SURFACE_TYPE_, EGL_PBUFFER_BIT_, EGL_NONE_};
  EGLConfig cfg;
  EGLint ncfg = 0;
  if (!gl.eglChooseConfig(dpy, cattr, &cfg, 1, &ncfg) || ncfg < 1) die("eglChooseConfig");
  EGLint sattr[] = {EGL_WIDTH_, 1, EGL_HEIGHT_, 1, EGL_NONE_};
  EGLSurface surf = gl.eglCreatePbufferSurface(dpy, cfg, sattr);
  EGLint xattr[] = {EGL_CONTEXT_CLIENT_VERSION_, 3, EGL_NONE_};
  EGLContext ctx = gl.eglCreateContext(dpy, cfg, nullptr, xattr);
  if (!ctx || !gl.eglMakeCurrent(dpy, surf, surf, ctx)) die("eglCreateContext/MakeCurrent");
  fprintf(stderr, "GL: %s | %s\n", (const char*)gl.glGetString(GL_VERSION_), (const char*)gl.glGetString(GL_RENDERER_));

  // ---------------------------------------------------------------- program
  GLuint vs = compile(GL_VERTEX_SHADER_, slurp(vsPath));
  GLuint fs = compile(GL_FRAGMENT_SHADER_, slurp(fsPath));
  GLuint prog = gl.glCreateProgram();
  gl.glAttachShader(prog, vs);
  gl.glAttachShader(prog, fs);
  gl.glBindAttribLocation(prog, 0, "Position");
  gl.glBindAttribLocation(prog, 1, "TexCoord");
  gl.glLinkProgram(prog);
  GLint linked = 0;
  gl.glGetProgramiv(prog, GL_LINK_STATUS_, &linked);
  if (!linked) {
    std::vector<char> log(1 << 16);
    gl.glGetProgramInfoLog(prog, (GLsizei)log.size(), nullptr, log.data());
    fprintf(stderr, "link failed:\n%s\n", log.data());
    return 3;
  }
  gl.glUseProgram(prog);

  // ---------------------------------------------------------------- quad
  const float quad[] = {1.0f, 1.0f, 0.0f, 1.0f, 1.0f, -1.0f, 1.0f,  0.0f, 0.0f, 1.0f,
                        -1.0f, -1.0f, 0.0f, 0.0f, 0.0f, 1.0f, -1.0f, 0.0f, 1.0f, 0.0f};
  GLuint vbo = 0;
  gl.glGenBuffers(1, &vbo);
  gl.glBindBuffer(GL_ARRAY_BUFFER_, vbo);
  gl.glBufferData(GL_ARRAY_BUFFER_, sizeof(quad), quad, GL_STATIC_DRAW_);
  gl.glEnableVertexAttribArray(0);
  gl.glVertexAttribPointer(0, 3, GL_FLOAT_, 0, 5 * sizeof(float), (const void*)0);
  gl.glEnableVertexAttribArray(1);
  gl.glVertexAttribPointer(1, 2, GL_FLOAT_, 0, 5 * sizeof(float), (const void*)(3 * sizeof(float)));

  // ---------------------------------------------------------------- textures
  std::vector<Tex> texs;
  for (auto& ts : texSpecs) {
    const auto& f = ts.second;
    if (ts.first < 0) {  // texture_srgb8: unit is -1 - UNIT
      int w = atoi(f[2].c_str()), h = atoi(f[3].c_str());
      std::string bytes = slurp(f[4]);
      if ((3 * w) % 4 != 0) die("texture_srgb8: rows must be whole 4-byte words");
      if (bytes.size() != (size_t)w * h * 3) die("texture_srgb8: size mismatch");
      GLuint id = makeSrgb8Texture(w, h, (const unsigned char*)bytes.data());
      texs.push_back({-1 - ts.first, f[1], id});
      continue;
    }
    int w = atoi(f[2].c_str()), h = atoi(f[3].c_str()), ch = atoi(f[4].c_str());
    std::vector<float> data = readFloats(f[7], (size_t)w * h * ch);
    GLuint id = makeTexture(w, h, ch, f[5] == "1", f[6] == "1", data.data());
    texs.push_back({ts.first, f[1], id});
  }
  size_t npx = (size_t)W * H;
  std::vector<float> rad0 = readFloats(radIn, npx * 4), rng0 = readFloats(rngIn, npx * 4);
  GLuint rad[2] = {makeTexture(W, H, 4, false, true, rad0.data()), makeTexture(W, H, 4, false, true, nullptr)};
  GLuint rng[2] = {makeTexture(W, H, 4, false, true, rng0.data()), makeTexture(W, H, 4, false, true, nullptr)};
  GLuint fbo = 0;
  gl.glGenFramebuffers(1, &fbo);

  // ---------------------------------------------------------------- uniforms
  for (const Uni& u : unis) {
    GLint loc = gl.glGetUniformLocation(prog, u.name.c_str());
    if (loc < 0) continue;  // unknown names are ignored, as WebGL ignores a null location
    std::vector<float> f(u.v.begin(), u.v.end());
    if (u.kind == "u1i") gl.glUniform1i(loc, (GLint)u.v[0]);
    else if (u.kind == "u1f") gl.glUniform1f(loc, f[0]);
    else if (u.kind == "u2f") gl.glUniform2f(loc, f[0], f[1]);
    else if (u.kind == "u3f") gl.glUniform3f(loc, f[0], f[1], f[2]);
    else if (u.kind == "u4f") gl.glUniform4f(loc, f[0], f[1], f[2], f[3]);
    else if (u.kind == "u1fv") gl.glUniform1fv(loc, (GLsizei)f[0], f.data() + 1);
    else if (u.kind == "u3fv") gl.glUniform3fv(loc, (GLsizei)f[0], f.data() + 1);
    else { fprintf(stderr, "unknown uniform kind %s\n", u.kind.c_str()); return 1; }
  }
  for (const Tex& t : texs) {
    if (t.name == "-") continue;
    GLint loc = gl.glGetUniformLocation(prog, t.name.c_str());
    if (loc >= 0) gl.glUniform1i(loc, t.unit);
  }
  GLint locRad = gl.glGetUniformLocation(prog, "Radiance"), locRng = gl.glGetUniformLocation(prog, "RngData");
  if (locRad >= 0) gl.glUniform1i(locRad, 0);
  if (locRng >= 0) gl.glUniform1i(locRng, 1);

  // ---------------------------------------------------------------- frames
  gl.glDisable(GL_BLEND_);
  gl.glViewport(0, 0, W, H);
  int cur = 0;
  if (pickMode) {
    gl.glBindFramebuffer(GL_FRAMEBUFFER_, fbo);
    gl.glFramebufferTexture2D(GL_FRAMEBUFFER_, GL_COLOR_ATTACHMENT0_, GL_TEXTURE_2D_, rad[1], 0);
    GLenum buf0 = GL_COLOR_ATTACHMENT0_;
    gl.glDrawBuffers(1, &buf0);
    if (gl.glCheckFramebufferStatus(GL_FRAMEBUFFER_) != GL_FRAMEBUFFER_COMPLETE_) die("framebuffer incomplete");
    for (const Tex& t : texs) {
      gl.glActiveTexture(GL_TEXTURE0_ + t.unit);
      gl.glBindTexture(GL_TEXTURE_2D_, t.id);
    }
    GLint locPick = gl.glGetUniformLocation(prog, "mousePick");
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        gl.glUniform2f(locPick, x + 0.5f, y + 0.5f);
        gl.glViewport(x, y, 1, 1);
        gl.glDrawArrays(GL_TRIANGLE_FAN_, 0, 4);
      }
    gl.glFinish();
    frames = 0;
    cur = 1;
  }
  for (int f = 0; f < frames; ++f) {
    int nxt = 1 - cur;
    gl.glBindFramebuffer(GL_FRAMEBUFFER_, fbo);
    gl.glFramebufferTexture2D(GL_FRAMEBUFFER_, GL_COLOR_ATTACHMENT0_, GL_TEXTURE_2D_, rad[nxt], 0);
    gl.glFramebufferTexture2D(GL_FRAMEBUFFER_, GL_COLOR_ATTACHMENT0_ + 1, GL_TEXTURE_2D_, outputs > 1 ? rng[nxt] : 0, 0);
    GLenum bufs[2] = {GL_COLOR_ATTACHMENT0_, GL_COLOR_ATTACHMENT0_ + 1};
    gl.glDrawBuffers(outputs, bufs);
    if (gl.glCheckFramebufferStatus(GL_FRAMEBUFFER_) != GL_FRAMEBUFFER_COMPLETE_) die("framebuffer incomplete");
    gl.glActiveTexture(GL_TEXTURE0_ + 0);
    gl.glBindTexture(GL_TEXTURE_2D_, rad[cur]);
    gl.glActiveTexture(GL_TEXTURE0_ + 1);
    gl.glBindTexture(GL_TEXTURE_2D_, rng[cur]);
    for (const Tex& t : texs) {
      gl.glActiveTexture(GL_TEXTURE0_ + t.unit);
      gl.glBindTexture(GL_TEXTURE_2D_, t.id);
    }
    gl.glDrawArrays(GL_TRIANGLE_FAN_, 0, 4);
    gl.glFinish();
    if (outputs < 2) std::swap(rng[0], rng[1]);  // the rng input stays the rng input
    cur = nxt;
    fprintf(stderr, "frame %d/%d\n", f + 1, frames);
  }

  // ---------------------------------------------------------------- read back
  gl.glBindFramebuffer(GL_FRAMEBUFFER_, fbo);
  gl.glFramebufferTexture2D(GL_FRAMEBUFFER_, GL_COLOR_ATTACHMENT0_, GL_TEXTURE_2D_, rad[cur], 0);
  gl.glFramebufferTexture2D(GL_FRAMEBUFFER_, GL_COLOR_ATTACHMENT0_ + 1, GL_TEXTURE_2D_, rng[cur], 0);
  std::vector<float> out(npx * 4);
  auto dump = [&](GLenum att, const std::string& path) {
    gl.glReadBuffer(att);
    gl.glReadPixels(0, 0, W, H, GL_RGBA_, GL_FLOAT_, out.data());
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) die("cannot write output");
    fwrite(out.data(), sizeof(float), out.size(), fp);
    fclose(fp);
  };
  dump(GL_COLOR_ATTACHMENT0_, radOut);
  if (!rngOut.empty()) dump(GL_COLOR_ATTACHMENT0_ + 1, rngOut);
  return 0;
}
