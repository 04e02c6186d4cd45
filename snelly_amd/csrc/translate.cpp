// translate.cpp -- GLSL ES 3.00 scene snippets -> C++ for hipRTC.
//
// Input is what the reference splices into its fragment shaders at __SHADER__
// and __IOR_FUNC__ (js/renderer.js:217-359, js/gl.js:157-168): the scene's
// shader() text with the default hooks appended, followed by the loaded
// materials' IOR code.  Output is one C++ struct, SnlScene, whose members are
// the snippet's functions and global variables, so GLSL "globals" (e.g. the
// `mat3 rot;` that INIT() fills in exampleScenes/fractals/menger.html) are
// per-invocation as in a fragment shader.  Uniform declarations are lifted into
// a reflected block (struct SnlUser) read through a const pointer; references
// to them and to the built-in uniforms become snlU_->name / snlP_->name.
//
// The rewrite is token-level: qualifiers in/out/inout become by-value/by-ref
// parameters, float literals get an f suffix (GLSL float literals are single
// precision), precision qualifiers are dropped, C++-only keywords used as GLSL
// identifiers are renamed, and GLSL array constructors in initialisers become
// brace lists.
#include "translate.h"

#include <cctype>
#include <cstring>
#include <map>
#include <set>
#include <sstream>

namespace snl {

namespace {

enum TokKind { T_WS, T_PP, T_ID, T_NUM, T_PUNCT };
struct Tok {
  TokKind k;
  std::string s;
  int line;
};

std::vector<Tok> tokenize(const std::string& src, std::string& err) {
  std::vector<Tok> out;
  size_t i = 0, n = src.size();
  int line = 1;
  bool lineStart = true;
  auto push = [&](TokKind k, size_t b, size_t e) { out.push_back({k, src.substr(b, e - b), line}); };
  while (i < n) {
    char c = src[i];
    if (c == '\n') { push(T_WS, i, i + 1); ++line; ++i; lineStart = true; continue; }
    if (isspace((unsigned char)c)) { size_t b = i; while (i < n && src[i] != '\n' && isspace((unsigned char)src[i])) ++i; push(T_WS, b, i); continue; }
    if (c == '/' && i + 1 < n && src[i + 1] == '/') {  // line comment
      while (i < n && src[i] != '\n') ++i;
      continue;
    }
    if (c == '/' && i + 1 < n && src[i + 1] == '*') {  // block comment: keep its newlines
      size_t e = src.find("*/", i + 2);
      if (e == std::string::npos) { err = "unterminated comment"; return out; }
      for (size_t k = i; k < e; ++k) if (src[k] == '\n') { out.push_back({T_WS, "\n", line}); ++line; }
      out.push_back({T_WS, " ", line});
      i = e + 2;
      continue;
    }
    if (c == '#' && lineStart) {  // preprocessor line (with continuations)
      size_t b = i;
      while (i < n && src[i] != '\n') {
        if (src[i] == '\\' && i + 1 < n && src[i + 1] == '\n') { i += 2; ++line; continue; }
        ++i;
      }
      push(T_PP, b, i);
      continue;
    }
    lineStart = false;
    if (isalpha((unsigned char)c) || c == '_') {
      size_t b = i;
      while (i < n && (isalnum((unsigned char)src[i]) || src[i] == '_')) ++i;
      push(T_ID, b, i);
      continue;
    }
    if (isdigit((unsigned char)c) || (c == '.' && i + 1 < n && isdigit((unsigned char)src[i + 1]))) {
      size_t b = i;
      if (c == '0' && i + 1 < n && (src[i + 1] == 'x' || src[i + 1] == 'X')) {
        i += 2;
        while (i < n && isxdigit((unsigned char)src[i])) ++i;
        while (i < n && (src[i] == 'u' || src[i] == 'U')) ++i;
        push(T_NUM, b, i);
        continue;
      }
      while (i < n && isdigit((unsigned char)src[i])) ++i;
      if (i < n && src[i] == '.') { ++i; while (i < n && isdigit((unsigned char)src[i])) ++i; }
      if (i < n && (src[i] == 'e' || src[i] == 'E')) {
        size_t s = i + 1;
        if (s < n && (src[s] == '+' || src[s] == '-')) ++s;
        if (s < n && isdigit((unsigned char)src[s])) { i = s; while (i < n && isdigit((unsigned char)src[i])) ++i; }
      }
      while (i < n && (src[i] == 'f' || src[i] == 'F' || src[i] == 'u' || src[i] == 'U')) ++i;
      push(T_NUM, b, i);
      continue;
    }
    static const char* ops[] = {"<<=", ">>=", "++", "--", "+=", "-=", "*=", "/=", "%=", "==", "!=", "<=", ">=",
                                "&&", "||", "^^", "<<", ">>", "&=", "|=", "^=", "->"};
    bool matched = false;
    for (const char* op : ops) {
      size_t L = strlen(op);
      if (src.compare(i, L, op) == 0) { push(T_PUNCT, i, i + L); i += L; matched = true; break; }
    }
    if (!matched) { push(T_PUNCT, i, i + 1); ++i; }
  }
  return out;
}

const std::set<std::string>& builtinTypes() {
  static const std::set<std::string> s = {
      "void", "bool", "int", "uint", "float", "vec2", "vec3", "vec4", "ivec2", "ivec3", "ivec4",
      "uvec2", "uvec3", "uvec4", "bvec2", "bvec3", "bvec4", "mat2", "mat3", "mat4",
      "sampler2D", "sampler3D"};
  return s;
}

// C++ keywords (and names glsl.h reserves) that are ordinary identifiers in GLSL
const std::set<std::string>& renamed() {
  static const std::set<std::string> s = {
      "alignas", "alignof", "and", "and_eq", "asm", "auto", "bitand", "bitor", "catch", "char", "char16_t",
      "char32_t", "class", "compl", "concept", "constexpr", "const_cast", "decltype", "delete", "double",
      "dynamic_cast", "explicit", "export", "extern", "friend", "goto", "inline", "long", "mutable",
      "namespace", "new", "noexcept", "not_eq", "nullptr", "operator", "or", "or_eq", "private",
      "protected", "public", "register", "reinterpret_cast", "requires", "short", "signed", "sizeof",
      "static", "static_assert", "static_cast", "template", "this", "thread_local", "throw", "try",
      "typedef", "typeid", "typename", "union", "unsigned", "using", "virtual", "volatile", "wchar_t",
      "xor", "xor_eq", "glsl", "std", "SnlScene", "SnlUser", "SnlParams"};
  return s;
}

const std::set<std::string>& dropped() {
  static const std::set<std::string> s = {"lowp", "mediump", "highp", "precise", "invariant", "flat", "smooth",
                                          "centroid", "noperspective"};
  return s;
}

bool isFloatLiteral(const std::string& t) {
  if (t.size() > 1 && t[0] == '0' && (t[1] == 'x' || t[1] == 'X')) return false;
  bool hasDot = t.find('.') != std::string::npos;
  bool hasExp = t.find_first_of("eE") != std::string::npos;
  return hasDot || hasExp;
}

struct Uniform {
  std::string type, name;
  int count;  // array length, 1 for scalars
};

int typeSize(const std::string& t, int& code, int& align) {
  align = 4;
  if (t == "float") { code = 1; return 4; }
  if (t == "vec2") { code = 2; return 8; }
  if (t == "vec3") { code = 3; return 12; }
  if (t == "vec4") { code = 4; return 16; }
  if (t == "int" || t == "bool" || t == "uint") { code = 5; return 4; }
  if (t == "ivec2" || t == "bvec2") { code = 5; return 8; }
  if (t == "ivec3" || t == "bvec3") { code = 5; return 12; }
  if (t == "ivec4" || t == "bvec4") { code = 5; return 16; }
  if (t == "mat2") { code = 6; return 16; }
  if (t == "mat3") { code = 7; return 36; }
  if (t == "mat4") { code = 8; return 64; }
  if (t == "sampler2D") { code = 9; align = 8; return 24; }
  if (t == "sampler3D") { code = 10; align = 8; return 32; }
  code = 0;
  return 0;
}

// C++ spelling of a GLSL type in the uniform block (bool/bvec stored as int)
std::string blockType(const std::string& t) {
  if (t == "bool" || t == "uint") return "int";
  if (t == "bvec2") return "ivec2";
  if (t == "bvec3") return "ivec3";
  if (t == "bvec4") return "ivec4";
  return t;
}

class Translator {
 public:
  Translator(const std::string& src, const std::set<std::string>& builtinUniforms, const std::set<std::string>& defines)
      : src_(src), builtinUniforms_(builtinUniforms), defines_(defines) {}

  bool run(TranslateResult& out) {
    std::string err;
    toks_ = tokenize(src_, err);
    if (!err.empty()) { out.error = err; return false; }
    // collect user struct names and uniforms first (a uniform may be used before its declaration)
    for (size_t i = 0; i < toks_.size(); ++i) {
      if (toks_[i].k == T_ID && toks_[i].s == "struct") {
        size_t j = next(i);
        if (j < toks_.size() && toks_[j].k == T_ID) structs_.insert(toks_[j].s);
      }
    }
    // and the uniform names, which #define bodies may mention before the declaration
    for (size_t i = 0; i < toks_.size(); ++i) {
      if (!(toks_[i].k == T_ID && toks_[i].s == "uniform")) continue;
      size_t j = next(i);
      while (j < toks_.size() && dropped().count(toks_[j].s)) j = next(j);
      j = next(j);  // past the type
      while (j < toks_.size() && toks_[j].k == T_ID) {
        if (!builtinUniforms_.count(toks_[j].s)) preUniforms_.insert(toks_[j].s);
        j = next(j);
        while (j < toks_.size() && !(toks_[j].k == T_PUNCT && (toks_[j].s == "," || toks_[j].s == ";"))) j = next(j);
        if (j >= toks_.size() || toks_[j].s == ";") break;
        j = next(j);
      }
    }
    std::ostringstream body;
    size_t i = 0;
    while (i < toks_.size()) {
      const Tok& t = toks_[i];
      if (t.k == T_WS) { body << t.s; ++i; continue; }
      if (t.k == T_PP) { noteMacro(t.s); body << ppLine(t.s) << "\n"; ++i; continue; }
      // find the extent of this top-level declaration
      size_t end = i;
      int depth = 0;
      bool sawParen = false;
      bool isFunc = false;
      size_t k = i;
      for (; k < toks_.size(); ++k) {
        const Tok& u = toks_[k];
        if (u.k != T_PUNCT) continue;
        if (u.s == "(" ) { if (depth == 0 && !sawParen) sawParen = true; ++depth; }
        else if (u.s == ")") --depth;
        else if (u.s == "{") {
          if (depth == 0) {
            // function body or struct body
            int d2 = 0;
            size_t m = k;
            for (; m < toks_.size(); ++m) {
              if (toks_[m].k != T_PUNCT) continue;
              if (toks_[m].s == "{") ++d2;
              else if (toks_[m].s == "}") { if (--d2 == 0) break; }
            }
            if (m >= toks_.size()) { out.error = "unbalanced braces near line " + std::to_string(u.line); return false; }
            bool isStruct = firstId(i) == "struct";
            if (isStruct) {
              // consume through the closing ';'
              size_t s = m + 1;
              while (s < toks_.size() && !(toks_[s].k == T_PUNCT && toks_[s].s == ";")) ++s;
              end = s;
            } else {
              isFunc = true;
              end = m;
            }
            break;
          }
          ++depth;
        } else if (u.s == "}") --depth;
        else if (u.s == ";" && depth == 0) { end = k; break; }
      }
      if (k >= toks_.size()) { out.error = "unterminated declaration near line " + std::to_string(t.line); return false; }
      std::string err2;
      if (isFunc) body << function(i, end, err2);
      else body << declaration(i, end, err2);
      if (!err2.empty()) { out.error = err2; return false; }
      i = end + 1;
    }
    out.uniforms.clear();
    // uniform block layout: 8-byte aligned samplers first, then the rest in order
    std::ostringstream blk;
    blk << "struct SnlUser {\n";
    int off = 0, pad = 0;
    std::vector<const Uniform*> ordered;
    for (auto& u : uniforms_) if (u.type == "sampler2D" || u.type == "sampler3D") ordered.push_back(&u);
    for (auto& u : uniforms_) if (!(u.type == "sampler2D" || u.type == "sampler3D")) ordered.push_back(&u);
    std::ostringstream asserts;
    for (const Uniform* u : ordered) {
      int code, align;
      int sz = typeSize(u->type, code, align);
      if (!sz) { out.error = "unsupported uniform type '" + u->type + "' for " + u->name; return false; }
      if (off % align) { int p = align - off % align; blk << "  char __pad" << pad++ << "[" << p << "];\n"; off += p; }
      blk << "  " << blockType(u->type) << " " << u->name;
      if (u->count > 1) blk << "[" << u->count << "]";
      blk << ";\n";
      out.uniforms.push_back({u->name, off, code, u->count, sz});
      off += sz * u->count;
    }
    if (ordered.empty()) blk << "  int __unused;\n";
    blk << "};\n";
    out.userBlockSize = off ? off : 4;
    std::ostringstream o;
    o << blk.str();
    o << "struct SnlScene {\n";
    o << "  const SnlParams* __restrict__ snlP_;\n  const SnlUser* __restrict__ snlU_;\n";
    o << "  SNL_HD SnlScene(const SnlParams* p, const SnlUser* u) : snlP_(p), snlU_(u) {}\n";
    auto flag = [&](const char* name, const char* def) {
      o << "  static constexpr bool " << name << " = " << (defines_.count(def) ? "true" : "false") << ";\n";
    };
    flag("kHasSurface", "HAS_SURFACE");
    flag("kHasMetal", "HAS_METAL");
    flag("kHasDielectric", "HAS_DIELECTRIC");
    flag("kHasAtmosphere", "HAS_ATMOSPHERE");
    flag("kHasFog", "HAS_FOG");
    flag("kHasVolumeEmission", "HAS_VOLUME_EMISSION");
    flag("kHasVolume", "HAS_VOLUME");
    flag("kHasNormalMap", "HAS_NORMALMAP");
    flag("kHasSurfaceNM", "HAS_SURFACE_NORMALMAP");
    flag("kHasMetalNM", "HAS_METAL_NORMALMAP");
    flag("kHasDielectricNM", "HAS_DIELECTRIC_NORMALMAP");
    flag("kHasCustomCamera", "HAS_CUSTOM_CAMERA");
    flag("kHasSphereLight", "HAS_SPHERE_LIGHT");
    flag("kInteractive", "INTERACTIVE_MODE");
    // the program's __MAX_*__ limits, from the macros its build defines where
    // this struct is compiled (one scene per block of a batch module)
    o << "  static constexpr int kMaxMarchSteps = SNL_MAX_MARCH_STEPS;\n";
    o << "  static constexpr int kMaxAtmosphereScatters = SNL_MAX_ATMOSPHERE_SCATTERS;\n";
    o << "  using User = SnlUser;\n";
    o << "  SNL_HD float rand(vec4& rnd) { return snl_rand(rnd); }\n";
    o << "#line 1 \"scene.glsl\"\n";
    o << body.str() << "\n};\n";
    // the scene's own macros end with the scene: kernel code that follows must
    // not see, say, a "#define PI 3.14159"
    for (const auto& m : macros_) o << "#undef " << m << "\n";
    out.code = o.str();
    return true;
  }

 private:
  size_t next(size_t i) const {
    ++i;
    while (i < toks_.size() && toks_[i].k == T_WS) ++i;
    return i;
  }
  std::string firstId(size_t i) const {
    while (i < toks_.size() && toks_[i].k != T_ID) ++i;
    return i < toks_.size() ? toks_[i].s : "";
  }
  std::string ppLine(const std::string& s) const {
    // drop directives that have no meaning for the C++ build
    size_t p = s.find_first_not_of("# \t");
    std::string w = p == std::string::npos ? "" : s.substr(p, s.find_first_of(" \t(", p) - p);
    if (w == "version" || w == "extension" || w == "pragma") return "";
    if (w == "define") return defineLine(s, p + 6);
    return s;
  }
  // "#define NAME[(params)] body": the body gets the same literal and uniform
  // rewriting as code (a "#define PI 3.14159" must expand to a float)
  std::string defineLine(const std::string& s, size_t p) const {
    size_t b = s.find_first_not_of(" \t", p);
    if (b == std::string::npos) return s;
    size_t e = b;
    while (e < s.size() && (isalnum((unsigned char)s[e]) || s[e] == '_')) ++e;
    std::set<std::string> params;
    if (e < s.size() && s[e] == '(') {
      size_t c = s.find(')', e);
      if (c == std::string::npos) return s;
      std::string plist = s.substr(e + 1, c - e - 1);
      std::string cur;
      for (char ch : plist + ",") {
        if (ch == ',') { if (!cur.empty()) params.insert(cur); cur.clear(); }
        else if (isalnum((unsigned char)ch) || ch == '_') cur += ch;
      }
      e = c + 1;
    }
    std::string err;
    std::vector<Tok> bt = tokenize(s.substr(e), err);
    if (!err.empty()) return s;
    std::ostringstream o;
    o << s.substr(0, e);
    for (size_t i = 0; i < bt.size(); ++i) {
      const Tok& t = bt[i];
      if (t.k == T_NUM && isFloatLiteral(t.s)) {
        std::string v = t.s;
        if (v.back() != 'f' && v.back() != 'F') v += "f";
        if (v[0] == '.') v = "0" + v;
        o << v;
        continue;
      }
      if (t.k == T_ID && !params.count(t.s)) {
        bool member = false;
        for (size_t q = i; q > 0;) {
          --q;
          if (bt[q].k == T_WS) continue;
          member = bt[q].k == T_PUNCT && (bt[q].s == "." || bt[q].s == "->");
          break;
        }
        if (!member && preUniforms_.count(t.s)) { o << "snlU_->" << t.s; continue; }
        if (!member && builtinUniforms_.count(t.s)) { o << "snlP_->" << t.s; continue; }
      }
      o << (t.s == "\\\n" ? " " : t.s);
    }
    return o.str();
  }
  void noteMacro(const std::string& s) {
    size_t p = s.find_first_not_of("# \t");
    if (p == std::string::npos || s.compare(p, 6, "define") != 0) return;
    size_t b = s.find_first_not_of(" \t", p + 6);
    if (b == std::string::npos) return;
    size_t e = b;
    while (e < s.size() && (isalnum((unsigned char)s[e]) || s[e] == '_')) ++e;
    std::string name = s.substr(b, e - b);
    // feature switches the kernels test stay defined
    if (name.empty() || name.compare(0, 4, "HAS_") == 0 || name == "DISPERSION_ENABLED" || name == "INTERACTIVE_MODE")
      return;
    macros_.insert(name);
  }
  bool isType(const std::string& s) const { return builtinTypes().count(s) || structs_.count(s); }

  // A local or parameter named like a uniform hides it inside its scope, as in
  // GLSL: names declared in the open scopes are never rewritten to block reads.
  bool shadowed(const std::string& n) const {
    for (const auto& sc : scopes_) if (sc.count(n)) return true;
    return false;
  }
  void declare(const std::string& n) const {
    if (!scopes_.empty()) scopes_.back().insert(n);
  }

  // rewrite a token range [b, e) of expressions / statements
  std::string rewrite(size_t b, size_t e) const {
    std::ostringstream o;
    // declaration tracking: after "TYPE name", names following top-level
    // commas up to ';' are declared too
    bool inDecl = false, expectName = false;
    int declDepth = 0;
    for (size_t i = b; i < e; ++i) {
      const Tok& t = toks_[i];
      if (t.k == T_PUNCT) {
        const std::string& p = t.s;
        if (p == "{") scopes_.emplace_back();
        else if (p == "}" && !scopes_.empty()) scopes_.pop_back();
        if (inDecl) {
          if (p == "(" || p == "[" || p == "{") ++declDepth;
          else if (p == ")" || p == "]" || p == "}") { if (--declDepth < 0) inDecl = false; }
          else if (p == ";") inDecl = false;
          else if (p == "," && declDepth == 0) expectName = true;
        }
      }
      if (t.k == T_PP) { o << "\n" << ppLine(t.s) << "\n"; continue; }
      if (t.k == T_NUM) {
        std::string s = t.s;
        if (isFloatLiteral(s)) {
          if (s.back() != 'f' && s.back() != 'F') s += "f";
          if (s[0] == '.') s = "0" + s;
        } else {
          while (!s.empty() && (s.back() == 'u' || s.back() == 'U')) s.pop_back();
          if (t.s != s) s += "u";
        }
        o << s;
        continue;
      }
      if (t.k == T_ID) {
        if (dropped().count(t.s)) continue;
        // member access: leave swizzles and fields alone
        size_t p = i;
        bool member = false;
        while (p > b) { --p; if (toks_[p].k == T_WS) continue; member = (toks_[p].k == T_PUNCT && (toks_[p].s == "." || toks_[p].s == "->")); break; }
        if (expectName && !member) {
          declare(t.s);
          expectName = false;
        }
        if (!member && isType(t.s)) {
          // "TYPE name" or "TYPE[N] name" declares name in the current scope
          size_t q = next(i);
          if (q < e && toks_[q].k == T_PUNCT && toks_[q].s == "[") {
            while (q < e && !(toks_[q].k == T_PUNCT && toks_[q].s == "]")) ++q;
            q = next(q);
          }
          if (q < e && toks_[q].k == T_ID) {
            inDecl = true;
            declDepth = 0;
            expectName = true;
          }
        }
        if (!member) {
          if (!shadowed(t.s)) {
            if (uniformNames_.count(t.s)) { o << "snlU_->" << t.s; continue; }
            if (builtinUniforms_.count(t.s)) { o << "snlP_->" << t.s; continue; }
          }
          // GLSL float(v) of a vector takes its first component
          if (t.s == "float") {
            size_t q = next(i);
            if (q < e && toks_[q].k == T_PUNCT && toks_[q].s == "(") { o << "snl_float"; continue; }
          }
          if (t.s == "not") { o << "not_"; continue; }
          if (t.s == "uint") { o << "unsigned"; continue; }
          if (renamed().count(t.s)) { o << t.s << "_"; continue; }
          // GLSL array constructor T[N](...) -> brace list
          if (isType(t.s)) {
            size_t q = next(i);
            if (q < e && toks_[q].k == T_PUNCT && toks_[q].s == "[") {
              size_t r = q;
              while (r < e && !(toks_[r].k == T_PUNCT && toks_[r].s == "]")) ++r;
              size_t s = next(r);
              if (s < e && toks_[s].k == T_PUNCT && toks_[s].s == "(") {
                int depth = 0;
                size_t m = s;
                for (; m < e; ++m) {
                  if (toks_[m].k != T_PUNCT) continue;
                  if (toks_[m].s == "(") ++depth;
                  else if (toks_[m].s == ")") { if (--depth == 0) break; }
                }
                o << "{" << rewrite(s + 1, m) << "}";
                i = m;
                continue;
              }
            }
          }
        } else if (renamed().count(t.s)) {
          o << t.s << "_";
          continue;
        }
        o << t.s;
        continue;
      }
      if (t.k == T_PUNCT && t.s == "^^") { o << "!=";  continue; }  // logical xor on bools
      o << t.s;
    }
    return o.str();
  }

  // a top-level declaration ending at the ';' token `e`
  std::string declaration(size_t b, size_t e, std::string& err) {
    std::vector<size_t> ids;
    for (size_t i = b; i < e; ++i) if (toks_[i].k != T_WS) ids.push_back(i);
    if (ids.empty()) return "";
    const std::string& w0 = toks_[ids[0]].s;
    if (w0 == "precision") return "";
    if (w0 == "struct") return rewrite(b, e + 1) + "\n";
    if (w0 == "uniform") {
      // uniform [prec] TYPE name [ '[' N ']' ] (, name [...])*
      size_t k = 1;
      while (k < ids.size() && dropped().count(toks_[ids[k]].s)) ++k;
      if (k >= ids.size()) { err = "bad uniform declaration at line " + std::to_string(toks_[b].line); return ""; }
      std::string type = toks_[ids[k]].s;
      ++k;
      while (k < ids.size()) {
        std::string name = toks_[ids[k]].s;
        int count = 1;
        ++k;
        if (k < ids.size() && toks_[ids[k]].s == "[") {
          if (k + 2 < ids.size()) count = atoi(toks_[ids[k + 1]].s.c_str());
          k += 3;
        }
        if (builtinUniforms_.count(name)) {
          // the built-in block already declares it (scenes may redeclare)
        } else if (!uniformNames_.count(name)) {
          uniforms_.push_back({type, name, count > 0 ? count : 1});
          uniformNames_.insert(name);
        }
        if (k < ids.size() && toks_[ids[k]].s == ",") ++k;
      }
      return "";
    }
    // global variable(s): [const] TYPE a [= init] (, b [= init])*
    size_t k = 0;
    while (k < ids.size() && (toks_[ids[k]].s == "const" || dropped().count(toks_[ids[k]].s))) ++k;
    if (k >= ids.size()) return "";
    // a function prototype (TYPE name(...);) -- members need no forward declaration
    if (k + 2 < ids.size() && toks_[ids[k + 1]].k == T_ID && toks_[ids[k + 2]].s == "(") return "";
    std::string type = toks_[ids[k]].s;
    size_t declStart = ids[k] + 1;
    // split declarators at top-level commas
    std::ostringstream o;
    size_t s = declStart;
    int depth = 0;
    for (size_t i = declStart; i <= e; ++i) {
      bool atEnd = (i == e);
      if (!atEnd && toks_[i].k == T_PUNCT) {
        if (toks_[i].s == "(" || toks_[i].s == "[" || toks_[i].s == "{") ++depth;
        else if (toks_[i].s == ")" || toks_[i].s == "]" || toks_[i].s == "}") --depth;
      }
      if (atEnd || (depth == 0 && toks_[i].k == T_PUNCT && toks_[i].s == ",")) {
        o << "  " << rewrite(ids[k], ids[k] + 1) << " " << declarator(s, i) << ";";
        s = i + 1;
      }
    }
    (void)type;
    return o.str();
  }

  // one declarator "name [N] = init"; an unsized array takes its size from a
  // GLSL array-constructor initialiser
  std::string declarator(size_t b, size_t e) const {
    std::string d = rewrite(b, e);
    size_t lb = d.find("[]");
    if (lb != std::string::npos) {
      size_t eq = d.find('=');
      if (eq != std::string::npos) {
        // count top-level commas in the brace list
        size_t ob = d.find('{', eq), cb = d.rfind('}');
        if (ob != std::string::npos && cb != std::string::npos) {
          int n = 1, depth = 0;
          for (size_t i = ob + 1; i < cb; ++i) {
            char c = d[i];
            if (c == '(' || c == '{' || c == '[') ++depth;
            else if (c == ')' || c == '}' || c == ']') --depth;
            else if (c == ',' && depth == 0) ++n;
          }
          d = d.substr(0, lb + 1) + std::to_string(n) + d.substr(lb + 1);
        }
      }
    }
    return d;
  }

  // function definition tokens [b .. closing brace e]
  std::string function(size_t b, size_t e, std::string& err) {
    // header up to the '(' that follows the function name
    size_t lp = b;
    while (lp < e && !(toks_[lp].k == T_PUNCT && toks_[lp].s == "(")) ++lp;
    int depth = 0;
    size_t rp = lp;
    for (; rp < e; ++rp) {
      if (toks_[rp].k != T_PUNCT) continue;
      if (toks_[rp].s == "(") ++depth;
      else if (toks_[rp].s == ")") { if (--depth == 0) break; }
    }
    size_t lb = rp;
    while (lb < e && !(toks_[lb].k == T_PUNCT && toks_[lb].s == "{")) ++lb;
    if (lb >= e) { err = "malformed function near line " + std::to_string(toks_[b].line); return ""; }
    std::ostringstream o;
    o << "SNL_HD " << rewrite(b, lp) << "(";
    // parameters
    std::vector<std::pair<size_t, size_t>> params;
    size_t s = lp + 1;
    depth = 0;
    for (size_t i = lp + 1; i <= rp; ++i) {
      bool atEnd = (i == rp);
      if (!atEnd && toks_[i].k == T_PUNCT) {
        if (toks_[i].s == "(" || toks_[i].s == "[") ++depth;
        else if (toks_[i].s == ")" || toks_[i].s == "]") --depth;
      }
      if (atEnd || (depth == 0 && toks_[i].k == T_PUNCT && toks_[i].s == ",")) { params.push_back({s, i}); s = i + 1; }
    }
    // the parameters' names open the function's scope
    scopes_.clear();
    scopes_.emplace_back();
    for (auto& pr : params) {
      std::string name;
      for (size_t i = pr.first; i < pr.second; ++i) {
        if (toks_[i].k == T_PUNCT && toks_[i].s == "[") break;
        if (toks_[i].k == T_ID) name = toks_[i].s;
      }
      if (!name.empty()) scopes_.back().insert(name);
    }
    struct PInfo { std::string type, name; bool ref, vec, array; };
    std::vector<PInfo> pinfo;
    bool first = true;
    for (auto& pr : params) {
      std::vector<size_t> w;
      for (size_t i = pr.first; i < pr.second; ++i) if (toks_[i].k != T_WS) w.push_back(i);
      if (w.empty()) continue;
      if (w.size() == 1 && toks_[w[0]].s == "void") continue;
      bool ref = false, isConst = false;
      size_t k = 0;
      for (; k < w.size(); ++k) {
        const std::string& q = toks_[w[k]].s;
        if (q == "in") continue;
        if (q == "out" || q == "inout") { ref = true; continue; }
        if (q == "const") { isConst = true; continue; }
        if (dropped().count(q)) continue;
        break;
      }
      if (k >= w.size()) continue;
      std::string type = rewrite(w[k], w[k] + 1);
      std::string rest = (k + 1 < w.size()) ? rewrite(w[k + 1], pr.second) : std::string("");
      if (!first) o << ", ";
      first = false;
      if (isConst && !ref) o << "const ";
      o << type << (ref ? "& " : " ") << rest;
      bool arr = rest.find('[') != std::string::npos;
      pinfo.push_back({type, rest, ref, type == "vec2" || type == "vec3" || type == "vec4", arr});
    }
    o << ")";
    o << rewrite(rp + 1, e + 1) << "\n";
    scopes_.clear();
    // GLSL passes a swizzle (p.xz) to an out/inout vector parameter by copy-in /
    // copy-out; a C++ reference cannot bind to it, so add an overload that takes
    // the swizzle, runs the function on a copy and writes the copy back
    bool wrap = false;
    for (auto& pi : pinfo) {
      if (pi.array) return o.str();
      wrap |= pi.ref && pi.vec;
    }
    if (wrap) {
      size_t nameIdx = lp;
      while (nameIdx > b && toks_[nameIdx].k != T_ID) --nameIdx;
      std::string ret = rewrite(b, nameIdx), fname = toks_[nameIdx].s;
      std::ostringstream tp, args, pre, post, call;
      int nt = 0;
      for (size_t q = 0; q < pinfo.size(); ++q) {
        const PInfo& pi = pinfo[q];
        std::string an = "snl_a" + std::to_string(q);
        if (q) { args << ", "; call << ", "; }
        if (pi.ref && pi.vec) {
          std::string tn = "SnlA" + std::to_string(q);
          tp << (nt++ ? ", " : "") << "class " << tn;
          args << tn << "& " << an;
          pre << " " << pi.type << " snl_t" << q << " = " << an << ";";
          post << " " << an << " = snl_t" << q << ";";
          call << "snl_t" << q;
        } else {
          args << pi.type << (pi.ref ? "& " : " ") << an;
          call << an;
        }
      }
      std::ostringstream cond;
      int c = 0;
      for (size_t q = 0; q < pinfo.size(); ++q)
        if (pinfo[q].ref && pinfo[q].vec)
          cond << (c++ ? " && " : "") << "snl_lv<SnlA" << q << ", " << pinfo[q].type << ">::ok";
      bool isVoid = ret.find("void") != std::string::npos;
      o << "template <" << tp.str() << ", typename snl_if<(" << cond.str() << ")>::type = 0>\n"
        << "SNL_HD " << ret << " " << fname << "(" << args.str() << ") {" << pre.str() << " "
        << (isVoid ? "" : "auto snl_r = ") << fname << "(" << call.str() << ");" << post.str()
        << (isVoid ? "" : " return snl_r;") << " }\n";
    }
    return o.str();
  }

  std::string src_;
  std::set<std::string> builtinUniforms_, defines_;
  std::vector<Tok> toks_;
  std::set<std::string> structs_;
  std::vector<Uniform> uniforms_;
  std::set<std::string> uniformNames_;
  std::set<std::string> macros_, preUniforms_;
  mutable std::vector<std::set<std::string>> scopes_;
};

}  // namespace

bool translateScene(const std::string& glsl, const std::set<std::string>& builtinUniforms,
                    const std::set<std::string>& defines, TranslateResult& out) {
  Translator t(glsl, builtinUniforms, defines);
  return t.run(out);
}

std::set<std::string> parseDefines(const std::string& defines) {
  std::set<std::string> s;
  std::istringstream in(defines);
  std::string line;
  while (std::getline(in, line)) {
    size_t p = line.find("#define");
    if (p == std::string::npos) continue;
    std::istringstream w(line.substr(p + 7));
    std::string name;
    w >> name;
    if (!name.empty()) s.insert(name);
  }
  return s;
}

}  // namespace snl
