// translate.h -- GLSL scene snippet -> C++ (see translate.cpp)
#pragma once
#include <set>
#include <string>
#include <vector>

namespace snl {

struct UniformInfo {
  std::string name;
  int offset;  // byte offset in the user uniform block
  int type;    // SNL_T_* code
  int count;   // array length
  int size;    // bytes per element
};

struct TranslateResult {
  std::string code;                  // struct SnlUser + struct SnlScene
  std::vector<UniformInfo> uniforms; // user uniform block reflection
  int userBlockSize = 4;
  std::string error;
};

bool translateScene(const std::string& glsl, const std::set<std::string>& builtinUniforms,
                    const std::set<std::string>& defines, TranslateResult& out);
std::set<std::string> parseDefines(const std::string& defines);

}  // namespace snl
