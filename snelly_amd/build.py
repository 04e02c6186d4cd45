"""Build the native parts of snelly_amd in-tree.

  libsnelly_hip.so   C ABI runtime (runtime.hip + translate.cpp), hipcc for gfx950
  _build/embedded_sources.h  the kernel headers embedded for hipRTC

Usage: python3 snelly_amd/build.py [--precompile]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsnelly_hip.so")

# headers that hipRTC programs #include, in dependency order
EMBEDDED = ["glsl.h", "glsl_swz2.h", "glsl_swz3.h", "glsl_swz4.h", "snl_params.h", "snl_prelude.h",
            "snl_wavefront.h", "snl_pt.h", "snl_kernels.h", "snl_persist.h", "snl_mega.h", "snl_modes.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled for every gfx950 code object we build
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def gen_embedded():
    os.makedirs(BUILD, exist_ok=True)
    out = ["// generated by snelly_amd/build.py -- do not edit", "#pragma once"]
    names = []
    for i, name in enumerate(EMBEDDED):
        with open(os.path.join(CSRC, name)) as f:
            text = f.read()
        assert ")SNLSRC\"" not in text
        out.append(f'static const char kEmb{i}[] = R"SNLSRC({text})SNLSRC";')
        names.append(name)
    out.append(f"static const int kNumEmbedded = {len(EMBEDDED)};")
    out.append("static const char* kEmbeddedSources[] = {" + ", ".join(f"kEmb{i}" for i in range(len(EMBEDDED))) + "};")
    out.append("static const char* kEmbeddedNames[] = {" + ", ".join(f'"{n}"' for n in names) + "};")
    path = os.path.join(BUILD, "embedded_sources.h")
    new = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != new:
        with open(path, "w") as f:
            f.write(new)
    return path


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_runtime():
    gen_embedded()
    srcs = [os.path.join(CSRC, "runtime.hip"), os.path.join(CSRC, "translate.cpp")]
    objs = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-fPIC", "-std=c++17", "-I", CSRC, "-I", BUILD,
               "-Wno-unused-command-line-argument", "-c", s, "-o", o]
        if s.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
            cmd += NO_PACKED_FP32
        else:
            cmd[1:1] = ["-x", "c++"]
            cmd.remove("--offload-arch=gfx950")
        run(cmd)
        objs.append(o)
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"])
    os.replace(tmp, LIB)


NODE_INCLUDE = "/usr/include/node"
ADDON = os.path.join(HERE, "snelly_napi.node")


def build_addon():
    """The Node-API addon (snelly_napi.node) over libsnelly_hip.so, for the
    JavaScript host in js/.  Skipped when Node's headers are not installed."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        print("build: node_api.h not found, skipping the Node addon")
        return
    tmp = ADDON + ".tmp"
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DNODE_GYP_MODULE_NAME=snelly_napi", "-I", NODE_INCLUDE,
         os.path.join(HERE, "napi", "snelly_napi.cpp"), "-o", tmp, "-L", HERE, "-lsnelly_hip",
         "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, ADDON)


def main():
    build_runtime()
    build_addon()


if __name__ == "__main__":
    main()
