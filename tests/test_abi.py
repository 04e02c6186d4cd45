"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol include/snelly.h declares, and the scene translator accepts the
reference's scene snippets (no GPU calls)."""
import ctypes
import os
import re

import pytest

from snelly_amd import host, native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "snelly.h")).read()
    return sorted(set(re.findall(r"\b(snl_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    assert native.lib().snl_version().startswith(b"snelly-mi355x")


def test_mega_variant_calls_reject_null():
    L = native.lib()
    assert L.snl_mega_variants(None) == -1  # SNL_E_ARG
    assert L.snl_set_mega_variant(None, 0) == -1


@pytest.mark.parametrize("make", [scenes.sphere_plane, scenes.menger_metal, scenes.mandelbox_dielectric])
def test_scene_compiles_for_gfx950(make):
    """translate + hipRTC compile of the benchmark scenes (cached code objects)"""
    s = make(8, 8, 1)
    sh, ior, d = host.assemble(s.glsl, s.renderer, s.materials)
    r = s.renderer
    native.precompile(sh, ior, d, r.maxMarchSteps, r.maxBounces, r.maxAtmosphereScatters, r.maxSamplesPerFrame)


def test_compile_error_is_reported():
    with pytest.raises(native.SnellyError) as e:
        native.precompile("float SDF_SURFACE(vec3 X) { return lenght(X) - 1.0; }", "", "#define HAS_SURFACE\n",
                          64, 1, 1, 1)
    assert "lenght" in str(e.value)


def test_scene_without_geometry_fails_like_the_reference():
    s = scenes.sphere_plane(8, 8, 1)
    with pytest.raises(native.SnellyError):
        host.assemble("float foo(vec3 X) { return 1.0; }", s.renderer, s.materials)


def kernel_private_bytes(path, kd):
    """the private-segment (scratch) bytes per lane in the kernel descriptor
    `kd` of an AMDGPU code object (ELF64: the descriptor's second u32)"""
    import struct
    b = open(path, "rb").read()
    assert b[:4] == b"\x7fELF" and b[4] == 2
    shoff, = struct.unpack_from("<Q", b, 0x28)
    shentsize, shnum = struct.unpack_from("<HH", b, 0x3A)
    sh = [struct.unpack_from("<IIQQQQIIQQ", b, shoff + i * shentsize) for i in range(shnum)]
    for name, typ, flags, addr, off, size, link, info, align, entsize in sh:
        if typ not in (2, 11):  # SHT_SYMTAB, SHT_DYNSYM
            continue
        stroff = sh[link][4]
        for k in range(size // 24):
            st_name, st_info, st_other, st_shndx, st_value, st_size = struct.unpack_from("<IBBHQQ", b, off + 24 * k)
            end = b.index(b"\0", stroff + st_name)
            if b[stroff + st_name:end].decode() == kd:
                sec = sh[st_shndx]
                priv, = struct.unpack_from("<I", b, sec[4] + (st_value - sec[3]) + 4)
                return priv
    raise KeyError(kd)


def test_megakernel_rebuilt_without_occupancy_bound_where_it_spills(tmp_path, monkeypatch):
    """the megakernel asks for four waves per SIMD; a program whose megakernel
    then spills (a Menger-4 dielectric batch job) is built again without the bound,
    and the program keeps that scratch-free build; one that does not spill (the
    Menger-4 benchmark scene) keeps the first build"""
    import numpy as np
    from snelly_amd import warm
    monkeypatch.setenv("SNL_KCACHE", str(tmp_path / "spills"))
    spills = scenes.batch_job("menger4", "DIELECTRIC", 32, 4, 6, np.random.RandomState(0))
    native.precompile(*warm.program_key(spills, native.BUILD_FAST))
    objs = sorted(tmp_path.glob("spills/*.hsaco"))
    assert len(objs) == 2
    assert sorted(kernel_private_bytes(str(p), "snl_mega.kd") for p in objs)[0] == 0
    assert max(kernel_private_bytes(str(p), "snl_mega.kd") for p in objs) > 0
    monkeypatch.setenv("SNL_KCACHE", str(tmp_path / "clean"))
    native.precompile(*warm.program_key(scenes.menger_metal(64, 64, 4), native.BUILD_FAST))
    objs = list(tmp_path.glob("clean/*.hsaco"))
    assert len(objs) == 1 and kernel_private_bytes(str(objs[0]), "snl_mega.kd") == 0


def test_state_bytes_counts_every_buffer_of_a_state():
    """radiance + RNG (32 B/px over all chains), the queue counters, the tile
    order's costs / histogram / order and the error estimate's partials"""
    from snelly_amd import native
    W, H, K = 1024, 768, 3
    tiles = (W // 8) * (H // 8)
    assert native.state_bytes(W, H, K) == W * H * K * 32 + 5 * 16 * 4 + (2 * tiles + 128) * 4 + 514 * 8
    assert native.state_bytes(0, 10, 1) == 0
