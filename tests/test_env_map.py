"""Environment maps: the image decode and sRGB conversion of the hosts.

The lighting itself (environmentRadianceRGB, pathtracer-fragment-shader.glsl:
1051-1070, and the normals program's lookup) is pinned by the `env_map`
reference fixture in test_reference_fixtures.py.
"""
import json
import os
import shutil
import struct
import subprocess
import zlib

import numpy as np
import pytest

from snelly_amd import host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_srgb_closed_form():
    """the sRGB transfer function at its two branches and the ends"""
    b = np.array([0, 10, 11, 128, 255], np.uint8)
    got = host.srgb_to_linear(b)
    c = b / 255.0
    want = [0.0, c[1] / 12.92, ((c[2] + 0.055) / 1.055) ** 2.4, ((c[3] + 0.055) / 1.055) ** 2.4, 1.0]
    assert np.allclose(got, want, rtol=1e-7, atol=0)
    assert got[0] == 0.0 and got[-1] == 1.0
    assert abs(got[3] - 0.2158605) < 1e-6  # sRGB 128 -> linear 0.2158605


def encode_png(img, bpp):
    """a PNG encoder with filter type y % 5 on row y, so that the decoder is
    exercised on all five filters"""
    h, w = img.shape[:2]
    rows = img.reshape(h, w * bpp).astype(np.int32)
    raw = bytearray()
    prev = np.zeros(w * bpp, np.int32)
    for y in range(h):
        ft = y % 5
        cur = rows[y]
        out = np.zeros_like(cur)
        for x in range(w * bpp):
            a = cur[x - bpp] if x >= bpp else 0
            bb = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            if ft == 0: pred = 0
            elif ft == 1: pred = a
            elif ft == 2: pred = bb
            elif ft == 3: pred = (a + bb) // 2
            else:
                p = a + bb - c
                pa, pb, pc = abs(p - a), abs(p - bb), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (bb if pb <= pc else c)
            out[x] = (cur[x] - pred) & 255
        raw.append(ft)
        raw += bytes(out.astype(np.uint8))
        prev = cur

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xffffffff)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2 if bpp == 3 else 6, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


@pytest.mark.parametrize("bpp", [3, 4])
def test_png_decode_all_filters(tmp_path, bpp):
    rs = np.random.RandomState(bpp)
    img = rs.randint(0, 256, (11, 7, bpp)).astype(np.uint8)
    p = tmp_path / "t.png"
    p.write_bytes(encode_png(img, bpp))
    w, h, rgb = host.load_png(str(p))
    assert (w, h) == (7, 11)
    assert np.array_equal(rgb, img[:, :, :3])


def js_decode(path):
    code = ("const h = require(%r); const r = h.loadPng(%r); const f = h.srgbToLinear(r.rgb);"
            "console.log(JSON.stringify({w: r.width, h: r.height, rgb: Array.from(r.rgb), lin: Array.from(f)}));"
            % (os.path.join(ROOT, "snelly_amd", "js", "snelly_host.js"), path))
    out = subprocess.run(["node", "-e", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout)


@pytest.mark.skipif(shutil.which("node") is None or
                    not os.path.exists(os.path.join(ROOT, "snelly_amd", "snelly_napi.node")),
                    reason="needs node and the built addon")
def test_js_and_python_decoders_agree(tmp_path):
    """the two hosts decode the same PNG (all five filters) to the same bytes
    and the same float32 texels"""
    rs = np.random.RandomState(9)
    img = rs.randint(0, 256, (9, 13, 3)).astype(np.uint8)
    p = tmp_path / "e.png"
    p.write_bytes(encode_png(img, 3))
    j = js_decode(str(p))
    w, h, rgb = host.load_png(str(p))
    assert (j["w"], j["h"]) == (w, h)
    assert np.array_equal(np.array(j["rgb"], np.uint8), rgb.reshape(-1))
    assert np.array_equal(np.array(j["lin"], np.float32), host.srgb_to_linear(rgb).reshape(-1))


def test_save_png_round_trip(tmp_path):
    """save_png writes row 0 (the bottom row) last; load_png reads it back"""
    rs = np.random.RandomState(4)
    img = rs.randint(0, 256, (6, 5, 4)).astype(np.uint8)
    p = tmp_path / "o.png"
    host.save_png(str(p), img)
    w, h, rgb = host.load_png(str(p))
    assert (w, h) == (5, 6)
    assert np.array_equal(rgb, img[::-1, :, :3])
