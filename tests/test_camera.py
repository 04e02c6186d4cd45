"""The camera uniforms (row a3): camDir / camX / camY of js/renderer.js:520-527,
computed by the hosts (host.Camera.basis and the JS host's PerspectiveCamera)
against the oracle's restatement of the three.js r79 chain the reference runs
through (oracle/three_camera.py)."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from snelly_amd import host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import three_camera  # noqa: E402

NODE = shutil.which("node")

# exampleScenes/animation/bounce.html:40-42: an up vector far from (0, 1, 0)
BOUNCE = ((3.2296072966816123, 5.458482159181951, 539.5128069568301),
          (-6.234516366353484, 4.031668852780276, 549.5128069568301),
          (-0.7150442285646916, -0.10309210646584782, -0.6914360192966225))


def cameras(n=200, seed=5):
    rs = np.random.RandomState(seed)
    out = [((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)), BOUNCE,
           ((2.6, 2.0, 3.4), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))]
    for _ in range(n):
        pos = rs.uniform(-10, 10, 3)
        tgt = rs.uniform(-1, 1, 3)
        up = rs.normal(size=3) if rs.rand() < 0.5 else np.array([0.0, 1.0, 0.0])
        out.append((tuple(pos), tuple(tgt), tuple(up)))
    return out


def host_basis(pos, tgt, up):
    c = host.Camera()
    c.position, c.target, c.up = list(pos), list(tgt), list(up)
    return c.basis()


def test_default_camera_known_answer():
    """camera on +z looking at the origin, up +y: looks down -z, camY = +y and
    camX = camY x camDir = -x"""
    _, d, x, y = host_basis((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    assert list(d) == [0.0, 0.0, -1.0] and list(y) == [0.0, 1.0, 0.0] and list(x) == [-1.0, 0.0, 0.0]


def test_host_matches_the_three_js_chain():
    """the host follows the chain with r79's float32 matrix elements and agrees
    with the restatement exactly; without the float32 stores the uniforms would
    differ (so the stores are visible in float32 and the check has power).  The
    chain against the executed reference is tests/test_host_vectors.py."""
    cams = cameras()
    differ = 0
    for pos, tgt, up in cams:
        h = host_basis(pos, tgt, up)
        r = three_camera.camera_uniforms(pos, tgt, up)
        d = three_camera.camera_uniforms(pos, tgt, up, float32_matrices=False)
        for a, b, c in zip(h[1:], r[1:], d[1:]):
            assert list(np.asarray(a)) == list(np.asarray(b))
            differ += not np.array_equal(np.asarray(b, np.float32), np.asarray(c, np.float32))
    assert differ >= 0.2 * 3 * len(cams)


def test_up_is_carried_through_the_camera_rotation():
    """renderer.js:524-527: camY = up.transformDirection(matrixWorld), which is
    the lookAt frame's y only for up = (0, 1, 0); camX = camY x camDir is then
    not a unit vector.  bounce.html's camera shows it"""
    pos, tgt, up = BOUNCE
    _, d, x, y = host_basis(pos, tgt, up)
    z = np.asarray(pos) - np.asarray(tgt)
    z /= np.linalg.norm(z)
    xl = np.cross(up, z)
    xl /= np.linalg.norm(xl)
    y_lookat = np.cross(z, xl)
    assert np.abs(y - y_lookat).max() > 0.1
    assert abs(np.linalg.norm(x) - 1.0) > 0.1
    r = three_camera.camera_uniforms(pos, tgt, up)
    assert np.allclose(x, r[2], rtol=0, atol=1e-15) and np.allclose(y, r[3], rtol=0, atol=1e-15)


@pytest.mark.skipif(NODE is None, reason="node is missing")
def test_js_host_computes_the_same_doubles():
    cams = cameras(40, seed=9)
    code = ("const h = require(%s); const cams = %s; const out = cams.map(([p, t, u]) => {"
            " const c = new h.THREE.PerspectiveCamera(45, 1, 0.1, 100);"
            " c.position.set(...p); c.up.set(...u); c.lookAt(new h.THREE.Vector3(...t));"
            " const b = c.basis(); return [b.dir.toArray(), b.camX.toArray(), b.camY.toArray()]; });"
            " console.log(JSON.stringify(out));") % (
        json.dumps(os.path.join(ROOT, "snelly_amd", "js", "snelly_host.js")),
        json.dumps([[list(p), list(t), list(u)] for p, t, u in cams]))
    out = subprocess.run([NODE, "-e", code], capture_output=True, text=True, check=True).stdout
    js = json.loads(out.strip().splitlines()[-1])
    for (pos, tgt, up), (d, x, y) in zip(cams, js):
        h = host_basis(pos, tgt, up)
        assert list(h[1]) == d and list(h[2]) == x and list(h[3]) == y
