"""The reference's camera uniforms through three.js r79, restated (test
infrastructure: the checker of the hosts' Camera.basis, never the product).

js/renderer.js:520-527 derives camDir / camX / camY from a THREE.PerspectiveCamera
that OrbitControls points at its target (Camera.lookAt).  The reference vendors
three.js r79 (js/thirdparty/three/three.js); this file restates the functions
that chain runs through, in doubles as JavaScript computes them:

  Camera.lookAt(target)            three.js:17025-17036  m1.lookAt(position, target, up);
                                                          quaternion.setFromRotationMatrix(m1)
  Matrix4.lookAt                   three.js:5105-5150
  Quaternion.setFromRotationMatrix three.js:1048-1105
  Object3D.updateMatrix            three.js:8999-9005    matrix.compose(position, quaternion, scale)
  Matrix4.compose                  three.js:5585-5593    makeRotationFromQuaternion, scale, setPosition
  Matrix4.makeRotationFromQuaternion three.js:5068-5100
  Matrix4.decompose                three.js:5595-5650    (getWorldQuaternion, 8883-8900)
  Camera.getWorldDirection         three.js:17007-17021  (0, 0, -1).applyQuaternion(q)
  Vector3.applyQuaternion          three.js:2240-2260
  Vector3.transformDirection       three.js:2292-2306    camUp = up.transformDirection(matrixWorld)
  Vector3.normalize / divideScalar three.js:2460-2464, 2318-2322

Matrix4.elements is a Float32Array in r79 (three.js:4806-4815): every matrix
write rounds to float32 (`_store`), while vectors and quaternions stay doubles.
camera_uniforms(..., float32_matrices=False) gives the all-doubles variant, which
tests use to show that the rounding is visible in the uniforms.
"""
import math

import numpy as np

_store = [lambda v: float(np.float32(v))]


def _length(v):
    return math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])


def _normalize(v):
    s = 1.0 / _length(v) if _length(v) != 0.0 else math.inf
    return [c * s for c in v]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _look_at(eye, target, up):
    """Matrix4.lookAt: the 16 column-major elements (rotation part only set)"""
    z = _normalize([eye[k] - target[k] for k in range(3)])
    if z[0] * z[0] + z[1] * z[1] + z[2] * z[2] == 0:
        z[2] = 1.0
    x = _normalize(_cross(up, z))
    if x[0] * x[0] + x[1] * x[1] + x[2] * x[2] == 0:
        z[2] += 0.0001
        x = _normalize(_cross(up, z))
    y = _cross(z, x)
    st = _store[0]
    te = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]
    te[0], te[4], te[8] = st(x[0]), st(y[0]), st(z[0])
    te[1], te[5], te[9] = st(x[1]), st(y[1]), st(z[1])
    te[2], te[6], te[10] = st(x[2]), st(y[2]), st(z[2])
    return te


def _quaternion_from_rotation(te):
    m11, m12, m13 = te[0], te[4], te[8]
    m21, m22, m23 = te[1], te[5], te[9]
    m31, m32, m33 = te[2], te[6], te[10]
    trace = m11 + m22 + m33
    if trace > 0:
        s = 0.5 / math.sqrt(trace + 1.0)
        return [(m32 - m23) * s, (m13 - m31) * s, (m21 - m12) * s, 0.25 / s]
    if m11 > m22 and m11 > m33:
        s = 2.0 * math.sqrt(1.0 + m11 - m22 - m33)
        return [0.25 * s, (m12 + m21) / s, (m13 + m31) / s, (m32 - m23) / s]
    if m22 > m33:
        s = 2.0 * math.sqrt(1.0 + m22 - m11 - m33)
        return [(m12 + m21) / s, 0.25 * s, (m23 + m32) / s, (m13 - m31) / s]
    s = 2.0 * math.sqrt(1.0 + m33 - m11 - m22)
    return [(m13 + m31) / s, (m23 + m32) / s, 0.25 * s, (m21 - m12) / s]


def _compose(position, q):
    x, y, z, w = q
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    st = _store[0]
    te = [0.0] * 16
    te[0] = st(1 - (yy + zz)); te[4] = st(xy - wz); te[8] = st(xz + wy)
    te[1] = st(xy + wz); te[5] = st(1 - (xx + zz)); te[9] = st(yz - wx)
    te[2] = st(xz - wy); te[6] = st(yz + wx); te[10] = st(1 - (xx + yy))
    te[15] = 1.0
    # scale (1, 1, 1) multiplies by one: no change
    te[12], te[13], te[14] = [st(c) for c in position]
    return te


def _determinant3(te):
    return (te[0] * (te[5] * te[10] - te[9] * te[6]) - te[4] * (te[1] * te[10] - te[9] * te[2])
            + te[8] * (te[1] * te[6] - te[5] * te[2]))


def _decompose_quaternion(te):
    sx = _length([te[0], te[1], te[2]])
    sy = _length([te[4], te[5], te[6]])
    sz = _length([te[8], te[9], te[10]])
    if _determinant3(te) < 0:
        sx = -sx
    st = _store[0]
    m = list(te)
    for k, inv in ((0, 1 / sx), (4, 1 / sy), (8, 1 / sz)):
        m[k] = st(m[k] * inv)
        m[k + 1] = st(m[k + 1] * inv)
        m[k + 2] = st(m[k + 2] * inv)
    return _quaternion_from_rotation(m)


def _apply_quaternion(v, q):
    x, y, z = v
    qx, qy, qz, qw = q
    ix = qw * x + qy * z - qz * y
    iy = qw * y + qz * x - qx * z
    iz = qw * z + qx * y - qy * x
    iw = -qx * x - qy * y - qz * z
    return [ix * qw + iw * -qx + iy * -qz - iz * -qy,
            iy * qw + iw * -qy + iz * -qx - ix * -qz,
            iz * qw + iw * -qz + ix * -qy - iy * -qx]


def _transform_direction(v, te):
    x, y, z = v
    return _normalize([te[0] * x + te[4] * y + te[8] * z, te[1] * x + te[5] * y + te[9] * z,
                       te[2] * x + te[6] * y + te[10] * z])


def camera_uniforms(position, target, up, float32_matrices=True):
    """(camPos, camDir, camX, camY) as js/renderer.js:520-527 computes them for
    a camera at `position` looking at `target` with `up`"""
    _store[0] = (lambda v: float(np.float32(v))) if float32_matrices else float
    position, target, up = [float(c) for c in position], [float(c) for c in target], [float(c) for c in up]
    q = _quaternion_from_rotation(_look_at(position, target, up))
    world = _compose(position, q)
    cam_dir = _apply_quaternion([0.0, 0.0, -1.0], _decompose_quaternion(world))
    cam_up = _transform_direction(up, world)
    cam_x = _cross(cam_up, cam_dir)
    return position, cam_dir, cam_x, cam_up
