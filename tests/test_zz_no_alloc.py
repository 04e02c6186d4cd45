"""snl_render_noalloc, the render entry point that allocates nothing, against
snl_render through the C ABI: the same program rendered into two identically
seeded states, one by each entry point, must give bit-identical radiance and
RNG states.  Two calls in a row, the second on a smaller image, so the second
call finds the first call's contents in the shared scratch array."""
import ctypes

import numpy as np
import pytest

from snelly_amd import host, native, scenes

pytestmark = pytest.mark.gpu


def _render_pair(ctx, sess, w, h, frames, seed):
    """the program of `sess` rendered `frames` frames into two fresh w x h
    states seeded alike: one through snl_render, one through snl_render_noalloc"""
    a = native.State(ctx, w, h, seed)
    b = native.State(ctx, w, h, seed)
    try:
        sess.scene.camera.aspect = w / h
        u = host.frame_uniforms(sess.scene.renderer, sess.scene.camera, sess.scene.scales, sess.scene.materials, w, h)
        for k in ("resolution", "camAspect"):
            sess.prog.set(k, u[k])
        ctx.check(native.lib().snl_render(sess.prog.h, a.h, frames))
        ctx.check(native.lib().snl_render_noalloc(sess.prog.h, b.h, frames))
        return (a.read(0), a.read(1)), (b.read(0), b.read(1))
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("engine", [native.ENGINE_MEGAKERNEL, native.ENGINE_PERSISTENT], ids=["megakernel", "persistent"])
def test_noalloc_render_matches_render(ctx, engine):
    scene = scenes.menger_metal(96, 64, 6)
    sess = host.Session(scene, 96, 64, ctx, flags=native.BUILD_FAST)
    try:
        sess.prog.engine(engine)
        # first call: a 96 x 64 image (12 x 8 tiles), tile cost order on
        (ra, ga), (rb, gb) = _render_pair(ctx, sess, 96, 64, 6, 1234)
        assert np.array_equal(ra, rb) and np.array_equal(ga, gb)
        assert np.all(ra[..., 3] == 7.0)
        # second call right after: a smaller image (5 x 4 tiles), fewer frames
        (ra, ga), (rb, gb) = _render_pair(ctx, sess, 40, 32, 3, 99)
        assert np.array_equal(ra, rb) and np.array_equal(ga, gb)
        assert np.all(ra[..., 3] == 4.0)
        # an image whose sides are not multiples of 8 renders in queue order
        (ra, ga), (rb, gb) = _render_pair(ctx, sess, 37, 23, 2, 5)
        assert np.array_equal(ra, rb) and np.array_equal(ga, gb)
    finally:
        sess.close()


def test_noalloc_render_rejects_what_it_cannot_hold(ctx):
    scene = scenes.sphere_plane(16, 16, 2)
    sess = host.Session(scene, 16, 16, ctx)
    try:
        # no metal: the default tabulated metal's kTex is never read, so the
        # program renders without a unit-0 snapshot and is accepted
        (ra, ga), (rb, gb) = _render_pair(ctx, sess, 16, 16, 3, 7)
        assert np.array_equal(ra, rb) and np.array_equal(ga, gb)
        # more 8x8 tiles than the scratch array holds: 4104 x 4096 = 513 x 512 tiles
        big = native.State(ctx, 4104, 4096, 1)
        try:
            assert 513 * 512 > 262144
            rc = native.lib().snl_render_noalloc(sess.prog.h, big.h, 1)
            assert rc == -1
            assert b"at most 262144" in native.lib().snl_last_error(ctx.h)
        finally:
            big.close()
        # the wavefront engine's path buffers are allocations
        sess.prog.engine(native.ENGINE_WAVEFRONT)
        rc = native.lib().snl_render_noalloc(sess.prog.h, sess.state.h, 1)
        assert rc == -5
        assert b"wavefront" in native.lib().snl_last_error(ctx.h)
    finally:
        sess.close()
    # a tabulated metal's K_METAL samples texture unit 0 (js/materials.js:665):
    # its per-frame radiance snapshot is an allocation
    scene = scenes.sphere_plane(16, 16, 2)
    scene.glsl = scenes.TORUS_GLSL.replace("@MAT@", "METAL")
    scene.materials.metal = host.metal("Gold")
    sess = host.Session(scene, 16, 16, ctx)
    try:
        rc = native.lib().snl_render_noalloc(sess.prog.h, sess.state.h, 1)
        assert rc == -5
        assert b"texture unit 0" in native.lib().snl_last_error(ctx.h)
    finally:
        sess.close()
