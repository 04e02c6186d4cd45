"""Runtime object lifetimes and accounting through the C ABI: a texture
destroyed while a program's sampler is bound to it, a batch that repeats a
program, and the device bytes admission control counts."""
import numpy as np
import pytest

from snelly_amd import host, native, scenes

pytestmark = pytest.mark.gpu

TEXTURED = """
uniform sampler2D albedoTex;
float SDF_SURFACE(vec3 X) { return length(X) - 1.0; }
vec3 SURFACE_DIFFUSE_REFLECTANCE(in vec3 C, in vec3 X, in vec3 N, in vec3 V)
{
    return texture(albedoTex, vec2(0.5 * X.x + 0.5, 0.5 * X.y + 0.5)).rgb;
}
"""


def textured_scene(W=24, H=24):
    s = scenes.furnace(W, H, 4)
    s.glsl = TEXTURED
    s.renderer.filterRadius = 0.0  # no jitter: a pixel sees the sphere or the sky in every sample
    return s


def test_texture_destroyed_while_bound_reads_an_empty_unit(ctx):
    """after snl_texture_destroy the sampler reads (0, 0, 0, 1), as a WebGL unit
    without a texture does: the render equals one with an all-(0,0,0,1)
    texture bound, and never reads freed memory"""
    W = H = 24
    s = textured_scene(W, H)
    a = host.Session(s, W, H, ctx, seed=9, flags=native.BUILD_PRECISE)
    tex = native.Texture(ctx, np.full((4, 4, 4), 0.7, np.float32), 4, 4, 1, 4, True, True)
    a.prog.set_texture("albedoTex", tex)
    a.render(2)
    lit = a.radiance().copy()
    assert lit[..., 1].mean() > 0.1
    tex.close()                     # destroyed while still bound
    a.state.seed(9)
    a.render(2)
    gone = a.radiance().copy()
    b = host.Session(s, W, H, ctx, seed=9, flags=native.BUILD_PRECISE)
    empty = np.zeros((4, 4, 4), np.float32)
    empty[..., 3] = 1.0
    tex0 = native.Texture(ctx, empty, 4, 4, 1, 4, True, True)
    b.prog.set_texture("albedoTex", tex0)
    b.render(2)
    assert np.array_equal(gone, b.radiance())
    # the sphere's pixels go black (albedo 0); the ones that miss it see the sky
    sky = lit[..., 1] == lit[..., 1].max()
    assert (~sky).sum() > 100
    assert np.all(gone[~sky][:, :3] == 0.0)
    assert np.array_equal(gone[sky], lit[sky])
    for x in (a, b, tex0):
        x.close()


def test_texture_changed_while_a_submitted_job_samples_it(ctx):
    """a texture re-uploaded, then destroyed, right after a batch that samples
    it was submitted: the job first finishes with the texels it was submitted
    with (it equals the same render done before either call)"""
    W = H = 24
    s = textured_scene(W, H)
    ref = host.Session(s, W, H, ctx, seed=9, flags=native.BUILD_PRECISE)
    t0 = native.Texture(ctx, np.full((4, 4, 4), 0.7, np.float32), 4, 4, 1, 4, True, True)
    ref.prog.set_texture("albedoTex", t0)
    ref.render(3)
    want = ref.radiance().copy()
    a = host.Session(s, W, H, ctx, seed=9, flags=native.BUILD_PRECISE)
    tex = native.Texture(ctx, np.full((4, 4, 4), 0.7, np.float32), 4, 4, 1, 4, True, True)
    a.prog.set_texture("albedoTex", tex)
    b = native.Batch([(a.prog, a.state, 3)])
    tex.upload(np.full((4, 4, 4), 0.1, np.float32))
    tex.close()
    b.wait()
    b.release()
    assert np.array_equal(a.radiance(), want)
    for x in (a, ref, t0):
        x.close()


def test_batch_rejects_a_repeated_program(ctx):
    s = scenes.sphere_plane(16, 16, 2)
    a = host.Session(s, 16, 16, ctx)
    other = native.State(ctx, 16, 16, 5)
    with pytest.raises(native.SnellyError, match="program used twice"):
        native.render_batch([(a.prog, a.state, 1), (a.prog, other, 1)])
    other.close()
    a.close()


def test_program_state_bytes_counts_the_unit0_snapshot(ctx):
    """a tabulated metal samples texture unit 0 (js/materials.js:664-665): its
    states need a 16 B/px radiance snapshot; a program without a metal does not"""
    W, H, K = 64, 40, 2
    metal = scenes.sphere_plane(W, H, 2)
    metal.glsl = scenes.TORUS_GLSL.replace("@MAT@", "METAL")
    metal.materials.metal = host.metal("Gold")
    assert host.reads_texture_unit0(metal)
    m = host.Session(metal, W, H, ctx)
    plain = host.Session(scenes.sphere_plane(W, H, 2), W, H, ctx)
    assert not host.reads_texture_unit0(plain.scene)
    base = native.state_bytes(W, H, K)
    assert native.state_bytes(W, H, K, m.prog) == base + 16 * W * H * K
    assert native.state_bytes(W, H, K, plain.prog) == base
    m.close()
    plain.close()
