"""Sphere, spot and triangle lights and the light grid on the GPU, through
the C API, against the closed forms of light_cases.py (the same checks the
oracle passes in test_oracle_lights.py): one light over a plane, ReSTIR DI
reuse off, four frames of a still camera."""
import pytest

import light_cases as lc
from raytracedgl1_amd import ffi, rtgl

pytestmark = pytest.mark.gpu

W, H = 160, 90
P = ffi.PLANE


@pytest.mark.parametrize("name", sorted(lc.CASES))
def test_light_closed_form(name):
    print(f"lights: {name} over a plane, closed form, 4 frames at {W}x{H}", flush=True)
    lights, over = lc.CASES[name]
    r = rtgl.Renderer(W, H)
    try:
        r.path.restirTemporalSamples = 0
        r.path.restirSpatialSamples = 0
        r.set_illumination(cellWorldSize=1.0, polygonalLightSpotlightFactor=over.get("polyLightSpotlightFactor", 2.0))
        up = rtgl.SceneUploader(lc.plane_scene(lights))
        seen_in = False
        for f in range(4):
            r.frame(up, lc.CAM)
            p = r.frame_params()
            fb = r.framebuffers()
            m_in, _ = lc.check(name, fb.download(P["UNFILTERED_DIRECT"]), fb.download(P["SURFACE_POSITION"]), fb.download(P["IS_SKY"]),
                               f, p.cameraPosition, 1.0, p.polyLightSpotlightFactor, "gpu ")
            seen_in = seen_in or m_in == m_in
        assert seen_in
    finally:
        r.close()


def test_direct_reuse_keeps_the_mean_gpu():
    """test_oracle_reuse.test_direct_reuse_keeps_the_mean on the GPU: a sun
    over a square inside the light grid, ReSTIR DI 8/1/8 and 32/1/1; the
    mean direct light is colour cos(theta) / pi in every frame."""
    import math
    import numpy as np
    import test_oracle_reuse as tr
    from raytracedgl1_amd import scenes
    s = tr._square(1.5, 3.5, 4.0)
    scene = scenes.Scene(s.tris, s.normals, s.colors, s.objects, {"sun": tr.SUN})
    cos_t = -tr.SUN_D[1] / np.linalg.norm(tr.SUN_D)
    expect = 3.0 * cos_t / math.pi
    r_ = math.sin(max(0.01, math.radians(0.25)))
    jitter = (r_ * math.sqrt(1.0 - cos_t * cos_t) + 0.5 * r_ * r_) * 3.0 / math.pi
    for cand, temporal, spatial in ((8, 1, 8), (32, 1, 1)):
        print(f"lights: direct reuse {cand}/{temporal}/{spatial} on a sunlit square", flush=True)
        r = rtgl.Renderer(W, H)
        try:
            r.path.restirInitialCandidates, r.path.restirTemporalSamples, r.path.restirSpatialSamples = cand, temporal, spatial
            r.set_illumination(cellWorldSize=2.0)
            up = rtgl.SceneUploader(scene)
            for f in range(4):
                r.frame(up, tr.CAM)
                fb = r.framebuffers()
                ok = fb.download(P["IS_SKY"])[..., 0] == 0
                d = fb.download(P["UNFILTERED_DIRECT"])[..., 0][ok]
                empty = (fb.download(P["RESERVOIRS"])[..., 0][ok] & 0xFFFF) == tr.LIGHT_INDEX_NONE
                print(f"  frame {f}: mean / expected {d.mean() / expect:.6f}, {empty.sum()} empty reservoirs", flush=True)
                assert abs(d.mean() / expect - 1.0) < 0.2 * jitter / expect
                assert empty.sum() <= tr.MAX_EMPTY_FRACTION * ok.sum()
                assert np.abs(d[~empty] - expect).max() <= 4.0 * jitter
                assert (d[empty] == 0.0).all()
        finally:
            r.close()


def test_gradients_still_scene_and_removed_light_gpu():
    """test_oracle_reuse.test_gradients_still_scene_and_removed_light on the
    GPU: ~0 while nothing changes; 0.25 everywhere once the only light goes."""
    import math
    import numpy as np
    import test_oracle_reuse as tr
    from raytracedgl1_amd import scenes
    s = tr._square(0.0, 0.0, 50.0)
    lit = rtgl.SceneUploader(scenes.Scene(s.tris, s.normals, s.colors, s.objects, {"sun": tr.SUN}))
    dark = rtgl.SceneUploader(scenes.Scene(s.tris, s.normals, s.colors, s.objects, {}))
    cam = scenes.Camera(np.array([0.0, 6.0, 0.0]), np.array([0.5, 0.0, 1.5]), fovy=math.radians(50.0))
    print("lights: gradient samples of a still scene, then with the light removed", flush=True)
    r = rtgl.Renderer(W, H)
    try:
        for f in range(5):
            r.frame(dark if f == 4 else lit, cam)
            p = r.frame_params()
            fb = r.framebuffers()
            assert (fb.download(P["IS_SKY"]) == 0).all()
            g = fb.download(P["DIS_GRADIENT_HISTORY"])[..., :3]
            print(f"  frame {f}: diffuse {g[..., 0].min():.6f}..{g[..., 0].max():.6f}", flush=True)
            if 1 <= f <= 3:
                assert np.abs(g).max() < 1e-3
            if f == 4:
                assert np.abs(g[..., 0] - 0.5 * p.gradientMultDiffuse).max() < 1e-3
                assert np.abs(g[..., 2] - 0.5 * p.gradientMultSpecular).max() < 1e-3
    finally:
        r.close()
