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
