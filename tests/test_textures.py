"""Albedo textures and the alpha test of a frame through the C API on the GPU
against closed forms (the cases and their derivation: texture_cases.py)."""
import numpy as np
import pytest

from raytracedgl1_amd import rtgl
from raytracedgl1_amd.ffi import PLANE
import reflrefr_cases as rc
import texture_cases as tc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", list(tc.CASES))
def test_texture_closed_forms(case):
    pixels, filt, addr, repeat, alpha, back, check = tc.CASES[case]
    print(f"texture case {case}", flush=True)
    scene, uv = tc.quad_scene(repeat, back)
    r = rtgl.Renderer(rc.W, rc.H)
    try:
        r.provide_texture("tex", pixels(), filt, addr, addr)
        flags = {0: rtgl.RG_MESH_PRIMITIVE_ALPHA_TESTED} if alpha else {}
        r.frame(rtgl.SceneUploader(scene, flags, {0: "tex"}, uv), rc.camera(60.0))
        fb = r.framebuffers()
        out = {k: fb.download(PLANE[k]) for k in ("ALBEDO", "DEPTH_WORLD", "VISIBILITY_BUFFER")}
        out["params"] = r.frame_params()
    finally:
        r.close()
    check(out)


def test_alpha_tested_pane_shadows_only_where_kept():
    scene, uv = tc.shadow_scene()
    r = rtgl.Renderer(rc.W, rc.H)
    try:
        r.provide_texture("pane", tc.alpha_strip(), rtgl.RG_SAMPLER_FILTER_NEAREST, rtgl.RG_SAMPLER_ADDRESS_MODE_CLAMP,
                          rtgl.RG_SAMPLER_ADDRESS_MODE_CLAMP)
        r.frame(rtgl.SceneUploader(scene, {1: rtgl.RG_MESH_PRIMITIVE_ALPHA_TESTED}, {1: "pane"}, uv), tc.SHADOW_CAMERA)
        fb = r.framebuffers()
        tc.check_alpha_shadow(fb.download(PLANE["UNFILTERED_DIRECT"]), fb.download(PLANE["SURFACE_POSITION"]))
    finally:
        r.close()


def test_deleted_texture_falls_back_to_the_colour_and_names_resolve_at_upload():
    scene, uv = tc.quad_scene()
    r = rtgl.Renderer(rc.W, rc.H)
    try:
        r.provide_texture("tex", tc.checker8(), rtgl.RG_SAMPLER_FILTER_NEAREST)
        up = rtgl.SceneUploader(scene, {}, {0: "tex"}, uv)
        r.frame(up, rc.camera(60.0))
        textured = r.framebuffers().download(PLANE["ALBEDO"])[..., :3].copy()
        r.delete_texture("tex")
        r.frame(up, rc.camera(60.0))
        plain = r.framebuffers().download(PLANE["ALBEDO"])[..., :3].copy()
        r.provide_texture("tex", tc.checker8(), rtgl.RG_SAMPLER_FILTER_NEAREST)
        r.frame(up, rc.camera(60.0))
        again = r.framebuffers().download(PLANE["ALBEDO"])[..., :3].copy()
    finally:
        r.close()
    assert textured.std() > 0.05
    assert abs(plain - 1.0).max() < 1e-6  # white primitive colour, no texture
    assert (again == textured).all()


def test_material_maps(oracle):
    """tex_orm (UNORM) sets roughness and metallic, tex_n bends the shading
    normal (closed forms: texture_cases.check_orm / check_normal_map)."""
    scene, uv = tc.quad_scene()
    r = rtgl.Renderer(rc.W, rc.H)
    try:
        r.provide_texture("tex", tc.checker8(), rtgl.RG_SAMPLER_FILTER_NEAREST)
        r.provide_texture("tex_orm", tc.orm_map(), rtgl.RG_SAMPLER_FILTER_NEAREST)
        r.provide_texture("tex_n", tc.normal_map(), rtgl.RG_SAMPLER_FILTER_NEAREST)
        r.frame(rtgl.SceneUploader(scene, {}, {0: "tex"}, uv), rc.camera(60.0))
        fb = r.framebuffers()
        mr, nrm = fb.download(PLANE["METALLIC_ROUGHNESS"]), fb.download(PLANE["NORMAL"])
    finally:
        r.close()
    tc.check_orm(mr)
    tc.check_normal_map(nrm, lambda a: np.array([oracle.decode_normal(int(x)) for x in a.ravel()]).reshape(a.shape[:2] + (3,)))
