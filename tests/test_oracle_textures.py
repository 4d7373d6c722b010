"""The oracle's albedo textures (oracle/oracle_texture.cpp) and alpha test on
the CPU against what the mathematics fixes (the frame cases are shared with
the GPU tests: texture_cases.py):
  * mip chains: a constant texture stays constant at every level; a black /
    white checker averages to linear 0.5 (sRGB 188) from level 1 on; a 3x3
    level blits to its centre texel; level sizes halve with a floor of 1;
  * sampling: nearest and bilinear texels under magnification equal their
    closed forms; between two levels the result is the linear blend of the
    two levels' samples; the LOD of textureGrad is log2 of the longer scaled
    gradient;
  * frames: the nearest / bilinear / minified / alpha-tested cases and the
    alpha-tested pane's shadow."""
import numpy as np
import pytest

from raytracedgl1_amd import rtgl
from raytracedgl1_amd.ffi import PLANE
from oracle_lib import OracleFramebuffers, OraclePathTracer, OracleTexture
import reflrefr_cases as rc
import texture_cases as tc

NEAREST, LINEAR = rtgl.RG_SAMPLER_FILTER_NEAREST, rtgl.RG_SAMPLER_FILTER_LINEAR


def test_mip_chains(oracle):
    const = np.zeros((16, 8, 4), np.uint8)
    const[...] = (10, 200, 77, 128)
    m = OracleTexture(oracle, const).mips()
    assert [a.shape[:2] for a in m] == [(16, 8), (8, 4), (4, 2), (2, 1), (1, 1)]
    for a in m:
        assert (a == (10, 200, 77, 128)).all()
    m = OracleTexture(oracle, tc.fine_checker(16)).mips()
    for a in m[1:]:
        assert (a[..., :3] == 188).all() and (a[..., 3] == 255).all()
    odd = np.zeros((3, 3, 4), np.uint8)
    odd[1, 1] = (9, 99, 199, 255)
    m = OracleTexture(oracle, odd).mips()
    assert len(m) == 2 and (m[1][0, 0] == (9, 99, 199, 255)).all()


def test_sampling_closed_forms(oracle):
    px = tc.checker8()
    rng = np.random.default_rng(5)
    tn = OracleTexture(oracle, px, NEAREST)
    tl = OracleTexture(oracle, px, LINEAR)
    for u, v in rng.uniform(-2, 3, (200, 2)):
        if min((u * 8) % 1, 1 - (u * 8) % 1, (v * 8) % 1, 1 - (v * 8) % 1) < 1e-6:
            continue
        np.testing.assert_allclose(tn.sample_lod(u, v, -1.0)[:3], tc.expect_nearest(px, np.array(u), np.array(v)), atol=1e-12)
        np.testing.assert_allclose(tl.sample_lod(u, v, -1.0)[:3], tc.expect_bilinear(px, np.array([u]), np.array([v]))[0], atol=1e-12)
    # between levels 1 and 2: the linear blend of the two levels' samples
    big = OracleTexture(oracle, rng.integers(0, 256, (32, 32, 4), dtype=np.uint8), LINEAR)
    for u, v in rng.uniform(0, 1, (50, 2)):
        a, b = big.sample_lod(u, v, 1.0), big.sample_lod(u, v, 2.0)
        np.testing.assert_allclose(big.sample_lod(u, v, 1.25), 0.75 * a + 0.25 * b, atol=1e-12)
        # gradients of 4 and 3 texels along x and y: lod = log2(4) = 2
        np.testing.assert_allclose(big.sample_grad(u, v, 4 / 32, 0, 0, 3 / 32), b, atol=1e-12)


def _frame(oracle, scene, flags, textures, uv, pixels, params):
    fb = OracleFramebuffers(oracle, rc.W, rc.H)
    pt = OraclePathTracer(oracle, scene, rc.W, rc.H, flags=flags, textures=textures, texcoords=uv, texture_pixels=pixels)
    pt.frame(params, fb, threads=8)
    return fb.cur


@pytest.mark.parametrize("case", list(tc.CASES))
def test_oracle_texture_frames(oracle, case):
    pixels, filt, addr, repeat, alpha, back, check = tc.CASES[case]
    scene, uv = tc.quad_scene(repeat, back)
    flags = {0: rtgl.RG_MESH_PRIMITIVE_ALPHA_TESTED} if alpha else {}
    p = rc.frame_params(rc.camera(60.0))
    cur = _frame(oracle, scene, flags, {0: "tex"}, uv, {"tex": (pixels(), filt, addr, addr)}, p)
    out = {k: cur[PLANE[k]] for k in ("ALBEDO", "DEPTH_WORLD", "VISIBILITY_BUFFER")}
    out["params"] = p
    check(out)


def test_oracle_alpha_tested_pane_shadow(oracle):
    scene, uv = tc.shadow_scene()
    clamp = rtgl.RG_SAMPLER_ADDRESS_MODE_CLAMP
    cur = _frame(oracle, scene, {1: rtgl.RG_MESH_PRIMITIVE_ALPHA_TESTED}, {1: "pane"}, uv,
                 {"pane": (tc.alpha_strip(), NEAREST, clamp, clamp)}, rc.frame_params(tc.SHADOW_CAMERA))
    tc.check_alpha_shadow(cur[PLANE["UNFILTERED_DIRECT"]], cur[PLANE["SURFACE_POSITION"]])


def test_oracle_material_maps(oracle):
    scene, uv = tc.quad_scene()
    pixels = {"tex": (tc.checker8(), NEAREST, 0, 0), "tex_orm": (tc.orm_map(), NEAREST, 0, 0), "tex_n": (tc.normal_map(), NEAREST, 0, 0)}
    cur = _frame(oracle, scene, {}, {0: "tex"}, uv, pixels, rc.frame_params(rc.camera(60.0)))
    tc.check_orm(cur[PLANE["METALLIC_ROUGHNESS"]])
    tc.check_normal_map(cur[PLANE["NORMAL"]],
                        lambda a: np.array([oracle.decode_normal(int(x)) for x in a.ravel()]).reshape(a.shape[:2] + (3,)))
