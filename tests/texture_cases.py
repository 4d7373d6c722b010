"""Closed-form cases of albedo textures and the alpha test (HitInfo.inl:28-128,
RtAlphaTest.rahit), shared by the GPU tests and the oracle's CPU tests.

A camera at the origin looks down -Z at a quad in the plane z = -5 whose
texture coordinates are an affine function of x and y, so the texture
coordinate of every pixel's hit is known in closed form:
  * nearest filtering under magnification returns the texel's sRGB-decoded
    value (IEC 61966-2-1), times the primitive colour;
  * bilinear filtering under magnification is the textbook bilinear weights
    of the four texels around u*w - 0.5, v*h - 0.5 (repeat wrap);
  * a checker of black and white texels seen from far away is sampled from
    the mips, which average in linear space: 0.5, stored as sRGB 188, read
    back as 0.5029 (an average of the sRGB codes would read back as 0.21);
  * the alpha test discards where (r+g+b)/3 * a + a < 0.5 (a texel of white
    at alpha 60/255 is discarded, at 66/255 kept), in primary and shadow rays.
"""
import math

import numpy as np

from raytracedgl1_amd import rtgl, scenes
from reflrefr_cases import W, H, _albedo, _quad, _to_regular, _camera_dirs, camera

QX, QY = 6.0, 4.0  # the quad spans [-QX, QX] x [-QY, QY] at z = -5


def srgb_to_linear(c8):
    c = np.asarray(c8, np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def quad_scene(repeat=1.0, with_back=False):
    """The textured quad (object 0) and, optionally, a far wall (object 1,
    colour (0.9, 0.3, 0.1)) at z = -10.  u = repeat * (x + QX) / 2QX,
    v = repeat * (QY - y) / 2QY."""
    q = _quad((0, 0, -5), (QX, 0, 0), (0, QY, 0), (0, 0, 1), (1.0, 1.0, 1.0))
    parts, objects = [q], [("q", 0, 2)]
    if with_back:
        parts.append(_quad((0, 0, -10), (100, 0, 0), (0, 100, 0), (0, 0, 1), (0.9, 0.3, 0.1)))
        objects.append(("far", 2, 2))
    s = scenes._finish(parts, objects)
    pos = s.tris[0:2].reshape(-1, 3).astype(np.float64)
    uv = np.stack([repeat * (pos[:, 0] + QX) / (2 * QX), repeat * (QY - pos[:, 1]) / (2 * QY)], 1)
    return s, {0: uv.astype(np.float32)}


def hit_uv(d, repeat=1.0):
    """Texture coordinates of each pixel's hit on the quad (camera ray d)."""
    t = 5.0 / np.abs(d[..., 2])
    x, y = d[..., 0] * t, d[..., 1] * t
    return repeat * (x + QX) / (2 * QX), repeat * (QY - y) / (2 * QY)


def checker8():
    """8x8 texels, texel (row j, column i) = sRGB (32 i, 32 j, 128), alpha 255."""
    j, i = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    px = np.zeros((8, 8, 4), np.uint8)
    px[..., 0], px[..., 1], px[..., 2], px[..., 3] = 32 * i, 32 * j, 128, 255
    return px


def expect_nearest(px, u, v):
    h, w = px.shape[:2]
    i = np.floor(u * w).astype(int) % w
    j = np.floor(v * h).astype(int) % h
    return srgb_to_linear(px[j, i, :3])


def expect_bilinear(px, u, v):
    h, w = px.shape[:2]
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    ax, ay = (x - x0)[..., None], (y - y0)[..., None]
    x0, y0 = x0.astype(int), y0.astype(int)
    lin = srgb_to_linear(px[..., :3])
    t = lambda dx, dy: lin[(y0 + dy) % h, (x0 + dx) % w]
    return (t(0, 0) * (1 - ax) + t(1, 0) * ax) * (1 - ay) + (t(0, 1) * (1 - ax) + t(1, 1) * ax) * ay


def boundary_mask(u, v, w, h, margin):
    """Pixels whose hit is at least `margin` texels from a texel edge."""
    fu, fv = (u * w) % 1.0, (v * h) % 1.0
    return (fu > margin) & (fu < 1 - margin) & (fv > margin) & (fv < 1 - margin)


def fine_checker(n=256):
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    c = np.where((i + j) % 2 == 0, 0, 255).astype(np.uint8)
    return np.stack([c, c, c, np.full_like(c, 255)], -1)


def alpha_strip():
    """Two texels: left white at alpha 60 (discarded), right white at 66 (kept)."""
    return np.array([[[255, 255, 255, 60], [255, 255, 255, 66]]], np.uint8)


# ---- checks over the planes (albedo regular layout; others checkerboard)

def check_nearest(out):
    d = _camera_dirs(out["params"])
    u, v = hit_uv(d)
    px = checker8()
    m = boundary_mask(u, v, 8, 8, 0.02)
    np.testing.assert_allclose(out["ALBEDO"][..., :3][m], expect_nearest(px, u, v)[m], atol=2e-6)


def check_bilinear(out):
    d = _camera_dirs(out["params"])
    u, v = hit_uv(d)
    np.testing.assert_allclose(out["ALBEDO"][..., :3], expect_bilinear(checker8(), u, v), atol=2e-4)


def check_minified(out):
    # every pixel covers dozens of texels: lod >= 1, where every level is 188
    got = out["ALBEDO"][..., :3]
    np.testing.assert_allclose(got, np.full_like(got, srgb_to_linear(188)), atol=2e-6)


def check_alpha(out):
    d = _camera_dirs(out["params"])
    u, _ = hit_uv(d)
    az = np.abs(d[..., 2])
    depth = _to_regular(out["DEPTH_WORLD"].reshape(H, W))
    keep, drop = u > 0.5 + 1e-3, u < 0.5 - 1e-3
    assert keep.sum() > 100 and drop.sum() > 100
    np.testing.assert_allclose(depth[keep], 5.0 / az[keep], rtol=1e-4)
    np.testing.assert_allclose(depth[drop], 10.0 / az[drop], rtol=1e-4)
    np.testing.assert_allclose(out["ALBEDO"][..., :3][drop], np.broadcast_to(_albedo((0.9, 0.3, 0.1)), out["ALBEDO"][..., :3][drop].shape), atol=1e-6)


CASES = {
    # name: (texture pixels, filter, address, repeat, alpha tested, with back wall, check)
    "nearest": (checker8, rtgl.RG_SAMPLER_FILTER_NEAREST, rtgl.RG_SAMPLER_ADDRESS_MODE_REPEAT, 1.0, False, False, check_nearest),
    "bilinear": (checker8, rtgl.RG_SAMPLER_FILTER_LINEAR, rtgl.RG_SAMPLER_ADDRESS_MODE_REPEAT, 1.0, False, False, check_bilinear),
    "minified": (fine_checker, rtgl.RG_SAMPLER_FILTER_LINEAR, rtgl.RG_SAMPLER_ADDRESS_MODE_REPEAT, 64.0, False, False, check_minified),
    "alpha": (alpha_strip, rtgl.RG_SAMPLER_FILTER_NEAREST, rtgl.RG_SAMPLER_ADDRESS_MODE_CLAMP, 1.0, True, True, check_alpha),
}


def shadow_scene():
    """A floor under a straight-down sun and an alpha-tested pane above the
    camera's view whose left half (x < 0) is discarded."""
    floor = _quad((0, 0, 0), (50, 0, 0), (0, 0, 50), (0, 1, 0), (0.8, 0.8, 0.8))
    pane = _quad((0, 6, 0), (1, 0, 0), (0, 0, 1), (0, 1, 0), (1.0, 1.0, 1.0))
    sun = {"color": (1.0, 1.0, 1.0), "intensity": 5.0, "direction": (0.0, -1.0, 0.0), "angularDiameterDegrees": 0.5}
    s = scenes._finish([floor, pane], [("floor", 0, 2), ("pane", 2, 2)], {"sun": sun})
    pos = s.tris[2:4].reshape(-1, 3).astype(np.float64)
    uv = np.stack([(pos[:, 0] + 1) / 2, (pos[:, 2] + 1) / 2], 1).astype(np.float32)
    return s, {1: uv}


SHADOW_CAMERA = scenes.Camera(np.array([0.0, 3.0, 3.0]), np.array([0.0, 0.0, 0.0]), fovy=math.radians(60.0))


def check_alpha_shadow(direct_cb, pos_cb):
    direct = _to_regular(direct_cb)[..., :3].sum(-1)
    pos = _to_regular(pos_cb)[..., :3]
    flat = np.abs(pos[..., 1]) < 1e-3
    under_kept = flat & (pos[..., 0] > 0.2) & (pos[..., 0] < 0.8) & (np.abs(pos[..., 2]) < 0.8)
    under_dropped = flat & (pos[..., 0] < -0.2) & (pos[..., 0] > -0.8) & (np.abs(pos[..., 2]) < 0.8)
    around = flat & (np.abs(pos[..., 0]) > 1.3) & (np.abs(pos[..., 0]) < 3) & (np.abs(pos[..., 2]) < 3)
    assert under_kept.sum() > 30 and under_dropped.sum() > 30 and around.sum() > 30
    lit = direct[around].mean()
    print(f"alpha shadow: kept {direct[under_kept].mean() / lit:.4f}, discarded {direct[under_dropped].mean() / lit:.4f}")
    assert direct[under_kept].mean() / lit < 0.05
    assert direct[under_dropped].mean() / lit > 0.9


# ---- material maps: N_orm (UNORM), N_n (UNORM), N_e (sRGB) beside N

def orm_map():
    """Roughness 128 / 255 and metallic 1 everywhere (read as UNORM: 0.502,
    not the sRGB-decoded 0.216)."""
    px = np.zeros((4, 4, 4), np.uint8)
    px[..., 0], px[..., 1], px[..., 2], px[..., 3] = 255, 128, 255, 255
    return px


def normal_map():
    """A constant tangent-space normal (255, 128) -> (1, 0.0039)."""
    px = np.zeros((4, 4, 4), np.uint8)
    px[..., 0], px[..., 1], px[..., 2], px[..., 3] = 255, 128, 255, 255
    return px


def check_orm(mr_cb):
    mr = _to_regular(mr_cb)
    np.testing.assert_allclose(mr[..., 0], 1.0, atol=1e-6)        # metallic = b
    np.testing.assert_allclose(mr[..., 1], 128 / 255, atol=1e-6)  # roughness = g


def check_normal_map(normal_cb, decode):
    """The quad's texture u grows with +x and v with -y, so the tangent is +x
    and the bitangent cross(n, t) * w = (0, -1, 0) (w = -1: v runs against
    cross(n, t)); the bent normal is normalize(t * 1 + b * 0.0039 + n)."""
    nx, ny = 255 / 255 * 2 - 1, 128 / 255 * 2 - 1
    want = np.array([nx, -ny, 1.0])
    want /= np.linalg.norm(want)
    got = decode(_to_regular(normal_cb))
    np.testing.assert_allclose(got.reshape(-1, 3), np.broadcast_to(want, (got.size // 3, 3)), atol=2e-4)
