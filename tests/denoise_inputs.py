"""Seeded synthetic G-buffers for the A-SVGF parity tests.

Depth and normals are constant over 8x8 blocks with values chosen so that
every reprojection test (|dz|/z < 0.1, dot(n, n') > 0.95) is far from its
threshold; motion vectors move by a whole number of pixels plus a fraction
in [0.25, 0.75].  Then FP32 and FP64 take the same branch everywhere and the
outputs can be compared with a tight tolerance.
"""
from __future__ import annotations

import numpy as np

from raytracedgl1_amd.ffi import PLANE, PLANES, plane_shape

SKY_DEPTH = 20000.0  # MAX_RAY_LENGTH * 2, written for sky pixels (RaygenPrimary.inl:122)


def encode_normal(n: np.ndarray) -> np.ndarray:
    """Utils.h:108-124, vectorised in float64."""
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)
    phi = np.arccos(np.clip(n[..., 2], -1, 1))
    theta = np.arctan2(n[..., 1], n[..., 0])
    theta = np.where(theta < 0, theta + 2 * np.pi, theta)
    j = np.round(phi * 65535 / np.pi).astype(np.uint64)
    k = np.round(theta * 65536 / (2 * np.pi)).astype(np.uint64) % 65536
    return ((j << np.uint64(16)) | k).astype(np.uint32)


# pairwise dot products of these directions are all below 0.8
_NORMALS = np.array([[0, 1, 0], [0.6, 0.8, 0], [-0.6, 0.8, 0], [0, 0.8, 0.6], [0, 0.8, -0.6],
                     [0.9, 0.1, 0.42], [-0.5, 0.2, -0.84]], np.float64)
_DEPTHS = np.array([3.0, 4.5, 7.0, 11.0, 17.0, 26.0])


def make_gbuffer(width: int, height: int, seed: int, sky_fraction: float = 0.05, block: int = 8,
                 motion_px=(1, 0), motion_frac=(0.4, 0.6)):
    r = np.random.default_rng(seed)
    H, W = height, width
    by, bx = (H + block - 1) // block, (W + block - 1) // block
    bn = r.integers(0, len(_NORMALS), (by, bx))
    bd = r.integers(0, len(_DEPTHS), (by, bx))
    bsky = r.uniform(0, 1, (by, bx)) < sky_fraction
    up = lambda a: np.repeat(np.repeat(a, block, 0), block, 1)[:H, :W]
    n_idx, d_idx, sky = up(bn), up(bd), up(bsky)
    normals = encode_normal(_NORMALS[n_idx])
    depth = np.where(sky, SKY_DEPTH, _DEPTHS[d_idx]).astype(np.float32)

    cur, prev = {}, {}

    def put(name, arr, p=None):
        cur[PLANE[name]] = np.ascontiguousarray(arr).reshape(plane_shape(PLANE[name], W, H))
        if p is not None:
            prev[PLANE[name]] = np.ascontiguousarray(p).reshape(plane_shape(PLANE[name], W, H))

    # prev frame: the same blocks displaced by motion_px (in checkerboard pixels)
    dx, dy = motion_px
    prev_normals = np.roll(normals, (dy, dx), (0, 1))
    prev_depth = np.roll(depth, (dy, dx), (0, 1))
    put("NORMAL", normals, prev_normals)
    put("DEPTH_WORLD", depth, prev_depth)
    put("DEPTH_GRAD", r.uniform(0.01, 0.3, (H, W)).astype(np.float32))
    mr = np.stack([r.uniform(0, 1, (H, W)), r.uniform(0.02, 1, (H, W))], -1).astype(np.float32)
    put("METALLIC_ROUGHNESS", mr, mr[::-1].copy())
    # motion in UV: the checkerboard half is W/2 wide (getPrevScreenPos)
    frac = r.uniform(motion_frac[0], motion_frac[1], (H, W))
    mot = np.zeros((H, W, 4), np.float32)
    mot[..., 0] = (dx + frac) / (W / 2)
    mot[..., 1] = (dy + r.uniform(motion_frac[0], motion_frac[1], (H, W))) / H
    put("MOTION", mot)
    for name in ("UNFILTERED_DIRECT", "UNFILTERED_SPECULAR", "UNFILTERED_INDIR"):
        c = np.zeros((H, W, 4), np.float32)
        c[..., :3] = np.exp(r.normal(0, 1.0, (H, W, 3)))
        put(name, c)
    pos = np.zeros((H, W, 4), np.float32)
    pos[..., :3] = r.uniform(-10, 10, (H, W, 3))
    pos[..., 2] -= 20.0
    put("SURFACE_POSITION", pos, pos + np.float32(0.05))
    vd = r.normal(size=(H, W, 3))
    vd /= np.linalg.norm(vd, axis=-1, keepdims=True)
    view = np.zeros((H, W, 4), np.float32)
    view[..., :3] = vd
    view[..., 3] = r.uniform(0.5, 30, (H, W))
    put("VIEW_DIRECTION", view)
    th = np.zeros((H, W, 4), np.float32)
    th[..., :3] = r.uniform(0.5, 1.0, (H, W, 3))
    put("THROUGHPUT", th)
    put("IS_SKY", sky.astype(np.uint32))
    hl = np.zeros((H, W, 4), np.float32)
    hl[..., 0] = r.uniform(1, 30, (H, W))
    hl[..., 1] = r.uniform(1, 30, (H, W))
    hl[..., 2] = r.uniform(1, 30, (H, W))
    hl[..., 3] = r.uniform(0.5, 30, (H, W))
    put("ACCUM_HISTORY_LENGTH", hl, hl[:, ::-1].copy())
    m0 = r.uniform(0.1, 3, (H, W))
    mom = np.stack([m0, m0 * m0 * r.uniform(1.0, 2.0, (H, W))], -1).astype(np.float32)
    put("DIFF_ACCUM_MOMENTS", mom, mom[::-1].copy())
    for name in ("DIFF_COLOR_HISTORY", "SPEC_ACCUM_COLOR", "INDIR_ACCUM"):
        c = np.zeros((H, W, 4), np.float32)
        c[..., :3] = np.exp(r.normal(0, 1.0, (H, W, 3)))
        if PLANES[PLANE[name]][4]:
            put(name, np.zeros_like(c), c)
        else:
            put(name, c)
    alb = np.zeros((H, W, 4), np.float32)
    alb[..., :3] = r.uniform(0.0, 1.0, (H, W, 3))
    put("ALBEDO", alb)
    gw, gh = max(1, (W + 1) // 3), max(1, (H + 1) // 3)
    g = np.zeros((gh, gw, 4), np.float32)
    g[..., :3] = r.uniform(0, 1, (gh, gw, 3))
    put("DIS_PING_GRADIENT", g)
    return cur, prev


def fill_planes(cur, prev, gpu_fb=None, oracle_fb=None):
    """Uploads the same values to a GPU FramebufferSet and an oracle set.
    Planes not given are zero on both sides.  A plane the GPU set stores as
    half is rounded to half for the oracle too, so both start equal."""
    def stored(p, a):
        if gpu_fb is None:
            return a
        return np.asarray(a).astype(gpu_fb.dtype(p, gpu_fb.plane_ptr(p)[1]))

    for p, a in cur.items():
        a = stored(p, a)
        if gpu_fb is not None:
            gpu_fb.upload(p, a)
        if oracle_fb is not None:
            oracle_fb.set(p, a)
    for p, a in prev.items():
        a = stored(p, a)
        if gpu_fb is not None:
            gpu_fb.upload(p, a, prev=True)
        if oracle_fb is not None:
            oracle_fb.set(p, a, prev=True)


def rel_err(got: np.ndarray, ref: np.ndarray) -> float:
    """max |got - ref| / (|ref| + 1e-6 max|ref|) over the plane."""
    got = got.astype(np.float64)
    ref = ref.astype(np.float64)
    scale = np.abs(ref).max() if ref.size else 0.0
    return float((np.abs(got - ref) / (np.abs(ref) + 1e-6 * scale + 1e-30)).max(initial=0.0))
