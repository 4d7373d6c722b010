"""The oracle's hashing, ReSTIR DI reuse, gradient samples and ReSTIR GI
reuse against what the mathematics fixes, on the CPU.

- murmurHash33 (Random.h:277-284) is MurmurHash2 (Appleby's published
  32-bit routine, constants m = 0x5bd1e995, r = 24) of the 12-byte key
  (x, y, z) run three times, lane i with the seed that makes h = seed ^ 12
  start at the lane's constant (1190494759, 2147483647, 3559788179).
- ReSTIR DI with reuse (RtRaygenDirect.rgen, RaygenCommon.h:503-599) on a
  plane under a sun only: every candidate has the same target function up
  to the sun disk's jitter, and the reservoir weight sum over M / target of
  the selected sample has expectation 1 for RIS with a constant integrand,
  so the mean of the direct light equals colour cos(theta) / pi at 8/1/8 and
  32/1/1 in every frame.
- Gradient samples (RtGradients.rgen): a still scene re-shades each sample
  as it was (the only difference is the 16+16-bit normal codec), so every
  antilag gradient is ~0; with the only light removed, the diffuse and
  specular gradients are 1 (getAntilagAlpha of g = -prev, n = prev,
  ShaderCommonGLSLFunc.h:383-390), times the sensitivity (0.5), mixed 1:1
  with the near-zero history, and a constant stays constant through the
  normalised gradient a-trous: 0.25 everywhere.
- ReSTIR GI (RtRaygenIndirect.inl:257-434) under a uniform sky with no
  lights: every sample's radiance is the (E5B9G9R9) sky colour times a
  product of grey albedos, so the indirect light scales exactly with the
  sky and keeps its colour channel for channel.
"""
import math
import struct

import numpy as np
import pytest

from raytracedgl1_amd import ffi, scenes
from oracle_lib import OracleFramebuffers, OraclePathTracer
from test_oracle_pathtrace import params_for, plane_scene, W, H

P = ffi.PLANE
M32 = 0xFFFFFFFF


def murmurhash2(data: bytes, seed: int) -> int:
    m, r = 0x5BD1E995, 24
    h = (seed ^ len(data)) & M32
    n = len(data) // 4
    for i in range(n):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (k * m) & M32
        k ^= k >> r
        k = (k * m) & M32
        h = (h * m) & M32
        h ^= k
    tail = data[4 * n:]
    if len(tail) == 3:
        h ^= tail[2] << 16
    if len(tail) >= 2:
        h ^= tail[1] << 8
    if len(tail) >= 1:
        h ^= tail[0]
        h = (h * m) & M32
    h ^= h >> 13
    h = (h * m) & M32
    h ^= h >> 15
    return h


def test_murmur_hash33_is_murmurhash2(oracle):
    lanes = (1190494759, 2147483647, 3559788179)
    rng = np.random.default_rng(7)
    keys = [(0, 0, 0), (M32, M32, M32), (1, 2, 3), (1920, 1080, 59)] + [tuple(int(v) for v in rng.integers(0, 2 ** 32, 3)) for _ in range(200)]
    for x, y, z in keys:
        got = oracle.murmur33(x, y, z)
        want = [murmurhash2(struct.pack("<3I", x, y, z), lane ^ 12) for lane in lanes]
        assert list(got) == want, (x, y, z)


def _square(cx, cz, half):
    a = [(cx - half, 0, cz - half), (cx - half, 0, cz + half), (cx + half, 0, cz + half), (cx + half, 0, cz - half)]
    tris = np.array([np.concatenate([a[0], a[1], a[2]]), np.concatenate([a[0], a[2], a[3]])], np.float32)
    return scenes.Scene(tris, np.tile(np.array([0, 1, 0], np.float32), (2, 3)), np.full(2, scenes.pack_rgba8(0.8, 0.8, 0.8), np.uint32),
                        [("square", 0, 2)], {})


SUN_D = np.array([0.3, -1.0, 0.2])
LIGHT_INDEX_NONE = 0x7FFF
# rnd16 is (x & 0xFFFF) / 65535, so it is exactly 1 once in 65536 draws, and
# the RIS test rnd * weightSum < weight (Reservoir.h:66-97) then keeps the
# reservoir's previous choice even for its only candidate: an initial
# reservoir can hold the sun's weight with no light selected, and a pixel
# that reuses it ends with an empty reservoir (RaygenCommon.h:595-598) and
# no direct light.  Such pixels are rare.
MAX_EMPTY_FRACTION = 1e-3
SUN = {"color": (1.0, 1.0, 1.0), "intensity": 3.0, "direction": tuple(SUN_D), "angularDiameterDegrees": 0.5}
CAM = scenes.Camera(np.array([0.0, 5.0, 0.0]), np.array([1.5, 0.0, 3.5]))


@pytest.mark.parametrize("counts", [(8, 1, 8), (32, 1, 1)])
def test_direct_reuse_keeps_the_mean(oracle, counts):
    cand, temporal, spatial = counts
    s = _square(1.5, 3.5, 4.0)
    scene = scenes.Scene(s.tris, s.normals, s.colors, s.objects, {"sun": SUN})
    fb = OracleFramebuffers(oracle, W, H)
    pt = OraclePathTracer(oracle, scene, W, H)
    cos_t = -SUN_D[1] / np.linalg.norm(SUN_D)
    expect = 3.0 * cos_t / math.pi
    r = math.sin(max(0.01, math.radians(0.25)))
    sin_t = math.sqrt(1.0 - cos_t * cos_t)
    jitter = (r * sin_t + 0.5 * r * r) * 3.0 / math.pi
    for f in range(4):
        if f:
            fb.swap()
        p = params_for(CAM, f, restirInitialCandidates=cand, restirTemporalSamples=temporal, restirSpatialSamples=spatial,
                       cellWorldSize=2.0)
        pt.frame(p, fb, threads=8)
        o = fb.cur
        ok = o[P["IS_SKY"]][..., 0] == 0
        pos = o[P["SURFACE_POSITION"]][..., :3][ok]
        center = np.array(p.cameraPosition[:3]) + 1.0
        # the square lies inside the light grid, whose cells hold no light
        assert np.all(np.abs(pos - center) < 16.0)
        d = o[P["UNFILTERED_DIRECT"]][..., 0][ok]
        m = o[P["RESERVOIRS"]][..., 0][ok].astype(np.uint64) >> 16
        print(f"reuse {counts} frame {f}: {ok.sum()} pixels, mean / expected {d.mean() / expect:.6f}, max M {m.max()},"
              f" max |d - e| {np.abs(d - expect).max() / jitter:.2f} jitter bounds", flush=True)
        if f:
            assert m.max() > 1  # reuse happened
        empty = (o[P["RESERVOIRS"]][..., 0][ok].astype(np.uint64) & 0xFFFF) == LIGHT_INDEX_NONE
        assert abs(d.mean() / expect - 1.0) < 0.2 * jitter / expect
        assert empty.sum() <= MAX_EMPTY_FRACTION * ok.sum()
        assert np.abs(d[~empty] - expect).max() <= 4.0 * jitter
        assert (d[empty] == 0.0).all()


def test_gradients_still_scene_and_removed_light(oracle):
    s = _square(0.0, 0.0, 50.0)
    lit = scenes.Scene(s.tris, s.normals, s.colors, s.objects, {"sun": SUN})
    dark = scenes.Scene(s.tris, s.normals, s.colors, s.objects, {})
    cam = scenes.Camera(np.array([0.0, 6.0, 0.0]), np.array([0.5, 0.0, 1.5]), fovy=math.radians(50.0))
    fb = OracleFramebuffers(oracle, W, H)
    pt = OraclePathTracer(oracle, lit, W, H)
    for f in range(5):
        if f:
            fb.swap()
        if f == 4:
            pt.scene_lights(dark)
        p = params_for(cam, f)
        pt.frame(p, fb, threads=8)
        fb.run("", p)
        o = fb.cur
        assert (o[P["IS_SKY"]] == 0).all()
        g = o[P["DIS_GRADIENT_HISTORY"]][..., :3]
        print(f"gradients frame {f}: diffuse {g[..., 0].min():.6f}..{g[..., 0].max():.6f}, specular {g[..., 2].min():.6f}..{g[..., 2].max():.6f}",
              flush=True)
        if 1 <= f <= 3:
            assert np.abs(g).max() < 1e-3
        if f == 4:
            want = 0.5 * 1.0 * p.gradientMultDiffuse
            assert np.abs(g[..., 0] - want).max() < 1e-3
            want = 0.5 * 1.0 * p.gradientMultSpecular
            assert np.abs(g[..., 2] - want).max() < 1e-3


def test_gi_reuse_is_linear_and_keeps_the_sky_colour(oracle):
    out = {}
    for mult in (1.0, 2.0):
        fb = OracleFramebuffers(oracle, W, H)
        pt = OraclePathTracer(oracle, plane_scene({}), W, H)
        for f in range(3):
            if f:
                fb.swap()
            p = params_for(CAM, f, skyColorMultiplier=mult)
            p.skyColorDefault[:3] = [0.2, 0.5, 0.9]
            pt.frame(p, fb, threads=8)
        out[mult] = fb.cur[P["UNFILTERED_INDIR"]][..., :3].copy()
        sky = fb.cur[P["IS_SKY"]][..., 0] != 0
    ind = out[1.0][~sky]
    assert (ind > 0).mean() > 0.5
    # exact linearity: the E5B9G9R9 radiance, the RIS choices (rnd * wsum < w)
    # and the weights are all invariant under a power-of-two scale
    assert np.array_equal(out[2.0], 2.0 * out[1.0])
    # the sky colour, channel for channel, as its E5B9G9R9 code holds it
    e5 = np.array(oracle.decode_e5(oracle.encode_e5((0.2, 0.5, 0.9))))
    nz = ind[:, 0] > 0
    ratio = ind[nz] / e5
    assert np.allclose(ratio[:, 1], ratio[:, 0], rtol=1e-9) and np.allclose(ratio[:, 2], ratio[:, 0], rtol=1e-9)
