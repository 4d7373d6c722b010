"""Pins the CPU oracle's codecs, RNG, math, LBVH and traversal against closed
forms, invariants and an independent FP64 brute-force intersector.

The reference ships no golden vectors for this path (its only test is an
interactive demo, Tests/RtglTest.cpp), so each check here is chosen so that
a plausible mistake (wrong shift, dropped term, swapped operand, wrong split)
fails it.
"""
import math

import numpy as np
import pytest

from raytracedgl1_amd import scenes
from raytracedgl1_amd.ffi import HIT_DTYPE, RAY_DTYPE, RG_EXT_MISS, RG_EXT_TRACE_ANY_HIT, RG_EXT_TRACE_CULL_FRONT_FACE

RNG = np.random.default_rng(1234)


# ----------------------------------------------------------------- math

def test_log2_exact_on_powers_of_two(oracle):
    for k in range(-140, 128):
        x = math.ldexp(1.0, k)
        if x < 1.5e-45:
            continue
        assert oracle.log2f(x) == float(k)


def test_log2_accuracy(oracle):
    xs = np.concatenate([np.exp(RNG.uniform(-80, 80, 4000)), RNG.uniform(0.5, 2.0, 2000)]).astype(np.float32)
    for x in xs:
        ref = math.log2(float(x))
        got = oracle.log2f(float(x))
        assert abs(got - ref) <= 4e-7 * max(1.0, abs(ref)), (x, got, ref)


def test_exp2_accuracy_and_exact_integers(oracle):
    for k in range(-125, 127):
        assert oracle.exp2f(float(k)) == math.ldexp(1.0, k)
    for x in RNG.uniform(-120, 120, 4000).astype(np.float32):
        ref = 2.0 ** float(x)
        assert abs(oracle.exp2f(float(x)) - ref) <= 6e-7 * ref
    assert oracle.exp2f(-200.0) == 0.0
    assert math.isinf(oracle.exp2f(200.0))


# ------------------------------------------------------------------ RNG

M32 = 0xFFFFFFFF


def _wellons_inverse(y: int) -> int:
    """Inverse of lowbias32 built from the inverse of each step: xorshift by s
    is undone by repeated xorshift, odd multipliers by their inverse mod 2^32."""
    def unxorshift(v, s):
        r = v
        for _ in range(32 // s + 1):
            r = v ^ (r >> s)
        return r & M32
    y = unxorshift(y, 16)
    y = (y * pow(0x846CA68B, -1, 1 << 32)) & M32
    y = unxorshift(y, 15)
    y = (y * pow(0x7FEB352D, -1, 1 << 32)) & M32
    y = unxorshift(y, 16)
    return y


def test_wellons_is_the_published_bijection(oracle):
    for x in [0, 1, 2, 0xDEADBEEF, M32] + RNG.integers(0, 1 << 32, 500).tolist():
        y = oracle.l.or_wellons_low_bias32(int(x))
        assert _wellons_inverse(y) == int(x)
    assert oracle.l.or_wellons_low_bias32(0) == 0  # every step maps 0 to 0


def test_wellons_avalanche(oracle):
    xs = RNG.integers(0, 1 << 32, 2000)
    flips = []
    for x in xs:
        y = oracle.l.or_wellons_low_bias32(int(x))
        b = int(RNG.integers(0, 32))
        y2 = oracle.l.or_wellons_low_bias32(int(x) ^ (1 << b))
        flips.append(bin(y ^ y2).count("1"))
    assert 15.0 < np.mean(flips) < 17.0


def test_random_seed_layout_and_uniformity(oracle):
    layers = np.zeros(128, np.int64)
    ox = np.zeros(128, np.int64)
    for y in range(64):
        for x in range(64):
            s = oracle.l.or_random_seed(x, y, 7)
            assert s < (1 << 21)
            layers[s >> 14] += 1
            ox[s & 127] += 1
    # chi-square with 127 dof: mean 127, sd ~16
    for hist in (layers, ox):
        e = hist.sum() / 128
        chi = ((hist - e) ** 2 / e).sum()
        assert chi < 127 + 6 * 16


def test_murmur_mixes_each_input(oracle):
    base = oracle.murmur33(10, 20, 30)
    for args in [(11, 20, 30), (10, 21, 30), (10, 20, 31)]:
        h = oracle.murmur33(*args)
        for a, b in zip(base, h):
            assert 8 < bin(a ^ b).count("1") < 24


# --------------------------------------------------------------- codecs

def test_normal_codec_closed_forms(oracle):
    assert oracle.encode_normal([0, 0, 1]) == 0
    assert oracle.encode_normal([0, 0, -1]) == 65535 << 16
    # phi = pi/2 -> j = round(32767.5) = 32768; theta = 0 -> k = 0
    assert oracle.encode_normal([1, 0, 0]) == 32768 << 16
    # theta = pi/2 -> k = 16384
    assert oracle.encode_normal([0, 1, 0]) == (32768 << 16) | 16384


def test_normal_codec_roundtrip(oracle):
    v = RNG.normal(size=(3000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for n in v:
        d = oracle.decode_normal(oracle.encode_normal(n))
        ang = math.acos(min(1.0, float(np.dot(n, d))))
        assert ang < 2.5 * math.pi / 65535


def test_e5b9g9r9_closed_forms(oracle):
    one = (16 << 27) | (256 << 18) | (256 << 9) | 256
    assert oracle.encode_e5([1, 1, 1]) == one
    assert np.all(oracle.decode_e5(one) == [1, 1, 1])
    assert oracle.encode_e5([0, 0, 0]) == 0
    # max representable (65408) saturates, negatives clamp to 0
    p = oracle.encode_e5([1e9, -5, 0])
    assert np.allclose(oracle.decode_e5(p), [65408, 0, 0])


def test_e5b9g9r9_roundtrip_error(oracle):
    for rgb in np.exp(RNG.uniform(-10, 10, (2000, 3))):
        d = oracle.decode_e5(oracle.encode_e5(rgb))
        m = rgb.max()
        assert np.all(np.abs(d - rgb) <= m * 2.0 ** -9 + 1e-30)


@pytest.mark.parametrize("w,h", [(8, 4), (640, 9), (34, 3)])
def test_checkerboard_is_a_bijection(oracle, w, h):
    seen = set()
    for y in range(h):
        for x in range(w):
            cx, cy = oracle.cb_from_regular(x, y, w)
            assert 0 <= cx < w and cy == y
            assert oracle.cb_to_regular(cx, cy, w) == (x, y)
            # odd pixels (x + y) land in the right half
            assert (cx >= w // 2) == ((x + y % 2) % 2 == 1)
            seen.add((cx, cy))
    assert len(seen) == w * h


# ----------------------------------------------------------------- LBVH

def _numpy_morton(tris: np.ndarray):
    """Independent vectorised restatement of the quantisation (for pinning)."""
    t = tris.astype(np.float64).reshape(-1, 3, 3)
    c = ((t[:, 0] + t[:, 1]) + t[:, 2]) / 3.0
    lo, hi = c.min(0), c.max(0)
    ext = hi - lo
    q = np.zeros_like(c, dtype=np.uint64)
    for k in range(3):
        if ext[k] > 0:
            f = np.floor(((c[:, k] - lo[k]) / ext[k]) * 2097152.0)
            q[:, k] = np.clip(f, 0, 2097151).astype(np.uint64)
    code = np.zeros(len(c), np.uint64)
    for bit in range(21):
        for k, sh in ((0, 2), (1, 1), (2, 0)):
            code |= ((q[:, k] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + sh)
    return code, lo, hi


def _random_tris(n, seed, spread=10.0, size=1.0):
    r = np.random.default_rng(seed)
    c = r.uniform(-spread, spread, (n, 1, 3))
    return (c + r.uniform(-size, size, (n, 3, 3))).reshape(n, 9).astype(np.float32)


def _delta(keys, n, i, j):
    if j < 0 or j >= n:
        return -1
    a, b = int(keys[i]), int(keys[j])
    if a == b:
        return 64 + (32 - (i ^ j).bit_length())
    return 64 - (a ^ b).bit_length()


def _check_bvh_invariants(tris, out, leaf):
    n = out["n"]
    code, lo, hi = _numpy_morton(tris)
    assert np.array_equal(out["codes"], code)
    assert np.array_equal(out["scene_min"], lo) and np.array_equal(out["scene_max"], hi)
    # sorted order is the stable sort of the codes
    assert np.array_equal(out["sorted_idx"], np.argsort(code, kind="stable").astype(np.uint32))
    keys = code[out["sorted_idx"]]
    ch = out["children"]
    b = out["bounds"]
    # leaf boxes are the triangle boxes
    t = tris.reshape(-1, 3, 3)[out["sorted_idx"]]
    assert np.array_equal(b[n - 1:, :3], t.min(1)) and np.array_equal(b[n - 1:, 3:], t.max(1))
    if n > 1:
        # ranges: every leaf reached once; split at the unique minimum of delta
        rng = {}

        def walk(node, a, z):
            rng[node] = (a, z)
            if node >= n - 1:
                assert a == z == node - (n - 1)
                return
            d = [_delta(keys, n, i, i + 1) for i in range(a, z)]
            s = a + int(np.argmin(d))
            assert d.count(min(d)) == 1
            l, r = ch[node]
            walk(int(l), a, s)
            walk(int(r), s + 1, z)
            # parent box is the union of the children boxes
            assert np.array_equal(b[node, :3], np.minimum(b[l, :3], b[r, :3]))
            assert np.array_equal(b[node, 3:], np.maximum(b[l, 3:], b[r, 3:]))

        import sys
        sys.setrecursionlimit(100000)
        walk(0, 0, n - 1)
        assert len(rng) == 2 * n - 1
    # 8-wide nodes: every triangle in exactly one leaf; child boxes contain the subtree
    nodes = out["nodes"]
    seen = np.zeros(n, np.int64)
    visited = np.zeros(len(nodes), np.int64)

    def visit(idx, box):
        visited[idx] += 1
        nd = nodes[idx]
        for k in range(8):
            c = int(nd["child"][k])
            if c == 0xFFFFFFFF:
                continue
            cb = np.array([nd["lo_x"][k], nd["lo_y"][k], nd["lo_z"][k], nd["hi_x"][k], nd["hi_y"][k], nd["hi_z"][k]])
            if box is not None:
                assert np.all(cb[:3] >= box[:3]) and np.all(cb[3:] <= box[3:])
            if c & 0x80000000:
                first, cnt = (c >> 3) & 0x0FFFFFFF, (c & 7) + 1
                assert cnt <= leaf
                tt = t[first:first + cnt]
                assert np.all(tt.min(1) >= cb[:3]) and np.all(tt.max(1) <= cb[3:])
                seen[first:first + cnt] += 1
            else:
                visit(c, cb)

    visit(0, None)
    assert np.all(seen == 1)
    assert np.all(visited == 1)  # the 8-wide nodes form one tree rooted at 0
    assert out["node_count"] <= max(1, n - 1)


@pytest.mark.parametrize("n,leaf", [(1, 3), (2, 1), (3, 3), (7, 1), (64, 3), (500, 4), (2000, 2)])
def test_lbvh_invariants_random(oracle, n, leaf):
    tris = _random_tris(n, seed=n)
    _check_bvh_invariants(tris, oracle.bvh(tris, leaf).download(), leaf)


def test_lbvh_invariants_duplicates(oracle):
    # identical centroids -> identical codes -> index tie-break in Karras
    base = _random_tris(1, seed=5)
    tris = np.repeat(base, 300, axis=0)
    tris[:, 0] += np.linspace(-1e-3, 1e-3, 300, dtype=np.float32)
    tris[:, 3] -= np.linspace(-1e-3, 1e-3, 300, dtype=np.float32)
    out = oracle.bvh(tris, 3).download()
    assert len(np.unique(out["codes"])) < 300
    _check_bvh_invariants(tris, out, 3)


def test_lbvh_invariants_terrain_boxes(oracle):
    s = scenes.terrain_boxes_scene()
    _check_bvh_invariants(s.tris, oracle.bvh(s.tris, 3).download(), 3)


# ------------------------------------------------------------ traversal

def _compare_with_bruteforce(hits, t_bf, p_bf, tol_t=1e-5, tie=1e-6):
    miss = p_bf == RG_EXT_MISS
    assert np.array_equal(hits["prim"] == RG_EXT_MISS, miss)
    h = ~miss
    rel = np.abs(hits["t"][h] - t_bf[h]) / np.abs(t_bf[h])
    assert rel.max(initial=0) < tol_t
    diff = h & (hits["prim"] != p_bf)
    # a different primitive is only allowed where two hits tie in t
    if diff.any():
        assert np.all(np.abs(hits["t"][diff] - t_bf[diff]) <= tie * np.abs(t_bf[diff]) + 1e-7)
    return int(diff.sum())


@pytest.mark.parametrize("n", [1, 2, 5, 64])
def test_traversal_matches_bruteforce_small(oracle, n):
    tris = _random_tris(n, seed=100 + n, spread=3.0)
    r = np.random.default_rng(n)
    rays = np.zeros(4000, RAY_DTYPE)
    rays["origin"] = r.uniform(-8, 8, (4000, 3))
    tgt = r.uniform(-3, 3, (4000, 3))
    d = tgt - rays["origin"]
    rays["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmin"], rays["tmax"] = 0.0, 100.0
    bvh = oracle.bvh(tris, 1 if n < 8 else 3)
    for flags in (0, RG_EXT_TRACE_CULL_FRONT_FACE):
        hits = bvh.trace(rays, flags)
        t, u, v, p = oracle.bruteforce(tris, rays, flags)
        assert _compare_with_bruteforce(hits, t, p) <= 2


def test_traversal_matches_bruteforce_terrain_boxes(oracle):
    s = scenes.terrain_boxes_scene()
    rays = scenes.primary_rays(scenes.terrain_boxes_camera(), 640, 360)
    sub = rays[:: 23]  # 10k rays, spread over the image
    bvh = oracle.bvh(s.tris, 3)
    hits = bvh.trace(sub, 0)
    t, u, v, p = oracle.bruteforce(s.tris, sub, 0)
    ndiff = _compare_with_bruteforce(hits, t, p)
    assert ndiff <= len(sub) * 1e-3
    # barycentrics agree where the primitive agrees
    same = (hits["prim"] == p) & (p != RG_EXT_MISS)
    assert np.abs(hits["u"][same] - u[same]).max() < 1e-3
    assert np.abs(hits["v"][same] - v[same]).max() < 1e-3
    assert (p != RG_EXT_MISS).mean() > 0.3


def test_shadow_rays_any_hit(oracle):
    s = scenes.terrain_boxes_scene()
    r = np.random.default_rng(7)
    rays = np.zeros(5000, RAY_DTYPE)
    rays["origin"] = np.stack([r.uniform(-30, 30, 5000), r.uniform(0.5, 6, 5000), r.uniform(-30, 30, 5000)], 1)
    d = np.array([1.0, 8.0, 1.0]) / np.linalg.norm([1.0, 8.0, 1.0])
    rays["direction"] = d
    rays["tmin"], rays["tmax"] = 0.001, 50.0
    bvh = oracle.bvh(s.tris, 3)
    anyhit = bvh.trace(rays, RG_EXT_TRACE_ANY_HIT)
    closest = bvh.trace(rays, 0)
    assert np.array_equal(anyhit["prim"] == RG_EXT_MISS, closest["prim"] == RG_EXT_MISS)


def test_watertight_shared_edges_and_vertices(oracle):
    # rays from outside through vertices and edge midpoints of a closed
    # icosphere must hit it: no cracks between neighbouring triangles.  Only
    # points that face the origin are used (a grazing ray at the silhouette
    # may legitimately miss after its direction is rounded to FP32).
    unit = scenes.icosphere(2)
    tris = (unit * 2.0).reshape(-1, 9).astype(np.float32)
    v = tris.reshape(-1, 3, 3).astype(np.float64)
    pts = np.concatenate([v.reshape(-1, 3), ((v + np.roll(v, 1, axis=1)) / 2).reshape(-1, 3)])
    org = np.array([0.3, 7.0, -5.0])
    d = pts - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    facing = (-(d * pts).sum(1) / np.linalg.norm(pts, axis=1)) > 0.2
    pts, d = pts[facing], d[facing]
    rays = np.zeros(len(pts), RAY_DTYPE)
    rays["origin"] = org
    rays["direction"] = d
    rays["tmin"], rays["tmax"] = 0.0, 100.0
    assert len(rays) > 200
    hits = oracle.bvh(tris, 3).trace(rays, 0)
    assert np.all(hits["prim"] != RG_EXT_MISS)


def test_cull_front_face_convention(oracle):
    # counter-clockwise seen from the origin (normal towards the origin) = front
    ccw = np.array([[-1, -1, 5, 1, -1, 5, 0, 1, 5]], np.float32)  # n = +z... seen from z<5 looking +z
    n = np.cross(ccw[0, 3:6] - ccw[0, :3], ccw[0, 6:9] - ccw[0, :3])
    ray = np.zeros(1, RAY_DTYPE)
    ray["origin"] = (0, 0, 0)
    ray["direction"] = (0, 0, 1)
    ray["tmax"] = 100
    # n points along +z, the ray travels along +z: dot(n, d) > 0, so this
    # winding is clockwise from the origin (a back face): not culled
    assert n[2] > 0
    b = oracle.bvh(ccw, 1)
    assert b.trace(ray, RG_EXT_TRACE_CULL_FRONT_FACE)["prim"][0] == 0
    cw = ccw[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]].copy()  # swapped: front face, culled
    b2 = oracle.bvh(cw, 1)
    assert b2.trace(ray, RG_EXT_TRACE_CULL_FRONT_FACE)["prim"][0] == RG_EXT_MISS
    assert b2.trace(ray, 0)["prim"][0] == 0
    _, _, _, p = oracle.bruteforce(cw, ray, RG_EXT_TRACE_CULL_FRONT_FACE)
    assert p[0] == RG_EXT_MISS


def test_traversal_barycentrics_reconstruct_hit_point(oracle):
    s = scenes.terrain_boxes_scene()
    rays = scenes.primary_rays(scenes.terrain_boxes_camera(), 160, 90)
    hits = oracle.bvh(s.tris, 3).trace(rays, 0)
    h = hits["prim"] != RG_EXT_MISS
    t = s.tris.reshape(-1, 3, 3)[hits["prim"][h]].astype(np.float64)
    u, v = hits["u"][h, None], hits["v"][h, None]
    p_bary = (1 - u - v) * t[:, 0] + u * t[:, 1] + v * t[:, 2]
    p_ray = rays["origin"][h] + rays["direction"][h] * hits["t"][h, None]
    assert np.abs(p_bary - p_ray).max() < 1e-3


# ------------------------------------------------------- eye adaptation

def test_histogram_closed_form_bins(oracle):
    # minLog = -8, maxLog = 8: lum = 2^k -> x = (k + 8) / 16 -> bin = uint(x*254 + 1)
    lum = np.array([2.0 ** k for k in range(-9, 10)] + [0.0, 0.0009, np.nan, np.inf, 1e-3], np.float32)
    h = oracle.lum_histogram(lum, -8.0, 8.0)
    expect = np.zeros(256, np.uint32)
    for k in range(-9, 10):
        x = min(max((k + 8) / 16.0, 0.0), 1.0)
        expect[int(np.float32(x) * np.float32(254.0) + np.float32(1.0))] += 1
    expect[0] += 3      # 0, 0.0009, NaN
    expect[255] += 1    # +inf
    # 1e-3 >= LUMINANCE_EPS: log2(1e-3) = -9.97 clamps to bin 1
    expect[1] += 1
    assert np.array_equal(h, expect)
    assert h.sum() == len(lum)


def test_average_luminance_one_bin_closed_form(oracle):
    hist = np.zeros(256, np.uint32)
    hist[129] = 1000
    min_log, max_log = -8.0, 8.0
    # weighted average bin = 129 -> avg = 2^((129-1)/254 * 16 - 8)
    expect = 2.0 ** ((129 - 1) / 254.0 * 16.0 - 8.0)
    got = oracle.lum_average(hist, 1000, min_log, max_log, 1000.0, False, 0.0)
    assert abs(got - expect) <= 2e-6 * expect
    # adaptation: t = 1 - exp(-dt * 1.1)
    got2 = oracle.lum_average(hist, 1000, min_log, max_log, 0.1, False, 1.0)
    t = 1 - math.exp(-0.1 * 1.1)
    assert abs(got2 - (1.0 + (expect - 1.0) * t)) < 1e-6
    # stop adaptation keeps the last value; black pixels are excluded
    assert oracle.lum_average(hist, 1000, min_log, max_log, 0.1, True, 0.25) == 0.25
    hist2 = hist.copy()
    hist2[0] = 500
    got3 = oracle.lum_average(hist2, 1500, min_log, max_log, 1000.0, False, 0.0)
    assert abs(got3 - expect) <= 2e-6 * expect
