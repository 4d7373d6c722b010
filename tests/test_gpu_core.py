"""GPU parity of the LBVH build, traversal and luminance histogram against the
CPU oracle, called through the C ABI (librtgl1_amd.so).

Integer and index outputs (Morton codes, sort order, topology, node child
references, primitive ids, histogram bins) must be bit-exact; FP32 outputs of
the same IEEE operation sequence (boxes, hit t/u/v, average luminance) must be
equal as floats.
"""
import numpy as np
import pytest

from raytracedgl1_amd import ffi, scenes
from raytracedgl1_amd.ffi import (HIT_DTYPE, RAY_DTYPE, RG_EXT_MISS, RG_EXT_TRACE_ANY_HIT, RG_EXT_TRACE_COLLECT_STATS,
                                  RG_EXT_TRACE_CULL_FRONT_FACE, Bvh, DeviceArray)

pytestmark = pytest.mark.gpu


def _random_tris(n, seed, spread=10.0, size=1.0):
    r = np.random.default_rng(seed)
    c = r.uniform(-spread, spread, (n, 1, 3))
    return (c + r.uniform(-size, size, (n, 3, 3))).reshape(n, 9).astype(np.float32)


def _gpu_bvh(tris, leaf=3):
    d = DeviceArray.from_numpy(tris)
    b = Bvh()
    b.build(d, len(tris), leaf)
    return b, d


def _assert_same_bvh(g, o):
    assert g["n"] == o["n"]
    assert np.array_equal(g["codes"], o["codes"])
    assert np.array_equal(g["scene_min"], o["scene_min"])
    assert np.array_equal(g["scene_max"], o["scene_max"])
    assert np.array_equal(g["sorted_idx"], o["sorted_idx"])
    assert np.array_equal(g["children"], o["children"])
    assert np.array_equal(g["bounds"], o["bounds"])  # float equality (+0 == -0)
    assert g["node_count"] == o["node_count"]
    for f in ("lo_x", "hi_x", "lo_y", "hi_y", "lo_z", "hi_z"):
        assert np.array_equal(g["nodes"][f], o["nodes"][f]), f
    assert np.array_equal(g["nodes"]["child"], o["nodes"]["child"])


@pytest.mark.parametrize("n,leaf", [(1, 3), (2, 1), (3, 3), (7, 1), (64, 3), (1000, 4), (4097, 2), (70000, 3)])
def test_lbvh_bit_exact_random(gpu, oracle, n, leaf):
    print(f"lbvh random n={n} leaf={leaf}", flush=True)
    tris = _random_tris(n, seed=n)
    b, _ = _gpu_bvh(tris, leaf)
    _assert_same_bvh(b.download(), oracle.bvh(tris, leaf).download())


def test_lbvh_bit_exact_duplicates(gpu, oracle):
    base = _random_tris(1, seed=5)
    tris = np.repeat(base, 5000, axis=0)
    tris[:, 0] += np.linspace(-1e-3, 1e-3, 5000, dtype=np.float32)
    b, _ = _gpu_bvh(tris, 3)
    _assert_same_bvh(b.download(), oracle.bvh(tris, 3).download())


def test_lbvh_bit_exact_terrain_boxes(gpu, oracle):
    s = scenes.terrain_boxes_scene()
    b, _ = _gpu_bvh(s.tris, 3)
    _assert_same_bvh(b.download(), oracle.bvh(s.tris, 3).download())


def test_lbvh_bit_exact_terrain_spheres_1m(gpu, oracle):
    print("lbvh terrain-spheres 1M tris", flush=True)
    s = scenes.terrain_spheres_scene()
    b, _ = _gpu_bvh(s.tris, 3)
    _assert_same_bvh(b.download(), oracle.bvh(s.tris, 3).download())


def test_lbvh_rebuild_is_deterministic_and_order_invariant(gpu):
    # Tests/RtglTest.cpp:510-511 shuffles primitives because upload order must
    # not change the image: the set of leaf triangles and the tree shape are
    # the same after a permutation that keeps ids stable by construction.
    s = scenes.terrain_boxes_scene()
    b, d = _gpu_bvh(s.tris, 3)
    first = b.download()
    b.build(d, s.n, 3)
    again = b.download()
    _assert_same_bvh(first, again)


def _trace_gpu(b, rays, flags):
    dr = DeviceArray.from_numpy(rays)
    dh = DeviceArray(len(rays), HIT_DTYPE)
    b.trace(dr, len(rays), dh, flags)
    out = dh.download()
    assert b.trace_errors() == 0, "traversal reported a dropped stack entry or a step-limit stop"
    return out


def _assert_same_hits(g, o, any_hit=False):
    assert np.array_equal(g["prim"] == RG_EXT_MISS, o["prim"] == RG_EXT_MISS)
    if any_hit:
        return
    assert np.array_equal(g["prim"], o["prim"])
    for f in ("t", "u", "v"):
        assert np.array_equal(g[f], o[f]), f


@pytest.mark.parametrize("flags", [0, RG_EXT_TRACE_CULL_FRONT_FACE, RG_EXT_TRACE_ANY_HIT])
def test_trace_bit_exact_terrain_boxes_primary(gpu, oracle, flags):
    print(f"trace terrain-boxes flags={flags}", flush=True)
    s = scenes.terrain_boxes_scene()
    rays = scenes.primary_rays(scenes.terrain_boxes_camera(), 640, 360)
    b, _ = _gpu_bvh(s.tris, 3)
    _assert_same_hits(_trace_gpu(b, rays, flags), oracle.bvh(s.tris, 3).trace(rays, flags),
                      any_hit=bool(flags & RG_EXT_TRACE_ANY_HIT))


@pytest.mark.parametrize("flags", [0, RG_EXT_TRACE_ANY_HIT])
def test_trace_visit_counts_match_oracle(gpu, oracle, flags):
    # Node visits and triangle tests are the traversal's algorithmic work
    # (they price its memory traffic); the GPU must visit exactly what the
    # oracle's traversal of the same tree visits, for primary and random rays.
    print(f"trace stats flags={flags}", flush=True)
    s = scenes.terrain_boxes_scene()
    r = np.random.default_rng(5)
    rnd = np.zeros(20000, RAY_DTYPE)
    rnd["origin"] = r.uniform(-30, 30, (20000, 3)) + np.array([0, 4, 0])
    d = r.normal(size=(20000, 3))
    rnd["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rnd["tmax"] = 1000.0
    rays = np.concatenate([scenes.primary_rays(scenes.terrain_boxes_camera(), 640, 360), rnd])
    b, _ = _gpu_bvh(s.tris, 3)
    g = _trace_gpu(b, rays, flags | RG_EXT_TRACE_COLLECT_STATS)
    nodes, tris = b.trace_stats()
    o, onodes, otris = oracle.bvh(s.tris, 3).trace(rays, flags, stats=True)
    _assert_same_hits(g, o, any_hit=bool(flags & RG_EXT_TRACE_ANY_HIT))
    assert (nodes, tris) == (onodes, otris)
    assert nodes >= len(rays)
    # without the flag the counters are not touched
    _trace_gpu(b, rays[:100], flags)
    assert b.trace_stats() == (nodes, tris)


def test_trace_matches_bruteforce_fp64(gpu, oracle):
    s = scenes.terrain_boxes_scene()
    rays = scenes.primary_rays(scenes.terrain_boxes_camera(), 640, 360)[::29]
    b, _ = _gpu_bvh(s.tris, 3)
    g = _trace_gpu(b, rays, 0)
    t, u, v, p = oracle.bruteforce(s.tris, rays, 0)
    miss = p == RG_EXT_MISS
    assert np.array_equal(g["prim"] == RG_EXT_MISS, miss)
    h = ~miss
    assert (np.abs(g["t"][h] - t[h]) / t[h]).max() < 1e-5
    diff = h & (g["prim"] != p)
    assert np.all(np.abs(g["t"][diff] - t[diff]) <= 1e-6 * t[diff] + 1e-7)


def test_trace_random_rays_and_empty(gpu, oracle):
    s = scenes.terrain_boxes_scene()
    r = np.random.default_rng(3)
    rays = np.zeros(50000, RAY_DTYPE)
    rays["origin"] = r.uniform(-40, 40, (50000, 3))
    d = r.normal(size=(50000, 3))
    rays["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmin"] = r.uniform(0, 1, 50000)
    rays["tmax"] = r.uniform(1, 100, 50000)
    b, _ = _gpu_bvh(s.tris, 3)
    _assert_same_hits(_trace_gpu(b, rays, 0), oracle.bvh(s.tris, 3).trace(rays, 0))
    # zero rays is a no-op
    dh = DeviceArray(1, HIT_DTYPE)
    dr = DeviceArray(1, RAY_DTYPE)
    b.trace(dr, 0, dh, 0)


def test_trace_terrain_spheres_1080p(gpu, oracle):
    print("trace terrain-spheres 1080p", flush=True)
    s = scenes.terrain_spheres_scene()
    rays = scenes.primary_rays(scenes.orbit_camera(0), 1920, 1080)
    b, _ = _gpu_bvh(s.tris, 3)
    g = _trace_gpu(b, rays, 0)
    ob = oracle.bvh(s.tris, 3)
    sub = np.arange(0, len(rays), 7)
    _assert_same_hits(g[sub], ob.trace(rays[sub], 0))
    # size-independent property over the full image: hit points lie on the hit triangle
    h = g["prim"] != RG_EXT_MISS
    assert h.mean() > 0.5
    tri = s.tris.reshape(-1, 3, 3)[g["prim"][h]].astype(np.float64)
    u, v = g["u"][h, None].astype(np.float64), g["v"][h, None].astype(np.float64)
    pb = (1 - u - v) * tri[:, 0] + u * tri[:, 1] + v * tri[:, 2]
    pr = rays["origin"][h] + rays["direction"][h] * g["t"][h, None]
    assert np.abs(pb - pr).max() < 1e-2


def test_histogram_bit_exact(gpu, oracle):
    r = np.random.default_rng(11)
    w, h = 1920, 1080
    lum = np.exp(r.uniform(-12, 12, w * h)).astype(np.float32)
    lum[::97] = 0.0
    lum[::101] = np.nan
    lum[::103] = np.inf
    lum[1::89] = 0.001
    min_log, max_log = np.log2(np.exp2(-4.0) * 12.5 / 100), np.log2(np.exp2(12.0) * 12.5 / 100)
    d = DeviceArray.from_numpy(lum)
    dh = DeviceArray(256, np.uint32, zero=True)
    ffi.call("rgExtLuminanceHistogram", d.ptr, w, h, float(min_log), float(max_log), dh.ptr, None)
    g = dh.download()
    o = oracle.lum_histogram(lum, float(np.float32(min_log)), float(np.float32(max_log)))
    assert np.array_equal(g, o)
    assert g.sum() == w * h
    # average: equal to the oracle's FP32 evaluation, and the histogram is cleared
    davg = DeviceArray.from_numpy(np.array([0.5], np.float32))
    ffi.call("rgExtLuminanceAverage", dh.ptr, w * h, float(min_log), float(max_log), 0.016, 0, davg.ptr, None)
    ga = davg.download()[0]
    oa = oracle.lum_average(o, w * h, float(np.float32(min_log)), float(np.float32(max_log)), 0.016, False, 0.5)
    assert ga == np.float32(oa)
    assert np.all(dh.download() == 0)


def test_device_copy_kernel(gpu):
    a = np.random.default_rng(2).integers(0, 2**32, 4096 * 4 + 8, dtype=np.uint32)
    src = DeviceArray.from_numpy(a)
    dst = DeviceArray(a.shape, a.dtype, zero=True)
    ffi.call("rgExtCopyDeviceToDevice", dst.ptr, src.ptr, a.nbytes, None)
    assert np.array_equal(dst.download(), a)
    ffi.call("rgExtCopyDeviceToDevice", dst.ptr, src.ptr, 0, None)
    with pytest.raises(ffi.RgError):
        ffi.call("rgExtCopyDeviceToDevice", dst.ptr, src.ptr, 20, None)



def _decode_half_nodes(nodes):
    """Child boxes of FP16 nodes as the traversal decodes them: origin +
    float(offset), rounded to FP32."""
    o = nodes["origin"].astype(np.float32)
    out = {}
    for f, axis in (("lo_x", 0), ("hi_x", 0), ("lo_y", 1), ("hi_y", 1), ("lo_z", 2), ("hi_z", 2)):
        out[f] = o[:, axis:axis + 1] + nodes[f].view(np.float16).astype(np.float32)
    return out


@pytest.mark.parametrize("case", ["terrain_boxes", "far_soup"])
def test_fp16_nodes_contain_fp32_boxes_and_hit_the_same(gpu, case):
    """FP16 nodes (FP32 origin + half offsets rounded outward) have the FP32
    build's topology and boxes that contain its boxes, so closest-hit traversal
    finds the same hits bit for bit (the hit rule is order-independent)."""
    print(f"fp16 nodes {case}", flush=True)
    if case == "terrain_boxes":
        tris = scenes.terrain_boxes_scene().tris
        rays = np.concatenate([scenes.primary_rays(scenes.terrain_boxes_camera(), 640, 360)])
    else:
        # 200k small triangles 20,000 units from the origin: absolute FP16
        # coordinates there are 16 units apart
        tris = _random_tris(200_000, 9, spread=500.0, size=0.3) + np.float32(20000.0)
        r = np.random.default_rng(4)
        rays = np.zeros(100_000, RAY_DTYPE)
        rays["origin"] = r.uniform(-600, 600, (100_000, 3)) + 20000.0
        d = r.normal(size=(100_000, 3))
        rays["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
        rays["tmax"] = 2000.0
    d = DeviceArray.from_numpy(tris)
    b32, b16 = Bvh(), Bvh()
    b32.build(d, len(tris), 1, ffi.RG_EXT_BVH_NODE_FP32)
    b16.build(d, len(tris), 1, ffi.RG_EXT_BVH_NODE_FP16)
    g32, g16 = b32.download(), b16.download()
    assert b16.view().nodeStrideBytes == 144
    assert g16["node_count"] == g32["node_count"]
    assert np.array_equal(g16["nodes"]["child"], g32["nodes"]["child"])
    assert np.array_equal(g16["sorted_tris"], g32["sorted_tris"])
    dec = _decode_half_nodes(g16["nodes"])
    full = g32["nodes"]["child"] != 0xFFFFFFFF
    for f in ("lo_x", "lo_y", "lo_z"):
        assert (dec[f][full] <= g32["nodes"][f][full]).all(), f
    for f in ("hi_x", "hi_y", "hi_z"):
        assert (dec[f][full] >= g32["nodes"][f][full]).all(), f
    # outward rounding costs ~2^-11 of the node's extent, not of the coordinates
    ext32 = g32["nodes"]["hi_x"][full].astype(np.float64) - g32["nodes"]["lo_x"][full]
    ext16 = dec["hi_x"][full].astype(np.float64) - dec["lo_x"][full]
    grow = (ext16 - ext32) / np.maximum(ext32, 1e-6)
    print(f"fp16 box growth (x): median {np.median(grow):.2e}, p99 {np.quantile(grow, 0.99):.2e}", flush=True)
    assert np.median(grow) < 2e-2
    for flags in (0, RG_EXT_TRACE_ANY_HIT):
        h32 = _trace_gpu(b32, rays, flags | RG_EXT_TRACE_COLLECT_STATS)
        n32 = b32.trace_stats()
        h16 = _trace_gpu(b16, rays, flags | RG_EXT_TRACE_COLLECT_STATS)
        n16 = b16.trace_stats()
        print(f"flags {flags}: node visits fp32 {n32[0]} fp16 {n16[0]}, triangle tests {n32[1]} {n16[1]}", flush=True)
        assert b16.trace_errors() == 0
        _assert_same_hits(h16, h32, any_hit=bool(flags & RG_EXT_TRACE_ANY_HIT))
