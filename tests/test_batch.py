"""Batched scenes on the GPU: one BVH object over several scenes built by a
single pass sequence, and one persistent traversal launch over all of their
rays.

Closest-hit traversal is order-independent (a hit replaces the best one if
nearer, or equally near with a smaller primitive id), so every scene's hits
in the batched launch must equal, bit for bit, the hits of that scene's own
single-scene BVH (which the oracle tests pin), with primitive ids local to
the scene.  Each scene's subtree must reference only its own triangles.
The autotuner's choice must not change any result.
"""
import numpy as np
import pytest

from raytracedgl1_amd import ffi, scenes
from raytracedgl1_amd.ffi import HIT_DTYPE, RAY_DTYPE, RG_EXT_TRACE_ANY_HIT, Bvh, DeviceArray

pytestmark = pytest.mark.gpu


def _soup(n, seed, spread, size, center):
    r = np.random.default_rng(seed)
    c = r.uniform(-spread, spread, (n, 1, 3)) + np.asarray(center, np.float64)
    return (c + r.uniform(-size, size, (n, 3, 3))).reshape(n, 9).astype(np.float32)


def _rays(n, seed, lo, hi):
    r = np.random.default_rng(seed)
    out = np.zeros(n, RAY_DTYPE)
    out["origin"] = r.uniform(lo, hi, (n, 3))
    d = r.normal(size=(n, 3))
    out["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    out["tmax"] = 1000.0
    return out


def _scenes():
    s = [
        (_soup(2, 1, 1.0, 1.0, (0, 0, 0)), (-3, 3)),               # the smallest scene
        (_soup(10_000, 2, 20.0, 0.5, (100, 0, 0)), (80, 120)),
        (scenes.terrain_boxes_scene().tris, (-30, 30)),
        (_soup(150_000, 3, 300.0, 0.4, (0, 5000, 0)), (4700, 5300)),  # far from the origin
        (_soup(3, 4, 1.0, 1.0, (0, 0, 0)), (-3, 3)),
        (_soup(40_000, 5, 50.0, 2.0, (-500, 0, 0)), (-550, -450)),
    ]
    return s


def _trace(b, rays, flags, batch_offsets=None):
    dr = DeviceArray.from_numpy(rays)
    dh = DeviceArray(len(rays), HIT_DTYPE)
    if batch_offsets is None:
        b.trace(dr, len(rays), dh, flags)
    else:
        b.trace_batch(dr, batch_offsets, dh, flags)
    out = dh.download()
    assert b.trace_errors() == 0
    return out


def _subtree_tris(nodes, root, leaf_bit=0x80000000):
    seen, stack = [], [root]
    while stack:
        nd = stack.pop()
        for c in nodes["child"][nd]:
            c = int(c)
            if c == 0xFFFFFFFF:
                continue
            if c & leaf_bit:
                first, cnt = (c >> 3) & 0x0FFFFFFF, (c & 7) + 1
                seen.extend(range(first, first + cnt))
            else:
                stack.append(c)
    return np.array(sorted(seen))


@pytest.mark.parametrize("fmt", [ffi.RG_EXT_BVH_NODE_FP32, ffi.RG_EXT_BVH_NODE_FP16])
def test_batch_equals_single_scene_builds(gpu, fmt):
    print(f"batched build + trace, node format {fmt}", flush=True)
    sc = _scenes()
    tris = np.concatenate([t for t, _ in sc])
    offs = np.concatenate([[0], np.cumsum([len(t) for t, _ in sc])]).astype(np.uint32)
    rays = [_rays(5000 + 1000 * k, 10 + k, lo, hi) for k, (_, (lo, hi)) in enumerate(sc)]
    roffs = np.concatenate([[0], np.cumsum([len(r) for r in rays])]).astype(np.uint32)
    allrays = np.concatenate(rays)
    d = DeviceArray.from_numpy(tris)
    bb = Bvh()
    bb.build_batch(d, offs, 1, fmt)
    roots = bb.scene_roots()
    assert len(roots) == len(sc) and len(set(roots.tolist())) == len(sc)
    g = bb.download()
    for s in range(len(sc)):
        t = _subtree_tris(g["nodes"], int(roots[s]))
        assert np.array_equal(t, np.arange(offs[s], offs[s + 1])), s
    for flags in (0, RG_EXT_TRACE_ANY_HIT):
        hb = _trace(bb, allrays, flags, roffs)
        for s, (t, _) in enumerate(sc):
            b1 = Bvh()
            b1.build(DeviceArray.from_numpy(t), len(t), 1, fmt)
            h1 = _trace(b1, rays[s], flags)
            part = hb[roffs[s]:roffs[s + 1]]
            assert np.array_equal(part["prim"] == ffi.RG_EXT_MISS, h1["prim"] == ffi.RG_EXT_MISS), (s, flags)
            if flags == 0:
                for f in ("prim", "t", "u", "v"):
                    assert np.array_equal(part[f], h1[f]), (s, f)
            assert (part["prim"][part["prim"] != ffi.RG_EXT_MISS] < len(t)).all()


def test_autotune_keeps_results(gpu):
    print("trace autotune on a batch", flush=True)
    sc = _scenes()
    tris = np.concatenate([t for t, _ in sc])
    offs = np.concatenate([[0], np.cumsum([len(t) for t, _ in sc])]).astype(np.uint32)
    rays = [_rays(20000, 30 + k, lo, hi) for k, (_, (lo, hi)) in enumerate(sc)]
    roffs = np.concatenate([[0], np.cumsum([len(r) for r in rays])]).astype(np.uint32)
    allrays = np.concatenate(rays)
    bb = Bvh()
    bb.build_batch(DeviceArray.from_numpy(tris), offs, 1)
    before = _trace(bb, allrays, 0, roffs)
    dr = DeviceArray.from_numpy(allrays)
    dh = DeviceArray(len(allrays), HIT_DTYPE)
    t = bb.autotune(dr, len(allrays), dh, 0, roffs)
    print(f"tuning: stack {t.stackLdsEntries}, chunk {t.rayChunk}, {t.msPerLaunch:.3f} ms vs default {t.msDefault:.3f} ms",
          flush=True)
    assert t.stackLdsEntries in (16, 24, 32) and t.rayChunk in (32, 64, 128)
    assert 0 < t.msPerLaunch <= t.msDefault
    after = _trace(bb, allrays, 0, roffs)
    for f in ("prim", "t", "u", "v"):
        assert np.array_equal(before[f], after[f]), f
    # a batched BVH is traced only through the batched entry point
    with pytest.raises(ffi.RgError):
        bb.trace(dr, len(allrays), dh, 0)
