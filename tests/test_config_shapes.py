"""The batched-scenes and serving workloads at their full shapes on the GPU.

Batched scenes: the 64 scenes of the batched-scenes workload (triangle
counts log-uniform in [10k, 2M], seeded soups of icospheres and boxes,
20.6 M triangles; the same seeds as the bench) in one batched BVH.
- Every scene's hits from the one batched launch equal, bit for bit, the
  hits of that scene's own single-scene BVH (on a subsample of rays per
  scene: primary rays of the scene's camera plus random rays), and every
  hit lies on its ray and its triangle (t, barycentrics).
- The oracle builds and traces every scene (seven of them above 1M
  triangles) with the same rays: primitive ids, t, u, v bit-exact.

Serving: sessions at 960x540 and 2560x1440 over a 2M-triangle scene.
- Each session's persistent planes are the same alone and mixed with the
  other session on the server's streams.
- The 2560x1440 session's primary visibility equals rgDrawFrame's for the
  same cameras at the same size.
"""
from __future__ import annotations


import numpy as np
import pytest

import oracle_lib
from raytracedgl1_amd import ffi, rtgl, scenes, serving
from raytracedgl1_amd.ffi import HIT_DTYPE, RAY_DTYPE, Bvh, DeviceArray

pytestmark = pytest.mark.gpu

BATCH_SCENES = 64
RAYS_PER_SCENE = 4096


def _batch_workload():
    rng = np.random.default_rng(scenes.SEED_SCENE_BATCH)
    counts = np.exp(rng.uniform(np.log(1e4), np.log(2e6), BATCH_SCENES)).astype(np.int64)
    return rng, counts


def _scene_rays(rng, k, tris_k, n):
    """Half primary rays of a camera circling the scene, half random rays
    from inside its bounds."""
    ang = rng.uniform(0, 2 * np.pi)
    eye = np.array([180 * np.cos(ang), rng.uniform(20, 80), 180 * np.sin(ang)])
    prim = scenes.primary_rays(scenes.Camera(eye, np.zeros(3)), 128, 128)
    sel = rng.choice(len(prim), n // 2, replace=False)
    out = np.zeros(n, RAY_DTYPE)
    out[: n // 2] = prim[np.sort(sel)]
    lo = tris_k.reshape(-1, 3).min(0)
    hi = tris_k.reshape(-1, 3).max(0)
    m = n - n // 2
    out["origin"][n // 2:] = rng.uniform(lo, hi, (m, 3))
    d = rng.normal(size=(m, 3))
    out["direction"][n // 2:] = d / np.linalg.norm(d, axis=1, keepdims=True)
    out["tmin"][n // 2:] = 0.0
    out["tmax"][n // 2:] = 1000.0
    return out


def _on_ray_and_triangle(rays, hits, tris):
    """Hit point from (t, ray) and from (u, v, triangle) agree; t within the
    ray interval; barycentrics inside the triangle."""
    h = hits["prim"] != ffi.RG_EXT_MISS
    if not h.any():
        return 0
    r, hh = rays[h], hits[h]
    tri = tris[hh["prim"]].reshape(-1, 3, 3).astype(np.float64)
    u, v, t = hh["u"].astype(np.float64), hh["v"].astype(np.float64), hh["t"].astype(np.float64)
    assert (t >= r["tmin"]).all() and (t <= r["tmax"]).all()
    assert (u >= -1e-6).all() and (v >= -1e-6).all() and (u + v <= 1 + 1e-6).all()
    p_ray = r["origin"].astype(np.float64) + t[:, None] * r["direction"].astype(np.float64)
    p_tri = (1 - u - v)[:, None] * tri[:, 0] + u[:, None] * tri[:, 1] + v[:, None] * tri[:, 2]
    scale = np.maximum(1.0, np.abs(p_ray).max(axis=1))
    assert (np.abs(p_ray - p_tri).max(axis=1) <= 1e-4 * scale).all()
    return int(h.sum())


def test_batched_scenes_full_shape(gpu):
    rng, counts = _batch_workload()
    assert counts.max() > 1_000_000
    print(f"config-4 shape: {BATCH_SCENES} scenes, {int(counts.sum())} triangles", flush=True)
    tris = [scenes.soup_tris(scenes.SEED_SCENE_BATCH + 1 + i, int(c)) for i, c in enumerate(counts)]
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    d_tris = DeviceArray.from_numpy(np.concatenate(tris))
    bb = Bvh()
    bb.build_batch(d_tris, offs, 1)
    roots = bb.scene_roots()
    assert len(set(roots.tolist())) == BATCH_SCENES

    rrng = np.random.default_rng(0xC0FFEE04)
    rays = [_scene_rays(rrng, k, tris[k], RAYS_PER_SCENE) for k in range(BATCH_SCENES)]
    roffs = np.concatenate([[0], np.cumsum([len(r) for r in rays])]).astype(np.uint32)
    dr = DeviceArray.from_numpy(np.concatenate(rays))
    dh = DeviceArray(int(roffs[-1]), HIT_DTYPE)
    bb.trace_batch(dr, roffs, dh, 0)
    hb = dh.download()
    assert bb.trace_errors() == 0
    print("batched launch traced; comparing with single-scene builds", flush=True)

    hit_total = 0
    for k in range(BATCH_SCENES):
        b1 = Bvh()
        b1.build(DeviceArray.from_numpy(tris[k]), len(tris[k]), 1)
        d1 = DeviceArray.from_numpy(rays[k])
        h1d = DeviceArray(len(rays[k]), HIT_DTYPE)
        b1.trace(d1, len(rays[k]), h1d, 0)
        h1 = h1d.download()
        assert b1.trace_errors() == 0
        part = hb[roffs[k]:roffs[k + 1]]
        for f in ("prim", "t", "u", "v"):
            assert np.array_equal(part[f], h1[f]), (k, f)
        hit_total += _on_ray_and_triangle(rays[k], part, tris[k])
        del b1
    assert hit_total > BATCH_SCENES * RAYS_PER_SCENE // 8
    print(f"all scenes equal; {hit_total} hits checked on ray and triangle", flush=True)

    o = oracle_lib.Oracle()
    for k in range(BATCH_SCENES):
        oh = o.bvh(tris[k], 1).trace(rays[k], 0)
        part = hb[roffs[k]:roffs[k + 1]]
        for f in ("prim", "t", "u", "v"):
            assert np.array_equal(part[f], oh[f]), (k, int(counts[k]), f)
    print("every scene equal to the oracle", flush=True)


def _big_scene():
    """The 2M-triangle end of the serving scenes: the terrain-and-icospheres
    scene with a 724 x 724-cell heightfield and 2.0M triangles."""
    return scenes.terrain_spheres_scene(seed=scenes.SEED_TERRAIN_SPHERES + 5, cells=724, target_tris=2_000_000, lights=16)


PERSISTENT = ("FINAL", "RESERVOIRS", "ACCUM_HISTORY_LENGTH", "DIFF_COLOR_HISTORY", "VISIBILITY_BUFFER")
FRAMES = 2


def _serve(srv, sess, frames):
    for f in range(frames):
        for k, s in sess:
            srv.submit(s, serving.session_camera(k, 0.1 * f), (f + 1) / 60.0, f)
        srv.wait_idle()
    res = srv.poll()
    assert all(r[2] == serving.RG_EXT_FRAME_DONE for r in res), res
    return {k: {p: srv.framebuffers(s).download(ffi.PLANE[p]).copy() for p in PERSISTENT} for k, s in sess}


def test_serving_full_shape(gpu):
    sc = _big_scene()
    assert len(sc.tris) >= 1_990_000
    sizes = [(960, 540), (2560, 1440)]
    print(f"config-5 shape: {len(sc.tris)} triangles, sessions {sizes}", flush=True)
    alone = {}
    for k, (w, h) in enumerate(sizes):
        srv = serving.Server(2560, 1440, streams=2)
        try:
            alone.update(_serve(srv, [(k, srv.session(srv.scene(sc), w, h))], FRAMES))
        finally:
            srv.close()
        print(f"session {w}x{h} alone done", flush=True)
    srv = serving.Server(2560, 1440, streams=2, frames_in_flight=2)
    try:
        scn = srv.scene(sc)
        mixed = _serve(srv, [(k, srv.session(scn, w, h)) for k, (w, h) in enumerate(sizes)], FRAMES)
    finally:
        srv.close()
    for k in range(len(sizes)):
        for p in PERSISTENT:
            assert np.array_equal(alone[k][p], mixed[k][p]), (k, p)
    print("alone == mixed", flush=True)

    w, h = sizes[1]
    r = rtgl.Renderer(w, h)
    try:
        up = rtgl.SceneUploader(sc)
        for f in range(FRAMES):
            r.frame(up, serving.session_camera(1, 0.1 * f), geometry=True)
        vis = r.framebuffers().download(ffi.PLANE["VISIBILITY_BUFFER"]).copy()
    finally:
        r.close()
    assert np.array_equal(vis, alone[1]["VISIBILITY_BUFFER"])
    fin = alone[1]["FINAL"][..., :3]
    assert np.isfinite(fin).all() and 0.0 < fin.mean() < 1.0
