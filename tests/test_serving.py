"""The render server (rgExtServer*) on the GPU.

- A session renders the same pixels alone and among other sessions: four
  sessions over two scenes and two streams, with frames of different
  sessions in flight together and sharing their stream's within-frame
  planes, are compared bit for bit with each session rendered alone.
- A session's frames are the instance's frames: FINAL of a session after a
  few frames is close to rgDrawFrame's for the same cameras (the instance's
  first frame has no previous lights, the session's static scene maps every
  light to itself, so the two may differ in a few samples).
- Sessions are admitted by the memory budget, and a waiting request is
  replaced by a newer one.
- Batched launches (batchFrames > 1: one traversal launch per stage over
  the rays of several sessions of a scene) render every session's frames
  bit for bit as alone.
- Scene groups: sessions of two scenes in one group share launches; every
  session's frames are bit for bit its frames rendered alone through the
  group and its frames through its scene's own BVH (the group's BVH is
  another tree over the same triangles, and the closest-hit rule orders
  hits by distance, then primitive, whatever the visiting order).  Group
  creation refuses scenes with other light kinds, and a grouped scene
  cannot be destroyed.
- Readback: a finished frame's pImage holds FINAL as RGBA8 with the sRGB
  transfer (within one code of a float64 encoding of the FINAL plane).
- Scheduling: NEWEST_FIRST starts the newest waiting request first, and
  requests that wait longer than maxQueueSeconds are reported expired.
"""
import ctypes as C
import time

import numpy as np
import pytest

from raytracedgl1_amd import ffi, rtgl, scenes, serving

pytestmark = pytest.mark.gpu

PERSISTENT = ("FINAL", "RESERVOIRS", "ACCUM_HISTORY_LENGTH", "DIFF_COLOR_HISTORY", "DIFF_ACCUM_MOMENTS",
              "INDIR_RESERVOIR_WEIGHT", "DIS_GRADIENT_HISTORY", "VISIBILITY_BUFFER")
FRAMES = 4


def _scenes():
    a = scenes.terrain_boxes_scene()
    b = scenes.terrain_boxes_scene(seed=scenes.SEED_TERRAIN_BOXES + 7)
    r = np.random.default_rng(5)
    lights = {"sun": a.lights["sun"], "spheres": [
        {"color": (1.0, 0.8, 0.6), "intensity": float(r.uniform(50, 300)),
         "position": (float(r.uniform(-20, 20)), float(r.uniform(4, 8)), float(r.uniform(-20, 20))), "radius": 0.5}
        for _ in range(6)]}
    b = scenes.Scene(b.tris, b.normals, b.colors, b.objects, lights)
    return a, b


# (scene index, width, height)
SPECS = [(1, 128, 72), (0, 160, 90), (0, 200, 100), (1, 160, 90)]


def _render(server, sess, frames, batches=None):
    """Frames for every session in `sess`, one request per session per round,
    all of a round in flight together; persistent planes after the last.
    `batches` collects the batch size of every finished frame."""
    res = []
    for f in range(frames):
        for k, s in sess:
            server.submit(s, serving.session_camera(k, 0.1 * f), (f + 1) / 60.0, f)
        server.wait_idle()
        res += server.poll()
    assert all(r[2] == serving.RG_EXT_FRAME_DONE for r in res), res
    if batches is not None:
        batches.extend(r[8] for r in res)
    out = {}
    for k, s in sess:
        fb = server.framebuffers(s)
        out[k] = {p: fb.download(ffi.PLANE[p]).copy() for p in PERSISTENT}
    return out


def _alone(sc, specs):
    alone = {}
    for k, (si, w, h) in enumerate(specs):
        srv = serving.Server(256, 128, streams=2)
        try:
            scn = srv.scene(sc[si])
            alone.update(_render(srv, [(k, srv.session(scn, w, h))], FRAMES))
        finally:
            srv.close()
    return alone


def test_session_same_alone_and_in_the_mix():
    print("serving: 4 sessions alone vs mixed, 2 streams", flush=True)
    sc = _scenes()
    alone = _alone(sc, SPECS)
    srv = serving.Server(256, 128, streams=2, frames_in_flight=2)
    try:
        handles = [srv.scene(s) for s in sc]
        sess = [(k, srv.session(handles[si], w, h)) for k, (si, w, h) in enumerate(SPECS)]
        mixed = _render(srv, sess, FRAMES)
        st = srv.stats()
        assert st.sessions == 4 and st.framesDone == 4 * FRAMES and st.framesDropped == 0
    finally:
        srv.close()
    for k in range(len(SPECS)):
        for p in PERSISTENT:
            assert np.array_equal(alone[k][p], mixed[k][p]), (k, p)
        fin = mixed[k]["FINAL"][..., :3]
        assert np.isfinite(fin).all() and 0.0 < fin.mean() < 1.0


def test_session_frames_are_the_instance_frames():
    print("serving: session vs rgDrawFrame, terrain-boxes 160x90", flush=True)
    a, _ = _scenes()
    cams = [serving.session_camera(1, 0.1 * f) for f in range(FRAMES)]
    r = rtgl.Renderer(160, 90)
    try:
        up = rtgl.SceneUploader(a)
        for f in range(FRAMES):
            r.frame(up, cams[f])
        ref = r.framebuffers().download(ffi.PLANE["FINAL"]).copy()
        ref_vis = r.framebuffers().download(ffi.PLANE["VISIBILITY_BUFFER"]).copy()
    finally:
        r.close()
    srv = serving.Server(160, 90, streams=1)
    try:
        s = srv.session(srv.scene(a), 160, 90)
        got = _render(srv, [(1, s)], FRAMES)[1]
    finally:
        srv.close()
    # the same primary visibility, and a tone-mapped image within a few codes
    assert np.array_equal(got["VISIBILITY_BUFFER"], ref_vis)
    d = np.abs(got["FINAL"][..., :3].astype(np.float64) - ref[..., :3])
    print(f"session vs instance FINAL: mean |d| {d.mean():.2e}, p99 {np.quantile(d, 0.99):.2e}", flush=True)
    assert d.mean() < 2e-3 and np.quantile(d, 0.99) < 5e-2


def test_admission_by_memory_budget():
    print("serving: admission by budget", flush=True)
    a, _ = _scenes()
    probe = serving.Server(160, 90, streams=1)
    try:
        scn = probe.scene(a)
        probe.session(scn, 160, 90)
        per = probe.stats().sessionBytes
    finally:
        probe.close()
    srv = serving.Server(160, 90, streams=2, budget_bytes=2 * per)
    try:
        scn = srv.scene(a)
        s1 = srv.session(scn, 160, 90)
        srv.session(scn, 160, 90)
        with pytest.raises(serving.OverBudget):
            srv.session(scn, 160, 90)
        assert srv.stats().sessionBytes == 2 * per
        srv.destroy_session(s1)
        srv.session(scn, 160, 90)
        assert srv.stats().sessions == 2
    finally:
        srv.close()


def test_waiting_request_is_replaced():
    print("serving: newest request wins", flush=True)
    a, _ = _scenes()
    srv = serving.Server(640, 360, streams=1, frames_in_flight=1)
    try:
        s = srv.session(srv.scene(a), 640, 360)
        for tag in range(3):
            srv.submit(s, serving.session_camera(0, 0.1 * tag), (tag + 1) / 60.0, tag)
        srv.wait_idle()
        res = sorted(srv.poll(), key=lambda r: r[1])
    finally:
        srv.close()
    assert [r[1] for r in res] == [0, 1, 2]
    status = [r[2] for r in res]
    assert status[2] == serving.RG_EXT_FRAME_DONE
    assert status.count(serving.RG_EXT_FRAME_DROPPED) >= 1
    done = [r for r in res if r[2] == serving.RG_EXT_FRAME_DONE]
    assert [r[3] for r in done] == list(range(len(done)))  # frame ids count rendered frames
    assert all(r[4] <= r[5] <= r[6] for r in done)


# six sessions over two scenes; batches form among the sessions of a scene
BATCH_SPECS = SPECS + [(0, 96, 64), (0, 256, 128)]


def test_batched_sessions_match_alone():
    print("serving: 6 sessions in batches of up to 4, one stream", flush=True)
    sc = _scenes()
    alone = _alone(sc, BATCH_SPECS)
    srv = serving.Server(256, 128, streams=1, frames_in_flight=1, batch=4)
    batches = []
    try:
        handles = [srv.scene(s) for s in sc]
        sess = [(k, srv.session(handles[si], w, h)) for k, (si, w, h) in enumerate(BATCH_SPECS)]
        got = _render(srv, sess, FRAMES, batches)
    finally:
        srv.close()
    print(f"  batch sizes of the finished frames: {sorted(batches)}", flush=True)
    assert max(batches) > 1 and max(batches) <= 4
    for k in range(len(BATCH_SPECS)):
        for p in PERSISTENT:
            assert np.array_equal(alone[k][p], got[k][p]), (k, p)


GROUP_SPECS = [(0, 160, 90), (1, 128, 72), (0, 200, 100), (1, 160, 90), (0, 128, 72), (1, 200, 100)]


def _group_scenes():
    """Two plain scenes with the same light kinds (a sun each)."""
    return scenes.terrain_boxes_scene(), scenes.terrain_boxes_scene(seed=scenes.SEED_TERRAIN_BOXES + 7)


def _grouped(specs, batch):
    """specs rendered on one server whose two scenes form a group; returns
    (planes per session index, batch sizes)."""
    sc = _group_scenes()
    srv = serving.Server(256, 128, streams=1, frames_in_flight=1, batch=batch)
    batches = []
    try:
        handles = [srv.scene(s) for s in sc]
        g = srv.group(handles)
        sess = [(k, srv.session(handles[si], w, h)) for k, (si, w, h) in specs]
        got = _render(srv, sess, FRAMES, batches)
        with pytest.raises(ffi.RgError):  # a grouped scene stays
            ffi.check("rgExtSceneDestroy", serving._lib().rgExtSceneDestroy(handles[0]))
        srv.destroy_group(g)
    finally:
        srv.close()
    return got, batches


def test_scene_group_batches_match_alone():
    print("serving: 6 sessions of two grouped scenes in batches of up to 4", flush=True)
    specs = list(enumerate(GROUP_SPECS))
    alone = {}
    for k, spec in specs:
        alone.update(_grouped([(k, spec)], 1)[0])
    got, batches = _grouped(specs, 4)
    print(f"  batch sizes of the finished frames: {sorted(batches)}", flush=True)
    # 6 sessions, 3 per scene: a batch of 4 holds frames of both scenes
    assert max(batches) == 4
    for k in range(len(GROUP_SPECS)):
        for p in PERSISTENT:
            assert np.array_equal(alone[k][p], got[k][p]), (k, p)
    # the same sessions through each scene's own BVH
    sc = _group_scenes()
    own = {}
    for k, (si, w, h) in specs:
        srv = serving.Server(256, 128, streams=1)
        try:
            own.update(_render(srv, [(k, srv.session(srv.scene(sc[si]), w, h))], FRAMES))
        finally:
            srv.close()
    for k in range(len(GROUP_SPECS)):
        for p in PERSISTENT:
            assert np.array_equal(own[k][p], got[k][p]), (k, p)


def test_scene_group_fp16_nodes():
    """The precision form's FP16 nodes in a group's batched BVH: grouped and
    batched frames equal each session's frames through its own FP16-node
    BVH, bit for bit."""
    print("serving: grouped scenes with FP16 nodes, batches of up to 4", flush=True)
    specs = list(enumerate(GROUP_SPECS[:4]))
    sc = _group_scenes()
    prec = ffi.RG_EXT_PRECISION_FP16_NODES
    srv = serving.Server(256, 128, streams=1, frames_in_flight=1, batch=4, precision=prec)
    batches = []
    try:
        handles = [srv.scene(s) for s in sc]
        srv.group(handles)
        got = _render(srv, [(k, srv.session(handles[si], w, h)) for k, (si, w, h) in specs], FRAMES, batches)
    finally:
        srv.close()
    print(f"  batch sizes of the finished frames: {sorted(batches)}", flush=True)
    assert max(batches) > 1
    for k, (si, w, h) in specs:
        srv = serving.Server(256, 128, streams=1, precision=prec)
        try:
            own = _render(srv, [(k, srv.session(srv.scene(sc[si]), w, h))], FRAMES)
        finally:
            srv.close()
        for p in PERSISTENT:
            assert np.array_equal(own[k][p], got[k][p]), (k, p)


def test_scene_group_refuses_other_light_kinds():
    a, _ = _scenes()
    _, b = _scenes()  # a sun and sphere lights
    srv = serving.Server(160, 90, streams=1)
    try:
        ha, hb = srv.scene(a), srv.scene(b)
        with pytest.raises(ffi.RgError):
            srv.group([ha, hb])
        with pytest.raises(ffi.RgError):
            srv.group([ha])
        g = srv.group([ha, srv.scene(scenes.terrain_boxes_scene(seed=scenes.SEED_TERRAIN_BOXES + 9))])
        with pytest.raises(ffi.RgError):
            srv.group([ha, hb])  # ha is in a group already
        srv.destroy_group(g)
        ffi.check("rgExtSceneDestroy", serving._lib().rgExtSceneDestroy(ha))
    finally:
        srv.close()


def _srgb8(c):
    """RGBA8 code of linear values (IEC 61966-2-1), float64."""
    c = np.clip(c.astype(np.float64), 0.0, 1.0)
    e = np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1.0 / 2.4) - 0.055)
    return np.floor(np.clip(e, 0.0, 1.0) * 255.0 + 0.5).astype(np.int64)


def test_readback_image_is_the_final_plane():
    print("serving: readback of FINAL as RGBA8", flush=True)
    a, _ = _scenes()
    srv = serving.Server(160, 90, streams=1, readback=True)
    try:
        s = srv.session(srv.scene(a), 160, 90)
        res = []
        for f in range(3):
            srv.submit(s, serving.session_camera(1, 0.1 * f), (f + 1) / 60.0, f)
            srv.wait_idle()
            res += srv.poll()
        assert all(r[2] == serving.RG_EXT_FRAME_DONE and r[7] for r in res)
        assert res[-1][7] != res[-2][7]  # the two images are used in turn
        img = np.frombuffer(C.string_at(res[-1][7], 160 * 90 * 4), np.uint8).reshape(90, 160, 4).astype(np.int64)
        fin = srv.framebuffers(s).download(ffi.PLANE["FINAL"])
    finally:
        srv.close()
    ref = _srgb8(fin[..., :3])
    d = np.abs(img[..., :3] - ref)
    print(f"  readback vs FINAL: max code difference {d.max()}, differing {np.mean(d > 0):.4f}", flush=True)
    assert d.max() <= 1 and np.mean(d > 0) < 0.01
    assert (img[..., 3] == 255).all()
    assert 0 < img[..., :3].mean() < 255


def test_newest_first_and_expiry():
    print("serving: NEWEST_FIRST order and maxQueueSeconds", flush=True)
    a, _ = _scenes()
    srv = serving.Server(2560, 1440, streams=1, frames_in_flight=1, schedule=serving.RG_EXT_SCHEDULE_NEWEST_FIRST)
    try:
        scn = srv.scene(a)
        big = srv.session(scn, 2560, 1440)
        small = [srv.session(scn, 160, 90) for _ in range(3)]
        srv.submit(big, serving.session_camera(0, 0.0), 1 / 60.0, 100)
        time.sleep(0.0005)  # the dispatcher takes the 2560x1440 frame; the others wait behind it
        for i, s in enumerate(small):
            srv.submit(s, serving.session_camera(i + 1, 0.0), 1 / 60.0, i)
        srv.wait_idle()
        res = {r[1]: r for r in srv.poll()}
    finally:
        srv.close()
    assert all(r[2] == serving.RG_EXT_FRAME_DONE for r in res.values())
    # of two requests that were both waiting when the first of them started,
    # the newer one starts first (how many waited together depends on when
    # the 2560x1440 frame finished)
    sub = {i: res[i][4] for i in range(3)}
    start = {i: res[i][5] for i in range(3)}
    print(f"  start order of tags 0, 1, 2: {sorted(range(3), key=lambda i: start[i])}", flush=True)
    pairs = [(i, j) for i in range(3) for j in range(3) if sub[j] > sub[i] and sub[j] < min(start[i], start[j])]
    assert pairs, "no two requests waited together"
    assert all(start[j] < start[i] for i, j in pairs)
    srv = serving.Server(160, 90, streams=1, frames_in_flight=1, max_queue_s=1e-9)
    try:
        scn = srv.scene(a)
        sess = [srv.session(scn, 160, 90) for _ in range(4)]
        for i, s in enumerate(sess):
            srv.submit(s, serving.session_camera(i, 0.0), 1 / 60.0, i)
        srv.wait_idle()
        res = srv.poll()
        st = srv.stats()
    finally:
        srv.close()
    status = [r[2] for r in res]
    assert len(res) == 4 and all(x in (serving.RG_EXT_FRAME_DONE, serving.RG_EXT_FRAME_EXPIRED) for x in status)
    assert status.count(serving.RG_EXT_FRAME_EXPIRED) >= 1
    assert st.framesDropped == status.count(serving.RG_EXT_FRAME_EXPIRED)
