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
"""
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


def _render(server, sess, frames):
    """Frames for every session in `sess`, one request per session per round,
    all of a round in flight together; persistent planes after the last."""
    for f in range(frames):
        for k, s in sess:
            server.submit(s, serving.session_camera(k, 0.1 * f), (f + 1) / 60.0, f)
        server.wait_idle()
    res = server.poll()
    assert all(r[2] == serving.RG_EXT_FRAME_DONE for r in res), res
    out = {}
    for k, s in sess:
        fb = server.framebuffers(s)
        out[k] = {p: fb.download(ffi.PLANE[p]).copy() for p in PERSISTENT}
    return out


def test_session_same_alone_and_in_the_mix():
    print("serving: 4 sessions alone vs mixed, 2 streams", flush=True)
    sc = _scenes()
    alone = {}
    for k, (si, w, h) in enumerate(SPECS):
        srv = serving.Server(256, 128, streams=2)
        try:
            scn = srv.scene(sc[si])
            alone.update(_render(srv, [(k, srv.session(scn, w, h))], FRAMES))
        finally:
            srv.close()
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
