#!/usr/bin/env python3
"""Frames/s and Mrays/s of the full ray-traced 1080p frame on one MI355X.

Workload: 1920x1080, a 999,168-triangle scene (512x512-cell fBm heightfield
plus 500 icospheres of subdivision 2-5), 64 sphere lights, the 60-frame
camera orbit (radius 80, height 25, 6 degrees per frame).  One step is one
rgDrawFrame through the RTGL1 C API: primary G-buffer, light grid, ReSTIR DI
(32 initial candidates, 1 temporal + 1 spatial sample) with shadow rays, two
diffuse bounces, A-SVGF with 5 a-trous iterations and eye adaptation.  The
geometry is uploaded once before the timed region and stays resident in HBM
(static geometry); lights and camera are passed every frame.

Prints one JSON line.  `roofline` prices the dominant kernel (the traversal
kernel k_trace) by algorithmic bytes: 32 B ray in + 16 B hit out per ray,
256 B per BVH8 node visit, 36 B per triangle test, divided by the kernel time
measured with HIP events recorded on the instance stream around every launch.
`cpu_baseline` times the CPU oracle on a bounded sample of the same frame.

Beside the FP32 line (`value`), the same JSON object carries
- `fp16`: the same frames with FP16 BVH nodes and FP16 storage of the
  denoiser's temporal and a-trous planes (frames/s, Mrays/s, and the last
  frame's denoised-radiance difference from the FP32 run);
- `reflrefr`: the same frames with every third icosphere a mirror, glass or
  water (reflection / refraction depth 2);
- `serving`: 256 camera sessions (960x540 .. 2560x1440) over 16 scenes of
  200k-2M triangles, admitted against a 200 GiB budget, with Poisson frame
  requests at 30 Hz per session for a few seconds: p50 / p99 frame latency
  and frames/s at that load, then at a sustainable number of sessions.
- `batch`: 64 scenes of 10k-2M triangles built as one batch and as 64
  builds, and their primary rays (one camera per scene, 256x256 .. 1080p)
  traced in one launch and as 64, before and after the traversal autotuner.

Usage: python bench.py [--gpus 1] [--steps 60] [--warmup 5] [--no-cpu-baseline] [--no-fp16] [--no-reflrefr]
                       [--no-serving] [--no-batch]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))

METRIC = "Mrays/s & 1080p frames/s (FP16 vs FP32); p99 frame latency at 256 sessions"
WIDTH, HEIGHT = 1920, 1080
ORBIT_FRAMES = 60
BYTES_PER_RAY = 32 + 16      # RgExtRay in, RgExtHit out
BYTES_PER_NODE = 256         # one FP32 BVH8 node
BYTES_PER_TRI = 36           # three FP32 vertices
HBM_PEAK_GBPS = 8000.0       # MI355X HBM3E
TRAFFIC_FILE = REPO / "profiles" / "k_trace_traffic.json"
DENOISE_TRAFFIC_FILE = REPO / "profiles" / "denoise_1080p_traffic.json"
SERVING_BATCH = 4            # frames of one scene per serving launch
# serving loads at 256 sessions: (schedule, frames per launch, launches in
# flight per stream, streams, expiry seconds, seed, description)
SERVING_POLICIES = {
    "full_load_unbatched": (0, 1, 2, 4, 0.0, 1, "one frame per launch, 2 launches queued on each of 4 streams, oldest request "
                                               "first; a waiting request is replaced by its session's newer one"),
    "full_load": (0, SERVING_BATCH, 2, 4, 0.0, 2, f"up to {SERVING_BATCH} frames of one scene per launch (one traversal launch "
                                                  "per stage), 2 launches queued on each of 4 streams, oldest request first; a "
                                                  "waiting request is replaced by its session's newer one"),
    "full_load_latency": (1, 1, 1, 4, 0.02, 3, "one frame per launch, 1 launch queued on each of 4 streams, newest request "
                                               "first, requests waiting over 20 ms expire; a waiting request is replaced by "
                                               "its session's newer one"),
    "full_load_batched_latency": (1, 2, 1, 4, 0.02, 4, "up to 2 frames of one scene per launch, 1 launch queued on each of "
                                                       "4 streams, newest request first, requests waiting over 20 ms expire; "
                                                       "a waiting request is replaced by its session's newer one"),
}


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (default: min(16, cpus))")
    ap.add_argument("--dynamic-upload", action="store_true",
                    help="re-upload every primitive each frame (the host-to-device-inclusive rate; not the headline)")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16 precision-form run")
    ap.add_argument("--no-reflrefr", action="store_true", help="skip the reflection / refraction run")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving run")
    ap.add_argument("--serving-sessions", type=int, default=256)
    ap.add_argument("--serving-scenes", type=int, default=16)
    ap.add_argument("--serving-seconds", type=float, default=5.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes run")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="instead of the frame benchmark, run 5 plain device copies of MIB MiB (k_copy16): the "
                         "known-size reference for scaling HBM counter readings")
    return ap.parse_args(argv)


CPU_SAMPLE = (960, 540)  # a quarter of the 1080p frame's pixels


def _oracle_params(w, h, frame, prev_cam, cam):
    """Frame constants of orbit frame `frame` at w x h for the oracle: the
    bench workload's settings (ReSTIR DI 32/1/1, 5 a-trous iterations)."""
    import numpy as np
    from raytracedgl1_amd import ffi, scenes

    p = ffi.frame_params(w, h)

    def mats(c):
        view = c.view().astype(np.float32)
        proj = c.proj(w, h)
        iv = np.linalg.inv(scenes.mat_cm(view)).T.reshape(16).astype(np.float32)
        ip = np.linalg.inv(scenes.mat_cm(proj)).T.reshape(16).astype(np.float32)
        return view, proj, iv, ip

    view, proj, iv, ip = mats(cam)
    pview, pproj, piv, _ = mats(prev_cam)
    p.view[:], p.projection[:] = view.tolist(), proj.tolist()
    p.viewPrev[:], p.projectionPrev[:] = pview.tolist(), pproj.tolist()
    p.invView[:], p.invProjection[:] = iv.tolist(), ip.tolist()
    p.cameraPosition[:] = [float(iv[12]), float(iv[13]), float(iv[14]), 0.0]
    p.cameraPositionPrev[:] = [float(piv[12]), float(piv[13]), float(piv[14]), 0.0]
    p.frameId = frame
    p.rayLength = 10000.0
    p.primaryRayMinDist = 0.1
    p.restirInitialCandidates, p.restirTemporalSamples, p.restirSpatialSamples = 32, 1, 1
    p.atrousIterations = 5
    return p


def cpu_baseline(scene, threads: int, frames: int = 5) -> dict:
    """The CPU oracle running whole frames of the bench workload on a bounded
    sample: the traced passes (G-buffer, light grid, ReSTIR DI 32/1/1 with
    shadow rays, two bounces, ReSTIR GI, gradient samples; or_pt_frame) and
    the A-SVGF chain with 5 a-trous iterations (or_denoise), both on
    `threads` host threads, over the same 999,168-triangle scene, 64 sphere
    lights and orbit cameras at 960x540.  The median of `frames` timed
    frames (after one untimed frame) per stage, scaled to 1920x1080 by pixel
    count.  The oracle computes in FP64 with FP32 storage of its planes'
    inputs; its BVH is built once, untimed, as the GPU's stays resident."""
    import statistics

    sys.path.insert(0, str(REPO / "tests"))
    import oracle_lib
    from raytracedgl1_amd import scenes

    w, h = CPU_SAMPLE
    o = oracle_lib.Oracle()
    pt = oracle_lib.OraclePathTracer(o, scene, w, h, leaf=1)
    fb = oracle_lib.OracleFramebuffers(o, w, h)
    o.l.or_denoise_set_threads(threads)
    traced, denoise = [], []
    t_all = time.perf_counter()
    try:
        for f in range(frames + 1):
            if f:
                fb.swap()
            p = _oracle_params(w, h, f, scenes.orbit_camera(max(0, f - 1) % ORBIT_FRAMES), scenes.orbit_camera(f % ORBIT_FRAMES))
            t0 = time.perf_counter()
            pt.frame(p, fb, threads=threads)
            t1 = time.perf_counter()
            fb.run("", p)
            t2 = time.perf_counter()
            if f:
                traced.append(t1 - t0)
                denoise.append(t2 - t1)
            print(f"cpu baseline frame {f}: traced passes {t1 - t0:.2f} s, a-trous chain {t2 - t1:.2f} s", file=sys.stderr, flush=True)
    finally:
        o.l.or_denoise_set_threads(1)
    scale = (WIDTH * HEIGHT) / (w * h)
    med_t, med_d = statistics.median(traced), statistics.median(denoise)
    frame_s = (med_t + med_d) * scale
    return {
        "value": 1.0 / frame_s, "unit": "frames/s", "cores": threads, "kind": "port",
        "sample": (f"oracle frames of the same workload at {w}x{h} (1/{scale:.0f} of the pixels) on {threads} host threads: "
                   f"traced passes (G-buffer, ReSTIR DI 32/1/1 + shadow rays, 2 bounces, ReSTIR GI, gradients) and "
                   f"A-SVGF (5 a-trous iterations); median of {frames} frames after 1 untimed, scaled to 1920x1080 by "
                   f"pixel count; FP64 arithmetic"),
        "stage_s_per_1080p_frame": {"traced_passes": med_t * scale, "denoise": med_d * scale},
        "seconds": time.perf_counter() - t_all,
    }


def calibrate_copy(mib: int) -> int:
    import ctypes as C
    from raytracedgl1_amd import ffi
    nbytes = mib << 20
    src, dst = ffi.DeviceArray(nbytes, "u1"), ffi.DeviceArray(nbytes, "u1")
    ffi.call("rgExtMemsetDevice", src.ptr, 1, nbytes)
    t = ffi.Timer()
    times = []
    for _ in range(5):
        t.start()
        ffi.call("rgExtCopyDeviceToDevice", dst.ptr, src.ptr, nbytes, None)
        t.stop()
        times.append(t.elapsed_ms())
    ms = sorted(times)[2]
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "dispatches": 5,
                      "median_ms": ms, "gbps_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
    return 0


def run_frames(a, scene, precision: int, flags=None, planes=("PRE_FINAL",)):
    """Warm-up, then a.steps timed rgDrawFrame calls of the orbit; returns the
    per-frame stats, the wall time of the timed frames and the last frame's
    PRE_FINAL plane (with `planes` beyond it: a dict of them).  flags maps
    object indices to RgMeshPrimitiveFlags."""
    from raytracedgl1_amd import ffi, rtgl, scenes

    up = rtgl.SceneUploader(scene, flags)
    r = rtgl.Renderer(WIDTH, HEIGHT)
    try:
        r.set_static_geometry(not a.dynamic_upload)
        r.path.atrousIterations = 5
        r.path.restirInitialCandidates = 32
        r.path.restirTemporalSamples = 1
        r.path.restirSpatialSamples = 1
        r.path.collectTraversalStats = 1
        r.path.precisionFlags = precision
        frame = [0]

        def step(geometry=False, stats=True):
            geometry = geometry or a.dynamic_upload
            st = r.frame(up, scenes.orbit_camera(frame[0] % ORBIT_FRAMES), geometry=geometry, stats=stats)
            frame[0] += 1
            return st

        # warm-up; the first step uploads the scene, which then stays resident
        for i in range(max(1, a.warmup)):
            step(geometry=(i == 0))
            print(f"warmup {i + 1}/{max(1, a.warmup)} (precision flags {precision})", file=sys.stderr, flush=True)
        ffi.call("rgExtSynchronize")
        # the a-trous workgroup tile, tuned on the warm frame's planes (the
        # timed frames use the kept width; results do not depend on it)
        import ctypes as C
        width, ms = C.c_uint32(0), (C.c_float * 4)()
        ffi.call("rgExtDenoiseAutotune", r.framebuffers().h, r.frame_params(), None, C.byref(width), ms)
        tune = {"tile_width": width.value, "tile": f"{width.value}x{256 // width.value}",
                "iteration1_ms": {f"{w}x{256 // w}": round(float(m), 4) for w, m in zip((16, 32, 64, 128), ms)}}
        if precision == 0:
            a.atrous_tune = tune
        print(f"a-trous tile autotune (precision flags {precision}): {tune}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step(stats=False)  # two frames in flight, as an application runs them
        ffi.call("rgExtSynchronize")
        wall = time.perf_counter() - t0
        print(f"timed {a.steps} frames in {wall:.3f} s (precision flags {precision})", file=sys.stderr, flush=True)
        # per-stage GPU times and traversal counts of the next a.steps orbit
        # frames, each read once its frame has finished (outside the timed
        # region)
        stats = [step() for _ in range(a.steps)]
        assert all(st.bvhRebuilt == 0 for st in stats), "static geometry was rebuilt inside the timed region"
        fb = r.framebuffers()
        if planes == ("PRE_FINAL",):
            return stats, wall, fb.download(ffi.PLANE["PRE_FINAL"]).copy()
        return stats, wall, {k: fb.download(ffi.PLANE[k]).copy() for k in planes}
    finally:
        r.close()


def serving_run(a) -> dict:
    """The serving form: sessions over shared scenes, Poisson requests at
    30 Hz per session, FP16 storage and nodes, every finished frame read
    back to host memory as RGBA8.  The loads of SERVING_POLICIES on one
    server at every session (unbatched and batched throughput, and a
    latency-bounded policy), then a sustainable load of sessions sized from
    the batched throughput.  Latency = submit -> readback done, over finished
    frames; replaced and expired requests are counted beside it."""
    import numpy as np

    from raytracedgl1_amd import ffi, scenes, serving

    rng = np.random.default_rng(scenes.SEED_SERVING)
    counts = np.exp(rng.uniform(np.log(200_000), np.log(2_000_000), a.serving_scenes)).astype(int)
    res = [serving.SERVING_RESOLUTIONS[i] for i in rng.integers(0, 4, a.serving_sessions)]
    precision = ffi.RG_EXT_PRECISION_FP16_STORAGE | ffi.RG_EXT_PRECISION_FP16_NODES
    srv = serving.Server(2560, 1440, streams=4, frames_in_flight=2, precision=precision, budget_bytes=200 << 30,
                         atrous=5, candidates=32, temporal=1, spatial=1, batch=SERVING_BATCH, readback=True)
    try:
        t = time.perf_counter()
        handles = []
        for i, c in enumerate(counts):
            handles.append(srv.scene(serving.serving_scene(i, int(c))))
            print(f"serving scene {i + 1}/{len(counts)}: {int(c)} triangles", file=sys.stderr, flush=True)
        sessions, refused = [], 0
        for k, (w, h) in enumerate(res):
            try:
                sessions.append(srv.session(handles[k % len(handles)], w, h))
            except serving.OverBudget:
                refused += 1
        st = srv.stats()
        print(f"serving: {len(sessions)} sessions admitted, {refused} refused, set-up {time.perf_counter() - t:.1f} s",
              file=sys.stderr, flush=True)
        # warm every session once (first frames allocate the traversal stacks)
        for k, s in enumerate(sessions):
            srv.submit(s, serving.session_camera(k, 0.0), 0.0, 0)
        srv.wait_idle()
        srv.poll()
        runs = {}
        for name, (sched, batch, fif, streams, expiry, seed, desc) in SERVING_POLICIES.items():
            srv.set_policy(sched, batch, fif, streams, expiry)
            r = serving.run_poisson(srv, sessions, 30.0, a.serving_seconds, seed=seed, t_offset=50.0 * seed)
            r["policy"] = desc
            runs[name] = r
            print(f"serving at {len(sessions)} sessions, {name}: {r}", file=sys.stderr, flush=True)
        # sustainable load: the first sessions, as many as 70% of the batched
        # throughput can serve at 30 Hz, under the latency policy
        m = max(1, min(len(sessions), int(0.7 * runs["full_load"]["frames_per_s"] / 30.0)))
        sched, batch, fif, streams, expiry, _, desc = SERVING_POLICIES["full_load_latency"]
        srv.set_policy(sched, batch, fif, streams, expiry)
        light = serving.run_poisson(srv, sessions[:m], 30.0, min(3.0, a.serving_seconds), seed=9, t_offset=900.0)
        light["policy"] = desc
        print(f"serving at {m} sessions: {light}", file=sys.stderr, flush=True)
        mpx = float(np.mean([s.width * s.height for s in sessions])) / 1e6 if sessions else 0.0
        return {
            "sessions_requested": a.serving_sessions, "sessions_admitted": len(sessions), "sessions_refused": refused,
            "scenes": len(handles), "scene_triangles": [int(c) for c in counts], "mean_mpix_per_session": mpx,
            "memory_budget_gib": 200, "session_gib": st.sessionBytes / 2**30, "shared_gib": st.sharedBytes / 2**30,
            "scene_gib": st.sceneBytes / 2**30, "streams": st.streams, "precision": "fp16 storage + fp16 nodes",
            "readback": "FINAL as RGBA8 (sRGB) to pinned host memory; latency ends when the copy is done",
            "request_rate_hz_per_session": 30.0, **runs, "sustainable_load": light,
        }
    finally:
        srv.close()


def batch_run(a) -> dict:
    """Batched scenes: 64 soups of 10k-2M triangles (log-uniform), one camera
    each at a resolution from {256x256, 640x360, 1280x720, 1920x1080}.  The
    scenes are built as one batch (one pass sequence over every triangle) and
    as 64 separate builds; their primary rays are traced in one persistent
    launch over all scenes and as 64 launches; the traversal autotuner then
    picks the batch's stack depth and ray chunk."""
    import ctypes as C

    import numpy as np

    from raytracedgl1_amd import ffi, scenes

    rng = np.random.default_rng(scenes.SEED_SCENE_BATCH)
    S = a.batch_scenes
    counts = np.exp(rng.uniform(np.log(1e4), np.log(2e6), S)).astype(np.int64)
    resl = ((256, 256), (640, 360), (1280, 720), (1920, 1080))
    res = [resl[k] for k in rng.integers(0, 4, S)]
    t = time.perf_counter()
    soups = [scenes.soup_tris(scenes.SEED_SCENE_BATCH + 1 + i, int(c)) for i, c in enumerate(counts)]
    tris = np.concatenate(soups)
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    print(f"batch: {S} scenes, {len(tris)} triangles generated in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    d_tris = ffi.DeviceArray.from_numpy(tris)
    del tris
    params = ffi.RgExtBvhBuildParams(1, ffi.RG_EXT_BVH_NODE_FP32)

    bb = ffi.Bvh()
    bb.build_batch(d_tris, offs, 1)  # warm (allocates)
    ffi.call("rgExtSynchronize")
    t0 = time.perf_counter()
    bb.build_batch(d_tris, offs, 1)
    ffi.call("rgExtSynchronize")
    t_batch_build = time.perf_counter() - t0

    singles = [ffi.Bvh() for _ in range(S)]
    for k in range(S):  # warm
        ffi.call("rgExtBvhBuild", singles[k].h, d_tris.ptr + int(offs[k]) * 36, int(counts[k]), C.byref(params), None)
    ffi.call("rgExtSynchronize")
    t0 = time.perf_counter()
    for k in range(S):
        ffi.call("rgExtBvhBuild", singles[k].h, d_tris.ptr + int(offs[k]) * 36, int(counts[k]), C.byref(params), None)
    ffi.call("rgExtSynchronize")
    t_single_build = time.perf_counter() - t0
    print(f"batch build {t_batch_build * 1e3:.1f} ms, 64 single builds {t_single_build * 1e3:.1f} ms", file=sys.stderr, flush=True)

    roffs = np.concatenate([[0], np.cumsum([w * h for w, h in res])]).astype(np.uint32)
    n_rays = int(roffs[-1])
    d_rays = ffi.DeviceArray(n_rays, ffi.RAY_DTYPE)
    d_hits = ffi.DeviceArray(n_rays, ffi.HIT_DTYPE)
    eyes = []
    for k, (w, h) in enumerate(res):
        ang = rng.uniform(0, 2 * np.pi)
        eye = np.array([180 * np.cos(ang), rng.uniform(20, 80), 180 * np.sin(ang)])
        eyes.append(eye)
        cam = scenes.Camera(eye, np.zeros(3))
        p = ffi.frame_params(w, h)
        view = cam.view().astype(np.float32)
        proj = cam.proj(w, h)
        iv = np.linalg.inv(scenes.mat_cm(view)).T.reshape(16).astype(np.float32)
        ip = np.linalg.inv(scenes.mat_cm(proj)).T.reshape(16).astype(np.float32)
        p.invView[:] = iv.tolist()
        p.invProjection[:] = ip.tolist()
        p.cameraPosition[:] = [float(iv[12]), float(iv[13]), float(iv[14]), 0.0]
        p.rayLength = 10000.0
        p.primaryRayMinDist = 0.1
        ffi.call("rgExtPrimaryRays", C.byref(p), d_rays.ptr + int(roffs[k]) * 32, None)

    def timed(fn, reps=3):
        fn()
        ffi.call("rgExtSynchronize")
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ffi.call("rgExtSynchronize")
            best = min(best, time.perf_counter() - t0)
        return best

    t_batch_trace = timed(lambda: bb.trace_batch(d_rays, roffs, d_hits, 0))

    def singles_trace():
        for k in range(S):
            ffi.call("rgExtTraceRays", singles[k].h, d_rays.ptr + int(roffs[k]) * 32, int(roffs[k + 1] - roffs[k]),
                     d_hits.ptr + int(roffs[k]) * 16, 0, None)

    t_single_trace = timed(singles_trace)
    tune = bb.autotune(d_rays, n_rays, d_hits, 0, roffs)
    t_tuned_trace = timed(lambda: bb.trace_batch(d_rays, roffs, d_hits, 0))
    hits = np.frombuffer(ffi.download_raw(d_hits.ptr, n_rays * 16, np.uint32).tobytes(), np.uint32).reshape(-1, 4)
    hit_frac = float((hits[:, 3] != ffi.RG_EXT_MISS).mean())
    print(f"batch trace {t_batch_trace * 1e3:.1f} ms (tuned {t_tuned_trace * 1e3:.1f}), 64 launches {t_single_trace * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    del d_rays, d_hits, d_tris, bb, singles
    frames = group_frames(soups, res, eyes)
    return {
        "scenes": S, "triangles": int(offs[-1]), "rays": n_rays, "hit_fraction": hit_frac,
        "resolutions": [f"{w}x{h}" for w, h in res],
        "build_ms": {"batched": t_batch_build * 1e3, "separate": t_single_build * 1e3},
        "build_mtris_per_s": {"batched": offs[-1] / t_batch_build / 1e6, "separate": offs[-1] / t_single_build / 1e6},
        "trace_ms": {"batched": t_batch_trace * 1e3, "batched_tuned": t_tuned_trace * 1e3, "separate": t_single_trace * 1e3},
        "trace_mrays_per_s": {"batched": n_rays / t_batch_trace / 1e6, "batched_tuned": n_rays / t_tuned_trace / 1e6,
                              "separate": n_rays / t_single_trace / 1e6},
        "autotune": {"stack_lds_entries": tune.stackLdsEntries, "ray_chunk": tune.rayChunk},
        "frames": frames,
    }


def group_frames(soups, res, eyes, rounds: int = 3) -> dict:
    """Whole frames of the batched scenes: each soup as a plain scene (a sun,
    grey surfaces) of a render server, the scenes in one scene group, one
    session per scene at its resolution and camera.  A round submits every
    session's next frame and waits for all of them; frames/s over `rounds`
    rounds, with up to 16 frames of different scenes per launch (each traced
    stage one traversal launch over all their rays, through the group's
    batched BVH; then A-SVGF and compose per frame) and with one frame per
    launch through the same BVH."""
    import numpy as np

    from raytracedgl1_amd import scenes, serving

    srv = serving.Server(max(w for w, _ in res), max(h for _, h in res), streams=2, frames_in_flight=2, batch=16, atrous=5,
                         candidates=8, temporal=1, spatial=8)
    try:
        t = time.perf_counter()
        handles = [srv.scene(scenes.soup_scene(s)) for s in soups]
        t_scenes = time.perf_counter() - t
        t = time.perf_counter()
        g = srv.group(handles)
        t_group = time.perf_counter() - t
        sess = [srv.session(handles[k], w, h) for k, (w, h) in enumerate(res)]
        cams = [scenes.Camera(e, np.zeros(3)) for e in eyes]
        print(f"group frames: {len(handles)} scenes built in {t_scenes:.1f} s, group BVH {t_group * 1e3:.1f} ms", file=sys.stderr, flush=True)
        frame = [0]

        def run(batch):
            srv.set_policy(serving.RG_EXT_SCHEDULE_OLDEST_FIRST, batch, 2, 2, 0.0)

            def one_round():
                for k, s in enumerate(sess):
                    srv.submit(s, cams[k], frame[0] / 60.0, frame[0])
                frame[0] += 1
                srv.wait_idle()
                return srv.poll()
            one_round()  # warm
            t0 = time.perf_counter()
            done = []
            for _ in range(rounds):
                done += one_round()
            dt = time.perf_counter() - t0
            assert all(r[2] == serving.RG_EXT_FRAME_DONE for r in done), "a frame of the group run failed"
            px = sum(s.width * s.height for s in sess) * rounds
            return {"frames_per_s": len(done) / dt, "mpix_per_s": px / dt / 1e6, "mean_batch": float(np.mean([r[8] for r in done]))}
        batched = run(16)
        single = run(1)
        srv.destroy_group(g)
        print(f"group frames: batched {batched}, one per launch {single}", file=sys.stderr, flush=True)
        return {"sessions": len(sess), "rounds": rounds, "group_build_ms": t_group * 1e3, "batched_16": batched,
                "one_per_launch": single, "speedup": batched["frames_per_s"] / single["frames_per_s"]}
    finally:
        srv.close()


def bvh_build_time(scene, reps: int = 5) -> dict:
    """The LBVH build of the frame's triangles (one triangle per leaf, FP32
    nodes, as the instance builds it), best of `reps` synchronised builds
    after a warm one.  Off the timed frames: static geometry is built once."""
    import numpy as np

    from raytracedgl1_amd import ffi

    d_tris = ffi.DeviceArray.from_numpy(np.ascontiguousarray(scene.tris, np.float32))
    b = ffi.Bvh()
    b.build(d_tris, len(scene.tris), 1)
    ffi.call("rgExtSynchronize")
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        b.build(d_tris, len(scene.tris), 1)
        ffi.call("rgExtSynchronize")
        best = min(best, time.perf_counter() - t0)
    return {"triangles": len(scene.tris), "ms": best * 1e3, "mtris_per_s": len(scene.tris) / best / 1e6}


def load_profile(path):
    """A counter summary committed under profiles/ (None when absent)."""
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return None


def main(argv=None) -> int:
    a = parse_args(argv)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU (asked for --gpus {a.gpus}, WORLD_SIZE {world}): this build has "
              "no multi-GPU path; frames and sessions would shard across GPUs one process each with no "
              "exchange, so the whole-job rate is N times the one-GPU line.", file=sys.stderr)
        return 2
    if a.steps < 1:
        print("--steps must be >= 1", file=sys.stderr)
        return 2
    if a.calibrate_copy:
        return calibrate_copy(a.calibrate_copy)

    from raytracedgl1_amd import scenes

    scene = scenes.terrain_spheres_scene()
    stats, wall, final32 = run_frames(a, scene, precision=0)
    n = len(stats)
    rays = sum(st.raysTraced for st in stats)
    trav_ms = sum(st.gpuMsTraversal for st in stats)
    launches = sum(st.traversalLaunches for st in stats)
    nodes = sum(st.nodeVisits for st in stats)
    tris = sum(st.triangleTests for st in stats)
    alg_bytes = rays * BYTES_PER_RAY + nodes * BYTES_PER_NODE + tris * BYTES_PER_TRI
    achieved = alg_bytes / (trav_ms * 1e-3) / 1e9
    prof = load_profile(TRAFFIC_FILE) or {}
    traffic = prof.get("traffic_bytes_per_launch")

    def avg(field):
        return sum(getattr(st, field) for st in stats) / n

    den_ms = avg("gpuMsDenoise")
    out = {
        "metric": METRIC,
        "value": n / wall,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": max(1, a.warmup),
        "ms_per_step": wall / n * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "1920x1080 full frame, 999,168-triangle terrain + 500 icospheres, 64 sphere lights, "
                        "60-frame orbit, " + ("every primitive re-uploaded each frame" if a.dynamic_upload
                                              else "static geometry resident in HBM"),
            "width": WIDTH, "height": HEIGHT, "triangles": int(stats[0].triangleCount), "sphere_lights": 64,
            "bvh8_nodes": int(stats[0].bvhNodeCount),
            "spp": 1, "bounces": 2, "restir_di": "32 candidates / 1 temporal / 1 spatial",
            "atrous_iterations": 5, "parallelism": "single GPU",
        },
        "mrays_per_s": rays / wall / 1e6,
        "rays_per_frame": rays / n,
        "gpu_ms_per_frame": avg("gpuMsTotal"),
        "stage_ms": {k: avg(f) for k, f in (("upload", "gpuMsUpload"), ("assemble", "gpuMsAssemble"),
                                            ("bvh", "gpuMsBvh"), ("trace_and_shade", "gpuMsTrace"),
                                            ("traversal", "gpuMsTraversal"), ("denoise", "gpuMsDenoise"),
                                            ("tonemap", "gpuMsTonemap"))},
        "denoised_mpix_per_s": WIDTH * HEIGHT / (den_ms * 1e-3) / 1e6 if den_ms > 0 else None,
        "roofline": {
            "kernel": "k_trace",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes / launches,
            "avg_launch_ms": trav_ms / launches,
            "launches": launches,
            "node_visits_per_ray": nodes / rays,
            "triangle_tests_per_ray": tris / rays,
            # `achieved` counts every node fetch as 256 B although the node
            # array stays L2-resident: a work rate.  What the counters of
            # profiles/k_trace_traffic.json show for the same kernel:
            "measured": None if not prof else {
                "hbm_gbps": prof["hbm_gbps_measured"], "hbm_frac": prof["hbm_frac_measured"], "l2_hit_rate": prof["l2_hit_rate"],
                "limiter": "VALU issue",
                "valu_issue": {"achieved": prof["issue"]["valu_g_winst_per_s"], "peak": prof["issue"]["valu_peak_g_winst_per_s"],
                               "unit": "G wave64 VALU instructions/s", "frac": prof["issue"]["valu_frac"]},
                "source": "profiles/k_trace_traffic.json (rocprofv3 counters, same kernel and workload)",
            },
        },
        "denoise_kernels": None,
        "atrous_autotune": getattr(a, "atrous_tune", None),
        "bvh_build": bvh_build_time(scene),
        "cpu_baseline": None,
    }
    dprof = load_profile(DENOISE_TRAFFIC_FILE)
    if dprof:
        out["denoise_kernels"] = {
            k: {"avg_us": v["avg_us"], "launches_per_frame": v["launches_per_frame"], "hbm_frac": v["hbm_frac"],
                "valu_frac": v["valu_frac"], "l2_hit": v["l2_hit"]} for k, v in dprof["kernels"].items()}
        out["denoise_kernels"]["source"] = "profiles/denoise_1080p_traffic.json (rocprofv3 counters)"
    if not a.no_fp16:
        from raytracedgl1_amd import ffi
        import numpy as np

        st16, wall16, final16 = run_frames(a, scene, ffi.RG_EXT_PRECISION_FP16_NODES | ffi.RG_EXT_PRECISION_FP16_STORAGE)
        n16 = len(st16)
        rays16 = sum(s_.raysTraced for s_ in st16)
        g, e = final16[..., :3].astype(np.float64), final32[..., :3].astype(np.float64)
        rel = (np.abs(g - e) / (np.abs(e) + 1e-3 * np.abs(e).mean())).max(axis=-1)
        out["fp16"] = {
            "precision": "FP16 BVH nodes (144 B) + FP16 temporal/a-trous storage; FP32 arithmetic",
            "frames_per_s": n16 / wall16,
            "mrays_per_s": rays16 / wall16 / 1e6,
            "speedup_vs_fp32": (n16 / wall16) / (n / wall),
            "gpu_ms_per_frame": sum(s_.gpuMsTotal for s_ in st16) / n16,
            "traversal_ms_per_frame": sum(s_.gpuMsTraversal for s_ in st16) / n16,
            "denoise_ms_per_frame": sum(s_.gpuMsDenoise for s_ in st16) / n16,
            "node_visits_per_ray": sum(s_.nodeVisits for s_ in st16) / rays16,
            "pre_final_rel_err_vs_fp32": {"p50": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
                                          "max": float(rel.max())},
        }
    if not a.no_reflrefr:
        from raytracedgl1_amd import rtgl
        import numpy as np

        # every third icosphere a mirror, glass or water (depth 2, the default)
        kinds = (rtgl.RG_MESH_PRIMITIVE_MIRROR, rtgl.RG_MESH_PRIMITIVE_GLASS, rtgl.RG_MESH_PRIMITIVE_WATER)
        flags = {i: kinds[i % 3] for i in range(1, len(scene.objects))}
        st_r, wall_r, pl = run_frames(a, scene, 0, flags, planes=("THROUGHPUT", "IS_SKY"))
        n_r = len(st_r)
        rays_r = sum(s_.raysTraced for s_ in st_r)
        chain = np.abs(pl["THROUGHPUT"][..., 3]) == 1.0
        out["reflrefr"] = {
            "workload": "the same frames with every third icosphere a mirror, glass or water; reflection / refraction "
                        "depth 2; glass and water invisible to shadow and bounce rays",
            "frames_per_s": n_r / wall_r,
            # every chain launch spans all pixels; finished pixels pass null rays
            "ray_slots_per_frame": rays_r / n_r,
            "vs_plain_frames": (n_r / wall_r) / (n / wall),
            "gpu_ms_per_frame": sum(s_.gpuMsTotal for s_ in st_r) / n_r,
            "traversal_ms_per_frame": sum(s_.gpuMsTraversal for s_ in st_r) / n_r,
            "traversal_launches_per_frame": sum(s_.traversalLaunches for s_ in st_r) / n_r,
            "pixels_with_chain_surface": float(chain.mean()),
        }
    if not a.no_serving:
        out["serving"] = serving_run(a)
        fp32_mpix = out["value"] * WIDTH * HEIGHT / 1e6
        out["serving"]["vs_fp32_frame_mpix"] = {k: out["serving"][k]["mpix_per_s"] / fp32_mpix for k in SERVING_POLICIES}
        out["serving"]["vs_fp32_frame_mpix"]["fp32_frame_mpix_per_s"] = fp32_mpix
    if not a.no_batch:
        out["batch"] = batch_run(a)
    if not a.no_cpu_baseline:
        # all host cores of the box's share (16 per GPU)
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(scene, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
