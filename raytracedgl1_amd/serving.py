"""The render server of RTGL1_ext.h (rgExtServer*, rgExtScene*, rgExtSession*)
from Python: shared static scenes, camera sessions with their history
resident in HBM, asynchronous frame requests, and a Poisson load generator
that reports frame latency percentiles beside throughput.

The server itself (admission, dispatch over streams, completion tracking) is
C++ in the library; this module only builds requests and reads results."""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import ffi, rtgl, scenes

_P, _U32, _F = C.c_void_p, C.c_uint32, C.c_float

RG_EXT_RESULT_OVER_BUDGET = 0x52470101
RG_EXT_FRAME_DONE, RG_EXT_FRAME_DROPPED, RG_EXT_FRAME_FAILED, RG_EXT_FRAME_EXPIRED = 0, 1, 2, 3
RG_EXT_SCHEDULE_OLDEST_FIRST, RG_EXT_SCHEDULE_NEWEST_FIRST = 0, 1
RG_EXT_SERVER_READBACK = 1


class RgExtServerCreateInfo(C.Structure):
    _fields_ = [("streamCount", _U32), ("framesInFlightPerStream", _U32), ("maxRenderWidth", _U32),
                ("maxRenderHeight", _U32), ("precisionFlags", _U32), ("batchFrames", _U32), ("memoryBudgetBytes", C.c_uint64),
                ("maxQueueSeconds", C.c_double), ("schedule", _U32), ("flags", _U32)]


class RgExtFrameResult(C.Structure):
    _fields_ = [("session", _P), ("tag", C.c_uint64), ("status", _U32), ("frameId", _U32), ("submitTime", C.c_double),
                ("startTime", C.c_double), ("endTime", C.c_double), ("pImage", _P), ("batchSize", _U32), ("_pad", _U32)]


class RgExtServerPolicy(C.Structure):
    _fields_ = [("schedule", _U32), ("batchFrames", _U32), ("framesInFlightPerStream", _U32), ("streams", _U32),
                ("maxQueueSeconds", C.c_double)]


class RgExtServerStats(C.Structure):
    _fields_ = [("sessions", _U32), ("streams", _U32), ("sessionBytes", C.c_uint64), ("sharedBytes", C.c_uint64),
                ("sceneBytes", C.c_uint64), ("framesDone", C.c_uint64), ("framesDropped", C.c_uint64),
                ("framesFailed", C.c_uint64), ("now", C.c_double)]


assert C.sizeof(RgExtServerPolicy) == 24 and C.sizeof(RgExtServerCreateInfo) == 48 and C.sizeof(RgExtFrameResult) == 64 and C.sizeof(RgExtServerStats) == 64

_SIGS = {
    "rgExtServerCreate": [C.POINTER(RgExtServerCreateInfo), C.POINTER(_P)],
    "rgExtServerDestroy": [_P],
    "rgExtSceneCreate": [_P, C.POINTER(_P)],
    "rgExtSceneBuild": [_P],
    "rgExtSceneDestroy": [_P],
    "rgExtSceneGroupCreate": [_P, C.POINTER(_P), _U32, C.POINTER(_P)],
    "rgExtSceneGroupDestroy": [_P],
    "rgExtSessionCreate": [_P, _U32, _U32, C.POINTER(_P)],
    "rgExtSessionDestroy": [_P],
    "rgExtSessionSubmit": [_P, C.POINTER(rtgl.RgDrawFrameInfo), C.c_uint64],
    "rgExtServerPoll": [_P, C.POINTER(RgExtFrameResult), _U32, C.POINTER(_U32)],
    "rgExtServerWaitIdle": [_P],
    "rgExtServerGetStats": [_P, C.POINTER(RgExtServerStats)],
    "rgExtSessionGetFramebuffers": [_P, C.POINTER(_P)],
    "rgExtServerSetPolicy": [_P, C.c_void_p],
}
_bound = False


def _lib() -> C.CDLL:
    global _bound
    l = rtgl.api()
    if not _bound:
        for name, args in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = C.c_int
            fn.argtypes = args
        _bound = True
    return l


class OverBudget(RuntimeError):
    pass


@dataclass
class Session:
    handle: int
    width: int
    height: int
    scene: int


class Server:
    """rgExtServer with the sky and path parameters every request carries."""

    def __init__(self, max_width: int, max_height: int, streams: int = 4, frames_in_flight: int = 2,
                 precision: int = 0, budget_bytes: int = 0, atrous: int = 4, candidates: int = 8,
                 temporal: int = 1, spatial: int = 8, batch: int = 1, schedule: int = RG_EXT_SCHEDULE_OLDEST_FIRST,
                 max_queue_s: float = 0.0, readback: bool = False):
        l = _lib()
        ci = RgExtServerCreateInfo(streams, frames_in_flight, max_width, max_height, precision, batch, budget_bytes,
                                   max_queue_s, schedule, RG_EXT_SERVER_READBACK if readback else 0)
        h = _P()
        ffi.check("rgExtServerCreate", l.rgExtServerCreate(C.byref(ci), C.byref(h)))
        self.h = h.value
        self.path = rtgl.RgExtDrawFramePathParams()
        self.path.sType = rtgl.RG_EXT_STRUCTURE_TYPE_PATH_PARAMS
        self.path.atrousIterations = atrous
        self.path.restirInitialCandidates = candidates
        self.path.restirTemporalSamples = temporal
        self.path.restirSpatialSamples = spatial
        self.path.precisionFlags = precision
        self.sky = rtgl.RgDrawFrameSkyParams()
        self.sky.sType = rtgl.RG_STRUCTURE_TYPE_SKY
        self.sky.skyColorDefault.data[:] = (0.71, 0.88, 1.0)
        self.sky.skyColorMultiplier = 1000.0
        self.sky.skyColorSaturation = 1.0
        self.sky.pNext = C.cast(C.pointer(self.path), _P)
        self.chain = C.cast(C.pointer(self.sky), _P)
        self._results = (RgExtFrameResult * 4096)()

    def scene(self, scene: scenes.Scene) -> int:
        l = _lib()
        h = _P()
        ffi.check("rgExtSceneCreate", l.rgExtSceneCreate(self.h, C.byref(h)))
        rtgl.SceneUploader(scene).upload(h.value, True, prefix="rgExtScene")
        ffi.check("rgExtSceneBuild", l.rgExtSceneBuild(h.value))
        return h.value

    def group(self, scene_handles: list) -> int:
        """rgExtSceneGroupCreate: frames of these scenes may share launches."""
        arr = (_P * len(scene_handles))(*scene_handles)
        h = _P()
        ffi.check("rgExtSceneGroupCreate", _lib().rgExtSceneGroupCreate(self.h, arr, len(scene_handles), C.byref(h)))
        return h.value

    def destroy_group(self, group: int) -> None:
        ffi.check("rgExtSceneGroupDestroy", _lib().rgExtSceneGroupDestroy(group))

    def session(self, scene: int, width: int, height: int) -> Session:
        h = _P()
        r = _lib().rgExtSessionCreate(scene, width, height, C.byref(h))
        if r == RG_EXT_RESULT_OVER_BUDGET:
            msg = ffi.lib().rgExtGetLastErrorMessage()
            raise OverBudget(msg.decode() if msg else "over budget")
        ffi.check("rgExtSessionCreate", r)
        return Session(h.value, width, height, scene)

    def destroy_session(self, s: Session) -> None:
        ffi.check("rgExtSessionDestroy", _lib().rgExtSessionDestroy(s.handle))

    def submit(self, s: Session, cam: scenes.Camera, t: float, tag: int) -> None:
        di = rtgl.draw_info(cam, t, self.chain)
        ffi.check("rgExtSessionSubmit", _lib().rgExtSessionSubmit(s.handle, C.byref(di), tag))

    def poll(self) -> list:
        out = []
        n = _U32()
        while True:
            ffi.check("rgExtServerPoll", _lib().rgExtServerPoll(self.h, self._results, len(self._results), C.byref(n)))
            out.extend((r.session, r.tag, r.status, r.frameId, r.submitTime, r.startTime, r.endTime, r.pImage or 0, r.batchSize)
                       for r in self._results[:n.value])
            if n.value < len(self._results):
                return out

    def set_policy(self, schedule: int, batch: int, frames_in_flight: int, streams: int, max_queue_s: float = 0.0) -> None:
        p = RgExtServerPolicy(schedule, batch, frames_in_flight, streams, max_queue_s)
        ffi.check("rgExtServerSetPolicy", _lib().rgExtServerSetPolicy(self.h, C.byref(p)))

    def wait_idle(self) -> None:
        ffi.check("rgExtServerWaitIdle", _lib().rgExtServerWaitIdle(self.h))

    def stats(self) -> RgExtServerStats:
        st = RgExtServerStats()
        ffi.check("rgExtServerGetStats", _lib().rgExtServerGetStats(self.h, C.byref(st)))
        return st

    def framebuffers(self, s: Session) -> ffi.FramebufferView:
        p = _P()
        ffi.check("rgExtSessionGetFramebuffers", _lib().rgExtSessionGetFramebuffers(s.handle, C.byref(p)))
        return ffi.FramebufferView(p.value, s.width, s.height)

    def close(self) -> None:
        if self.h:
            ffi.check("rgExtServerDestroy", _lib().rgExtServerDestroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- load

SERVING_RESOLUTIONS = ((960, 540), (1280, 720), (1920, 1080), (2560, 1440))


def serving_scene(i: int, target_tris: int) -> scenes.Scene:
    """Scene i of the serving workload: a heightfield with icospheres and 64
    sphere lights, about target_tris triangles."""
    cells = 256 if target_tris < 600_000 else 512
    spheres = max(50, min(500, target_tris // 4000))
    return scenes.terrain_spheres_scene(seed=scenes.SEED_SERVING + 1000 * i, cells=cells, spheres=spheres,
                                        target_tris=target_tris)


def session_camera(k: int, t: float) -> scenes.Camera:
    """Session k's path: an orbit or a dolly over the terrain, by its index."""
    r = np.random.default_rng(scenes.SEED_SERVING + k)
    radius, height, speed, phase = r.uniform(30, 90), r.uniform(8, 30), r.uniform(0.2, 0.8), r.uniform(0, 2 * math.pi)
    if k % 2 == 0:
        a = phase + speed * t
        eye = np.array([radius * math.cos(a), height, radius * math.sin(a)])
        return scenes.Camera(eye, np.array([0.0, 2.0, 0.0]))
    d = np.array([math.cos(phase), 0.0, math.sin(phase)])
    eye = -d * radius + d * (10.0 * math.sin(speed * t)) + np.array([0.0, height, 0.0])
    return scenes.Camera(eye, eye + d * 10.0 - np.array([0.0, 3.0, 0.0]))


def run_poisson(server: Server, sessions: list, rate_hz: float, duration_s: float, seed: int = 0,
                t_offset: float = 0.0, tail_s: float = 1.5) -> dict:
    """Submits frame requests with Poisson arrivals at rate_hz per session.
    Requests submitted in [0, duration_s) are measured; arrivals go on for
    tail_s more seconds so that measured requests still meet newer ones (a
    request left waiting at the end of the load would otherwise count the
    drain as latency), then the server drains.  Latency of a finished frame
    = its completion (after the readback when the server reads frames back)
    - its request's submit; requests replaced by a newer one of their
    session, or expired in the queue, are counted apart."""
    rng = np.random.default_rng(seed)
    n = len(sessions)
    span = duration_s + tail_s
    expected = int(rate_hz * span * 1.3) + 16
    gaps = rng.exponential(1.0 / rate_hz, size=(n, expected))
    times = np.cumsum(gaps, axis=1)
    events = sorted((float(tt), k) for k in range(n) for tt in times[k] if tt < span)
    server.poll()
    t0 = time.perf_counter()
    results = []
    tag = 0
    measured_tags = 0
    i = 0
    next_poll = 0.0
    while i < len(events):
        now = time.perf_counter() - t0
        while i < len(events) and events[i][0] <= now:
            tt, k = events[i]
            server.submit(sessions[k], session_camera(k, tt + t_offset), tt + t_offset, tag)
            if tt < duration_s:
                measured_tags = tag + 1
            tag += 1
            i += 1
        if i < len(events):
            wait = events[i][0] - (time.perf_counter() - t0)
            if wait > 0.0005:
                time.sleep(min(wait, 0.002))
        if time.perf_counter() - t0 >= next_poll:
            results.extend(server.poll())
            next_poll += 0.01
    server.wait_idle()
    results.extend(server.poll())
    results = [r for r in results if r[1] < measured_tags]  # requests of the measured window
    done = [r for r in results if r[2] == RG_EXT_FRAME_DONE]
    lat = np.array([r[6] - r[4] for r in done]) * 1000.0
    first = min((r[4] for r in done), default=0.0)
    last = max((r[6] for r in done), default=0.0)
    px = {s.handle: s.width * s.height for s in sessions}
    return {
        "sessions": n,
        "rate_hz_per_session": rate_hz,
        "offered_fps": rate_hz * n,
        "requests": measured_tags,
        "frames_done": len(done),
        "frames_dropped": sum(1 for r in results if r[2] == RG_EXT_FRAME_DROPPED),
        "frames_expired": sum(1 for r in results if r[2] == RG_EXT_FRAME_EXPIRED),
        "frames_failed": sum(1 for r in results if r[2] == RG_EXT_FRAME_FAILED),
        "mean_batch": float(np.mean([r[8] for r in done])) if done else None,
        "frames_per_s": len(done) / max(last - first, 1e-9),
        "mpix_per_s": sum(px.get(r[0], 0) for r in done) / max(last - first, 1e-9) / 1e6,
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "max_ms": float(lat.max()) if len(lat) else None,
    }
