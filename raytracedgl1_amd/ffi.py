"""ctypes binding of librtgl1_amd.so (include/RTGL1/RTGL1.h and RTGL1_ext.h).

The library is loaded from the package directory; a missing or unloadable
library raises immediately - there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "librtgl1_amd.so"

# RgResult (RTGL1.h:79-95)
RG_RESULT_SUCCESS = 0
RG_RESULT_WRONG_INSTANCE = 3
RG_RESULT_ALREADY_INITIALIZED = 4
RG_RESULT_GRAPHICS_API_ERROR = 5
RG_RESULT_INTERNAL_ERROR = 6
RG_RESULT_CANT_FIND_SUPPORTED_PHYSICAL_DEVICE = 7
RG_RESULT_FRAME_WASNT_STARTED = 8
RG_RESULT_FRAME_WASNT_ENDED = 9
RG_RESULT_WRONG_FUNCTION_CALL = 10
RG_RESULT_WRONG_FUNCTION_ARGUMENT = 11

RG_EXT_MISS = 0xFFFFFFFF
RG_EXT_TRACE_CLOSEST = 0
RG_EXT_TRACE_ANY_HIT = 1
RG_EXT_TRACE_CULL_FRONT_FACE = 2
RG_EXT_TRACE_COLLECT_STATS = 4
RG_EXT_BVH_NODE_FP32 = 0
RG_EXT_BVH_NODE_FP16 = 1
RG_EXT_BVH_CHILD_EMPTY = 0xFFFFFFFF
RG_EXT_BVH_CHILD_LEAF_BIT = 0x80000000

RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4")])
NODE8_DTYPE = np.dtype([("lo_x", "<f4", 8), ("hi_x", "<f4", 8), ("lo_y", "<f4", 8), ("hi_y", "<f4", 8),
                        ("lo_z", "<f4", 8), ("hi_z", "<f4", 8), ("child", "<u4", 8), ("pad", "<u4", 8)])
NODE8H_DTYPE = np.dtype([("origin", "<f4", 3), ("pad", "<u4"), ("lo_x", "<u2", 8), ("hi_x", "<u2", 8),
                         ("lo_y", "<u2", 8), ("hi_y", "<u2", 8), ("lo_z", "<u2", 8), ("hi_z", "<u2", 8),
                         ("child", "<u4", 8)])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16 and NODE8_DTYPE.itemsize == 256
assert NODE8H_DTYPE.itemsize == 144


class RgExtTraceTuning(C.Structure):
    _fields_ = [("stackLdsEntries", C.c_uint32), ("rayChunk", C.c_uint32), ("msPerLaunch", C.c_float),
                ("msDefault", C.c_float)]


class RgExtBvhBuildParams(C.Structure):
    _fields_ = [("maxLeafSize", C.c_uint32), ("nodeFormat", C.c_uint32)]


class RgExtBvhView(C.Structure):
    _fields_ = [
        ("triangleCount", C.c_uint32),
        ("nodeCount", C.c_uint32),
        ("nodeFormat", C.c_uint32),
        ("nodeStrideBytes", C.c_uint32),
        ("d_nodes", C.c_void_p),
        ("d_sortedTriangles", C.c_void_p),
        ("d_mortonCodes", C.c_void_p),
        ("d_sortedIndices", C.c_void_p),
        ("d_binaryChildren", C.c_void_p),
        ("d_binaryBounds", C.c_void_p),
        ("sceneCentroidMin", C.c_double * 3),
        ("sceneCentroidMax", C.c_double * 3),
    ]


class RgExtFrameParams(C.Structure):
    _fields_ = [
        ("renderWidth", C.c_uint32), ("renderHeight", C.c_uint32), ("frameId", C.c_uint32), ("timeDelta", C.c_float),
        ("view", C.c_float * 16), ("invView", C.c_float * 16), ("viewPrev", C.c_float * 16),
        ("projection", C.c_float * 16), ("invProjection", C.c_float * 16), ("projectionPrev", C.c_float * 16),
        ("cameraPosition", C.c_float * 4), ("cameraPositionPrev", C.c_float * 4),
        ("minLogLuminance", C.c_float), ("maxLogLuminance", C.c_float),
        ("stopEyeAdaptation", C.c_uint32), ("antiFireflyEnabled", C.c_uint32),
        ("jitterX", C.c_float), ("jitterY", C.c_float), ("primaryRayMinDist", C.c_float), ("rayLength", C.c_float),
        ("rayCullBackFaces", C.c_uint32), ("lightCount", C.c_uint32), ("lightCountPrev", C.c_uint32),
        ("directionalLightExists", C.c_uint32),
        ("skyColorDefault", C.c_float * 4), ("skyColorMultiplier", C.c_float), ("skyColorSaturation", C.c_float),
        ("cellWorldSize", C.c_float), ("maxBounceShadowsLights", C.c_uint32),
        ("gradientMultDiffuse", C.c_float), ("gradientMultIndirect", C.c_float), ("gradientMultSpecular", C.c_float),
        ("indirSecondBounce", C.c_uint32),
        ("minRoughness", C.c_float), ("polyLightSpotlightFactor", C.c_float), ("atrousIterations", C.c_uint32),
        ("restirInitialCandidates", C.c_uint32),
        ("restirTemporalSamples", C.c_uint32), ("restirSpatialSamples", C.c_uint32), ("debugShowFlags", C.c_uint32),
        ("emissionMapBoost", C.c_float),
        ("reflectRefractMaxDepth", C.c_uint32), ("cameraMediaType", C.c_uint32),
        ("indexOfRefractionGlass", C.c_float), ("indexOfRefractionWater", C.c_float), ("thinMediaWidth", C.c_float),
        ("cameraRayConeSpreadAngle", C.c_float), ("normalMapStrength", C.c_float), ("waterWaveStrength", C.c_float), ("waterColorAndDensity", C.c_float * 4), ("acidColorAndDensity", C.c_float * 4),
        ("worldUp", C.c_float * 4), ("emissionMaxScreenColor", C.c_float),
        ("rayCullMaskWorld", C.c_uint32), ("rayCullMaskWorldShadow", C.c_uint32), ("lightIndexIgnoreFPVShadows", C.c_uint32),
        ("skyGeometry", C.c_uint32), ("pad2", C.c_uint32 * 3),
    ]


class RgError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with RgResult {code}: {msg}")
        self.code = code


_lib: C.CDLL | None = None

_P = C.c_void_p
_U32 = C.c_uint32
_F = C.c_float
_RES = C.c_int

# name -> (restype, argtypes)
_SIGS = {
    "rgExtDeviceAlloc": (_RES, [C.c_size_t, C.POINTER(_P)]),
    "rgExtDeviceFree": (_RES, [_P]),
    "rgExtCopyHostToDevice": (_RES, [_P, _P, C.c_size_t]),
    "rgExtCopyDeviceToHost": (_RES, [_P, _P, C.c_size_t]),
    "rgExtMemsetDevice": (_RES, [_P, C.c_int, C.c_size_t]),
    "rgExtCopyDeviceToDevice": (_RES, [_P, _P, C.c_size_t, _P]),
    "rgExtSynchronize": (_RES, []),
    "rgExtGetDeviceCount": (C.c_int, []),
    "rgExtGetLastErrorMessage": (C.c_char_p, []),
    "rgExtTimerCreate": (_RES, [C.POINTER(_P)]),
    "rgExtTimerDestroy": (_RES, [_P]),
    "rgExtTimerStart": (_RES, [_P, _P]),
    "rgExtTimerStop": (_RES, [_P, _P]),
    "rgExtTimerElapsedMs": (_RES, [_P, C.POINTER(_F)]),
    "rgExtBvhCreate": (_RES, [C.POINTER(_P)]),
    "rgExtBvhDestroy": (_RES, [_P]),
    "rgExtBvhBuild": (_RES, [_P, _P, _U32, C.POINTER(RgExtBvhBuildParams), _P]),
    "rgExtBvhGetView": (_RES, [_P, C.POINTER(RgExtBvhView)]),
    "rgExtTraceRays": (_RES, [_P, _P, _U32, _P, _U32, _P]),
    "rgExtTraceRaysNoAlloc": (_RES, [_P, _P, _U32, _P, _U32, _P]),
    "rgExtBvhBuildBatch": (_RES, [_P, _P, _P, _U32, C.POINTER(RgExtBvhBuildParams), _P]),
    "rgExtBvhGetSceneRoots": (_RES, [_P, _P, _U32, C.POINTER(_U32)]),
    "rgExtTraceRaysBatch": (_RES, [_P, _P, _P, _P, _U32, _P]),
    "rgExtPrimaryRays": (_RES, [C.POINTER(RgExtFrameParams), _P, _P]),
    "rgExtTraceAutotune": (_RES, [_P, _P, _U32, _P, _P, _U32, _P, C.POINTER(RgExtTraceTuning)]),
    "rgExtBvhGetTraceStats": (_RES, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rgExtBvhGetTraceErrors": (_RES, [_P, C.POINTER(C.c_uint32)]),
    "rgExtLuminanceHistogram": (_RES, [_P, _U32, _U32, _F, _F, _P, _P]),
    "rgExtLuminanceAverage": (_RES, [_P, _U32, _F, _F, _F, _U32, _P, _P]),
    "rgExtComposeFinal": (_RES, [_P, _P, _P, _P]),
    "rgExtComposeFinalEx": (_RES, [_P, _P, _P, _P, _P]),
    "rgExtFramebuffersCreate": (_RES, [_U32, _U32, C.POINTER(_P)]),
    "rgExtFramebuffersCreateWithStorage": (_RES, [_U32, _U32, _U32, C.POINTER(_P)]),
    "rgExtFramebuffersDestroy": (_RES, [_P]),
    "rgExtFramebuffersGetPlane": (_RES, [_P, C.c_int, _U32, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "rgExtFramebuffersSwap": (_RES, [_P]),
    "rgExtFrameParamsDefaults": (_RES, [_U32, _U32, C.POINTER(RgExtFrameParams)]),
    "rgExtDenoiseGradientAtrous": (_RES, [_P, C.POINTER(RgExtFrameParams), _U32, _P]),
    "rgExtDenoiseTemporal": (_RES, [_P, C.POINTER(RgExtFrameParams), _P]),
    "rgExtDenoiseAntiFirefly": (_RES, [_P, C.POINTER(RgExtFrameParams), _P]),
    "rgExtDenoiseVariance": (_RES, [_P, C.POINTER(RgExtFrameParams), _P]),
    "rgExtDenoiseAtrous": (_RES, [_P, C.POINTER(RgExtFrameParams), _U32, _P]),
    "rgExtDenoise": (_RES, [_P, C.POINTER(RgExtFrameParams), _P]),
    "rgExtFramebuffersSetAtrousTile": (_RES, [_P, _U32]),
    "rgExtDenoiseAutotune": (_RES, [_P, C.POINTER(RgExtFrameParams), _P, C.POINTER(_U32), C.POINTER(C.c_float)]),
}


def lib() -> C.CDLL:
    """Loads librtgl1_amd.so (once).  Raises if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m raytracedgl1_amd.build`")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(fn: str, code: int) -> None:
    if code != RG_RESULT_SUCCESS:
        msg = lib().rgExtGetLastErrorMessage()
        raise RgError(fn, code, msg.decode() if msg else "")


def call(fn: str, *args) -> None:
    check(fn, getattr(lib(), fn)(*args))


def device_count() -> int:
    return int(lib().rgExtGetDeviceCount())


class DeviceArray:
    """A device allocation holding a numpy-described array (shape + dtype)."""

    def __init__(self, shape, dtype, zero: bool = False):
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = _P()
        call("rgExtDeviceAlloc", max(self.nbytes, 1), C.byref(p))
        self.ptr = p.value
        if zero:
            self.zero()

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        call("rgExtCopyHostToDevice", self.ptr, a.ctypes.data, self.nbytes)

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        call("rgExtCopyDeviceToHost", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def zero(self) -> None:
        call("rgExtMemsetDevice", self.ptr, 0, self.nbytes)

    def free(self) -> None:
        if self.ptr:
            lib().rgExtDeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_raw(ptr: int, nbytes: int, dtype, shape=None) -> np.ndarray:
    dt = np.dtype(dtype)
    out = np.empty(nbytes // dt.itemsize, dt)
    call("rgExtCopyDeviceToHost", out.ctypes.data, ptr, nbytes)
    return out if shape is None else out.reshape(shape)


class Timer:
    """HIP-event timer on a stream (null = default stream)."""

    def __init__(self):
        p = _P()
        call("rgExtTimerCreate", C.byref(p))
        self.h = p.value

    def start(self, stream=None):
        call("rgExtTimerStart", self.h, stream)

    def stop(self, stream=None):
        call("rgExtTimerStop", self.h, stream)

    def elapsed_ms(self) -> float:
        v = _F()
        call("rgExtTimerElapsedMs", self.h, C.byref(v))
        return float(v.value)

    def __del__(self):
        try:
            lib().rgExtTimerDestroy(self.h)
        except Exception:
            pass


class Bvh:
    """An LBVH on the device (rgExtBvh*)."""

    def __init__(self):
        p = _P()
        call("rgExtBvhCreate", C.byref(p))
        self.h = p.value

    def build(self, d_tris: DeviceArray, n: int, leaf_size: int = 3, node_format: int = RG_EXT_BVH_NODE_FP32,
              stream=None) -> None:
        params = RgExtBvhBuildParams(leaf_size, node_format)
        call("rgExtBvhBuild", self.h, d_tris.ptr, n, C.byref(params), stream)

    def view(self) -> RgExtBvhView:
        v = RgExtBvhView()
        call("rgExtBvhGetView", self.h, C.byref(v))
        return v

    def trace(self, d_rays: DeviceArray, count: int, d_hits: DeviceArray, flags: int = 0, stream=None) -> None:
        call("rgExtTraceRays", self.h, d_rays.ptr, count, d_hits.ptr, flags, stream)

    def build_batch(self, d_tris: DeviceArray, offsets, leaf_size: int = 1, node_format: int = RG_EXT_BVH_NODE_FP32,
                    stream=None) -> None:
        """Scenes concatenated in d_tris; scene s is triangles [offsets[s], offsets[s+1])."""
        offs = np.ascontiguousarray(offsets, np.uint32)
        params = RgExtBvhBuildParams(leaf_size, node_format)
        call("rgExtBvhBuildBatch", self.h, d_tris.ptr, offs.ctypes.data, len(offs) - 1, C.byref(params), stream)

    def scene_roots(self) -> np.ndarray:
        n = C.c_uint32()
        call("rgExtBvhGetSceneRoots", self.h, None, 0, C.byref(n))
        out = np.zeros(max(1, n.value), np.uint32)
        call("rgExtBvhGetSceneRoots", self.h, out.ctypes.data, n.value, C.byref(n))
        return out[:n.value]

    def trace_batch(self, d_rays: DeviceArray, ray_offsets, d_hits: DeviceArray, flags: int = 0, stream=None) -> None:
        offs = np.ascontiguousarray(ray_offsets, np.uint32)
        call("rgExtTraceRaysBatch", self.h, d_rays.ptr, offs.ctypes.data, d_hits.ptr, flags, stream)

    def autotune(self, d_rays: DeviceArray, count: int, d_hits: DeviceArray, flags: int = 0, ray_offsets=None,
                 stream=None) -> RgExtTraceTuning:
        t = RgExtTraceTuning()
        offs = None if ray_offsets is None else np.ascontiguousarray(ray_offsets, np.uint32)
        call("rgExtTraceAutotune", self.h, d_rays.ptr, count, None if offs is None else offs.ctypes.data, d_hits.ptr, flags,
             stream, C.byref(t))
        return t

    def trace_stats(self) -> tuple:
        """(node visits, triangle tests) of the last trace with RG_EXT_TRACE_COLLECT_STATS."""
        nodes, tris = C.c_uint64(), C.c_uint64()
        call("rgExtBvhGetTraceStats", self.h, C.byref(nodes), C.byref(tris))
        return int(nodes.value), int(tris.value)

    def trace_errors(self) -> int:
        """RG_EXT_TRACE_ERROR_* bits of the last trace (0: every ray complete)."""
        e = C.c_uint32()
        call("rgExtBvhGetTraceErrors", self.h, C.byref(e))
        return int(e.value)

    def download(self) -> dict:
        """Copies every build output to host numpy arrays."""
        v = self.view()
        n, m = v.triangleCount, v.nodeCount
        out = {
            "n": n,
            "node_count": m,
            "nodes": download_raw(v.d_nodes, m * v.nodeStrideBytes,
                                  NODE8H_DTYPE if v.nodeFormat == RG_EXT_BVH_NODE_FP16 else NODE8_DTYPE),
            "sorted_tris": download_raw(v.d_sortedTriangles, n * 48, np.float32, (n, 12)),
            "codes": download_raw(v.d_mortonCodes, n * 8, np.uint64),
            "sorted_idx": download_raw(v.d_sortedIndices, n * 4, np.uint32),
            "bounds": download_raw(v.d_binaryBounds, (2 * n - 1) * 24, np.float32, (2 * n - 1, 6)),
            "scene_min": np.array(v.sceneCentroidMin[:]),
            "scene_max": np.array(v.sceneCentroidMax[:]),
        }
        if n > 1:
            out["children"] = download_raw(v.d_binaryChildren, (n - 1) * 8, np.uint32, (n - 1, 2))
        else:
            out["children"] = np.zeros((0, 2), np.uint32)
        return out

    def __del__(self):
        try:
            lib().rgExtBvhDestroy(self.h)
        except Exception:
            pass


# ------------------------------------------------------------ framebuffers

# RgExtPlane: (name, layout, components, dtype, has_prev); layout 0 =
# checkerboard W x H, 1 = regular W x H, 2 = strata max(1,(W+1)/3) x max(1,(H+1)/3)
PLANES = [
    ("NORMAL", 0, 1, np.uint32, True),
    ("METALLIC_ROUGHNESS", 0, 2, np.float32, True),
    ("DEPTH_WORLD", 0, 1, np.float32, True),
    ("DEPTH_GRAD", 0, 1, np.float32, False),
    ("MOTION", 0, 4, np.float32, False),
    ("UNFILTERED_DIRECT", 0, 4, np.float32, False),
    ("UNFILTERED_SPECULAR", 0, 4, np.float32, False),
    ("UNFILTERED_INDIR", 0, 4, np.float32, False),
    ("SURFACE_POSITION", 0, 4, np.float32, True),
    ("VISIBILITY_BUFFER", 0, 4, np.uint32, True),
    ("VIEW_DIRECTION", 0, 4, np.float32, False),
    ("THROUGHPUT", 0, 4, np.float32, False),
    ("IS_SKY", 0, 1, np.uint32, False),
    ("ACCUM_HISTORY_LENGTH", 0, 4, np.float32, True),
    ("DIFF_TEMPORARY", 0, 4, np.float32, False),
    ("DIFF_ACCUM_COLOR", 0, 4, np.float32, False),
    ("DIFF_ACCUM_MOMENTS", 0, 2, np.float32, True),
    ("DIFF_COLOR_HISTORY", 0, 4, np.float32, False),
    ("DIFF_PING", 0, 4, np.float32, False),
    ("DIFF_PONG", 0, 4, np.float32, False),
    ("SPEC_ACCUM_COLOR", 0, 4, np.float32, True),
    ("SPEC_PING", 0, 4, np.float32, False),
    ("SPEC_PONG", 0, 4, np.float32, False),
    ("INDIR_ACCUM", 0, 4, np.float32, True),
    ("INDIR_PING", 0, 4, np.float32, False),
    ("INDIR_PONG", 0, 4, np.float32, False),
    ("ATROUS_FILTERED_VARIANCE", 0, 1, np.float32, False),
    ("RESERVOIRS", 0, 4, np.uint32, True),
    ("RESERVOIRS_INITIAL", 0, 4, np.uint32, False),
    ("GRADIENT_INPUTS", 0, 2, np.float32, True),
    ("ALBEDO", 1, 4, np.float32, False),
    ("HISTOGRAM_INPUT", 1, 1, np.float32, False),
    ("PRE_FINAL", 1, 4, np.float32, False),
    ("DIS_PING_GRADIENT", 2, 4, np.float32, False),
    ("DIS_PONG_GRADIENT", 2, 4, np.float32, False),
    ("DIS_GRADIENT_HISTORY", 2, 4, np.float32, False),
    ("GRADIENT_PREV_PIX", 2, 1, np.uint32, False),
    ("INDIR_INITIAL_SAMPLE", 0, 4, np.uint32, False),
    ("INDIR_INITIAL_RADIANCE", 0, 2, np.uint32, False),
    ("INDIR_RESERVOIR_SAMPLE", 0, 4, np.uint32, True),
    ("INDIR_RESERVOIR_WEIGHT", 0, 4, np.uint32, True),
    ("FINAL", 1, 4, np.float32, False),
    ("SCREEN_EMISSION", 1, 4, np.float32, False),
    ("ACID_FOG", 1, 4, np.float32, False),
    ("DENOISE_GUIDE", 0, 4, np.float32, False),
    ("GI_REUSE", 0, 16, np.uint32, False),
]
PLANE = {name: i for i, (name, *_rest) in enumerate(PLANES)}


RG_EXT_STORAGE_FP32, RG_EXT_STORAGE_FP16 = 0, 1
RG_EXT_PRECISION_FP16_STORAGE, RG_EXT_PRECISION_FP16_NODES = 1, 2


def is_half_plane(plane: int) -> bool:
    """Planes stored as IEEE half by RG_EXT_STORAGE_FP16."""
    return PLANE["ACCUM_HISTORY_LENGTH"] <= plane <= PLANE["INDIR_PONG"] or plane == PLANE["ATROUS_FILTERED_VARIANCE"]


def plane_shape(plane: int, width: int, height: int) -> tuple:
    _, layout, comps, _, _ = PLANES[plane]
    if layout == 2:
        w, h = max(1, (width + 1) // 3), max(1, (height + 1) // 3)
    else:
        w, h = width, height
    return (h, w, comps)


def frame_params(width: int, height: int) -> RgExtFrameParams:
    p = RgExtFrameParams()
    call("rgExtFrameParamsDefaults", width, height, C.byref(p))
    return p


class FramebufferSet:
    """rgExtFramebuffers*: the device planes of one render size."""

    def __init__(self, width: int, height: int, storage: int = 0):
        p = _P()
        call("rgExtFramebuffersCreateWithStorage", width, height, storage, C.byref(p))
        self.h = p.value
        self.width, self.height = width, height

    def dtype(self, plane: int, nbytes: int):
        """Element type of a plane: float16 for the planes the FP16 storage
        mode halves (their byte size tells), else the FP32-mode type."""
        full = int(np.prod(plane_shape(plane, self.width, self.height))) * 4
        if is_half_plane(plane) and nbytes * 2 == full:
            return np.float16
        return PLANES[plane][3]

    def plane_ptr(self, plane: int, prev: bool = False) -> tuple:
        d = _P()
        nbytes = C.c_size_t()
        call("rgExtFramebuffersGetPlane", self.h, plane, int(prev), C.byref(d), C.byref(nbytes))
        return d.value, nbytes.value

    def upload(self, plane: int, arr: np.ndarray, prev: bool = False) -> None:
        ptr, nbytes = self.plane_ptr(plane, prev)
        a = np.ascontiguousarray(arr, dtype=self.dtype(plane, nbytes))
        assert a.nbytes == nbytes, (PLANES[plane][0], a.nbytes, nbytes)
        call("rgExtCopyHostToDevice", ptr, a.ctypes.data, nbytes)

    def download(self, plane: int, prev: bool = False) -> np.ndarray:
        ptr, nbytes = self.plane_ptr(plane, prev)
        shape = plane_shape(plane, self.width, self.height)
        return download_raw(ptr, nbytes, self.dtype(plane, nbytes), shape)

    def swap(self) -> None:
        call("rgExtFramebuffersSwap", self.h)

    def __del__(self):
        try:
            lib().rgExtFramebuffersDestroy(self.h)
        except Exception:
            pass


class FramebufferView(FramebufferSet):
    """Non-owning view of an instance's framebuffer set (rgExtGetFramebuffers)."""

    def __init__(self, handle: int, width: int, height: int):
        self.h = handle
        self.width, self.height = width, height

    def __del__(self):
        pass
