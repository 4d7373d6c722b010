"""The RTGL1 C API (include/RTGL1/RTGL1.h) through ctypes, and a small
driver that feeds a synthetic scene to it the way an application does
(rgStartFrame -> rgUpload* -> rgDrawFrame, Tests/RtglTest.cpp:439-703).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import ffi
from .ffi import RgExtFrameParams, call, lib

_P = C.c_void_p
_U32 = C.c_uint32
_F = C.c_float


class RgFloat3D(C.Structure):
    _fields_ = [("data", C.c_float * 3)]


class RgInstanceCreateInfo(C.Structure):
    _fields_ = [
        ("pAppName", C.c_char_p), ("pAppGUID", C.c_char_p),
        ("pWin32SurfaceInfo", _P), ("pMetalSurfaceCreateInfo", _P), ("pWaylandSurfaceCreateInfo", _P),
        ("pXcbSurfaceCreateInfo", _P), ("pXlibSurfaceCreateInfo", _P),
        ("pConfigPath", C.c_char_p), ("pOverrideFolderPath", C.c_char_p),
        ("pfnPrint", _P), ("pUserPrintData", _P), ("allowedMessages", _U32),
        ("primaryRaysMaxAlbedoLayers", _U32), ("indirectIlluminationMaxAlbedoLayers", _U32),
        ("rayCullBackFacingTriangles", _U32), ("allowGeometryWithSkyFlag", _U32),
        ("allowTexCoordLayer1", _U32), ("allowTexCoordLayer2", _U32), ("allowTexCoordLayer3", _U32),
        ("lightmapTexCoordLayerIndex", _U32),
        ("rasterizedMaxVertexCount", _U32), ("rasterizedMaxIndexCount", _U32), ("rasterizedVertexColorGamma", _U32),
        ("rasterizedSkyCubemapSize", _U32),
        ("textureSamplerForceMinificationFilterLinear", _U32), ("textureSamplerForceNormalMapFilterLinear", _U32),
        ("pbrTextureSwizzling", C.c_int), ("effectWipeIsUsed", _U32),
        ("worldUp", RgFloat3D), ("worldForward", RgFloat3D), ("worldScale", _F),
    ]


class RgTransform(C.Structure):
    _fields_ = [("matrix", (C.c_float * 4) * 3)]


class RgMeshInfo(C.Structure):
    _fields_ = [("uniqueObjectID", _U32), ("pMeshName", C.c_char_p), ("transform", RgTransform),
                ("isExportable", _U32), ("animationName", C.c_char_p), ("animationTime", _F)]


class RgMeshPrimitiveInfo(C.Structure):
    _fields_ = [("pPrimitiveNameInMesh", C.c_char_p), ("primitiveIndexInMesh", _U32), ("flags", _U32),
                ("pVertices", _P), ("vertexCount", _U32), ("pIndices", _P), ("indexCount", _U32),
                ("pTextureName", C.c_char_p), ("textureFrame", _U32), ("color", _U32), ("emissive", _F),
                ("pEditorInfo", _P)]


class RgLightExtraInfo(C.Structure):
    _fields_ = [("exists", _U32), ("lightstyle", C.c_int), ("isVolumetric", _U32)]


class RgDirectionalLightUploadInfo(C.Structure):
    _fields_ = [("uniqueID", C.c_uint64), ("isExportable", _U32), ("extra", RgLightExtraInfo), ("color", _U32),
                ("intensity", _F), ("direction", RgFloat3D), ("angularDiameterDegrees", _F)]


class RgSphericalLightUploadInfo(C.Structure):
    _fields_ = [("uniqueID", C.c_uint64), ("isExportable", _U32), ("extra", RgLightExtraInfo), ("color", _U32),
                ("intensity", _F), ("position", RgFloat3D), ("radius", _F)]


class RgSpotLightUploadInfo(C.Structure):
    _fields_ = [("uniqueID", C.c_uint64), ("isExportable", _U32), ("extra", RgLightExtraInfo), ("color", _U32),
                ("intensity", _F), ("position", RgFloat3D), ("direction", RgFloat3D), ("radius", _F), ("angleOuter", _F),
                ("angleInner", _F)]


class RgPolygonalLightUploadInfo(C.Structure):
    _fields_ = [("uniqueID", C.c_uint64), ("isExportable", _U32), ("extra", RgLightExtraInfo), ("color", _U32),
                ("intensity", _F), ("positions", RgFloat3D * 3)]


class RgStartFrameInfo(C.Structure):
    _fields_ = [("pMapName", C.c_char_p), ("ignoreExternalGeometry", _U32)]


class RgDrawFrameSkyParams(C.Structure):
    _fields_ = [("sType", C.c_int), ("pNext", _P), ("skyType", C.c_int), ("skyColorDefault", RgFloat3D),
                ("skyColorMultiplier", _F), ("skyColorSaturation", _F), ("skyViewerPosition", RgFloat3D),
                ("pSkyCubemapTextureName", C.c_char_p), ("skyCubemapRotationTransform", C.c_float * 9)]


class RgDrawFrameTonemappingParams(C.Structure):
    _fields_ = [("sType", C.c_int), ("pNext", _P), ("disableEyeAdaptation", _U32), ("ev100Min", _F), ("ev100Max", _F),
                ("luminanceWhitePoint", _F), ("saturation", RgFloat3D), ("crosstalk", RgFloat3D)]


class RgDrawFrameIlluminationParams(C.Structure):
    _fields_ = [("sType", C.c_int), ("pNext", _P), ("maxBounceShadows", _U32), ("enableSecondBounceForIndirect", _U32),
                ("cellWorldSize", _F), ("directDiffuseSensitivityToChange", _F), ("indirectDiffuseSensitivityToChange", _F),
                ("specularSensitivityToChange", _F), ("polygonalLightSpotlightFactor", _F),
                ("lightUniqueIdIgnoreFirstPersonViewerShadows", _P), ("lightstyleValuesCount", _U32),
                ("pLightstyleValues", _P)]


class RgExtent2D(C.Structure):
    _fields_ = [("width", _U32), ("height", _U32)]


class RgOriginalTextureInfo(C.Structure):
    _fields_ = [("pTextureName", C.c_char_p), ("pPixels", _P), ("size", RgExtent2D), ("filter", C.c_int),
                ("addressModeU", C.c_int), ("addressModeV", C.c_int)]


# RgSamplerFilter / RgSamplerAddressMode
RG_SAMPLER_FILTER_AUTO, RG_SAMPLER_FILTER_LINEAR, RG_SAMPLER_FILTER_NEAREST = 0, 1, 2
RG_SAMPLER_ADDRESS_MODE_REPEAT, RG_SAMPLER_ADDRESS_MODE_CLAMP = 0, 1
RG_MESH_PRIMITIVE_ALPHA_TESTED = 1


class RgOriginalCubemapInfo(C.Structure):
    _fields_ = [("pTextureName", C.c_char_p), ("pPixelsPositiveX", _P), ("pPixelsNegativeX", _P),
                ("pPixelsPositiveY", _P), ("pPixelsNegativeY", _P), ("pPixelsPositiveZ", _P),
                ("pPixelsNegativeZ", _P), ("sideSize", _U32)]


class RgDrawFrameRenderResolutionParams(C.Structure):
    _fields_ = [("sType", C.c_int), ("pNext", _P), ("upscaleTechnique", C.c_int), ("sharpenTechnique", C.c_int),
                ("resolutionMode", C.c_int), ("customRenderSize", RgExtent2D), ("pPixelizedRenderSize", _P),
                ("resetUpscalerHistory", _U32)]


RG_STRUCTURE_TYPE_RENDER_RESOLUTION = 1
RG_RENDER_UPSCALE_TECHNIQUE_LINEAR = 0
RG_RENDER_RESOLUTION_MODE_CUSTOM = 0
RG_SKY_TYPE_COLOR = 0


class RgDrawFrameTexturesParams(C.Structure):
    _fields_ = [("sType", C.c_int), ("pNext", _P), ("dynamicSamplerFilter", C.c_int), ("normalMapStrength", _F),
                ("emissionMapBoost", _F), ("emissionMaxScreenColor", _F), ("minRoughness", _F)]


RG_STRUCTURE_TYPE_TEXTURES = 8


class RgDrawFrameReflectRefractParams(C.Structure):
    _fields_ = [("sType", C.c_int), ("pNext", _P), ("maxReflectRefractDepth", _U32), ("typeOfMediaAroundCamera", C.c_int),
                ("indexOfRefractionGlass", _F), ("indexOfRefractionWater", _F), ("thinMediaWidth", _F),
                ("waterWaveSpeed", _F), ("waterWaveNormalStrength", _F), ("waterColor", RgFloat3D), ("acidColor", RgFloat3D),
                ("acidDensity", _F), ("waterWaveTextureDerivativesMultiplier", _F), ("waterTextureAreaScale", _F),
                ("portalNormalTwirl", _U32)]


def reflect_refract_params(**kw) -> RgDrawFrameReflectRefractParams:
    """RgDrawFrameReflectRefractParams with the reference defaults
    (Source/DrawFrameInfo.h:125-146), overridden by keyword."""
    r = RgDrawFrameReflectRefractParams()
    r.sType = RG_STRUCTURE_TYPE_REFLECTREFRACT
    r.maxReflectRefractDepth = 2
    r.typeOfMediaAroundCamera = RG_MEDIA_TYPE_VACUUM
    r.indexOfRefractionGlass = 1.52
    r.indexOfRefractionWater = 1.33
    r.thinMediaWidth = 0.1
    r.waterWaveSpeed = 1.0
    r.waterWaveNormalStrength = 1.0
    r.waterColor.data[:] = (0.3, 0.73, 0.63)
    r.acidColor.data[:] = (0.0, 0.66, 0.55)
    r.acidDensity = 10.0
    r.waterWaveTextureDerivativesMultiplier = 1.0
    r.waterTextureAreaScale = 1.0
    for k, v in kw.items():
        if k in ("waterColor", "acidColor"):
            getattr(r, k).data[:] = v
        else:
            setattr(r, k, v)
    return r


class RgExtDrawFramePathParams(C.Structure):
    _fields_ = [("sType", C.c_int), ("pNext", _P), ("atrousIterations", _U32), ("restirInitialCandidates", _U32),
                ("restirTemporalSamples", _U32), ("restirSpatialSamples", _U32),
                ("renderWidth", _U32), ("renderHeight", _U32), ("collectTraversalStats", _U32), ("precisionFlags", _U32)]


class RgDrawFrameInfo(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("fovYRadians", _F), ("cameraNear", _F), ("cameraFar", _F),
                ("rayLength", _F), ("rayCullMaskWorld", _U32), ("disableRayTracedGeometry", _U32),
                ("disableRasterization", _U32), ("presentPrevFrame", _U32), ("currentTime", C.c_double),
                ("vsync", _U32), ("pParams", _P)]


class RgExtFrameStats(C.Structure):
    _fields_ = [("frameId", _U32), ("renderWidth", _U32), ("renderHeight", _U32), ("triangleCount", _U32),
                ("bvhRebuilt", _U32), ("bvhNodeCount", _U32), ("raysTraced", C.c_uint64),
                ("gpuMsTotal", _F), ("gpuMsAssemble", _F), ("gpuMsBvh", _F), ("gpuMsTrace", _F),
                ("gpuMsDenoise", _F), ("gpuMsTonemap", _F), ("avgLuminance", _F), ("gpuMsUpload", _F),
                ("gpuMsTraversal", _F), ("traversalLaunches", _U32), ("traversalErrors", _U32), ("nodeVisits", C.c_uint64),
                ("triangleTests", C.c_uint64)]


assert C.sizeof(RgInstanceCreateInfo) == 184 and C.sizeof(RgMeshInfo) == 88
assert C.sizeof(RgMeshPrimitiveInfo) == 80 and C.sizeof(RgDrawFrameInfo) == 120
assert C.sizeof(RgDirectionalLightUploadInfo) == 48 and C.sizeof(RgSphericalLightUploadInfo) == 48
assert C.sizeof(RgDrawFrameTonemappingParams) == 56 and C.sizeof(RgDrawFrameIlluminationParams) == 72
assert C.sizeof(RgExtDrawFramePathParams) == 48 and C.sizeof(RgExtFrameStats) == 96

RG_STRUCTURE_TYPE_ILLUMINATION = 2
RG_STRUCTURE_TYPE_TONEMAPPING = 4
RG_STRUCTURE_TYPE_SKY = 7
RG_STRUCTURE_TYPE_REFLECTREFRACT = 6
# RgMeshPrimitiveFlags (RTGL1.h) used by the reflection / refraction chain
RG_MESH_PRIMITIVE_FIRST_PERSON = 4
RG_MESH_PRIMITIVE_FIRST_PERSON_VIEWER = 8
RG_MESH_PRIMITIVE_MIRROR = 32
RG_MESH_PRIMITIVE_GLASS = 64
RG_MESH_PRIMITIVE_WATER = 128
RG_MESH_PRIMITIVE_FORCE_EXACT_NORMALS = 512
RG_MESH_PRIMITIVE_MIRROR_IF_SMOOTH = 1024
RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH = 2048
RG_MESH_PRIMITIVE_ACID = 8192
RG_MESH_PRIMITIVE_THIN_MEDIA = 16384
RG_MESH_PRIMITIVE_DONT_GENERATE_NORMALS = 256
RG_MESH_PRIMITIVE_IGNORE_REFRACT_AFTER = 4096
RG_MESH_PRIMITIVE_SKY_VISIBILITY = 32768
# RgMediaType
RG_MEDIA_TYPE_VACUUM, RG_MEDIA_TYPE_WATER, RG_MEDIA_TYPE_GLASS, RG_MEDIA_TYPE_ACID = 0, 1, 2, 3
RG_EXT_STRUCTURE_TYPE_PATH_PARAMS = 0x52470001
RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT = 1
RG_DRAW_FRAME_RAY_CULL_WORLD_1_BIT = 2
RG_DRAW_FRAME_RAY_CULL_WORLD_2_BIT = 4
RG_DRAW_FRAME_RAY_CULL_SKY_BIT = 8

VERTEX_DTYPE = np.dtype([("position", "<f4", 3), ("_p0", "<u4"), ("normal", "<f4", 3), ("_p1", "<u4"),
                         ("tangent", "<f4", 4), ("texCoord", "<f4", 2), ("color", "<u4"), ("_p2", "<u4")])
assert VERTEX_DTYPE.itemsize == 64

_API = {
    "rgProvideOriginalTexture": (C.c_int, [_P, C.POINTER(RgOriginalTextureInfo)]),
    "rgMarkOriginalTextureAsDeleted": (C.c_int, [_P, C.c_char_p]),
    "rgProvideOriginalCubemapTexture": (C.c_int, [_P, C.POINTER(RgOriginalCubemapInfo)]),
    "rgCreateInstance": (C.c_int, [C.POINTER(RgInstanceCreateInfo), C.POINTER(_P)]),
    "rgDestroyInstance": (C.c_int, [_P]),
    "rgUploadMeshPrimitive": (C.c_int, [_P, C.POINTER(RgMeshInfo), C.POINTER(RgMeshPrimitiveInfo)]),
    "rgUploadDirectionalLight": (C.c_int, [_P, C.POINTER(RgDirectionalLightUploadInfo)]),
    "rgUploadSphericalLight": (C.c_int, [_P, C.POINTER(RgSphericalLightUploadInfo)]),
    "rgUploadSpotLight": (C.c_int, [_P, C.POINTER(RgSpotLightUploadInfo)]),
    "rgUploadPolygonalLight": (C.c_int, [_P, C.POINTER(RgPolygonalLightUploadInfo)]),
    "rgStartFrame": (C.c_int, [_P, C.POINTER(RgStartFrameInfo)]),
    "rgDrawFrame": (C.c_int, [_P, C.POINTER(RgDrawFrameInfo)]),
    "rgUtilGetResultDescription": (C.c_char_p, [C.c_int]),
    "rgUtilPackColorByte4D": (_U32, [C.c_uint8] * 4),
    "rgUtilPackColorFloat4D": (_U32, [_F] * 4),
    "rgExtGetFrameStats": (C.c_int, [_P, C.POINTER(RgExtFrameStats)]),
    "rgExtGetFramebuffers": (C.c_int, [_P, C.POINTER(_P)]),
    "rgExtGetFrameParams": (C.c_int, [_P, C.POINTER(RgExtFrameParams)]),
    "rgExtReadbackHistogram": (C.c_int, [_P, _P]),
    "rgExtGetStream": (C.c_int, [_P, C.POINTER(_P)]),
    "rgExtSetStaticGeometry": (C.c_int, [_P, _U32]),
    "rgExtSceneUploadMeshPrimitive": (C.c_int, [_P, C.POINTER(RgMeshInfo), C.POINTER(RgMeshPrimitiveInfo)]),
    "rgExtSceneUploadDirectionalLight": (C.c_int, [_P, C.POINTER(RgDirectionalLightUploadInfo)]),
    "rgExtSceneUploadSphericalLight": (C.c_int, [_P, C.POINTER(RgSphericalLightUploadInfo)]),
}
_bound = False


def api() -> C.CDLL:
    global _bound
    l = lib()
    if not _bound:
        for name, (res, args) in _API.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return l


def identity_transform() -> RgTransform:
    t = RgTransform()
    for r in range(3):
        for c in range(4):
            t.matrix[r][c] = 1.0 if r == c else 0.0
    return t


def _style(info, light: dict) -> None:
    """A light dict's optional "lightstyle" index -> RgLightExtraInfo."""
    if light.get("lightstyle") is not None:
        info.extra.exists = 1
        info.extra.lightstyle = int(light["lightstyle"])


def draw_info(cam, current_time: float, params_ptr, cull_mask: int = RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT) -> RgDrawFrameInfo:
    """RgDrawFrameInfo for a camera (ray length 10000; by default only
    world 0 is traced, as Tests/RtglTest.cpp:689 does)."""
    di = RgDrawFrameInfo()
    di.view[:] = cam.view().tolist()
    di.fovYRadians = cam.fovy
    di.cameraNear = cam.znear
    di.cameraFar = cam.zfar
    di.rayLength = 10000.0
    di.rayCullMaskWorld = cull_mask
    di.currentTime = current_time
    di.pParams = params_ptr
    return di


class SceneUploader:
    """Prebuilt per-object vertex arrays of a scenes.Scene, uploaded each frame
    with rgUploadMeshPrimitive (non-indexed triangle lists, identity transform)."""

    def __init__(self, scene, flags: dict | None = None, textures: dict | None = None, texcoords: dict | None = None,
                 emissive: dict | None = None):
        """flags: object index -> RgMeshPrimitiveFlags; textures: object
        index -> pTextureName; texcoords: object index -> (tris * 3, 2)
        float32 vertex texture coordinates."""
        self.scene = scene
        self.flags = dict(flags or {})  # object index -> RgMeshPrimitiveFlags
        self.textures = {k: v.encode() for k, v in (textures or {}).items()}
        self.emissive = dict(emissive or {})  # object index -> RgMeshPrimitiveInfo::emissive
        self.objects = []
        for oid, (name, first, count) in enumerate(scene.objects):
            v = np.zeros(count * 3, VERTEX_DTYPE)
            v["position"] = scene.tris[first:first + count].reshape(-1, 3)
            v["normal"] = scene.normals[first:first + count].reshape(-1, 3)
            if texcoords and oid in texcoords:
                v["texCoord"] = np.asarray(texcoords[oid], np.float32).reshape(-1, 2)
            color = int(scene.colors[first])
            self.objects.append((oid, v, color))
        self.mesh = RgMeshInfo()
        self.mesh.transform = identity_transform()
        self.prim = RgMeshPrimitiveInfo()

    def upload(self, inst, geometry: bool = True, prefix: str = "rg", transforms: dict | None = None,
               order=None) -> None:
        """Uploads the primitives (unless geometry is False) and the lights
        to an instance (prefix "rg": rgUploadMeshPrimitive, ...) or to a
        server scene (prefix "rgExtScene": rgExtSceneUploadMeshPrimitive, ...).
        transforms maps an object index to its 3x4 row-major model matrix
        (identity otherwise); order lists the object indices in upload order."""
        l = api()
        mesh_fn = getattr(l, prefix + "UploadMeshPrimitive")
        dir_fn = getattr(l, prefix + "UploadDirectionalLight")
        sph_fn = getattr(l, prefix + "UploadSphericalLight")
        objs = [self.objects[i] for i in order] if order is not None else self.objects
        for oid, v, color in (objs if geometry else []):
            self.mesh.uniqueObjectID = oid
            m = (transforms or {}).get(oid)
            if m is None:
                self.mesh.transform = identity_transform()
            else:
                for r_ in range(3):
                    for c_ in range(4):
                        self.mesh.transform.matrix[r_][c_] = float(m[r_][c_])
            self.prim.primitiveIndexInMesh = 0
            self.prim.pVertices = v.ctypes.data
            self.prim.vertexCount = len(v)
            self.prim.pIndices = None
            self.prim.indexCount = 0
            self.prim.color = color | 0xFF000000
            self.prim.emissive = float(self.emissive.get(oid, 0.0))
            self.prim.flags = self.flags.get(oid, 0)
            self.prim.pTextureName = self.textures.get(oid)
            ffi.check(prefix + "UploadMeshPrimitive", mesh_fn(inst, C.byref(self.mesh), C.byref(self.prim)))
        lights = self.scene.lights or {}
        sun = lights.get("sun")
        if sun:
            d = RgDirectionalLightUploadInfo()
            d.uniqueID = 1
            d.color = l.rgUtilPackColorFloat4D(*sun["color"], 1.0)
            d.intensity = sun["intensity"]
            d.direction.data[:] = sun["direction"]
            d.angularDiameterDegrees = sun["angularDiameterDegrees"]
            _style(d, sun)
            ffi.check(prefix + "UploadDirectionalLight", dir_fn(inst, C.byref(d)))
        for i, sp in enumerate(lights.get("spheres", [])):
            s = RgSphericalLightUploadInfo()
            s.uniqueID = 100 + i
            s.color = l.rgUtilPackColorFloat4D(*sp["color"], 1.0)
            s.intensity = sp["intensity"]
            s.position.data[:] = sp["position"]
            s.radius = sp["radius"]
            _style(s, sp)
            ffi.check(prefix + "UploadSphericalLight", sph_fn(inst, C.byref(s)))
        for i, sp in enumerate(lights.get("spots", [])):
            ffi.check("rgUploadSpotLight", l.rgUploadSpotLight(inst, C.byref(spot_info(sp, 200 + i))))
        for i, pl in enumerate(lights.get("polys", [])):
            ffi.check("rgUploadPolygonalLight", l.rgUploadPolygonalLight(inst, C.byref(poly_info(pl, 300 + i))))


def spot_info(sp: dict, uid: int) -> RgSpotLightUploadInfo:
    """A spot light dict (color, intensity, position, direction, radius,
    angleOuter, angleInner in radians) -> RgSpotLightUploadInfo."""
    s = RgSpotLightUploadInfo()
    s.uniqueID = uid
    s.color = api().rgUtilPackColorFloat4D(*sp["color"], 1.0)
    s.intensity = sp["intensity"]
    s.position.data[:] = sp["position"]
    s.direction.data[:] = sp["direction"]
    s.radius, s.angleOuter, s.angleInner = sp["radius"], sp["angleOuter"], sp["angleInner"]
    _style(s, sp)
    return s


def poly_info(pl: dict, uid: int) -> RgPolygonalLightUploadInfo:
    """A triangle light dict (color, intensity, positions: 3 points) ->
    RgPolygonalLightUploadInfo."""
    s = RgPolygonalLightUploadInfo()
    s.uniqueID = uid
    s.color = api().rgUtilPackColorFloat4D(*pl["color"], 1.0)
    s.intensity = pl["intensity"]
    for k in range(3):
        s.positions[k].data[:] = pl["positions"][k]
    _style(s, pl)
    return s


class Renderer:
    """One RTGL1 instance rendering a synthetic scene at a fixed render size."""

    def __init__(self, width: int, height: int, cull_back: bool = False, allow_sky: bool = False):
        l = api()
        ci = RgInstanceCreateInfo()
        ci.pAppName = b"raytracedgl1_amd"
        ci.rayCullBackFacingTriangles = 1 if cull_back else 0
        ci.allowGeometryWithSkyFlag = 1 if allow_sky else 0
        h = _P()
        ffi.check("rgCreateInstance", l.rgCreateInstance(C.byref(ci), C.byref(h)))
        self.h = h.value
        self.width, self.height = width, height
        self.path = RgExtDrawFramePathParams()
        self.path.sType = RG_EXT_STRUCTURE_TYPE_PATH_PARAMS
        self.path.renderWidth, self.path.renderHeight = width, height
        self.path.restirTemporalSamples = 0xFFFFFFFF
        self.path.restirSpatialSamples = 0xFFFFFFFF
        self.sky = RgDrawFrameSkyParams()
        self.sky.sType = RG_STRUCTURE_TYPE_SKY
        self.sky.skyColorDefault.data[:] = (0.71, 0.88, 1.0)
        self.sky.skyColorMultiplier = 1000.0
        self.sky.skyColorSaturation = 1.0
        self.sky.pNext = C.cast(C.pointer(self.path), _P)
        self.illum = None
        self.rr = None
        self.time = 0.0
        self.cull_mask = RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT  # RgDrawFrameInfo::rayCullMaskWorld
        self.styles = None
        self.ignore_fpv = None
        self.illum_over = {}

    def _link(self) -> None:
        chain = [self.sky, self.path] + [x for x in (self.illum, self.rr) if x is not None]
        for a, b in zip(chain, chain[1:] + [None]):
            a.pNext = C.cast(C.pointer(b), _P) if b is not None else None

    def set_reflect_refract(self, params: RgDrawFrameReflectRefractParams | None) -> None:
        """Chains RgDrawFrameReflectRefractParams (None: the defaults apply)."""
        self.rr = params
        self._link()

    def set_lightstyles(self, values) -> None:
        """Chains RgDrawFrameIlluminationParams (defaults of DrawFrameInfo.h)
        with these light style values into every following rgDrawFrame."""
        self._illumination(values, self.ignore_fpv)

    def set_ignore_fpv_shadows(self, light_unique_id) -> None:
        """RgDrawFrameIlluminationParams::lightUniqueIdIgnoreFirstPersonViewerShadows
        (None: no such light)."""
        self._illumination(self.styles_values if self.styles is not None else None, light_unique_id)

    def set_illumination(self, **fields) -> None:
        """Overrides RgDrawFrameIlluminationParams fields (cellWorldSize,
        polygonalLightSpotlightFactor, ...) in every following rgDrawFrame."""
        self.illum_over.update(fields)
        self._illumination(self.styles_values if self.styles is not None else None, self.ignore_fpv)

    def _illumination(self, values, ignore_fpv) -> None:
        self.styles_values = values
        self.ignore_fpv = ignore_fpv
        if values is None and ignore_fpv is None and not self.illum_over:
            self.illum = None
            self.styles = None
            self._link()
            return
        self.styles = (C.c_float * max(1, len(values or [])))(*(values or []))
        self.fpv_id = C.c_uint64(ignore_fpv or 0)
        il = RgDrawFrameIlluminationParams()
        il.sType = RG_STRUCTURE_TYPE_ILLUMINATION
        il.maxBounceShadows = 2
        il.enableSecondBounceForIndirect = 1
        il.cellWorldSize = 1.0
        il.directDiffuseSensitivityToChange = 0.5
        il.indirectDiffuseSensitivityToChange = 0.2
        il.specularSensitivityToChange = 0.5
        il.polygonalLightSpotlightFactor = 2.0
        il.lightstyleValuesCount = len(values or [])
        il.pLightstyleValues = C.cast(self.styles, _P) if values else None
        il.lightUniqueIdIgnoreFirstPersonViewerShadows = C.cast(C.pointer(self.fpv_id), _P) if ignore_fpv is not None else None
        for k, v in self.illum_over.items():
            setattr(il, k, v)
        self.illum = il
        self._link()

    def provide_texture(self, name: str, rgba8: np.ndarray, filter: int = RG_SAMPLER_FILTER_AUTO,
                        address_u: int = RG_SAMPLER_ADDRESS_MODE_REPEAT, address_v: int = RG_SAMPLER_ADDRESS_MODE_REPEAT) -> None:
        """rgProvideOriginalTexture with (height, width, 4) uint8 sRGB pixels."""
        px = np.ascontiguousarray(rgba8, np.uint8)
        info = RgOriginalTextureInfo()
        info.pTextureName = name.encode()
        info.pPixels = px.ctypes.data
        info.size.width, info.size.height = px.shape[1], px.shape[0]
        info.filter, info.addressModeU, info.addressModeV = filter, address_u, address_v
        ffi.check("rgProvideOriginalTexture", api().rgProvideOriginalTexture(self.h, C.byref(info)))

    def delete_texture(self, name: str) -> None:
        ffi.check("rgMarkOriginalTextureAsDeleted", api().rgMarkOriginalTextureAsDeleted(self.h, name.encode()))

    def set_static_geometry(self, enable: bool) -> None:
        ffi.check("rgExtSetStaticGeometry", api().rgExtSetStaticGeometry(self.h, 1 if enable else 0))

    def frame(self, uploader: SceneUploader, cam, dt: float = 1.0 / 60.0, geometry: bool = True,
              stats: bool = True):
        """rgStartFrame, uploads, rgDrawFrame; geometry=False uploads lights only.
        Returns the frame's statistics (waiting for the frame), or None with
        stats=False (the frame stays in flight)."""
        l = api()
        sfi = RgStartFrameInfo()
        ffi.check("rgStartFrame", l.rgStartFrame(self.h, C.byref(sfi)))
        uploader.upload(self.h, geometry)
        self.time += dt
        di = draw_info(cam, self.time, C.cast(C.pointer(self.sky), _P), self.cull_mask)
        ffi.check("rgDrawFrame", l.rgDrawFrame(self.h, C.byref(di)))
        return self.stats() if stats else None

    def stats(self) -> RgExtFrameStats:
        st = RgExtFrameStats()
        ffi.check("rgExtGetFrameStats", api().rgExtGetFrameStats(self.h, C.byref(st)))
        return st

    def framebuffers(self) -> "ffi.FramebufferView":
        p = _P()
        ffi.check("rgExtGetFramebuffers", api().rgExtGetFramebuffers(self.h, C.byref(p)))
        return ffi.FramebufferView(p.value, self.width, self.height)

    def frame_params(self) -> RgExtFrameParams:
        p = RgExtFrameParams()
        ffi.check("rgExtGetFrameParams", api().rgExtGetFrameParams(self.h, C.byref(p)))
        return p

    def histogram(self) -> np.ndarray:
        h = np.zeros(256, np.uint32)
        ffi.check("rgExtReadbackHistogram", api().rgExtReadbackHistogram(self.h, h.ctypes.data))
        return h

    def close(self):
        if self.h:
            api().rgDestroyInstance(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
