"""The C ABI called from compiled C: tests/c_host/rtgl_host.c runs the frame
loop of the reference's test application (Tests/RtglTest.cpp:439-703)
headless against librtgl1_amd.so, checks every result code and the
protocol's error codes, and writes the last frame's FINAL plane.  The GPU
test renders the same frames through the Python binding, uploading the
ground's primitives in their natural order where the C program shuffles
them, and requires the two FINAL planes to be bit-identical."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from raytracedgl1_amd import ffi, rtgl

REPO = Path(__file__).resolve().parent.parent
HOST_DIR = REPO / "tests" / "c_host"
HOST_BIN = HOST_DIR / "rtgl_host"

FRAMES, W, H = 4, 160, 90
# column-major view matrix of rtgl_host.c (camera at (0, 6, 14), pitched down)
VIEW = [1.0, 0.0, 0.0, 0.0, 0.0, 0.8, 0.6, 0.0, 0.0, -0.6, 0.8, 0.0, 0.0, 3.6, -14.8, 1.0]


def test_c_host_compiles_against_headers(tmp_path):
    """The C program compiles as C11 with every warning an error against
    include/RTGL1 (no GPU involved)."""
    res = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                          f"-I{REPO / 'include'}", str(HOST_DIR / "rtgl_host.c")],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def _ground_tile(tile: int):
    x0 = np.float32(-8.0 + 8.0 * (tile % 2))
    z0 = np.float32(-8.0 + 8.0 * (tile // 2))
    xs = np.array([x0, x0 + 8, x0 + 8, x0], np.float32)
    zs = np.array([z0, z0, z0 + 8, z0 + 8], np.float32)
    v = np.zeros(4, rtgl.VERTEX_DTYPE)
    v["position"][:, 0] = xs
    v["position"][:, 2] = zs
    v["normal"][:, 1] = 1.0
    v["texCoord"][:, 0] = xs * np.float32(0.25)
    v["texCoord"][:, 1] = zs * np.float32(0.25)
    v["color"] = 0xFFFFFFFF
    return v, np.array([0, 2, 1, 0, 3, 2], np.uint32)


def _box():
    faces = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    out = np.zeros(36, rtgl.VERTEX_DTYPE)
    k = 0
    for f, n in enumerate(faces):
        n = np.array(n, np.float32)
        a = f // 2
        u, w = (a + 1) % 3, (a + 2) % 3
        t = np.zeros(3, np.float32)
        b = np.zeros(3, np.float32)
        t[u] = 1.0
        b[w] = n[a]
        su = np.array([-0.5, 0.5, 0.5, -0.5], np.float32)
        sw = np.array([-0.5, -0.5, 0.5, 0.5], np.float32)
        c = [np.float32(0.5) * n + su[i] * t + sw[i] * b for i in range(4)]
        for i in (0, 1, 2, 0, 2, 3):
            out["position"][k] = c[i]
            out["normal"][k] = n
            out["color"][k] = 0xFFFFFFFF
            k += 1
    return out


def _transform(x, y, z) -> rtgl.RgTransform:
    t = rtgl.RgTransform()
    for i in range(3):
        t.matrix[i][i] = 1.0
    t.matrix[0][3], t.matrix[1][3], t.matrix[2][3] = x, y, z
    return t


def _python_final() -> np.ndarray:
    """The C program's frames through ctypes, ground tiles unshuffled."""
    l = rtgl.api()
    ci = rtgl.RgInstanceCreateInfo()
    ci.pAppName = b"rtgl_host"
    ci.pAppGUID = b"c4e3b8a0-0000-4000-8000-rtglhost0001"
    ci.worldUp.data[1] = 1.0
    ci.worldForward.data[2] = 1.0
    ci.worldScale = 1.0
    h = C.c_void_p()
    ffi.check("rgCreateInstance", l.rgCreateInstance(C.byref(ci), C.byref(h)))
    keep = []
    try:
        white = C.c_uint32(0xFFFFFFFF)
        cube = rtgl.RgOriginalCubemapInfo()
        cube.pTextureName = b"_external_/cubemap/0"
        for side in ("PositiveX", "NegativeX", "PositiveY", "NegativeY", "PositiveZ", "NegativeZ"):
            setattr(cube, "pPixels" + side, C.cast(C.pointer(white), C.c_void_p))
        cube.sideSize = 1
        ffi.check("rgProvideOriginalCubemapTexture", l.rgProvideOriginalCubemapTexture(h, C.byref(cube)))
        px = np.array([l.rgUtilPackColorByte4D(200, 200, 200, 255) if (i % 4 + i // 4) % 2
                       else l.rgUtilPackColorByte4D(90, 110, 90, 255) for i in range(16)], np.uint32)
        tex = rtgl.RgOriginalTextureInfo()
        tex.pTextureName = b"ground"
        tex.pPixels = px.ctypes.data
        tex.size.width, tex.size.height = 4, 4
        ffi.check("rgProvideOriginalTexture", l.rgProvideOriginalTexture(h, C.byref(tex)))

        tiles = [_ground_tile(t) for t in range(4)]
        box = _box()
        quad = box[:6].copy()
        keep += [tiles, box, quad]
        for frame in range(FRAMES):
            sfi = rtgl.RgStartFrameInfo()
            sfi.pMapName = b"untitled"
            ffi.check("rgStartFrame", l.rgStartFrame(h, C.byref(sfi)))

            mesh = rtgl.RgMeshInfo()
            mesh.uniqueObjectID = 1
            mesh.pMeshName = b"ground"
            mesh.transform = _transform(0.0, 0.0, 0.0)
            mesh.isExportable = 1
            for t, (v, idx) in enumerate(tiles):
                prim = rtgl.RgMeshPrimitiveInfo()
                prim.primitiveIndexInMesh = t
                prim.pVertices = v.ctypes.data
                prim.vertexCount = 4
                prim.pIndices = idx.ctypes.data
                prim.indexCount = 6
                prim.pTextureName = b"ground"
                prim.color = 0xFFFFFFFF
                ffi.check("rgUploadMeshPrimitive", l.rgUploadMeshPrimitive(h, C.byref(mesh), C.byref(prim)))

            x = np.float32(5.0) - np.float32(0.05) * np.float32((frame + 30) % 200)
            mesh = rtgl.RgMeshInfo()
            mesh.uniqueObjectID = 10
            mesh.pMeshName = b"test"
            mesh.transform = _transform(float(x), 1.0, 0.0)
            prim = rtgl.RgMeshPrimitiveInfo()
            prim.pVertices = box.ctypes.data
            prim.vertexCount = 36
            prim.color = l.rgUtilPackColorByte4D(128, 255, 128, 255)
            ffi.check("rgUploadMeshPrimitive", l.rgUploadMeshPrimitive(h, C.byref(mesh), C.byref(prim)))

            mesh = rtgl.RgMeshInfo()
            mesh.uniqueObjectID = 12
            mesh.pMeshName = b"test_raster"
            mesh.transform = _transform(-0.5, 1.0, 1.0)
            prim = rtgl.RgMeshPrimitiveInfo()
            prim.pVertices = quad.ctypes.data
            prim.vertexCount = 6
            prim.color = l.rgUtilPackColorByte4D(255, 128, 128, 128)
            ffi.check("rgUploadMeshPrimitive", l.rgUploadMeshPrimitive(h, C.byref(mesh), C.byref(prim)))

            sun = rtgl.RgDirectionalLightUploadInfo()
            sun.uniqueID = 0
            sun.isExportable = 1
            sun.color = l.rgUtilPackColorByte4D(255, 255, 255, 255)
            sun.intensity = 7000.0
            sun.direction.data[:] = (-1.0, -8.0, -1.0)
            sun.angularDiameterDegrees = 0.5
            ffi.check("rgUploadDirectionalLight", l.rgUploadDirectionalLight(h, C.byref(sun)))

            sky = rtgl.RgDrawFrameSkyParams()
            sky.sType = rtgl.RG_STRUCTURE_TYPE_SKY
            sky.skyType = rtgl.RG_SKY_TYPE_COLOR
            sky.skyColorDefault.data[:] = (0.71, 0.88, 1.0)
            sky.skyColorMultiplier = 1000.0
            sky.skyColorSaturation = 1.0
            sky.pSkyCubemapTextureName = b"_external_/cubemap/0"
            res = rtgl.RgDrawFrameRenderResolutionParams()
            res.sType = rtgl.RG_STRUCTURE_TYPE_RENDER_RESOLUTION
            res.pNext = C.cast(C.pointer(sky), C.c_void_p)
            res.upscaleTechnique = rtgl.RG_RENDER_UPSCALE_TECHNIQUE_LINEAR
            res.resolutionMode = rtgl.RG_RENDER_RESOLUTION_MODE_CUSTOM
            res.customRenderSize.width, res.customRenderSize.height = W, H
            fi = rtgl.RgDrawFrameInfo()
            fi.view[:] = VIEW
            fi.fovYRadians = 1.3089969
            fi.cameraNear = 0.1
            fi.cameraFar = 10000.0
            fi.rayLength = 10000.0
            fi.rayCullMaskWorld = rtgl.RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT
            fi.currentTime = (frame + 1) / 60.0
            fi.pParams = C.cast(C.pointer(res), C.c_void_p)
            ffi.check("rgDrawFrame", l.rgDrawFrame(h, C.byref(fi)))
        p = C.c_void_p()
        ffi.check("rgExtGetFramebuffers", l.rgExtGetFramebuffers(h, C.byref(p)))
        return ffi.FramebufferView(p.value, W, H).download(ffi.PLANE["FINAL"]).copy()
    finally:
        l.rgDestroyInstance(h)


@pytest.mark.gpu
def test_c_host_frames_match_python(tmp_path):
    assert HOST_BIN.exists(), "tests/c_host/rtgl_host is built by __graft_entry__.build()"
    out = tmp_path / "final.bin"
    print(f"running {HOST_BIN.name} {FRAMES} frames {W}x{H}", flush=True)
    res = subprocess.run([str(HOST_BIN), str(out), str(FRAMES), str(W), str(H)], capture_output=True,
                         text=True, timeout=300, env=dict(os.environ))
    assert res.returncode == 0, res.stdout + res.stderr
    assert "rtgl_host ok" in res.stdout
    c_final = np.fromfile(out, np.float32).reshape(H, W, 4)
    py_final = _python_final()
    assert py_final.shape == (H, W, 4)
    assert np.isfinite(c_final).all()
    # the image is not trivial: ground, box and sky differ
    assert np.unique(c_final[..., :3].round(3).reshape(-1, 3), axis=0).shape[0] > 16
    assert np.array_equal(c_final.view(np.uint32), py_final.view(np.uint32))
