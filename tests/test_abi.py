"""The C-ABI boundary without a GPU: librtgl1_amd.so loads, exports every
function include/RTGL1/*.h declares, its structures have the reference's
sizes, and the calls that need no device behave like the reference's
(Source/RTGL1.cpp:31-361: result codes, colour packing, descriptions).
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from raytracedgl1_amd import ffi, rtgl

REPO = Path(__file__).resolve().parents[1]
HEADERS = sorted((REPO / "include" / "RTGL1").glob("*.h"))

# RgResult values (RTGL1.h, enum RgResult)
RG_RESULT_SUCCESS = 0
RG_RESULT_WRONG_INSTANCE = 3
RG_RESULT_CANT_FIND_SUPPORTED_PHYSICAL_DEVICE = 7
RG_RESULT_FRAME_WASNT_STARTED = 8


def declared_functions():
    names = []
    for h in HEADERS:
        text = h.read_text()
        names += re.findall(r"RGAPI\s+[\w\s\*]+?RGCONV\s+(\w+)\s*\(", text)
    return names


def test_headers_declare_the_reference_entry_points():
    names = set(declared_functions())
    # the minimum set a drop-in for the reference's ray-traced path must export
    for n in ("rgCreateInstance", "rgDestroyInstance", "rgUploadMeshPrimitive", "rgUploadDirectionalLight",
              "rgUploadSphericalLight", "rgUploadSpotLight", "rgUploadPolygonalLight", "rgStartFrame",
              "rgDrawFrame", "rgUtilGetResultDescription", "rgUtilPackColorByte4D", "rgUtilPackColorFloat4D"):
        assert n in names, n


def test_library_exports_every_declared_symbol():
    lib = ffi.lib()
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(ffi.LIB_PATH)], capture_output=True, text=True)
    if out.returncode == 0:
        exported = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
        assert set(names) <= exported


def test_struct_sizes_match_the_reference_abi():
    # sizes of the reference's structures on x86-64 (RTGL1.h)
    assert C.sizeof(rtgl.RgInstanceCreateInfo) == 184
    assert C.sizeof(rtgl.RgMeshInfo) == 88
    assert C.sizeof(rtgl.RgMeshPrimitiveInfo) == 80
    assert C.sizeof(rtgl.RgDrawFrameInfo) == 120
    assert C.sizeof(rtgl.RgTransform) == 48
    assert C.sizeof(rtgl.RgDirectionalLightUploadInfo) == 48
    assert C.sizeof(rtgl.RgSphericalLightUploadInfo) == 48
    assert rtgl.VERTEX_DTYPE.itemsize == 64
    assert rtgl.RgDrawFrameInfo.pParams.offset == 112
    assert rtgl.RgMeshPrimitiveInfo.pVertices.offset == 16


def test_pack_color_known_answers():
    l = rtgl.api()
    # Utils::PackColor: r | g<<8 | b<<16 | a<<24 (Source/Utils.h:300-304)
    assert l.rgUtilPackColorByte4D(1, 2, 3, 4) == 0x04030201
    assert l.rgUtilPackColorByte4D(255, 0, 0, 255) == 0xFF0000FF
    # PackColorFromFloat truncates c*255 and clamps to [0, 255] (Source/Utils.h:306-313)
    assert l.rgUtilPackColorFloat4D(1.0, 0.5, 0.0, 1.0) == 0xFF007FFF
    assert l.rgUtilPackColorFloat4D(2.0, -1.0, 0.999, 0.0) == 0x00FE00FF


def test_result_descriptions():
    l = rtgl.api()
    for code in range(0, 16):
        d = l.rgUtilGetResultDescription(code)
        assert isinstance(d, bytes) and len(d) > 0
    assert l.rgUtilGetResultDescription(RG_RESULT_SUCCESS) != l.rgUtilGetResultDescription(RG_RESULT_WRONG_INSTANCE)


def test_calls_on_a_missing_instance_return_wrong_instance():
    l = rtgl.api()
    assert l.rgStartFrame(None, None) == RG_RESULT_WRONG_INSTANCE
    assert l.rgDrawFrame(None, None) == RG_RESULT_WRONG_INSTANCE
    assert l.rgStartFrame(C.c_void_p(0x1234), None) == RG_RESULT_WRONG_INSTANCE
    assert l.rgDestroyInstance(C.c_void_p(0x1234)) == RG_RESULT_WRONG_INSTANCE


def test_create_without_device_fails_cleanly():
    if ffi.device_count() > 0:
        pytest.skip("a HIP device is visible; the no-device path is not reachable")
    l = rtgl.api()
    ci = rtgl.RgInstanceCreateInfo()
    h = C.c_void_p()
    r = l.rgCreateInstance(C.byref(ci), C.byref(h))
    assert r == RG_RESULT_CANT_FIND_SUPPORTED_PHYSICAL_DEVICE
    assert not h.value
    assert ffi.lib().rgExtGetLastErrorMessage()


def test_frame_params_defaults():
    p = ffi.RgExtFrameParams()
    ffi.call("rgExtFrameParamsDefaults", 1920, 1080, C.byref(p))
    assert (p.renderWidth, p.renderHeight) == (1920, 1080)
    assert p.atrousIterations == 4
    assert p.restirInitialCandidates == 8 and p.restirSpatialSamples == 8
