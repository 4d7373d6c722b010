"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: used only by tests, the smoke check and the CPU
baseline of bench.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from raytracedgl1_amd.ffi import PLANES as _PLANES

REPO = Path(__file__).resolve().parents[1]
ORACLE_DIR = REPO / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

_P = C.c_void_p
_U32 = C.c_uint32
_F = C.c_float


def build_oracle() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "-j8"], check=True)
    return ORACLE_LIB


class Oracle:
    def __init__(self):
        if not ORACLE_LIB.exists():
            build_oracle()
        l = C.CDLL(str(ORACLE_LIB))
        sig = {
            "or_log2f": (_F, [_F]),
            "or_exp2f": (_F, [_F]),
            "or_wellons_low_bias32": (_U32, [_U32]),
            "or_murmur_hash33": (None, [_U32, _U32, _U32, _P]),
            "or_random_seed": (_U32, [C.c_int, C.c_int, _U32]),
            "or_encode_normal": (_U32, [_P]),
            "or_decode_normal": (None, [_U32, _P]),
            "or_encode_e5b9g9r9": (_U32, [_P]),
            "or_decode_e5b9g9r9": (None, [_U32, _P]),
            "or_checkerboard_from_regular": (None, [C.c_int, C.c_int, C.c_int, _P, _P]),
            "or_checkerboard_to_regular": (None, [C.c_int, C.c_int, C.c_int, _P, _P]),
            "or_bvh_build": (_P, [_P, _U32, _U32]),
            "or_bvh_free": (None, [_P]),
            "or_bvh_node_count": (_U32, [_P]),
            "or_bvh_scene_bounds": (None, [_P, _P, _P]),
            "or_bvh_copy_codes": (None, [_P, _P]),
            "or_bvh_copy_sorted_indices": (None, [_P, _P]),
            "or_bvh_copy_binary": (None, [_P, _P, _P]),
            "or_bvh_copy_nodes": (None, [_P, _P]),
            "or_trace_rays": (None, [_P, _P, _U32, _P, _U32, C.c_int, _P, _P]),
            "or_trace_bruteforce_f64": (None, [_P, _U32, _P, _U32, _U32, _P, _P, _P, _P, C.c_int]),
            "or_lum_histogram": (None, [_P, _U32, _F, _F, _P]),
            "or_lum_average": (_F, [_P, _U32, _F, _F, _F, C.c_int, _F]),
            "or_plane_components": (C.c_int, [C.c_int]),
            "or_plane_layout": (C.c_int, [C.c_int]),
            "or_plane_has_prev": (C.c_int, [C.c_int]),
            "or_denoise_gradient_atrous": (None, [_P, _P, _U32]),
            "or_lpm_luma_coefficients": (None, [_P]),
            "or_lpm_map": (None, [_P, _P, _P, _P]),
            "or_compose_final": (None, [_P, _P, _P, C.c_double, C.c_double]),
            "or_pt_create": (_P, [C.c_int, C.c_int]),
            "or_pt_destroy": (None, [_P]),
            "or_pt_frame": (None, [_P, _P, _P, _P, _P, C.c_int]),
            "or_flat_normals": (None, [_P, _U32, _P, _U32, _P]),
            "or_denoise_temporal": (None, [_P, _P]),
            "or_denoise_antifirefly": (None, [_P, _P]),
            "or_denoise_variance": (None, [_P, _P]),
            "or_denoise_atrous": (None, [_P, _P, _U32]),
            "or_denoise": (None, [_P, _P]),
            "or_denoise_set_threads": (None, [C.c_int]),
            "or_texture_create": (_P, [_P, _U32, _U32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "or_texture_free": (None, [_P]),
            "or_texture_mip_count": (_U32, [_P]),
            "or_texture_copy_mip": (None, [_P, _U32, _P]),
            "or_texture_sample_lod": (None, [_P, C.c_double, C.c_double, C.c_double, _P]),
            "or_texture_sample_grad": (None, [_P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        self.l = l

    # ---- math / codecs
    def log2f(self, x: float) -> float:
        return self.l.or_log2f(x)

    def exp2f(self, x: float) -> float:
        return self.l.or_exp2f(x)

    def murmur33(self, x, y, z):
        out = (C.c_uint32 * 3)()
        self.l.or_murmur_hash33(x, y, z, out)
        return tuple(out)

    def encode_normal(self, n) -> int:
        a = np.ascontiguousarray(n, np.float64)
        return self.l.or_encode_normal(a.ctypes.data)

    def decode_normal(self, p: int) -> np.ndarray:
        a = np.zeros(3, np.float64)
        self.l.or_decode_normal(p, a.ctypes.data)
        return a

    def encode_e5(self, rgb) -> int:
        a = np.ascontiguousarray(rgb, np.float64)
        return self.l.or_encode_e5b9g9r9(a.ctypes.data)

    def decode_e5(self, p: int) -> np.ndarray:
        a = np.zeros(3, np.float64)
        self.l.or_decode_e5b9g9r9(p, a.ctypes.data)
        return a

    def cb_from_regular(self, x, y, w):
        cx, cy = C.c_int(), C.c_int()
        self.l.or_checkerboard_from_regular(x, y, w, C.byref(cx), C.byref(cy))
        return cx.value, cy.value

    def cb_to_regular(self, cx, cy, w):
        x, y = C.c_int(), C.c_int()
        self.l.or_checkerboard_to_regular(cx, cy, w, C.byref(x), C.byref(y))
        return x.value, y.value

    # ---- LBVH
    def bvh(self, tris: np.ndarray, leaf: int = 3) -> "OracleBvh":
        return OracleBvh(self, tris, leaf)

    def bruteforce(self, tris: np.ndarray, rays: np.ndarray, flags: int = 0, threads: int = 8):
        tris = np.ascontiguousarray(tris, np.float32)
        rays = np.ascontiguousarray(rays)
        n = len(rays)
        t, u, v = (np.zeros(n, np.float64) for _ in range(3))
        p = np.zeros(n, np.uint32)
        self.l.or_trace_bruteforce_f64(tris.ctypes.data, len(tris), rays.ctypes.data, n, flags,
                                       t.ctypes.data, u.ctypes.data, v.ctypes.data, p.ctypes.data, threads)
        return t, u, v, p

    # ---- eye adaptation
    def lum_histogram(self, lum: np.ndarray, min_log: float, max_log: float) -> np.ndarray:
        lum = np.ascontiguousarray(lum, np.float32).reshape(-1)
        h = np.zeros(256, np.uint32)
        self.l.or_lum_histogram(lum.ctypes.data, lum.size, min_log, max_log, h.ctypes.data)
        return h

    def lum_average(self, hist: np.ndarray, pixels: int, min_log, max_log, dt, stop, last) -> float:
        h = np.ascontiguousarray(hist, np.uint32).copy()
        return self.l.or_lum_average(h.ctypes.data, pixels, min_log, max_log, dt, int(stop), last)


class OracleBvh:
    def __init__(self, o: Oracle, tris: np.ndarray, leaf: int):
        self.o = o
        self.tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.n = len(self.tris)
        self.h = o.l.or_bvh_build(self.tris.ctypes.data, self.n, leaf)
        assert self.h, "or_bvh_build failed"

    def __del__(self):
        try:
            self.o.l.or_bvh_free(self.h)
        except Exception:
            pass

    def download(self) -> dict:
        from raytracedgl1_amd.ffi import NODE8_DTYPE
        l, n = self.o.l, self.n
        m = l.or_bvh_node_count(self.h)
        codes = np.zeros(n, np.uint64)
        idx = np.zeros(n, np.uint32)
        children = np.zeros((max(n - 1, 0), 2), np.uint32)
        bounds = np.zeros((2 * n - 1, 6), np.float32)
        nodes = np.zeros(m, NODE8_DTYPE)
        lo, hi = np.zeros(3), np.zeros(3)
        l.or_bvh_copy_codes(self.h, codes.ctypes.data)
        l.or_bvh_copy_sorted_indices(self.h, idx.ctypes.data)
        l.or_bvh_copy_binary(self.h, children.ctypes.data if n > 1 else None, bounds.ctypes.data)
        l.or_bvh_copy_nodes(self.h, nodes.ctypes.data)
        l.or_bvh_scene_bounds(self.h, lo.ctypes.data, hi.ctypes.data)
        return {"n": n, "node_count": m, "codes": codes, "sorted_idx": idx, "children": children,
                "bounds": bounds, "nodes": nodes, "scene_min": lo, "scene_max": hi}

    def trace(self, rays: np.ndarray, flags: int = 0, threads: int = 8, stats: bool = False):
        from raytracedgl1_amd.ffi import HIT_DTYPE
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), HIT_DTYPE)
        nv, tt = C.c_uint64(), C.c_uint64()
        self.o.l.or_trace_rays(self.h, rays.ctypes.data, len(rays), hits.ctypes.data, flags, threads,
                               C.byref(nv), C.byref(tt))
        if stats:
            return hits, int(nv.value), int(tt.value)
        return hits


# the product's planes, then OR_PLANE_GI_TIE (oracle/oracle_denoise.h)
OR_PLANE_GI_TIE = len(_PLANES)
PLANE_COUNT = len(_PLANES) + 1


class OrFramebuffersC(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int),
                ("cur", C.c_void_p * PLANE_COUNT), ("prev", C.c_void_p * PLANE_COUNT)]


class OracleFramebuffers:
    """FP64 host planes for the denoiser oracle (components interleaved)."""

    def __init__(self, o: Oracle, width: int, height: int):
        self.o, self.width, self.height = o, width, height
        self.cur, self.prev = {}, {}
        gw, gh = max(1, (width + 1) // 3), max(1, (height + 1) // 3)
        self.c = OrFramebuffersC()
        self.c.width, self.c.height = width, height
        for p in range(PLANE_COUNT):
            comps = o.l.or_plane_components(p)
            layout = o.l.or_plane_layout(p)
            shape = (gh, gw, comps) if layout == 2 else (height, width, comps)
            self.cur[p] = np.zeros(shape, np.float64)
            self.c.cur[p] = self.cur[p].ctypes.data
            if o.l.or_plane_has_prev(p):
                self.prev[p] = np.zeros(shape, np.float64)
                self.c.prev[p] = self.prev[p].ctypes.data

    def set(self, plane: int, arr, prev: bool = False):
        dst = self.prev[plane] if prev else self.cur[plane]
        dst[...] = np.asarray(arr).reshape(dst.shape).astype(np.float64)

    def run(self, name: str, params, *args):
        fn = getattr(self.o.l, "or_denoise" + ("_" + name if name else ""))
        fn(C.byref(self.c), C.byref(params), *args)

    def swap(self):
        """Start of a frame: previous-frame planes become the current frame's prev."""
        for p in list(self.prev):
            self.cur[p], self.prev[p] = self.prev[p], self.cur[p]
            self.c.cur[p] = self.cur[p].ctypes.data
            self.c.prev[p] = self.prev[p].ctypes.data


class OrGeomC(C.Structure):
    _fields_ = [("prevFromCur", C.c_float * 12), ("isStatic", _U32), ("color", _U32), ("emissive", _F),
                ("roughness", _F), ("metallic", _F), ("flags", _U32), ("firstTri", _U32), ("triCount", _U32),
                ("texture", _U32), ("texOrm", _U32), ("texNormal", _U32), ("texEmissive", _U32)]


class OracleTexture:
    """The oracle's albedo texture (oracle/oracle_texture.cpp) of
    (height, width, 4) uint8 sRGB pixels."""

    def __init__(self, o: "Oracle", rgba8, filter: int = 0, address_u: int = 0, address_v: int = 0, force_min_linear: bool = False,
                 srgb: bool = True):
        self.o = o
        self.px = np.ascontiguousarray(rgba8, np.uint8)
        self.h = o.l.or_texture_create(self.px.ctypes.data, self.px.shape[1], self.px.shape[0], filter, address_u, address_v,
                                       1 if force_min_linear else 0, 1 if srgb else 0)

    def mips(self) -> list:
        out, w, h = [], self.px.shape[1], self.px.shape[0]
        for lvl in range(self.o.l.or_texture_mip_count(self.h)):
            a = np.zeros((h, w, 4), np.uint8)
            self.o.l.or_texture_copy_mip(self.h, lvl, a.ctypes.data)
            out.append(a)
            w, h = max(1, w // 2), max(1, h // 2)
        return out

    def sample_lod(self, u, v, lod):
        r = np.zeros(4)
        self.o.l.or_texture_sample_lod(self.h, u, v, lod, r.ctypes.data)
        return r

    def sample_grad(self, u, v, dudx, dvdx, dudy, dvdy):
        r = np.zeros(4)
        self.o.l.or_texture_sample_grad(self.h, u, v, dudx, dvdx, dudy, dvdy, r.ctypes.data)
        return r

    def __del__(self):
        try:
            self.o.l.or_texture_free(self.h)
        except Exception:
            pass


class OrPtSceneC(C.Structure):
    _fields_ = [("tris", _P), ("normals", _P), ("triGeom", _P), ("triCount", _U32), ("geoms", _P), ("geomIds", _P),
                ("geomCount", _U32), ("dirLight", _P), ("spheres", _P), ("sphereCount", _U32),
                ("spots", _P), ("spotCount", _U32), ("polys", _P), ("polyCount", _U32),
                ("textures", _P), ("textureCount", _U32), ("texcoords", _P)]

# RgMeshPrimitiveFlags of the refracting primitives that are not alpha tested
# (VertexCollectorFilterType.cpp:215-229): water 128, glass 64,
# glass-if-smooth 2048, acid 8192; alpha tested 1.
_REFRACT_FLAGS, _ALPHA_TESTED = 128 | 64 | 2048 | 8192, 1
FIRST_PERSON, FIRST_PERSON_VIEWER, DONT_GENERATE_NORMALS, SKY_VISIBILITY = 4, 8, 256, 32768


def refracts(flags: int) -> bool:
    return (flags & _ALPHA_TESTED) == 0 and (flags & _REFRACT_FLAGS) != 0


def frame_class(flags: int) -> int:
    """The position of a primitive's triangles in the library's frame order:
    world 0, world 2 (sky visibility), refracting of world 0, refracting of
    world 2, first person, first-person viewer; upload order within each."""
    if flags & FIRST_PERSON:
        return 4
    if flags & FIRST_PERSON_VIEWER:
        return 5
    w2 = (flags & SKY_VISIBILITY) != 0
    return (3 if w2 else 2) if refracts(flags) else (1 if w2 else 0)


def flat_normals(o: "Oracle", tris: np.ndarray) -> np.ndarray:
    """or_flat_normals over a non-indexed triangle list ((n, 9) floats)."""
    pos = np.ascontiguousarray(tris, np.float32).reshape(-1, 3)
    nrm = np.zeros_like(pos)
    o.l.or_flat_normals(pos.ctypes.data, len(pos), None, 0, nrm.ctypes.data)
    return nrm.reshape(-1, 9)


class OraclePathTracer:
    """The FP64 oracle of a frame's traced passes (or_pt_frame) over a
    scenes.Scene uploaded the way rtgl.SceneUploader uploads it: one
    non-indexed primitive per object, identity transform, opaque colour,
    roughness 1, metallic 0."""

    def __init__(self, o: Oracle, scene, width: int, height: int, leaf: int = 3, flags: dict | None = None,
                 textures: dict | None = None, texcoords: dict | None = None, texture_pixels: dict | None = None,
                 emissive: dict | None = None):
        """flags, textures (object index -> texture name) and texcoords
        (object index -> (tris * 3, 2) floats) as rtgl.SceneUploader takes
        them; texture_pixels: name -> (rgba8, filter, address u, address v)
        as rgProvideOriginalTexture got them.  Refracting objects follow the
        others, in upload order, as the frame assembles them."""
        self.o, self.scene = o, scene
        flags = dict(flags or {})
        textures, texcoords = dict(textures or {}), dict(texcoords or {})
        names = sorted(texture_pixels or {})
        # the _orm and _n maps are UNORM, the rest sRGB (as the library stores them)
        self.tex_objs = [OracleTexture(o, *texture_pixels[nm], srgb=not nm.endswith(("_orm", "_n"))) for nm in names]
        self.tex_ptrs = (_P * max(1, len(names)))(*[t.h for t in self.tex_objs])
        tex_id = {nm: i + 1 for i, nm in enumerate(names)}
        self.h = o.l.or_pt_create(width, height)
        objs = list(enumerate(scene.objects))
        objs = sorted(objs, key=lambda ob: frame_class(flags.get(ob[0], 0)))  # stable: upload order within a class
        order = np.concatenate([np.arange(first, first + count) for _, (_, first, count) in objs]) if objs else np.zeros(0, np.int64)
        n = len(order)
        self.tris = np.ascontiguousarray(scene.tris[order], np.float32)
        normals = []
        for oid, (_, first, count) in objs:
            # flat normals unless the primitive opts out (GeomInfoManager.cpp:173-176)
            if flags.get(oid, 0) & DONT_GENERATE_NORMALS:
                normals.append(scene.normals[first:first + count])
            else:
                normals.append(flat_normals(o, scene.tris[first:first + count]))
        self.normals = np.ascontiguousarray(np.concatenate(normals) if normals else np.zeros((0, 9)), np.float32)
        self.bvh = OracleBvh(o, self.tris, leaf)
        self.tri_geom = np.zeros(n, np.uint32)
        self.texcoords = np.zeros((max(1, n), 6), np.float32)
        self.geoms = (OrGeomC * max(1, len(scene.objects)))()
        self.ids = np.zeros(max(1, len(scene.objects)), np.uint64)
        pos, opaque = 0, 0
        for g, (oid, (_, first, count)) in enumerate(objs):
            self.tri_geom[pos:pos + count] = g
            gr = self.geoms[g]
            for k in range(12):
                gr.prevFromCur[k] = 1.0 if k % 5 == 0 else 0.0
            gr.isStatic, gr.color = 1, int(scene.colors[first]) | 0xFF000000
            gr.emissive, gr.roughness, gr.metallic = min(max(float((emissive or {}).get(oid, 0.0)), 0.0), 1.0), 1.0, 0.0
            gr.flags, gr.firstTri, gr.triCount = flags.get(oid, 0), pos, count
            base = textures.get(oid)
            if base:
                gr.texture = tex_id.get(base, 0)
                gr.texOrm, gr.texNormal, gr.texEmissive = (tex_id.get(base + sfx, 0) for sfx in ("_orm", "_n", "_e"))
            if oid in texcoords:
                self.texcoords[pos:pos + count] = np.asarray(texcoords[oid], np.float32).reshape(count, 6)
            self.ids[g] = oid << 32
            pos += count
        self.c = OrPtSceneC(self.tris.ctypes.data, self.normals.ctypes.data, self.tri_geom.ctypes.data, n,
                            C.cast(self.geoms, _P), self.ids.ctypes.data, len(scene.objects), None, None, 0, None, 0, None, 0,
                            C.cast(self.tex_ptrs, _P) if names else None, len(names), self.texcoords.ctypes.data)
        self.scene_lights(scene)

    def scene_lights(self, scene):
        """Uses scene.lights for the next frames (same uniqueIDs as SceneUploader)."""
        from raytracedgl1_amd import rtgl
        lights = scene.lights or {}
        l = rtgl.api()
        self.sun = None
        if lights.get("sun"):
            sun = lights["sun"]
            d = rtgl.RgDirectionalLightUploadInfo()
            d.uniqueID = 1
            d.color = l.rgUtilPackColorFloat4D(*sun["color"], 1.0)
            d.intensity = sun["intensity"]
            d.direction.data[:] = sun["direction"]
            d.angularDiameterDegrees = sun["angularDiameterDegrees"]
            self.sun = d
        sph = lights.get("spheres", [])
        self.spheres = (rtgl.RgSphericalLightUploadInfo * max(1, len(sph)))()
        for i, sp in enumerate(sph):
            s = self.spheres[i]
            s.uniqueID = 100 + i
            s.color = l.rgUtilPackColorFloat4D(*sp["color"], 1.0)
            s.intensity = sp["intensity"]
            s.position.data[:] = sp["position"]
            s.radius = sp["radius"]
        spots, polys = lights.get("spots", []), lights.get("polys", [])
        self.spots = (rtgl.RgSpotLightUploadInfo * max(1, len(spots)))(*[rtgl.spot_info(sp, 200 + i) for i, sp in enumerate(spots)])
        self.polys = (rtgl.RgPolygonalLightUploadInfo * max(1, len(polys)))(*[rtgl.poly_info(pl, 300 + i) for i, pl in enumerate(polys)])
        self.c.dirLight = C.cast(C.pointer(self.sun), _P) if self.sun else None
        self.c.spheres = C.cast(self.spheres, _P)
        self.c.sphereCount = len(sph)
        self.c.spots, self.c.spotCount = C.cast(self.spots, _P), len(spots)
        self.c.polys, self.c.polyCount = C.cast(self.polys, _P), len(polys)

    def frame(self, params, fb: "OracleFramebuffers", threads: int = 8):
        self.o.l.or_pt_frame(self.h, C.byref(self.c), self.bvh.h, C.byref(params), C.byref(fb.c), threads)

    def __del__(self):
        try:
            self.o.l.or_pt_destroy(self.h)
        except Exception:
            pass
