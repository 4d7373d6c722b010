"""Whole frames through the RTGL1 C API (rgCreateInstance / rgStartFrame /
rgUploadMeshPrimitive / rgUpload*Light / rgDrawFrame) on the GPU.

Follows the protocol checks of Source/RTGL1.cpp:31-194 and
Source/VulkanDevice.cpp:828-847, and the frame of VulkanDevice::Render
(:540-588) as rebuilt here: scene assembly, LBVH, ray-traced passes, A-SVGF,
eye adaptation.  Primary visibility is checked bit-exact against the oracle's
traversal of the same rays; whole frames are checked for determinism,
static-geometry equivalence and sanity of every statistic.
"""
import ctypes as C

import numpy as np
import pytest

from raytracedgl1_amd import ffi, rtgl, scenes
from raytracedgl1_amd.ffi import PLANE, RAY_DTYPE, RG_EXT_MISS

pytestmark = pytest.mark.gpu

RG_RESULT_WRONG_INSTANCE = 3
RG_RESULT_ALREADY_INITIALIZED = 4
RG_RESULT_FRAME_WASNT_STARTED = 8
RG_RESULT_FRAME_WASNT_ENDED = 9
RG_RESULT_WRONG_FUNCTION_ARGUMENT = 11

W1, H1 = 640, 360


@pytest.fixture(scope="module")
def boxes():
    s = scenes.terrain_boxes_scene()
    return s, rtgl.SceneUploader(s)


def _render(uploader, cams, width=W1, height=H1, static=False, collect=False, precision=0):
    r = rtgl.Renderer(width, height)
    try:
        if static:
            r.set_static_geometry(True)
        r.path.collectTraversalStats = 1 if collect else 0
        r.path.precisionFlags = precision
        stats = [r.frame(uploader, c, geometry=(not static) or i == 0) for i, c in enumerate(cams)]
        fb = r.framebuffers()
        out = {name: fb.download(PLANE[name]) for name in ("PRE_FINAL", "VISIBILITY_BUFFER", "UNFILTERED_DIRECT",
                                                             "UNFILTERED_INDIR", "DEPTH_WORLD", "MOTION")}
        out["params"] = r.frame_params()
        out["histogram"] = r.histogram()
        return stats, out
    finally:
        r.close()


def _primary_rays_fp32(p, width, height):
    """k_primary_rays restated in float32 with the same operation order
    (getRayDir, ShaderCommonGLSLFunc.h:460-468)."""
    f = np.float32
    ip = np.array(p.invProjection[:], f)
    iv = np.array(p.invView[:], f)
    x, y = np.meshgrid(np.arange(width, dtype=f), np.arange(height, dtype=f))
    u = (x + f(0.5) + f(p.jitterX)) / f(width)
    v = (y + f(0.5) + f(p.jitterY)) / f(height)
    u = u * f(2) - f(1)
    v = v * f(2) - f(1)
    one = f(1)
    t = [ip[0 + i] * u + ip[4 + i] * v + ip[8 + i] * one + ip[12 + i] * one for i in range(3)]
    tw = ip[3] * u + ip[7] * v + ip[11] * one + ip[15] * one
    div = np.abs(tw) >= f(0.001)
    loc = [np.where(div, c / np.where(div, tw, one), c) for c in t]
    ln = np.sqrt(loc[0] * loc[0] + loc[1] * loc[1] + loc[2] * loc[2])
    n = [c / ln for c in loc]
    zero = f(0)
    d = [iv[0 + i] * n[0] + iv[4 + i] * n[1] + iv[8 + i] * n[2] + iv[12 + i] * zero for i in range(3)]
    rays = np.zeros(width * height, RAY_DTYPE)
    rays["origin"] = np.array(p.cameraPosition[:3], f)
    rays["direction"] = np.stack([c.ravel() for c in d], axis=1)
    rays["tmin"] = p.primaryRayMinDist
    rays["tmax"] = p.rayLength
    return rays


def _regular_from_checkerboard(plane, width):
    """Reorders a checkerboard-layout plane (ShaderCommonGLSLFunc.h:306-357)
    into regular pixel order."""
    h = plane.shape[0]
    x, y = np.meshgrid(np.arange(width), np.arange(h))
    odd = (x + y % 2) % 2
    cx = odd * (width // 2) + x // 2
    return plane[y, cx]


def test_primary_visibility_bit_exact(boxes, oracle):
    print("api: primary visibility terrain-boxes 640x360", flush=True)
    s, up = boxes
    _, out = _render(up, [scenes.terrain_boxes_camera()])
    p = out["params"]
    assert (p.jitterX, p.jitterY) == (0.0, 0.0)
    rays = _primary_rays_fp32(p, W1, H1)
    hits = oracle.bvh(s.tris, 3).trace(rays, 0)
    vis = _regular_from_checkerboard(out["VISIBILITY_BUFFER"], W1).reshape(-1, 4)
    hit = vis[:, 0] == 1
    assert np.array_equal(hit, hits["prim"] != RG_EXT_MISS)
    # geometry index -> first triangle of the uploaded object (upload order)
    first = np.array([o[1] for o in s.objects], np.uint32)
    prim = first[vis[hit, 1] & 0xFFF] + (vis[hit, 1] >> 12)
    assert np.array_equal(prim, hits["prim"][hit])
    assert np.array_equal(vis[hit, 2].view(np.float32), hits["u"][hit])
    assert np.array_equal(vis[hit, 3].view(np.float32), hits["v"][hit])
    assert 0.2 < hit.mean() < 1.0


def test_frame_statistics_and_outputs_are_sane(boxes):
    print("api: frame stats terrain-boxes", flush=True)
    s, up = boxes
    cams = [scenes.terrain_boxes_camera()] * 3
    stats, out = _render(up, cams, collect=True)
    for i, st in enumerate(stats):
        assert st.frameId == i
        assert (st.renderWidth, st.renderHeight) == (W1, H1)
        assert st.triangleCount == len(s.tris)
        assert st.bvhRebuilt == (1 if i == 0 else 0)
        assert st.raysTraced >= 5 * W1 * H1
        assert st.traversalLaunches == 7 and st.traversalErrors == 0  # 6 ray sets + gradient samples
        assert 0.0 < st.gpuMsTraversal <= st.gpuMsTrace <= st.gpuMsTotal
        assert st.nodeVisits >= st.raysTraced // 2 and st.triangleTests > 0
        assert st.avgLuminance > 0.0
    pre = out["PRE_FINAL"][..., :3]
    assert np.isfinite(pre).all() and (pre >= 0).all() and pre.max() > 0
    assert out["histogram"].sum() == W1 * H1
    assert np.isfinite(out["UNFILTERED_DIRECT"]).all() and np.isfinite(out["UNFILTERED_INDIR"]).all()
    # a still camera has (almost) zero screen-space motion
    assert np.abs(out["MOTION"][..., :2]).max() < 1e-4


def test_frames_are_deterministic_and_stats_do_not_change_them(boxes):
    print("api: determinism terrain-boxes", flush=True)
    _, up = boxes
    cams = [scenes.terrain_boxes_camera()] * 3
    _, a = _render(up, cams)
    _, b = _render(up, cams, collect=True)
    for k in ("PRE_FINAL", "VISIBILITY_BUFFER", "UNFILTERED_DIRECT", "UNFILTERED_INDIR", "histogram"):
        assert np.array_equal(a[k], b[k]), k


def test_static_geometry_matches_per_frame_upload(boxes):
    print("api: static geometry terrain-boxes", flush=True)
    _, up = boxes
    cams = [scenes.terrain_boxes_camera()] * 3
    sa, a = _render(up, cams)
    sb, b = _render(up, cams, static=True)
    for k in ("PRE_FINAL", "VISIBILITY_BUFFER", "UNFILTERED_DIRECT", "UNFILTERED_INDIR", "histogram"):
        assert np.array_equal(a[k], b[k]), k
    assert all(st.bvhRebuilt == 0 for st in sb[1:])
    assert all(st.triangleCount == sa[0].triangleCount for st in sb)


def test_moving_camera_has_motion_and_finite_output():
    print("api: terrain-spheres orbit 480x270", flush=True)
    s = scenes.terrain_spheres_scene()
    up = rtgl.SceneUploader(s)
    cams = [scenes.orbit_camera(f) for f in range(3)]
    stats, out = _render(up, cams, width=480, height=270, static=True)
    assert stats[0].triangleCount == len(s.tris)
    assert np.isfinite(out["PRE_FINAL"]).all()
    assert np.abs(out["MOTION"][..., :2]).max() > 0.0
    assert out["histogram"].sum() == 480 * 270


def test_protocol_errors(boxes):
    print("api: protocol errors", flush=True)
    l = rtgl.api()
    r = rtgl.Renderer(64, 32)
    try:
        h = C.c_void_p(r.h)
        di = rtgl.RgDrawFrameInfo()
        assert l.rgDrawFrame(h, C.byref(di)) == RG_RESULT_FRAME_WASNT_STARTED
        sfi = rtgl.RgStartFrameInfo()
        assert l.rgStartFrame(h, C.byref(sfi)) == 0
        assert l.rgStartFrame(h, C.byref(sfi)) == RG_RESULT_FRAME_WASNT_ENDED
        # a second instance is refused while one exists (singleton)
        ci = rtgl.RgInstanceCreateInfo()
        h2 = C.c_void_p()
        assert l.rgCreateInstance(C.byref(ci), C.byref(h2)) == RG_RESULT_ALREADY_INITIALIZED
        assert l.rgStartFrame(C.c_void_p(r.h + 8), C.byref(sfi)) == RG_RESULT_WRONG_INSTANCE
        # indices past the vertex range are rejected before anything is copied
        v = np.zeros(3, rtgl.VERTEX_DTYPE)
        idx = np.array([0, 1, 7], np.uint32)
        mesh = rtgl.RgMeshInfo()
        mesh.transform = rtgl.identity_transform()
        prim = rtgl.RgMeshPrimitiveInfo()
        prim.pVertices, prim.vertexCount = v.ctypes.data, 3
        prim.pIndices, prim.indexCount = idx.ctypes.data, 3
        prim.color = 0xFFFFFFFF
        assert l.rgUploadMeshPrimitive(h, C.byref(mesh), C.byref(prim)) == RG_RESULT_WRONG_FUNCTION_ARGUMENT
        # an empty frame (no geometry, no lights) renders sky
        di.view[:] = scenes.terrain_boxes_camera().view().tolist()
        di.fovYRadians, di.cameraNear, di.cameraFar, di.rayLength = 1.3, 0.1, 1000.0, 1000.0
        di.pParams = C.cast(C.pointer(r.sky), C.c_void_p)
        assert l.rgDrawFrame(h, C.byref(di)) == 0
        assert r.stats().triangleCount == 0
    finally:
        r.close()
    # after destruction the handle is no longer valid, and a new instance can be made
    assert l.rgStartFrame(C.c_void_p(h.value), None) == RG_RESULT_WRONG_INSTANCE
    r2 = rtgl.Renderer(64, 32)
    r2.close()


def _render_seq(uploader_seq, cams, width=W1, height=H1):
    """Frames with a different uploader per frame (geometry or lights change)."""
    r = rtgl.Renderer(width, height)
    try:
        stats = [r.frame(up, c) for up, c in zip(uploader_seq, cams)]
        fb = r.framebuffers()
        out = {name: fb.download(PLANE[name]) for name in ("DIS_GRADIENT_HISTORY", "IS_SKY", "INDIR_INITIAL_RADIANCE",
                                                             "INDIR_RESERVOIR_WEIGHT", "INDIR_RESERVOIR_SAMPLE",
                                                             "UNFILTERED_INDIR", "GRADIENT_PREV_PIX")}
        return stats, out
    finally:
        r.close()


def test_restir_gi_reservoirs_are_well_formed(boxes, oracle):
    print("api: ReSTIR GI reservoirs terrain-boxes", flush=True)
    _, up = boxes
    _, out = _render_seq([up] * 4, [scenes.terrain_boxes_camera()] * 4)
    sky = out["IS_SKY"][..., 0] != 0
    w = out["INDIR_RESERVOIR_WEIGHT"]
    M = w[..., 3]
    wsum = w[..., 2].view(np.float32)
    tp = w[..., 1].view(np.float32)
    # sky pixels store the empty reservoir; others carry at least their own sample
    assert (M[sky] == 0).all() and (wsum[sky] == 0).all()
    assert (M[~sky] >= 1).all()
    # temporal history is renormalised to 20 before adding 1 initial + at most 7 spatial
    assert M.max() <= 20 + 1 + 7 * 21
    assert np.isfinite(wsum).all() and (wsum >= 0).all() and (tp >= 0).all()
    ind = out["UNFILTERED_INDIR"][..., :3]
    assert np.isfinite(ind).all() and (ind >= 0).all() and ind[~sky].max() > 0
    # the kernels' E5B9G9R9 encodings are canonical: re-encoding their decoded
    # value with the oracle's EXT_texture_shared_exponent encoder reproduces them
    codes = np.unique(np.concatenate([out["INDIR_INITIAL_RADIANCE"][..., 0].ravel(), w[..., 0].ravel()]))
    codes = codes[:: max(1, len(codes) // 4000)]
    for c in codes:
        assert oracle.encode_e5(oracle.decode_e5(int(c))) == int(c), hex(int(c))


def test_gradients_stay_low_when_static_and_rise_when_light_goes(boxes):
    print("api: A-SVGF gradient samples terrain-boxes", flush=True)
    s, up = boxes
    dark = rtgl.SceneUploader(scenes.Scene(s.tris, s.normals, s.colors, s.objects, {}))
    cam = scenes.terrain_boxes_camera()
    _, still = _render_seq([up] * 5, [cam] * 5)
    _, lost = _render_seq([up] * 4 + [dark], [cam] * 5)
    g_still = still["DIS_GRADIENT_HISTORY"][..., 0]
    g_lost = lost["DIS_GRADIENT_HISTORY"][..., 0]
    assert np.isfinite(g_still).all() and (g_still >= 0).all() and (g_still <= 1).all()
    # a static frame re-shades each stratum's sample to (almost) its old value
    assert np.quantile(g_still, 0.99) < 0.05
    # removing the sun makes the forward-projected direct light vanish
    assert g_lost.mean() > 5 * max(g_still.mean(), 1e-4) and g_lost.max() > 0.2
    # stratum positions are 0..8 or 10 (rejected)
    pp = still["GRADIENT_PREV_PIX"][..., 0]
    assert np.isin(pp, list(range(9)) + [10]).all()


def test_lightstyles_scale_lights_one_frame_late(boxes):
    """RgLightExtraInfo::lightstyle picks a multiplier from the last
    rgDrawFrame's RgDrawFrameIlluminationParams::pLightstyleValues
    (CalculateLightStyle, Source/LightManager.cpp:366-383; the values are set
    in DrawFrame after the frame's lights were encoded, VulkanDevice.cpp:504).
    A styled sun at 0.5 renders exactly as an unstyled sun of half the
    intensity, starting with the frame after the values were first given."""
    print("api: light styles terrain-boxes 160x90", flush=True)
    s, _ = boxes
    sun = dict(s.lights["sun"])
    cam = scenes.terrain_boxes_camera()
    planes = ("UNFILTERED_DIRECT", "PRE_FINAL", "FINAL")

    def run(suns, styles):
        r = rtgl.Renderer(160, 90)
        outs = []
        try:
            if styles is not None:
                r.set_lightstyles(styles)
            for sn in suns:
                r.frame(rtgl.SceneUploader(scenes.Scene(s.tris, s.normals, s.colors, s.objects, {"sun": sn})), cam)
                fb = r.framebuffers()
                outs.append({k: fb.download(PLANE[k]).copy() for k in planes})
        finally:
            r.close()
        return outs

    styled = run([dict(sun, lightstyle=0)] * 3, [0.5, 3.0])
    half = dict(sun, intensity=sun["intensity"] * 0.5)
    plain = run([sun, half, half], None)
    for f in range(3):
        for k in planes:
            assert np.array_equal(styled[f][k], plain[f][k]), (f, k)
    # an index outside the table is no style
    out_of_range = run([dict(sun, lightstyle=7)] * 2, [0.5])
    full = run([sun] * 2, None)
    assert np.array_equal(out_of_range[1]["UNFILTERED_DIRECT"], full[1]["UNFILTERED_DIRECT"])


def test_fp16_precision_form(boxes):
    """FP16 nodes change only the order boxes are visited in, so frames are
    bit-identical to the FP32 path's; FP16 storage of the denoiser planes
    rounds the temporal and a-trous buffers, so the denoised image stays
    within a few 1e-3."""
    print("api: FP16 nodes / FP16 storage vs FP32, terrain-boxes", flush=True)
    _, up = boxes
    cams = [scenes.terrain_boxes_camera()] * 3
    _, ref = _render(up, cams)
    _, nodes16 = _render(up, cams, precision=ffi.RG_EXT_PRECISION_FP16_NODES)
    for k in ("PRE_FINAL", "VISIBILITY_BUFFER", "UNFILTERED_DIRECT", "UNFILTERED_INDIR", "DEPTH_WORLD"):
        assert np.array_equal(nodes16[k], ref[k]), k
    _, both = _render(up, cams, precision=ffi.RG_EXT_PRECISION_FP16_NODES | ffi.RG_EXT_PRECISION_FP16_STORAGE)
    assert np.array_equal(both["VISIBILITY_BUFFER"], ref["VISIBILITY_BUFFER"])
    g, e = both["PRE_FINAL"][..., :3].astype(np.float64), ref["PRE_FINAL"][..., :3].astype(np.float64)
    rel = (np.abs(g - e) / (np.abs(e) + 1e-3 * np.abs(e).mean())).max(axis=-1)
    print(f"FP16 storage frame vs FP32: PRE_FINAL rel p50 {np.median(rel):.2e} p99 {np.quantile(rel, 0.99):.2e} "
          f"max {rel.max():.2e}", flush=True)
    assert np.quantile(rel, 0.99) < 1e-2
    assert abs(g.mean() / e.mean() - 1.0) < 1e-3


def test_moving_object_motion_vectors(boxes):
    """A primitive whose transform changes between frames (matched by
    uniqueObjectID and primitive index, UniqueID::MakeForPrimitive) gets
    motion vectors of its own movement: the pixel's current world position
    moved back by the previous transform, projected with the previous camera
    (HitInfo.inl:303-313); static geometry under a static camera has none."""
    print("api: motion vectors of a moving object", flush=True)
    s, _ = boxes
    up = rtgl.SceneUploader(s)
    cam = scenes.terrain_boxes_camera()
    W, H = 320, 180
    r = rtgl.Renderer(W, H)
    try:
        r.frame(up, cam)
        # move the box that covers the most pixels
        fb0 = r.framebuffers()
        # visibility buffer .y: local primitive << 12 | geometry index, and
        # the uploader makes object k geometry k
        v0 = fb0.download(PLANE["VISIBILITY_BUFFER"])[..., 1]
        hit0 = fb0.download(PLANE["IS_SKY"])[..., 0] == 0
        owner = (v0[hit0] & 0xFFF).astype(np.int64)
        cnt = np.bincount(owner, minlength=len(s.objects))
        cnt[0] = 0  # not the terrain
        moved = int(np.argmax(cnt))
        shift = [[1, 0, 0, 0.5], [0, 1, 0, 0.0], [0, 0, 1, 0.25]]
        l = rtgl.api()
        ffi.check("rgStartFrame", l.rgStartFrame(r.h, C.byref(rtgl.RgStartFrameInfo())))
        up.upload(r.h, True, transforms={moved: shift})
        di = rtgl.draw_info(cam, r.time + 1.0 / 60.0, C.cast(C.pointer(r.sky), C.c_void_p))
        ffi.check("rgDrawFrame", l.rgDrawFrame(r.h, C.byref(di)))
        p = r.frame_params()
        fb = r.framebuffers()
        vis, mot, pos = (fb.download(PLANE[k]) for k in ("VISIBILITY_BUFFER", "MOTION", "SURFACE_POSITION"))
        sky = fb.download(PLANE["IS_SKY"])[..., 0] != 0
    finally:
        r.close()
    on = (~sky) & ((vis[..., 1] & 0xFFF) == moved)
    off = (~sky) & ~on
    assert on.sum() > 20
    assert np.abs(mot[off][:, :2]).max() < 1e-6
    view, proj = scenes.mat_cm(p.view), scenes.mat_cm(p.projection)
    vprev, pprev = scenes.mat_cm(p.viewPrev), scenes.mat_cm(p.projectionPrev)
    P = pos[on][:, :3].astype(np.float64)
    prev = P - np.array([0.5, 0.0, 0.25])

    def screen(m_v, m_p, x):
        h = np.concatenate([x, np.ones((len(x), 1))], 1) @ m_v.T @ m_p.T
        return h[:, :2] / h[:, 3:4] * 0.5 + 0.5

    expect = screen(vprev, pprev, prev) - screen(view, proj, P)
    assert np.abs(expect).max() > 1e-3  # the object did move on screen
    assert np.abs(mot[on][:, :2] - expect).max() < 2e-5


def test_upload_order_does_not_change_the_image(boxes):
    """The same primitives uploaded in reverse order give the same frame:
    only triangle and geometry numbering change, and closest hits break
    exact ties by primitive id, so at most a few edge pixels may differ."""
    print("api: upload order invariance", flush=True)
    s, up = boxes
    cam = scenes.terrain_boxes_camera()
    out = []
    for order in (None, list(range(len(s.objects)))[::-1]):
        r = rtgl.Renderer(320, 180)
        try:
            for f in range(2):
                l = rtgl.api()
                ffi.check("rgStartFrame", l.rgStartFrame(r.h, C.byref(rtgl.RgStartFrameInfo())))
                up.upload(r.h, True, order=order)
                r.time += 1.0 / 60.0
                di = rtgl.draw_info(cam, r.time, C.cast(C.pointer(r.sky), C.c_void_p))
                ffi.check("rgDrawFrame", l.rgDrawFrame(r.h, C.byref(di)))
            fb = r.framebuffers()
            out.append({k: fb.download(PLANE[k]).copy() for k in ("PRE_FINAL", "DEPTH_WORLD", "UNFILTERED_DIRECT")})
        finally:
            r.close()
    a, b = out
    for k in ("PRE_FINAL", "DEPTH_WORLD", "UNFILTERED_DIRECT"):
        same = np.all(a[k] == b[k], axis=-1) if a[k].ndim == 3 else a[k] == b[k]
        assert same.mean() >= 0.999, (k, same.mean())
    assert abs(a["PRE_FINAL"][..., :3].mean() / b["PRE_FINAL"][..., :3].mean() - 1.0) < 1e-4


def test_fp16_nodes_with_refracting_split_and_alpha_test():
    """The frame BVH's top node over the opaque and refracting subtrees and
    the alpha-testing traversal kernels also exist in the FP16 node format:
    frames with mirror / glass / water boxes and alpha-tested, textured boxes
    are bit-identical to the FP32 nodes' (FP16 boxes only change the visit
    order)."""
    print("api: FP16 nodes with the refracting split and the alpha test", flush=True)
    s = scenes.terrain_boxes_scene()
    kinds = (rtgl.RG_MESH_PRIMITIVE_MIRROR, rtgl.RG_MESH_PRIMITIVE_GLASS, rtgl.RG_MESH_PRIMITIVE_WATER,
             rtgl.RG_MESH_PRIMITIVE_ALPHA_TESTED)
    flags = {i: kinds[i % 4] for i in range(1, len(s.objects))}
    texcoords, textures = {}, {}
    for k, (_, first, count) in enumerate(s.objects):
        pos = s.tris[first:first + count].reshape(-1, 3).astype(np.float64)
        texcoords[k] = np.stack([pos[:, 0] / 4.0, pos[:, 2] / 4.0], 1).astype(np.float32)
        textures[k] = "t"
    j, i = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    px = np.zeros((8, 8, 4), np.uint8)
    px[..., 0], px[..., 1], px[..., 2] = 40 * (i % 6), 30 * (j % 8), 150
    px[..., 3] = np.where((i + j) % 3 == 0, 0, 255)
    up = rtgl.SceneUploader(s, flags, textures, texcoords)
    cams = [scenes.terrain_boxes_camera()] * 2
    out = {}
    for prec in (0, ffi.RG_EXT_PRECISION_FP16_NODES):
        r = rtgl.Renderer(W1, H1)
        try:
            r.provide_texture("t", px, rtgl.RG_SAMPLER_FILTER_NEAREST)
            r.set_reflect_refract(rtgl.reflect_refract_params(maxReflectRefractDepth=3))
            r.path.precisionFlags = prec
            for c in cams:
                r.frame(up, c)
            fb = r.framebuffers()
            out[prec] = {k: fb.download(PLANE[k]).copy() for k in ("VISIBILITY_BUFFER", "THROUGHPUT", "PRE_FINAL", "DEPTH_WORLD")}
        finally:
            r.close()
    chain = np.abs(out[0]["THROUGHPUT"][..., 3]) == 1.0
    assert chain.mean() > 0.01
    for k in out[0]:
        assert np.array_equal(out[0][k], out[ffi.RG_EXT_PRECISION_FP16_NODES][k]), k
