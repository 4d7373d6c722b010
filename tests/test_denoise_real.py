"""A-SVGF on rendered inputs: the planes of a frame rendered through the C
API are given to the GPU denoiser passes (rgExtDenoiseTemporal,
rgExtDenoiseAntiFirefly, rgExtDenoiseVariance, rgExtDenoiseAtrous on a
standalone framebuffer set) and to the FP64 oracle's passes, and the
denoised radiance PRE_FINAL is compared.

Workloads: the 640x360 terrain-and-boxes frame (directional sun, 1 spp) and
the 1920x1080 terrain-and-icospheres orbit frame (64 sphere lights, ReSTIR
DI 32 candidates / 1 temporal / 1 spatial, two bounces), each the third
frame of a moving camera, with 5 a-trous iterations.

The inputs of frame N's denoiser come from two readbacks of the instance.
After frame N: the traced planes of frame N (G-buffer, unfiltered radiance,
view directions, the filtered gradient the temporal pass reads) and the
previous-frame planes (history length, moments, accumulated specular and
indirect, normals, depths, positions of frame N-1).  After frame N-1:
DIFF_COLOR_HISTORY, the diffuse history that frame N's temporal pass reads
before its first a-trous iteration rewrites it.

Bound: PRE_FINAL within 1e-3 relative (|got - ref| / (|ref| + 1e-6 max|ref|))
on every pixel outside the tie class, the pixels whose FP64 result FP32
arithmetic cannot be expected to reproduce:
- ill-conditioned pixels: the oracle chain is run again on inputs whose
  every float value is moved by 4 FP32 ulps (relative 2^-22, random signs,
  three trials) and whose reprojected positions are moved by one FP32 ulp:
  the position ((x + 0.5) / (W/2) + motion) * W/2 is a number near 1000
  that FP32 holds to 6e-5 .. 1.2e-4 pixels whatever the motion's own
  precision, and next to a firefly of the previous frame (the accumulated
  specular and indirect planes are unfiltered) a bilinear weight that far
  off moves the result by 0.5%.  A pixel whose PRE_FINAL moves by more
  than 2.5e-4 relative (a quarter of the bound) under these moves is in
  the class.  This covers every branch decided on computed values (the
  reprojection depth test against 0.1, the bilinear tap position and
  weights, the accepted weight against 1e-4, the history length against
  the variance pass's 4) and ill-conditioned arithmetic, at every pixel the
  decision reaches through the later passes.
  The previous frame's packed normals are moved by one code of their
  azimuth in the same runs: they only enter the reprojection tests
  (CmSVGFTemporalAccumulation.comp testReprojectedNormal, dot > 0.95), and
  one code (1e-4 rad) flips every decision whose dot lies within about
  1e-4 of the threshold, a hundred times the FP32 decode's error, so the
  pixels those decisions reach are in the class too.
The class's size is printed and bounded; the count of taps whose FP64 dot
lies within 1e-5 of 0.95 is printed beside it.
"""
import math

import numpy as np
import pytest

from raytracedgl1_amd import ffi, rtgl, scenes
from raytracedgl1_amd.ffi import PLANE, FramebufferSet
from denoise_inputs import fill_planes
from oracle_lib import OracleFramebuffers

pytestmark = pytest.mark.gpu

TOL = 1e-3
TIE_EPS = 1e-5
PERTURB = 2.0 ** -22   # 4 FP32 ulps, relative
SENSITIVITY = 2.5e-4   # PRE_FINAL moves more than this: ill-conditioned

# what the temporal .. a-trous passes read: frame N's planes ...
CUR = ("NORMAL", "METALLIC_ROUGHNESS", "DEPTH_WORLD", "DEPTH_GRAD", "MOTION", "UNFILTERED_DIRECT", "UNFILTERED_SPECULAR",
       "UNFILTERED_INDIR", "SURFACE_POSITION", "VIEW_DIRECTION", "THROUGHPUT", "IS_SKY", "ALBEDO", "DIS_PING_GRADIENT")
# ... and frame N-1's
PREV = ("NORMAL", "DEPTH_WORLD", "SURFACE_POSITION", "ACCUM_HISTORY_LENGTH", "DIFF_ACCUM_MOMENTS", "SPEC_ACCUM_COLOR",
        "INDIR_ACCUM")


def _workload(cfg):
    if cfg == "terrain_boxes_640x360":
        cam0 = scenes.terrain_boxes_camera()
        cams = [scenes.Camera(cam0.eye + np.array([0.05 * f, 0, 0.03 * f]), cam0.target + np.array([0.05 * f, 0, 0.03 * f]))
                for f in range(3)]
        return 640, 360, scenes.terrain_boxes_scene(), cams, None
    cams = [scenes.orbit_camera(f) for f in range(3)]
    return 1920, 1080, scenes.terrain_spheres_scene(), cams, (32, 1, 1)


def render_inputs(cfg):
    """Renders the workload's frames; returns (width, height, cur planes,
    prev planes, frame N-1's DIFF_COLOR_HISTORY, frame N's parameters)."""
    w, h, scene, cams, restir = _workload(cfg)
    r = rtgl.Renderer(w, h)
    try:
        r.set_static_geometry(True)  # geometry uploaded with the first frame stays resident
        r.path.atrousIterations = 5
        if restir:
            r.path.restirInitialCandidates, r.path.restirTemporalSamples, r.path.restirSpatialSamples = restir
        up = rtgl.SceneUploader(scene)
        hist = None
        for f, cam in enumerate(cams):
            r.frame(up, cam, geometry=(f == 0))
            if f == len(cams) - 2:
                hist = r.framebuffers().download(PLANE["DIFF_COLOR_HISTORY"]).copy()
        fb = r.framebuffers()
        cur = {PLANE[k]: fb.download(PLANE[k]).copy() for k in CUR}
        cur[PLANE["DIFF_COLOR_HISTORY"]] = hist
        prev = {PLANE[k]: fb.download(PLANE[k], prev=True).copy() for k in PREV}
        return w, h, cur, prev, r.frame_params()
    finally:
        r.close()


def oracle_chain(oracle, w, h, cur, prev, p, o=None):
    """The oracle's temporal, anti-firefly, variance and a-trous passes on
    the given planes (into `o` when given)."""
    o = o or OracleFramebuffers(oracle, w, h)
    fill_planes(cur, prev, None, o)
    oracle.l.or_denoise_set_threads(8)
    try:
        o.run("temporal", p)
        o.run("antifirefly", p)
        o.run("variance", p)
        for k in range(p.atrousIterations):
            o.run("atrous", p, k)
    finally:
        oracle.l.or_denoise_set_threads(1)
    return o


def denoise_both(oracle, w, h, cur, prev, p):
    """Runs temporal, anti-firefly, variance and the a-trous iterations on
    both sides; returns (GPU PRE_FINAL, oracle framebuffers)."""
    g = FramebufferSet(w, h)
    fill_planes(cur, prev, g, None)
    ffi.call("rgExtDenoiseTemporal", g.h, p, None)
    ffi.call("rgExtDenoiseAntiFirefly", g.h, p, None)
    ffi.call("rgExtDenoiseVariance", g.h, p, None)
    for k in range(p.atrousIterations):
        ffi.call("rgExtDenoiseAtrous", g.h, p, k, None)
    got = g.download(PLANE["PRE_FINAL"]).copy()
    return got, oracle_chain(oracle, w, h, cur, prev, p)


def _rel(a, b):
    a, b = a.astype(np.float64), b.astype(np.float64)
    return (np.abs(a - b) / (np.abs(b) + 1e-6 * np.abs(b).max() + 1e-30)).max(-1)


def _position_ulps(w, h, motion):
    """One FP32 ulp of each pixel's reprojected position
    ((x + 0.5) / sx + motion) * sx, in UV units (sx = W/2, sy = H)."""
    ys, xs = np.mgrid[0:h, 0:w]
    sx, sy = w / 2.0, float(h)
    px = ((xs + 0.5) / sx + motion[..., 0].astype(np.float64)) * sx
    py = ((ys + 0.5) / sy + motion[..., 1].astype(np.float64)) * sy
    ux = np.spacing(np.abs(px).astype(np.float32)).astype(np.float64) / sx
    uy = np.spacing(np.abs(py).astype(np.float32)).astype(np.float64) / sy
    return np.stack([ux, uy], -1)


def ill_conditioned(oracle, w, h, cur, prev, p, ref, trials=3):
    """PRE_FINAL pixels that move by more than SENSITIVITY (relative) when
    every float input moves by PERTURB relative and the reprojected
    positions by one FP32 ulp, random signs."""
    ulps = _position_ulps(w, h, cur[PLANE["MOTION"]])
    sens = np.zeros(ref.shape[:2])
    o = OracleFramebuffers(oracle, w, h)
    for t in range(trials):
        rng = np.random.default_rng(1000 + t)

        def jitter(planes):
            return {k: (v.astype(np.float64) * (1.0 + PERTURB * rng.choice((-1.0, 1.0), v.shape))
                        if np.issubdtype(v.dtype, np.floating) else v) for k, v in planes.items()}
        moved = jitter(cur)
        mo = moved[PLANE["MOTION"]]
        mo[..., :2] += ulps * rng.choice((-1.0, 1.0), ulps.shape)
        moved_prev = jitter(prev)
        code = prev[PLANE["NORMAL"]].astype(np.int64)
        step = rng.choice((-1, 1), code.shape)
        moved_prev[PLANE["NORMAL"]] = ((code & ~0xFFFF) | ((code + step) & 0xFFFF)).astype(prev[PLANE["NORMAL"]].dtype)
        oracle_chain(oracle, w, h, moved, moved_prev, p, o)
        sens = np.maximum(sens, _rel(o.cur[PLANE["PRE_FINAL"]][..., :3], ref))
    return sens > SENSITIVITY


def _decode_normal(packed):
    """decodeNormal (Utils.h:126-140): 16-bit polar and azimuth angles."""
    packed = np.asarray(packed, np.uint64)
    phi = (packed >> 16).astype(np.float64) / 65535.0 * math.pi
    theta = (packed & 0xFFFF).astype(np.float64) / 32768.0 * math.pi
    return np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)], -1)


def _normal_ties(w, h, cur, prev):
    """Checkerboard pixels with a reprojection tap whose decoded normals'
    dot lies within TIE_EPS of the threshold 0.95 (FP64 decode)."""
    mo = cur[PLANE["MOTION"]].astype(np.float64)
    n = _decode_normal(cur[PLANE["NORMAL"]][..., 0])
    npv = _decode_normal(prev[PLANE["NORMAL"]][..., 0])
    ys, xs = np.mgrid[0:h, 0:w]
    sx, sy = w / 2.0, float(h)
    px = ((xs + 0.5) / sx + mo[..., 0]) * sx
    py = ((ys + 0.5) / sy + mo[..., 1]) * sy
    ppx, ppy = np.floor(px - 0.5).astype(np.int64), np.floor(py - 0.5).astype(np.int64)
    near = np.zeros((h, w), bool)
    for dy in (0, 1):
        for dx in (0, 1):
            qx, qy = ppx + dx, ppy + dy
            ok = (qx >= 0) & (qx < w) & (qy >= 0) & (qy < h)
            qxc, qyc = np.clip(qx, 0, w - 1), np.clip(qy, 0, h - 1)
            near |= ok & (np.abs((n * npv[qyc, qxc]).sum(-1) - 0.95) <= TIE_EPS)
    return near


@pytest.mark.parametrize("cfg", ["terrain_boxes_640x360", "terrain_spheres_1920x1080"])
def test_denoiser_on_rendered_frames(gpu, oracle, cfg):
    print(f"denoiser on rendered inputs: {cfg}", flush=True)
    w, h, cur, prev, p = render_inputs(cfg)
    got, o = denoise_both(oracle, w, h, cur, prev, p)
    got = got[..., :3].astype(np.float64)
    ref = o.cur[PLANE["PRE_FINAL"]][..., :3].copy()
    scale = np.abs(ref).max()
    err = _rel(got, ref)
    del o
    seeds = _normal_ties(w, h, cur, prev)
    tie = ill_conditioned(oracle, w, h, cur, prev, p, ref)
    print(f"  PRE_FINAL max rel err {err.max():.3e}; outside the tie class {err[~tie].max(initial=0.0):.3e}; "
          f"pixels > {TOL}: {(err > TOL).sum()} ({(err[~tie] > TOL).sum()} outside); tie class {tie.mean():.4f} of the "
          f"frame; reprojection taps with |dot - 0.95| <= {TIE_EPS}: {seeds.sum()}", flush=True)
    # the frame has surfaces and light: most pixels are lit
    assert np.isfinite(got).all() and scale > 0.0 and np.mean(ref.max(-1) > 1e-3 * scale) > 0.5
    assert tie.mean() < 0.05
    assert err[~tie].max(initial=0.0) <= TOL
