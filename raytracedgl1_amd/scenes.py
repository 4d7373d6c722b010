"""Seeded synthetic scenes and cameras of the benchmark workloads.

All randomness comes from a counter-based SplitMix64 stream (Steele, Lea,
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014), so a
(seed, stream, index) triple gives the same value on every machine and numpy
version.

Scenes are flat arrays of world-space triangles (n, 9) float32 plus per-vertex
normals and per-triangle packed RGBA8 colours, the data an application passes
through rgUploadMeshPrimitive (RTGL1.h:385-423) after applying its transforms.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SEED_TERRAIN_BOXES = 0xC0FFEE01
SEED_TERRAIN_SPHERES = 0xC0FFEE02
SEED_SCENE_BATCH = 0xC0FFEE04
SEED_SERVING = 0xC0FFEE05

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """count 64-bit outputs of SplitMix64 for (seed, stream), starting at offset."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = base + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, count: int, lo=0.0, hi=1.0) -> np.ndarray:
    """count doubles uniform in [lo, hi) (53-bit mantissa)."""
    u = (splitmix64(seed, stream, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


# ----------------------------------------------------------------- geometry

@dataclass
class Scene:
    tris: np.ndarray                 # (n, 9) float32 world-space positions
    normals: np.ndarray              # (n, 9) float32 per-vertex normals
    colors: np.ndarray               # (n,) uint32 packed RGBA8 (little-endian r first)
    objects: list = field(default_factory=list)  # (name, first_tri, tri_count)
    lights: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.tris.shape[0])


def pack_rgba8(r, g, b, a=1.0) -> np.ndarray:
    """RgColor4DPacked32 / rgUtilPackColorFloat4D layout: r in the low byte."""
    c = [np.clip(np.asarray(x, np.float64) * 255.0, 0, 255).astype(np.uint32) for x in (r, g, b, a)]
    return c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24)


def _value_noise(seed: int, stream: int, res: int, x: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Value noise on a (res+1)^2 lattice over [0,1]^2 with smoothstep blending."""
    lat = uniform(seed, stream, (res + 1) * (res + 1)).reshape(res + 1, res + 1)
    fx, fz = x * res, z * res
    ix = np.clip(np.floor(fx).astype(np.int64), 0, res - 1)
    iz = np.clip(np.floor(fz).astype(np.int64), 0, res - 1)
    tx, tz = fx - ix, fz - iz
    sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
    a = lat[iz, ix] * (1 - sx) + lat[iz, ix + 1] * sx
    b = lat[iz + 1, ix] * (1 - sx) + lat[iz + 1, ix + 1] * sx
    return a * (1 - sz) + b * sz


def heightfield(seed: int, cells: int, size: float, hmax: float):
    """(cells x cells) grid over [-size/2, size/2]^2, 5-octave fBm heights in [0, hmax].

    Returns (vertices (cells+1)^2 x 3, heights (cells+1, cells+1), tris, normals)."""
    g = np.linspace(0.0, 1.0, cells + 1)
    X, Z = np.meshgrid(g, g)
    h = np.zeros_like(X)
    amp, res = 1.0, 4
    for octave in range(5):
        h += amp * _value_noise(seed, 100 + octave, res, X, Z)
        amp *= 0.5
        res *= 2
    h = (h - h.min()) / max(h.max() - h.min(), 1e-12) * hmax
    xs = (X - 0.5) * size
    zs = (Z - 0.5) * size
    P = np.stack([xs, h, zs], -1)  # (cells+1, cells+1, 3) indexed [iz, ix]
    # smooth normals from central differences
    dx = size / cells
    hp = np.pad(h, 1, mode="edge")
    gx = (hp[1:-1, 2:] - hp[1:-1, :-2]) / (2 * dx)
    gz = (hp[2:, 1:-1] - hp[:-2, 1:-1]) / (2 * dx)
    N = np.stack([-gx, np.ones_like(h), -gz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    v00 = P[:-1, :-1].reshape(-1, 3); v10 = P[:-1, 1:].reshape(-1, 3)
    v01 = P[1:, :-1].reshape(-1, 3); v11 = P[1:, 1:].reshape(-1, 3)
    n00 = N[:-1, :-1].reshape(-1, 3); n10 = N[:-1, 1:].reshape(-1, 3)
    n01 = N[1:, :-1].reshape(-1, 3); n11 = N[1:, 1:].reshape(-1, 3)
    # counter-clockwise seen from +y: (v00, v01, v10) and (v10, v01, v11)
    t1 = np.concatenate([v00, v01, v10], 1)
    t2 = np.concatenate([v10, v01, v11], 1)
    m1 = np.concatenate([n00, n01, n10], 1)
    m2 = np.concatenate([n10, n01, n11], 1)
    tris = np.stack([t1, t2], 1).reshape(-1, 9)
    nrm = np.stack([m1, m2], 1).reshape(-1, 9)
    return h, tris.astype(np.float32), nrm.astype(np.float32)


def height_at(h: np.ndarray, size: float, x, z):
    cells = h.shape[0] - 1
    fx = np.clip((np.asarray(x) / size + 0.5) * cells, 0, cells - 1e-9)
    fz = np.clip((np.asarray(z) / size + 0.5) * cells, 0, cells - 1e-9)
    ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
    tx, tz = fx - ix, fz - iz
    a = h[iz, ix] * (1 - tx) + h[iz, ix + 1] * tx
    b = h[iz + 1, ix] * (1 - tx) + h[iz + 1, ix + 1] * tx
    return a * (1 - tz) + b * tz


_BOX_FACES = np.array([
    # each face: 4 corner indices (of the 8 corners, bit0=x bit1=y bit2=z), CCW from outside
    [0, 4, 6, 2],  # -x
    [1, 3, 7, 5],  # +x
    [0, 1, 5, 4],  # -y
    [2, 6, 7, 3],  # +y
    [0, 2, 3, 1],  # -z
    [4, 5, 7, 6],  # +z
])
_BOX_NORMALS = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], np.float64)


def boxes(centers: np.ndarray, sizes: np.ndarray):
    """Axis-aligned boxes, 12 triangles each, outward counter-clockwise faces."""
    corners = np.array([[(i >> 0) & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], np.float64) - 0.5
    tris, nrms = [], []
    for f, nrm in zip(_BOX_FACES, _BOX_NORMALS):
        a, b, c, d = (corners[k] for k in f)
        for (p, q, r) in ((a, b, c), (a, c, d)):
            tri = np.concatenate([centers + p * sizes, centers + q * sizes, centers + r * sizes], 1)
            tris.append(tri)
            nrms.append(np.tile(nrm, (len(centers), 3)))
    t = np.stack(tris, 1).reshape(-1, 9)
    n = np.stack(nrms, 1).reshape(-1, 9)
    return t.astype(np.float32), n.astype(np.float32)


def _icosphere_unit(level: int):
    p = (1 + 5 ** 0.5) / 2
    v = [[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
         [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = np.array(v, np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(f, np.int64)
    for _ in range(level):
        a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
        ab, bc, ca = a + b, b + c, c + a
        for m in (ab, bc, ca):
            m /= np.linalg.norm(m, axis=1, keepdims=True)
        T = np.stack([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                      np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
        # re-index (vertices duplicated; only positions matter)
        V = T.reshape(-1, 3)
        F = np.arange(len(V)).reshape(-1, 3)
    return V[F]  # (ntri, 3, 3) unit-sphere triangles, outward CCW


_ICO_CACHE: dict = {}


def icosphere(level: int) -> np.ndarray:
    if level not in _ICO_CACHE:
        _ICO_CACHE[level] = _icosphere_unit(level)
    return _ICO_CACHE[level]


def soup_tris(seed: int, n: int, extent: float = 100.0) -> np.ndarray:
    """n triangles of a random soup in a 2*extent cube: icospheres of
    subdivision 2 (radii U[0.5,3]) for about half, boxes (sizes U[0.2,2])
    for the rest (the batched-scenes workload)."""
    r = np.random.default_rng(seed)
    unit = icosphere(2).astype(np.float32)  # (320, 3, 3)
    ns = (n // 2) // len(unit)
    nb = (n - ns * len(unit)) // 12
    parts = []
    if ns:
        c = r.uniform(-extent, extent, (ns, 1, 1, 3)).astype(np.float32)
        rad = r.uniform(0.5, 3.0, (ns, 1, 1, 1)).astype(np.float32)
        parts.append((unit[None] * rad + c).reshape(-1, 9))
    if nb:
        ub = boxes(np.zeros((1, 3)), np.ones((1, 3)))[0].reshape(12, 3, 3)
        c = r.uniform(-extent, extent, (nb, 1, 1, 3)).astype(np.float32)
        sz = r.uniform(0.2, 2.0, (nb, 1, 1, 3)).astype(np.float32)
        parts.append((ub[None] * sz + c).reshape(-1, 9))
    left = n - sum(len(p) for p in parts)
    if left:
        c = r.uniform(-extent, extent, (left, 1, 3))
        parts.append((c + r.uniform(-1.0, 1.0, (left, 3, 3))).reshape(-1, 9))
    return np.concatenate(parts).astype(np.float32)


def soup_scene(tris: np.ndarray) -> Scene:
    """A soup (soup_tris) as a plain scene: face normals, grey surfaces, a
    sun; one object per 65,536 triangles."""
    t = tris.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    nrm = np.repeat(n[:, None, :], 3, 1).reshape(-1, 9).astype(np.float32)
    cols = np.full(len(tris), pack_rgba8(0.6, 0.6, 0.6), np.uint32)
    step = 65536
    objects = [(f"soup{i // step}", i, min(step, len(tris) - i)) for i in range(0, len(tris), step)]
    sun = {"direction": (-1.0, -6.0, -2.0), "angularDiameterDegrees": 0.5, "intensity": 7000.0, "color": (1.0, 1.0, 1.0)}
    return Scene(np.ascontiguousarray(tris, np.float32), nrm, cols, objects, {"sun": sun, "spheres": []})


def _finish(parts, objects, lights=None) -> Scene:
    tris = np.concatenate([p[0] for p in parts]).astype(np.float32)
    nrms = np.concatenate([p[1] for p in parts]).astype(np.float32)
    cols = np.concatenate([p[2] for p in parts]).astype(np.uint32)
    return Scene(tris, nrms, cols, objects, lights or {})


def terrain_boxes_scene(seed: int = SEED_TERRAIN_BOXES) -> Scene:
    """64x64-cell heightfield (8,192 tris, fBm in [0,4]) + 200 boxes (2,400 tris)."""
    size = 64.0
    h, ht, hn = heightfield(seed, 64, size, 4.0)
    nb = 200
    sz = uniform(seed, 1, nb * 3, 0.2, 2.0).reshape(nb, 3)
    xz = uniform(seed, 2, nb * 2, -30.0, 30.0).reshape(nb, 2)
    cy = height_at(h, size, xz[:, 0], xz[:, 1]) + sz[:, 1] * 0.5
    centers = np.stack([xz[:, 0], cy, xz[:, 1]], 1)
    bt, bn = boxes(centers, sz)
    hc = np.full(len(ht), pack_rgba8(0.45, 0.55, 0.35), np.uint32)
    bcol = uniform(seed, 3, nb * 3, 0.2, 0.9).reshape(nb, 3)
    bc = np.repeat(pack_rgba8(bcol[:, 0], bcol[:, 1], bcol[:, 2]), 12)
    objects = [("terrain", 0, len(ht))] + [(f"box{i}", len(ht) + 12 * i, 12) for i in range(nb)]
    sun = {"direction": (-1.0, -8.0, -1.0), "angularDiameterDegrees": 0.5, "intensity": 7000.0,
           "color": (1.0, 1.0, 1.0)}
    return _finish([(ht, hn, hc), (bt, bn, bc)], objects, {"sun": sun, "spheres": []})


def _sphere_levels(seed: int, count: int, budget: int) -> np.ndarray:
    """Levels 2..5 drawn with seeded weights, then capped greedily so the
    total triangle count lands within 0.5% of budget."""
    tri = {2: 320, 3: 1280, 4: 5120, 5: 20480}
    u = uniform(seed, 10, count)
    cdf = np.cumsum([0.55, 0.30, 0.12, 0.03])
    drawn = 2 + np.searchsorted(cdf, u, side="right")
    drawn = np.minimum(drawn, 5)
    levels = np.full(count, 2, np.int64)
    total = count * tri[2]
    for i in range(count):  # accept draws while the budget allows
        extra = tri[int(drawn[i])] - tri[2]
        if total + extra <= budget:
            levels[i] = drawn[i]
            total += extra
    i = 0
    while total < budget * 0.995 and i < count:  # top up level-2 spheres
        if levels[i] == 2 and total + (tri[3] - tri[2]) <= budget * 1.005:
            levels[i] = 3
            total += tri[3] - tri[2]
        i += 1
    return levels


def terrain_spheres_scene(seed: int = SEED_TERRAIN_SPHERES, cells: int = 512, spheres: int = 500,
                          target_tris: int = 1_000_000, lights: int = 64) -> Scene:
    """512x512 heightfield (524,288 tris) + 500 icospheres (levels 2-5, radii
    U[0.5,3]) totalling 1.0M +- 0.5% triangles, and 64 sphere lights."""
    size = 256.0
    h, ht, hn = heightfield(seed, cells, size, 4.0)
    levels = _sphere_levels(seed, spheres, target_tris - len(ht))
    radii = uniform(seed, 11, spheres, 0.5, 3.0)
    xz = uniform(seed, 12, spheres * 2, -100.0, 100.0).reshape(spheres, 2)
    cy = height_at(h, size, xz[:, 0], xz[:, 1]) + radii
    scol = uniform(seed, 13, spheres * 3, 0.2, 0.95).reshape(spheres, 3)
    parts = [(ht, hn, np.full(len(ht), pack_rgba8(0.45, 0.55, 0.35), np.uint32))]
    objects = [("terrain", 0, len(ht))]
    first = len(ht)
    for i in range(spheres):
        unit = icosphere(int(levels[i]))
        c = np.array([xz[i, 0], cy[i], xz[i, 1]])
        t = (unit * radii[i] + c).reshape(-1, 9)
        parts.append((t, unit.reshape(-1, 9), np.full(len(t), pack_rgba8(*scol[i]), np.uint32)))
        objects.append((f"sphere{i}", first, len(t)))
        first += len(t)
    lxz = uniform(seed, 14, lights * 2, -100.0, 100.0).reshape(lights, 2)
    ly = height_at(h, size, lxz[:, 0], lxz[:, 1]) + uniform(seed, 15, lights, 2.0, 10.0)
    lrad = uniform(seed, 16, lights, 0.2, 1.0)
    lint = uniform(seed, 17, lights, 500.0, 5000.0)
    lcol = uniform(seed, 18, lights * 3, 0.6, 1.0).reshape(lights, 3)
    sph = [{"position": (float(lxz[i, 0]), float(ly[i]), float(lxz[i, 1])), "radius": float(lrad[i]),
            "intensity": float(lint[i]), "color": tuple(map(float, lcol[i]))} for i in range(lights)]
    return _finish(parts, objects, {"sun": None, "spheres": sph})


# ------------------------------------------------------------------ camera

def look_at_view(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Column-major 4x4 view matrix (RgDrawFrameInfo::view, RTGL1.h:935-936)."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    M = np.eye(4)
    M[0, :3], M[1, :3], M[2, :3] = s, u, -f
    M[0, 3], M[1, 3], M[2, 3] = -s @ eye, -u @ eye, f @ eye
    return M.T.reshape(16)  # column-major


def projection(aspect: float, fovy: float, znear: float, zfar: float) -> np.ndarray:
    """Matrix::MakeProjectionMatrix (Source/Matrix.cpp:380-395): Vulkan clip
    space with a negated y axis and [0,1] depth; column-major, float32 math."""
    t = np.float32(math.tan(np.float32(fovy) * np.float32(0.5)))
    m = np.zeros(16, np.float32)
    m[0] = np.float32(1.0) / (np.float32(aspect) * t)
    m[5] = np.float32(-1.0) / t
    m[10] = np.float32(zfar) / (np.float32(znear) - np.float32(zfar))
    m[11] = -1.0
    m[14] = (np.float32(zfar) * np.float32(znear)) / (np.float32(znear) - np.float32(zfar))
    return m


def mat_cm(m16) -> np.ndarray:
    """Column-major 16 -> 4x4 (row, col)."""
    return np.asarray(m16, np.float64).reshape(4, 4).T


@dataclass
class Camera:
    eye: np.ndarray
    target: np.ndarray
    fovy: float = math.radians(75.0)
    znear: float = 0.1
    zfar: float = 10000.0

    def view(self) -> np.ndarray:
        return look_at_view(self.eye, self.target).astype(np.float32)

    def proj(self, width: int, height: int) -> np.ndarray:
        return projection(width / height, self.fovy, self.znear, self.zfar)


def terrain_boxes_camera(seed: int = SEED_TERRAIN_BOXES) -> Camera:
    """Seeded pose over the terrain, height 6-10, looking down 10-30 degrees."""
    u = uniform(seed, 20, 5)
    x, z = -20 + 40 * u[0], -20 + 40 * u[1]
    y = 6 + 4 * u[2]
    yaw = 2 * math.pi * u[3]
    pitch = math.radians(10 + 20 * u[4])
    d = np.array([math.cos(pitch) * math.cos(yaw), -math.sin(pitch), math.cos(pitch) * math.sin(yaw)])
    eye = np.array([x, y, z])
    return Camera(eye, eye + d)


def orbit_camera(frame: int) -> Camera:
    """Orbit of radius 80 at height 25, 6 degrees per frame, looking at the centre."""
    a = math.radians(6.0 * frame)
    eye = np.array([80 * math.cos(a), 25.0, 80 * math.sin(a)])
    return Camera(eye, np.array([0.0, 2.0, 0.0]))


def primary_rays(cam: Camera, width: int, height: int, tmin: float = 0.1, tmax: float = 10000.0) -> np.ndarray:
    """Pinhole camera rays through pixel centres, as getRayDir of
    ShaderCommonGLSLFunc.h:460-478 (no jitter), evaluated in float64 and
    rounded to float32.  Row-major over (y, x) of the regular image."""
    from .ffi import RAY_DTYPE
    view = mat_cm(cam.view())
    proj = mat_cm(cam.proj(width, height))
    inv_view, inv_proj = np.linalg.inv(view), np.linalg.inv(proj)
    xs = (np.arange(width) + 0.5) / width
    ys = (np.arange(height) + 0.5) / height
    U, V = np.meshgrid(xs, ys)
    ndc = np.stack([U * 2 - 1, V * 2 - 1, np.ones_like(U), np.ones_like(U)], -1).reshape(-1, 4)
    tgt = ndc @ inv_proj.T
    local = tgt[:, :3] / tgt[:, 3:4]
    local /= np.linalg.norm(local, axis=1, keepdims=True)
    d = np.concatenate([local, np.zeros((len(local), 1))], 1) @ inv_view.T
    rays = np.zeros(len(d), RAY_DTYPE)
    rays["origin"] = inv_view[:3, 3].astype(np.float32)
    rays["direction"] = d[:, :3].astype(np.float32)
    rays["tmin"] = tmin
    rays["tmax"] = tmax
    return rays
