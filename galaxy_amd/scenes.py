"""Scene descriptions and builders shared by the tests, the bench and smoke().

A scene is described once (plain numpy + dataclasses) and then instantiated on
a backend: the HIP library (galaxy_amd.lib) or, in tests only, the CPU oracle.
Both expose the same entry-point shapes (volume / TF / model creation and the
Visualization_ispc struct of src/renderer/Visualization.ih:30-47).

State files follow Galaxy's JSON schema (docs/state_files.md): Datasets,
Visualizations (Lighting + operators) and Cameras, loaded as in
Lighting::LoadStateFromValue (Lighting.cpp:58-115), VolumeVis::LoadFromJSON
(VolumeVis.cpp:119-163), MappedVis::LoadFromJSON (MappedVis.cpp:86-203) and
Camera::LoadCamerasFromJSON (Camera.cpp:79-115, 240-262).
"""
import ctypes as C
import json
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import abi

f32 = np.float32


# --------------------------------------------------------------------- volumes
@dataclass
class VolumeDesc:
    voxels: Optional[np.ndarray]       # float32, shape (nz, ny, nx), x fastest
    origin: tuple                      # ghosted brick origin (OSPRay gridOrigin)
    spacing: tuple                     # OSPRay gridSpacing
    sampling_rate: float = 1.0         # OSPRay samplingRate (OsprayVolume.cpp:45)
    local_box: Optional[tuple] = None  # (lo, hi) of the partition without ghosts
    global_box: Optional[tuple] = None
    neighbors: tuple = (-1, -1, -1, -1, -1, -1)
    shape: Optional[tuple] = None      # (nx, ny, nz) when voxels is None
    noise: Optional[tuple] = None      # (seed, octaves, base_frequency): filled on the device

    @property
    def dims(self):
        if self.voxels is None:
            return tuple(self.shape)
        nz, ny, nx = self.voxels.shape
        return (nx, ny, nz)


def radial_fields(res=256):
    """src/apps/radial.cpp:109-230 at one timestep (T = 1): the five float32
    fields of a res^3 grid over [-1,1]^3, spacing d = 2/(res-1)."""
    d = f32(2.0 / (res - 1))
    c = f32(-1) + np.arange(res, dtype=f32) * d
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    one = np.sqrt(X * X + Y * Y + Z * Z).astype(f32)
    ax = (np.abs(X) - f32(0.5)).astype(f32)
    ay = (np.abs(Y) - f32(0.5)).astype(f32)
    az = (np.abs(Z) - f32(0.5)).astype(f32)
    eight = np.sqrt(ax * ax + ay * ay + az * az).astype(f32)
    return {"xramp": np.ascontiguousarray(X), "yramp": np.ascontiguousarray(Y),
            "zramp": np.ascontiguousarray(Z), "oneBall": one, "eightBalls": eight}


def vol_header_from_vti(res=256):
    """scripts/vti2vol:72-79: origin -1, delta 2/(res-1) written with '%f'."""
    origin = float("%f" % -1.0)
    delta = float("%f" % (2.0 / (res - 1)))
    return (origin,) * 3, (delta,) * 3


def factor(ijk):
    """Volume.cpp:88-122: the i*j*k factorisation with the smallest i+j+k."""
    if ijk == 1:
        return (1, 1, 1)
    mm = ijk + 3
    best = (1, 1, ijk)
    for i in range(1, (ijk >> 1) + 1):
        jk = ijk // i
        if ijk == i * jk:
            for j in range(1, (jk >> 1) + 1):
                k = jk // j
                if jk == j * k:
                    m = i + j + k
                    if m < mm:
                        mm = m
                        best = (i, j, k)
    return best


def partition(size, factors, grid):
    """Volume.cpp:133-172: per-rank ijk, counts, ghosted counts, offsets."""
    ni, nj, nk = grid[0] - 2, grid[1] - 2, grid[2] - 2
    di, dj, dk = ni // factors[0], nj // factors[1], nk // factors[2]
    parts = []
    for k in range(factors[2]):
        for j in range(factors[1]):
            for i in range(factors[0]):
                off = (1 + i * di, 1 + j * dj, 1 + k * dk)
                cnt = (1 + (ni - off[0] if i == factors[0] - 1 else di),
                       1 + (nj - off[1] if j == factors[1] - 1 else dj),
                       1 + (nk - off[2] if k == factors[2] - 1 else dk))
                parts.append(dict(ijk=(i, j, k), offsets=off, counts=cnt,
                                  goffsets=tuple(o - 1 for o in off),
                                  gcounts=tuple(c + 2 for c in cnt)))
    return parts


def box_from(o, n, d):
    """Box::set(o, n, d) (Box.cpp:68-80) in float32."""
    lo = tuple(f32(o[k]) for k in range(3))
    hi = tuple(f32(lo[k] + f32(n[k] - 1) * f32(d[k])) for k in range(3))
    return lo, hi


def partitioned_volume(voxels, origin, deltas, rank=0, size=1, partitions=None,
                       sampling_rate=1.0):
    """Volume::local_import (Volume.cpp:174-392) for one rank of `size`: the
    ghosted brick, its OSPRay grid origin, the local and global boxes and the
    six face neighbours (-1 where none)."""
    nz, ny, nx = voxels.shape
    gcounts = (nx, ny, nz)
    fac = partitions or factor(size)
    parts = partition(size, fac, gcounts)
    p = parts[rank]
    go, gc = p["goffsets"], p["gcounts"]
    brick = np.ascontiguousarray(voxels[go[2]:go[2] + gc[2], go[1]:go[1] + gc[1], go[0]:go[0] + gc[0]])
    d = tuple(f32(x) for x in deltas)
    o = tuple(f32(x) for x in origin)
    gorigin = tuple(f32(o[k] + f32(go[k]) * d[k]) for k in range(3))
    gbox = box_from(tuple(f32(o[k] + d[k]) for k in range(3)), tuple(c - 2 for c in gcounts), d)
    lo = tuple(f32(o[k] + f32(p["offsets"][k]) * d[k]) for k in range(3))
    lbox = box_from(lo, p["counts"], d)
    i, j, k = p["ijk"]
    fx, fy, fz = fac

    def rk(a, b, c):
        return a + b * fx + c * fx * fy
    nb = (rk(i - 1, j, k) if i > 0 else -1, rk(i + 1, j, k) if i < fx - 1 else -1,
          rk(i, j - 1, k) if j > 0 else -1, rk(i, j + 1, k) if j < fy - 1 else -1,
          rk(i, j, k - 1) if k > 0 else -1, rk(i, j, k + 1) if k < fz - 1 else -1)
    return VolumeDesc(brick, gorigin, d, sampling_rate, lbox, gbox, nb)


# --------------------------------------------------------------------- vis / lights
DEFAULT_COLORMAP = np.array([[0.0, 0.4, 0.4, 0.4], [1.0, 1.0, 1.0, 1.0]], f32)   # MappedVis.cpp:58-59
DEFAULT_OPACITYMAP = np.array([[0.0, 1.0], [1.0, 1.0]], f32)                      # MappedVis.cpp:61-62


@dataclass
class VolumeVisDesc:
    volume: VolumeDesc
    colormap: np.ndarray = field(default_factory=lambda: DEFAULT_COLORMAP.copy())
    opacitymap: np.ndarray = field(default_factory=lambda: DEFAULT_OPACITYMAP.copy())
    data_range: Optional[tuple] = None
    slices: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), f32))
    isovalues: np.ndarray = field(default_factory=lambda: np.zeros((0,), f32))
    volume_render: bool = False


@dataclass
class MeshDesc:
    vertices: np.ndarray                       # (nv, 3) float32
    indices: np.ndarray                        # (nt, 3) uint32
    normals: Optional[np.ndarray] = None       # (nv, 3)
    vertex_data: Optional[np.ndarray] = None   # (nv,)
    colormap: Optional[np.ndarray] = None      # TF for vertex_data


@dataclass
class SpheresDesc:
    centers: np.ndarray                        # (n, 3)
    data: Optional[np.ndarray] = None
    radius0: float = 0.025                     # ParticlesVis.cpp:46-54
    radius1: float = 0.025
    value0: float = 0.0
    value1: float = 0.0
    colormap: Optional[np.ndarray] = None


@dataclass
class PathLinesDesc:
    vertices: np.ndarray                       # (nv, 3)
    data: np.ndarray                           # (nv,) per-vertex value -> radius
    indices: np.ndarray                        # (nseg,) segment i joins indices[i], indices[i] + 1
    radius0: float = 0.01                      # DataDrivenPathLines.cpp:58-63
    radius1: float = 0.0
    value0: float = 0.0
    value1: float = 0.0
    colormap: Optional[np.ndarray] = None


@dataclass
class LightingDesc:
    sources: np.ndarray = field(default_factory=lambda: np.array([[1, 1, 1]], f32))   # Lighting.ispc:39-43
    types: np.ndarray = field(default_factory=lambda: np.array([2], np.int32))
    Ka: float = 0.5
    Kd: float = 0.5
    n_ao: int = 0
    ao_radius: float = 1.0
    shadows: bool = False


@dataclass
class CameraDesc:
    eye: tuple
    center: tuple
    up: tuple
    aov: float
    width: int = 512
    height: int = 512


@dataclass
class SceneDesc:
    volume_vis: List[VolumeVisDesc] = field(default_factory=list)
    meshes: List[MeshDesc] = field(default_factory=list)
    spheres: List[SpheresDesc] = field(default_factory=list)
    pathlines: List[PathLinesDesc] = field(default_factory=list)
    local_box: tuple = ((-1, -1, -1), (1, 1, 1))
    global_box: tuple = ((-1, -1, -1), (1, 1, 1))
    neighbors: tuple = (-1, -1, -1, -1, -1, -1)
    lighting: LightingDesc = field(default_factory=LightingDesc)
    cameras: List[CameraDesc] = field(default_factory=list)
    epsilon: float = 0.001


def load_lighting(v):
    """Lighting::LoadStateFromValue (Lighting.cpp:58-115)."""
    L = LightingDesc()
    if "Sources" in v:
        src, typ = [], []
        for s in v["Sources"]:
            x, y, z = f32(s[0]), f32(s[1]), f32(s[2])
            t = 0 if len(s) == 3 else int(s[3])
            if t == 0:
                d = f32(np.sqrt(f32(f32(x * x) + f32(y * y)) + f32(z * z)))
                if d == 0:
                    x = y = z = f32(0.577350)
                else:
                    x, y, z = f32(x / d), f32(y / d), f32(z / d)
            src.append((x, y, z))
            typ.append(t)
        L.sources = np.array(src, f32)
        L.types = np.array(typ, np.int32)
    if "shadows" in v:
        L.shadows = bool(v["shadows"])
    L.n_ao = int(v.get("ao count", 0))
    L.ao_radius = float(v.get("ao radius", 1.0))
    L.Ka = float(v.get("Ka", 0.5))
    L.Kd = float(v.get("Kd", 0.5))
    return L


def resolve_lights(L, cam):
    """Rendering::resolve_lights (Rendering.cpp:157-216)."""
    def norm(a):
        d = f32(np.sqrt(f32(f32(a[0] * a[0]) + f32(a[1] * a[1])) + f32(a[2] * a[2])))
        if d != 0:
            d = f32(1.0 / float(d))
            a = np.array([a[0] * d, a[1] * d, a[2] * d], f32)
        return a

    def cross(a, b):
        return np.array([f32(a[1] * b[2]) - f32(a[2] * b[1]), f32(a[2] * b[0]) - f32(a[0] * b[2]),
                         f32(a[0] * b[1]) - f32(a[1] * b[0])], f32)
    vp = np.array(cam.eye, f32)
    vu = norm(np.array(cam.up, f32))
    vd = norm(np.array([f32(cam.center[k]) - f32(cam.eye[k]) for k in range(3)], f32))
    right = norm(cross(vd, vu))
    up2 = cross(vd, right)
    src = L.sources.astype(f32).copy()
    typ = L.types.astype(np.int32).copy()
    for i in range(len(src)):
        if typ[i] == 1:
            l = src[i].copy()
            o = vp + f32(l[0]) * right
            o = o + f32(l[1]) * up2
            o = o + f32(l[2]) * vd
            src[i] = o
            typ[i] = 2
    out = LightingDesc(src, typ, L.Ka, L.Kd, L.n_ao, L.ao_radius, L.shadows)
    return out


def camera_from_json(c, width=512, height=512):
    """One camera object of a state file or a viewer "Camera" command
    (Camera.cpp LoadFromJSON: viewpoint, viewcenter or viewdirection, viewup,
    aov)."""
    eye = tuple(float(x) for x in c["viewpoint"])
    if "viewcenter" in c:
        center = tuple(float(x) for x in c["viewcenter"])
    else:
        center = tuple(eye[k] + float(c["viewdirection"][k]) for k in range(3))
    return CameraDesc(eye, center, tuple(float(x) for x in c["viewup"]), float(c.get("aov", 30.0)), width, height)


def vis_from_json(v, dsets, eps=0.001, cams=()):
    """One Visualization object (operators + Lighting) over the datasets
    `dsets` (name -> VolumeDesc) as a SceneDesc (Visualization.cpp
    LoadFromJSON / local_commit)."""
    sc = SceneDesc(epsilon=eps, cameras=list(cams))
    sc.lighting = load_lighting(v.get("Lighting", {}))
    first = True
    for op in v.get("operators", []):
        if op["type"] not in ("Volume", "VolumeVis"):
            continue
        vol = dsets[op["dataset"]]
        vv = VolumeVisDesc(vol)
        if "colormap" in op:
            vv.colormap = np.array(op["colormap"], f32)
        if "opacitymap" in op:
            vv.opacitymap = np.array(op["opacitymap"], f32)
        if "data range" in op:
            vv.data_range = tuple(op["data range"])
        vv.isovalues = np.array(op.get("isovalues", []), f32)
        if "slices" in op:
            vv.slices = np.array(op["slices"], f32).reshape(-1, 4)
        elif "plane" in op:
            vv.slices = np.array(op["plane"], f32).reshape(1, 4)
        vv.volume_render = bool(op.get("volume rendering", False))
        sc.volume_vis.append(vv)
        if first:   # Visualization::local_commit takes the first vis's boxes (Visualization.cpp:150-162)
            sc.local_box, sc.global_box, sc.neighbors = vol.local_box, vol.global_box, vol.neighbors
            first = False
    return sc


def load_state(path_or_dict, volumes, width=512, height=512):
    """Loads a Galaxy state file.  `volumes` maps a dataset file name or
    dataset name to a VolumeDesc.  Returns a list of (SceneDesc, [CameraDesc])
    per Visualization."""
    doc = path_or_dict if isinstance(path_or_dict, dict) else json.load(open(path_or_dict))
    dsets = {}
    for d in doc.get("Datasets", []):
        key = d.get("filename", d["name"])
        vol = volumes.get(key, volumes.get(d["name"]))
        dsets[d["name"]] = vol
    cams = [camera_from_json(c, width, height) for c in doc.get("Cameras", [])]
    eps = float(doc.get("Renderer", {}).get("epsilon", 0.001))
    return [vis_from_json(v, dsets, eps, cams) for v in doc.get("Visualizations", [])]


# --------------------------------------------------------------------- images
def fb_to_rgba8(fb, width, height):
    """ColorImageWriter::Write (ImageWriter.cpp:30-48): (unsigned char)(255*v)
    with no clamp (x86: truncate to int32, keep the low byte), rows flipped."""
    v = np.asarray(fb, f32).reshape(height, width, 4)
    q = (np.trunc(v[..., :3] * f32(255)).astype(np.int64) & 0xFF).astype(np.uint8)
    img = np.empty((height, width, 4), np.uint8)
    img[..., :3] = q[::-1]
    img[..., 3] = 255
    return img


# --------------------------------------------------------------------- meshes
def uv_sphere(radius=1.0, nlon=128, nlat=80):
    """UV sphere with nlon*nlat*2 triangles (one triangle of each pole quad is
    degenerate), outward vertex normals and vertex data = z."""
    th = np.linspace(0, np.pi, nlat + 1)
    ph = np.linspace(0, 2 * np.pi, nlon, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    V = (radius * n).astype(f32)
    idx = []
    for i in range(nlat):
        for j in range(nlon):
            a = i * nlon + j
            b = i * nlon + (j + 1) % nlon
            c = (i + 1) * nlon + j
            d = (i + 1) * nlon + (j + 1) % nlon
            idx.append((a, c, b))
            idx.append((b, c, d))
    return MeshDesc(V, np.array(idx, np.uint32), n.astype(f32), V[:, 2].copy())


def uv_grid_indices(nlon, nlat):
    """Triangles of a UV sphere grid ((nlat+1) rings of nlon vertices), two per
    quad, in the order uv_sphere builds them."""
    i = np.arange(nlat, dtype=np.int64)[:, None]
    j = np.arange(nlon, dtype=np.int64)[None, :]
    a = i * nlon + j
    b = i * nlon + (j + 1) % nlon
    c = (i + 1) * nlon + j
    d = (i + 1) * nlon + (j + 1) % nlon
    t = np.stack([np.stack([a, c, b], -1), np.stack([b, c, d], -1)], 2)   # (nlat, nlon, 2, 3)
    return t.reshape(-1, 3).astype(np.uint32)


def blob_mesh(ntris, seed, center=(0.5, 0.5, 0.5), radius=0.32, amp=0.15):
    """A random blob: a UV sphere with about ntris triangles (4 nlat^2, nlon =
    2 nlat) whose radius is modulated by six seeded sinusoids of the direction,
    r = radius (1 + amp sum_k sin(f_k (u . a_k) + p_k) / (k+1)); vertex data =
    the modulation scaled to [0, 1]; no vertex normals (face normals shade)."""
    nlat = max(8, int(round((ntris / 4.0) ** 0.5)))
    nlon = 2 * nlat
    rng = np.random.default_rng(seed)
    axes = rng.normal(size=(6, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    phase = rng.uniform(0, 2 * np.pi, 6)
    th = np.linspace(0, np.pi, nlat + 1)
    ph = np.linspace(0, 2 * np.pi, nlon, endpoint=False)
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    u = np.stack([np.broadcast_to(st * np.cos(ph)[None, :], (nlat + 1, nlon)),
                  np.broadcast_to(st * np.sin(ph)[None, :], (nlat + 1, nlon)),
                  np.broadcast_to(ct, (nlat + 1, nlon))], -1).reshape(-1, 3)
    f = np.zeros(len(u))
    for k in range(6):
        f += np.sin(1.5 * (k + 1) * (u @ axes[k]) + phase[k]) / (k + 1)
    r = radius * (1.0 + amp * f)
    V = (u * r[:, None] + np.asarray(center)).astype(f32)
    data = ((f - f.min()) / max(1e-9, f.max() - f.min())).astype(f32)
    return MeshDesc(V, uv_grid_indices(nlon, nlat), None, data)


def icosphere(subdiv, radius=1.0, center=(0, 0, 0)):
    """Icosahedron subdivided `subdiv` times (20*4^subdiv triangles), projected
    to the sphere; vertex normals radial; vertex data = distance to the
    icosphere center along z, scaled to [0, 1]."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = np.array(V, np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(F, np.int64)
    for _ in range(subdiv):
        nv = len(V)
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (nv + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // (nv + 1), uk % (nv + 1)
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        V = np.concatenate([V, mid])
        m = inv.reshape(3, -1).T + nv      # midpoints of edges (01, 12, 20)
        f0, f1, f2 = F[:, 0], F[:, 1], F[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        F = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
                            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    n = V.astype(f32)
    P = (V * radius + np.array(center)).astype(f32)
    data = (0.5 * (V[:, 2] + 1)).astype(f32)
    return MeshDesc(P, F.astype(np.uint32), n, data)


# --------------------------------------------------------------------- backends
class Built:
    """Keeps every ctypes object and array of one instantiated scene alive."""

    def __init__(self):
        self.keep = []
        self.vis = None
        self.lights = None
        self.handles = []


def build_scene(be, sc: SceneDesc, lights: Optional[LightingDesc] = None):
    """Instantiates `sc` on backend `be` (galaxy_amd.lib.HipBackend or the
    tests' oracle backend).  Returns a Built with .vis (Visualization_ispc) and
    .lights (Lighting_ispc)."""
    B = Built()
    vvs = []
    made = {}
    for vv in sc.volume_vis:
        vol = vv.volume
        if id(vol) in made:
            hv = made[id(vol)]
        elif vol.voxels is None:
            hv = be.volume_create(vol.dims, vol.origin, vol.spacing, None, vol.sampling_rate)
            be.volume_fill_noise(hv, *vol.noise)
        else:
            vox = np.ascontiguousarray(vol.voxels, f32)
            B.keep.append(vox)
            hv = be.volume_create(vol.dims, vol.origin, vol.spacing, vox, vol.sampling_rate)
        made[id(vol)] = hv
        cm = np.ascontiguousarray(vv.colormap, f32)
        om = np.ascontiguousarray(vv.opacitymap, f32)
        rng = None if vv.data_range is None else np.array(vv.data_range, f32)
        ht = be.tf_create_from_maps(cm, om, rng)
        B.handles += [hv, ht]
        sl = np.ascontiguousarray(vv.slices, f32).reshape(-1)
        iso = np.ascontiguousarray(vv.isovalues, f32)
        B.keep += [sl, iso]
        s = abi.VolumeVis_ispc()
        s.vis.vis.data = hv
        s.vis.transferFunction = ht
        s.nSlices = len(sl) // 4
        s.slices = sl.ctypes.data if len(sl) else None
        s.nIsovalues = len(iso)
        s.isovalues = iso.ctypes.data if len(iso) else None
        s.volume_render = bool(vv.volume_render)
        vvs.append(s)
    arr = (C.POINTER(abi.VolumeVis_ispc) * max(1, len(vvs)))()
    for i, s in enumerate(vvs):
        arr[i] = C.pointer(s)
    B.keep += [vvs, arr]
    model = None
    if sc.meshes or sc.spheres or sc.pathlines:
        model = be.model_create()
        for m in sc.meshes:
            tf = None
            if m.colormap is not None and m.vertex_data is not None:
                tf = be.tf_create_from_maps(np.ascontiguousarray(m.colormap, f32), DEFAULT_OPACITYMAP.copy(), None)
                B.handles.append(tf)
            be.model_add_triangles(model, m.vertices, m.indices, m.normals, m.vertex_data, tf)
        for s in sc.spheres:
            tf = None
            if s.colormap is not None:
                tf = be.tf_create_from_maps(np.ascontiguousarray(s.colormap, f32), DEFAULT_OPACITYMAP.copy(), None)
                B.handles.append(tf)
            be.model_add_spheres(model, s.centers, s.data, s.radius0, s.radius1, s.value0, s.value1, tf)
        for p in sc.pathlines:
            tf = None
            if p.colormap is not None:
                tf = be.tf_create_from_maps(np.ascontiguousarray(p.colormap, f32), DEFAULT_OPACITYMAP.copy(), None)
                B.handles.append(tf)
            be.model_add_pathlines(model, p.vertices, p.data, p.indices, p.radius0, p.radius1, p.value0, p.value1,
                                   tf)
        be.model_commit(model)
        B.handles.append(model)
    vis = abi.Visualization_ispc()
    vis.model = model
    vis.nVolumeVis = len(vvs)
    vis.volumeVis = C.cast(arr, C.POINTER(C.POINTER(abi.VolumeVis_ispc)))
    vis.nMappedVis = 0
    vis.mappedVis = None
    vis.global_bb = abi.make_box(*sc.global_box)
    vis.local_bb = abi.make_box(*sc.local_box)
    B.vis = vis
    B.lights = make_lighting(lights or sc.lighting, B.keep)
    return B


def make_lighting(L: LightingDesc, keep):
    src = np.ascontiguousarray(L.sources, f32).reshape(-1)
    typ = np.ascontiguousarray(L.types, np.int32)
    keep += [src, typ]
    l = abi.Lighting_ispc()
    l.n_ao_rays = int(L.n_ao)
    l.do_shadows = int(bool(L.shadows))
    l.ao_radius = float(L.ao_radius)
    l.Ka = float(L.Ka)
    l.Kd = float(L.Kd)
    l.types = typ.ctypes.data
    l.lights = src.ctypes.data
    l.nLights = len(typ)
    return l
