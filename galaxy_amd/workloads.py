"""The benchmark workloads (synthetic, seeded), as scene descriptions.

  sphere_mesh        512x512 pinhole camera, primary + one hard shadow ray per pixel from one
      point light, one UV-sphere mesh (radius 1, 128x80 segments, 20,480 FP32
      triangles, Lambertian shading), camera on a +z orbit at distance 3
  volume_render      1920x1080 rays through a 512^3 FP32 field of seeded 4-octave gradient
      noise in [0,1]; 8-point piecewise-linear RGBA transfer function; step 0.5
      voxel; front-to-back compositing with early termination at alpha 0.99;
      isosurface at 0.5 with shadow rays
  nine_icospheres    3x3 grid of icospheres at subdivision 8 (1,310,720 triangles each,
      11,796,480 total), 3840x2160 primary rays, 1 shadow ray and 16
      cosine-weighted AO rays per hit, every secondary ray spawned through the
      device RayList queue
  batched_lists      256 ray lists of log-uniform sizes in [1k, 1M] (origins uniform
      in the unit cube, isotropic directions) over 64 resident datasets, each a
      random blob mesh of 50k-5M triangles or a 128^3-1024^3 FP32 noise volume
      (log-uniform sizes), traced together pass by pass with shadow and AO
      secondary rays
"""
import math

import numpy as np

from . import scenes

f32 = np.float32


def volume_boxes(n, origin, delta):
    """Local and global boxes of a single-rank volume with Galaxy's one-voxel
    ghost ring (Volume.cpp:174-392 with size 1)."""
    o = tuple(f32(x) for x in origin)
    d = tuple(f32(x) for x in delta)
    lo = tuple(f32(o[k] + d[k]) for k in range(3))
    box = scenes.box_from(lo, tuple(c - 2 for c in n), d)
    return box, box


def volume_render(n=512, width=1920, height=1080, seed=1234, octaves=4, base_frequency=4.0):
    """Volume render: noise volume + DVR + iso 0.5 + shadows."""
    delta = 2.0 / (n - 1)
    origin = (-1.0, -1.0, -1.0)
    lbox, gbox = volume_boxes((n, n, n), origin, (delta,) * 3)
    # step 0.5 voxel: TraceRays.ispc:354 uses samplingStep * samplingRate
    vol = scenes.VolumeDesc(None, tuple(f32(x) for x in origin), (f32(delta),) * 3, sampling_rate=0.5,
                            local_box=lbox, global_box=gbox, shape=(n, n, n),
                            noise=(seed, octaves, base_frequency))
    d = np.linspace(0.0, 1.0, 8)
    rgb = np.array([[0.23, 0.30, 0.75], [0.34, 0.51, 0.95], [0.55, 0.69, 1.00], [0.75, 0.82, 0.93],
                    [0.93, 0.82, 0.76], [0.97, 0.66, 0.54], [0.90, 0.45, 0.34], [0.71, 0.02, 0.15]])
    op = np.array([0.02, 0.0, 0.04, 0.0, 0.0, 0.04, 0.0, 0.03])
    vv = scenes.VolumeVisDesc(vol, colormap=np.c_[d, rgb].astype(f32), opacitymap=np.c_[d, op].astype(f32),
                              isovalues=np.array([0.5], f32), volume_render=True)
    L = scenes.LightingDesc(np.array([[3.0, 4.0, 5.0]], f32), np.array([2], np.int32), Ka=0.3, Kd=0.7,
                            n_ao=0, ao_radius=1.0, shadows=True)
    cam = scenes.CameraDesc((2.2, 1.6, 2.8), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 35.0, width, height)
    return scenes.SceneDesc(volume_vis=[vv], local_box=lbox, global_box=gbox, lighting=L, cameras=[cam]), 0.99


def sphere_mesh(width=512, height=512, seed=1):
    """Sphere mesh with one point light and hard shadows."""
    m = scenes.uv_sphere(1.0, 128, 80)
    m.colormap = np.array([[-1.0, 0.8, 0.3, 0.2], [1.0, 0.2, 0.5, 0.9]], f32)
    rng = np.random.default_rng(seed)
    az = rng.uniform(0, 2 * math.pi)
    eye = (3.0 * math.sin(az), 0.0, 3.0 * math.cos(az))
    box = ((-1.25,) * 3, (1.25,) * 3)
    L = scenes.LightingDesc(np.array([[4.0, 3.0, 5.0]], f32), np.array([2], np.int32), Ka=0.3, Kd=0.7,
                            shadows=True)
    cam = scenes.CameraDesc(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, width, height)
    return scenes.SceneDesc(meshes=[m], local_box=box, global_box=box, lighting=L, cameras=[cam]), 0.999


# seed of the nine-icosphere frame's AO directions (PCG32 streams per pixel,
# gxy_set_ao_directions(1, seed)); parity runs of the frame also use the
# reference's UV table
NINE_ICOSPHERES_AO_SEED = 8


def nine_icospheres(width=3840, height=2160, subdiv=8, n_ao=16, ao_radius=0.5):
    """Nine icospheres, shadows and 16 AO rays per hit (AO directions seeded:
    NINE_ICOSPHERES_AO_SEED)."""
    base = scenes.icosphere(subdiv, 0.45)
    meshes = []
    for i in range(3):
        for j in range(3):
            c = np.array([(i - 1) * 1.1, (j - 1) * 1.1, 0.0], f32)
            m = scenes.MeshDesc(base.vertices + c, base.indices, base.normals, base.vertex_data,
                                np.array([[0.0, 0.9, 0.4, 0.3], [1.0, 0.3, 0.6, 0.9]], f32))
            meshes.append(m)
    box = ((-1.8, -1.8, -0.6), (1.8, 1.8, 0.6))
    L = scenes.LightingDesc(np.array([[2.0, 3.0, 6.0]], f32), np.array([2], np.int32), Ka=0.4, Kd=0.6,
                            n_ao=n_ao, ao_radius=ao_radius, shadows=True)
    cam = scenes.CameraDesc((0.0, -1.2, 4.2), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, width, height)
    return scenes.SceneDesc(meshes=meshes, local_box=box, global_box=box, lighting=L, cameras=[cam]), 0.999


VOLUME_RENDER_TF = (np.array([[0.23, 0.30, 0.75], [0.34, 0.51, 0.95], [0.55, 0.69, 1.00], [0.75, 0.82, 0.93],
                   [0.93, 0.82, 0.76], [0.97, 0.66, 0.54], [0.90, 0.45, 0.34], [0.71, 0.02, 0.15]]),
         np.array([0.02, 0.0, 0.04, 0.0, 0.0, 0.04, 0.0, 0.03]))

UNIT_BOX = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def pathlines(width=1024, height=1024, nlines=4000, nverts=64, seed=7):
    """Data-driven path lines: seeded smooth random curves (a walk with
    momentum) through the unit cube, per-vertex data mapping to radii
    0.002-0.006 and to a two-colour TF, one point light with shadows."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.1, 0.9, (nlines, 3))
    v = rng.normal(size=(nlines, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    pts = np.empty((nlines, nverts, 3))
    for k in range(nverts):
        pts[:, k] = p
        v = v + 0.35 * rng.normal(size=v.shape)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        p = np.clip(p + 0.012 * v, 0.02, 0.98)
    V = pts.reshape(-1, 3).astype(f32)
    D = np.repeat(rng.uniform(0, 1, nlines), nverts).astype(f32) * np.tile(np.linspace(0.3, 1.0, nverts), nlines)
    I = (np.arange(nlines)[:, None] * nverts + np.arange(nverts - 1)[None, :]).reshape(-1).astype(np.int32)
    pl = scenes.PathLinesDesc(V, D.astype(f32), I, radius0=0.002, radius1=0.006, value0=0.0, value1=1.0,
                              colormap=np.array([[0.0, 0.2, 0.4, 0.9], [1.0, 0.95, 0.5, 0.1]], f32))
    box = UNIT_BOX
    L = scenes.LightingDesc(np.array([[2.0, 3.0, 4.0]], f32), np.array([2], np.int32), Ka=0.3, Kd=0.7,
                            shadows=True)
    cam = scenes.CameraDesc((0.5, 0.6, 2.3), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 35.0, width, height)
    return scenes.SceneDesc(pathlines=[pl], local_box=box, global_box=box, lighting=L, cameras=[cam]), 0.999


def unit_volume_scene(edge, seed, octaves=4, base_frequency=4.0):
    """A noise volume of edge^3 voxels spanning the unit cube, with the volume
    render's 8-point TF, iso 0.5 and compositing on; step 0.5 voxel."""
    d = np.linspace(0.0, 1.0, 8)
    rgb, op = VOLUME_RENDER_TF
    delta = f32(1.0 / (edge - 1))
    vol = scenes.VolumeDesc(None, (f32(0.0),) * 3, (delta,) * 3, sampling_rate=0.5, local_box=UNIT_BOX,
                            global_box=UNIT_BOX, shape=(edge, edge, edge), noise=(seed, octaves, base_frequency))
    vv = scenes.VolumeVisDesc(vol, colormap=np.c_[d, rgb].astype(f32), opacitymap=np.c_[d, op].astype(f32),
                              isovalues=np.array([0.5], f32), volume_render=True)
    return scenes.SceneDesc(volume_vis=[vv], local_box=UNIT_BOX, global_box=UNIT_BOX)


def blob_scene(ntris, seed):
    m = scenes.blob_mesh(ntris, seed)
    m.colormap = np.array([[0.0, 0.9, 0.5, 0.3], [1.0, 0.3, 0.6, 0.9]], f32)
    return scenes.SceneDesc(meshes=[m], local_box=UNIT_BOX, global_box=UNIT_BOX)


def batch_lighting():
    return scenes.LightingDesc(np.array([[2.0, 3.0, 4.0]], f32), np.array([2], np.int32), Ka=0.3, Kd=0.7,
                               n_ao=2, ao_radius=0.1, shadows=True)


def batched_lists(nlists=256, ndatasets=64, rays=(1_000, 1_000_000), edges=(128, 1024), tris=(50_000, 5_000_000), seed=4):
    """Returns (dataset scene descriptions, [(rays, dataset, seed), ...] per
    list, lighting).  Dataset k is a volume when k is even, a blob mesh when
    odd; sizes and list sizes are log-uniform in the given ranges."""
    rng = np.random.default_rng(seed)

    def logu(lo, hi):
        return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))

    descs = []
    for k in range(ndatasets):
        if k % 2 == 0:
            descs.append(("volume", int(round(logu(*edges))), 1000 + k))
        else:
            descs.append(("mesh", int(round(logu(*tris))), 2000 + k))
    lists = [(int(round(logu(*rays))), int(rng.integers(ndatasets)), 5000 + i) for i in range(nlists)]
    return descs, lists, batch_lighting()


def dataset_scene(desc):
    kind, size, s = desc
    return unit_volume_scene(size, s) if kind == "volume" else blob_scene(size, s)
