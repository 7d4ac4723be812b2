"""Closed-form pins of the geometry shading a primary ray's surface hit carries
into the RayList (sr sg sb, nx ny nz):

* triangles: DataDrivenTriangleMesh_postIntersect
  (src/ospray/DataDrivenTriangleMesh.ispc:34-121): Ns is the barycentric
  interpolation of the vertex normals with weights (1-u-v, u, v) on the
  triangle's vertices (index.x, index.y, index.z); the colour is the TF of the
  barycentric interpolation of the vertex data;
* spheres: DataDrivenSpheres_postIntersect (src/ospray/DataDrivenSpheres.ispc:46-63):
  Ns = Ng = hit point - centre; the colour is the TF of the sphere's datum;
* then OSPRay's DG_NORMALIZE | DG_FACEFORWARD (src/renderer/TraceRays.ispc:129):
  unit normal, facing against the ray.

What the mathematics fixes:
* vertex data that is an affine function of position is reproduced exactly by
  barycentric interpolation at any point of the triangle, so the hit colour
  through a two-point linear colormap is the colormap at the affine function
  of the hit point;
* vertex normals (v - c)/R of a mesh inscribed in a sphere interpolate to
  (P - c)/R at any point P of a triangle, so the normalised Ns is exactly the
  radial direction at the hit point (not merely within O(h^2) of it), while
  the facet normal departs from it by O(h/R);
* the face-forwarded normal satisfies N . dir <= 0 for rays from outside and
  from inside alike.

Each check runs on the CPU oracle and, marked gpu, on libgxy_hip.so.  Planted
mistakes in the oracle (swapped barycentric weights, the facet normal for Ns,
the first vertex's datum for all three, no face-forward, the next sphere's
datum) must each fail a check."""
import numpy as np
import pytest

from galaxy_amd import abi, scenes

f32 = np.float32

CEN = np.array([0.1, -0.2, 0.05])
RAD = 1.2
GRAD = np.array([0.3, -0.2, 0.25])
OFF = 0.5
C0 = np.array([0.1, 0.8, 0.3])
C1 = np.array([0.9, 0.2, 0.6])
BOX = ((-3.0,) * 3, (3.0,) * 3)
NTOL = 2e-5     # unit-normal components
CTOL = 2e-5     # colour channels

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


def _mesh_scene():
    m = scenes.icosphere(2, RAD, tuple(CEN))
    V = np.asarray(m.vertices, np.float64)
    m.normals = ((V - CEN) / RAD).astype(f32)
    data = V @ GRAD + OFF
    m.vertex_data = data.astype(f32)
    lo, hi = float(data.min()) - 0.05, float(data.max()) + 0.05
    m.colormap = np.array([[lo, *C0], [hi, *C1]], f32)
    return scenes.SceneDesc(meshes=[m], local_box=BOX, global_box=BOX), (lo, hi)


def _mesh_rays(n, seed):
    """Half the rays start outside the sphere aimed at points inside it, half
    start inside it (their hits see the back faces, so face-forward flips)."""
    rng = np.random.default_rng(seed)
    k = n // 2
    d_out = rng.normal(size=(k, 3))
    d_out /= np.linalg.norm(d_out, axis=1, keepdims=True)
    o_out = CEN + 2.5 * d_out
    tgt = CEN + rng.uniform(-0.7, 0.7, (k, 3))
    o_in = CEN + rng.uniform(-0.5, 0.5, (n - k, 3))
    tgt_in = o_in + rng.normal(size=(n - k, 3))
    org = np.concatenate([o_out, o_in])
    d = np.concatenate([tgt - o_out, tgt_in - o_in])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return abi.HostRays.from_arrays(org.astype(f32), d.astype(f32))


def _trace(name, be, sc, hr):
    """Traces primary rays `hr` through scene `sc` on backend `be`; returns the
    RayList after the pass and the hit primitive per ray (-1 = none)."""
    B = scenes.build_scene(be, sc)
    if name == "oracle":
        o = hr.copy()
        res = be.trace(B.vis, o, geom_ids=True)
        return o, res["prim"]
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    n = hr.n
    gd, pd = glib.DeviceArray(4 * n), glib.DeviceArray(4 * n)
    glib.check(L_.gxy_model_set_hit_output(B.vis.model, gd.ptr, pd.ptr))
    try:
        d = glib.DeviceRays.from_host(hr)
        be.trace(B.vis, d)
        return d.download(), pd.to_host(np.int32)
    finally:
        glib.check(L_.gxy_model_set_hit_output(B.vis.model, None, None))


def _backend(name, request):
    return request.getfixturevalue(name)


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _mesh_errors(R, prim, rng_lohi):
    """Largest departures of the traced normals and colours from the closed
    forms, over the rays that hit."""
    lo, hi = rng_lohi
    hit = prim >= 0
    org = np.stack([R["ox"], R["oy"], R["oz"]], 1).astype(np.float64)[hit]
    d = np.stack([R["dx"], R["dy"], R["dz"]], 1).astype(np.float64)[hit]
    P = org + R["t"].astype(np.float64)[hit, None] * d
    N = _unit(P - CEN)
    N = np.where((N * d).sum(1, keepdims=True) >= 0, -N, N)
    got_n = np.stack([R["nx"], R["ny"], R["nz"]], 1).astype(np.float64)[hit]
    w = np.clip((P @ GRAD + OFF - lo) / (hi - lo), 0, 1)[:, None]
    col = (1 - w) * C0 + w * C1
    got_c = np.stack([R["sr"], R["sg"], R["sb"]], 1).astype(np.float64)[hit]
    facing = (got_n * d).sum(1).max()
    return np.abs(got_n - N).max(), np.abs(got_c - col).max(), facing, int(hit.sum())


@pytest.mark.parametrize("name", BACKENDS)
def test_triangle_shading_closed_forms(name, request):
    be = _backend(name, request)
    sc, lohi = _mesh_scene()
    hr = _mesh_rays(4000, seed=21)
    R, prim = _trace(name, be, sc, hr)
    en, ec, facing, nhit = _mesh_errors(R, prim, lohi)
    assert nhit >= 3900                      # every ray meets the closed mesh
    assert en <= NTOL, en
    assert ec <= CTOL, ec
    assert facing <= 1e-6
    # the surface flag and the unit surface alpha (TraceRays.ispc:581-617)
    hit = prim >= 0
    assert np.all(R["term"][hit] & abi.RAY_SURFACE) and np.all(R["so"][hit] == 1.0)


@pytest.mark.parametrize("fault", [1, 2, 3, 4])
def test_triangle_shading_pins_fail_on_planted_mistakes(oracle, fault):
    sc, lohi = _mesh_scene()
    hr = _mesh_rays(2000, seed=22)
    oracle.L.oracle_plant_fault(fault)
    try:
        R, prim = _trace("oracle", oracle, sc, hr)
    finally:
        oracle.L.oracle_plant_fault(0)
    en, ec, facing, _ = _mesh_errors(R, prim, lohi)
    assert en > 100 * NTOL or ec > 100 * CTOL or facing > 1e-3


def _sphere_scene(seed=5):
    rng = np.random.default_rng(seed)
    g = np.arange(4) * 0.9 - 1.35
    cen = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + rng.uniform(-0.05, 0.05, (64, 3))
    data = rng.uniform(0, 1, 64)
    sp = scenes.SpheresDesc(cen.astype(f32), data.astype(f32), radius0=0.3, radius1=0.3, value0=0.0, value1=1.0,
                            colormap=np.array([[0, *C0], [1, *C1]], f32))
    return scenes.SceneDesc(spheres=[sp], local_box=BOX, global_box=BOX), cen.astype(f32).astype(np.float64), data


def _sphere_rays(cen, n, seed):
    rng = np.random.default_rng(seed)
    k = n - n // 8
    tgt = cen[rng.integers(0, len(cen), k)] + rng.uniform(-0.25, 0.25, (k, 3))
    o = rng.uniform(-2.8, 2.8, (k, 3))
    j = rng.integers(0, len(cen), n - k)
    o_in = cen[j] + rng.uniform(-0.1, 0.1, (n - k, 3))      # inside a sphere: exits through it
    org = np.concatenate([o, o_in])
    d = np.concatenate([tgt - o, rng.normal(size=(n - k, 3))])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return abi.HostRays.from_arrays(org.astype(f32), d.astype(f32))


def _sphere_closed_form(org, d, cen, r):
    """FP64 nearest hit of each ray over all spheres: the entry root if it
    lies ahead of t0 = 0, else the exit root if it lies beyond t0 + epsilon,
    epsilon = min(r/100, -1/log r) for r < 1 (DataDrivenSpheres.ispc:134-141,
    212-220)."""
    eps = min(r / 100.0, -1.0 / np.log(r))
    oc = org[:, None, :] - cen[None]
    b = (oc * d[:, None, :]).sum(-1)
    c = (oc * oc).sum(-1) - r * r
    disc = b * b - c
    s = np.sqrt(np.maximum(disc, 0))
    t_in, t_out = -b - s, -b + s
    t = np.where(t_in > 0, t_in, np.where(t_out > eps, t_out, np.inf))
    t = np.where(disc >= 0, t, np.inf)
    k = t.argmin(1)
    return t[np.arange(len(org)), k], k


def _sphere_errors(R, prim, cen, data):
    hit = prim >= 0
    org = np.stack([R["ox"], R["oy"], R["oz"]], 1).astype(np.float64)
    d = np.stack([R["dx"], R["dy"], R["dz"]], 1).astype(np.float64)
    t_cf, k = _sphere_closed_form(org, d, cen, 0.3)
    assert np.array_equal(hit, np.isfinite(t_cf))
    org, d, k, p = org[hit], d[hit], k[hit], prim[hit]
    ids = (p == k).mean()
    t = R["t"].astype(np.float64)[hit]
    et = np.abs(t - t_cf[hit]) / np.maximum(t_cf[hit], 1.0)
    P = org + t[:, None] * d
    N = _unit(P - cen[p])
    N = np.where((N * d).sum(1, keepdims=True) >= 0, -N, N)
    got_n = np.stack([R["nx"], R["ny"], R["nz"]], 1).astype(np.float64)[hit]
    w = np.clip(data[p], 0, 1)[:, None]
    col = (1 - w) * C0 + w * C1
    got_c = np.stack([R["sr"], R["sg"], R["sb"]], 1).astype(np.float64)[hit]
    return ids, et.max(), np.abs(got_n - N).max(), np.abs(got_c - col).max(), int(hit.sum())


@pytest.mark.parametrize("name", BACKENDS)
def test_sphere_shading_closed_forms(name, request):
    be = _backend(name, request)
    sc, cen, data = _sphere_scene()
    hr = _sphere_rays(cen, 4000, seed=23)
    R, prim = _trace(name, be, sc, hr)
    ids, et, en, ec, nhit = _sphere_errors(R, prim, cen, data)
    assert nhit >= 3000
    assert ids >= 0.9999
    assert et <= 1e-5, et          # relative to max(t, 1): a shorter t carries the O(1) coordinates' rounding
    assert en <= 1e-4, en          # the normal of an FP32 hit point 1e-6 off the surface, over r = 0.3
    assert ec <= CTOL, ec


def test_sphere_shading_pin_fails_on_planted_mistake(oracle):
    sc, cen, data = _sphere_scene()
    hr = _sphere_rays(cen, 2000, seed=24)
    oracle.L.oracle_plant_fault(5)
    try:
        R, prim = _trace("oracle", oracle, sc, hr)
    finally:
        oracle.L.oracle_plant_fault(0)
    _, _, _, ec, _ = _sphere_errors(R, prim, cen, data)
    assert ec > 100 * CTOL
