"""Pins of the CPU oracle's functions against closed forms, the in-tree
reference formulas, hand-enumerated reference rules and brute force, so that a
dropped term, a wrong sign or index, or a transposed operand fails a test."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, random_rays
from galaxy_amd import abi, scenes

f32 = np.float32


def test_uv_table_matches_reference_values(oracle):
    U = np.zeros(256, f32)
    V = np.zeros(256, f32)
    oracle.L.oracle_uv_tables(U.ctypes.data, V.ctypes.data)
    ref = json.load(open(os.path.join(GOLDEN, "uv_table.json")))
    assert np.array_equal(U, np.array(ref["randomU"], f32))
    assert np.array_equal(V, np.array(ref["randomV"], f32))


def _vol(oracle, data, origin=(0, 0, 0), spacing=(1, 1, 1), rate=1.0):
    nz, ny, nx = data.shape
    return oracle.volume_create((nx, ny, nz), origin, spacing, np.ascontiguousarray(data, f32), rate)


def test_trilinear_reproduces_linear_fields(oracle):
    # trilinear interpolation is exact on affine fields: v = 0.25x - 0.5y + 0.125z + 1
    n = 9
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sp = 0.5
    data = (0.25 * x * sp - 0.5 * y * sp + 0.125 * z * sp + 1.0).astype(f32)
    h = _vol(oracle, data, (0, 0, 0), (sp, sp, sp))
    rng = np.random.default_rng(3)
    for p in rng.uniform(0, (n - 1) * sp, (500, 3)):
        v = oracle.L.oracle_volume_sample(h, *map(float, p))
        assert abs(v - (0.25 * p[0] - 0.5 * p[1] + 0.125 * p[2] + 1.0)) < 2e-6


def test_sample_matches_in_tree_volume_sample(oracle, radial_vols_small):
    """Galaxy's own CPU trilinear (Volume::Sample, Volume.cpp:546-590: divide
    by delta, floor, lerp x then y then z) against the oracle's OSPRay-style
    sample (multiply by 1/spacing, clamp, truncate)."""
    vol = radial_vols_small["radial-eightBalls.vol"]
    h = _vol(oracle, vol.voxels, vol.origin, vol.spacing)
    d = np.array(vol.spacing, np.float64)
    o = np.array(vol.origin, np.float64)
    D = vol.voxels.astype(np.float64)
    rng = np.random.default_rng(5)
    pts = rng.uniform(o + 1e-3, o + (np.array(vol.dims) - 1.01) * d, (400, 3))
    for p in pts:
        g = (p - o) / d
        ll = np.floor(g).astype(int)
        f = g - ll
        c = D[ll[2]:ll[2] + 2, ll[1]:ll[1] + 2, ll[0]:ll[0] + 2]
        t0 = (1 - f[0]) * c[:, :, 0] + f[0] * c[:, :, 1]
        t1 = (1 - f[1]) * t0[:, 0] + f[1] * t0[:, 1]
        ref = (1 - f[2]) * t1[0] + f[2] * t1[1]
        assert abs(oracle.L.oracle_volume_sample(h, *map(float, p)) - ref) < 1e-5


def test_sample_clamps_outside_the_grid(oracle):
    data = np.arange(27, dtype=f32).reshape(3, 3, 3)
    h = _vol(oracle, data)
    assert oracle.L.oracle_volume_sample(h, -5.0, -5.0, -5.0) == 0.0
    assert abs(oracle.L.oracle_volume_sample(h, 9.0, 9.0, 9.0) - 26.0) < 1e-5


def test_gradient_of_distance_field_is_radial(oracle, radial_vols):
    """oneBall = |p|: the forward-difference gradient must point along p (the
    isosurface normal) to within the O(h) stencil error."""
    vol = radial_vols["radial-oneBall.vol"]
    h = _vol(oracle, vol.voxels, vol.origin, vol.spacing)
    rng = np.random.default_rng(7)
    g = np.zeros(3, f32)
    worst = 0.0
    for _ in range(300):
        p = rng.normal(size=3)
        p = 0.5 * p / np.linalg.norm(p)
        oracle.L.oracle_volume_gradient(h, *map(float, p), g.ctypes.data)
        cosang = float(np.dot(g, p) / (np.linalg.norm(g) * np.linalg.norm(p)))
        worst = max(worst, math.acos(min(1.0, cosang)))
    assert worst < 0.02


def test_tf_resample_and_lookup_closed_forms(oracle):
    cmap = np.array([[0.0, 0.0, 0.0, 0.0], [1.0, 1.0, 0.5, 0.25]], f32)
    omap = np.array([[0.0, 0.0], [1.0, 1.0]], f32)
    t = oracle.tf_create_from_maps(cmap, omap, None)
    col = np.zeros(768, f32)
    op = np.zeros(256, f32)
    rng2 = np.zeros(2, f32)
    oracle.L.oracle_tf_get_tables(t, col.ctypes.data, op.ctypes.data, rng2.ctypes.data)
    i = np.arange(256)
    assert np.allclose(col[0::3], i / 255.0, atol=1e-6)
    assert np.allclose(col[1::3], 0.5 * i / 255.0, atol=1e-6)
    assert np.allclose(op, i / 255.0, atol=1e-6)
    assert tuple(rng2) == (0.0, 1.0)
    rgb = np.zeros(3, f32)
    o = C.c_float()
    for v, ex in ((-1.0, 0.0), (0.0, 0.0), (0.5, 0.5), (0.3, 0.3), (2.0, 1.0)):
        oracle.L.oracle_tf_lookup(t, v, rgb.ctypes.data, C.byref(o))
        assert abs(rgb[0] - ex) < 1e-6 and abs(o.value - ex) < 1e-6


def test_tf_opacity_uses_colormap_value_range(oracle):
    """MappedVis.cpp:330-333: the value range comes from the colormap (or the
    data range); the opacity table is resampled over its own points but looked
    up over that range."""
    cmap = np.array([[0.0, 1, 1, 1], [2.0, 1, 1, 1]], f32)
    omap = np.array([[0.0, 0.0], [1.0, 1.0]], f32)
    t = oracle.tf_create_from_maps(cmap, omap, None)
    rgb = np.zeros(3, f32)
    o = C.c_float()
    oracle.L.oracle_tf_lookup(t, 1.0, rgb.ctypes.data, C.byref(o))
    assert abs(o.value - 0.5) < 1e-6      # 1.0 is half-way through [0, 2]


def _classify(oracle, types, terms, reverse=1):
    hr = abi.HostRays(len(types))
    hr["type"] = np.array(types, np.int32)
    hr["term"] = np.array(terms, np.int32)
    oracle.L.oracle_classify(hr.n, C.byref(hr.ispc), reverse)
    return list(hr["classification"])


def test_classify_rules():
    from oracle_backend import OracleBackend
    o = OracleBackend()
    P, S, A = abi.RAY_PRIMARY, abi.RAY_SHADOW, abi.RAY_AO
    SU, OP, BO, TI = abi.RAY_SURFACE, abi.RAY_OPAQUE, abi.RAY_BOUNDARY, abi.RAY_TIMEOUT
    T, D, K = abi.TERMINATED, abi.DROP_ON_FLOOR, abi.KEEP_HERE
    cases = [  # Renderer.cpp:318-420 under GXY_REVERSE_LIGHTING
        (P, BO, BO), (P, OP, T), (P, TI, T), (P, SU, K), (P, SU | OP, T), (P, 0, K),
        (S, SU, T), (S, OP, T), (S, BO, BO), (S, TI, D), (S, 0, D),
        (A, SU, T), (A, BO, BO), (A, TI, D), (A, 0, D)]
    got = _classify(o, [c[0] for c in cases], [c[1] for c in cases])
    assert got == [c[2] for c in cases]
    # without reverse lighting the shadow/AO rules flip (Renderer.cpp:366-411)
    got = _classify(o, [S, A, A], [SU, SU, TI], reverse=0)
    assert got == [D, D, T]


def test_exit_face_and_assign(oracle):
    box = abi.make_box((0, 0, 0), (1, 1, 1))
    ef = lambda *a: oracle.L.oracle_exit_face(C.byref(box), *map(float, a))  # noqa: E731
    assert ef(0.5, 0.5, 0.5, -1, 0, 0) == 0
    assert ef(0.5, 0.5, 0.5, 1, 0, 0) == 1
    assert ef(0.5, 0.5, 0.5, 0, -1, 0) == 2
    assert ef(0.5, 0.5, 0.5, 0, 1, 0) == 3
    assert ef(0.5, 0.5, 0.5, 0, 0, -1) == 4
    assert ef(0.5, 0.5, 0.5, 0, 0, 1) == 5
    assert ef(0.9, 0.5, 0.5, 1, 1, 0) == 1        # nearer x face
    assert ef(0.5, 0.5, 0.5, 0.00005, 0, 0) == 5  # no axis moves more than 1e-4: falls to the z test
    hr = abi.HostRays(4)
    hr["ox"], hr["oy"], hr["oz"] = 0.5, 0.5, 0.5
    hr["dx"] = np.array([1, -1, 1, -1], f32)
    hr["type"] = np.array([1, 1, 2, 4], np.int32)
    hr["classification"] = abi.RAY_BOUNDARY
    nb = np.array([-1, 7, -1, -1, -1, -1], np.int32)
    oracle.L.oracle_assign_destinations(4, C.byref(hr.ispc), C.byref(box), nb.ctypes.data, 1)
    assert list(hr["classification"]) == [7, abi.TERMINATED, 7, abi.DROP_ON_FLOOR]


def test_spawn_offsets_ao_block_then_shadow_block(oracle):
    hr = abi.HostRays(6)
    hr["type"] = np.array([1, 1, 2, 1, 1, 1], np.int32)
    hr["term"] = np.array([1, 0, 1, 3, 4, 1], np.int32)
    ao = np.zeros(6, np.int32)
    sh = np.zeros(6, np.int32)
    na, ns = C.c_int(), C.c_int()
    tot = oracle.L.oracle_spawn_offsets(6, C.byref(hr.ispc), 3, 2, 1, ao.ctypes.data, sh.ctypes.data, C.byref(na),
                                        C.byref(ns))
    # parents: rays 0, 3, 5 -> AO at 0, 3, 6 (running count for every ray); shadows after 9 AO slots
    assert list(ao) == [0, 3, 3, 3, 6, 6]
    assert list(sh) == [9, -1, -1, 11, -1, 13]
    assert (na.value, ns.value, tot) == (9, 6, 15)


def test_lighting_sequence_on_one_parent(oracle):
    """ambient -> AO spawn -> diffuse -> shadow spawn on one hand-built parent
    (TraceRays.ispc:625-923, reverse lighting), against the formulas evaluated
    here in float64."""
    hr = abi.HostRays(1)
    hr["type"], hr["term"] = 1, abi.RAY_SURFACE | abi.RAY_OPAQUE
    hr["ox"], hr["oy"], hr["oz"] = 0.0, 0.0, 3.0
    hr["dx"], hr["dy"], hr["dz"] = 0.0, 0.0, -1.0
    hr["t"] = 2.0
    hr["nx"], hr["ny"], hr["nz"] = 0.0, 0.0, 1.0
    hr["sr"], hr["sg"], hr["sb"] = 0.5, 0.25, 1.0
    hr["r"], hr["g"], hr["b"], hr["o"] = 0.1, 0.1, 0.1, 0.2
    Ld = scenes.LightingDesc(np.array([[0.0, 0.0, 5.0], [1.0, 0.0, 0.0]], f32), np.array([2, 0], np.int32),
                             Ka=0.4, Kd=0.6, n_ao=4, ao_radius=0.7, shadows=True)
    keep = []
    L = scenes.make_lighting(Ld, keep)
    out = abi.HostRays(6)
    ao = np.zeros(1, np.int32)
    sh = np.array([4], np.int32)
    oracle.L.oracle_ambientLighting(C.byref(L), 1, C.byref(hr.ispc))
    oracle.L.oracle_generateAORays(C.byref(L), 1, C.byref(hr.ispc), ao.ctypes.data, C.byref(out.ispc), 0.001)
    oracle.L.oracle_diffuseLighting(C.byref(L), 1, C.byref(hr.ispc))
    oracle.L.oracle_generateShadowRays(C.byref(L), 1, C.byref(hr.ispc), sh.ctypes.data, C.byref(out.ispc), 0.001)
    o0, sc = 0.2, np.array([0.5, 0.25, 1.0])
    amb = 0.1 + 0.4 * (1 - o0) * sc
    # diffuse: point light straight above (n.l = 1), directional (-L = (-1,0,0)) gives n.l = 0
    dif = amb + (0.6 / 2) * (1 - o0) * ((1 - o0) * 1.0 * sc)
    o1 = o0 + (0.6 / 2) * (1 - o0) * o0
    assert np.allclose([hr["r"][0], hr["g"][0], hr["b"][0]], dif, atol=1e-6)
    assert abs(hr["o"][0] - o1) < 1e-6
    # AO children: colour -Ka/nAO*(1-o0)*sc, origin at the hit + eps*n, cosine-weighted about +z
    assert np.allclose(out["r"][:4], -0.4 / 4 * (1 - o0) * sc[0], atol=1e-7)
    assert np.allclose(out["oz"][:4], 1.0 + 0.001, atol=1e-6)
    assert np.all(out["dz"][:4] > 0) and np.all(out["type"][:4] == abi.RAY_AO)
    assert np.allclose(out["tMax"][:4], 0.7)
    # shadow children: first towards the point light with -(Kd/nL)*(1-o1)*1*sc, second n.l = 0
    assert np.allclose([out["dx"][4], out["dy"][4], out["dz"][4]], [0, 0, 1], atol=1e-6)
    assert np.allclose(out["r"][4], -(0.6 / 2) * (1 - o1) * sc[0], atol=1e-6)
    assert out["r"][5] == 0.0 and out["dx"][5] == -1.0
    assert np.all(out["type"][4:] == abi.RAY_SHADOW) and np.all(np.isinf(out["tMax"][4:]))


def test_camera_rays_point_through_the_image_plane(oracle):
    cam = scenes.CameraDesc((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 64, 48)
    box = ((-1.0,) * 3, (1.0,) * 3)
    R = oracle.camera_rays(cam, box, box)
    assert 0 < R.n < 64 * 48
    d = np.stack([R["dx"], R["dy"], R["dz"]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # pixel -> image plane at distance 1/tan(aov/2), scaled by (min(w,h)-1)/2
    dist = 1 / math.tan(math.radians(22.5))
    px = (R["x"] - 31.5) / 23.5
    py = (R["y"] - 23.5) / 23.5
    exp = np.stack([-px, py, -dist * np.ones_like(px)], 1)   # right = cross(dir, up) = (-1,0,0)... mirrored
    exp /= np.linalg.norm(exp, axis=1, keepdims=True)
    right = np.cross([0, 0, -1], [0, 1, 0])
    exp = np.stack([px * right[0], py * 1.0, -dist * np.ones_like(px)], 1)
    exp /= np.linalg.norm(exp, axis=1, keepdims=True)
    assert np.allclose(d, exp, atol=2e-6)
    # every generated ray meets the box; every pixel whose ray meets it is generated
    t_hit = np.abs(d[:, 2]) > 0
    assert t_hit.all()


def test_triangle_hits_against_fp64_bruteforce(oracle):
    m = scenes.uv_sphere(1.0, 32, 20)
    V = np.ascontiguousarray(m.vertices, f32)
    I = np.ascontiguousarray(m.indices, np.uint32)
    n = 4000
    rng = np.random.default_rng(11)
    org = rng.uniform(-2, 2, (n, 3)).astype(f32)
    tgt = rng.uniform(-0.9, 0.9, (n, 3)).astype(f32)
    d = (tgt - org).astype(np.float64)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f32)
    t0 = np.zeros(n, f32)
    t64 = np.zeros(n)
    p64 = np.zeros(n, np.int32)
    oracle.L.oracle_triangles_bruteforce64(V.ctypes.data, I.ctypes.data, len(I), n, org.ctypes.data, d.ctypes.data,
                                           t0.ctypes.data, t64.ctypes.data, p64.ctypes.data)
    for use_bvh in (False, True):
        M = oracle.model_create()
        oracle.model_add_triangles(M, V, I, None, None, None)
        oracle.model_commit(M, use_bvh)
        t = np.zeros(n, f32)
        g = np.zeros(n, np.int32)
        p = np.zeros(n, np.int32)
        u = np.zeros(n, f32)
        v = np.zeros(n, f32)
        tf = np.full(n, 1e30, f32)
        oracle.L.oracle_model_intersect(M, n, org.ctypes.data, d.ctypes.data, t0.ctypes.data, tf.ctypes.data,
                                        t.ctypes.data, g.ctypes.data, p.ctypes.data, u.ctypes.data, v.ctypes.data)
        hit = p64 >= 0
        assert (p >= 0).sum() == hit.sum()
        agree = (p == p64).mean()
        assert agree >= 0.9999 or (p != p64).sum() <= 1
        assert np.allclose(t[hit], t64[hit], rtol=1e-5)
        # barycentrics reproduce the hit point
        pts = org[hit] + t[hit, None] * d[hit]
        tri = V[I[p[hit]]]
        rec = (1 - u[hit] - v[hit])[:, None] * tri[:, 0] + u[hit, None] * tri[:, 1] + v[hit, None] * tri[:, 2]
        assert np.abs(rec - pts).max() < 1e-4


def test_sphere_hit_closed_form(oracle):
    M = oracle.model_create()
    cen = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]], f32)
    D = np.array([0.0, 1.0], f32)
    oracle.model_add_spheres(M, cen, D, 0.5, 1.0, 0.0, 1.0, None)   # radii 0.5 and 1.0 by data
    oracle.model_commit(M, True)
    org = np.array([[0, 0, 5], [3, 0, 5], [0, 0, 0]], f32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, 1]], f32)
    n = 3
    t0 = np.zeros(n, f32)
    tf = np.full(n, 1e30, f32)
    t = np.zeros(n, f32)
    g = np.zeros(n, np.int32)
    p = np.zeros(n, np.int32)
    u = np.zeros(n, f32)
    v = np.zeros(n, f32)
    oracle.L.oracle_model_intersect(M, n, org.ctypes.data, d.ctypes.data, t0.ctypes.data, tf.ctypes.data,
                                    t.ctypes.data, g.ctypes.data, p.ctypes.data, u.ctypes.data, v.ctypes.data)
    assert list(p) == [0, 1, 0]
    assert np.allclose(t, [4.5, 4.0, 0.5], atol=1e-5)   # the ray from inside exits at t_out


def test_sampler_crossings_on_a_ramp(oracle):
    """A z-ramp v = z: every crossing of isovalue c along a +z ray is at
    z = c; the sampler emits one particle per crossing, t = tHit + 0.001."""
    n = 33
    z = np.linspace(0, 1, n, dtype=f32)
    data = np.broadcast_to(z[:, None, None], (n, n, n)).astype(f32).copy()
    vol = scenes.VolumeDesc(data, (0.0, 0.0, 0.0), (1.0 / (n - 1),) * 3, local_box=((0,) * 3, (1,) * 3),
                            global_box=((0,) * 3, (1,) * 3))
    vv = scenes.VolumeVisDesc(vol, isovalues=np.array([0.37], f32))
    sc = scenes.SceneDesc(volume_vis=[vv], local_box=vol.local_box, global_box=vol.global_box)
    B = scenes.build_scene(oracle, sc)
    hr = abi.HostRays.from_arrays(np.array([[0.5, 0.5, -0.5], [0.25, 0.7, 1.5]], f32),
                                  np.array([[0, 0, 1], [0, 0, -1]], f32))
    xyz = np.zeros((16, 3), f32)
    passes = C.c_longlong()
    cnt = oracle.L.oracle_sample_raylist(C.byref(B.vis), hr.n, C.byref(hr.ispc), xyz.ctypes.data, 16,
                                         C.byref(passes))
    assert cnt == 2
    assert abs(xyz[0, 2] - (0.37 + 0.001)) < 1e-4 and abs(xyz[1, 2] - (0.37 - 0.001)) < 1e-4


def test_fp32_compositing_within_1e4_of_fp64(oracle, radial_vols_small):
    vol = radial_vols_small["radial-eightBalls.vol"]
    vv = scenes.VolumeVisDesc(vol, colormap=np.array([[0, 1, .5, .5], [.6, .5, 1, .5], [1.2, .5, .5, 1]], f32),
                              opacitymap=np.array([[0, .08], [.4, .01], [1.2, .05]], f32), volume_render=True,
                              isovalues=np.array([0.15], f32))
    sc = scenes.SceneDesc(volume_vis=[vv], local_box=vol.local_box, global_box=vol.global_box)
    B = scenes.build_scene(oracle, sc)
    hr = random_rays(20000, seed=2)
    h32 = hr.copy()
    oracle.trace(B.vis, h32)
    h64 = hr.copy()
    res = oracle.trace(B.vis, h64, fp64=True)
    rgba = np.stack([h32["r"], h32["g"], h32["b"], h32["o"]], 1)
    assert np.abs(rgba - res["rgba64"]).max() < 1e-4
    assert (h32["term"] == h64["term"]).mean() >= 0.9999


def test_factor_and_partition():
    assert scenes.factor(1) == (1, 1, 1)
    assert scenes.factor(8) == (2, 2, 2)
    assert scenes.factor(2) == (1, 1, 2)
    assert scenes.factor(7) == (1, 1, 7)
    parts = scenes.partition(8, (2, 2, 2), (10, 10, 10))
    assert parts[0]["offsets"] == (1, 1, 1) and parts[0]["counts"] == (5, 5, 5)
    assert parts[7]["offsets"] == (5, 5, 5) and parts[7]["counts"] == (4, 4, 4)
    assert parts[7]["gcounts"] == (6, 6, 6)


def test_walk_counts_fixture_reproduces():
    """The recorded BVH work per ray of the sphere-mesh frame (the geometry
    roofline's algorithmic flops) is what the oracle's walk counts today."""
    import sys
    sys.path.insert(0, os.path.join(GOLDEN))
    import make_walk_counts as mw
    ref = json.load(open(os.path.join(GOLDEN, "walk_counts.json")))
    make, stride, ao = mw.scenes_to_count()["sphere_mesh_512"]
    assert mw.walk_counts(make(), stride, ao) == ref["sphere_mesh_512"]
    # every recorded ray kind did some work: a box test per walk at least, no more primitive tests than boxes
    for scene in ref.values():
        for kind in ("primary", "shadow", "ao"):
            if kind in scene:
                c = scene[kind]
                assert c["rays"] > 0 and c["box_tests"] >= c["rays"] and c["prim_tests"] <= c["box_tests"]
