"""Path lines on the GPU (gxy_model_add_pathlines): hits, depths, normals and
colours equal the oracle's on random rays; a straight constant-radius path
line is the closed-form capsule; every hit lies on the FP64 swept-sphere
surface and within the stated tolerance of the true swept-sphere surface of
the curves; a lit frame with shadow and AO rays matches the oracle's."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_rays
from galaxy_amd import abi, scenes
from pathlines_common import (capsule_hits, controls, curve_hit_error, helix_desc, pieces64, rays_toward,
                              segment_controls, swept_distance)

pytestmark = pytest.mark.gpu
f32 = np.float32


def gpu_hits(hip, built, hr):
    """Traces host rays on the GPU; returns (downloaded rays, geomID, primID)."""
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    n = hr.n
    gd, pd = glib.DeviceArray(4 * n), glib.DeviceArray(4 * n)
    glib.check(L_.gxy_model_set_hit_output(built.vis.model, gd.ptr, pd.ptr))
    try:
        d = glib.DeviceRays.from_host(hr)
        assert hip.trace(built.vis, d) == ""
        g = d.download()
        return g, gd.to_host(np.int32), pd.to_host(np.int32)
    finally:
        glib.check(L_.gxy_model_set_hit_output(built.vis.model, None, None))


@pytest.fixture(params=[(2, 0), (4, 0), (2, 8)], ids=["binary", "wide", "binary-8-waves"])
def bvh_width(request):
    """Runs the test with path lines walked in the binary and the 4-wide trees,
    and in the binary tree at 8 waves per SIMD."""
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    saved = L_.gxy_get_bvh_width()
    glib.check(L_.gxy_set_bvh_width(request.param[0]))
    glib.check(L_.gxy_set_geometry_waves(request.param[1]))
    yield request.param
    L_.gxy_set_bvh_width(saved)
    L_.gxy_set_geometry_waves(0)


def scene_with(pl, meshes=(), box=1.5):
    b = ((-box,) * 3, (box,) * 3)
    return scenes.SceneDesc(pathlines=[pl], meshes=list(meshes), local_box=b, global_box=b)


def test_pathlines_match_oracle(hip, oracle, bvh_width):
    mesh = scenes.icosphere(3, 0.25, (0.9, -0.9, 0.0))
    sc = scene_with(helix_desc(seed=3), [mesh])
    Bh, Bo = scenes.build_scene(hip, sc), scenes.build_scene(oracle, sc)
    hr = random_rays(60000, seed=12, lo=-1.4, hi=1.4)
    g, gg, gp = gpu_hits(hip, Bh, hr)
    o = hr.copy()
    res = oracle.trace(Bo.vis, o, geom_ids=True)
    hits = res["geom"] >= 0
    # geomID 0 is the mesh (meshes are added first), 1 the path lines
    assert (res["geom"] == 1).sum() > 500 and (res["geom"] == 0).sum() > 100
    agree = (gg == res["geom"]) & (gp == res["prim"])
    assert agree.mean() >= 0.9999
    assert np.allclose(g["t"][hits & agree], o["t"][hits & agree], rtol=1e-5)
    prim = agree & hits & (o["type"] == 1)
    for f in ("nx", "ny", "nz", "sr", "sg", "sb"):
        assert np.abs(g[f][prim] - o[f][prim]).max() <= 1e-5, f
    assert (g["term"] == o["term"]).mean() >= 0.9999


def test_straight_pathline_is_a_capsule(hip):
    V = np.array([[0.25 * k, 0.0, 0.0] for k in range(5)], f32)
    pl = scenes.PathLinesDesc(V, np.zeros(5, f32), np.arange(4, dtype=np.int32), radius0=0.1, radius1=0.3)
    Bh = scenes.build_scene(hip, scene_with(pl, box=3.0))
    org, d = rays_toward(8000, 5, (-0.15, -0.15, -0.15), (1.15, 0.15, 0.15))
    hr = abi.HostRays.from_arrays(org, d, tmax=np.full(len(org), 1e30, f32))
    hr["type"] = np.full(hr.n, 2, np.int32)      # shadow type: no shading, depth only
    g, gg, gp = gpu_hits(hip, Bh, hr)
    ref = capsule_hits(org, d, (0, 0, 0), (1, 0, 0), 0.1)
    hit = np.isfinite(ref)
    got = gg >= 0
    assert hit.sum() > 2000 and (got == hit).mean() >= 0.999
    both = got & hit
    assert np.allclose(g["t"][both], ref[both], rtol=2e-5, atol=2e-6)


def test_hits_lie_on_the_swept_surface(hip, oracle):
    pl = helix_desc(seed=4, nlines=4, nverts=10)
    Bh = scenes.build_scene(hip, scene_with(pl))
    org, d = rays_toward(4000, 21, (-0.9, -0.9, -0.7), (0.9, 0.9, 0.7))
    hr = abi.HostRays.from_arrays(org, d, tmax=np.full(len(org), 1e30, f32))
    g, gg, gp = gpu_hits(hip, Bh, hr)
    hit = gg >= 0
    assert hit.sum() > 300
    # the control points are built on the host by the same rules in both: take the oracle's
    M = oracle.model_create()
    oracle.model_add_pathlines(M, pl.vertices, pl.data, pl.indices, pl.radius0, pl.radius1, pl.value0, pl.value1,
                               None)
    cp, ic = controls(oracle, M, 0, len(pl.indices))
    a, b = pieces64(cp, ic)
    x = org[hit].astype(np.float64) + g["t"][hit, None].astype(np.float64) * d[hit]
    assert np.abs(swept_distance(x, a[gp[hit]], b[gp[hit]])).max() <= 2e-5
    # unit normals facing the ray
    nrm = np.c_[g["nx"][hit], g["ny"][hit], g["nz"][hit]]
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-5)
    assert (np.einsum("ij,ij->i", nrm, d[hit]) < 0).all()


def test_hits_within_tolerance_of_the_true_sweep(hip, oracle):
    """GPU hits on the first 150 lines of the 4000-line path-line workload lie
    within CURVE_TOL x (least control radius of the hit segment) of the
    surface the spheres (p(u), r(u)) of the round Bezier segments sweep
    (DataDrivenPathLines.ispc:319-331), measured in FP64."""
    from galaxy_amd import workloads
    full = workloads.pathlines()[0].pathlines[0]
    nl, nv = 150, 64
    pl = scenes.PathLinesDesc(full.vertices[:nl * nv], full.data[:nl * nv], full.indices[:nl * (nv - 1)],
                              radius0=full.radius0, radius1=full.radius1, value0=full.value0, value1=full.value1,
                              colormap=full.colormap)
    Bh = scenes.build_scene(hip, scene_with(pl))
    org, d = rays_toward(12000, 4, pl.vertices.min(0), pl.vertices.max(0), dist=2.0)
    hr = abi.HostRays.from_arrays(org, d, tmax=np.full(len(org), 1e30, f32))
    hr["type"] = np.full(hr.n, 2, np.int32)
    g, gg, gp = gpu_hits(hip, Bh, hr)
    hit = gg >= 0
    assert hit.sum() > 3000
    M = oracle.model_create()
    oracle.model_add_pathlines(M, pl.vertices, pl.data, pl.indices, pl.radius0, pl.radius1, pl.value0, pl.value1,
                               None)
    cp, ic = controls(oracle, M, 0, len(pl.indices))
    x = org[hit].astype(np.float64) + g["t"][hit, None].astype(np.float64) * d[hit].astype(np.float64)
    err, tol = curve_hit_error(x, gp[hit], segment_controls(cp, ic), pl.indices)
    print("%d GPU hits: distance to the swept surface max %.3g mean %.3g, max err/tol %.3f"
          % (hit.sum(), err.max(), err.mean(), (err / tol).max()), flush=True)
    assert (err <= tol).all()


def test_pathline_frame_with_shadows_and_ao_matches_oracle(hip, oracle, bvh_width):
    box = ((-1.2,) * 3, (1.2,) * 3)
    L = scenes.LightingDesc(np.array([[3.0, 4.0, 5.0]], f32), np.array([2], np.int32), Ka=0.3, Kd=0.7, n_ao=2,
                            ao_radius=0.3, shadows=True)
    cam = scenes.CameraDesc((0.4, 0.3, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 160, 120)
    sc = scenes.SceneDesc(pathlines=[helix_desc(seed=6)], local_box=box, global_box=box, lighting=L, cameras=[cam])
    fb_g, st_g = hip.render_frame(scenes.build_scene(hip, sc), cam)
    fb_o, st_o = oracle.render_frame(scenes.build_scene(oracle, sc), cam)
    assert st_g.primary_rays == st_o.primary_rays and st_g.secondary_rays > 1000
    assert abs(st_g.secondary_rays - st_o.secondary_rays) <= max(3, st_o.secondary_rays // 5000)
    assert np.mean(np.abs(fb_g - fb_o) <= 1e-4) >= 0.999


def test_batch_rejects_pathline_datasets(hip):
    from galaxy_amd import batch as gb
    from galaxy_amd import lib as glib
    Bh = scenes.build_scene(hip, scene_with(helix_desc(seed=7, nlines=1, nverts=4)))
    B = gb.Batch(256)
    with pytest.raises(glib.GxyError, match="path lines"):
        B.add_dataset(Bh)
    B.destroy()
