"""The served form at the sizes of the served-stream workload: a 1M-ray
sample request over a 512^3 noise volume, a 1024x1024 render of a
1M-triangle mesh, a 512x512 render of the volume and a smaller sample
request, arriving together and served with continuous batching.  Every
request gives the result it gets alone, and each checks against the oracle on
a sampled subset: render rows (the oracle traces those rows' camera rays and
their shadow and AO children through the same scene), and the particles of
every 100th sample ray (each oracle particle has a served particle within
1e-5)."""
import ctypes as C

import numpy as np
import pytest

from galaxy_amd import abi, scenes, workloads

pytestmark = pytest.mark.gpu
f32 = np.float32

N_SAMPLE = 1_000_000


def _requests():
    from galaxy_amd import serve
    outs = [np.zeros(24_000_000 * 3, f32), np.zeros(1024 * 1024 * 4, f32), np.zeros(512 * 512 * 4, f32),
            np.zeros(2_000_000 * 3, f32)]
    reqs = (abi.Request * 4)()
    reqs[0] = serve.sample_request(0, N_SAMPLE, 4242, 24_000_000, arrival=0.0, out=outs[0])
    reqs[1] = serve.render_request(1, 1024, 1024, (1.9, 1.2, 1.6), arrival=0.002, out=outs[1])
    reqs[2] = serve.render_request(0, 512, 512, (-1.2, 0.9, 1.8), arrival=0.004, out=outs[2])
    reqs[3] = serve.sample_request(0, 200_000, 77, 2_000_000, arrival=0.006, out=outs[3])
    return reqs, outs


def _particles(req, res, out):
    c = int(res.particles)
    assert c <= req.capacity
    a = out[:3 * c].reshape(-1, 3).copy()
    return a[np.lexsort(a.T)]


def _rows_close(img, ref, w, h, every):
    g = img.reshape(h, w, 4)[::every].astype(np.float64)
    o = ref.reshape(h, w, 4)[::every]
    d = np.abs(g - o)
    print("rows: %d values, within 1e-4 %.6f, max diff %.3g" % (d.size, np.mean(d <= 1e-4), d.max()), flush=True)
    return np.mean(d <= 1e-4) >= 0.9999 and np.sqrt(np.mean(d ** 2)) <= 1e-3


def test_served_requests_at_configured_sizes(hip, oracle):
    from scipy.spatial import cKDTree

    from galaxy_amd import batch as gb
    from galaxy_amd import lib as glib
    from galaxy_amd import serve
    vol_sc = workloads.unit_volume_scene(512, 91)
    mesh_sc = workloads.blob_scene(1_000_000, 92)
    S, built = serve.build_server(hip, [vol_sc, mesh_sc], budget=64 << 30, slots=16, layout=8)
    reqs, outs = _requests()
    res, st = S.run(reqs, abi.SERVE_CONTINUOUS)
    assert all(res[k].status == abi.RQ_DONE for k in range(4))
    assert st.max_in_flight >= 2
    print("served: %d ray passes, %d particles + %d particles" % (st.rays, res[0].particles, res[3].particles),
          flush=True)
    assert res[0].particles > N_SAMPLE and res[1].rays > 500_000
    mix = [_particles(reqs[0], res[0], outs[0]), outs[1].copy(), outs[2].copy(), _particles(reqs[3], res[3], outs[3])]
    # each request alone gives what it gets in the mix
    for k in range(4):
        one = (abi.Request * 1)()
        one[0] = reqs[k]
        one[0].arrival = 0.0
        outs[k][:] = 0
        r1, _ = S.run(one, abi.SERVE_CONTINUOUS)
        assert r1[0].status == abi.RQ_DONE
        if reqs[k].kind == 1:
            assert int(r1[0].particles) == int(res[k].particles), k
            assert np.array_equal(_particles(one[0], r1[0], outs[k]), mix[k]), k
        else:
            assert np.abs(outs[k] - mix[k]).max() <= 1e-5, k
    # the oracle on the same scenes
    lights = workloads.batch_lighting()
    vol = vol_sc.volume_vis[0].volume
    vol.voxels = hip.volume_download(built[0].handles[0], vol.dims)
    Bo = [scenes.build_scene(oracle, sc, lights) for sc in (vol_sc, mesh_sc)]
    nb = np.full(6, -1, np.int32)
    for k, every in ((1, 16), (2, 8)):
        r = reqs[k]
        sc = (vol_sc, mesh_sc)[r.dataset]
        cam = scenes.CameraDesc(tuple(r.eye), tuple(r.center), tuple(r.up), r.aov, r.width, r.height)
        rays = oracle.camera_rays(cam, sc.local_box, sc.global_box)
        sub = rays.subset(np.nonzero(rays["y"] % every == 0)[0])
        fbo = np.zeros(r.width * r.height * 4, f32)
        so = abi.FrameStats()
        oracle.L.oracle_process_raylist(C.byref(Bo[r.dataset].vis), C.byref(Bo[r.dataset].lights), sub.n,
                                        C.byref(sub.ispc), nb.ctypes.data, 0.001, 1 << 30, fbo.ctypes.data, r.width,
                                        r.height, C.byref(so))
        assert so.ray_passes > 50_000
        assert _rows_close(mix[k], fbo, r.width, r.height, every), k
    # sample rays: the generator the server uses for the request (same seed and box)
    lo, hi = workloads.UNIT_BOX
    d = gb.random_rays(N_SAMPLE, reqs[0].seed, width=N_SAMPLE, lo=lo, hi=hi)
    hr = d.download()
    sub = hr.subset(np.arange(0, N_SAMPLE, 100))
    cap = 400_000
    xyz = np.zeros((cap, 3), f32)
    passes = C.c_longlong()
    cnt = oracle.L.oracle_sample_raylist(C.byref(Bo[0].vis), sub.n, C.byref(sub.ispc), xyz.ctypes.data, cap,
                                         C.byref(passes))
    assert 0 < cnt < cap
    dist, _ = cKDTree(mix[0].astype(np.float64)).query(xyz[:cnt].astype(np.float64))
    print("sample subset: %d oracle particles, within 1e-5 of a served one: %.6f" % (cnt, np.mean(dist <= 1e-5)),
          flush=True)
    assert np.mean(dist <= 1e-5) >= 0.9999
