"""The batched form at the configured sizes of the batched-lists workload:
256 ray lists of 1k-1M rays (log-uniform, 43 M rays) over 64 resident
datasets (noise volumes of 128^3-1024^3 voxels and blob meshes of 50 k-5 M
triangles), shadow + 2 AO rays per surface hit.  Every list's framebuffer
equals the per-list loop's on the same rays (only the order of FP32
additions into a pixel differs), the batch totals equal the loop's summed
statistics exactly, and the smallest lists on the smallest datasets of each
kind equal the oracle's."""
import ctypes as C

import numpy as np
import pytest

from galaxy_amd import abi, scenes, workloads

pytestmark = pytest.mark.gpu
f32 = np.float32


def test_batched_lists_at_configured_sizes(hip, oracle):
    from galaxy_amd import batch as gb
    from galaxy_amd import lib as glib
    L = glib.lib()
    descs, lists, Ld = workloads.batched_lists()
    built = [scenes.build_scene(hip, workloads.dataset_scene(d)) for d in descs]
    glib.check(L.gxy_synchronize())
    keep = []
    lights = scenes.make_lighting(Ld, keep)
    B = gb.Batch(256)
    ids = [B.add_dataset(b) for b in built]
    B.set_lighting(lights)
    rays = []
    for n, ds, seed in lists:
        d = gb.random_rays(n, seed, width=n)
        l, fb = B.add_render_list(ids[ds], d, n, 1)
        B.set_seed(l, seed)
        rays.append((d, fb, n, ds, seed))
    st = B.run(2)
    assert st.overflow == 0
    total = sum(n for n, _, _ in lists)
    print("batched: %d lists, %d primary rays, %d ray passes, %d samples" % (len(lists), total, st.rays, st.samples),
          flush=True)
    assert total > 30_000_000 and st.spawned > 1_000_000
    got = [fb.to_host(f32) for (_, fb, _, _, _) in rays]
    # the per-list loop on the same rays
    nb = np.full(6, -1, np.int32)
    lo, hi = glib.a3((0, 0, 0)), glib.a3((1, 1, 1))
    tot = dict(rays=0, samples=0, spawned=0, terminated=0)
    worst = 0.0
    for k, (d, fb, n, ds, seed) in enumerate(rays):
        glib.check(L.gxy_rays_random(C.byref(d.ispc), n, seed, n, 1, lo.ctypes.data, hi.ctypes.data))
        fb2 = glib.DeviceArray(16 * n).zero()
        fs = abi.FrameStats()
        glib.check(L.gxy_process_raylist(C.byref(built[ds].vis), C.byref(lights), n, C.byref(d.ispc), nb.ctypes.data,
                                          0.001, 1 << 30, fb2.ptr, n, 1, C.byref(fs)))
        ref = fb2.to_host(f32)
        fb2.free()
        worst = max(worst, float(np.abs(got[k] - ref).max()))
        tot["rays"] += fs.ray_passes
        tot["samples"] += fs.samples
        tot["spawned"] += fs.secondary_rays
        tot["terminated"] += fs.terminated
    print("per-list loop: max framebuffer difference %.3g" % worst, flush=True)
    assert worst <= 1e-5
    assert (st.rays, st.samples, st.spawned, st.terminated) == (tot["rays"], tot["samples"], tot["spawned"],
                                                               tot["terminated"])
    # the oracle on the smallest list of the smallest dataset of each kind
    for kind in ("volume", "mesh"):
        cand = [k for k, (n, ds, _) in enumerate(lists) if descs[ds][0] == kind]
        smallest_ds = min({lists[k][1] for k in cand}, key=lambda ds: descs[ds][1])
        k = min((k for k in cand if lists[k][1] == smallest_ds), key=lambda k: lists[k][0])
        n, ds, seed = lists[k]
        sc = workloads.dataset_scene(descs[ds])
        if kind == "volume":
            vol = sc.volume_vis[0].volume
            vol.voxels = hip.volume_download(built[ds].handles[0], vol.dims)
        Bo = scenes.build_scene(oracle, sc)
        hr = gb.random_rays(n, seed, width=n).download()
        fbo = np.zeros(4 * n, f32)
        so = abi.FrameStats()
        oracle.L.oracle_process_raylist(C.byref(Bo.vis), C.byref(lights), hr.n, C.byref(hr.ispc), nb.ctypes.data,
                                        0.001, 1 << 30, fbo.ctypes.data, n, 1, C.byref(so))
        print("%s dataset %s, list of %d rays: oracle ray passes %d" % (kind, descs[ds][1], n, so.ray_passes),
              flush=True)
        assert np.abs(got[k] - fbo).max() <= 1e-4, kind
    B.destroy()
    for b in built:
        for h in b.handles:
            for fn in (L.gxy_volume_destroy, L.gxy_tf_destroy, L.gxy_model_destroy):
                fn(h)
