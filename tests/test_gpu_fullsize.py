"""The volume-render frame at its full size (1920x1080 rays, 512^3 noise
volume, DVR + iso 0.5 + shadow rays, opacity threshold 0.99): the GPU frame
equals the oracle's on every 8th image row (the oracle traces those rows'
primary rays and their shadow rays through the same voxels), and rendering
the frame twice gives the identical image."""
import ctypes as C

import numpy as np
import pytest

from galaxy_amd import abi, scenes, workloads

pytestmark = pytest.mark.gpu
f32 = np.float32


def test_full_size_volume_frame_matches_oracle_on_sampled_rows(hip, oracle):
    from galaxy_amd import lib as glib
    from oracle_backend import lib as olib
    L = glib.lib()
    sc, alpha = workloads.volume_render()
    cam = sc.cameras[0]
    saved = L.gxy_get_opacity_threshold()
    glib.check(L.gxy_set_opacity_threshold(alpha))
    olib().oracle_set_opacity_threshold(alpha)
    try:
        Bh = scenes.build_scene(hip, sc)
        fb1, st1 = hip.render_frame(Bh, cam)
        fb2, st2 = hip.render_frame(Bh, cam)
        assert np.array_equal(fb1, fb2) and st1.as_dict() == st2.as_dict()
        assert st1.primary_rays > 1_000_000 and st1.secondary_rays > 500_000
        vol = sc.volume_vis[0].volume
        vol.voxels = hip.volume_download(Bh.handles[0], vol.dims)
        Bo = scenes.build_scene(oracle, sc)
        rays = oracle.camera_rays(cam, sc.local_box, sc.global_box)
        assert rays.n == st1.primary_rays
        sub = rays.subset(np.nonzero(rays["y"] % 8 == 0)[0])
        fbo = np.zeros(cam.width * cam.height * 4, f32)
        so = abi.FrameStats()
        nb = np.full(6, -1, np.int32)
        oracle.L.oracle_process_raylist(C.byref(Bo.vis), C.byref(Bo.lights), sub.n, C.byref(sub.ispc),
                                        nb.ctypes.data, sc.epsilon, 1 << 30, fbo.ctypes.data, cam.width, cam.height,
                                        C.byref(so))
        g = fb1.reshape(cam.height, cam.width, 4)[::8]
        o = fbo.reshape(cam.height, cam.width, 4)[::8]
        assert so.ray_passes > 200_000
        assert np.abs(g - o).max() <= 1e-4
        assert np.sqrt(np.mean((g - o) ** 2)) <= 1e-5
    finally:
        L.gxy_set_opacity_threshold(saved)
        olib().oracle_set_opacity_threshold(0.999)
