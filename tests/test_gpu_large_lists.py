"""Lists above 4 M rays, whose per-workgroup count arrays go through the
device-wide scan instead of the one-workgroup scan: spawn offsets and the
destination partition of 5 M rays equal the oracle's exactly, and a 4K
frame traced as whole lists (camera 8.3 M pixels, 4.7 M primary rays) gives
the counts and framebuffer of the same frame in lists of at most 2 M rays."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_rays
from galaxy_amd import scenes, workloads

pytestmark = pytest.mark.gpu
N = 5_000_000


def big_rays(seed):
    hr = random_rays(N, seed)
    rng = np.random.default_rng(seed + 1)
    hr["term"] = rng.choice(np.array([0, 1, 2, 4], np.int32), N)                  # RAY_SURFACE = 1
    hr["classification"] = rng.choice(np.array([-1, -2, -3, 0, 1, 2, 3, 4, 5], np.int32), N)
    return hr


def test_spawn_offsets_of_5m_rays(oracle):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    hr = big_rays(31)
    d = glib.DeviceRays.from_host(hr)
    ao_d, sh_d = glib.DeviceArray(4 * N), glib.DeviceArray(4 * N)
    na, ns = C.c_int(), C.c_int()
    glib.check(L_.gxy_spawn_offsets(N, C.byref(d.ispc), 4, 2, 1, ao_d.ptr, sh_d.ptr, C.byref(na), C.byref(ns)))
    ao_o, sh_o = np.zeros(N, np.int32), np.zeros(N, np.int32)
    na_o, ns_o = C.c_int(), C.c_int()
    oracle.L.oracle_spawn_offsets(N, C.byref(hr.ispc), 4, 2, 1, ao_o.ctypes.data, sh_o.ctypes.data, C.byref(na_o),
                                  C.byref(ns_o))
    assert (na.value, ns.value) == (na_o.value, ns_o.value) and na.value > 1_000_000
    assert np.array_equal(ao_d.to_host(np.int32), ao_o) and np.array_equal(sh_d.to_host(np.int32), sh_o)


def test_partition_of_5m_rays(oracle):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    hr = big_rays(32)
    d = glib.DeviceRays.from_host(hr)
    ndest = 6
    counts, oc = np.zeros(ndest + 1, np.int32), np.zeros(ndest + 1, np.int32)
    oracle.L.oracle_partition_counts(N, C.byref(hr.ispc), ndest, oc.ctypes.data)
    lists = [glib.DeviceRays(max(1, int(c))) for c in oc]
    ptrs = (C.c_void_p * ndest)(*[C.cast(C.pointer(lists[k].ispc), C.c_void_p) for k in range(ndest)])
    glib.check(L_.gxy_partition_rays(N, C.byref(d.ispc), ndest, ptrs, C.byref(lists[ndest].ispc),
                                     counts.ctypes.data))
    assert np.array_equal(counts, oc) and counts.min() > 100_000
    cls = hr["classification"]
    for k, want in ((0, 0), (5, 5), (ndest, -3)):
        lists[k].n = int(counts[k])
        got = lists[k].download(fields=("ox", "dy", "x", "classification"))
        idx = np.nonzero(cls == want)[0]
        for f in ("ox", "dy", "x", "classification"):
            assert np.array_equal(got[f], hr[f][idx]), (k, f)


def test_4k_frame_in_whole_lists_equals_split_lists(hip):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    sc, alpha = workloads.nine_icospheres(3840, 2160, subdiv=5, n_ao=2)
    glib.check(L_.gxy_set_opacity_threshold(alpha))
    built = scenes.build_scene(hip, sc)
    cam = sc.cameras[0]
    fb1, st1 = hip.render_frame(built, cam, eps=sc.epsilon, maxlist=1 << 30)
    fb2, st2 = hip.render_frame(built, cam, eps=sc.epsilon, maxlist=1 << 21)
    assert st1.primary_rays > 4_500_000 and st1.secondary_rays > st1.primary_rays
    for f in ("primary_rays", "secondary_rays", "ray_passes", "terminated", "dropped", "samples"):
        assert getattr(st1, f) == getattr(st2, f), f
    assert np.abs(fb1 - fb2).max() <= 1e-5 * max(1.0, np.abs(fb1).max())
