"""gxy_spawn_offsets_noalloc — no counterpart in the reference; the spawn
offsets of TraceRays.cpp:89-117 for callers that cannot allocate — against
gxy_spawn_offsets through the C-ABI, exactly, on calls in a row of falling
size (the largest list the static scratch takes first, so a smaller call after
it would expose any entry it failed to rewrite), with the device bytes in use
unchanged by the call, and the host rejecting a list beyond the limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_rays

pytestmark = pytest.mark.gpu

MAX_RAYS = 16383 * 256       # GXY_NOALLOC_MAX_RAYS in include/gxy_hip.h


def _rays(n, seed):
    hr = random_rays(n, seed)
    rng = np.random.default_rng(seed + 1)
    hr["term"] = rng.choice(np.array([0, 1, 2, 4], np.int32), n)      # RAY_SURFACE = 1
    return hr


def _offsets(fn, d, n, nao, nl, sh):
    from galaxy_amd import lib as glib
    ao, so = glib.DeviceArray(4 * max(n, 1)), glib.DeviceArray(4 * max(n, 1))
    na, ns = C.c_int(-7), C.c_int(-7)
    glib.check(fn(n, C.byref(d.ispc), nao, nl, sh, ao.ptr, so.ptr, C.byref(na), C.byref(ns)))
    return ao.to_host(np.int32, n), so.to_host(np.int32, n), na.value, ns.value


def test_noalloc_offsets_equal_the_allocating_entry_point(oracle):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    for n, seed, nao, nl, sh in [(MAX_RAYS, 41, 4, 2, 1), (70_001, 42, 16, 1, 1), (1000, 43, 3, 2, 0),
                                 (1, 44, 2, 1, 1)]:
        print("noalloc offsets: %d rays, nAO %d, nLights %d, shadows %d" % (n, nao, nl, sh), flush=True)
        hr = _rays(n, seed)
        d = glib.DeviceRays.from_host(hr)
        ref = _offsets(L_.gxy_spawn_offsets, d, n, nao, nl, sh)
        ao, so = glib.DeviceArray(4 * n), glib.DeviceArray(4 * n)
        na, ns = C.c_int(-7), C.c_int(-7)
        before = L_.gxy_device_bytes_in_use()
        glib.check(L_.gxy_spawn_offsets_noalloc(n, C.byref(d.ispc), nao, nl, sh, ao.ptr, so.ptr, C.byref(na),
                                                C.byref(ns)))
        assert L_.gxy_device_bytes_in_use() == before
        got = (ao.to_host(np.int32, n), so.to_host(np.int32, n), na.value, ns.value)
        assert got[2:] == ref[2:]
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        ao_o, sh_o = np.zeros(n, np.int32), np.zeros(n, np.int32)
        na_o, ns_o = C.c_int(), C.c_int()
        oracle.L.oracle_spawn_offsets(n, C.byref(hr.ispc), nao, nl, sh, ao_o.ctypes.data, sh_o.ctypes.data,
                                      C.byref(na_o), C.byref(ns_o))
        assert got[2:] == (na_o.value, ns_o.value)
        assert np.array_equal(got[0], ao_o) and np.array_equal(got[1], sh_o)


def test_noalloc_empty_list_and_limit():
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    d = glib.DeviceRays.from_host(_rays(64, 45))
    na, ns = C.c_int(-7), C.c_int(-7)
    ao, so = glib.DeviceArray(256), glib.DeviceArray(256)
    glib.check(L_.gxy_spawn_offsets_noalloc(0, C.byref(d.ispc), 4, 1, 1, ao.ptr, so.ptr, C.byref(na), C.byref(ns)))
    assert (na.value, ns.value) == (0, 0)
    # rejected on the host before any launch: the list is never read
    rc = L_.gxy_spawn_offsets_noalloc(MAX_RAYS + 1, C.byref(d.ispc), 4, 1, 1, ao.ptr, so.ptr, C.byref(na),
                                      C.byref(ns))
    assert rc != 0 and b"static scratch limit" in L_.gxy_last_error()
