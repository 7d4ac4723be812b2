"""Pixel messages on the GPU (gxy_pixels_pack / gxy_pixels_apply): the packed
SendPixelsMsg bytes equal the oracle's bit for bit (empty, one-ray, ragged
and large lists), an undersized buffer is refused with nothing written, and
applying a message (duplicate pixels, frame stamps, older frames, pixels
outside the framebuffer) matches the oracle and direct accumulation."""
import ctypes as C

import numpy as np
import pytest

from test_oracle_pixels import HDR, PIX, classified_rays, pack

pytestmark = pytest.mark.gpu
f32 = np.float32


def gpu_pack(hr, cap, rkey=0x1122334455667788, rskey=-5, frame=7, source=3, fill=0xAB):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    d = glib.DeviceRays.from_host(hr)
    msg = glib.DeviceArray.from_host(np.full(cap, fill, np.uint8))
    cnt = C.c_int(-1)
    r = L_.gxy_pixels_pack(hr.n, C.byref(d.ispc), rkey, rskey, frame, source, msg.ptr, cap, C.byref(cnt))
    return r, cnt.value, msg.to_host(np.uint8)


@pytest.mark.parametrize("n", [0, 1, 777, 100000])
def test_pack_equals_oracle_bytes(oracle, n):
    hr = classified_rays(n, 11 + n)
    cap = 32 + 24 * max(n, 1)
    nbytes, ref = pack(oracle, hr, cap=cap)
    r, cnt, got = gpu_pack(hr, cap)
    assert r == 0 and cnt == (hr["classification"] == -1).sum() and nbytes == 32 + 24 * cnt
    assert np.array_equal(got[:nbytes], ref[:nbytes])
    # the rest of the buffer is untouched
    assert (got[nbytes:] == 0xAB).all()


def test_pack_refuses_a_short_buffer():
    hr = classified_rays(3000, 2)
    cnt = int((hr["classification"] == -1).sum())
    cap = 32 + 24 * (cnt - 1)
    r, c, got = gpu_pack(hr, cap)
    assert r != 0 and c == -1 and (got == 0xAB).all()


def gpu_apply(buf, fb0, w, h, kbuf0, frame0):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    msg = glib.DeviceArray.from_host(buf)
    fb = glib.DeviceArray.from_host(fb0)
    kb = glib.DeviceArray.from_host(kbuf0) if kbuf0 is not None else None
    frame, applied = C.c_int(frame0), C.c_int(-1)
    glib.check(L_.gxy_pixels_apply(msg.ptr, buf.nbytes, fb.ptr, w, h, kb.ptr if kb else None, C.byref(frame),
                                   C.byref(applied)))
    return fb.to_host(f32), (kb.to_host(np.int32) if kb else None), frame.value, applied.value


@pytest.mark.parametrize("frame0", [3, 5, 6])
def test_apply_equals_oracle(oracle, frame0):
    w, h = 16, 12
    hr = classified_rays(40000, 21, w + 2, h + 1)     # duplicates and pixels outside the framebuffer
    nbytes, buf = pack(oracle, hr, frame=5)
    buf = buf[:nbytes]
    rng = np.random.default_rng(4)
    fb0 = rng.uniform(0, 1, w * h * 4).astype(f32)
    kb0 = rng.integers(0, 7, w * h).astype(np.int32)
    fo, ko, fro = fb0.copy(), kb0.copy(), C.c_int(frame0)
    ao = oracle.L.oracle_pixels_apply(buf.ctypes.data, fo.ctypes.data, w, h, ko.ctypes.data, C.byref(fro))
    fg, kg, frg, ag = gpu_apply(buf, fb0, w, h, kb0, frame0)
    assert ag == ao and frg == fro.value and np.array_equal(kg, ko)
    if frame0 > 5:
        assert ao == 0 and np.array_equal(fg, fb0)
    else:
        assert ao > 9000        # a third of the rays are TERMINATED, some fall outside
        assert np.abs(fg - fo).max() <= 1e-4 * max(1.0, np.abs(fo).max())


def test_pack_then_apply_equals_accumulate_terminated():
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    w, h = 64, 64
    hr = classified_rays(50000, 8, w, h)
    cap = 32 + 24 * hr.n
    r, cnt, msg = gpu_pack(hr, cap)
    assert r == 0
    fg, _, _, applied = gpu_apply(msg[:32 + 24 * cnt], np.zeros(w * h * 4, f32), w, h, None, 0)
    d = glib.DeviceRays.from_host(hr)
    fb = glib.DeviceArray(w * h * 16).zero()
    glib.check(L_.gxy_accumulate_terminated(hr.n, C.byref(d.ispc), fb.ptr, w, h))
    assert applied == cnt
    assert np.abs(fg - fb.to_host(f32)).max() <= 1e-4
