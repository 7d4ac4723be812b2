"""Pins of the oracle's pixel messages (SendPixelsMsg, Renderer.h:246-336;
AddLocalPixels, Rendering.cpp:104-153): the byte layout against the
reference's struct declarations (two 64-bit Keys, three ints, tail padding;
pixels of two ints and four floats), the selection and order of the
TERMINATED rays, accumulation equal to direct accumulation of the rays, and
the frame-stamp rules on a hand-worked case."""
import ctypes as C

import numpy as np

from conftest import random_rays

f32 = np.float32
# struct hdr { Key rkey; Key rskey; int count; int frame; int source; } with
# Key = long (KeyedObject.h:120): 8-byte aligned, so 28 bytes pad to 32
HDR = np.dtype({"names": ["rkey", "rskey", "count", "frame", "source"],
                "formats": ["<i8", "<i8", "<i4", "<i4", "<i4"], "offsets": [0, 8, 16, 20, 24], "itemsize": 32})
# struct Pixel { int x, y; float r, g, b, o; } (Pixel.h:33-37)
PIX = np.dtype([("x", "<i4"), ("y", "<i4"), ("r", "<f4"), ("g", "<f4"), ("b", "<f4"), ("o", "<f4")])


def classified_rays(n, seed, w=64, h=64):
    hr = random_rays(n, seed)
    rng = np.random.default_rng(seed + 1)
    hr["classification"] = rng.choice(np.array([-1, -1, -2, -3, 0, 3], np.int32), n)
    for f in ("r", "g", "b", "o"):
        hr[f] = rng.uniform(0, 1, n).astype(f32)
    hr["x"] = rng.integers(0, w, n).astype(np.int32)
    hr["y"] = rng.integers(0, h, n).astype(np.int32)
    return hr


def pack(oracle, hr, rkey=0x1122334455667788, rskey=-5, frame=7, source=3, cap=None):
    cap = 32 + 24 * hr.n if cap is None else cap
    buf = np.zeros(max(cap, 32), np.uint8)
    nbytes = oracle.L.oracle_pixels_pack(hr.n, C.byref(hr.ispc), rkey, rskey, frame, source, buf.ctypes.data, cap)
    return nbytes, buf


def test_pack_layout_selection_and_order(oracle):
    hr = classified_rays(5000, 3)
    nbytes, buf = pack(oracle, hr)
    term = np.nonzero(hr["classification"] == -1)[0]
    assert nbytes == 32 + 24 * len(term) and len(term) > 1000
    h = buf[:32].view(HDR)[0]
    assert (h["rkey"], h["rskey"], h["count"], h["frame"], h["source"]) == (0x1122334455667788, -5, len(term), 7, 3)
    px = buf[32:nbytes].view(PIX)
    for f in PIX.names:
        assert np.array_equal(px[f], hr[f][term]), f
    # too small by one pixel: nothing packed
    assert pack(oracle, hr, cap=nbytes - 24)[0] == -1
    # no terminated rays: the header alone
    hr["classification"] = 0
    nb0, buf0 = pack(oracle, hr)
    assert nb0 == 32 and buf0[:32].view(HDR)[0]["count"] == 0


def test_apply_equals_direct_accumulation(oracle):
    hr = classified_rays(20000, 5, 40, 30)
    nbytes, buf = pack(oracle, hr)
    fb = np.zeros(40 * 30 * 4, f32)
    frame = C.c_int(0)
    k = oracle.L.oracle_pixels_apply(buf.ctypes.data, fb.ctypes.data, 40, 30, None, C.byref(frame))
    ref = np.zeros_like(fb)
    kr = oracle.L.oracle_accumulate(hr.n, C.byref(hr.ispc), ref.ctypes.data, 40, 30)
    assert k == kr == (hr["classification"] == -1).sum() and frame.value == 7
    assert np.array_equal(fb, ref)


def test_apply_frame_stamp_rules(oracle):
    w, h = 4, 3
    msg = np.zeros(1, HDR)
    pix = np.array([(1, 1, 1, 2, 3, 4), (1, 1, 10, 20, 30, 40), (2, 0, 0.5, 0.5, 0.5, 0.5),
                    (-1, 0, 9, 9, 9, 9), (4, 2, 9, 9, 9, 9)], PIX)
    msg["count"], msg["frame"] = len(pix), 5
    buf = np.concatenate([msg.view(np.uint8), pix.view(np.uint8)])
    fb = np.full(w * h * 4, 7.0, f32)
    kbuf = np.zeros(w * h, np.int32)
    kbuf[1 * w + 1], kbuf[0 * w + 2] = 3, 5
    frame = C.c_int(4)
    k = oracle.L.oracle_pixels_apply(buf.ctypes.data, fb.ctypes.data, w, h, kbuf.ctypes.data, C.byref(frame))
    exp = np.full((h, w, 4), 7.0, f32)
    exp[1, 1] = [11, 22, 33, 44]            # stamp 3 < 5: cleared once, then both pixels added
    exp[0, 2] = 7.5                         # stamp 5: current, added to what is there
    kexp = np.zeros((h, w), np.int32)
    kexp[1, 1], kexp[0, 2] = 5, 5
    assert k == 3 and frame.value == 5      # the two pixels outside the 4x3 framebuffer are skipped
    assert np.array_equal(fb.reshape(h, w, 4), exp) and np.array_equal(kbuf.reshape(h, w), kexp)
    # a message of an older frame is dropped whole
    msg["frame"] = 4
    old = np.concatenate([msg.view(np.uint8), pix.view(np.uint8)])
    before = fb.copy()
    assert oracle.L.oracle_pixels_apply(old.ctypes.data, fb.ctypes.data, w, h, kbuf.ctypes.data, C.byref(frame)) == 0
    assert frame.value == 5 and np.array_equal(fb, before)
    # one of the current frame is applied, over current stamps (no clearing)
    assert oracle.L.oracle_pixels_apply(buf.ctypes.data, fb.ctypes.data, w, h, kbuf.ctypes.data, C.byref(frame)) == 3
    exp[1, 1] = [22, 44, 66, 88]
    exp[0, 2] = 8.0
    assert np.array_equal(fb.reshape(h, w, 4), exp) and np.array_equal(kbuf.reshape(h, w), kexp)
