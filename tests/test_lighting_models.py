"""Galaxy's two lighting models (GXY_REVERSE_LIGHTING, CMakeLists.txt:66).

Reverse lighting (the default) adds ambient and diffuse to the parent and
casts AO and shadow rays that carry the negated contribution, which an
occluded ray puts back (Renderer.cpp:365-411, TraceRays.cpp:124-131).  The
direct model casts rays carrying the positive contribution, which an
unoccluded ray delivers (TraceRays.cpp:133-141, TraceRays.ispc:639-643,
778-782), and lights the parent itself only for the rays it does not cast.

Pins of the direct model: the spawned colours on a hand-built parent against
the formulas in float64, and a property the two models share by design: where
the primary ray reaches its surface with alpha 0 (no volume in front), every
pixel sums to ambient x (escaping AO fraction) + diffuse x (unshadowed), so a
geometry scene with a directional light renders the same frame under either
model.  The GPU follows one library switch (gxy_set_reverse_lighting), and
its classify/assign entry points refuse a flag that disagrees with it."""
import ctypes as C

import numpy as np
import pytest

from galaxy_amd import abi, scenes, workloads

f32 = np.float32


def _parent():
    hr = abi.HostRays(1)
    hr["type"], hr["term"] = 1, abi.RAY_SURFACE | abi.RAY_OPAQUE
    hr["ox"], hr["oy"], hr["oz"] = 0.0, 0.0, 3.0
    hr["dx"], hr["dy"], hr["dz"] = 0.0, 0.0, -1.0
    hr["t"] = 2.0
    hr["nx"], hr["ny"], hr["nz"] = 0.0, 0.0, 1.0
    hr["sr"], hr["sg"], hr["sb"] = 0.5, 0.25, 1.0
    hr["r"], hr["g"], hr["b"], hr["o"] = 0.1, 0.1, 0.1, 0.2
    return hr


def test_direct_model_on_one_parent(oracle):
    """TraceRays.cpp:133-141 with AO and shadow rays: the parent is not lit,
    the AO children carry +Ka/nAO*(1-o)*sc and the shadow children
    +(Kd/nL)*(1-o)*max(0, n.l)*sc with the parent's own alpha."""
    hr = _parent()
    Ld = scenes.LightingDesc(np.array([[0.0, 0.0, 5.0], [1.0, 0.0, 0.0]], f32), np.array([2, 0], np.int32),
                             Ka=0.4, Kd=0.6, n_ao=4, ao_radius=0.7, shadows=True)
    keep = []
    L = scenes.make_lighting(Ld, keep)
    out = abi.HostRays(6)
    ao = np.zeros(1, np.int32)
    sh = np.array([4], np.int32)
    oracle.L.oracle_set_reverse_lighting(0)
    try:
        oracle.L.oracle_generateAORays(C.byref(L), 1, C.byref(hr.ispc), ao.ctypes.data, C.byref(out.ispc), 0.001)
        oracle.L.oracle_generateShadowRays(C.byref(L), 1, C.byref(hr.ispc), sh.ctypes.data, C.byref(out.ispc),
                                           0.001)
    finally:
        oracle.L.oracle_set_reverse_lighting(1)
    o0, sc = 0.2, np.array([0.5, 0.25, 1.0])
    assert [hr["r"][0], hr["g"][0], hr["b"][0], hr["o"][0]] == pytest.approx([0.1, 0.1, 0.1, 0.2], abs=1e-7)
    for c, f in enumerate("rgb"):
        assert np.allclose(out[f][:4], 0.4 / 4 * (1 - o0) * sc[c], atol=1e-7)
    assert np.allclose(out["r"][4], (0.6 / 2) * (1 - o0) * 1.0 * sc[0], atol=1e-6)
    assert out["r"][5] == 0.0
    assert np.all(out["type"][:4] == abi.RAY_AO) and np.all(out["type"][4:] == abi.RAY_SHADOW)


def _opaque_scene(w=96, h=72):
    sc, _ = workloads.nine_icospheres(w, h, subdiv=2, n_ao=8, ao_radius=0.8)
    d = np.array([-0.3, -0.5, -1.0], f32)
    sc.lighting = scenes.LightingDesc((d / np.linalg.norm(d))[None, :].astype(f32), np.array([0], np.int32),
                                      Ka=0.4, Kd=0.6, n_ao=8, ao_radius=0.8, shadows=True)
    return sc


def _frame(be, sc, reverse, set_reverse):
    set_reverse(reverse)
    try:
        cam = sc.cameras[0]
        return be.render_frame(scenes.build_scene(be, sc, scenes.resolve_lights(sc.lighting, cam)), cam)
    finally:
        set_reverse(1)


def test_models_render_the_same_opaque_geometry(oracle):
    sc = _opaque_scene()
    fb_r, st_r = _frame(oracle, sc, 1, oracle.L.oracle_set_reverse_lighting)
    fb_d, st_d = _frame(oracle, sc, 0, oracle.L.oracle_set_reverse_lighting)
    assert st_r.primary_rays == st_d.primary_rays and st_r.secondary_rays == st_d.secondary_rays
    assert st_r.secondary_rays > 10000
    # the models route the children oppositely: occluded ones terminate under reverse lighting only
    assert st_r.terminated != st_d.terminated
    assert np.abs(fb_r - fb_d).max() <= 1e-5
    # and the frame is not trivially dark or unshadowed
    assert fb_r.reshape(-1, 4)[:, 0].max() > 0.3 and (st_r.dropped > 0 and st_d.dropped > 0)


@pytest.mark.gpu
def test_gpu_direct_model_matches_oracle(hip, oracle):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    sc = _opaque_scene()
    fb_o, st_o = _frame(oracle, sc, 0, oracle.L.oracle_set_reverse_lighting)
    fb_g, st_g = _frame(hip, sc, 0, L_.gxy_set_reverse_lighting)
    fb_r, _ = _frame(hip, sc, 1, L_.gxy_set_reverse_lighting)
    assert st_g.as_dict() == st_o.as_dict()
    assert np.abs(fb_g - fb_o).max() <= 1e-4
    assert np.abs(fb_g - fb_r).max() <= 1e-4
    assert L_.gxy_get_reverse_lighting() == 1


@pytest.mark.gpu
def test_gpu_classify_refuses_the_other_model():
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    d = glib.DeviceRays.from_host(_parent())
    assert L_.gxy_classify(1, C.byref(d.ispc), 0) != 0
    assert b"lighting model" in L_.gxy_last_error()
    glib.check(L_.gxy_set_reverse_lighting(0))
    try:
        glib.check(L_.gxy_classify(1, C.byref(d.ispc), 0))
        assert L_.gxy_classify(1, C.byref(d.ispc), 1) != 0
    finally:
        glib.check(L_.gxy_set_reverse_lighting(1))
    glib.check(L_.gxy_classify(1, C.byref(d.ispc), 1))
