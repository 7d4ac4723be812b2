"""AO directions from seeded PCG32 streams (gxy_set_ao_directions mode 1, the
nine-icosphere workload's "1 shadow + 16 AO with PCG seeds"), beside the
reference's UV table (mode 0, TraceRays.ispc:625-733 with UV.ih:21-58).

Pins: the oracle's PCG32 against the outputs the generator's published demo
prints for pcg32_srandom(42, 54); the streams' (U, V) against the same
generator written here in Python (64-bit integer arithmetic); the AO
directions they give against the moments of the cosine-weighted hemisphere
(E[z] = 2/3 about the normal, E[x] = E[y] = 0, |d| within the reference's
un-renormalised epsilon); and the GPU's children and a lit frame against the
oracle in both modes."""
import ctypes as C

import numpy as np
import pytest

from galaxy_amd import abi, scenes

f32 = np.float32
M64 = (1 << 64) - 1


def pcg32_stream(initstate, initseq, n):
    """PCG32 (O'Neill 2014): pcg32_srandom_r then n x pcg32_random_r."""
    state, inc = 0, ((initseq << 1) | 1) & M64

    def step():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + inc) & M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF
    step()
    state = (state + initstate) & M64
    step()
    return [step() for _ in range(n)]


def test_pcg32_known_answers(oracle):
    out = np.zeros(6, np.uint32)
    oracle.L.oracle_pcg32(42, 54, 6, out.ctypes.data)
    demo = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert list(out) == demo
    assert pcg32_stream(42, 54, 6) == demo
    out2 = np.zeros(5, np.uint32)
    oracle.L.oracle_pcg32((7 << 32) | 3, (11 << 32) | 5, 5, out2.ctypes.data)
    assert list(out2) == pcg32_stream((7 << 32) | 3, (11 << 32) | 5, 5)


def _parents(n, seed=3):
    """n primary surface hits facing +z at distinct pixels."""
    rng = np.random.default_rng(seed)
    hr = abi.HostRays(n)
    hr["type"], hr["term"] = 1, abi.RAY_SURFACE | abi.RAY_OPAQUE
    hr["ox"], hr["oy"], hr["oz"] = 0.0, 0.0, 3.0
    hr["dx"], hr["dy"], hr["dz"] = 0.0, 0.0, -1.0
    hr["t"] = 2.0
    hr["nx"], hr["ny"], hr["nz"] = 0.0, 0.0, 1.0
    hr["sr"], hr["sg"], hr["sb"] = 0.5, 0.25, 1.0
    hr["r"], hr["g"], hr["b"], hr["o"] = 0.0, 0.0, 0.0, 0.0
    hr["x"] = rng.integers(0, 4096, n).astype(np.int32)
    hr["y"] = rng.integers(0, 4096, n).astype(np.int32)
    return hr


def _ao_children(oracle, hr, nao, mode, seed):
    Ld = scenes.LightingDesc(np.array([[0.0, 0.0, 5.0]], f32), np.array([2], np.int32), Ka=0.4, Kd=0.6,
                             n_ao=nao, ao_radius=0.7, shadows=False)
    keep = []
    L = scenes.make_lighting(Ld, keep)
    out = abi.HostRays(hr.n * nao)
    off = (np.arange(hr.n) * nao).astype(np.int32)
    oracle.L.oracle_set_ao_directions(mode, seed)
    try:
        oracle.L.oracle_generateAORays(C.byref(L), hr.n, C.byref(hr.ispc), off.ctypes.data, C.byref(out.ispc),
                                       0.001)
    finally:
        oracle.L.oracle_set_ao_directions(0, 0)
    return out


def test_pcg_streams_give_the_stated_uv_pairs(oracle):
    hr = _parents(40)
    out = _ao_children(oracle, hr, 4, 1, 1234)
    for i in range(hr.n):
        for j in range(4):
            a, b = pcg32_stream((1234 << 32) | j, (int(hr["y"][i]) << 32) | int(hr["x"][i]), 2)
            u, v = (a >> 8) / 2.0 ** 24, (b >> 8) / 2.0 ** 24
            k = 4 * i + j
            # normal +z: the reference's basis is b0 = (-1, 0, 0), b1 = (0, -1, 0)
            # (x-hat crossed with n, then that crossed with n), so
            # d = (-cos(2 pi U) sqrt(1 - V), -sin(2 pi U) sqrt(1 - V), sqrt(V) + eps)
            w = np.sqrt(1 - v)
            assert out["dx"][k] == pytest.approx(-np.cos(2 * np.pi * u) * w, abs=3e-6)
            assert out["dy"][k] == pytest.approx(-np.sin(2 * np.pi * u) * w, abs=3e-6)
            assert out["dz"][k] == pytest.approx(np.sqrt(v) + 0.001, abs=2e-6)


def test_pcg_directions_are_cosine_weighted(oracle):
    hr = _parents(5000, seed=9)
    out = _ao_children(oracle, hr, 16, 1, 77)
    d = np.c_[out["dx"], out["dy"], out["dz"]].astype(np.float64)
    n = len(d)
    # cosine-weighted hemisphere about +z: E[z] = 2/3, E[z^2] = 1/2, E[x] = E[y] = 0
    assert abs(d[:, 2].mean() - 0.001 - 2 / 3) < 4 / np.sqrt(n)
    assert abs(((d[:, 2] - 0.001) ** 2).mean() - 0.5) < 4 / np.sqrt(n)
    assert abs(d[:, 0].mean()) < 4 / np.sqrt(n) and abs(d[:, 1].mean()) < 4 / np.sqrt(n)
    assert (d[:, 2] > 0).all()
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1.5e-3          # not renormalised: +eps in z
    # distinct from the table: the reference hash takes 256 directions at most
    assert len(np.unique(np.round(d, 6), axis=0)) > 0.99 * n
    tab = _ao_children(oracle, hr, 16, 0, 0)
    assert len(np.unique(np.round(np.c_[tab["dx"], tab["dy"], tab["dz"]], 6), axis=0)) <= 256
    # another seed, other directions
    other = _ao_children(oracle, hr, 16, 1, 78)
    assert np.mean(other["dz"] != out["dz"]) > 0.99


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_ao_children_and_frame_match_oracle(hip, oracle, mode):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    hr = _parents(3000, seed=5)
    ref = _ao_children(oracle, hr, 16, mode, 4242)
    Ld = scenes.LightingDesc(np.array([[0.0, 0.0, 5.0]], f32), np.array([2], np.int32), Ka=0.4, Kd=0.6,
                             n_ao=16, ao_radius=0.7, shadows=False)
    keep = []
    L = scenes.make_lighting(Ld, keep)
    glib.check(L_.gxy_set_ao_directions(mode, 4242))
    try:
        d = glib.DeviceRays.from_host(hr)
        out = glib.DeviceRays(hr.n * 16)
        off = glib.DeviceArray.from_host((np.arange(hr.n) * 16).astype(np.int32))
        L_.TraceRays_generateAORays(None, C.byref(L), hr.n, C.byref(d.ispc), off.ptr, C.byref(out.ispc), 0.001)
        glib.check(L_.gxy_synchronize())
        got = out.download()
        for f in ("ox", "oy", "oz", "r", "g", "b", "tMax"):
            assert np.array_equal(got[f], ref[f]), f
        for f in ("dx", "dy", "dz"):
            assert np.abs(got[f] - ref[f]).max() <= 2e-6, f
        # a lit icosphere frame with 16 AO rays per hit in this mode
        box = ((-1.5,) * 3, (1.5,) * 3)
        Lf = scenes.LightingDesc(np.array([[3.0, 4.0, 5.0]], f32), np.array([2], np.int32), Ka=0.3, Kd=0.7,
                                 n_ao=16, ao_radius=0.4, shadows=True)
        cam = scenes.CameraDesc((0.3, 0.4, 3.2), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 96, 72)
        sc = scenes.SceneDesc(meshes=[scenes.icosphere(4, 0.6, (-0.4, 0.0, 0.0)),
                                      scenes.icosphere(4, 0.5, (0.6, 0.2, -0.2))],
                              local_box=box, global_box=box, lighting=Lf, cameras=[cam])
        oracle.L.oracle_set_ao_directions(mode, 4242)
        try:
            fb_o, st_o = oracle.render_frame(scenes.build_scene(oracle, sc), cam)
        finally:
            oracle.L.oracle_set_ao_directions(0, 0)
        fb_g, st_g = hip.render_frame(scenes.build_scene(hip, sc), cam)
        assert st_g.primary_rays == st_o.primary_rays and st_g.secondary_rays > 10000
        assert abs(st_g.secondary_rays - st_o.secondary_rays) <= max(3, st_o.secondary_rays // 5000)
        assert np.mean(np.abs(fb_g - fb_o) <= 1e-4) >= 0.999
    finally:
        L_.gxy_set_ao_directions(0, 0)
