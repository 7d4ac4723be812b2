"""Writes tests/golden/walk_counts.json: the algorithmic BVH work per ray of
the sphere-mesh frame (512x512, one shadow ray per hit) and the
nine-icosphere frame (3840x2160, 16 AO + 1 shadow rays per hit), counted by
the CPU oracle's binned-SAH tree (leaves of at most 4 primitives) walked
nearest first: child-box tests and primitive tests per primary, shadow and AO
ray (occlusion rays end at their first hit, as the frame loops do; the
nine-icosphere frame's AO directions from its seeded PCG32 streams).  The
bench prices them at 20 flops per box test and 51 per triangle test for the
geometry roofline.  Run from the repository root:
    python tests/golden/make_walk_counts.py"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

OUT = os.path.join(HERE, "walk_counts.json")


def walk_counts(sc, stride, ao=(0, 0)):
    """Per-ray walk work of the camera rays of every stride-th image row and of
    the AO and shadow rays they spawn (AO directions: ao = (mode, seed) of
    gxy_set_ao_directions)."""
    from galaxy_amd import abi, scenes
    from oracle_backend import OracleBackend
    ob = OracleBackend()
    L = ob.L
    cam = sc.cameras[0]
    lights = scenes.resolve_lights(sc.lighting, cam)
    B = scenes.build_scene(ob, sc, lights)
    rays = ob.camera_rays(cam, sc.local_box, sc.global_box)
    sub = rays.subset(np.nonzero(rays["y"] % stride == 0)[0])
    cnt = (C.c_longlong * 3)()

    def traced(hr):
        L.oracle_walk_counters(1, None)
        ob.trace(B.vis, hr)
        L.oracle_walk_counters(0, cnt)
        return dict(rays=int(cnt[2]), box_tests=int(cnt[0]), prim_tests=int(cnt[1]))

    L.oracle_set_occlusion_any(1)
    L.oracle_set_ao_directions(*ao)
    try:
        o = sub.copy()
        out = {"stride": stride, "primary": traced(o)}
        n = o.n
        ao, sh = np.zeros(n, np.int32), np.zeros(n, np.int32)
        na, ns = C.c_int(), C.c_int()
        nout = L.oracle_spawn_offsets(n, C.byref(o.ispc), lights.n_ao, len(lights.types), int(lights.shadows),
                                      ao.ctypes.data, sh.ctypes.data, C.byref(na), C.byref(ns))
        kids = abi.HostRays(max(nout, 1))
        kids.n = nout
        if na.value:
            L.oracle_generateAORays(C.byref(B.lights), n, C.byref(o.ispc), ao.ctypes.data, C.byref(kids.ispc),
                                    sc.epsilon)
        if ns.value:
            L.oracle_generateShadowRays(C.byref(B.lights), n, C.byref(o.ispc), sh.ctypes.data, C.byref(kids.ispc),
                                        sc.epsilon)
        if na.value:
            out["ao"] = traced(kids.subset(np.arange(0, na.value)))
        if ns.value:
            out["shadow"] = traced(kids.subset(np.arange(na.value, nout)))
    finally:
        L.oracle_set_occlusion_any(0)
        L.oracle_set_ao_directions(0, 0)
    return out


def scenes_to_count():
    from galaxy_amd import workloads
    return {"sphere_mesh_512": (lambda: workloads.sphere_mesh(512, 512)[0], 1, (0, 0)),
            "nine_icospheres_4k_ao16": (lambda: workloads.nine_icospheres()[0], 16,
                                        (1, workloads.NINE_ICOSPHERES_AO_SEED))}


def main():
    res = {}
    for name, (make, stride, ao) in scenes_to_count().items():
        res[name] = walk_counts(make(), stride, ao)
        print(name, json.dumps(res[name]), flush=True)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
