"""ctypes binding of the CPU oracle (oracle/_build/libgxy_oracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg only, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from galaxy_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "libgxy_oracle.so")

vp = C.c_void_p
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int)


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.oracle_volume_create.restype = vp
        L.oracle_volume_create.argtypes = [vp, vp, vp, vp, C.c_float]
        L.oracle_volume_sample.restype = C.c_float
        L.oracle_volume_sample.argtypes = [vp, C.c_float, C.c_float, C.c_float]
        L.oracle_volume_gradient.argtypes = [vp, C.c_float, C.c_float, C.c_float, vp]
        L.oracle_tf_create_from_maps.restype = vp
        L.oracle_tf_create_from_maps.argtypes = [vp, C.c_int, vp, C.c_int, vp]
        L.oracle_tf_create.restype = vp
        L.oracle_tf_create.argtypes = [vp, C.c_int, vp, C.c_int, C.c_float, C.c_float]
        L.oracle_tf_get_tables.argtypes = [vp, vp, vp, vp]
        L.oracle_tf_lookup.argtypes = [vp, C.c_float, vp, vp]
        L.oracle_model_create.restype = vp
        L.oracle_model_add_triangles.argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, vp, vp]
        L.oracle_model_add_spheres.argtypes = [vp, vp, vp, C.c_int, C.c_float, C.c_float, C.c_float,
                                               C.c_float, vp]
        L.oracle_model_commit.argtypes = [vp, C.c_int]
        L.oracle_model_add_pathlines.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_float, C.c_float,
                                                 C.c_float, C.c_float, vp]
        L.oracle_pathline_controls.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int]
        L.oracle_model_intersect.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_triangles_bruteforce64.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.oracle_TraceRays.restype = C.c_longlong
        L.oracle_TraceRays.argtypes = [vp, C.c_int, vp, C.c_float, C.c_int, vp, vp, vp]
        L.oracle_spawn_offsets.restype = C.c_int
        L.oracle_spawn_offsets.argtypes = [C.c_int, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
        for nm in ("oracle_ambientLighting", "oracle_diffuseLighting"):
            getattr(L, nm).argtypes = [vp, C.c_int, vp]
        L.oracle_generateAORays.argtypes = [vp, C.c_int, vp, vp, vp, C.c_float]
        L.oracle_generateShadowRays.argtypes = [vp, C.c_int, vp, vp, vp, C.c_float]
        L.oracle_classify.argtypes = [C.c_int, vp, C.c_int]
        L.oracle_exit_face.restype = C.c_int
        L.oracle_exit_face.argtypes = [vp] + [C.c_float] * 6
        L.oracle_assign_destinations.argtypes = [C.c_int, vp, vp, vp, C.c_int]
        L.oracle_accumulate.restype = C.c_int
        L.oracle_accumulate.argtypes = [C.c_int, vp, vp, C.c_int, C.c_int]
        L.oracle_pixels_pack.restype = C.c_longlong
        L.oracle_pixels_pack.argtypes = [C.c_int, vp, C.c_longlong, C.c_longlong, C.c_int, C.c_int, vp,
                                         C.c_longlong]
        L.oracle_pixels_apply.restype = C.c_int
        L.oracle_pixels_apply.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp]
        L.oracle_partition_counts.restype = C.c_int
        L.oracle_partition_counts.argtypes = [C.c_int, vp, C.c_int, vp]
        L.oracle_SamplerTraceRays.argtypes = [vp, C.c_int, vp]
        L.oracle_camera_generate.restype = C.c_int
        L.oracle_camera_generate.argtypes = [vp, vp, C.c_int, vp, C.c_float, C.c_int, C.c_int, vp, vp, vp,
                                             C.c_int]
        L.oracle_resolve_lights.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, vp]
        L.oracle_process_raylist.restype = C.c_int
        L.oracle_process_raylist.argtypes = [vp, vp, C.c_int, vp, vp, C.c_float, C.c_int, vp, C.c_int,
                                             C.c_int, vp]
        L.oracle_render_frame.restype = C.c_int
        L.oracle_render_frame.argtypes = [vp, vp, vp, vp, vp, C.c_float, C.c_int, C.c_int, vp, C.c_float,
                                          C.c_int, vp, vp]
        L.oracle_sample_raylist.restype = C.c_longlong
        L.oracle_sample_raylist.argtypes = [vp, C.c_int, vp, vp, C.c_longlong, vp]
        L.oracle_uv_tables.argtypes = [vp, vp]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_set_opacity_threshold.argtypes = [C.c_float]
        L.oracle_plant_fault.argtypes = [C.c_int]
        L.oracle_set_reverse_lighting.argtypes = [C.c_int]
        L.oracle_set_ao_directions.argtypes = [C.c_int, C.c_uint32]
        L.oracle_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_void_p]
        L.oracle_walk_counters.argtypes = [C.c_int, vp]
        L.oracle_set_occlusion_any.argtypes = [C.c_int]
        L.oracle_get_threads.restype = C.c_int
        _lib = L
    return _lib


def a3(v, dt=np.float32):
    return np.ascontiguousarray(np.asarray(v, dt).reshape(-1))


class OracleBackend:
    """Scene-building interface shared with galaxy_amd.lib.HipBackend."""

    name = "oracle"

    def __init__(self):
        self.L = lib()
        self.keep = []

    def volume_create(self, dims, origin, spacing, voxels, rate):
        d, o, s = a3(dims, np.int32), a3(origin), a3(spacing)
        self.keep += [d, o, s, voxels]
        return self.L.oracle_volume_create(d.ctypes.data, o.ctypes.data, s.ctypes.data, voxels.ctypes.data,
                                           float(rate))

    def tf_create_from_maps(self, cmap, omap, rng):
        cmap, omap = a3(cmap), a3(omap)
        r = None if rng is None else a3(rng)
        return self.L.oracle_tf_create_from_maps(cmap.ctypes.data, len(cmap) // 4, omap.ctypes.data,
                                                 len(omap) // 2, None if r is None else r.ctypes.data)

    def model_create(self):
        return self.L.oracle_model_create()

    def model_add_triangles(self, m, V, I, N, D, tf):
        V, I = a3(V), a3(I, np.uint32)
        N = None if N is None else a3(N)
        D = None if D is None else a3(D)
        self.keep += [V, I, N, D]
        return self.L.oracle_model_add_triangles(m, V.ctypes.data, len(V) // 3, I.ctypes.data, len(I) // 3,
                                                 None if N is None else N.ctypes.data,
                                                 None if D is None else D.ctypes.data, tf)

    def model_add_spheres(self, m, Cn, D, r0, r1, v0, v1, tf):
        Cn = a3(Cn)
        D = None if D is None else a3(D)
        self.keep += [Cn, D]
        return self.L.oracle_model_add_spheres(m, Cn.ctypes.data, None if D is None else D.ctypes.data,
                                               len(Cn) // 3, r0, r1, v0, v1, tf)

    def model_add_pathlines(self, m, V, D, I, r0, r1, v0, v1, tf):
        V, D, I = a3(V), a3(D), a3(I, np.int32)
        self.keep += [V, D, I]
        return self.L.oracle_model_add_pathlines(m, V.ctypes.data, D.ctypes.data, len(V) // 3, I.ctypes.data,
                                                 len(I), r0, r1, v0, v1, tf)

    def model_commit(self, m, use_bvh=True):
        return self.L.oracle_model_commit(m, 1 if use_bvh else 0)

    # ---------------------------------------------------------------- passes
    def trace(self, vis, rays, eps=0.001, fp64=False, geom_ids=False):
        n = rays.n
        rgba = np.zeros((n, 4), np.float64) if fp64 else None
        g = np.full(n, -1, np.int32) if geom_ids else None
        p = np.full(n, -1, np.int32) if geom_ids else None
        ns = self.L.oracle_TraceRays(C.byref(vis), n, C.byref(rays.ispc), eps, 1 if fp64 else 0,
                                     None if rgba is None else rgba.ctypes.data,
                                     None if g is None else g.ctypes.data, None if p is None else p.ctypes.data)
        return dict(samples=ns, rgba64=rgba, geom=g, prim=p)

    def render_frame(self, built, cam, neighbors=(-1,) * 6, eps=0.001, maxlist=1000000):
        w, h = cam.width, cam.height
        fb = np.zeros(w * h * 4, np.float32)
        st = abi.FrameStats()
        nb = np.array(neighbors, np.int32)
        eye, cen, up = a3(cam.eye), a3(cam.center), a3(cam.up)
        r = self.L.oracle_render_frame(C.byref(built.vis), C.byref(built.lights), eye.ctypes.data, cen.ctypes.data,
                                       up.ctypes.data, cam.aov, w, h, nb.ctypes.data, eps, maxlist, fb.ctypes.data,
                                       C.byref(st))
        assert r == 0
        return fb, st

    def camera_rays(self, cam, lbox, gbox):
        from galaxy_amd.abi import HostRays, make_box
        n = cam.width * cam.height
        R = HostRays(n)
        lb, gb = make_box(*lbox), make_box(*gbox)
        eye, cen, up = a3(cam.eye), a3(cam.center), a3(cam.up)
        k = self.L.oracle_camera_generate(eye.ctypes.data, cen.ctypes.data, 0, up.ctypes.data, cam.aov, cam.width,
                                          cam.height, C.byref(lb), C.byref(gb), C.byref(R.ispc), n)
        R.n = k
        return R
