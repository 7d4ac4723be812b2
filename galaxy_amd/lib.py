"""ctypes binding of libgxy_hip.so (the HIP path).  There is no CPU fallback:
every call goes to the GPU library, and loading fails with an error when the
library or a GPU is missing."""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgxy_hip.so")

vp = C.c_void_p
_lib = None


class GxyError(RuntimeError):
    pass


def _declare(L):
    sig = {
        "gxy_init": (C.c_int, [C.c_int]),
        "gxy_last_error": (C.c_char_p, []),
        "gxy_device_count": (C.c_int, []),
        "gxy_synchronize": (C.c_int, []),
        "gxy_set_stream": (C.c_int, [vp]),
        "gxy_get_stream": (vp, []),
        "gxy_context_create": (vp, []),
        "gxy_context_bind": (C.c_int, [vp]),
        "gxy_context_destroy": (C.c_int, [vp]),
        "gxy_device_alloc": (vp, [C.c_size_t]),
        "gxy_device_free": (C.c_int, [vp]),
        "gxy_copy_to_device": (C.c_int, [vp, vp, C.c_size_t]),
        "gxy_copy_to_host": (C.c_int, [vp, vp, C.c_size_t]),
        "gxy_copy_device": (C.c_int, [vp, vp, C.c_size_t]),
        "gxy_memset": (C.c_int, [vp, C.c_int, C.c_size_t]),
        "gxy_device_bytes_in_use": (C.c_size_t, []),
        "gxy_device_bytes_total": (C.c_size_t, []),
        "gxy_set_opacity_threshold": (C.c_int, [C.c_float]),
        "gxy_get_opacity_threshold": (C.c_float, []),
        "gxy_set_trace_config": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "gxy_get_trace_config": (C.c_int, [vp, vp, vp]),
        "gxy_timing_enable": (C.c_int, [C.c_int]),
        "gxy_timing_reset": (C.c_int, []),
        "gxy_timing_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "gxy_raylist_create": (vp, [C.c_int, C.c_int, C.c_int, vp]),
        "gxy_raylist_destroy": (C.c_int, [vp]),
        "gxy_raylist_size": (C.c_int, [vp]),
        "gxy_raylist_aligned_size": (C.c_int, [vp]),
        "gxy_raylist_split_part": (vp, [vp, C.c_int, C.c_int, vp]),
        "gxy_raylist_truncate": (C.c_int, [vp, C.c_int, vp]),
        "gxy_raylist_bytes": (C.c_size_t, [vp]),
        "gxy_raylist_download": (C.c_int, [vp, vp]),
        "gxy_raylist_upload": (vp, [vp, vp]),
        "gxy_volume_create": (vp, [vp, vp, vp, vp, C.c_float]),
        "gxy_volume_destroy": (C.c_int, [vp]),
        "gxy_volume_device_data": (vp, [vp]),
        "gxy_volume_set_sampling_rate": (C.c_int, [vp, C.c_float]),
        "gxy_volume_fill_noise": (C.c_int, [vp, C.c_uint32, C.c_int, C.c_float]),
        "gxy_volume_download": (C.c_int, [vp, vp]),
        "gxy_tf_create_from_maps": (vp, [vp, C.c_int, vp, C.c_int, vp]),
        "gxy_tf_create": (vp, [vp, C.c_int, vp, C.c_int, C.c_float, C.c_float]),
        "gxy_tf_destroy": (C.c_int, [vp]),
        "gxy_tf_get_tables": (C.c_int, [vp, vp, vp, vp]),
        "gxy_model_create": (vp, []),
        "gxy_model_destroy": (C.c_int, [vp]),
        "gxy_model_add_triangles": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp, vp, vp]),
        "gxy_model_add_spheres": (C.c_int, [vp, vp, vp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, vp]),
        "gxy_set_occlusion_rays": (C.c_int, [C.c_int]),
        "gxy_get_occlusion_rays": (C.c_int, []),
        "gxy_model_add_pathlines": (C.c_int, [vp, vp, vp, C.c_int, vp, C.c_int, C.c_float, C.c_float, C.c_float,
                                              C.c_float, vp]),
        "gxy_model_commit": (C.c_int, [vp]),
        "gxy_model_bvh_stats": (C.c_int, [vp, vp, vp, vp, vp]),
        "gxy_set_bvh_prim_cost": (C.c_int, [C.c_int, C.c_float]),
        "gxy_get_bvh_prim_cost": (C.c_float, [C.c_int]),
        "gxy_model_set_hit_output": (C.c_int, [vp, vp, vp]),
        "gxy_spawn_offsets": (C.c_int, [C.c_int, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]),
        "gxy_spawn_offsets_noalloc": (C.c_int, [C.c_int, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]),
        "gxy_classify": (C.c_int, [C.c_int, vp, C.c_int]),
        "gxy_set_reverse_lighting": (C.c_int, [C.c_int]),
        "gxy_get_reverse_lighting": (C.c_int, []),
        "gxy_set_ao_directions": (C.c_int, [C.c_int, C.c_uint]),
        "gxy_get_ao_directions": (C.c_int, [vp]),
        "gxy_assign_destinations": (C.c_int, [C.c_int, vp, vp, vp, C.c_int]),
        "gxy_accumulate_terminated": (C.c_int, [C.c_int, vp, vp, C.c_int, C.c_int]),
        "gxy_pixels_pack": (C.c_int, [C.c_int, vp, C.c_int64, C.c_int64, C.c_int, C.c_int, vp, C.c_size_t, vp]),
        "gxy_pixels_apply": (C.c_int, [vp, C.c_size_t, vp, C.c_int, C.c_int, vp, vp, vp]),
        "gxy_partition_rays": (C.c_int, [C.c_int, vp, C.c_int, vp, vp, vp]),
        "gxy_camera_generate": (C.c_int, [vp, vp, C.c_int, vp, C.c_float, C.c_int, C.c_int, vp, vp, vp, C.c_int,
                                          C.c_int, vp]),
        "gxy_sampler_collect": (C.c_int, [C.c_int, vp, vp, C.c_int, vp]),
        "gxy_render_frame": (C.c_int, [vp, vp, vp, vp, vp, C.c_float, C.c_int, C.c_int, vp, C.c_float, C.c_int,
                                       vp, vp]),
        "gxy_process_raylist": (C.c_int, [vp, vp, C.c_int, vp, vp, C.c_float, C.c_int, vp, C.c_int, C.c_int,
                                          vp]),
        "gxy_sample_raylist": (C.c_int, [vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp]),
        "gxy_set_geometry_stack": (C.c_int, [C.c_int]),
        "gxy_set_bvh_width": (C.c_int, [C.c_int]),
        "gxy_get_bvh_width": (C.c_int, []),
        "gxy_set_geometry_waves": (C.c_int, [C.c_int]),
        "gxy_get_geometry_waves": (C.c_int, []),
        "gxy_volume_set_layout": (C.c_int, [vp, C.c_int]),
        "gxy_volume_get_layout": (C.c_int, [vp]),
        "gxy_get_geometry_stack": (C.c_int, []),
        "gxy_batch_create": (vp, [C.c_int]),
        "gxy_batch_destroy": (C.c_int, [vp]),
        "gxy_batch_add_dataset": (C.c_int, [vp, vp, C.c_float]),
        "gxy_batch_update_dataset": (C.c_int, [vp, C.c_int, vp, C.c_float]),
        "gxy_batch_set_lighting": (C.c_int, [vp, vp, C.c_float]),
        "gxy_batch_add_list": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp, C.c_int,
                                         vp]),
        "gxy_batch_set_list_seed": (C.c_int, [vp, C.c_int, C.c_uint, vp, vp]),
        "gxy_batch_remove_list": (C.c_int, [vp, C.c_int]),
        "gxy_batch_set_config": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "gxy_batch_get_config": (C.c_int, [vp, vp, vp, vp, vp]),
        "gxy_batch_set_split": (C.c_int, [vp, C.c_int]),
        "gxy_batch_get_split": (C.c_int, [vp]),
        "gxy_batch_set_step_rays": (C.c_int, [vp, C.c_longlong]),
        "gxy_batch_set_max_segments": (C.c_int, [vp, C.c_int]),
        "gxy_batch_get_step_rays": (C.c_longlong, [vp]),
        "gxy_batch_run": (C.c_int, [vp, C.c_int, vp]),
        "gxy_batch_step": (C.c_int, [vp, vp]),
        "gxy_batch_pending": (C.c_int, [vp, vp, C.c_int]),
        "gxy_batch_get_stats": (C.c_int, [vp, vp]),
        "gxy_batch_bytes": (C.c_longlong, [vp]),
        "gxy_rays_random": (C.c_int, [vp, C.c_int, C.c_uint, C.c_int, C.c_int, vp, vp]),
        "gxy_server_create": (vp, [C.c_longlong, C.c_int, C.c_int]),
        "gxy_server_destroy": (C.c_int, [vp]),
        "gxy_server_add_dataset": (C.c_int, [vp, vp, C.c_float]),
        "gxy_server_set_lighting": (C.c_int, [vp, vp, C.c_float]),
        "gxy_server_batch": (vp, [vp]),
        "gxy_server_run": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, vp]),
        "gxy_server_start": (C.c_int, [vp]),
        "gxy_server_submit": (C.c_int, [vp, vp]),
        "gxy_server_poll": (C.c_int, [vp, C.c_int, vp]),
        "gxy_server_cancel": (C.c_int, [vp, C.c_int]),
        "gxy_server_drain": (C.c_int, [vp, C.c_double]),
        "gxy_server_stop": (C.c_int, [vp]),
        "gxy_viewer_create": (vp, [C.c_longlong, C.c_int]),
        "gxy_viewer_destroy": (C.c_int, [vp]),
        "gxy_viewer_add_volume": (C.c_int, [vp, C.c_char_p, vp, vp, vp]),
        "gxy_viewer_handle": (C.c_int, [vp, C.c_char_p, vp, C.c_size_t]),
        "gxy_viewer_frame": (C.c_int, [vp, vp, vp, vp, vp]),
        "gxy_viewer_settings": (C.c_int, [vp, vp, vp, vp, vp]),
        "gxy_outbound_count": (C.c_int, [C.c_int]),
        "gxy_outbound_take": (vp, [C.c_int, vp, vp]),
        "gxy_outbound_clear": (C.c_int, []),
        "TraceRays_allocate": (vp, []),
        "TraceRays_initialize": (None, [vp]),
        "TraceRays_destroy": (None, [vp]),
        "TraceRays_TraceRays": (vp, [vp, vp, C.c_int, vp, C.c_float]),
        "TraceRays_generateAORays": (None, [vp, vp, C.c_int, vp, vp, vp, C.c_float]),
        "TraceRays_generateShadowRays": (None, [vp, vp, C.c_int, vp, vp, vp, C.c_float]),
        "TraceRays_ambientLighting": (None, [vp, vp, C.c_int, vp]),
        "TraceRays_diffuseLighting": (None, [vp, vp, C.c_int, vp]),
        "SamplerTraceRays_SamplerTraceRays": (vp, [vp, vp, C.c_int, vp]),
        "Lighting_allocate": (vp, []),
        "Lighting_initialize": (None, [vp]),
        "Lighting_destroy": (None, [vp]),
        "Lighting_SetLights": (None, [vp, C.c_int, vp, vp]),
        "Lighting_GetLights": (None, [vp, vp, vp, vp]),
        "Lighting_SetK": (None, [vp, C.c_float, C.c_float]),
        "Lighting_GetK": (None, [vp, vp, vp]),
        "Lighting_SetAO": (None, [vp, C.c_int, C.c_float]),
        "Lighting_GetAO": (None, [vp, vp, vp]),
        "Lighting_SetShadowFlag": (None, [vp, C.c_int]),
        "Lighting_GetShadowFlag": (None, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return sig


EXPORTS = None


def load():
    """Loads the in-tree library (no GPU needed just to load)."""
    global _lib, EXPORTS
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GxyError("libgxy_hip.so is not built: run __graft_entry__.build() or python galaxy_amd/build.py")
        L = C.CDLL(LIB_PATH)
        EXPORTS = _declare(L)
        _lib = L
    return _lib


def lib():
    """Loads the library and initialises the GPU; raises without a GPU."""
    L = load()
    if L.gxy_init(0) != 0:
        raise GxyError("gxy_init failed: " + L.gxy_last_error().decode())
    return L


def check(rc, what="gxy call"):
    if rc is None or (isinstance(rc, int) and rc < 0):
        raise GxyError("%s failed: %s" % (what, load().gxy_last_error().decode()))
    return rc


def checkp(p, what="gxy call"):
    if not p:
        raise GxyError("%s failed: %s" % (what, load().gxy_last_error().decode()))
    return p


def a3(v, dt=np.float32):
    return np.ascontiguousarray(np.asarray(v, dt).reshape(-1))


class DeviceRays:
    """A device RayList (reference block layout) with host round-trips."""

    def __init__(self, n, frame=0, typ=0):
        self.L = lib()
        self.ispc = abi.RayList_ispc()
        self.n = int(n)
        self.h = checkp(self.L.gxy_raylist_create(self.n, frame, typ, C.byref(self.ispc)), "raylist_create")

    def free(self):
        if self.h:
            self.L.gxy_raylist_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @staticmethod
    def from_host(hr):
        d = DeviceRays(hr.n)
        d.upload(hr)
        return d

    def upload(self, hr, fields=abi.FIELDS):
        for f in fields:
            a = np.ascontiguousarray(hr.a[f][:self.n])
            if self.n:
                check(self.L.gxy_copy_to_device(getattr(self.ispc, f), a.ctypes.data, a.nbytes), "upload")

    def download(self, fields=abi.FIELDS):
        hr = abi.HostRays(self.n)
        for f in fields:
            if self.n:
                check(self.L.gxy_copy_to_host(hr.a[f].ctypes.data, getattr(self.ispc, f), 4 * self.n), "download")
        return hr


class DeviceArray:
    def __init__(self, nbytes):
        self.L = lib()
        self.nbytes = int(nbytes)
        self.ptr = checkp(self.L.gxy_device_alloc(max(1, self.nbytes)), "device_alloc")

    def zero(self):
        check(self.L.gxy_memset(self.ptr, 0, self.nbytes), "memset")
        return self

    def to_host(self, dtype, count=None):
        count = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(count, dtype)
        if count:
            check(self.L.gxy_copy_to_host(out.ctypes.data, self.ptr, out.nbytes), "copy_to_host")
        return out

    @staticmethod
    def from_host(a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.nbytes)
        if a.nbytes:
            check(d.L.gxy_copy_to_device(d.ptr, a.ctypes.data, a.nbytes), "copy_to_device")
        return d

    def free(self):
        if self.ptr:
            self.L.gxy_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipBackend:
    """Scene-building interface shared with the oracle backend used in tests."""

    name = "hip"

    def __init__(self):
        self.L = lib()
        self.keep = []

    def volume_create(self, dims, origin, spacing, voxels, rate):
        d, o, s = a3(dims, np.int32), a3(origin), a3(spacing)
        return checkp(self.L.gxy_volume_create(d.ctypes.data, o.ctypes.data, s.ctypes.data,
                                               None if voxels is None else voxels.ctypes.data, float(rate)),
                      "volume_create")

    def volume_fill_noise(self, h, seed, octaves, freq):
        return check(self.L.gxy_volume_fill_noise(h, int(seed), int(octaves), float(freq)), "fill_noise")

    def volume_download(self, h, dims):
        out = np.empty(int(dims[0]) * int(dims[1]) * int(dims[2]), np.float32)
        check(self.L.gxy_volume_download(h, out.ctypes.data), "volume_download")
        return out.reshape(dims[2], dims[1], dims[0])

    def tf_create_from_maps(self, cmap, omap, rng):
        cmap, omap = a3(cmap), a3(omap)
        r = None if rng is None else a3(rng)
        return checkp(self.L.gxy_tf_create_from_maps(cmap.ctypes.data, len(cmap) // 4, omap.ctypes.data,
                                                     len(omap) // 2, None if r is None else r.ctypes.data),
                      "tf_create")

    def model_create(self):
        return checkp(self.L.gxy_model_create(), "model_create")

    def model_add_triangles(self, m, V, I, N, D, tf):
        V, I = a3(V), a3(I, np.uint32)
        N = None if N is None else a3(N)
        D = None if D is None else a3(D)
        return check(self.L.gxy_model_add_triangles(m, V.ctypes.data, len(V) // 3, I.ctypes.data, len(I) // 3,
                                                    None if N is None else N.ctypes.data,
                                                    None if D is None else D.ctypes.data, tf), "add_triangles")

    def model_add_spheres(self, m, Cn, D, r0, r1, v0, v1, tf):
        Cn = a3(Cn)
        D = None if D is None else a3(D)
        return check(self.L.gxy_model_add_spheres(m, Cn.ctypes.data, None if D is None else D.ctypes.data,
                                                  len(Cn) // 3, r0, r1, v0, v1, tf), "add_spheres")

    def model_add_pathlines(self, m, V, D, I, r0, r1, v0, v1, tf):
        V, D, I = a3(V), a3(D), a3(I, np.int32)
        return check(self.L.gxy_model_add_pathlines(m, V.ctypes.data, D.ctypes.data, len(V) // 3, I.ctypes.data,
                                                    len(I), r0, r1, v0, v1, tf), "add_pathlines")

    def model_commit(self, m, use_bvh=True):
        return check(self.L.gxy_model_commit(m), "model_commit")

    # ------------------------------------------------------------------ passes
    def trace(self, vis, drays, eps=0.001):
        self.L.TraceRays_TraceRays(None, C.byref(vis), drays.n, C.byref(drays.ispc), eps)
        err = self.L.gxy_last_error().decode()
        check(self.L.gxy_synchronize(), "trace")
        return err

    def render_frame(self, built, cam, neighbors=(-1,) * 6, eps=0.001, maxlist=1000000):
        w, h = cam.width, cam.height
        fb = DeviceArray(w * h * 16).zero()
        st = abi.FrameStats()
        nb = np.array(neighbors, np.int32)
        eye, cen, up = a3(cam.eye), a3(cam.center), a3(cam.up)
        check(self.L.gxy_render_frame(C.byref(built.vis), C.byref(built.lights), eye.ctypes.data, cen.ctypes.data,
                                      up.ctypes.data, cam.aov, w, h, nb.ctypes.data, eps, maxlist, fb.ptr,
                                      C.byref(st)), "render_frame")
        out = fb.to_host(np.float32)
        fb.free()
        return out, st

    def camera_rays(self, cam, lbox, gbox, order=0):
        n = cam.width * cam.height
        d = DeviceRays(n)
        lb, gb = abi.make_box(*lbox), abi.make_box(*gbox)
        eye, cen, up = a3(cam.eye), a3(cam.center), a3(cam.up)
        k = C.c_int(0)
        check(self.L.gxy_camera_generate(eye.ctypes.data, cen.ctypes.data, 0, up.ctypes.data, cam.aov, cam.width,
                                         cam.height, C.byref(lb), C.byref(gb), C.byref(d.ispc), n, order,
                                         C.byref(k)), "camera_generate")
        d.n = k.value
        return d
