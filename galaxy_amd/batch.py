"""Host side of the batched form (gxy_batch_* in include/gxy_hip.h).

Many ray lists of uneven sizes over many resident datasets are traced pass by
pass: one persistent trace launch per pass works through a device chunk table
(at most `chunk` rays of one list per entry), and a post-pass applies each
list's lighting, spawn, classification and accumulation exactly as the
per-list loop does (TraceRays::Trace, Renderer::processRays_task in the
reference: src/renderer/TraceRays.cpp:68-146, src/renderer/Renderer.cpp:504-656).
"""
import ctypes as C

import numpy as np

from . import abi, scenes
from . import lib as glib

f32 = np.float32


class Batch:
    """A gxy batch handle plus everything its lists and datasets point at."""

    def __init__(self, chunk=256):
        self.L = glib.lib()
        self.h = glib.checkp(self.L.gxy_batch_create(chunk), "batch_create")
        self.keep = []          # Built scenes, ray lists, output arrays
        self.lists = {}

    def destroy(self):
        if self.h:
            self.L.gxy_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # ------------------------------------------------------------ set-up
    def add_dataset(self, built, eps=0.001):
        """built: scenes.Built of a scene with one volume vis or one model."""
        self.keep.append(built)
        return glib.check(self.L.gxy_batch_add_dataset(self.h, C.byref(built.vis), eps), "batch_add_dataset")

    def refresh_datasets(self, built, eps=0.001):
        """Re-reads every dataset (built[k] was added as dataset k), e.g. after
        gxy_volume_set_layout."""
        for k, b in enumerate(built):
            glib.check(self.L.gxy_batch_update_dataset(self.h, k, C.byref(b.vis), eps), "batch_update_dataset")

    def set_lighting(self, lights, eps=0.001):
        self.keep.append(lights)
        glib.check(self.L.gxy_batch_set_lighting(self.h, C.byref(lights), eps), "batch_set_lighting")

    def add_render_list(self, ds, drays, width=None, height=1):
        """drays: lib.DeviceRays (traced in place).  Returns (list id, fb)."""
        width = width or max(1, drays.n)
        fb = glib.DeviceArray(4 * 4 * width * height)
        l = glib.check(self.L.gxy_batch_add_list(self.h, ds, abi.BATCH_RENDER, drays.n, C.byref(drays.ispc),
                                                 fb.ptr, width, height, None, 0, None), "batch_add_list")
        self.lists[l] = (drays, fb)
        return l, fb

    def add_sample_list(self, ds, drays, capacity):
        """Returns (list id, particle array (capacity*3 floats), count (device int))."""
        xyz = glib.DeviceArray(max(12, 12 * capacity))
        cnt = glib.DeviceArray(4)
        l = glib.check(self.L.gxy_batch_add_list(self.h, ds, abi.BATCH_SAMPLE, drays.n, C.byref(drays.ispc), None,
                                                 0, 0, xyz.ptr, capacity, cnt.ptr), "batch_add_list")
        self.lists[l] = (drays, xyz, cnt)
        return l, xyz, cnt

    def set_seed(self, l, seed, lo=(0, 0, 0), hi=(1, 1, 1)):
        a, b = glib.a3(lo), glib.a3(hi)
        glib.check(self.L.gxy_batch_set_list_seed(self.h, l, seed, a.ctypes.data, b.ctypes.data), "set_list_seed")

    def remove_list(self, l):
        glib.check(self.L.gxy_batch_remove_list(self.h, l), "batch_remove_list")
        self.lists.pop(l, None)

    def set_config(self, chunk, stack, wpe, blocks_per_cu, split=0):
        glib.check(self.L.gxy_batch_set_config(self.h, chunk, stack, wpe, blocks_per_cu), "batch_set_config")
        glib.check(self.L.gxy_batch_set_split(self.h, split), "batch_set_split")

    def config(self):
        v = [C.c_int() for _ in range(4)]
        glib.check(self.L.gxy_batch_get_config(self.h, *[C.byref(x) for x in v]), "batch_get_config")
        return tuple(x.value for x in v) + (self.L.gxy_batch_get_split(self.h),)

    def set_max_segments(self, n):
        """Isovalue crossings a sampler ray follows within one pass (1..4096)."""
        glib.check(self.L.gxy_batch_set_max_segments(self.h, int(n)), "batch_set_max_segments")

    def set_step_rays(self, rays):
        """Primary rays of added lists per step (0 = all); fewest left first."""
        glib.check(self.L.gxy_batch_set_step_rays(self.h, int(rays)), "batch_set_step_rays")

    # ------------------------------------------------------------ running
    def run(self, graph_passes=0):
        st = abi.BatchStats()
        glib.check(self.L.gxy_batch_run(self.h, graph_passes, C.byref(st)), "batch_run")
        return st

    def step(self):
        n = C.c_longlong()
        glib.check(self.L.gxy_batch_step(self.h, C.byref(n)), "batch_step")
        return n.value

    def pending(self, n):
        a = np.zeros(max(1, n), np.int32)
        glib.check(self.L.gxy_batch_pending(self.h, a.ctypes.data, n), "batch_pending")
        return a[:n]

    def stats(self):
        st = abi.BatchStats()
        glib.check(self.L.gxy_batch_get_stats(self.h, C.byref(st)), "batch_get_stats")
        return st

    def nbytes(self):
        return self.L.gxy_batch_bytes(self.h)


def random_rays(n, seed, width=0, lo=(0, 0, 0), hi=(1, 1, 1), typ=1):
    """Device rays from gxy_rays_random."""
    L = glib.lib()
    d = glib.DeviceRays(n)
    a, b = glib.a3(lo), glib.a3(hi)
    glib.check(L.gxy_rays_random(C.byref(d.ispc), n, seed, width, typ, a.ctypes.data, b.ctypes.data), "rays_random")
    return d


def build_datasets(hip, descs):
    """descs: list of SceneDesc, each with one volume vis or meshes only.
    Returns their Built objects (device scenes)."""
    return [scenes.build_scene(hip, sc) for sc in descs]


# candidate launch configurations: (chunk, stack, wpe, blocks per CU, split)
DEFAULT_CANDIDATES = [(256, 0, 0, 4, 0), (256, 0, 4, 4, 0), (128, 0, 4, 8, 0), (512, 0, 4, 4, 0),
                      (256, 0, 5, 5, 0), (256, 16, 4, 4, 0), (256, 24, 0, 3, 0), (256, 32, 0, 3, 0),
                      (256, 0, 0, 4, 1), (256, 0, 5, 4, 1), (256, 24, 5, 4, 1), (128, 0, 5, 8, 1),
                      (64, 0, 0, 4, 0), (64, 0, 5, 4, 1), (256, 0, 6, 8, 1), (128, 0, 6, 8, 1),
                      (128, 0, 5, 5, 0), (256, 0, 6, 6, 0)]


def tune(batch, candidates=None, max_depth=64, graph_passes=2, reps=1):
    """Measures gxy_batch_run under each launch configuration whose traversal
    stack fits the deepest BVH (stack 0 = private memory, always valid) and
    keeps the fastest.  Results do not depend on the configuration; only the
    time does.  Returns (best, [(config, ms), ...])."""
    import time
    L = batch.L
    table = []
    for cfg in candidates or DEFAULT_CANDIDATES:
        if cfg[1] and max_depth + 2 > cfg[1]:
            continue
        try:
            batch.set_config(*cfg)
            batch.run(graph_passes)     # capture / warm
        except glib.GxyError:
            continue                    # e.g. a traversal stack too short for some model
        glib.check(L.gxy_synchronize())
        t = time.perf_counter()
        for _ in range(reps):
            batch.run(graph_passes)
        glib.check(L.gxy_synchronize())
        table.append((cfg, (time.perf_counter() - t) / reps * 1e3))
    best = min(table, key=lambda x: x[1])[0]
    batch.set_config(*best)
    return best, table
