"""ctypes mirrors of the structs declared in include/gxy_hip.h.

The layouts follow the reference's ISPC-exported structs:
RayList_ispc (src/renderer/Rays.ih:21-48), Lighting_ispc
(src/renderer/Lighting.ih:23-32), Vis_ispc / MappedVis_ispc / VolumeVis_ispc
(Vis.ih:23-26, MappedVis.ih:26-30, VolumeVis.ih:26-37) and
Visualization_ispc (Visualization.ih:30-47).
"""
import ctypes as C

import numpy as np

RAY_PRIMARY, RAY_SHADOW, RAY_AO, RAY_EMPTY = 1, 2, 4, 8
RAY_SURFACE, RAY_OPAQUE, RAY_BOUNDARY, RAY_TIMEOUT = 1, 2, 4, 8
TERMINATED, DROP_ON_FLOOR, KEEP_HERE = -1, -2, -3

FLOAT_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "nx", "ny", "nz", "sample",
                "r", "g", "b", "o", "sr", "sg", "sb", "so", "t", "tMax")
INT_FIELDS = ("x", "y", "type", "term", "classification")
FIELDS = FLOAT_FIELDS + INT_FIELDS


class RayList_ispc(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in FIELDS]


class Lighting_ispc(C.Structure):
    _fields_ = [("n_ao_rays", C.c_int), ("do_shadows", C.c_int), ("ao_radius", C.c_float),
                ("Ka", C.c_float), ("Kd", C.c_float), ("types", C.c_void_p),
                ("lights", C.c_void_p), ("nLights", C.c_uint32)]


class Vis_ispc(C.Structure):
    _fields_ = [("data", C.c_void_p)]


class MappedVis_ispc(C.Structure):
    _fields_ = [("vis", Vis_ispc), ("transferFunction", C.c_void_p)]


class VolumeVis_ispc(C.Structure):
    _fields_ = [("vis", MappedVis_ispc), ("nSlices", C.c_int), ("slices", C.c_void_p),
                ("nIsovalues", C.c_int), ("isovalues", C.c_void_p), ("volume_render", C.c_bool)]


class box3f(C.Structure):
    _fields_ = [("lower", C.c_float * 3), ("upper", C.c_float * 3)]


class Visualization_ispc(C.Structure):
    _fields_ = [("model", C.c_void_p), ("nVolumeVis", C.c_int32),
                ("volumeVis", C.POINTER(C.POINTER(VolumeVis_ispc))),
                ("nMappedVis", C.c_int32), ("mappedVis", C.c_void_p),
                ("global_bb", box3f), ("local_bb", box3f)]


class FrameStats(C.Structure):
    _fields_ = [("primary_rays", C.c_longlong), ("secondary_rays", C.c_longlong),
                ("ray_passes", C.c_longlong), ("terminated", C.c_longlong),
                ("dropped", C.c_longlong), ("kept", C.c_longlong),
                ("outbound", C.c_longlong * 6), ("samples", C.c_longlong),
                ("passes", C.c_int), ("lists", C.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "outbound"}
        d["outbound"] = list(self.outbound)
        return d


class BatchStats(C.Structure):
    _fields_ = [("rays", C.c_ulonglong), ("samples", C.c_ulonglong), ("spawned", C.c_ulonglong),
                ("terminated", C.c_ulonglong), ("dropped", C.c_ulonglong), ("kept", C.c_ulonglong),
                ("passes", C.c_int), ("overflow", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


BATCH_RENDER, BATCH_SAMPLE = 0, 1


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("dataset", C.c_int), ("width", C.c_int), ("height", C.c_int),
                ("nrays", C.c_longlong), ("seed", C.c_uint), ("arrival", C.c_double),
                ("eye", C.c_float * 3), ("center", C.c_float * 3), ("up", C.c_float * 3), ("aov", C.c_float),
                ("out", C.c_void_p), ("capacity", C.c_longlong)]


class RequestResult(C.Structure):
    _fields_ = [("admitted", C.c_double), ("finished", C.c_double), ("particles", C.c_longlong),
                ("rays", C.c_longlong), ("status", C.c_int), ("pad_", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("wall", C.c_double), ("rays", C.c_ulonglong), ("samples", C.c_ulonglong),
                ("peak_bytes", C.c_longlong), ("passes", C.c_int), ("max_in_flight", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


SERVE_CONTINUOUS, SERVE_ONE_AT_A_TIME = 0, 1
RQ_DONE, RQ_FAILED, RQ_REJECTED, RQ_CANCELLED, RQ_QUEUED, RQ_RUNNING = 0, -1, -2, -3, 1, 2


def make_box(lo, hi):
    b = box3f()
    for k in range(3):
        b.lower[k] = float(lo[k])
        b.upper[k] = float(hi[k])
    return b


def fptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)


class HostRays:
    """A RayList with host (numpy) storage: 20 float32 and 5 int32 arrays."""

    def __init__(self, n):
        self.n = int(n)
        self.a = {f: np.zeros(max(self.n, 1), np.float32) for f in FLOAT_FIELDS}
        self.a.update({f: np.zeros(max(self.n, 1), np.int32) for f in INT_FIELDS})
        self.ispc = RayList_ispc(**{f: self.a[f].ctypes.data for f in FIELDS})

    def __getitem__(self, f):
        return self.a[f][:self.n]

    def __setitem__(self, f, v):
        self.a[f][:self.n] = v

    def copy(self):
        h = HostRays(self.n)
        for f in FIELDS:
            h.a[f][:] = self.a[f]
        return h

    def subset(self, idx):
        idx = np.asarray(idx)
        h = HostRays(len(idx))
        for f in FIELDS:
            h.a[f][:len(idx)] = self.a[f][:self.n][idx]
        return h

    @staticmethod
    def from_arrays(org, dirs, t0=None, tmax=None, typ=RAY_PRIMARY, xy=None):
        n = len(org)
        h = HostRays(n)
        h["ox"], h["oy"], h["oz"] = org[:, 0], org[:, 1], org[:, 2]
        h["dx"], h["dy"], h["dz"] = dirs[:, 0], dirs[:, 1], dirs[:, 2]
        h["t"] = 0.0 if t0 is None else t0
        h["tMax"] = np.finfo(np.float32).max if tmax is None else tmax
        h["type"] = typ
        if xy is not None:
            h["x"], h["y"] = xy[:, 0], xy[:, 1]
        return h
