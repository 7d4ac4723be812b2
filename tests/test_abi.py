"""The C-ABI library: it loads without a GPU, exports every function that
include/gxy_hip.h declares, the ctypes mirrors match the C struct layouts (as
compiled by gcc from the header), and GPU entry points fail loudly when no GPU
is present instead of falling back to anything."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from conftest import ROOT
from galaxy_amd import abi

HEADER = os.path.join(ROOT, "include", "gxy_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_library_exports_every_declared_symbol():
    from galaxy_amd import lib
    L = lib.load()
    fns = declared_functions()
    assert len(fns) > 60
    missing = [f for f in fns if not hasattr(L, f)]
    assert not missing, missing
    # and the Python binding declares all of them
    assert set(fns) <= set(lib.EXPORTS), sorted(set(fns) - set(lib.EXPORTS))


def test_struct_layouts_match_the_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "gxy_hip.h"
#define P(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f))
#define S(T) printf("%s %zu\n", #T, sizeof(T))
int main(void) {
  S(RayList_ispc); P(RayList_ispc, tMax); P(RayList_ispc, x); P(RayList_ispc, classification);
  S(Lighting_ispc); P(Lighting_ispc, types); P(Lighting_ispc, lights); P(Lighting_ispc, nLights);
  S(VolumeVis_ispc); P(VolumeVis_ispc, nSlices); P(VolumeVis_ispc, slices); P(VolumeVis_ispc, isovalues);
  P(VolumeVis_ispc, volume_render);
  S(Visualization_ispc); P(Visualization_ispc, volumeVis); P(Visualization_ispc, global_bb);
  P(Visualization_ispc, local_bb);
  S(gxy_raylist_hdr); P(gxy_raylist_hdr, size); P(gxy_raylist_hdr, aligned_size);
  S(gxy_frame_stats); P(gxy_frame_stats, outbound); P(gxy_frame_stats, samples); P(gxy_frame_stats, lists);
  S(gxy_batch_stats); P(gxy_batch_stats, kept); P(gxy_batch_stats, passes); P(gxy_batch_stats, overflow);
  S(gxy_request); P(gxy_request, nrays); P(gxy_request, arrival); P(gxy_request, aov); P(gxy_request, out);
  P(gxy_request, capacity);
  S(gxy_request_result); P(gxy_request_result, rays); P(gxy_request_result, status);
  S(gxy_serve_stats); P(gxy_serve_stats, peak_bytes); P(gxy_serve_stats, max_in_flight);
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = dict(zip(out[0::2], map(int, out[1::2])))
    py = {
        "RayList_ispc": C.sizeof(abi.RayList_ispc), "RayList_ispc.tMax": abi.RayList_ispc.tMax.offset,
        "RayList_ispc.x": abi.RayList_ispc.x.offset,
        "RayList_ispc.classification": abi.RayList_ispc.classification.offset,
        "Lighting_ispc": C.sizeof(abi.Lighting_ispc), "Lighting_ispc.types": abi.Lighting_ispc.types.offset,
        "Lighting_ispc.lights": abi.Lighting_ispc.lights.offset,
        "Lighting_ispc.nLights": abi.Lighting_ispc.nLights.offset,
        "VolumeVis_ispc": C.sizeof(abi.VolumeVis_ispc), "VolumeVis_ispc.nSlices": abi.VolumeVis_ispc.nSlices.offset,
        "VolumeVis_ispc.slices": abi.VolumeVis_ispc.slices.offset,
        "VolumeVis_ispc.isovalues": abi.VolumeVis_ispc.isovalues.offset,
        "VolumeVis_ispc.volume_render": abi.VolumeVis_ispc.volume_render.offset,
        "Visualization_ispc": C.sizeof(abi.Visualization_ispc),
        "Visualization_ispc.volumeVis": abi.Visualization_ispc.volumeVis.offset,
        "Visualization_ispc.global_bb": abi.Visualization_ispc.global_bb.offset,
        "Visualization_ispc.local_bb": abi.Visualization_ispc.local_bb.offset,
        "gxy_frame_stats": C.sizeof(abi.FrameStats), "gxy_frame_stats.outbound": abi.FrameStats.outbound.offset,
        "gxy_frame_stats.samples": abi.FrameStats.samples.offset, "gxy_frame_stats.lists": abi.FrameStats.lists.offset,
        "gxy_batch_stats": C.sizeof(abi.BatchStats), "gxy_batch_stats.kept": abi.BatchStats.kept.offset,
        "gxy_batch_stats.passes": abi.BatchStats.passes.offset,
        "gxy_batch_stats.overflow": abi.BatchStats.overflow.offset,
        "gxy_request": C.sizeof(abi.Request), "gxy_request.nrays": abi.Request.nrays.offset,
        "gxy_request.arrival": abi.Request.arrival.offset, "gxy_request.aov": abi.Request.aov.offset,
        "gxy_request.out": abi.Request.out.offset, "gxy_request.capacity": abi.Request.capacity.offset,
        "gxy_request_result": C.sizeof(abi.RequestResult), "gxy_request_result.rays": abi.RequestResult.rays.offset,
        "gxy_request_result.status": abi.RequestResult.status.offset,
        "gxy_serve_stats": C.sizeof(abi.ServeStats), "gxy_serve_stats.peak_bytes": abi.ServeStats.peak_bytes.offset,
        "gxy_serve_stats.max_in_flight": abi.ServeStats.max_in_flight.offset,
    }
    for k, v in py.items():
        assert got[k] == v, (k, got[k], v)
    assert got["RayList_ispc"] == 25 * 8          # Rays.ih:21-48: 25 pointers
    assert got["gxy_raylist_hdr"] == 64           # Rays.cpp:37 HDRSZ
    assert got["gxy_raylist_hdr.size"] == 24      # after three 8-byte Keys (Rays.h:51-54)


def test_gpu_entry_points_fail_loudly_without_gpu():
    from galaxy_amd import lib
    L = lib.load()
    if L.gxy_device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(lib.GxyError):
        lib.lib()
    assert L.gxy_volume_create(None, None, None, None, 1.0) is None
    assert b"no HIP device" in L.gxy_last_error()
