"""The multiserver viewer's command front end (gxy_viewer_* in
include/gxy_hip.h, galaxy_amd/csrc/viewer.cpp).

`ViewerHandler.handle(line)` passes the command lines of Galaxy's viewer
handler (ViewerClientServer::handle, multiserver/ViewerClientServer.cpp:63-252)
to the C++ viewer and returns (handled, reply) with the reference's replies
and error texts:

    permute_pixels [on|off]     ok / error non-default arg ...
    max_rays_per_packet N       kept and reported
    commit                      the next frame rebuilds the served visualization
    json {...}                  Datasets / Visualization / Camera / Renderer
    window W H                  image size
    render / renderMany N       one frame (or N together) through the live server

Datasets are resolved by file name or name against the volumes the handler is
given (VolumeDesc on the host, uploaded once when a GPU is present, as
Galaxy's Datasets::LoadFromJSON would load them).  The last frame is `frame`
(width*height*4 floats) and `frames` counts frames; the socket transport of
the multiserver is not part of this library."""
import ctypes as C

import numpy as np

from . import abi, scenes
from . import lib as glib


class ViewerHandler:
    def __init__(self, datasets=None, budget=0, slots=0, hip=None):
        """datasets: key (file name or name) -> scenes.VolumeDesc.  Volumes are
        uploaded through `hip` (a lib.HipBackend) when given; without it only
        the names are registered (commands can be answered, frames not)."""
        self.L = glib.load()
        self.h = glib.checkp(self.L.gxy_viewer_create(int(budget), int(slots)), "viewer_create")
        self.keep = []
        for key, vol in (datasets or {}).items():
            handle = None
            if hip is not None and vol is not None:
                vox = np.ascontiguousarray(vol.voxels, np.float32)
                handle = hip.volume_create(vol.dims, vol.origin, vol.spacing, vox, vol.sampling_rate)
                self.keep += [vox, handle]
            lb = abi.make_box(*(vol.local_box if vol is not None else ((-1,) * 3, (1,) * 3)))
            gb = abi.make_box(*(vol.global_box if vol is not None else ((-1,) * 3, (1,) * 3)))
            glib.check(self.L.gxy_viewer_add_volume(self.h, key.encode(), handle, C.byref(lb), C.byref(gb)),
                       "viewer_add_volume")

    def destroy(self):
        if self.h:
            self.L.gxy_viewer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def handle(self, line):
        """Returns (handled, reply); unknown commands are not handled."""
        buf = C.create_string_buffer(4096)
        r = self.L.gxy_viewer_handle(self.h, line.encode(), buf, len(buf))
        if r < 0:
            raise glib.GxyError("viewer_handle failed")
        return bool(r), buf.value.decode()

    def _settings(self):
        v = [C.c_int() for _ in range(4)]
        self.L.gxy_viewer_settings(self.h, *[C.byref(x) for x in v])
        return [x.value for x in v]

    @property
    def permute(self):
        return bool(self._settings()[0])

    @property
    def maxlist(self):
        return self._settings()[1]

    @property
    def width(self):
        return self._settings()[2]

    @property
    def height(self):
        return self._settings()[3]

    @property
    def frames(self):
        n = C.c_int()
        self.L.gxy_viewer_frame(self.h, None, None, None, C.byref(n))
        return n.value

    @property
    def frame(self):
        w, h = C.c_int(), C.c_int()
        if self.L.gxy_viewer_frame(self.h, None, C.byref(w), C.byref(h), None) != 1:
            return None
        out = np.zeros(w.value * h.value * 4, np.float32)
        self.L.gxy_viewer_frame(self.h, out.ctypes.data, None, None, None)
        return out

    def image(self):
        """The last frame as 8-bit RGBA rows (ImageWriter's conversion)."""
        f = self.frame
        if f is None:
            return None
        w, h = C.c_int(), C.c_int()
        self.L.gxy_viewer_frame(self.h, None, C.byref(w), C.byref(h), None)
        return scenes.fb_to_rgba8(f, w.value, h.value)
