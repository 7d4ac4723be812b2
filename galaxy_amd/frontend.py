"""The multiserver viewer's command front end over this library.

`ViewerHandler.handle(line)` takes the command lines of Galaxy's viewer
handler (ViewerClientServer::handle, multiserver/ViewerClientServer.cpp:63-252)
and answers with the same replies and error texts:

    permute_pixels [on|off]     ok / error non-default arg ...
    max_rays_per_packet N       the list size limit of the rank loop
    commit                      rebuilds the device scene
    json {...}                  Datasets / Visualization / Camera / Renderer
    window W H                  image size
    render / renderMany N       one frame (or N) on the GPU

Datasets are resolved by name or file name through the `datasets` map the
handler is given (volumes already described on the host, as Galaxy's
Datasets::LoadFromJSON would load them); a `json` command holding Datasets
returns after loading them, as the reference does.  A frame goes through
gxy_render_frame (camera rays on the device, the per-list loop, accumulation
into a device framebuffer); the handler keeps the last frame in `frame`
(width*height*4 floats) and counts frames in `frames`, where the reference
hands the pixels to its socket handler.  The socket transport itself is not
part of this library."""
import json

from . import scenes


class ViewerHandler:
    def __init__(self, backend, datasets=None):
        self.be = backend
        self.catalog = dict(datasets or {})   # name or file name -> VolumeDesc
        self.dsets = {}                        # dataset name -> VolumeDesc (after a Datasets command)
        self.vis = None                        # Visualization JSON object
        self.cam = None                        # Camera JSON object
        self.eps = 0.001
        self.width, self.height = 512, 512
        self.permute = False
        self.maxlist = 1000000
        self.built = None
        self.scene = None
        self.frame = None
        self.frames = 0
        self.stats = None

    # ------------------------------------------------------------------ commands
    def handle(self, line):
        """Returns (handled, reply); unknown commands are not handled."""
        parts = line.strip().split(None, 1)
        cmd = parts[0] if parts else ""
        rest = parts[1] if len(parts) > 1 else ""
        if cmd == "permute_pixels":
            arg = rest.split()
            if not arg or arg[0] == "on":
                self.permute = True
                return True, "ok"
            if arg[0] == "off":
                self.permute = False
                return True, "ok"
            return True, "error non-default arg to permute_pixels must be on or off"
        if cmd == "max_rays_per_packet":
            try:
                self.maxlist = int(rest.split()[0])
            except (IndexError, ValueError):
                return True, "max_rays_per_packet requires integer arg"
            return True, "ok"
        if cmd == "commit":
            self.commit()
            return True, "ok"
        if cmd == "json":
            return True, self._json(rest)
        if cmd == "window":
            try:
                w, h = (int(x) for x in rest.split()[:2])
            except ValueError:
                return True, "error window command needs width and height as integers"
            self.width, self.height = w, h
            return True, "ok"
        if cmd == "render":
            self.render()
            return True, "ok"
        if cmd == "renderMany":
            try:
                n = int(rest.split()[0])
            except (IndexError, ValueError):
                n = 0
            for _ in range(n):
                self.render()
            return True, "ok"
        return False, ""

    def _json(self, text):
        try:
            doc = json.loads(text)
        except ValueError:
            return "error parsing json command"
        if not isinstance(doc, dict):
            return "error parsing json command"
        if "Datasets" in doc:
            try:
                for d in doc["Datasets"]:
                    vol = self.catalog.get(d.get("filename", d["name"]), self.catalog.get(d["name"]))
                    if vol is None:
                        raise KeyError(d["name"])
                    self.dsets[d["name"]] = vol
            except (KeyError, TypeError):
                return "error loading datasets from json command"
            self.built = None
            return "ok"
        if "Visualization" in doc:
            v = doc["Visualization"]
            try:
                scenes.vis_from_json(v, self.dsets, self.eps)
            except (KeyError, TypeError, ValueError):
                return "error loading visualization from json command"
            self.vis = v
            self.built = None
        if "Camera" in doc:
            try:
                scenes.camera_from_json(doc["Camera"], self.width, self.height)
            except (KeyError, TypeError, ValueError):
                return "error loading camera from json command"
            self.cam = doc["Camera"]
        if "Renderer" in doc:
            r = doc["Renderer"]
            if isinstance(r, dict) and "epsilon" in r:
                self.eps = float(r["epsilon"])
                self.built = None
        return "ok"

    # ------------------------------------------------------------------ state
    def commit(self):
        """Builds the device scene of the current visualization."""
        if self.vis is None:
            return
        self.scene = scenes.vis_from_json(self.vis, self.dsets, self.eps)
        self.built = None

    def render(self):
        if self.vis is None or self.cam is None:
            return
        cam = scenes.camera_from_json(self.cam, self.width, self.height)
        if self.built is None:
            self.scene = scenes.vis_from_json(self.vis, self.dsets, self.eps)
            # camera-relative lights follow the camera of each frame (Rendering.cpp:157-216)
            self.built = scenes.build_scene(self.be, self.scene, scenes.resolve_lights(self.scene.lighting, cam))
        else:
            # the device scene stays; only the lighting is re-resolved for this camera
            self.built.lights = scenes.make_lighting(scenes.resolve_lights(self.scene.lighting, cam), self.built.keep)
        self.frame, self.stats = self.be.render_frame(self.built, cam, self.scene.neighbors, self.scene.epsilon,
                                                      self.maxlist)
        self.frames += 1

    def image(self):
        """The last frame as 8-bit RGBA rows (ImageWriter's conversion)."""
        if self.frame is None:
            return None
        return scenes.fb_to_rgba8(self.frame, self.width, self.height)
