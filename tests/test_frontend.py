"""The viewer command front end (ViewerClientServer::handle,
multiserver/ViewerClientServer.cpp:63-252): replies and error texts of every
command, and a gold state driven through json / window / render giving the
frame of the same scene rendered directly (oracle backend here, the GPU in
tests/test_gpu_frontend.py)."""
import json
import os

import numpy as np

from conftest import GOLDEN
from galaxy_amd import frontend, scenes


def _handler(backend, vols):
    return frontend.ViewerHandler(backend, vols)


def test_command_replies(oracle, radial_vols_small):
    h = _handler(oracle, radial_vols_small)
    assert h.handle("permute_pixels") == (True, "ok") and h.permute
    assert h.handle("permute_pixels off") == (True, "ok") and not h.permute
    assert h.handle("permute_pixels maybe") == (True, "error non-default arg to permute_pixels must be on or off")
    assert h.handle("max_rays_per_packet 4096") == (True, "ok") and h.maxlist == 4096
    assert h.handle("max_rays_per_packet x") == (True, "max_rays_per_packet requires integer arg")
    assert h.handle("window 64 48") == (True, "ok") and (h.width, h.height) == (64, 48)
    assert h.handle("window 64") == (True, "error window command needs width and height as integers")
    assert h.handle("json {not json") == (True, "error parsing json command")
    assert h.handle('json {"Datasets": [{"name": "nope", "type": "Volume", "filename": "nope.vol"}]}') == \
        (True, "error loading datasets from json command")
    assert h.handle('json {"Visualization": {"operators": [{"type": "Volume", "dataset": "missing"}]}}') == \
        (True, "error loading visualization from json command")
    assert h.handle('json {"Camera": {"viewpoint": [0, 0, 3]}}') == (True, "error loading camera from json command")
    assert h.handle("commit") == (True, "ok")
    assert h.handle("render") == (True, "ok") and h.frames == 0     # nothing to render yet
    assert h.handle("frobnicate 3") == (False, "")


def drive(h, name, width=128, height=128):
    """The viewer's command sequence for a gold state: datasets, then the
    visualization, camera and renderer, the window, one frame."""
    doc = json.load(open(os.path.join(GOLDEN, "states", name + ".state")))
    assert h.handle("json " + json.dumps({"Datasets": doc["Datasets"]})) == (True, "ok")
    rest = {"Visualization": doc["Visualizations"][0], "Camera": doc["Cameras"][0]}
    if "Renderer" in doc:
        rest["Renderer"] = doc["Renderer"]
    assert h.handle("json " + json.dumps(rest)) == (True, "ok")
    assert h.handle("window %d %d" % (width, height)) == (True, "ok")
    assert h.handle("commit") == (True, "ok")
    assert h.handle("render") == (True, "ok")
    return doc


def test_gold_state_through_the_commands_equals_a_direct_frame(oracle, radial_vols_small):
    h = _handler(oracle, radial_vols_small)
    drive(h, "oneBall")
    assert h.frames == 1 and h.frame is not None
    sc = scenes.load_state(os.path.join(GOLDEN, "states", "oneBall.state"), radial_vols_small, 128, 128)[0]
    cam = sc.cameras[0]
    fb, st = oracle.render_frame(scenes.build_scene(oracle, sc, scenes.resolve_lights(sc.lighting, cam)), cam)
    assert np.array_equal(h.frame, fb)
    # a second frame reuses the device scene
    assert h.handle("renderMany 2") == (True, "ok") and h.frames == 3
    assert np.array_equal(h.frame, fb)
