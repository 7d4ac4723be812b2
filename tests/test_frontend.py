"""The viewer command front end (gxy_viewer_*, galaxy_amd/csrc/viewer.cpp;
ViewerClientServer::handle, multiserver/ViewerClientServer.cpp:63-252): replies
and error texts of every command, answered through the C-ABI without a GPU
(the frames of a gold state driven through the commands are checked in
tests/test_gpu_frontend.py)."""
import json
import os

from conftest import GOLDEN
from galaxy_amd import frontend


def _handler(vols):
    # names only: answering commands needs no device volumes
    return frontend.ViewerHandler({k: None for k in vols})


def test_command_replies(radial_vols_small):
    h = _handler(radial_vols_small)
    assert h.handle("permute_pixels") == (True, "ok") and h.permute
    assert h.handle("permute_pixels off") == (True, "ok") and not h.permute
    assert h.handle("permute_pixels maybe") == (True, "error non-default arg to permute_pixels must be on or off")
    assert h.handle("max_rays_per_packet 4096") == (True, "ok") and h.maxlist == 4096
    assert h.handle("max_rays_per_packet x") == (True, "max_rays_per_packet requires integer arg")
    assert h.handle("window 64 48") == (True, "ok") and (h.width, h.height) == (64, 48)
    assert h.handle("window 64") == (True, "error window command needs width and height as integers")
    assert h.handle("json {not json") == (True, "error parsing json command")
    assert h.handle("json [1, 2]") == (True, "error parsing json command")
    assert h.handle('json {"Datasets": [{"name": "nope", "type": "Volume", "filename": "nope.vol"}]}') == \
        (True, "error loading datasets from json command")
    assert h.handle('json {"Visualization": {"operators": [{"type": "Volume", "dataset": "missing"}]}}') == \
        (True, "error loading visualization from json command")
    assert h.handle('json {"Camera": {"viewpoint": [0, 0, 3]}}') == (True, "error loading camera from json command")
    assert h.handle("commit") == (True, "ok")
    assert h.handle("render") == (True, "ok") and h.frames == 0     # nothing to render yet
    assert h.handle("frobnicate 3") == (False, "")


def drive(h, name, width=128, height=128, render=True):
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
    if render:
        assert h.handle("render") == (True, "ok")
    return doc


def test_gold_states_load_through_the_commands(radial_vols_small):
    """Every gold state's datasets, visualization (operators and lighting),
    camera and renderer parse and resolve through the json commands."""
    for name in ("oneBall", "nineBalls", "camera-shadow", "infinite", "xyz", "absolute"):
        h = _handler(radial_vols_small)
        drive(h, name, render=False)
        assert (h.width, h.height) == (128, 128)
