"""The viewer command front end on the GPU: a gold state driven through
json / window / render commands is rendered by the live server into the gold
image (Galaxy's image gold bar) and the oracle's frame of the same scene;
renderMany submits its frames together."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, gold_image, image_error
from galaxy_amd import frontend, scenes
from test_frontend import drive

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["oneBall", "camera-shadow"])
def test_gold_state_through_the_commands(name, hip, oracle, radial_vols):
    h = frontend.ViewerHandler(radial_vols, hip=hip)
    drive(h, name, 512, 512)
    assert h.frames == 1
    rmse, bad = image_error(h.image(), gold_image(name, 0))
    assert rmse <= 1e-3 and bad <= 0.005
    sc = scenes.load_state(os.path.join(GOLDEN, "states", name + ".state"), radial_vols, 512, 512)[0]
    cam = sc.cameras[0]
    fb, st = oracle.render_frame(scenes.build_scene(oracle, sc, scenes.resolve_lights(sc.lighting, cam)), cam)
    d = np.abs(h.frame - fb)
    assert np.mean(d <= 1e-4) >= 0.9999 and np.sqrt(np.mean(d.astype(np.float64) ** 2)) <= 1e-3
    first = h.frame
    assert h.handle("renderMany 2") == (True, "ok") and h.frames == 3
    # the same frame again (up to the order of the FP32 additions into a pixel)
    assert np.abs(h.frame - first).max() <= 1e-5
    # a camera command moves the frame (camera-relative lights resolve again)
    assert h.handle('json {"Camera": {"viewpoint": [3, 1, 3], "viewcenter": [0, 0, 0], "viewup": [0, 1, 0], '
                    '"aov": 30}}') == (True, "ok")
    assert h.handle("render") == (True, "ok") and h.frames == 4
    assert np.abs(h.frame - first).max() > 0.01
