"""The viewer command front end on the GPU: a gold state driven through
json / window / render commands renders the gold image (Galaxy's image gold
bar) and the oracle's frame of the same scene."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, gold_image, image_error
from galaxy_amd import frontend, scenes
from test_frontend import drive

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["oneBall", "camera-shadow"])
def test_gold_state_through_the_commands(name, hip, oracle, radial_vols):
    h = frontend.ViewerHandler(hip, radial_vols)
    drive(h, name, 512, 512)
    rmse, bad = image_error(h.image(), gold_image(name, 0))
    assert rmse <= 1e-3 and bad <= 0.005
    sc = scenes.load_state(os.path.join(GOLDEN, "states", name + ".state"), radial_vols, 512, 512)[0]
    cam = sc.cameras[0]
    fb, st = oracle.render_frame(scenes.build_scene(oracle, sc, scenes.resolve_lights(sc.lighting, cam)), cam)
    assert np.abs(h.frame - fb).max() <= 1e-4
    assert h.handle("renderMany 2") == (True, "ok") and h.frames == 3
