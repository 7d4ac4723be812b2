"""The CPU oracle against Galaxy's own image gold tests (tests/golds/*.png from
tests/image-gold-tests.sh:40-174): state file -> radial volumes regenerated
from radial.cpp math with the vti2vol header rounding -> camera, trace,
lighting, shadow/AO spawn, routing and accumulation -> 8-bit image.

The reference compares with perceptualdiff -fov 85; here the bar is an RGB
RMSE of at most 1e-3 (the 11 images reach 0-6e-4) and fewer than 0.5% of
pixels off by more than 2 LSB."""
import pytest

from conftest import GOLD_TESTS, gold_image, gold_scenes, image_error

CASES = [(t, k) for t in GOLD_TESTS for k in range(3 if t == "nineBalls" else 1)]


@pytest.mark.parametrize("name,cam", CASES)
def test_oracle_matches_gold(name, cam, radial_vols, oracle):
    from galaxy_amd import scenes
    sc = gold_scenes(radial_vols, name)[0]
    c = sc.cameras[cam]
    L = scenes.resolve_lights(sc.lighting, c)
    B = scenes.build_scene(oracle, sc, L)
    fb, st = oracle.render_frame(B, c)
    img = scenes.fb_to_rgba8(fb, c.width, c.height)
    rmse, bad = image_error(img, gold_image(name, cam))
    assert st.primary_rays > 100000
    assert rmse <= 1e-3, rmse
    assert bad <= 0.005, bad


def test_gold_differs_when_oracle_is_perturbed(radial_vols, oracle):
    """A plausible mistake must fail the gold check: render oneBall with the
    shadow flag removed (drops the negative shadow contribution)."""
    from galaxy_amd import scenes
    sc = gold_scenes(radial_vols, "absolute-shadow")[0]
    c = sc.cameras[0]
    L = scenes.resolve_lights(sc.lighting, c)
    L.shadows = False
    fb, _ = oracle.render_frame(scenes.build_scene(oracle, sc, L), c)
    rmse, bad = image_error(scenes.fb_to_rgba8(fb, c.width, c.height), gold_image("absolute-shadow"))
    assert rmse > 5e-3 and bad > 0.01
