import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(HERE, "golden")
GOLD_TESTS = ["absolute", "absolute-shadow", "camera", "camera-shadow", "infinite", "infinite-shadow",
              "nineBalls", "oneBall", "xyz"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through libgxy_hip.so)")


@pytest.fixture(scope="session")
def radial_vols():
    from galaxy_amd import scenes
    F = scenes.radial_fields(256)
    org, dl = scenes.vol_header_from_vti(256)
    return {"radial-%s.vol" % k: scenes.partitioned_volume(v, org, dl) for k, v in F.items()}


@pytest.fixture(scope="session")
def radial_vols_small():
    from galaxy_amd import scenes
    F = scenes.radial_fields(64)
    org, dl = scenes.vol_header_from_vti(64)
    return {"radial-%s.vol" % k: scenes.partitioned_volume(v, org, dl) for k, v in F.items()}


@pytest.fixture(scope="session")
def oracle():
    from oracle_backend import OracleBackend
    return OracleBackend()


@pytest.fixture(scope="session")
def hip():
    from galaxy_amd import lib
    return lib.HipBackend()


def gold_scenes(vols, name, width=512, height=512):
    from galaxy_amd import scenes
    return scenes.load_state(os.path.join(GOLDEN, "states", name + ".state"), vols, width, height)


def gold_image(name, k=0):
    from PIL import Image
    return np.array(Image.open(os.path.join(GOLDEN, "golds", "%s_%05d.png" % (name, k))))


def image_error(img, gold):
    d = img[..., :3].astype(np.float64) / 255 - gold[..., :3].astype(np.float64) / 255
    return float(np.sqrt((d ** 2).mean())), float((np.abs(d).max(-1) > 2.5 / 255).mean())


def random_rays(n, seed=0, lo=-1.5, hi=1.5, typ_mix=(0.7, 0.2, 0.1), tmax=1e30):
    """Seeded random rays: origins uniform in [lo, hi]^3, isotropic directions,
    type PRIMARY/SHADOW/AO with the given mix."""
    from galaxy_amd import abi
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    hr = abi.HostRays.from_arrays(o, d.astype(np.float32), tmax=np.full(n, tmax, np.float32))
    u = rng.random(n)
    hr["type"] = np.where(u < typ_mix[0], 1, np.where(u < typ_mix[0] + typ_mix[1], 2, 4)).astype(np.int32)
    hr["x"] = rng.integers(0, 64, n).astype(np.int32)
    hr["y"] = rng.integers(0, 64, n).astype(np.int32)
    return hr
