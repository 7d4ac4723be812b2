"""Execution contexts (gxy_context_create / bind / destroy): a frame rendered
on a fresh context, on contexts created after others were destroyed (their
host readback areas reused), and on two contexts driven from two threads at
once equals the default context's frame: same ray-pass, sample and
termination counts, framebuffers within FP32 atomic-order rounding."""
import threading

import numpy as np
import pytest

from galaxy_amd import scenes, workloads

pytestmark = pytest.mark.gpu
STATS = ("primary_rays", "secondary_rays", "ray_passes", "terminated", "samples")


@pytest.fixture(scope="module")
def frame_setup(hip):
    from galaxy_amd import lib as glib
    L_ = glib.lib()
    sc, alpha = workloads.nine_icospheres(256, 192, subdiv=4, n_ao=4)
    glib.check(L_.gxy_set_opacity_threshold(alpha))
    built = scenes.build_scene(hip, sc)
    fb0, st0 = hip.render_frame(built, sc.cameras[0], eps=sc.epsilon)
    assert st0.secondary_rays > st0.primary_rays > 10000
    return L_, sc, built, fb0, st0


def same_frame(fb, st, fb0, st0):
    for f in STATS:
        assert getattr(st, f) == getattr(st0, f), f
    assert np.abs(fb - fb0).max() <= 1e-5 * max(1.0, np.abs(fb0).max())


def on_context(L_, fn):
    from galaxy_amd import lib as glib
    ctx = L_.gxy_context_create()
    assert ctx
    try:
        glib.check(L_.gxy_context_bind(ctx))
        return fn()
    finally:
        L_.gxy_context_bind(None)
        glib.check(L_.gxy_context_destroy(ctx))


def test_frames_on_new_and_recreated_contexts(hip, frame_setup):
    L_, sc, built, fb0, st0 = frame_setup
    render = lambda: hip.render_frame(built, sc.cameras[0], eps=sc.epsilon)
    for _ in range(3):
        # each context after the first takes the readback area its predecessor left
        frames = on_context(L_, lambda: [render() for _ in range(2)])
        for fb, st in frames:
            same_frame(fb, st, fb0, st0)
    fb, st = render()   # the default context again
    same_frame(fb, st, fb0, st0)


def test_two_threads_on_their_own_contexts(hip, frame_setup):
    L_, sc, built, fb0, st0 = frame_setup
    out, errors = {}, []

    def work(k):
        try:
            out[k] = on_context(L_, lambda: [hip.render_frame(built, sc.cameras[0], eps=sc.epsilon)
                                             for _ in range(3)])
        except Exception as e:      # reported by the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not errors and len(out) == 2
    for frames in out.values():
        for fb, st in frames:
            same_frame(fb, st, fb0, st0)
