"""Batched form: many uneven ray lists over several resident datasets traced
pass by pass by the persistent kernel give, list by list, the results of the
per-list loop (gxy_process_raylist / gxy_sample_raylist), and those of the
oracle; every launch configuration and the graph-captured run give the same
results; lists added while others are in flight (continuous form) get the
results they get alone."""
import ctypes as C

import numpy as np
import pytest

from batch_scenes import dataset_descs
from galaxy_amd import abi, scenes, workloads

pytestmark = pytest.mark.gpu
f32 = np.float32

SIZES = [1, 63, 64, 65, 200, 777, 1000, 2048, 3001, 5, 4096, 129]


@pytest.fixture(scope="module")
def datasets(hip):
    descs = dataset_descs()
    return descs, [scenes.build_scene(hip, sc) for sc in descs]


def lights_struct(keep):
    return scenes.make_lighting(workloads.batch_lighting(), keep)


def plain_list(L, built, lights, n, seed, ds_kind):
    """The per-list loop on the same rays: (fb, FrameStats)."""
    from galaxy_amd import batch as gb
    from galaxy_amd import lib as glib
    d = gb.random_rays(n, seed, width=n)
    fb = glib.DeviceArray(16 * n).zero()
    st = abi.FrameStats()
    nb = np.full(6, -1, np.int32)
    glib.check(L.gxy_process_raylist(C.byref(built.vis), C.byref(lights), n, C.byref(d.ispc), nb.ctypes.data, 0.001,
                                      1 << 30, fb.ptr, n, 1, C.byref(st)), "process_raylist")
    return fb.to_host(f32).reshape(-1, 4), st


def make_batch(built, lights, plan, chunk=256):
    from galaxy_amd import batch as gb
    B = gb.Batch(chunk)
    ids = [B.add_dataset(b) for b in built]
    B.set_lighting(lights)
    out = []
    for n, ds, seed in plan:
        d = gb.random_rays(n, seed, width=n)
        l, fb = B.add_render_list(ids[ds], d, n, 1)
        B.set_seed(l, seed)
        out.append((l, fb))
    return B, out


def plan_of(sizes, nds, seed0=100):
    return [(n, k % nds, seed0 + k) for k, n in enumerate(sizes)]


def test_batch_equals_per_list_loop(datasets, hip):
    from galaxy_amd import lib as glib
    L = glib.lib()
    descs, built = datasets
    keep = []
    lights = lights_struct(keep)
    plan = plan_of(SIZES, len(built))
    B, lists = make_batch(built, lights, plan)
    st = B.run(0)
    assert st.overflow == 0
    tot = dict(rays=0, samples=0, spawned=0, terminated=0, dropped=0, kept=0)
    for (n, ds, seed), (l, fb) in zip(plan, lists):
        ref, s = plain_list(L, built[ds], lights, n, seed, ds)
        got = fb.to_host(f32).reshape(-1, 4)
        # per-ray results are identical; only the order of float additions into a pixel may differ
        assert np.abs(got - ref).max() <= 1e-5, (n, ds)
        tot["rays"] += s.ray_passes
        tot["samples"] += s.samples
        tot["spawned"] += s.secondary_rays
        tot["terminated"] += s.terminated
        tot["dropped"] += s.dropped
        tot["kept"] += s.kept
    got = {k: getattr(st, k) for k in tot}
    assert got == tot, (got, tot)
    assert st.spawned > 1000 and st.samples > 100000


def test_graph_run_and_every_configuration_agree(datasets, hip):
    from galaxy_amd import batch as gb
    descs, built = datasets
    keep = []
    lights = lights_struct(keep)
    plan = plan_of(SIZES, len(built), 300)
    B, lists = make_batch(built, lights, plan)
    st0 = B.run(0)
    ref = [fb.to_host(f32) for _, fb in lists]
    depth = 0
    for b in built:
        if b.vis.model:
            nodes, leaves, dp, sah = C.c_int(), C.c_int(), C.c_int(), C.c_float()
            B.L.gxy_model_bvh_stats(b.vis.model, C.byref(nodes), C.byref(leaves), C.byref(dp), C.byref(sah))
            depth = max(depth, dp.value)
    runs = 0
    for cfg in gb.DEFAULT_CANDIDATES + [(64, 0, 0, 1), (1024, 0, 0, 2)]:
        if cfg[1] and depth + 2 > cfg[1]:
            continue
        B.set_config(*cfg)
        for gp in (0, 2):
            st = B.run(gp)
            assert st.overflow == 0
            for k in ("rays", "samples", "spawned", "terminated", "dropped", "kept"):
                assert getattr(st, k) == getattr(st0, k), (cfg, gp, k)
            for (l, fb), r in zip(lists, ref):
                assert np.abs(fb.to_host(f32) - r).max() <= 1e-5, (cfg, gp, l)
            runs += 1
    assert runs >= 8
    # graph replays after eager work in between give the same results again
    for _ in range(3):
        st = B.run(2)
        assert st.rays == st0.rays
        for (l, fb), r in zip(lists, ref):
            assert np.abs(fb.to_host(f32) - r).max() <= 1e-5


def test_batch_against_oracle(datasets, hip, oracle):
    """A few lists checked straight against the CPU restatement."""
    from galaxy_amd import batch as gb
    descs, built = datasets
    keep = []
    lights = lights_struct(keep)
    plan = plan_of([3000, 2500, 4000, 1500], len(built), 700)
    B, lists = make_batch(built, lights, plan)
    B.run(0)
    nb = np.full(6, -1, np.int32)
    for (n, ds, seed), (l, fb) in zip(plan, lists):
        hr = gb.random_rays(n, seed, width=n).download()
        Bo = scenes.build_scene(oracle, descs[ds])
        fbo = np.zeros(4 * n, f32)
        so = abi.FrameStats()
        oracle.L.oracle_process_raylist(C.byref(Bo.vis), C.byref(lights), hr.n,
                                        C.byref(hr.ispc), nb.ctypes.data, 0.001, 1 << 30, fbo.ctypes.data, n, 1,
                                        C.byref(so))
        got = fb.to_host(f32)
        assert np.abs(got - fbo).max() <= 1e-4, (n, ds)
        assert np.sqrt(np.mean((got - fbo) ** 2)) <= 1e-5


def test_sample_lists_equal_sampler_loop(datasets, hip):
    from galaxy_amd import batch as gb
    from galaxy_amd import lib as glib
    L = glib.lib()
    descs, built = datasets
    B = gb.Batch(128)
    ids = [B.add_dataset(b) for b in built]
    keep = []
    B.set_lighting(lights_struct(keep))
    cap = 400000
    runs = []
    for k, n in enumerate([5000, 777, 20000]):
        d = gb.random_rays(n, 900 + k, width=n)
        l, xyz, cnt = B.add_sample_list(ids[k % 2], d, cap)
        B.set_seed(l, 900 + k)
        runs.append((n, k % 2, 900 + k, xyz, cnt))
    # a render list in the same batch
    d = gb.random_rays(3000, 950, width=3000)
    lr, fbr = B.add_render_list(ids[3], d, 3000, 1)
    B.set_seed(lr, 950)
    st = B.run(0)
    # sampler rays follow all their crossings within a pass; the render list needs its shadow/AO pass
    assert st.passes == 2
    for n, ds, seed, xyz, cnt in runs:
        c = int(cnt.to_host(np.int32, 1)[0])
        d = gb.random_rays(n, seed, width=n)
        xr = glib.DeviceArray(cap * 12)
        cr = C.c_int()
        fs = abi.FrameStats()
        glib.check(L.gxy_sample_raylist(C.byref(built[ds].vis), n, C.byref(d.ispc), 1 << 30, xr.ptr, cap,
                                        C.byref(cr), C.byref(fs)))
        assert c == cr.value > n // 4
        a = xyz.to_host(f32, 3 * c).reshape(-1, 3)
        b = xr.to_host(f32, 3 * c).reshape(-1, 3)
        # the same particles; their order across passes and lists differs
        a = a[np.lexsort(a.T)]
        b = b[np.lexsort(b.T)]
        assert np.array_equal(a, b)


@pytest.mark.parametrize("step_rays", [0, 1500])
def test_continuous_batching_alone_equals_mix(datasets, hip, step_rays):
    """Lists added while others are in flight, with passes run one at a time,
    get the results they get in a batch of their own; with a step budget the
    primary rays of the lists join over several steps, fewest left first."""
    from galaxy_amd import batch as gb
    descs, built = datasets
    keep = []
    lights = lights_struct(keep)
    plan = plan_of([4000, 1500, 64, 3000, 999, 2500, 10, 4096], len(built), 1200)
    alone = []
    for p in plan:
        B1, ls = make_batch(built, lights, [p])
        B1.run(0)
        alone.append(ls[0][1].to_host(f32))
        B1.destroy()
    B = gb.Batch(256)
    B.set_step_rays(step_rays)
    assert B.L.gxy_batch_get_step_rays(B.h) == step_rays
    ids = [B.add_dataset(b) for b in built]
    B.set_lighting(lights)
    fbs = {}
    pending_plan = list(enumerate(plan))
    queued = []
    steps = 0
    done = set()
    started = {}
    while len(done) < len(plan):
        # admit up to two new lists per step
        for _ in range(2):
            if pending_plan:
                k, (n, ds, seed) = pending_plan.pop(0)
                d = gb.random_rays(n, seed, width=n)
                l, fb = B.add_render_list(ids[ds], d, n, 1)
                fbs[k] = fb
                started[k] = l
        queued.append(B.step())
        steps += 1
        pend = B.pending(len(B.lists) + 8)
        assert queued[-1] == pend.sum()
        for k, l in list(started.items()):
            if pend[l] == 0:
                done.add(k)
                del started[k]
                B.remove_list(l)
        assert steps < 100
    for k in range(len(plan)):
        assert np.abs(fbs[k].to_host(f32) - alone[k]).max() <= 1e-5, k
    if step_rays:
        # 16169 primary rays, at most 1500 + 255 (a whole last chunk) per step
        assert steps >= 9


def test_bricked_volume_datasets_give_the_same_results(datasets, hip):
    from galaxy_amd import lib as glib
    L = glib.lib()
    descs, built = datasets
    keep = []
    lights = lights_struct(keep)
    plan = plan_of(SIZES, len(built), 1500)
    B, lists = make_batch(built, lights, plan)
    st0 = B.run(0)
    ref = [fb.to_host(f32) for _, fb in lists]
    vols = [b.handles[0] for b, d in zip(built, descs) if d.volume_vis]
    try:
        for lay in (8, 4):
            for v in vols:
                glib.check(L.gxy_volume_set_layout(v, lay))
            B.refresh_datasets(built)
            for gp in (0, 2):
                st = B.run(gp)
                for k in ("rays", "samples", "spawned", "terminated", "dropped", "kept"):
                    assert getattr(st, k) == getattr(st0, k), (lay, gp, k)
                for (l, fb), r in zip(lists, ref):
                    assert np.abs(fb.to_host(f32) - r).max() <= 1e-5, (lay, gp, l)
    finally:
        for v in vols:
            L.gxy_volume_set_layout(v, 0)


def test_sampler_segment_cap_gives_the_same_particles(datasets, hip):
    """A sampler ray that reaches the per-pass crossing cap continues in the
    next pass: particle sets, ray passes and samples do not depend on the cap."""
    from galaxy_amd import batch as gb
    descs, built = datasets
    ref = None
    for cap_seg in (4096, 3, 1):
        B = gb.Batch(128)
        B.set_max_segments(cap_seg)
        ids = [B.add_dataset(b) for b in built]
        keep = []
        B.set_lighting(lights_struct(keep))
        lists = []
        for k, n in enumerate([5000, 20000]):
            d = gb.random_rays(n, 1900 + k, width=n)
            l, xyz, cnt = B.add_sample_list(ids[k % 2], d, 400000)
            B.set_seed(l, 1900 + k)
            lists.append((xyz, cnt))
        st = B.run(2)
        got = []
        for xyz, cnt in lists:
            c = int(cnt.to_host(np.int32, 1)[0])
            a = xyz.to_host(f32, 3 * c).reshape(-1, 3)
            got.append(a[np.lexsort(a.T)])
        if ref is None:
            ref = (got, st.rays, st.samples, st.passes)     # the graph's two passes, the second empty
        else:
            assert st.rays == ref[1] and st.samples == ref[2], cap_seg
            assert st.passes >= ref[3], cap_seg
            if cap_seg == 1:
                assert st.passes > ref[3]
            for a, b in zip(got, ref[0]):
                assert np.array_equal(a, b), cap_seg
        B.destroy()


def test_batch_direct_lighting_equals_per_list_loop(datasets, hip):
    """The direct lighting model (gxy_set_reverse_lighting(0), taken by the
    batch when its lighting is set) gives the per-list loop's frames."""
    from galaxy_amd import lib as glib
    L = glib.lib()
    descs, built = datasets
    keep = []
    glib.check(L.gxy_set_reverse_lighting(0))
    try:
        lights = lights_struct(keep)
        plan = plan_of([64, 777, 3001, 129, 2048, 5], len(built), 500)
        B, lists = make_batch(built, lights, plan)
        st = B.run(0)
        assert st.overflow == 0
        for (n, ds, seed), (l, fb) in zip(plan, lists):
            ref, s = plain_list(L, built[ds], lights, n, seed, ds)
            assert np.abs(fb.to_host(f32).reshape(-1, 4) - ref).max() <= 1e-5, (n, ds)
    finally:
        glib.check(L.gxy_set_reverse_lighting(1))


def test_pending_counts_lists_added_after_a_step(datasets, hip):
    """Lists added after a step report their waiting rays only: the pass
    counters cover the lists of the last passes and no entry beyond them."""
    from galaxy_amd import batch as gb
    descs, built = datasets
    keep = []
    lights = lights_struct(keep)
    B, lists = make_batch(built, lights, plan_of([500, 700, 900], len(built), 700))
    B.step()
    extra = []
    for k, n in enumerate((1000, 3000, 50, 4096, 7)):
        d = gb.random_rays(n, 800 + k, width=n)
        l, fb = B.add_render_list(k % len(built), d, n, 1)
        extra.append((l, n))
    p = B.pending(3 + len(extra))
    assert all(0 <= v <= 50 * 900 for v in p[:3])
    assert [int(p[l]) for l, _ in extra] == [n for _, n in extra]
