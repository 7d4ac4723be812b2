"""Served form: a Poisson stream of render and sample requests over resident
datasets gives every request the same result served with continuous batching,
served one at a time through the per-list loop, and served alone; an HBM
budget limits how many requests are in flight without changing any result."""
import numpy as np
import pytest

from batch_scenes import dataset_descs
from galaxy_amd import abi

pytestmark = pytest.mark.gpu
f32 = np.float32


@pytest.fixture(scope="module")
def server(hip):
    from galaxy_amd import serve
    S, built = serve.build_server(hip, dataset_descs(), budget=8 << 30, slots=64, chunk=256)
    return S


def stream(seed=3, n=14):
    from galaxy_amd import serve
    rng = np.random.default_rng(seed)
    return serve.request_stream(n, 200.0, 0.55, rng, render_px=(40 * 40, 110 * 110), sample_rays=(2000, 20000),
                                volume_ds=(0, 1), any_ds=(0, 1, 2, 3), keep_outputs=True)


def outputs(reqs, res, outs):
    got = []
    for k, o in enumerate(outs):
        r = reqs[k]
        if r.kind == 0:
            got.append(o.copy())
        else:
            c = int(res[k].particles)
            a = o[:3 * min(c, r.capacity)].reshape(-1, 3).copy()
            got.append((c, a[np.lexsort(a.T)]))
    return got


def same(a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        if isinstance(x, tuple):
            assert x[0] == y[0], k
            assert np.array_equal(x[1], y[1]), k
        else:
            assert np.abs(x - y).max() <= 1e-5, k


def test_stream_continuous_equals_one_at_a_time(server):
    from galaxy_amd import serve
    reqs, outs = stream()
    res, st = server.run(reqs, abi.SERVE_CONTINUOUS)
    a = outputs(reqs, res, outs)
    for o in outs:
        o[:] = 0
    res1, st1 = server.run(reqs, abi.SERVE_ONE_AT_A_TIME)
    b = outputs(reqs, res1, outs)
    same(a, b)
    assert st.rays == st1.rays and st.samples == st1.samples
    # (how many requests overlap depends on timing; the budget test checks a stream that arrives at once)
    assert st.max_in_flight >= 1
    for k in range(len(reqs)):
        assert res[k].status == 0 and res1[k].status == 0
        assert reqs[k].arrival <= res[k].admitted + 1e-6 <= res[k].finished + 2e-6
    lat = serve.latency_summary(reqs, res)
    assert lat["p99_ms"] >= lat["p50_ms"] > 0


def test_step_budget_gives_the_same_outputs(server):
    """Short steps (a few thousand primary rays, requests with the fewest rays
    left first) change when requests finish, not what they return."""
    reqs, outs = stream(seed=11, n=12)
    res, st = server.run(reqs, abi.SERVE_CONTINUOUS)
    a = outputs(reqs, res, outs)
    for o in outs:
        o[:] = 0
    try:
        server.set_step_rays(3000)
        res1, st1 = server.run(reqs, abi.SERVE_CONTINUOUS)
    finally:
        server.set_step_rays(0)
    same(a, outputs(reqs, res1, outs))
    assert st.rays == st1.rays and st.samples == st1.samples
    assert st1.passes > st.passes
    for k in range(len(reqs)):
        assert res1[k].status == 0


def test_each_request_alone_equals_the_mix(server):
    reqs, outs = stream(seed=5, n=10)
    res, _ = server.run(reqs, abi.SERVE_CONTINUOUS)
    mix = outputs(reqs, res, outs)
    for k in range(len(reqs)):
        one = (abi.Request * 1)()
        one[0] = reqs[k]
        one[0].arrival = 0.0
        outs[k][:] = 0
        r1, _ = server.run(one, abi.SERVE_CONTINUOUS)
        alone = outputs(one, r1, [outs[k]])
        same([mix[k]], alone)


def test_budget_limits_requests_in_flight(hip):
    from galaxy_amd import serve
    reqs, outs = stream(seed=7, n=8)
    for r in reqs:
        r.arrival = 0.0       # all at once: without a budget every request is in flight together
    S0, _ = serve.build_server(hip, dataset_descs(), budget=8 << 30, slots=64)
    res, st = S0.run(reqs, abi.SERVE_CONTINUOUS)
    ref = outputs(reqs, res, outs)
    assert st.max_in_flight == len(reqs)
    # the server's estimate: rays, outputs and two ray pools (1 + 3 secondaries per parent)
    est = [r.width * r.height * (116 + 2 * 4 * 100) if r.kind == 0 else r.nrays * 300 + r.capacity * 12
           for r in reqs]
    S1, _ = serve.build_server(hip, dataset_descs(), budget=max(est) + 1, slots=64)
    for o in outs:
        o[:] = 0
    res1, st1 = S1.run(reqs, abi.SERVE_CONTINUOUS)
    same(ref, outputs(reqs, res1, outs))
    assert st1.max_in_flight < len(reqs)
    assert st1.rays == st.rays


def test_over_budget_request_is_rejected_alone(hip):
    """A request over the HBM budget by itself, and one naming no dataset, are
    rejected (status -2) while the rest of the stream is served with its usual
    results; the next run on the same server starts clean."""
    from galaxy_amd import lib as glib
    from galaxy_amd import serve
    reqs, outs = stream(seed=13, n=6)
    for r in reqs:
        r.arrival = 0.0
    est = [r.width * r.height * (116 + 2 * 4 * 100) if r.kind == 0 else r.nrays * 300 + r.capacity * 12
           for r in reqs]
    S, _ = serve.build_server(hip, dataset_descs(), budget=max(est) + 1, slots=64)
    ref_res, _ = S.run(reqs, abi.SERVE_ONE_AT_A_TIME)
    ref = outputs(reqs, ref_res, outs)
    big = serve.sample_request(0, 4_000_000, 9, 100, arrival=0.0)       # 1.2 GB estimate, over the budget
    bad = serve.render_request(17, 32, 32, (0.5, 0.5, 3.0), arrival=0.0)  # no dataset 17
    mixed = (abi.Request * (len(reqs) + 2))()
    mixed[0] = big
    for k in range(len(reqs)):
        mixed[k + 1] = reqs[k]
    mixed[len(reqs) + 1] = bad
    for o in outs:
        o[:] = 0
    L = glib.lib()
    b = L.gxy_server_batch(S.h)

    def held():   # device bytes outside the batch's own pools and tables
        return L.gxy_device_bytes_in_use() - L.gxy_batch_bytes(b)
    S.run(reqs, abi.SERVE_CONTINUOUS)
    for o in outs:
        o[:] = 0
    before = held()
    res, st = S.run(mixed, abi.SERVE_CONTINUOUS)
    assert res[0].status == abi.RQ_REJECTED and res[len(reqs) + 1].status == abi.RQ_REJECTED
    assert all(res[k + 1].status == abi.RQ_DONE for k in range(len(reqs)))
    got = outputs(reqs, [res[k + 1] for k in range(len(reqs))], outs)
    same(ref, got)
    # no buffer of the run is left on the device
    assert held() == before
    res2, _ = S.run(reqs, abi.SERVE_CONTINUOUS)
    assert all(res2[k].status == abi.RQ_DONE for k in range(len(reqs)))


def _wait_final(S, rid, timeout=60.0):
    import time
    t = time.time()
    while time.time() - t < timeout:
        st, res = S.poll(rid)
        if st <= 0:
            return st, res
        time.sleep(0.002)
    raise AssertionError("request %d not final after %.0f s" % (rid, timeout))


def test_live_server_submit_poll_cancel_and_watchdog(hip):
    """The live form: requests submitted while others run give the results of
    the stream replay; a long request cancelled while admitted ends with
    status -3 and releases its buffers; a drain whose timeout passes cancels
    what is left and fails; the server takes new requests afterwards."""
    from galaxy_amd import serve
    from galaxy_amd import lib as glib
    S, _ = serve.build_server(hip, dataset_descs(), budget=8 << 30, slots=64)
    reqs, outs = stream(seed=17, n=8)
    ref_res, _ = S.run(reqs, abi.SERVE_ONE_AT_A_TIME)
    ref = outputs(reqs, ref_res, outs)
    for o in outs:
        o[:] = 0
    S.set_step_rays(20000)          # short steps, so a long request spans many of them
    L = glib.lib()
    b = L.gxy_server_batch(S.h)

    def held():   # device bytes outside the batch's own pools and tables
        return L.gxy_device_bytes_in_use() - L.gxy_batch_bytes(b)
    S.start()
    try:
        ids = [S.submit(reqs[k]) for k in range(len(reqs))]
        finals = [_wait_final(S, i) for i in ids]
        assert all(st == abi.RQ_DONE for st, _ in finals)
        same(ref, outputs(reqs, [r for _, r in finals], outs))
        base = held()
        # cancellation of an admitted request
        long_out = np.zeros(3 * 2_000_000, f32)
        lr = serve.sample_request(1, 2_000_000, 23, 2_000_000, out=long_out)
        lid = S.submit(lr)
        import time
        t = time.time()
        while S.poll(lid)[0] == abi.RQ_QUEUED and time.time() - t < 30:
            time.sleep(0.001)
        S.cancel(lid)
        st, res = _wait_final(S, lid)
        assert st == abi.RQ_CANCELLED
        # the watchdog: a drain that cannot finish in time cancels and fails
        wid = S.submit(serve.sample_request(1, 3_000_000, 29, 10, out=np.zeros(30, f32)))
        with pytest.raises(glib.GxyError, match="drain timed out"):
            S.drain(0.0005)
        assert S.poll(wid)[0] in (abi.RQ_CANCELLED, abi.RQ_DONE)
        assert held() <= base
        # and the server still serves
        again = S.submit(reqs[0])
        assert _wait_final(S, again)[0] == abi.RQ_DONE
        S.drain(60.0)
    finally:
        S.stop()
        S.set_step_rays(0)
