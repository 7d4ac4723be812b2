"""Host side of the served form (gxy_server_* in include/gxy_hip.h): a request
loop over resident datasets that admits render and sample requests in arrival
order under an HBM budget and traces all in-flight requests together, pass by
pass (continuous batching), or one request at a time through the per-list
loop for comparison."""
import ctypes as C

import numpy as np

from . import abi, scenes
from . import lib as glib

f32 = np.float32


class Server:
    def __init__(self, budget_bytes, slots=64, chunk=256):
        self.L = glib.lib()
        self.h = glib.checkp(self.L.gxy_server_create(int(budget_bytes), slots, chunk), "server_create")
        self.keep = []
        self.boxes = []

    def destroy(self):
        if self.h:
            self.L.gxy_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def add_dataset(self, built, eps=0.001):
        self.keep.append(built)
        return glib.check(self.L.gxy_server_add_dataset(self.h, C.byref(built.vis), eps), "server_add_dataset")

    def set_lighting(self, lights, eps=0.001):
        self.keep.append(lights)
        glib.check(self.L.gxy_server_set_lighting(self.h, C.byref(lights), eps), "server_set_lighting")

    def set_config(self, chunk, stack, wpe, blocks_per_cu, split=0):
        b = self.L.gxy_server_batch(self.h)
        glib.check(self.L.gxy_batch_set_config(b, chunk, stack, wpe, blocks_per_cu), "batch_set_config")
        glib.check(self.L.gxy_batch_set_split(b, split), "batch_set_split")

    def set_step_rays(self, rays):
        """Primary rays of admitted requests per continuous-batching step
        (0 = all of them): the requests with the fewest rays left go first."""
        glib.check(self.L.gxy_batch_set_step_rays(self.L.gxy_server_batch(self.h), int(rays)), "batch_set_step_rays")

    def set_max_segments(self, n):
        """Isovalue crossings a sampler ray follows within one step (1..4096)."""
        glib.check(self.L.gxy_batch_set_max_segments(self.L.gxy_server_batch(self.h), int(n)), "set_max_segments")

    # ---------------------------------------------------------- live service
    def start(self):
        glib.check(self.L.gxy_server_start(self.h), "server_start")

    def submit(self, req):
        """Queues a copy of abi.Request req (its out buffer must outlive it);
        returns the request id."""
        return glib.check(self.L.gxy_server_submit(self.h, C.byref(req)), "server_submit")

    def poll(self, rid):
        """(status, RequestResult): 1 queued, 2 admitted, 0 done, -1 failed,
        -2 rejected, -3 cancelled."""
        res = abi.RequestResult()
        st = self.L.gxy_server_poll(self.h, rid, C.byref(res))
        return st, res

    def cancel(self, rid):
        glib.check(self.L.gxy_server_cancel(self.h, rid), "server_cancel")

    def drain(self, timeout_s):
        """0 when everything finished in time; otherwise the rest is cancelled
        and the call fails (GxyError)."""
        return glib.check(self.L.gxy_server_drain(self.h, float(timeout_s)), "server_drain")

    def stop(self):
        glib.check(self.L.gxy_server_stop(self.h), "server_stop")

    def run(self, reqs, mode=abi.SERVE_CONTINUOUS):
        """reqs: a ctypes array of abi.Request sorted by arrival.  Returns
        (results array, ServeStats)."""
        n = len(reqs)
        res = (abi.RequestResult * max(1, n))()
        st = abi.ServeStats()
        glib.check(self.L.gxy_server_run(self.h, reqs, n, mode, res, C.byref(st)), "server_run")
        return res, st


def render_request(ds, width, height, eye, center=(0.5, 0.5, 0.5), up=(0, 1, 0), aov=40.0, arrival=0.0, out=None):
    r = abi.Request()
    r.kind, r.dataset, r.width, r.height = 0, ds, width, height
    r.arrival = arrival
    r.eye[:] = list(eye)
    r.center[:] = list(center)
    r.up[:] = list(up)
    r.aov = aov
    r.out = out.ctypes.data if out is not None else None
    return r


def sample_request(ds, nrays, seed, capacity, arrival=0.0, out=None):
    r = abi.Request()
    r.kind, r.dataset, r.nrays, r.seed, r.capacity = 1, ds, nrays, seed, capacity
    r.arrival = arrival
    r.out = out.ctypes.data if out is not None else None
    return r


def orbit_eye(rng, dist=1.9, center=(0.5, 0.5, 0.5)):
    az = rng.uniform(0, 2 * np.pi)
    el = rng.uniform(-0.5, 0.8)
    return (center[0] + dist * np.cos(el) * np.sin(az), center[1] + dist * np.sin(el),
            center[2] + dist * np.cos(el) * np.cos(az))


def request_stream(n, rate, kinds, rng, render_px=(256 * 256, 3840 * 2160), sample_rays=(1_000_000, 50_000_000),
                   volume_ds=(), any_ds=(), keep_outputs=False, particles_per_ray=2):
    """n requests with Poisson arrivals at `rate` per second: render requests
    (pixel counts log-uniform in render_px, aspect 16:9 above 512^2, else
    square) over any dataset and sample requests (ray counts log-uniform in
    sample_rays) over volume datasets, mixed by `kinds` (probability of a
    render request).  Returns (ctypes array of abi.Request, host outputs)."""
    t = np.cumsum(rng.exponential(1.0 / rate, n)) - 0.0
    t -= t[0]
    reqs = (abi.Request * n)()
    outs = []
    for k in range(n):
        if rng.uniform() < kinds or not volume_ds:
            px = float(np.exp(rng.uniform(np.log(render_px[0]), np.log(render_px[1]))))
            if px > 512 * 512:
                h = int(round((px * 9 / 16) ** 0.5))
                w = int(round(h * 16 / 9))
            else:
                w = h = int(round(px ** 0.5))
            out = np.zeros(w * h * 4, f32) if keep_outputs else None
            reqs[k] = render_request(int(rng.choice(any_ds)), w, h, orbit_eye(rng), arrival=float(t[k]), out=out)
        else:
            nr = int(np.exp(rng.uniform(np.log(sample_rays[0]), np.log(sample_rays[1]))))
            cap = particles_per_ray * nr
            out = np.zeros(cap * 3, f32) if keep_outputs else None
            reqs[k] = sample_request(int(rng.choice(volume_ds)), nr, int(rng.integers(1 << 30)), cap,
                                     arrival=float(t[k]), out=out)
        outs.append(out)
    return reqs, outs


def latency_summary(reqs, res):
    lat = np.array([res[k].finished - reqs[k].arrival for k in range(len(reqs))])
    return dict(p50_ms=round(float(np.percentile(lat, 50)) * 1e3, 3), p99_ms=round(float(np.percentile(lat, 99)) * 1e3,
                                                                                   3),
                mean_ms=round(float(lat.mean()) * 1e3, 3), max_ms=round(float(lat.max()) * 1e3, 3))


def build_server(hip, descs, budget, slots=64, chunk=256, lighting=None, layout=0, step_rays=0):
    """Datasets from scene descriptions (volumes in the given voxel layout);
    returns (Server, Built list)."""
    from . import workloads
    S = Server(budget, slots, chunk)
    S.set_step_rays(step_rays)
    built = [scenes.build_scene(hip, sc) for sc in descs]
    for b, sc in zip(built, descs):
        if layout and sc.volume_vis:
            glib.check(S.L.gxy_volume_set_layout(b.handles[0], layout), "volume_set_layout")
        S.add_dataset(b)
    keep = []
    S.set_lighting(scenes.make_lighting(lighting or workloads.batch_lighting(), keep))
    S.keep.append(keep)
    return S, built
