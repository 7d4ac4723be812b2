#!/usr/bin/env python3
"""Benchmark of the MI355X ray-processing path.

Headline workload: one frame of the volume render,
1920x1080 primary rays through a 512^3 FP32 noise volume (DVR with an 8-point
TF, 0.5-voxel step, early termination at alpha 0.99, isosurface 0.5) plus the
shadow rays the isosurface hits spawn, classified and accumulated into the
framebuffer.  A step is one whole frame: camera rays, every trace pass, spawn,
lighting, routing and accumulation, all on the GPU with the volume resident in
HBM.  value = ray passes (primary + shadow) / second; Gsamples/s beside it.

Also measured, as extra fields: the sphere-mesh frame and the nine-icosphere
AO frame; the batched form (256 ray lists over 64 resident datasets, one
persistent launch per pass) against the per-list loop; the served form (a
Poisson request stream with continuous batching, p50/p99 latency) against
serving one request at a time.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 (vector and matrix)
BYTES_PER_RAY_PASS = 104     # 52 B read + 52 B written per traced ray
FLOPS_PER_SAMPLE = 36        # trilinear sample incl. ray point and grid transform


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_time(L, name):
    ms, n = C.c_double(0), C.c_int(0)
    L.gxy_timing_get(name.encode(), C.byref(ms), C.byref(n))
    return ms.value, n.value


def frame_fn(L, built, cam, fb, eps=0.001):
    from galaxy_amd import abi, lib as glib
    w, h = cam.width, cam.height
    a3 = glib.a3
    eye, cen, up = a3(cam.eye), a3(cam.center), a3(cam.up)
    nb = np.full(6, -1, np.int32)
    st = abi.FrameStats()

    def frame():
        glib.check(L.gxy_memset(fb.ptr, 0, fb.nbytes), "memset")
        glib.check(L.gxy_render_frame(C.byref(built.vis), C.byref(built.lights), eye.ctypes.data, cen.ctypes.data,
                                      up.ctypes.data, cam.aov, w, h, nb.ctypes.data, eps, 1 << 30, fb.ptr,
                                      C.byref(st)), "render_frame")
        return st
    return frame


def run_frames(L, hip, built, cam, steps, warmup, eps=0.001):
    from galaxy_amd import abi, lib as glib
    w, h = cam.width, cam.height
    fb = glib.DeviceArray(w * h * 16)
    a3 = glib.a3
    eye, cen, up = a3(cam.eye), a3(cam.center), a3(cam.up)
    nb = np.full(6, -1, np.int32)
    st = abi.FrameStats()

    def frame():
        glib.check(L.gxy_memset(fb.ptr, 0, fb.nbytes), "memset")
        glib.check(L.gxy_render_frame(C.byref(built.vis), C.byref(built.lights), eye.ctypes.data, cen.ctypes.data,
                                      up.ctypes.data, cam.aov, w, h, nb.ctypes.data, eps, 1 << 30, fb.ptr,
                                      C.byref(st)), "render_frame")

    for _ in range(warmup):
        frame()
    glib.check(L.gxy_synchronize())
    L.gxy_timing_reset()
    # the timed frames record HIP events around the trace launches only (mode 2)
    L.gxy_timing_enable(2)
    totals = dict(ray_passes=0, samples=0, primary=0, secondary=0, trace_launches=0)
    glib.check(L.gxy_synchronize())
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
        totals["ray_passes"] += st.ray_passes
        totals["samples"] += st.samples
        totals["primary"] += st.primary_rays
        totals["secondary"] += st.secondary_rays
    glib.check(L.gxy_synchronize())
    t1 = time.perf_counter()
    L.gxy_timing_enable(0)
    kt = {"trace": kernel_time(L, "trace")}
    # per-kernel breakdown from a few more frames with every launch timed
    nb_frames = min(steps, 5)
    L.gxy_timing_reset()
    L.gxy_timing_enable(1)
    for _ in range(nb_frames):
        frame()
    glib.check(L.gxy_synchronize())
    L.gxy_timing_enable(0)
    breakdown = {k: round(kernel_time(L, k)[0] / nb_frames, 4)
                 for k in ("trace", "lighting", "route", "partition", "spawn_scan", "camera")}
    image = fb.to_host(np.float32)
    fb.free()
    return (t1 - t0), totals, kt, st, image, breakdown


def cpu_baseline(hip, sc, cam, alpha, target_s=15.0, ao=(0, 0)):
    """The CPU oracle (C++ restatement, std::thread on the host cores) on a
    bounded sample of the same frame: primary rays of every k-th image row,
    traced, lit, spawned, classified and accumulated like the GPU frame (AO
    directions: ao = (mode, seed) as the GPU frame's)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_backend import OracleBackend, lib as olib
    from galaxy_amd import abi, scenes
    threads = min(16, os.cpu_count() or 1)
    olib().oracle_set_threads(threads)
    olib().oracle_set_opacity_threshold(alpha)
    olib().oracle_set_ao_directions(*ao)
    ob = OracleBackend()
    vol = sc.volume_vis[0].volume if sc.volume_vis else None
    if vol is not None and vol.voxels is None:
        # the same voxels the GPU rendered (generated on the device, copied here)
        hv = hip.volume_create(vol.dims, vol.origin, vol.spacing, None, vol.sampling_rate)
        hip.volume_fill_noise(hv, *vol.noise)
        vox = hip.volume_download(hv, vol.dims)
        hip.L.gxy_volume_destroy(hv)
        vol.voxels = vox
    Bo = scenes.build_scene(ob, sc, scenes.resolve_lights(sc.lighting, cam))
    rays = ob.camera_rays(cam, sc.local_box, sc.global_box)
    nb = np.full(6, -1, np.int32)

    def run(stride):
        idx = np.nonzero(rays["y"] % stride == 0)[0]
        sub = rays.subset(idx)
        fb = np.zeros(cam.width * cam.height * 4, np.float32)
        st = abi.FrameStats()
        t = time.perf_counter()
        ob.L.oracle_process_raylist(C.byref(Bo.vis), C.byref(Bo.lights), sub.n, C.byref(sub.ispc), nb.ctypes.data,
                                    sc.epsilon, 1 << 30, fb.ctypes.data, cam.width, cam.height, C.byref(st))
        return time.perf_counter() - t, st, sub.n

    dt, st, n = run(64)
    stride = 64
    if dt > 0:
        want = int(64 * dt / target_s)
        stride = max(1, min(64, want if want > 0 else 1))
    reps, passes, samples = 1, 0, 0
    if stride != 64:
        dt, st, n = run(stride)
    passes, samples = st.ray_passes, st.samples
    if stride == 1 and 0 < dt < target_s / 2:
        # the whole frame takes less than the target: render it again until the target is reached
        reps = max(1, int(round(target_s / dt)))
        for _ in range(reps - 1):
            d2, s2, _ = run(1)
            dt += d2
            passes += s2.ray_passes
            samples += s2.samples
    if vol is not None:
        vol.voxels = None
    olib().oracle_set_ao_directions(0, 0)
    return dict(value=round(passes / dt / 1e6, 4), unit="Mrays/s", cores=threads, kind="port",
                gsamples_per_s=round(samples / dt / 1e9, 5), host_cores=os.cpu_count(), cpu_model=cpu_model(),
                sample="%d x primary rays of every %d-th image row (%d primary, %d ray passes in all, %.1f s), "
                       "same scene and seeds, host CPU %s" % (reps, stride, n, passes, dt,
                                                               platform.processor() or "x86_64"))


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's machine type."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.machine()


WALK_COUNTS = os.path.join(ROOT, "tests", "golden", "walk_counts.json")


def geometry_roofline(name, sc, tot, steps, trace_ms, trace_launches):
    """Roofline of a geometry frame's trace launches: algorithmic bytes (104 B
    per ray pass, the scene once per launch: 12 B per vertex and per
    triangle) and algorithmic flops (the recorded BVH work per primary,
    shadow and AO ray of the oracle's binned-SAH tree, 20 flops per child-box
    test and 51 per triangle test) over the measured trace time."""
    with open(WALK_COUNTS) as f:
        wc = json.load(f).get(name)
    if wc is None or trace_launches == 0 or trace_ms <= 0:
        return None
    per = {k: (v["box_tests"] * 20.0 + v["prim_tests"] * 51.0) / v["rays"] for k, v in wc.items() if k != "stride"}
    L_ = sc.lighting
    nao, nsh = int(L_.n_ao), (len(L_.types) if L_.shadows else 0)
    prim, sec = tot["primary"], tot["secondary"]
    ao = sec * nao // max(1, nao + nsh)
    sh = sec - ao
    flops = prim * per["primary"] + ao * per.get("ao", 0.0) + sh * per.get("shadow", 0.0)
    scene_bytes = sum(12 * len(m.vertices) + 12 * len(m.indices) for m in sc.meshes)
    launches = trace_launches
    nbytes = 104.0 * tot["ray_passes"] + scene_bytes * launches
    t = trace_ms * 1e-3
    gbs, tfl = nbytes / t / 1e9, flops / t / 1e12
    if gbs / HBM_PEAK_GBS >= tfl / FP32_PEAK_TFLOPS:
        roof = dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 5))
    else:
        roof = dict(bound="mfma", achieved=round(tfl, 4), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(tfl / FP32_PEAK_TFLOPS, 5), note="FP32 VALU work (no MFMA on this path); MI355X FP32 peak")
    prof = {"nine_icospheres_4k_ao16": "nine_icospheres_counters.md"}.get(name)
    roof.update(traffic=traffic_from_profile(prof) if prof else None,
                hbm_frac=round(gbs / HBM_PEAK_GBS, 5), fp32_frac=round(tfl / FP32_PEAK_TFLOPS, 5),
                kernel_ms_per_launch=round(trace_ms / launches, 4), alg_bytes_per_launch=int(nbytes / launches),
                alg_flops_per_launch=int(flops / launches),
                flops_per_ray={k: round(v, 1) for k, v in per.items()},
                counts="tests/golden/walk_counts.json (oracle binned-SAH tree, leaves <= 4, nearest-first walk)")
    return roof


def release(L, built):
    """Frees the device objects of a built scene (each destroy call checks the
    handle's kind and ignores handles of other kinds)."""
    for h in getattr(built, "handles", []):
        for fn in (L.gxy_volume_destroy, L.gxy_tf_destroy, L.gxy_model_destroy):
            fn(h)
    built.handles = []


def bvh_depth(L, built):
    if not built.vis.model:
        return 0
    nodes, leaves, dp, sah = C.c_int(), C.c_int(), C.c_int(), C.c_float()
    L.gxy_model_bvh_stats(built.vis.model, C.byref(nodes), C.byref(leaves), C.byref(dp), C.byref(sah))
    return dp.value


def bench_batched(L, hip, steps, tune=True, fixed=None):
    """256 ray lists (1k-1M rays, log-uniform) over 64 resident datasets (blob
    meshes of 50k-5M triangles, 128^3-1024^3 noise volumes), shadow + 2 AO rays
    per surface hit: the batched form (one persistent launch per pass, passes
    captured in a HIP graph) against the per-list loop over the same lists."""
    from galaxy_amd import abi, batch as gb, lib as glib, scenes, workloads
    descs, lists, Ld = workloads.batched_lists()
    log("building the batched lists: %d datasets, %d lists, %d rays" % (len(descs), len(lists), sum(n for n, _, _ in lists)))
    t = time.perf_counter()
    built = [scenes.build_scene(hip, workloads.dataset_scene(d)) for d in descs]
    glib.check(L.gxy_synchronize())
    build_s = time.perf_counter() - t
    B = gb.Batch(256)
    ids = [B.add_dataset(b) for b in built]
    keep = []
    lights = scenes.make_lighting(Ld, keep)
    B.set_lighting(lights)
    rays = []
    for k, (n, ds, seed) in enumerate(lists):
        d = glib.DeviceRays(n)
        l, fb = B.add_render_list(ids[ds], d, n, 1)
        B.set_seed(l, seed)
        rays.append((l, d, fb, n, ds, seed))
    depth = max(bvh_depth(L, b) for b in built)
    res = dict(lists=len(lists), datasets=len(descs), primary_rays=sum(n for n, _, _ in lists),
               triangles=int(sum(x[1] for x in descs if x[0] == "mesh")),
               volume_voxels=int(sum(x[1] ** 3 for x in descs if x[0] == "volume")), max_bvh_depth=depth,
               build_s=round(build_s, 2))
    vols = [b.handles[0] for b, d in zip(built, descs) if d[0] == "volume"]
    if fixed is not None:
        B.set_config(*fixed[:4])
        for v in vols:
            glib.check(L.gxy_volume_set_layout(v, fixed[4]))
        B.refresh_datasets(built)
    elif tune:
        # voxel layout of every volume dataset first (linear, 4^3 or 8^3 bricks; datasets
        # re-registered) under the default launch configuration, then the launch
        # configuration under that layout
        lay_ms = {}
        for lay in (0, 4, 8):
            for v in vols:
                glib.check(L.gxy_volume_set_layout(v, lay))
            B.refresh_datasets(built)
            B.run(2)
            glib.check(L.gxy_synchronize())
            t = time.perf_counter()
            B.run(2)
            glib.check(L.gxy_synchronize())
            lay_ms[lay] = (time.perf_counter() - t) * 1e3
        layout = min(lay_ms, key=lay_ms.get)
        for v in vols:
            glib.check(L.gxy_volume_set_layout(v, layout))
        B.refresh_datasets(built)
        best, table = gb.tune(B, max_depth=0, graph_passes=2)   # a stack too short for a model is skipped
        res["tuned"] = dict(best=list(best), ms=[[list(c), round(m, 2)] for c, m in table])
        res["tuned"]["volume_layout"] = layout
        res["tuned"]["layout_ms"] = {str(k): round(v, 2) for k, v in lay_ms.items()}
        log("batched lists tuned", res["tuned"])
    st = B.run(2)                      # warm (graph captured)
    glib.check(L.gxy_synchronize())
    t = time.perf_counter()
    tot = dict(rays=0, samples=0)
    for _ in range(steps):
        st = B.run(2)
        tot["rays"] += st.rays
        tot["samples"] += st.samples
    glib.check(L.gxy_synchronize())
    dt = time.perf_counter() - t
    res.update(mrays_per_s=round(tot["rays"] / dt / 1e6, 2), gsamples_per_s=round(tot["samples"] / dt / 1e9, 3),
               volume_layout=L.gxy_volume_get_layout(vols[0]) if vols else 0,
               ms_per_batch=round(dt / steps * 1e3, 3), ray_passes_per_batch=int(st.rays),
               spawned_per_batch=int(st.spawned), passes=int(st.passes), steps=steps,
               config=list(B.config()), pool_bytes=int(B.nbytes()))
    batched_fb = [fb.to_host(np.float32) for (_, _, fb, _, _, _) in rays[:16]]
    # the per-list loop over the same lists, one after another
    nb = np.full(6, -1, np.int32)
    fs = abi.FrameStats()
    prays = psamples = 0
    lo, hi = glib.a3((0, 0, 0)), glib.a3((1, 1, 1))
    glib.check(L.gxy_synchronize())
    t = time.perf_counter()
    for (l, d, fb, n, ds, seed) in rays:
        glib.check(L.gxy_rays_random(C.byref(d.ispc), n, seed, n, 1, lo.ctypes.data, hi.ctypes.data))
        glib.check(L.gxy_memset(fb.ptr, 0, fb.nbytes))
        glib.check(L.gxy_process_raylist(C.byref(built[ds].vis), C.byref(lights), n, C.byref(d.ispc),
                                          nb.ctypes.data, 0.001, 1 << 30, fb.ptr, n, 1, C.byref(fs)))
        prays += fs.ray_passes
        psamples += fs.samples
    glib.check(L.gxy_synchronize())
    pdt = time.perf_counter() - t
    diff = max(float(np.abs(fb.to_host(np.float32) - a).max()) for (_, _, fb, _, _, _), a in zip(rays, batched_fb))
    res["per_list_loop"] = dict(mrays_per_s=round(prays / pdt / 1e6, 2), ms=round(pdt * 1e3, 2),
                                ray_passes=int(prays), same_ray_passes=bool(prays == st.rays),
                                same_samples=bool(psamples == st.samples), max_abs_fb_diff_first16=diff)
    res["speedup_vs_per_list_loop"] = round(res["mrays_per_s"] / max(1e-9, res["per_list_loop"]["mrays_per_s"]), 3)
    B.destroy()
    for b in built:
        release(L, b)
    return res


def bench_served(L, hip, nreq, seed=5, layout=0, config=None, step_rays=4_000_000, tune=True):
    """Poisson stream (200 requests/s) of render requests (256^2 to 3840x2160)
    and sampler requests (1M-50M rays) over resident 512^3-1024^3 noise volumes
    and 1M-20M triangle blob meshes, admitted under an HBM budget and served
    with continuous batching; the same stream served one request at a time
    through the per-list loop beside it."""
    from galaxy_amd import abi, lib as glib, serve, workloads
    descs = [("volume", 512, 11), ("volume", 768, 12), ("volume", 1024, 13),
             ("mesh", 1_000_000, 21), ("mesh", 5_000_000, 22), ("mesh", 20_000_000, 23)]
    log("building c5 datasets")
    t = time.perf_counter()
    S, built = serve.build_server(hip, [workloads.dataset_scene(d) for d in descs], budget=160 << 30, slots=256,
                                  layout=layout)
    glib.check(L.gxy_synchronize())
    build_s = time.perf_counter() - t
    rng = np.random.default_rng(seed)
    reqs, _ = serve.request_stream(nreq, 200.0, 0.5, rng, volume_ds=(0, 1, 2), any_ds=(0, 1, 2, 3, 4, 5))
    # launch configuration for this mix: the batched form's choice against the split kernels and smaller
    # chunks, timed on a burst of the stream's first 12 requests arriving at once, served in steps
    # continuous batching steps take at most step_rays primary rays of admitted requests, fewest left first
    S.set_step_rays(step_rays)
    cands = ([tuple(config)] if config is not None else []) + [(256, 0, 0, 4, 1), (256, 0, 5, 4, 1),
                                                               (128, 0, 5, 4, 1), (64, 0, 5, 4, 1),
                                                               (64, 0, 5, 8, 1), (64, 0, 6, 8, 1),
                                                               (256, 0, 6, 8, 1)]
    if not tune and config is not None:
        cands = [tuple(config)]
    burst = (abi.Request * 12)()
    for k in range(12):
        burst[k] = reqs[k]
        burst[k].arrival = 0.0
    tune_ms = []
    for c in cands:
        S.set_config(*c)
        _, stb = S.run(burst, abi.SERVE_CONTINUOUS)
        tune_ms.append((c, stb.wall * 1e3))
    best = min(tune_ms, key=lambda x: x[1])[0]
    S.set_config(*best)
    # isovalue crossings a sampler ray follows per pass (the rest re-queued): shorter tails, more passes
    seg_ms = {}
    for nseg in (4096, 16, 4) if tune else ():
        S.set_max_segments(nseg)
        _, stb = S.run(burst, abi.SERVE_CONTINUOUS)
        seg_ms[nseg] = stb.wall * 1e3
    max_segments = min(seg_ms, key=seg_ms.get) if seg_ms else 4096
    S.set_max_segments(max_segments)
    nrender = sum(1 for r in reqs if r.kind == 0)
    res = dict(requests=nreq, render=nrender, sample=nreq - nrender, arrival_rate=200.0,
               arrival_span_s=round(reqs[nreq - 1].arrival, 3), budget_bytes=160 << 30, build_s=round(build_s, 2),
               datasets=["%s %d" % (k, s) for k, s, _ in descs], volume_layout=layout, config=list(best),
               config_ms=[[list(c), round(m, 1)] for c, m in tune_ms], step_rays=step_rays,
               max_segments=max_segments, max_segments_ms={str(k): round(v, 1) for k, v in seg_ms.items()})
    # warm-up: one request of each kind
    w = (abi.Request * 2)()
    w[0] = serve.render_request(0, 256, 256, (0.5, 0.5, 2.4))
    w[1] = serve.sample_request(0, 100000, 1, 200000)
    S.run(w, abi.SERVE_CONTINUOUS)
    for mode, name in ((abi.SERVE_CONTINUOUS, "continuous"), (abi.SERVE_ONE_AT_A_TIME, "one_at_a_time")):
        log("serving", name)
        r, st = S.run(reqs, mode)
        d = serve.latency_summary(reqs, r)
        d.update(mrays_per_s=round(st.rays / st.wall / 1e6, 2), gsamples_per_s=round(st.samples / st.wall / 1e9, 3),
                 wall_s=round(st.wall, 3), ray_passes=int(st.rays), passes=int(st.passes),
                 max_in_flight=int(st.max_in_flight), peak_bytes=int(st.peak_bytes))
        res[name] = d
    res["throughput_ratio"] = round(res["continuous"]["mrays_per_s"] / max(1e-9, res["one_at_a_time"]["mrays_per_s"]),
                                    3)
    # the same mix at 5 requests/s, which one GPU keeps up with: latency without queueing behind a backlog
    rng = np.random.default_rng(seed + 1)
    reqs, _ = serve.request_stream(24, 5.0, 0.5, rng, volume_ds=(0, 1, 2), any_ds=(0, 1, 2, 3, 4, 5))
    light = dict(requests=24, arrival_rate=5.0, arrival_span_s=round(reqs[23].arrival, 3))
    for mode, name, sr in ((abi.SERVE_CONTINUOUS, "continuous", step_rays),
                           (abi.SERVE_CONTINUOUS, "continuous_whole_requests", 0),
                           (abi.SERVE_ONE_AT_A_TIME, "one_at_a_time", step_rays)):
        log("serving at 5/s", name)
        S.set_step_rays(sr)
        r, st = S.run(reqs, mode)
        d = serve.latency_summary(reqs, r)
        d.update(mrays_per_s=round(st.rays / st.wall / 1e6, 2), wall_s=round(st.wall, 3),
                 ray_passes=int(st.rays), max_in_flight=int(st.max_in_flight))
        light[name] = d
    res["at_5_per_s"] = light
    S.destroy()
    for b in built:
        release(L, b)
    return res


def valu_from_profile(name):
    """(VALU instructions per wave, waves per launch) of the trace kernel from
    the SQ counter line of the committed profile summary."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    import re
    for line in open(p):
        m = re.search(r"VALU instructions per wave ([0-9.]+), waves per launch ([0-9.]+)", line)
        if m:
            return float(m.group(1)), float(m.group(2))
    return None


def traffic_from_profile(name):
    """HBM bytes per trace launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, copy-kernel calibrated) and committed under profiles/."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch"):
            try:
                return float(line.split("=")[1].split()[0])
            except (IndexError, ValueError):
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--icosphere-steps", type=int, default=3)
    ap.add_argument("--extra-only", default=None,
                    help="run only this mesh-frame workload (sphere_mesh_512, nine_icospheres_4k_ao16 or "
                         "pathlines_4000x63_segments_1024)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--serve-step-rays", type=int, default=4_000_000,
                    help="primary rays per continuous-batching step of the served form (0 = whole requests)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default trace launch configuration")
    ap.add_argument("--trace-config", default=None,
                    help="kind,chunk,refill: use this trace launch configuration (implies --no-tune)")
    ap.add_argument("--volume-layout", type=int, default=None,
                    help="voxel layout of the volume: 0 linear, 8 or 4 bricks (implies --no-tune)")
    ap.add_argument("--geometry-config", default=None,
                    help="BVH width, traversal stack, waves per SIMD of the mesh frames (skips their tuner), e.g. 2,0,8")
    ap.add_argument("--geometry-stack", type=int, default=None,
                    help="BVH traversal stack of the mesh frames: 0 private, 16/24/32 LDS (no tuning)")
    ap.add_argument("--batch-config", default=None,
                    help="chunk,stack,wpe,blocks_per_cu,layout of the batched and served forms (no tuning)")
    ap.add_argument("--serve-config", default=None,
                    help="chunk,stack,wpe,blocks_per_cu,split of the served form (no tuning); its voxel layout is "
                         "the last field of --batch-config")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: one partition traced on one MI355X; "
              "--gpus %d is not supported (use --gpus 1)" % args.gpus, flush=True)
        sys.exit(2)

    from galaxy_amd import lib as glib, scenes, workloads
    L = glib.lib()
    hip = glib.HipBackend()

    if args.copy_calibration:
        nbytes = 256 << 20
        a, b = glib.DeviceArray(nbytes).zero(), glib.DeviceArray(nbytes)
        glib.check(L.gxy_copy_device(b.ptr, a.ptr, nbytes))
        glib.check(L.gxy_synchronize())
        a.free()
        b.free()

    # ---------------------------------------------------------------- c2
    sc, alpha = workloads.volume_render()
    cam = sc.cameras[0]
    glib.check(L.gxy_set_opacity_threshold(alpha))
    log("building the volume render (512^3 noise volume)")
    built = scenes.build_scene(hip, sc)
    tuned = None
    if args.trace_config:
        glib.check(L.gxy_set_trace_config(*[int(x) for x in args.trace_config.split(",")]), "set_trace_config")
    if args.volume_layout is not None:
        glib.check(L.gxy_volume_set_layout(built.handles[0], args.volume_layout), "volume_set_layout")
    elif not args.no_tune and args.volume_layout is None:
        from galaxy_amd import autotune
        fbt = glib.DeviceArray(cam.width * cam.height * 16)
        lay, best, table = autotune.tune_volume_layout(L, [built.handles[0]], frame_fn(L, built, cam, fbt,
                                                                                        sc.epsilon))
        fbt.free()
        tuned = dict(best=list(best), volume_layout=lay,
                     trace_ms=[[l_, list(c), round(t, 4)] for l_, c, t in table])
        log("autotune", tuned)
    dt, tot, kt, st, image, breakdown = run_frames(L, hip, built, cam, args.steps, args.warmup, sc.epsilon)
    ms_step = dt / args.steps * 1e3
    mrays = tot["ray_passes"] / dt / 1e6
    gsamp = tot["samples"] / dt / 1e9
    tr_ms, tr_n = kt["trace"]
    # algorithmic work of the trace launches: 104 B per ray pass + the volume once per launch
    voxels = int(np.prod(np.array(sc.volume_vis[0].volume.dims, dtype=np.int64)))
    alg_bytes = BYTES_PER_RAY_PASS * tot["ray_passes"] + 4 * voxels * tr_n
    alg_flops = FLOPS_PER_SAMPLE * tot["samples"]
    gbs = alg_bytes / (tr_ms * 1e-3) / 1e9 if tr_ms > 0 else 0.0
    tfl = alg_flops / (tr_ms * 1e-3) / 1e12 if tr_ms > 0 else 0.0
    if gbs / HBM_PEAK_GBS >= tfl / FP32_PEAK_TFLOPS:
        roof = dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(gbs / HBM_PEAK_GBS, 5))
    else:
        roof = dict(bound="mfma", achieved=round(tfl, 4), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(tfl / FP32_PEAK_TFLOPS, 5),
                    note="FP32 VALU work (no MFMA on this path); MI355X FP32 peak")
    roof["traffic"] = traffic_from_profile("trace_volume_render_traffic.txt")
    # the bound the march actually meets: VALU issue.  A wave64 VALU instruction
    # occupies a 16-lane SIMD for 4 cycles; 1024 SIMDs at 2.4 GHz
    vp = valu_from_profile("trace_volume_render_traffic.txt")
    if vp and tr_n:
        ms = tr_ms / tr_n
        roof["valu_issue"] = dict(instr_per_launch=int(vp[0] * vp[1]),
                                  frac=round(vp[0] * vp[1] * 4 / (1024 * 2.4e9 * ms * 1e-3), 4),
                                  note="SQ_INSTS_VALU/SQ_WAVES from the committed profile x this run's launch time")
    kcfg = (C.c_int(), C.c_int(), C.c_int())
    L.gxy_get_trace_config(*[C.byref(x) for x in kcfg])
    lay = L.gxy_volume_get_layout(built.handles[0])
    sm = lay if lay else 1
    # the variant the launcher picks: one volume, bricked (32-bit offsets under 4 GiB), one isovalue
    one = bool(lay) and len(sc.volume_vis[0].isovalues) <= 1
    k = kcfg[0].value
    w = k - 10 if k >= 10 else 0
    if w > 6 and not one:
        w = 6                               # 17 / 18 act as 16 beyond one bricked volume and isovalue
    bs = kcfg[1].value if (one and w == 6 and kcfg[1].value in (64, 128)) else 256
    roof["kernel"] = ("k_trace_pl<1>" if k == 1 else
                      "k_trace<1,false,%d,%d,0,0,%d,%d>" % (w, sm, 1 if one else 8, bs))
    roof["volume_layout"] = lay
    roof["kernel_ms_per_launch"] = round(tr_ms / max(1, tr_n), 4)
    roof["alg_bytes_per_launch"] = int(alg_bytes / max(1, tr_n))
    roof["alg_flops_per_launch"] = int(alg_flops / max(1, tr_n))

    out = {
        "metric": "Mrays/s + volume Gsamples/s on 1 MI355X; p50/p99 request latency",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 4-octave gradient noise generated on the device)",
        "config": {"workload": "volume render: 1920x1080 rays, 512^3 FP32 noise, 8-point RGBA TF, step 0.5 voxel, "
                               "early termination alpha 0.99, iso 0.5 + shadow rays",
                   "rays_per_frame": tot["ray_passes"] // args.steps,
                   "primary_per_frame": tot["primary"] // args.steps,
                   "shadow_per_frame": tot["secondary"] // args.steps},
        "gsamples_per_s": round(gsamp, 3),
        "roofline": roof,
        "kernel_ms_per_step": breakdown,
        "trace_config": tuned,
    }
    built = None

    # ---------------------------------------------------------------- extras
    if not args.no_extra:
        extra = {}
        saved_width = L.gxy_get_bvh_width()
        for name, fn, steps in (("sphere_mesh_512", workloads.sphere_mesh, args.steps),
                                ("nine_icospheres_4k_ao16", workloads.nine_icospheres, args.icosphere_steps),
                                ("pathlines_4000x63_segments_1024", workloads.pathlines, args.steps)):
            if args.extra_only and name != args.extra_only:
                continue
            try:
                sc2, a2 = fn()
                glib.check(L.gxy_set_opacity_threshold(a2))
                # the nine-icosphere frame's AO directions: seeded PCG32 streams
                ao = (1, workloads.NINE_ICOSPHERES_AO_SEED) if name == "nine_icospheres_4k_ao16" else (0, 0)
                glib.check(L.gxy_set_ao_directions(*ao))
                log("building", name)
                t = time.perf_counter()
                b2 = scenes.build_scene(hip, sc2)
                build_s = time.perf_counter() - t
                gtab = []
                if args.geometry_config:
                    w_, s_, v_ = (int(x) for x in args.geometry_config.split(","))
                    glib.check(L.gxy_set_bvh_width(w_))
                    glib.check(L.gxy_set_geometry_stack(s_))
                    glib.check(L.gxy_set_geometry_waves(v_))
                elif args.geometry_stack is not None:
                    glib.check(L.gxy_set_geometry_stack(args.geometry_stack))
                elif not args.no_tune:
                    from galaxy_amd import autotune
                    fb2 = glib.DeviceArray(sc2.cameras[0].width * sc2.cameras[0].height * 16)
                    _, gtab = autotune.tune_geometry_stack(L, frame_fn(L, b2, sc2.cameras[0], fb2, sc2.epsilon),
                                                           reps=4)
                    fb2.free()
                dt2, tot2, kt2, st2, _, bd2 = run_frames(L, hip, b2, sc2.cameras[0], steps, 1, sc2.epsilon)
                extra[name] = dict(mrays_per_s=round(tot2["ray_passes"] / dt2 / 1e6, 2),
                                   ms_per_frame=round(dt2 / steps * 1e3, 3),
                                   rays_per_frame=tot2["ray_passes"] // steps,
                                   primary=tot2["primary"] // steps, secondary=tot2["secondary"] // steps,
                                   scene_build_s=round(build_s, 3),
                                   trace_ms_per_frame=round(kt2["trace"][0] / steps, 3),
                                   kernel_ms_per_frame=bd2,
                                   geometry_stack=L.gxy_get_geometry_stack(), bvh_width=L.gxy_get_bvh_width(),
                                   geometry_waves=L.gxy_get_geometry_waves(),
                                   geometry_tune=[[list(c), round(ms, 4)] for c, ms in gtab])
                if sc2.lighting is not None and sc2.lighting.n_ao:
                    extra[name]["ao_directions"] = ("PCG32 streams, seed %d" % ao[1]) if ao[0] else "UV table"
                roof2 = geometry_roofline(name, sc2, tot2, steps, kt2["trace"][0], kt2["trace"][1])
                if roof2:
                    extra[name]["roofline"] = roof2
                if not args.no_cpu and name in ("sphere_mesh_512", "nine_icospheres_4k_ao16"):
                    log("cpu baseline", name)
                    extra[name]["cpu_baseline"] = cpu_baseline(hip, sc2, sc2.cameras[0], a2,
                                                               target_s=10.0 if name.startswith("sphere") else 15.0,
                                                               ao=ao)
                    glib.check(L.gxy_set_opacity_threshold(alpha))
                release(L, b2)
                b2 = None
            except Exception as e:  # reported, not hidden
                extra[name] = dict(error=str(e))
            glib.check(L.gxy_set_ao_directions(0, 0))
        out["extra_workloads"] = extra
        # the batched and served forms start from the library's BVH width, not the last frame's tuned one
        glib.check(L.gxy_set_bvh_width(saved_width))
        glib.check(L.gxy_set_geometry_waves(0))
        glib.check(L.gxy_set_opacity_threshold(alpha))

    # ---------------------------------------------------------------- batched and served forms
    if not args.no_batched:
        try:
            fixed = [int(x) for x in args.batch_config.split(",")] if args.batch_config else None
            out["batched"] = bench_batched(L, hip, args.batch_steps, tune=not args.no_tune, fixed=fixed)
        except Exception as e:
            out["batched"] = dict(error=str(e))
    if not args.no_served:
        try:
            lay = out.get("batched", {}).get("volume_layout", 0)
            cfg = out.get("batched", {}).get("config")
            if args.batch_config:
                lay = int(args.batch_config.split(",")[4])
            if args.serve_config:
                cfg = [int(x) for x in args.serve_config.split(",")]
            out["served"] = bench_served(L, hip, args.serve_requests, layout=lay, config=cfg,
                                         step_rays=args.serve_step_rays, tune=not args.serve_config)
        except Exception as e:
            out["served"] = dict(error=str(e))

    # ---------------------------------------------------------------- CPU baseline
    if not args.no_cpu:
        try:
            log("cpu baseline")
            sc, alpha = workloads.volume_render()
            out["cpu_baseline"] = cpu_baseline(hip, sc, sc.cameras[0], alpha)
        except Exception as e:
            out["cpu_baseline"] = dict(error=str(e))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
