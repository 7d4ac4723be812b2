"""Autotuner for the trace launch configuration.

Every configuration computes identical results (the per-ray arithmetic does
not change; only how rays are mapped onto lanes does), so the tuner only
measures: it renders the caller's frame once per candidate with HIP-event
timing of the trace kernel enabled and keeps the fastest.

Candidates: one ray per work-item (kind 0, and its 5- and 6-waves-per-SIMD
occupancy variants 15 and 16), or persistent lanes (kind 1) with a per-wave
chunk of 128 or 256 consecutive rays and a refill threshold of 16 or 32 idle
lanes out of 64.  (Measured on the volume-render frame: the persistent form
was 2-5x slower than one ray per work-item; it stays a candidate for other
ray mixes.  The 7- and 8-waves-per-SIMD variants 17 and 18, which apply to
one bricked volume with at most one isovalue, are selectable but not
candidates: over two runs on that frame 17 timed 0.948-0.969 and 18
0.974-1.001 ms of trace per frame against 0.949-0.971 for 16, so they would
only make the pick noisy.  The same holds for kind 16 in 64-item workgroups
(chunk 64, for one bricked volume with at most one isovalue): 3.5 % faster
than 256-item workgroups before the march loop lost its scalar overhead,
0.879 against 0.873 ms of trace per frame after.)
"""
import ctypes as C

DEFAULT_CANDIDATES = ([(0, 256, 32), (15, 256, 32), (16, 256, 32)] +
                      [(1, chunk, refill) for chunk in (128, 256) for refill in (16, 32)])
# BVH traversal stacks of geometry-only traces: private memory (64 entries) or
# 16 / 24 / 32 entries in LDS (a model deeper than the stack keeps the private one)
GEOMETRY_STACKS = [0, 16, 24, 32]


def _trace_ms(L):
    ms, n = C.c_double(0), C.c_int(0)
    L.gxy_timing_get(b"trace", C.byref(ms), C.byref(n))
    return ms.value


def tune_trace(L, render_once, candidates=None, reps=2):
    """render_once(): renders one frame with the current configuration.
    Returns (best (kind, chunk, refill), [(config, trace ms per frame), ...])
    and leaves the best configuration set."""
    from . import lib as glib
    candidates = candidates or DEFAULT_CANDIDATES
    table = []
    for cfg in candidates:
        glib.check(L.gxy_set_trace_config(*cfg), "set_trace_config")
        render_once()                      # warm the configuration
        glib.check(L.gxy_synchronize())
        L.gxy_timing_reset()
        L.gxy_timing_enable(1)
        for _ in range(reps):
            render_once()
        glib.check(L.gxy_synchronize())
        L.gxy_timing_enable(0)
        table.append((cfg, _trace_ms(L) / reps))
    best = min(table, key=lambda x: x[1])[0]
    glib.check(L.gxy_set_trace_config(*best), "set_trace_config")
    L.gxy_timing_reset()
    return best, table


def geometry_candidates(stacks=None):
    """(BVH width, traversal stack entries, waves per SIMD) of the geometry
    tuner: both widths over the stacks at the compiler's occupancy, plus the
    binary private-stack walk at 8 waves per SIMD."""
    return [(w, s, 0) for w in (4, 2) for s in (stacks or GEOMETRY_STACKS)] + [(2, 0, 8)]


def tune_geometry_stack(L, render_once, stacks=None, reps=2):
    """The same measurement over the BVH width (binary or 4-wide), the
    geometry traversal stacks and the binary walk's occupancy.  Returns (best
    (width, entries, waves), [((width, entries, waves), trace ms per frame),
    ...]) and leaves the best set."""
    from . import lib as glib
    table = []
    for wst in geometry_candidates(stacks):
        glib.check(L.gxy_set_bvh_width(wst[0]), "set_bvh_width")
        glib.check(L.gxy_set_geometry_stack(wst[1]), "set_geometry_stack")
        glib.check(L.gxy_set_geometry_waves(wst[2]), "set_geometry_waves")
        st = wst
        render_once()
        glib.check(L.gxy_synchronize())
        L.gxy_timing_reset()
        L.gxy_timing_enable(1)
        for _ in range(reps):
            render_once()
        glib.check(L.gxy_synchronize())
        L.gxy_timing_enable(0)
        table.append((st, _trace_ms(L) / reps))
    best = min(table, key=lambda x: x[1])[0]
    glib.check(L.gxy_set_bvh_width(best[0]), "set_bvh_width")
    glib.check(L.gxy_set_geometry_stack(best[1]), "set_geometry_stack")
    glib.check(L.gxy_set_geometry_waves(best[2]), "set_geometry_waves")
    L.gxy_timing_reset()
    return best, table


VOLUME_LAYOUTS = [0, 8, 4]


def _refine(L, volumes, render_once, table, reps, within=0.03, rounds=4):
    """The pairs of the first pass within `within` of its fastest, timed again
    in turn over `rounds` rounds of 2 * reps frames each; returns the pair
    with the least total (a single pass cannot tell apart variants that
    differ by less than its frame-to-frame noise)."""
    from . import lib as glib
    fastest = min(ms for _, _, ms in table)
    close = [(lay, cfg) for lay, cfg, ms in table if ms <= fastest * (1 + within)]
    if len(close) == 1:
        return close[0]
    total = [0.0] * len(close)
    for _ in range(rounds):
        for k, (lay, cfg) in enumerate(close):
            for v in volumes:
                glib.check(L.gxy_volume_set_layout(v, lay), "volume_set_layout")
            glib.check(L.gxy_set_trace_config(*cfg), "set_trace_config")
            render_once()
            glib.check(L.gxy_synchronize())
            L.gxy_timing_reset()
            L.gxy_timing_enable(1)
            for _ in range(2 * reps):
                render_once()
            glib.check(L.gxy_synchronize())
            L.gxy_timing_enable(0)
            total[k] += _trace_ms(L)
    return close[min(range(len(close)), key=total.__getitem__)]


def tune_volume_layout(L, volumes, render_once, candidates=None, reps=2):
    """Joint search over the voxel layout of `volumes` (linear, 8^3 or 4^3
    bricks) and the trace kernel configuration.  Returns (best layout, best
    config, [(layout, config, ms), ...]) and leaves both set."""
    from . import lib as glib
    table = []
    for lay in VOLUME_LAYOUTS:
        for v in volumes:
            glib.check(L.gxy_volume_set_layout(v, lay), "volume_set_layout")
        best, t = tune_trace(L, render_once, candidates, reps)
        table += [(lay, cfg, ms) for cfg, ms in t]
    lay, cfg = _refine(L, volumes, render_once, table, reps)
    for v in volumes:
        glib.check(L.gxy_volume_set_layout(v, lay), "volume_set_layout")
    glib.check(L.gxy_set_trace_config(*cfg), "set_trace_config")
    L.gxy_timing_reset()
    return lay, cfg, table
