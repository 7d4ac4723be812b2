"""Path-line scenes and FP64 references shared by the oracle and GPU tests:
round-Bezier control points from the reference's construction rules
(DataDrivenPathLines.cpp:97-157) checked through their defining properties,
the closed-form capsule a straight constant-radius path line must be, the
signed distance to the round-cone pieces every hit must lie on, and the
distance to the true swept-sphere surface of the cubic (the union of the
spheres (p(u), r(u)), 0 <= u <= 1) that bounds the tessellation."""
import numpy as np

from galaxy_amd import scenes

f32 = np.float32
CURVE_TOL = 0.01    # traced solid within CURVE_TOL x the least control radius of the swept-sphere solid
MAX_PIECES = 32


def piece_counts(P):
    """Round cones per segment from control points P (nseg, 4, 4): the least
    n with (1/n)^2 / 8 * 6 (max|second difference of xyz| + max|second
    difference of r|) <= CURVE_TOL * least control radius (the interpolation
    bound of a cubic), 1 for a straight constant-radius segment."""
    P = np.asarray(P, np.float64)
    sd = np.stack([P[:, 0] - 2.0 * P[:, 1] + P[:, 2], P[:, 1] - 2.0 * P[:, 2] + P[:, 3]], 1)   # (nseg, 2, 4)
    ln = np.sqrt(sd[..., 0] * sd[..., 0] + sd[..., 1] * sd[..., 1] + sd[..., 2] * sd[..., 2]).max(1)
    bound = 6.0 * (ln + np.abs(sd[..., 3]).max(1))
    rlo = P[:, :, 3].min(1)
    with np.errstate(divide="ignore", invalid="ignore"):
        n = np.ceil(np.sqrt(bound / (8.0 * CURVE_TOL * rlo)))
    n = np.where(bound > 0, np.where(rlo > 0, np.clip(n, 1, MAX_PIECES), MAX_PIECES), 1)
    return n.astype(np.int64)


def helix_lines(seed=3, nlines=6, nverts=12):
    """Polylines (concatenated): vertices, per-vertex data in [0, 1], and the
    segment index list (one entry per segment; lines are not joined)."""
    rng = np.random.default_rng(seed)
    V, D, I = [], [], []
    for k in range(nlines):
        c = rng.uniform(-0.6, 0.6, 3)
        ph = rng.uniform(0, 2 * np.pi)
        s = np.linspace(0, 1, nverts)
        pts = np.c_[c[0] + 0.35 * np.cos(ph + 5 * s), c[1] + 0.35 * np.sin(ph + 5 * s), c[2] - 0.6 + 1.2 * s]
        pts += rng.normal(0, 0.02, pts.shape)
        base = len(V)
        V += list(pts)
        D += list(rng.uniform(0, 1, nverts))
        I += [base + j for j in range(nverts - 1)]
    return np.array(V, f32), np.array(D, f32), np.array(I, np.int32)


def helix_desc(seed=3, **kw):
    V, D, I = helix_lines(seed, **kw)
    return scenes.PathLinesDesc(V, D, I, radius0=0.02, radius1=0.06, value0=0.0, value1=1.0,
                                colormap=np.array([[0, 0.1, 0.3, 0.9], [1, 0.9, 0.4, 0.1]], f32))


def controls(oracle, model, which, nseg):
    cp = np.zeros((4 * nseg, 4), f32)
    ic = np.zeros(nseg, np.int32)
    n = oracle.L.oracle_pathline_controls(model, which, cp.ctypes.data, len(cp), ic.ctypes.data, nseg)
    assert n == 4 * nseg
    return cp, ic


def segment_controls(cp, ic):
    """Control points of every segment, (nseg, 4, 4) in FP64."""
    return cp.astype(np.float64)[ic[:, None] + np.arange(4)[None, :]]


def pieces64(cp, ic):
    """Round-cone pieces (a, ra, b, rb) of every segment in FP64 from the
    control points: Bernstein samples at u = j / n, n = piece_counts.
    Returns (a, b) of shape (nseg, MAX_PIECES, 4); a segment with fewer
    pieces repeats its last one."""
    P = segment_controls(cp, ic)
    n = piece_counts(P.astype(f32))
    j = np.minimum(np.arange(MAX_PIECES + 1)[None, :], n[:, None])            # (nseg, MAX_PIECES+1)
    u = j / n[:, None]
    q = 1 - u
    W = np.stack([q ** 3, 3 * q * q * u, 3 * q * u * u, u ** 3], 2)          # (nseg, MAX_PIECES+1, 4)
    X = np.einsum("sjk,skc->sjc", W, P)                                       # (nseg, MAX_PIECES+1, 4)
    a, b = X[:, :-1].copy(), X[:, 1:].copy()
    last = (np.arange(MAX_PIECES)[None, :] >= n[:, None])
    idx = np.minimum(np.arange(MAX_PIECES)[None, :], n[:, None] - 1)
    a[last] = np.take_along_axis(X, idx[..., None], 1)[last]
    b[last] = np.take_along_axis(X, (idx + 1)[..., None], 1)[last]
    return a, b


def _bezier(P, u):
    """Points (x y z r) of cubics P (n, 4, 4) at parameters u (n, k)."""
    q = 1 - u
    W = np.stack([q ** 3, 3 * q * q * u, 3 * q * u * u, u ** 3], -1)
    return np.einsum("nkj,njc->nkc", W, P)


def sweep_distance(x, P, grid=1025, iters=60):
    """FP64 signed distance from points x (n, 3) to the swept-sphere solid of
    each point's own cubic P (n, 4, 4): min over u of |x - p(u)| - r(u), by a
    dense grid in u refined by ternary search around its best sample.  Outside
    the solid this is the distance to it; just inside, minus the depth to first
    order in that depth over the radius."""
    x = np.asarray(x, np.float64)
    u = np.broadcast_to(np.linspace(0, 1, grid), (len(x), grid))
    X = _bezier(P, u)
    d = np.linalg.norm(x[:, None] - X[..., :3], axis=2) - X[..., 3]
    k = d.argmin(1)
    lo = np.clip((k - 1) / (grid - 1), 0, 1)
    hi = np.clip((k + 1) / (grid - 1), 0, 1)

    def f(uu):
        p = _bezier(P, uu[:, None])[:, 0]
        return np.linalg.norm(x - p[:, :3], axis=1) - p[:, 3]
    for _ in range(iters):
        m1 = lo + (hi - lo) / 3
        m2 = hi - (hi - lo) / 3
        left = f(m1) < f(m2)
        hi = np.where(left, m2, hi)
        lo = np.where(left, lo, m1)
    return np.minimum(f(0.5 * (lo + hi)), d.min(1))


def curve_hit_error(x, seg, P, I):
    """Distance of hit points x on segments seg to the true swept surface of
    their path line near the hit (the hit segment and its neighbours on the
    same polyline), and the tolerance CURVE_TOL x (least control radius of the
    hit segment) the tessellation guarantees."""
    best = np.full(len(x), np.inf)
    for off in (-1, 0, 1):
        s = np.clip(seg + off, 0, len(I) - 1)
        same = (I[s] - I[seg]) == off
        best = np.where(same, np.minimum(best, sweep_distance(x, P[s])), best)
    return np.abs(best), CURVE_TOL * P[seg, :, 3].min(1)


def swept_distance(x, a, b, grid=2049):
    """Signed distance (negative inside) from points x (n, 3) to the union of
    round cones a[i] -> b[i] of each point's own candidate set: a, b of shape
    (n, m, 4).  The union of the spheres along a piece is the piece's convex
    hull, so min over s of |x - c(s)| - r(s) is the distance outside."""
    s = np.linspace(0, 1, grid)
    best = np.full(len(x), np.inf)
    for j in range(a.shape[1]):
        A, B = a[:, j], b[:, j]
        c = A[:, None, :3] + s[None, :, None] * (B[:, None, :3] - A[:, None, :3])
        r = A[:, None, 3] + s[None, :] * (B[:, None, 3] - A[:, None, 3])
        d = np.linalg.norm(x[:, None, :] - c, axis=2) - r
        best = np.minimum(best, d.min(1))
    return best


def capsule_hits(org, dirs, A, B, r):
    """FP64 closest entry (> 1e-6) of rays into the capsule A-B of radius r;
    inf where missed."""
    org = org.astype(np.float64)
    d = dirs.astype(np.float64)
    A, B = np.asarray(A, np.float64), np.asarray(B, np.float64)
    ax = B - A
    L = np.linalg.norm(ax)
    w = ax / L
    best = np.full(len(org), np.inf)

    def take(t, ok):
        t = np.where(ok & (t > 1e-6), t, np.inf)
        return np.minimum(best, t)

    # cylinder
    oa = org - A
    dp = d - (d @ w)[:, None] * w
    op = oa - (oa @ w)[:, None] * w
    qa = np.einsum("ij,ij->i", dp, dp)
    qb = 2 * np.einsum("ij,ij->i", dp, op)
    qc = np.einsum("ij,ij->i", op, op) - r * r
    disc = qb * qb - 4 * qa * qc
    ok = (disc >= 0) & (qa > 0)
    sq = np.sqrt(np.where(ok, disc, 0))
    for sgn in (-1, 1):
        t = (-qb + sgn * sq) / np.where(qa > 0, 2 * qa, 1)
        y = ((oa + t[:, None] * d) @ w)
        best = take(t, ok & (y >= 0) & (y <= L))
    # caps
    for c in (A, B):
        oc = org - c
        bb = np.einsum("ij,ij->i", oc, d)
        cc = np.einsum("ij,ij->i", oc, oc) - r * r
        dd = np.einsum("ij,ij->i", d, d)
        disc = bb * bb - dd * cc
        ok = disc >= 0
        sq = np.sqrt(np.where(ok, disc, 0))
        for sgn in (-1, 1):
            best = take((-bb + sgn * sq) / dd, ok)
    return best


def rays_toward(n, seed, target_lo, target_hi, dist=2.5):
    """Rays from a sphere of radius `dist` aimed at points uniform in a box."""
    rng = np.random.default_rng(seed)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    tgt = rng.uniform(target_lo, target_hi, (n, 3))
    org = tgt - dist * dirs
    return org.astype(f32), dirs.astype(f32)


def oracle_intersect(oracle, M, org, dirs, t0=0.0):
    n = len(org)
    org = np.ascontiguousarray(org, f32)
    dirs = np.ascontiguousarray(dirs, f32)
    t0 = np.full(n, t0, f32)
    tf = np.full(n, 1e30, f32)
    t = np.zeros(n, f32)
    g = np.zeros(n, np.int32)
    p = np.zeros(n, np.int32)
    u = np.zeros(n, f32)
    v = np.zeros(n, f32)
    oracle.L.oracle_model_intersect(M, n, org.ctypes.data, dirs.ctypes.data, t0.ctypes.data, tf.ctypes.data,
                                    t.ctypes.data, g.ctypes.data, p.ctypes.data, u.ctypes.data, v.ctypes.data)
    return t, g, p
