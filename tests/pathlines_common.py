"""Path-line scenes and FP64 references shared by the oracle and GPU tests:
round-Bezier control points from the reference's construction rules
(DataDrivenPathLines.cpp:97-157) checked through their defining properties,
the closed-form capsule a straight constant-radius path line must be, and the
signed distance to the swept-sphere surface that every hit must lie on."""
import numpy as np

from galaxy_amd import scenes

f32 = np.float32
PIECES = 8          # round cones per curve segment in both implementations


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


def pieces64(cp, ic):
    """Round-cone pieces (a, ra, b, rb) of every segment in FP64 from the
    control points: Bernstein samples at u = j / PIECES.  Returns (a, b) of
    shape (nseg, PIECES, 4)."""
    P = cp.astype(np.float64)[ic[:, None] + np.arange(4)[None, :]]          # (nseg, 4, 4)
    u = np.arange(PIECES + 1) / PIECES
    q = 1 - u
    W = np.stack([q ** 3, 3 * q * q * u, 3 * q * u * u, u ** 3], 1)          # (PIECES+1, 4)
    X = np.einsum("jk,skc->sjc", W, P)                                        # (nseg, PIECES+1, 4)
    return X[:, :-1], X[:, 1:]


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
