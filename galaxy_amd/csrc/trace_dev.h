// trace_dev.h — per-ray device code of TraceRays_TraceRays shared by the
// one-ray-per-work-item kernels (trace.hip) and the batched persistent kernel
// (batch.hip): ray/triangle and ray/sphere tests, BVH traversal with a
// pluggable per-lane stack, geometry shading and the full per-ray trace.
#pragma once

#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

struct Hit {
  float sample, t;
  f3 point, normal, color;
  float opacity;
};

struct GeoHit {
  float t, u, v;
  int geom, prim;
};

// ------------------------------------------------------------- geometry
struct WRay {
  f3 org, dir;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

__device__ __forceinline__ float comp3(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

__device__ __forceinline__ WRay make_wray(f3 o, f3 d) {
  WRay r;
  r.org = o;
  r.dir = d;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = 0;
  float m = ax;
  if (ay > m) { kz = 1; m = ay; }
  if (az > m) { kz = 2; }
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (comp3(d, kz) < 0.f) { int t = kx; kx = ky; ky = t; }
  r.kx = kx; r.ky = ky; r.kz = kz;
  float dz = comp3(d, kz);
  r.Sx = comp3(d, kx) / dz;
  r.Sy = comp3(d, ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013); accepts t0 < t <= tfar.
__device__ __forceinline__ bool tri_hit(const WRay &r, f3 p0, f3 p1, f3 p2, float t0, float tfar, float &t,
                                        float &u, float &v) {
  f3 A = sub3(p0, r.org), B = sub3(p1, r.org), Cc = sub3(p2, r.org);
  float Az_ = comp3(A, r.kz), Bz_ = comp3(B, r.kz), Cz_ = comp3(Cc, r.kz);
  float Ax = comp3(A, r.kx) - r.Sx * Az_, Ay = comp3(A, r.ky) - r.Sy * Az_;
  float Bx = comp3(B, r.kx) - r.Sx * Bz_, By = comp3(B, r.ky) - r.Sy * Bz_;
  float Cx = comp3(Cc, r.kx) - r.Sx * Cz_, Cy = comp3(Cc, r.ky) - r.Sy * Cz_;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    U = (float)(CxBy - CyBx);
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    V = (float)(AxCy - AyCx);
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  float det = U + V + W;
  if (det == 0.f) return false;
  float Az = r.Sz * Az_, Bz = r.Sz * Bz_, Cz = r.Sz * Cz_;
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  float tt = T * rcp;
  if (!(tt > t0 && tt <= tfar)) return false;
  t = tt;
  u = V * rcp;
  v = W * rcp;
  return true;
}

// DataDrivenSpheres_intersect_kernel (DataDrivenSpheres.ispc:118-141).
__device__ __forceinline__ bool sphere_hit(f3 center, float radius, float eps, f3 org, f3 dir, float t0,
                                           float tfar, float &t) {
  float approxDist = dot3(sub3(center, org), dir);
  f3 closeOrg = add3(org, scl3(approxDist, dir));
  f3 A = sub3(center, closeOrg);
  float a = dot3(dir, dir);
  float b = 2.f * dot3(dir, A);
  float c = dot3(A, A) - radius * radius;
  float radical = b * b - 4.f * a * c;
  if (radical < 0.f) return false;
  float srad = sqrtf(radical);
  float r2a = 1.0f / (2.f * a);
  float t_in = (b - srad) * r2a + approxDist;
  float t_out = (b + srad) * r2a + approxDist;
  if (t_in > t0 && t_in < tfar) { t = t_in; return true; }
  if (t_out > (t0 + eps) && t_out < tfar) { t = t_out; return true; }
  return false;
}

// Path-line pieces.  A piece is the round cone of spheres (a, ra) and
// (b, rb): their convex hull, which is the surface a sphere sweeps when its
// centre and radius move linearly from one to the other.  Its boundary is the
// cap of sphere a behind the tangent circle at axial coordinate ra*sin(phi),
// the cone band up to the tangent circle on sphere b at l + rb*sin(phi), and
// the cap of sphere b beyond it (l = |b - a|, sin(phi) = (ra - rb) / l).  On
// the band, with y the axial coordinate from a and h the distance from the
// axis, h*cos(phi) + y*sin(phi) = ra, which along the ray is one quadratic.

// The two crossings of sphere (c, r), solved about the ray point nearest c.
__device__ __forceinline__ bool sphere_roots(f3 c, float r, f3 org, f3 dir, float dd, float &t1, float &t2) {
  const float tc = dot3(sub3(c, org), dir) / dd;
  const f3 q = sub3(add3(org, scl3(tc, dir)), c);
  const float disc = dd * (r * r - dot3(q, q));
  if (disc < 0.f) return false;
  const float h = sqrtf(disc) / dd;
  t1 = tc - h;
  t2 = tc + h;
  return true;
}

// Entry and exit of the ray through the round cone (a, ra, b, rb): the least
// and greatest boundary crossings.  t = the entry when it lies in (t0, tfar),
// else the exit when that does (the rule of the reference's cylinder kernel,
// DataDrivenPathLines.ispc:183-193).
__device__ __forceinline__ bool roundcone_hit(f3 a, float ra, f3 b, float rb, f3 org, f3 dir, float t0, float tfar,
                                              float &t) {
  const float dd = dot3(dir, dir);
  const f3 ba = sub3(b, a);
  const float l2 = dot3(ba, ba), rd = ra - rb;
  float tin = 3e38f, tout = -3e38f;
  float t1, t2;
  if (l2 <= rd * rd) {
    // one sphere holds the other: the larger one
    if (sphere_roots(ra >= rb ? a : b, ra >= rb ? ra : rb, org, dir, dd, t1, t2)) { tin = t1; tout = t2; }
  } else {
    const float l = sqrtf(l2);
    const f3 w = scl3(1.0f / l, ba);
    const float sn = rd / l, cs2 = 1.0f - sn * sn;
    const float ya = ra * sn, yb = l + rb * sn;
    if (sphere_roots(a, ra, org, dir, dd, t1, t2)) {
      if (dot3(sub3(add3(org, scl3(t1, dir)), a), w) <= ya) { tin = mnf(tin, t1); tout = mxf(tout, t1); }
      if (dot3(sub3(add3(org, scl3(t2, dir)), a), w) <= ya) { tin = mnf(tin, t2); tout = mxf(tout, t2); }
    }
    if (sphere_roots(b, rb, org, dir, dd, t1, t2)) {
      if (dot3(sub3(add3(org, scl3(t1, dir)), a), w) >= yb) { tin = mnf(tin, t1); tout = mxf(tout, t1); }
      if (dot3(sub3(add3(org, scl3(t2, dir)), a), w) >= yb) { tin = mnf(tin, t2); tout = mxf(tout, t2); }
    }
    // the band, solved about the ray point nearest a
    const float tc = dot3(sub3(a, org), dir) / dd;
    const f3 oa = sub3(add3(org, scl3(tc, dir)), a);
    const float y0 = dot3(oa, w), yd = dot3(dir, w);
    const float k = ra - sn * y0;
    const float A = cs2 * (dd - yd * yd) - sn * sn * (yd * yd);
    const float B = 2.0f * (cs2 * (dot3(oa, dir) - y0 * yd) + sn * yd * k);
    const float Cc = cs2 * (dot3(oa, oa) - y0 * y0) - k * k;
    const float disc = B * B - 4.0f * A * Cc;
    if (A != 0.f && disc >= 0.f) {
      const float sq = sqrtf(disc), i2a = 0.5f / A;
      const float r1 = (-B - sq) * i2a, r2 = (-B + sq) * i2a;
      const float y1 = y0 + r1 * yd, y2 = y0 + r2 * yd;
      if (y1 > ya && y1 < yb && k - sn * (r1 * yd) >= 0.f) { tin = mnf(tin, tc + r1); tout = mxf(tout, tc + r1); }
      if (y2 > ya && y2 < yb && k - sn * (r2 * yd) >= 0.f) { tin = mnf(tin, tc + r2); tout = mxf(tout, tc + r2); }
    }
  }
  if (tin > tout) return false;
  if (tin > t0 && tin < tfar) { t = tin; return true; }
  if (tout > t0 && tout < tfar) { t = tout; return true; }
  return false;
}

__device__ __forceinline__ bool better(float t, int g, int p, const GeoHit &h) {
  if (t < h.t) return true;
  if (t == h.t && (g < h.geom || (g == h.geom && p < h.prim))) return true;
  return false;
}

#define GXY_BVH_STACK 64
#define GXY_SENTINEL 0x7fffffff

// Per-lane traversal stacks.  PrivStack keeps the entries in private memory
// (scratch, cached per lane); LdsStack keeps them in LDS with entry k of a
// lane at base[k * stride] (stride = lanes of the workgroup, so lanes at the
// same depth touch consecutive words).  DEPTH entries; a model whose tree is
// deeper than DEPTH - 2 must not be traced with it (checked on the host).
template <int SD>
struct PrivStack {
  static constexpr int DEPTH = SD;
  int s[SD];
  __device__ __forceinline__ void put(int k, int v) { s[k] = v; }
  __device__ __forceinline__ int get(int k) const { return s[k]; }
};

template <int SD>
struct LdsStack {
  static constexpr int DEPTH = SD;
  int *base;
  int stride;
  __device__ __forceinline__ void put(int k, int v) { base[k * stride] = v; }
  __device__ __forceinline__ int get(int k) const { return base[k * stride]; }
};

// Slab test of one child box against [t0, tmax] with a few ulps of slack so
// that a primitive hit at exactly the current t still reaches its leaf.
__device__ __forceinline__ bool child_hit(const float *lo, const float *hi, f3 o, f3 inv, float t0, float tmax,
                                          float &tenter) {
  float ax = (lo[0] - o.x) * inv.x, bx = (hi[0] - o.x) * inv.x;
  float ay = (lo[1] - o.y) * inv.y, by = (hi[1] - o.y) * inv.y;
  float az = (lo[2] - o.z) * inv.z, bz = (hi[2] - o.z) * inv.z;
  float nx = inv.x < 0.f ? bx : ax, fx = inv.x < 0.f ? ax : bx;
  float ny = inv.y < 0.f ? by : ay, fy = inv.y < 0.f ? ay : by;
  float nz = inv.z < 0.f ? bz : az, fz = inv.z < 0.f ? az : bz;
  float tn = mxf(mxf(nx, ny), mxf(nz, t0));
  float tf = mnf(mnf(fx, fy), mnf(fz, tmax));
  tenter = tn;
  return tn <= tf * 1.0000005f + 1e-30f;
}

// any: the walk ends at the first leaf that records a hit (occlusion query).
template <class Stack, class LeafFn>
__device__ __forceinline__ void bvh_walk(const BVHNode *nodes, f3 o, f3 inv, float t0, const GeoHit &h,
                                         Stack &stack, LeafFn &&leaf, bool any = false) {
  int sp = 0;
  stack.put(sp++, GXY_SENTINEL);
  int cur = 0;
  while (cur != GXY_SENTINEL) {
    if (cur >= 0) {
      const float4 *np = (const float4 *)(gptr(nodes) + cur);
      float4 q0 = np[0], q1 = np[1], q2 = np[2];
      int4 q3 = ((const int4 *)np)[3];
      float lo0[3] = {q0.x, q0.y, q0.z}, hi0[3] = {q0.w, q1.x, q1.y};
      float lo1[3] = {q1.z, q1.w, q2.x}, hi1[3] = {q2.y, q2.z, q2.w};
      float e0, e1;
      bool h0 = child_hit(lo0, hi0, o, inv, t0, h.t, e0);
      bool h1 = child_hit(lo1, hi1, o, inv, t0, h.t, e1);
      if (h0 && h1) {
        int nearc = e0 <= e1 ? q3.x : q3.y;
        int farc = e0 <= e1 ? q3.y : q3.x;
        if (sp < Stack::DEPTH) stack.put(sp++, farc);
        cur = nearc;
      } else if (h0) {
        cur = q3.x;
      } else if (h1) {
        cur = q3.y;
      } else {
        cur = stack.get(--sp);
      }
    } else {
      int code = ~cur;
      leaf(code >> 4, (code & 15) + 1);
      if (any && h.geom >= 0) return;
      cur = stack.get(--sp);
    }
  }
}

// Nearest-first walk of a 4-wide BVH.  At a node, the children's boxes are
// tested against [t0, current hit t]; leaf children that are hit are
// intersected at once (the LBVH leaves hold one primitive), then the walk
// descends into the nearest inner child that is still hit and pushes one
// entry (node << 4 | mask of the other hit inner children), re-tested against
// the then-current hit t when popped, so the stack holds at most one entry per
// level (depth4 entries).
template <class Stack, class LeafFn>
__device__ __forceinline__ void bvh4_walk(const BVH4Node *nodes, f3 o, f3 inv, float t0, const GeoHit &h,
                                          Stack &stack, LeafFn &&leaf, bool any, int dlim) {
  int sp = 0;
  int cur = 0, mask = 0xF;
  const BVH4Node *gn = gptr(nodes);
  for (;;) {
    const float4 *np = (const float4 *)(gn + cur);
    const float4 lx = np[0], ly = np[1], lz = np[2], hx = np[3], hy = np[4], hz = np[5];
    const int4 ch = ((const int4 *)np)[6];
    const float lox[4] = {lx.x, lx.y, lx.z, lx.w}, loy[4] = {ly.x, ly.y, ly.z, ly.w}, loz[4] = {lz.x, lz.y, lz.z, lz.w};
    const float hix[4] = {hx.x, hx.y, hx.z, hx.w}, hiy[4] = {hy.x, hy.y, hy.z, hy.w}, hiz[4] = {hz.x, hz.y, hz.z, hz.w};
    const int cc[4] = {ch.x, ch.y, ch.z, ch.w};
    float te[4];
    int inner = 0, leaves = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      te[k] = 0.f;
      if ((mask >> k) & 1) {
        const float lo[3] = {lox[k], loy[k], loz[k]}, hi[3] = {hix[k], hiy[k], hiz[k]};
        if (child_hit(lo, hi, o, inv, t0, h.t, te[k])) {
          if (cc[k] >= 0) inner |= 1 << k;
          else leaves |= 1 << k;
        }
      }
    }
    // the hit leaf children in slot order, through one copy of the leaf code
    // (a wave runs it as often as its lanes' largest leaf count)
    while (leaves) {
      const int k = __builtin_ctz(leaves);
      leaves &= leaves - 1;
      int c = cc[0];
#pragma unroll
      for (int q = 1; q < 4; q++) c = (q == k) ? cc[q] : c;
      const int code = ~c;
      leaf(code >> 4, (code & 15) + 1);
      if (any && h.geom >= 0) return;
    }
    // the nearest inner child whose box still starts before the current hit
    int kn = -1;
    float tn = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (((inner >> k) & 1) && te[k] <= h.t * 1.0000005f + 1e-30f && (kn < 0 || te[k] < tn)) {
        kn = k;
        tn = te[k];
      }
    if (kn >= 0) {
      int rest = inner & ~(1 << kn);
      if (rest && sp <= dlim) {
        // the other children still hit, each its own entry, the nearest on top
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (((rest >> k) & 1) && !(te[k] <= h.t * 1.0000005f + 1e-30f)) rest &= ~(1 << k);
        while (rest) {
          int kf = __builtin_ctz(rest);
          float tf = te[0];
#pragma unroll
          for (int k = 1; k < 4; k++) tf = (k == kf) ? te[k] : tf;
#pragma unroll
          for (int k = 0; k < 4; k++)
            if (((rest >> k) & 1) && te[k] > tf) {
              kf = k;
              tf = te[k];
            }
          rest &= ~(1 << kf);
          int c2 = cc[0];
#pragma unroll
          for (int k = 1; k < 4; k++) c2 = (k == kf) ? cc[k] : c2;
          stack.put(sp++, c2 << 4);
        }
      } else if (rest && sp < Stack::DEPTH) {
        if ((rest & (rest - 1)) == 0) {
          // one other child: its own node, whole (mask 0)
          const int k2 = __builtin_ctz(rest);
          int c2 = cc[0];
#pragma unroll
          for (int k = 1; k < 4; k++) c2 = (k == k2) ? cc[k] : c2;
          stack.put(sp++, c2 << 4);
        } else {
          stack.put(sp++, (cur << 4) | rest);
        }
      }
      int c = cc[0];
#pragma unroll
      for (int k = 1; k < 4; k++) c = (k == kn) ? cc[k] : c;
      cur = c;
      mask = 0xF;
      continue;
    }
    if (sp == 0) return;
    const int e = stack.get(--sp);
    cur = e >> 4;
    mask = (e & 15) ? (e & 15) : 0xF;
  }
}

// W: 2 or 4, the tree width the walks use; 0, either (chosen per call from
// M.wide, both walks compiled into the caller); 16 + width: that width with
// the path-line pieces traced too.  Path lines are traced (and shaded) by the
// W = 0 and W > 16 variants only.
__host__ __device__ constexpr bool with_curves(int W) { return W == 0 || W > 16; }

template <class Stack, int W = 0>
// any: an occlusion query; the search ends at the first hit in (t0, tfar].
__device__ __forceinline__ void model_trace(const DModel &M, f3 org, f3 dir, float t0, float tfar, GeoHit &h,
                                            Stack &stack, bool any = false) {
  h.t = tfar;
  h.geom = -1;
  h.prim = -1;
  h.u = h.v = 0.f;
  if (tfar < t0) return;
  constexpr int TW = W & 15;
  f3 inv = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  // 4-wide walks: stack depth up to which other hit children get entries of their own
  const int dlim = Stack::DEPTH - M.depth4 - 3;
  if (M.ntris > 0) {
    WRay wr = make_wray(org, dir);
    auto tleaf = [&](int start, int count) {
      for (int k = start; k < start + count; k++) {
        const float *p = gptr(M.tverts) + 9 * (int64_t)k;
        f3 p0 = mk3(p[0], p[1], p[2]), p1 = mk3(p[3], p[4], p[5]), p2 = mk3(p[6], p[7], p[8]);
        float t, u, v;
        if (tri_hit(wr, p0, p1, p2, t0, h.t, t, u, v)) {
          int g = gptr(M.tgeom)[k], pr = gptr(M.tprim)[k];
          if (h.geom < 0 || better(t, g, pr, h)) { h.t = t; h.geom = g; h.prim = pr; h.u = u; h.v = v; }
        }
      }
    };
    if (TW == 4 || (TW == 0 && M.wide)) bvh4_walk(M.tnodes4, org, inv, t0, h, stack, tleaf, any, dlim);
    else bvh_walk(M.tnodes, org, inv, t0, h, stack, tleaf, any);
  }
  if (any && h.geom >= 0) return;
  if (M.nsph > 0) {
    auto sleaf = [&](int start, int count) {
      for (int k = start; k < start + count; k++) {
        const float *sc = gptr(M.scenter);
        f3 c = mk3(sc[3 * k], sc[3 * k + 1], sc[3 * k + 2]);
        float lim = h.geom < 0 ? h.t : nextafterf(h.t, 3e38f);
        float t;
        if (sphere_hit(c, gptr(M.sradius)[k], gptr(M.seps)[k], org, dir, t0, lim, t)) {
          int g = gptr(M.sgeom)[k], pr = gptr(M.sprim)[k];
          if (h.geom < 0 || better(t, g, pr, h)) { h.t = t; h.geom = g; h.prim = pr; h.u = h.v = 0.f; }
        }
      }
    };
    if (TW == 4 || (TW == 0 && M.wide)) bvh4_walk(M.snodes4, org, inv, t0, h, stack, sleaf, any, dlim);
    else bvh_walk(M.snodes, org, inv, t0, h, stack, sleaf, any);
  }
  if (any && h.geom >= 0) return;
  if (with_curves(W) && M.ncurve > 0) {
    // path-line pieces (models with path lines are dispatched to the variants
    // that trace them); h.u holds the hit piece's first curve parameter
    // (pieces of one segment that tie in t: the lower one wins, whatever the
    // traversal order), h.v the piece (its sorted index) for the shading
    auto cleaf = [&](int start, int count) {
      for (int k = start; k < start + count; k++) {
        const float *A = gptr(M.ca) + 4 * (int64_t)k, *B = gptr(M.cb) + 4 * (int64_t)k;
        float lim = h.geom < 0 ? h.t : nextafterf(h.t, 3e38f);
        float t;
        if (roundcone_hit(mk3(A[0], A[1], A[2]), A[3], mk3(B[0], B[1], B[2]), B[3], org, dir, t0, lim, t)) {
          const int g = gptr(M.cgeom)[k], pr = gptr(M.cprim)[k];
          const float u0 = gptr(M.cu)[2 * (int64_t)k];
          if (h.geom < 0 || better(t, g, pr, h) || (t == h.t && g == h.geom && pr == h.prim && u0 < h.u)) {
            h.t = t; h.geom = g; h.prim = pr; h.u = u0; h.v = __int_as_float(k);
          }
        }
      }
    };
    if (TW == 4 || (TW == 0 && M.wide)) bvh4_walk(M.cnodes4, org, inv, t0, h, stack, cleaf, any, dlim);
    else bvh_walk(M.cnodes, org, inv, t0, h, stack, cleaf, any);
  }
}

// The surface normal of round-cone piece k at P (unnormalised) and the curve
// parameter there: u0 or u1 on the caps, on the band the piece's (u0, u1) at
// the projection of P on its axis.
__device__ __forceinline__ f3 roundcone_normal(const DModel &M, int k, f3 P, float &u) {
  const float *A = gptr(M.ca) + 4 * (int64_t)k, *B = gptr(M.cb) + 4 * (int64_t)k, *U = gptr(M.cu) + 2 * (int64_t)k;
  const f3 a = mk3(A[0], A[1], A[2]), b = mk3(B[0], B[1], B[2]);
  const float ra = A[3], rb = B[3];
  const f3 ba = sub3(b, a), pa = sub3(P, a);
  const float l2 = dot3(ba, ba), rd = ra - rb;
  if (l2 <= rd * rd) {
    u = ra >= rb ? U[0] : U[1];
    return ra >= rb ? pa : sub3(P, b);
  }
  const float l = sqrtf(l2);
  const f3 w = scl3(1.0f / l, ba);
  const float sn = rd / l;
  const float y = dot3(pa, w);
  // caps: the end parameters; the band: the projection on the axis
  if (y <= ra * sn) { u = U[0]; return pa; }
  if (y >= l + rb * sn) { u = U[1]; return sub3(P, b); }
  float s = dot3(pa, ba) / l2;
  s = s < 0.f ? 0.f : (s > 1.f ? 1.f : s);
  u = s >= 1.f ? U[1] : U[0] + s * (U[1] - U[0]);
  const f3 q = sub3(pa, scl3(y, w));
  const float hq = sqrtf(dot3(q, q));
  if (!(hq > 0.f)) return w;
  return add3(scl3(sqrtf(1.0f - sn * sn) / hq, q), scl3(sn, w));
}

// postIntersect for the hit geometry, then DG_NORMALIZE | DG_FACEFORWARD.
// W: the walk variant of the kernel (model_trace); path lines are shaded by
// the variants that trace them only.
template <int W>
__device__ void geom_shade(const DModel &M, const GeoHit &gh, f3 org, f3 dir, Hit &hit) {
  const DGeomShade &S = gptr(M.shade)[gh.geom];
  f3 P = add3(org, scl3(gh.t, dir));
  f3 Ng = mk3(0.f, 0.f, 1.f), Ns = Ng, col = mk3(1.f, 1.f, 1.f);
  if (S.kind == 0) {
    const uint32_t *I = gptr(S.indices) + 3 * (int64_t)gh.prim;
    uint32_t i0 = I[0], i1 = I[1], i2 = I[2];
    const float *V = gptr(S.vertices);
    f3 p0 = mk3(V[3 * i0], V[3 * i0 + 1], V[3 * i0 + 2]);
    f3 p1 = mk3(V[3 * i1], V[3 * i1 + 1], V[3 * i1 + 2]);
    f3 p2 = mk3(V[3 * i2], V[3 * i2 + 1], V[3 * i2 + 2]);
    Ng = cross3(sub3(p0, p1), sub3(p2, p0));
    Ns = Ng;
    float b0 = 1.0f - gh.u - gh.v, b1 = gh.u, b2 = gh.v;
    if (S.has_normals) {
      const float *N = gptr(S.normals);
      f3 a = mk3(N[3 * i0], N[3 * i0 + 1], N[3 * i0 + 2]);
      f3 b = mk3(N[3 * i1], N[3 * i1 + 1], N[3 * i1 + 2]);
      f3 c = mk3(N[3 * i2], N[3 * i2 + 1], N[3 * i2 + 2]);
      Ns = add3(add3(scl3(b0, a), scl3(b1, b)), scl3(b2, c));
    }
    if (S.has_data && S.has_tf) {
      const float *vdat = gptr(S.vdata);
      float dv = b0 * vdat[i0] + b1 * vdat[i1] + b2 * vdat[i2];
      col = tf_color(gptr(S.tf.col), S.tf, dv);
    }
  } else if (S.kind == 1) {
    const float *cs = gptr(S.centers);
    f3 c = mk3(cs[3 * gh.prim], cs[3 * gh.prim + 1], cs[3 * gh.prim + 2]);
    Ng = sub3(P, c);
    Ns = Ng;
    if (S.has_tf && S.has_data) col = tf_color(gptr(S.tf.col), S.tf, gptr(S.vdata)[gh.prim]);
  } else if (with_curves(W)) {
    // DataDrivenPathLines_postIntersect (DataDrivenPathLines.ispc:212-269):
    // the radius interpolated between the segment's first two control points
    // at the curve parameter, mapped back to a data value for the TF
    float u;
    Ng = roundcone_normal(M, __float_as_int(gh.v), P, u);
    Ns = Ng;
    if (S.has_tf) {
      const uint32_t i0 = gptr(S.indices)[gh.prim];
      const float *cp = gptr(S.vertices);
      const float radius = ((1.f - u) * cp[4 * (int64_t)i0 + 3]) + (u * cp[4 * ((int64_t)i0 + 1) + 3]);
      const float r0 = S.r0, r1 = S.r1, v0 = S.v0, v1 = S.v1;
      float dv;
      if (r0 == r1) dv = v0;
      else if (r0 < r1 && radius < r0) dv = v0;
      else if (r0 < r1 && radius > r1) dv = v1;
      else if (r0 > r1 && radius < r1) dv = v1;
      else if (r0 > r1 && radius > r0) dv = v0;
      else dv = v0 + ((radius - r0) / (r1 - r0)) * (v1 - v0);
      col = tf_color(gptr(S.tf.col), S.tf, dv);
    }
  }
  Ng = normalize3(Ng);
  Ns = normalize3(Ns);
  if (dot3(dir, Ng) >= 0.f) Ng = neg3(Ng);
  if (dot3(dir, Ns) >= 0.f) Ns = neg3(Ns);
  hit.color = col;
  hit.opacity = 1.0f;
  hit.point = P;
  hit.normal = Ns;
  hit.t = gh.t;
}

// ------------------------------------------------------------- the trace kernel
// TraceRays.ispc:466,550,581 with the alpha threshold as a parameter
__device__ __forceinline__ bool opaque_of(float r, float g, float b, float a, float thr) {
  return mnf(mnf(r, g), b) >= 1.0f || a > thr;
}

// Front-to-back compositing of one weighted sample (TraceRays.ispc:512-533):
// C += (1 - C.a) * w * (rgb, 1), with (1 - C.a) * w formed once and one FMA
// per channel.
__device__ __forceinline__ void composite(float &cr, float &cg, float &cb, float &ca, f3 sc, float wo) {
  const float a = 1.0f - ca, aw = a * wo;
  cr = __builtin_fmaf(aw, sc.x, cr);
  cg = __builtin_fmaf(aw, sc.y, cg);
  cb = __builtin_fmaf(aw, sc.z, cb);
  ca = __builtin_fmaf(a, wo, ca);
}

// One ray of TraceRays_TraceRays (TraceRays.ispc:371-622): reads ray i of R,
// writes its colour, t, term and surface fields back.  s_col / s_op hold the
// NV transfer-function tables (LDS).  Returns the number of volume samples.
// MI: the isovalues per volume the iso test is unrolled for (the volume's
// nIso must not exceed it).
// SH: 1 when the caller has found every active lane's ray primary, 2 when it
// has found none primary (the choice is then known to the compiler), 0 per lane.
template <int NV, bool GEOM, class Stack, int SM = -1, int W = 0, int MI = GXY_MAX_ISO, int SH = 0>
__device__ __forceinline__ unsigned trace_one(const TraceParams &P, const RayList_ispc &R, const int i,
                                              const float (*s_col)[3 * GXY_TF_MAX],
                                              const float (*s_op)[GXY_TF_MAX], Stack &stack) {
  constexpr int NVS = NV > 0 ? NV : 1;
  unsigned nsamp = 0;
  {
    const bool shadeFlag = SH == 1 ? true : SH == 2 ? false : gptr(R.type)[i] == RAY_PRIMARY;
    f3 org = mk3(gptr(R.ox)[i], gptr(R.oy)[i], gptr(R.oz)[i]);
    f3 dir = mk3(gptr(R.dx)[i], gptr(R.dy)[i], gptr(R.dz)[i]);
    if (dir.x == 0.f) dir.x = 1e-6f;
    if (dir.y == 0.f) dir.y = 1e-6f;
    if (dir.z == 0.f) dir.z = 1e-6f;
    float ray_t0 = gptr(R.t)[i];
    float ray_t = gptr(R.tMax)[i];
    const float tTimeout = ray_t;
    float cr = gptr(R.r)[i], cg = gptr(R.g)[i], cb = gptr(R.b)[i], ca = gptr(R.o)[i];

    // box clip (TraceRays.ispc:406-418)
    float tEntry, tExitVolume;
    box_clip(org, dir, P.blo, P.bhi, tEntry, tExitVolume);
    if (tEntry < ray_t0) tEntry = ray_t0;
    else if (tEntry > ray_t0) ray_t0 = tEntry;
    ray_t = mnf(ray_t, tExitVolume);

    Hit hit;
    hit.sample = 0.f; hit.t = 0.f; hit.opacity = 0.f;
    hit.point = mk3(0.f, 0.f, 0.f); hit.normal = hit.point; hit.color = hit.point;

    // LookForSliceHit (TraceRays.ispc:143-250)
    bool surface_hit = false;
    int vid = -1;
#pragma unroll
    for (int major = 0; major < NV; major++) {
      const int ns = P.vv[major].nSlices;
      for (int minor = 0; minor < ns; minor++) {
        f3 pn = mk3(P.vv[major].slices[4 * minor], P.vv[major].slices[4 * minor + 1], P.vv[major].slices[4 * minor + 2]);
        float d = P.vv[major].slices[4 * minor + 3];
        float denom = dot3(dir, pn);
        if (fabsf(denom) > 0.0001f) {
          float t = (d - dot3(org, pn)) / denom;
          if (t >= ray_t0 && t <= ray_t) {
            hit.normal = denom > 0 ? neg3(pn) : pn;
            hit.opacity = 1.0f;
            hit.t = t;
            vid = major;
            ray_t = t;
            surface_hit = true;
          }
        }
      }
    }
    if (surface_hit && shadeFlag) {
      hit.point = add3(org, scl3(hit.t, dir));
#pragma unroll
      for (int major = 0; major < NV; major++)
        if (major == vid) {
          hit.sample = vol_sample_t<SM>(P.vv[major].vol, hit.point);
          hit.color = tf_color(s_col[major], P.vv[major].tf, hit.sample);
          hit.opacity = 1.0f;
        }
    }

    // LookForGeometryHit (TraceRays.ispc:108-141)
    if (GEOM) {
      GeoHit gh;
      // a shadow or AO ray in a scene without volumes needs only whether it
      // is blocked (any hit terminates it with its own colour) when the host
      // does not read its t (the rank loops; not TraceRays_TraceRays)
      const bool any = NV == 0 && P.occlusion && !shadeFlag;
      model_trace<Stack, W>(P.model, org, dir, ray_t0, ray_t, gh, stack, any);
      if (P.hit_geom) { P.hit_geom[i] = gh.geom; P.hit_prim[i] = gh.prim; }
      if (gh.geom >= 0) {
        ray_t = gh.t;
        surface_hit = true;
        if (shadeFlag) geom_shade<W>(P.model, gh, org, dir, hit);
      }
    }

    float tTermination = ray_t;

    if (NV > 0 && P.integrate) {
      // volume march (TraceRays.ispc:446-570)
      const float step = P.step, epsilon = P.eps;
      float sLast[NVS], sThis[NVS];
#pragma unroll
      for (int v = 0; v < NVS; v++) { sLast[v] = 0.f; sThis[v] = 0.f; }
      float tLast = tEntry + epsilon;
      bool hit_iso = false;
      int iso_vid = -1;
      const bool opaque0 = opaque_of(cr, cg, cb, ca, P.opaque_alpha);
      float tThis = tEntry;
      int iter = 0;
      // the gathers of the sample at tThis are issued one iteration ahead
      // (every fetch position is clamped into the volume, so a fetch that the
      // march does not use is harmless)
      // positions in each volume's voxel coordinates: one FMA per axis and sample
      VoxRay VR[NVS];
#pragma unroll
      for (int v = 0; v < NV; v++) VR[v] = vox_ray(P.vv[v].vol, org, dir);
      VFetch F[NVS];
#pragma unroll
      for (int v = 0; v < NV; v++) F[v] = vox_fetch<SM>(P.vv[v].vol, VR[v], tThis);
      // The first sample (at tEntry) ends no interval: it only becomes sLast,
      // and the next position is tEntry + epsilon.  Taken before the loop, it
      // leaves every pass of the loop with tThis > tEntry (t strictly grows,
      // or the march ends).
      bool go = tThis <= tTermination && !opaque0;
      iter = 1;
      // -inf held in a register the compiler cannot see into, so that the
      // three-way median below stays one v_med3_f32 (folded to a minimum it
      // would gain a NaN-quieting v_max_f32 per step)
      float ninf = -__builtin_inff();
      asm("" : "+s"(ninf));
      if (go) {
#pragma unroll
        for (int v = 0; v < NV; v++) sLast[v] = vol_finish_fma(F[v]);
        nsamp += NV;
        const float nx0 = tEntry + epsilon;
#pragma unroll
        for (int v = 0; v < NV; v++) F[v] = vox_fetch<SM>(P.vv[v].vol, VR[v], nx0);
        tLast = tThis;
        go = nx0 > tThis && nx0 <= tTermination;
        tThis = nx0;
      }
      // (every pass leaves tThis <= tTermination: the next position is cut
      // there, and a pass that moves tTermination ends the march; go is the
      // one exit condition, set at the end of each pass)
      // (unrolled by two: the loop's scalar bookkeeping is shared by two passes)
#pragma unroll 2
      while (go && iter++ < GXY_MAX_MARCH) {
        // (the ray is not opaque when a pass begins: only this pass's samples
        // can make it so)
        bool opaque = false;
#pragma unroll
        for (int v = 0; v < NV; v++) sThis[v] = vol_finish_fma(F[v]);
        nsamp += NV;
        // the next position when neither an iso hit nor opacity ends the march
        // here (the only cases in which the loop continues, and then
        // tTermination is still the value used here): tThis + step, cut at
        // tTermination (the median of the two and -inf is their minimum, in
        // one v_med3_f32).  (The reference steps past tTermination when tThis is
        // already there; the march ends then either way, with the same
        // samples, and the min ends it by the no-advance test below.  The
        // step is a number: the host rejects a NaN step.)
        const float nx0 = __builtin_amdgcn_fmed3f(tThis + step, tTermination, ninf);
#pragma unroll
        for (int v = 0; v < NV; v++) F[v] = vox_fetch<SM>(P.vv[v].vol, VR[v], nx0);
        {
          // The interval (tLast, tThis] counts only once tLast >= epsilon (a
          // ray entering before epsilon skips its first one).  Taken without
          // a branch: a skipped interval sees no isovalue crossing (sl = st)
          // and zero opacity, which leave the ray as it is.
          const bool live = tLast >= epsilon;
          // LookForIsoHit: the last matching isovalue in list order wins
          bool ih = false;
          int ivid = -1;
#pragma unroll
          for (int major = 0; major < NV; major++) {
            float sl = live ? sLast[major] : sThis[major], st = sThis[major];
            const int ni = P.vv[major].nIso;
            // unrolled to the fixed maximum: the isovalues stay in registers
#pragma unroll
            for (int m = 0; m < MI; m++) {
              if (m < ni) {
                float isoval = P.vv[major].iso[m];
                if (((isoval >= sl) && (isoval < st)) || ((isoval <= sl) && (isoval > st))) {
                  ih = true;
                  ivid = major;
                  hit.t = tLast + ((isoval - sl) / (st - sl)) * (tThis - tLast);
                  hit.sample = isoval;
                }
              }
            }
          }
          if (ih) {
            // the shading of the hit (gradient normal, TF colour) reads nothing
            // the rest of the march changes, so it is done once after the loop
            iso_vid = ivid;
            tTermination = hit.t;
            tThis = hit.t;
            hit_iso = true;
#pragma unroll
            for (int v = 0; v < NV; v++) sThis[v] = vox_sample<SM>(P.vv[v].vol, VR[v], tThis);
            nsamp += NV;
          }
#pragma unroll
          for (int major = 0; major < NV; major++) {
            if (P.vv[major].dvr) {
              const DTF &T = P.vv[major].tf;
              // the lookups of sVolume = (sLast + sThis) / 2 from the sum and
              // its offset (bit-identical to looking up sVolume)
              const float dVolume = (sLast[major] + sThis[major]) - T.lo2;
              float sop = tf_opacity_d2(s_op[major], T, dVolume);
              sop = live ? sop : 0.f;
              // a primary ray reads the colour beside the opacity, so that the
              // two LDS lookups wait once (the empty statement keeps the
              // colour reads ahead of the test on the opacity)
              f3 sc;
              if (shadeFlag) {
                sc = tf_color_num_d2(s_col[major], T, dVolume);
                asm volatile("" ::"v"(sc.x), "v"(sc.y), "v"(sc.z));
              }
              if (sop > 0) {
                if (shadeFlag) {
                  composite(cr, cg, cb, ca, sc, clamp01f(sop * P.vv[major].vol.irate));
                } else {
                  float wop = ((tThis - tLast) * P.istep) * clamp01f(sop * P.vv[major].vol.irate);
                  float f = 1.0f - wop;
                  cr *= f; cg *= f; cb *= f; ca *= f;
                }
                // only a sample that changes the colour can make the ray
                // opaque (it was not when this pass began)
                opaque = opaque_of(cr, cg, cb, ca, P.opaque_alpha);
              }
            }
          }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) sLast[v] = sThis[v];
        if (opaque) tTermination = tThis;
        tLast = tThis;
        // the march goes on unless opacity or an iso hit ended it here or the
        // next position no longer advances t (the reference would spin)
        go = !opaque && !hit_iso && nx0 > tThis;
        tThis = nx0;
      }
      ray_t = tTermination;
      if (hit_iso) {
        surface_hit = true;
        // LookForIsoHit shading (TraceRays.ispc:284-307)
        hit.point = add3(org, scl3(hit.t, dir));
        if (shadeFlag) {
#pragma unroll
          for (int major = 0; major < NV; major++)
            if (major == iso_vid) {
              hit.normal = safe_normalize3(vol_gradient(P.vv[major].vol, hit.point));
              if (dot3(dir, hit.normal) > 0) hit.normal = neg3(hit.normal);
              hit.color = tf_color(s_col[major], P.vv[major].tf, hit.sample);
              hit.opacity = 1.0f;
            }
        }
      }
    }

    gptr(R.r)[i] = cr; gptr(R.g)[i] = cg; gptr(R.b)[i] = cb; gptr(R.o)[i] = ca;
    int term = opaque_of(cr, cg, cb, ca, P.opaque_alpha) ? RAY_OPAQUE : 0;
    gptr(R.t)[i] = ray_t;
    if (surface_hit) {
      term |= RAY_SURFACE;
      if (hit.opacity > 0.999f) term |= RAY_OPAQUE;
      gptr(R.sr)[i] = hit.color.x; gptr(R.sg)[i] = hit.color.y; gptr(R.sb)[i] = hit.color.z; gptr(R.so)[i] = 1.0f;
      gptr(R.nx)[i] = hit.normal.x; gptr(R.ny)[i] = hit.normal.y; gptr(R.nz)[i] = hit.normal.z;
    } else if (tTermination == tExitVolume) {
      term |= RAY_BOUNDARY;
    } else if (tTermination == tTimeout) {
      term |= RAY_TIMEOUT;
    }
    gptr(R.term)[i] = term;
  }
  return nsamp;
}

// One ray of SamplerTraceRays_SamplerTraceRays (SamplerTraceRays.ispc:128-222)
// with the IsoSamplerVis interval test (IsoSamplerVis.ispc:36-64): marches to
// the next crossing of the first isovalue; t and term of ray i are written
// back.  Returns the number of volume samples.
template <int NV, int SM = -1>
__device__ __forceinline__ unsigned sample_one(const TraceParams &P, const RayList_ispc &R, const int i) {
  unsigned nsamp = 0;
  f3 org = mk3(gptr(R.ox)[i], gptr(R.oy)[i], gptr(R.oz)[i]);
  f3 dir = mk3(gptr(R.dx)[i], gptr(R.dy)[i], gptr(R.dz)[i]);
  float rt = gptr(R.t)[i];
  if (dir.x == 0.f) dir.x = 1e-6f;
  if (dir.y == 0.f) dir.y = 1e-6f;
  if (dir.z == 0.f) dir.z = 1e-6f;
  float tEntry, tExit;
  box_clip(org, dir, P.blo, P.bhi, tEntry, tExit);
  if (tEntry < rt) tEntry = rt;
  const float step = P.step;
  float tThis = tEntry + step;
  int hit = -1;
  float sLast[NV], tLast[NV];
  // positions in each volume's voxel coordinates: one FMA per axis and sample
  VoxRay VR[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    VR[k] = vox_ray(P.vv[k].vol, org, dir);
    sLast[k] = vox_sample<SM>(P.vv[k].vol, VR[k], tEntry);
    tLast[k] = tEntry;
  }
  nsamp += NV;
  int iter = 0;
  // gathers issued one step ahead, as in the trace march (positions clamp
  // into the volume, so an unused fetch is harmless)
  VFetch F[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) F[k] = vox_fetch<SM>(P.vv[k].vol, VR[k], tThis);
  // one exit condition, set at the end of each pass (every pass leaves
  // tThis <= tExit: the next position is cut there)
  bool go = tThis <= tExit;
  while (go && iter++ < GXY_MAX_MARCH) {
    float sCur[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) sCur[k] = vol_finish_fma(F[k]);
    // the next position when no crossing ends the march here
    float nt0 = tThis + step;
    if (nt0 > tExit) nt0 = tExit;
#pragma unroll
    for (int k = 0; k < NV; k++) F[k] = vox_fetch<SM>(P.vv[k].vol, VR[k], nt0);
#pragma unroll
    for (int k = 0; k < NV; k++) {
      if (hit == -1) {
        float sThis = sCur[k];
        nsamp++;
        float iso = P.vv[k].iso[0];
        bool h = false;
        float tHit = 0.f;
        if ((sLast[k] < iso && sThis >= iso) || (sLast[k] > iso && sThis <= iso)) {
          tHit = tLast[k] + (((iso - sLast[k]) / (sThis - sLast[k])) * (tThis - tLast[k]));
          h = true;
        }
        sLast[k] = sThis;
        tLast[k] = tThis;
        if (h) { tThis = tHit; hit = k; }
      }
    }
    // a crossing or the volume's exit ends the march, and so does a step that
    // no longer advances t
    go = hit == -1 && tThis != tExit && nt0 > tThis;
    if (go) tThis = nt0;
  }
  gptr(R.t)[i] = hit != -1 ? tThis + 0.001f : tThis;
  gptr(R.term)[i] = hit != -1 ? RAY_SURFACE : RAY_BOUNDARY;
  return nsamp;
}

// The sampler loop of one ray without re-queueing: SamplerTraceRays marches
// to the next isovalue crossing, Sampler::HandleTerminatedRays appends the
// particle at o + t*d (t = crossing + 0.001, Sampler.cpp:71-84), and the ray
// continues from there (the KEEP_HERE re-trace of Renderer.cpp:659-691) until
// it leaves the volume.  Every segment is one ray pass with the samples of
// sample_one, so counts and particles equal those of the per-list loop (the
// particles' order differs).  After max_segments segments the ray is left as
// a SURFACE ray for the caller to re-queue (its particle not yet emitted).
template <int NV, int SM>
__device__ __forceinline__ unsigned sample_all(const TraceParams &P, const RayList_ispc &R, const int i, float *xyz,
                                               int *pcount, int pcap, int max_segments, unsigned &segments) {
  unsigned nsamp = 0;
  segments = 0;
  for (;;) {
    nsamp += sample_one<NV, SM>(P, R, i);
    segments++;
    if (!(gptr(R.term)[i] & RAY_SURFACE) || (int)segments >= max_segments) break;
    const float t = gptr(R.t)[i];
    const float px = gptr(R.ox)[i] + t * gptr(R.dx)[i];
    const float py = gptr(R.oy)[i] + t * gptr(R.dy)[i];
    const float pz = gptr(R.oz)[i] + t * gptr(R.dz)[i];
    const int off = atomicAdd(pcount, 1);
    if (off < pcap) {
      xyz[3 * (int64_t)off + 0] = px;
      xyz[3 * (int64_t)off + 1] = py;
      xyz[3 * (int64_t)off + 2] = pz;
    }
  }
  return nsamp;
}

}  // namespace gxy
