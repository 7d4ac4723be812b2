// trace.hip — TraceRays_TraceRays and SamplerTraceRays_SamplerTraceRays on gfx950.
//
// One work-item per ray over the SoA RayList (coalesced 4-byte loads and stores
// of each field), 256 work-items (four wave64s) per workgroup.  Transfer-function
// tables are staged in LDS once per workgroup; volume samples are gathered from
// HBM through L2 with two x-adjacent voxels per 8-byte load.  The kernel is
// instantiated per volume count (loops unrolled, per-volume state in VGPRs) and
// with or without geometry, so volume-only launches carry no BVH code.
//
// Semantics follow src/renderer/TraceRays.ispc:326-623 step by step (see the
// comments at each stage); geometry follows Embree/OSPRay closest-hit and
// postIntersect rules with the deterministic tie rule (equal t -> lowest
// geomID, then primID).
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

__device__ __forceinline__ bool better(float t, int g, int p, const GeoHit &h) {
  if (t < h.t) return true;
  if (t == h.t && (g < h.geom || (g == h.geom && p < h.prim))) return true;
  return false;
}

#define GXY_BVH_STACK 64
#define GXY_SENTINEL 0x7fffffff

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

template <class LeafFn>
__device__ __forceinline__ void bvh_walk(const BVHNode *nodes, f3 o, f3 inv, float t0, const GeoHit &h,
                                         LeafFn &&leaf) {
  int stack[GXY_BVH_STACK];
  int sp = 0;
  stack[sp++] = GXY_SENTINEL;
  int cur = 0;
  while (cur != GXY_SENTINEL) {
    if (cur >= 0) {
      const float4 *np = (const float4 *)(nodes + cur);
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
        if (sp < GXY_BVH_STACK) stack[sp++] = farc;
        cur = nearc;
      } else if (h0) {
        cur = q3.x;
      } else if (h1) {
        cur = q3.y;
      } else {
        cur = stack[--sp];
      }
    } else {
      int code = ~cur;
      leaf(code >> 4, (code & 15) + 1);
      cur = stack[--sp];
    }
  }
}

__device__ void model_trace(const DModel &M, f3 org, f3 dir, float t0, float tfar, GeoHit &h) {
  h.t = tfar;
  h.geom = -1;
  h.prim = -1;
  h.u = h.v = 0.f;
  if (tfar < t0) return;
  f3 inv = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  if (M.ntris > 0) {
    WRay wr = make_wray(org, dir);
    bvh_walk(M.tnodes, org, inv, t0, h, [&](int start, int count) {
      for (int k = start; k < start + count; k++) {
        const float *p = M.tverts + 9 * (int64_t)k;
        f3 p0 = mk3(p[0], p[1], p[2]), p1 = mk3(p[3], p[4], p[5]), p2 = mk3(p[6], p[7], p[8]);
        float t, u, v;
        if (tri_hit(wr, p0, p1, p2, t0, h.t, t, u, v)) {
          int g = M.tgeom[k], pr = M.tprim[k];
          if (h.geom < 0 || better(t, g, pr, h)) { h.t = t; h.geom = g; h.prim = pr; h.u = u; h.v = v; }
        }
      }
    });
  }
  if (M.nsph > 0) {
    bvh_walk(M.snodes, org, inv, t0, h, [&](int start, int count) {
      for (int k = start; k < start + count; k++) {
        f3 c = mk3(M.scenter[3 * k], M.scenter[3 * k + 1], M.scenter[3 * k + 2]);
        float lim = h.geom < 0 ? h.t : nextafterf(h.t, 3e38f);
        float t;
        if (sphere_hit(c, M.sradius[k], M.seps[k], org, dir, t0, lim, t)) {
          int g = M.sgeom[k], pr = M.sprim[k];
          if (h.geom < 0 || better(t, g, pr, h)) { h.t = t; h.geom = g; h.prim = pr; h.u = h.v = 0.f; }
        }
      }
    });
  }
}

// postIntersect for the hit geometry, then DG_NORMALIZE | DG_FACEFORWARD.
__device__ void geom_shade(const DModel &M, const GeoHit &gh, f3 org, f3 dir, Hit &hit) {
  const DGeomShade &S = M.shade[gh.geom];
  f3 P = add3(org, scl3(gh.t, dir));
  f3 Ng, Ns, col = mk3(1.f, 1.f, 1.f);
  if (S.kind == 0) {
    const uint32_t *I = S.indices + 3 * (int64_t)gh.prim;
    uint32_t i0 = I[0], i1 = I[1], i2 = I[2];
    const float *V = S.vertices;
    f3 p0 = mk3(V[3 * i0], V[3 * i0 + 1], V[3 * i0 + 2]);
    f3 p1 = mk3(V[3 * i1], V[3 * i1 + 1], V[3 * i1 + 2]);
    f3 p2 = mk3(V[3 * i2], V[3 * i2 + 1], V[3 * i2 + 2]);
    Ng = cross3(sub3(p0, p1), sub3(p2, p0));
    Ns = Ng;
    float b0 = 1.0f - gh.u - gh.v, b1 = gh.u, b2 = gh.v;
    if (S.has_normals) {
      const float *N = S.normals;
      f3 a = mk3(N[3 * i0], N[3 * i0 + 1], N[3 * i0 + 2]);
      f3 b = mk3(N[3 * i1], N[3 * i1 + 1], N[3 * i1 + 2]);
      f3 c = mk3(N[3 * i2], N[3 * i2 + 1], N[3 * i2 + 2]);
      Ns = add3(add3(scl3(b0, a), scl3(b1, b)), scl3(b2, c));
    }
    if (S.has_data && S.has_tf) {
      float dv = b0 * S.vdata[i0] + b1 * S.vdata[i1] + b2 * S.vdata[i2];
      col = tf_color(S.tf.col, S.tf, dv);
    }
  } else {
    f3 c = mk3(S.centers[3 * gh.prim], S.centers[3 * gh.prim + 1], S.centers[3 * gh.prim + 2]);
    Ng = sub3(P, c);
    Ns = Ng;
    if (S.has_tf && S.has_data) col = tf_color(S.tf.col, S.tf, S.vdata[gh.prim]);
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

// Workgroup total of a per-item count written to out[blockIdx.x] (no atomics:
// same-address atomics from every wave serialise in one L2 channel).
__device__ __forceinline__ void block_sum_store(unsigned long long v, unsigned long long *out) {
  __shared__ unsigned long long ws[GXY_BLOCK / 64];
  v = wave_sum_u64(v);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < GXY_BLOCK / 64; w++) t += ws[w];
    out[blockIdx.x] = t;
  }
}

// adds the nb workgroup partials into *acc (one workgroup)
__global__ __launch_bounds__(1024) void k_reduce_add(const unsigned long long *part, int nb, unsigned long long *acc) {
  __shared__ unsigned long long ws[16];
  unsigned long long v = 0;
  for (int k = threadIdx.x; k < nb; k += 1024) v += part[k];
  v = wave_sum_u64(v);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < 16; w++) t += ws[w];
    *acc += t;
  }
}

hipError_t launch_reduce_add(const unsigned long long *part, int nb, unsigned long long *acc, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reduce_add, dim3(1), dim3(1024), 0, s, part, nb, acc);
  return hipGetLastError();
}

// ------------------------------------------------------------- the trace kernel
// TraceRays.ispc:466,550,581 with the alpha threshold as a parameter
__device__ __forceinline__ bool opaque_of(float r, float g, float b, float a, float thr) {
  return mnf(mnf(r, g), b) >= 1.0f || a > thr;
}

// WPE > 0 asks the compiler for at least WPE waves per SIMD (fewer VGPRs);
// the autotuner picks the variant.
template <int NV, bool GEOM, int WPE>
__global__ __launch_bounds__(GXY_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8)))
void k_trace(TraceParams P, RayList_ispc R, int n) {
  constexpr int NVS = NV > 0 ? NV : 1;
  __shared__ float s_col[NVS][3 * GXY_TF_MAX];
  __shared__ float s_op[NVS][GXY_TF_MAX];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    for (int k = threadIdx.x; k < 3 * P.vv[v].tf.nc; k += GXY_BLOCK) s_col[v][k] = P.vv[v].tf.col[k];
    for (int k = threadIdx.x; k < P.vv[v].tf.no; k += GXY_BLOCK) s_op[v][k] = P.vv[v].tf.op[k];
  }
  if (NV > 0) __syncthreads();

  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  unsigned long long nsamp = 0;
  if (i < n) {
    const bool shadeFlag = R.type[i] == RAY_PRIMARY;
    f3 org = mk3(R.ox[i], R.oy[i], R.oz[i]);
    f3 dir = mk3(R.dx[i], R.dy[i], R.dz[i]);
    if (dir.x == 0.f) dir.x = 1e-6f;
    if (dir.y == 0.f) dir.y = 1e-6f;
    if (dir.z == 0.f) dir.z = 1e-6f;
    float ray_t0 = R.t[i];
    float ray_t = R.tMax[i];
    const float tTimeout = ray_t;
    float cr = R.r[i], cg = R.g[i], cb = R.b[i], ca = R.o[i];

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
          hit.sample = vol_sample(P.vv[major].vol, hit.point);
          hit.color = tf_color(s_col[major], P.vv[major].tf, hit.sample);
          hit.opacity = 1.0f;
        }
    }

    // LookForGeometryHit (TraceRays.ispc:108-141)
    if (GEOM) {
      GeoHit gh;
      model_trace(P.model, org, dir, ray_t0, ray_t, gh);
      if (P.hit_geom) { P.hit_geom[i] = gh.geom; P.hit_prim[i] = gh.prim; }
      if (gh.geom >= 0) {
        ray_t = gh.t;
        surface_hit = true;
        if (shadeFlag) geom_shade(P.model, gh, org, dir, hit);
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
      bool opaque = opaque_of(cr, cg, cb, ca, P.opaque_alpha);
      float tThis = tEntry;
      int iter = 0;
      while (tThis <= tTermination && !opaque && !hit_iso && iter++ < GXY_MAX_MARCH) {
        f3 coord = add3(org, scl3(tThis, dir));
#pragma unroll
        for (int v = 0; v < NV; v++) sThis[v] = vol_sample(P.vv[v].vol, coord);
        nsamp += NV;
        if (tThis > tEntry && tLast >= epsilon) {
          // LookForIsoHit: the last matching isovalue in list order wins
          bool ih = false;
          int ivid = -1;
#pragma unroll
          for (int major = 0; major < NV; major++) {
            float sl = sLast[major], st = sThis[major];
            const int ni = P.vv[major].nIso;
            for (int m = 0; m < ni; m++) {
              float isoval = P.vv[major].iso[m];
              if (((isoval >= sl) && (isoval < st)) || ((isoval <= sl) && (isoval > st))) {
                ih = true;
                ivid = major;
                hit.t = tLast + ((isoval - sl) / (st - sl)) * (tThis - tLast);
                hit.sample = isoval;
              }
            }
          }
          if (ih) {
            // the shading of the hit (gradient normal, TF colour) reads nothing
            // the rest of the march changes, so it is done once after the loop
            iso_vid = ivid;
            tTermination = hit.t;
            tThis = hit.t;
            surface_hit = true;
            hit_iso = true;
            f3 c2 = add3(org, scl3(tThis, dir));
#pragma unroll
            for (int v = 0; v < NV; v++) sThis[v] = vol_sample(P.vv[v].vol, c2);
            nsamp += NV;
          }
#pragma unroll
          for (int major = 0; major < NV; major++) {
            if (P.vv[major].dvr) {
              const DTF &T = P.vv[major].tf;
              float sVolume = (sLast[major] + sThis[major]) / 2;
              float sop = tf_opacity(s_op[major], T.no, T.lo, T.hi, T.so, sVolume);
              if (sop > 0) {
                if (shadeFlag) {
                  f3 sc = tf_color(s_col[major], T, sVolume);
                  float wo = clamp01f(sop * P.vv[major].vol.irate);
                  float a = 1.0f - ca;
                  cr = cr + a * (wo * sc.x);
                  cg = cg + a * (wo * sc.y);
                  cb = cb + a * (wo * sc.z);
                  ca = ca + a * wo;
                } else {
                  float wop = ((tThis - tLast) * P.istep) * clamp01f(sop * P.vv[major].vol.irate);
                  float f = 1.0f - wop;
                  cr *= f; cg *= f; cb *= f; ca *= f;
                }
              }
            }
          }
        }
#pragma unroll
        for (int v = 0; v < NV; v++) sLast[v] = sThis[v];
        opaque = opaque_of(cr, cg, cb, ca, P.opaque_alpha);
        if (opaque) tTermination = tThis;
        tLast = tThis;
        float nxt = (tThis == tEntry) ? (tEntry + epsilon)
                    : (((tThis + step) > tTermination) && (tThis < tTermination)) ? tTermination : tThis + step;
        // a step that no longer advances t ends the march (the reference would spin)
        if (!(nxt > tThis) && !opaque && !hit_iso) break;
        tThis = nxt;
      }
      ray_t = tTermination;
      if (hit_iso) {
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

    R.r[i] = cr; R.g[i] = cg; R.b[i] = cb; R.o[i] = ca;
    int term = opaque_of(cr, cg, cb, ca, P.opaque_alpha) ? RAY_OPAQUE : 0;
    R.t[i] = ray_t;
    if (surface_hit) {
      term |= RAY_SURFACE;
      if (hit.opacity > 0.999f) term |= RAY_OPAQUE;
      R.sr[i] = hit.color.x; R.sg[i] = hit.color.y; R.sb[i] = hit.color.z; R.so[i] = 1.0f;
      R.nx[i] = hit.normal.x; R.ny[i] = hit.normal.y; R.nz[i] = hit.normal.z;
    } else if (tTermination == tExitVolume) {
      term |= RAY_BOUNDARY;
    } else if (tTermination == tTimeout) {
      term |= RAY_TIMEOUT;
    }
    R.term[i] = term;
  }
  if (P.sample_partials) block_sum_store(nsamp, P.sample_partials);
}

// ------------------------------------------------------------- persistent-lane march
// The same per-ray computation as k_trace (volumes, no geometry), organised so
// that a lane whose ray has finished takes the next ray of its wave's chunk
// instead of idling until the wave's longest ray is done.  Each wave owns
// `chunk` consecutive rays (camera tiles stay together); idle lanes are
// refilled once at least `refill` of them are idle (or all are), with ray
// indices handed out by a wave-wide ballot + popcount (no atomics).
struct MarchState {
  f3 org, dir;
  float tEntry, tExitVolume, tTimeout, tTerm, tThis, tLast;
  float cr, cg, cb, ca;
  float hit_t, hit_sample;
  int i;        // ray index, -1 when idle
  int flags;    // 1 shade, 2 surface_hit, 4 hit_iso, 8 opaque-surface (hit.opacity 1)
  int iso_vid;
};

template <int NV>
__device__ __forceinline__ void pl_finish(const TraceParams &P, const RayList_ispc &R, MarchState &m,
                                          const float (*s_col)[3 * GXY_TF_MAX]) {
  const int i = m.i;
  float ray_t = m.tTerm;
  if (m.flags & 4) {
    // LookForIsoHit shading (TraceRays.ispc:284-307)
    if (m.flags & 1) {
      f3 pt = add3(m.org, scl3(m.hit_t, m.dir));
      f3 nrm = mk3(0.f, 0.f, 0.f), col = nrm;
#pragma unroll
      for (int major = 0; major < NV; major++)
        if (major == m.iso_vid) {
          nrm = safe_normalize3(vol_gradient(P.vv[major].vol, pt));
          if (dot3(m.dir, nrm) > 0) nrm = neg3(nrm);
          col = tf_color(s_col[major], P.vv[major].tf, m.hit_sample);
        }
      R.sr[i] = col.x; R.sg[i] = col.y; R.sb[i] = col.z; R.so[i] = 1.0f;
      R.nx[i] = nrm.x; R.ny[i] = nrm.y; R.nz[i] = nrm.z;
      m.flags |= 8;
    } else if (!(m.flags & 8)) {
      // a non-primary iso hit with no earlier slice hit: zero colour/normal
      R.sr[i] = 0.f; R.sg[i] = 0.f; R.sb[i] = 0.f; R.so[i] = 1.0f;
      R.nx[i] = 0.f; R.ny[i] = 0.f; R.nz[i] = 0.f;
    }
  }
  R.r[i] = m.cr; R.g[i] = m.cg; R.b[i] = m.cb; R.o[i] = m.ca;
  int term = opaque_of(m.cr, m.cg, m.cb, m.ca, P.opaque_alpha) ? RAY_OPAQUE : 0;
  R.t[i] = ray_t;
  if (m.flags & 2) {
    term |= RAY_SURFACE;
    if (m.flags & 8) term |= RAY_OPAQUE;
  } else if (m.tTerm == m.tExitVolume) {
    term |= RAY_BOUNDARY;
  } else if (m.tTerm == m.tTimeout) {
    term |= RAY_TIMEOUT;
  }
  R.term[i] = term;
  m.i = -1;
}

// Loads ray i and runs everything before the march (TraceRays.ispc:371-442);
// returns false when the ray is already finished (no march needed).
template <int NV>
__device__ __forceinline__ bool pl_start(const TraceParams &P, const RayList_ispc &R, MarchState &m, int i,
                                         float *sLast, const float (*s_col)[3 * GXY_TF_MAX]) {
  m.i = i;
  m.flags = R.type[i] == RAY_PRIMARY ? 1 : 0;
  m.org = mk3(R.ox[i], R.oy[i], R.oz[i]);
  f3 dir = mk3(R.dx[i], R.dy[i], R.dz[i]);
  if (dir.x == 0.f) dir.x = 1e-6f;
  if (dir.y == 0.f) dir.y = 1e-6f;
  if (dir.z == 0.f) dir.z = 1e-6f;
  m.dir = dir;
  float ray_t0 = R.t[i];
  float ray_t = R.tMax[i];
  m.tTimeout = ray_t;
  m.cr = R.r[i]; m.cg = R.g[i]; m.cb = R.b[i]; m.ca = R.o[i];
  float tEntry, tExitVolume;
  box_clip(m.org, dir, P.blo, P.bhi, tEntry, tExitVolume);
  if (tEntry < ray_t0) tEntry = ray_t0;
  else if (tEntry > ray_t0) ray_t0 = tEntry;
  ray_t = mnf(ray_t, tExitVolume);
  m.tEntry = tEntry;
  m.tExitVolume = tExitVolume;
  m.iso_vid = -1;
  // LookForSliceHit (TraceRays.ispc:143-250); its outputs are written now and
  // replaced at the end if an isosurface is hit later
  bool sh = false;
  int vid = -1;
  f3 snrm = mk3(0.f, 0.f, 0.f);
  float st = 0.f;
#pragma unroll
  for (int major = 0; major < NV; major++) {
    const int ns = P.vv[major].nSlices;
    for (int minor = 0; minor < ns; minor++) {
      f3 pn = mk3(P.vv[major].slices[4 * minor], P.vv[major].slices[4 * minor + 1], P.vv[major].slices[4 * minor + 2]);
      float d = P.vv[major].slices[4 * minor + 3];
      float denom = dot3(dir, pn);
      if (fabsf(denom) > 0.0001f) {
        float t = (d - dot3(m.org, pn)) / denom;
        if (t >= ray_t0 && t <= ray_t) {
          snrm = denom > 0 ? neg3(pn) : pn;
          st = t;
          vid = major;
          ray_t = t;
          sh = true;
        }
      }
    }
  }
  if (sh) {
    f3 col = mk3(0.f, 0.f, 0.f);
    if (m.flags & 1) {
      f3 pt = add3(m.org, scl3(st, dir));
#pragma unroll
      for (int major = 0; major < NV; major++)
        if (major == vid) {
          float smp = vol_sample(P.vv[major].vol, pt);
          col = tf_color(s_col[major], P.vv[major].tf, smp);
        }
    }
    R.sr[i] = col.x; R.sg[i] = col.y; R.sb[i] = col.z; R.so[i] = 1.0f;
    R.nx[i] = snrm.x; R.ny[i] = snrm.y; R.nz[i] = snrm.z;
    m.flags |= 2 | 8;
  }
  m.tTerm = ray_t;
  if (!P.integrate) return false;
  m.tLast = tEntry + P.eps;
  m.tThis = tEntry;
#pragma unroll
  for (int v = 0; v < NV; v++) sLast[v] = 0.f;
  return m.tThis <= m.tTerm && !opaque_of(m.cr, m.cg, m.cb, m.ca, P.opaque_alpha);
}

// One march iteration (TraceRays.ispc:470-556); returns false when the ray is done.
template <int NV>
__device__ __forceinline__ bool pl_step(const TraceParams &P, MarchState &m, float *sLast, int &iter,
                                        const float (*s_col)[3 * GXY_TF_MAX], const float (*s_op)[GXY_TF_MAX],
                                        unsigned long long &nsamp) {
  const float step = P.step, epsilon = P.eps;
  float tThis = m.tThis;
  float sThis[NV];
  f3 coord = add3(m.org, scl3(tThis, m.dir));
#pragma unroll
  for (int v = 0; v < NV; v++) sThis[v] = vol_sample(P.vv[v].vol, coord);
  nsamp += NV;
  bool hit_iso = false;
  if (tThis > m.tEntry && m.tLast >= epsilon) {
    bool ih = false;
    int ivid = -1;
    float ht = 0.f, hs = 0.f;
#pragma unroll
    for (int major = 0; major < NV; major++) {
      float sl = sLast[major], st = sThis[major];
      const int ni = P.vv[major].nIso;
      for (int k = 0; k < ni; k++) {
        float isoval = P.vv[major].iso[k];
        if (((isoval >= sl) && (isoval < st)) || ((isoval <= sl) && (isoval > st))) {
          ih = true;
          ivid = major;
          ht = m.tLast + ((isoval - sl) / (st - sl)) * (tThis - m.tLast);
          hs = isoval;
        }
      }
    }
    if (ih) {
      m.iso_vid = ivid;
      m.hit_t = ht;
      m.hit_sample = hs;
      m.tTerm = ht;
      tThis = ht;
      m.flags |= 2 | 4;
      hit_iso = true;
      f3 c2 = add3(m.org, scl3(tThis, m.dir));
#pragma unroll
      for (int v = 0; v < NV; v++) sThis[v] = vol_sample(P.vv[v].vol, c2);
      nsamp += NV;
    }
#pragma unroll
    for (int major = 0; major < NV; major++) {
      if (P.vv[major].dvr) {
        const DTF &T = P.vv[major].tf;
        float sVolume = (sLast[major] + sThis[major]) / 2;
        float sop = tf_opacity(s_op[major], T.no, T.lo, T.hi, T.so, sVolume);
        if (sop > 0) {
          if (m.flags & 1) {
            f3 sc = tf_color(s_col[major], T, sVolume);
            float wo = clamp01f(sop * P.vv[major].vol.irate);
            float a = 1.0f - m.ca;
            m.cr = m.cr + a * (wo * sc.x);
            m.cg = m.cg + a * (wo * sc.y);
            m.cb = m.cb + a * (wo * sc.z);
            m.ca = m.ca + a * wo;
          } else {
            float wop = ((tThis - m.tLast) * P.istep) * clamp01f(sop * P.vv[major].vol.irate);
            float f = 1.0f - wop;
            m.cr *= f; m.cg *= f; m.cb *= f; m.ca *= f;
          }
        }
      }
    }
  }
#pragma unroll
  for (int v = 0; v < NV; v++) sLast[v] = sThis[v];
  bool opaque = opaque_of(m.cr, m.cg, m.cb, m.ca, P.opaque_alpha);
  if (opaque) m.tTerm = tThis;
  m.tLast = tThis;
  float nxt = (tThis == m.tEntry) ? (m.tEntry + epsilon)
              : (((tThis + step) > m.tTerm) && (tThis < m.tTerm)) ? m.tTerm : tThis + step;
  if (!(nxt > tThis) && !opaque && !hit_iso) return false;
  m.tThis = nxt;
  if (++iter >= GXY_MAX_MARCH) return false;
  return nxt <= m.tTerm && !opaque && !hit_iso;
}

template <int NV>
__global__ __launch_bounds__(GXY_BLOCK) void k_trace_pl(TraceParams P, RayList_ispc R, int n, int chunk,
                                                        int refill) {
  __shared__ float s_col[NV][3 * GXY_TF_MAX];
  __shared__ float s_op[NV][GXY_TF_MAX];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    for (int k = threadIdx.x; k < 3 * P.vv[v].tf.nc; k += GXY_BLOCK) s_col[v][k] = P.vv[v].tf.col[k];
    for (int k = threadIdx.x; k < P.vv[v].tf.no; k += GXY_BLOCK) s_op[v][k] = P.vv[v].tf.op[k];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int wave = (blockIdx.x * GXY_BLOCK + threadIdx.x) >> 6;
  const long long start = (long long)wave * chunk;
  const long long end = start + chunk < n ? start + chunk : n;
  long long next = start;
  unsigned long long nsamp = 0;
  MarchState m;
  m.i = -1;
  float sLast[NV];
  int iter = 0;
  for (;;) {
    unsigned long long idle = __ballot(m.i < 0);
    const int nidle = __popcll(idle);
    if (nidle == 64 && next >= end) break;
    if (next < end && (nidle >= refill || nidle == 64)) {
      if (m.i < 0) {
        long long j = next + __popcll(idle & lt);
        if (j < end) {
          iter = 0;
          if (!pl_start<NV>(P, R, m, (int)j, sLast, s_col)) pl_finish<NV>(P, R, m, s_col);
        }
      }
      next += nidle;
      continue;
    }
    if (m.i >= 0) {
      if (!pl_step<NV>(P, m, sLast, iter, s_col, s_op, nsamp)) pl_finish<NV>(P, R, m, s_col);
    }
  }
  if (P.sample_partials) block_sum_store(nsamp, P.sample_partials);
}

template <int NV>
static hipError_t launch_trace_pl_nv(const TraceParams &P, const RayList_ispc &R, int n, int chunk, int refill,
                                     hipStream_t s) {
  int waves = (n + chunk - 1) / chunk;
  dim3 g((waves + 3) / 4), b(GXY_BLOCK);
  hipLaunchKernelGGL(k_trace_pl<NV>, g, b, 0, s, P, R, n, chunk, refill);
  return hipGetLastError();
}

int trace_pl_blocks(int n, int chunk) { return ((n + chunk - 1) / chunk + 3) / 4; }

hipError_t launch_trace_pl(const TraceParams &P, const RayList_ispc &R, int n, int chunk, int refill,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (chunk < 64 || refill < 1 || refill > 64) return hipErrorInvalidValue;
  switch (P.nvv) {
    case 1: return launch_trace_pl_nv<1>(P, R, n, chunk, refill, s);
    case 2: return launch_trace_pl_nv<2>(P, R, n, chunk, refill, s);
    case 3: return launch_trace_pl_nv<3>(P, R, n, chunk, refill, s);
    case 4: return launch_trace_pl_nv<4>(P, R, n, chunk, refill, s);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------- sampler kernel
// SamplerTraceRays_SamplerTraceRays (SamplerTraceRays.ispc:128-222) with the
// IsoSamplerVis interval test (IsoSamplerVis.ispc:36-64).
template <int NV>
__global__ __launch_bounds__(GXY_BLOCK) void k_sampler(TraceParams P, RayList_ispc R, int n) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  unsigned long long nsamp = 0;
  if (i < n) {
    f3 org = mk3(R.ox[i], R.oy[i], R.oz[i]);
    f3 dir = mk3(R.dx[i], R.dy[i], R.dz[i]);
    float rt = R.t[i];
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
    f3 c0 = add3(org, scl3(tEntry, dir));
#pragma unroll
    for (int k = 0; k < NV; k++) { sLast[k] = vol_sample(P.vv[k].vol, c0); tLast[k] = tEntry; }
    nsamp += NV;
    int iter = 0;
    while (tThis <= tExit && hit == -1 && iter++ < GXY_MAX_MARCH) {
#pragma unroll
      for (int k = 0; k < NV; k++) {
        if (hit == -1) {
          float sThis = vol_sample(P.vv[k].vol, add3(org, scl3(tThis, dir)));
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
      if (hit != -1 || tThis == tExit) break;
      float nt = tThis + step;
      if (nt > tExit) nt = tExit;
      if (!(nt > tThis)) break;
      tThis = nt;
    }
    R.t[i] = hit != -1 ? tThis + 0.001f : tThis;
    R.term[i] = hit != -1 ? RAY_SURFACE : RAY_BOUNDARY;
  }
  if (P.sample_partials) block_sum_store(nsamp, P.sample_partials);
}

// ------------------------------------------------------------- launchers
template <int NV>
static hipError_t launch_trace_nv(const TraceParams &P, const RayList_ispc &R, int n, int wpe, hipStream_t s) {
  dim3 g((n + GXY_BLOCK - 1) / GXY_BLOCK), b(GXY_BLOCK);
  if (P.has_model) hipLaunchKernelGGL((k_trace<NV, true, 0>), g, b, 0, s, P, R, n);
  else if (wpe == 5) hipLaunchKernelGGL((k_trace<NV, false, 5>), g, b, 0, s, P, R, n);
  else if (wpe == 6) hipLaunchKernelGGL((k_trace<NV, false, 6>), g, b, 0, s, P, R, n);
  else hipLaunchKernelGGL((k_trace<NV, false, 0>), g, b, 0, s, P, R, n);
  return hipGetLastError();
}

hipError_t launch_trace(const TraceParams &P, const RayList_ispc &R, int n, int wpe, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  switch (P.nvv) {
    case 0: return launch_trace_nv<0>(P, R, n, wpe, s);
    case 1: return launch_trace_nv<1>(P, R, n, wpe, s);
    case 2: return launch_trace_nv<2>(P, R, n, wpe, s);
    case 3: return launch_trace_nv<3>(P, R, n, wpe, s);
    case 4: return launch_trace_nv<4>(P, R, n, wpe, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_sampler(const TraceParams &P, const RayList_ispc &R, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  dim3 g((n + GXY_BLOCK - 1) / GXY_BLOCK), b(GXY_BLOCK);
  switch (P.nvv) {
    case 1: hipLaunchKernelGGL(k_sampler<1>, g, b, 0, s, P, R, n); break;
    case 2: hipLaunchKernelGGL(k_sampler<2>, g, b, 0, s, P, R, n); break;
    case 3: hipLaunchKernelGGL(k_sampler<3>, g, b, 0, s, P, R, n); break;
    case 4: hipLaunchKernelGGL(k_sampler<4>, g, b, 0, s, P, R, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gxy
