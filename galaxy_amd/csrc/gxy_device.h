// gxy_device.h — device-side scene descriptors and ray math for gfx950.
//
// Everything a kernel needs about a Visualization is packed into one
// TraceParams value passed as a kernel argument (no per-launch allocation, no
// pointer chasing through host structs): up to GXY_MAX_VOLVIS volumes with
// their transfer-function tables, slice planes and isovalues, the partition
// box, the march step, and the device BVHs of the geometry model.
//
// The arithmetic below keeps the operation order of the reference kernels
// (src/renderer/TraceRays.ispc) so that results can be compared with the CPU
// restatement ray by ray; all HIP code is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gxy_hip.h"

#define GXY_MAX_VOLVIS 4
#define GXY_MAX_SLICES 8
#define GXY_MAX_ISO 8
// path-line curve segments are traced as round cones between curve samples at
// equal steps of u, as many per segment as keep the traced solid within
// GXY_CURVE_TOL x (the segment's least radius) of the swept-sphere solid
// (at most GXY_CURVE_MAX_PIECES)
#define GXY_CURVE_TOL 0.01
#define GXY_CURVE_MAX_PIECES 32
#define GXY_TF_MAX 256
#define GXY_BLOCK 256
// upper bound on march steps per ray and pass (a 4096^3 diagonal at half-voxel
// steps needs about 14k); the CPU restatement uses the same bound
#define GXY_MAX_MARCH (1 << 20)

// A structured volume brick resident in HBM (x fastest).
struct DVolume {
  const float *data;
  int n[3];
  float org[3];
  float isp[3];   // 1 / spacing (OSPRay transformWorldToLocal multiplies)
  float sp[3];
  float ub[3];    // nextafter(dims-1, 0): local-coordinate clamp
  float rate;     // samplingRate
  float step;     // samplingStep (minimum spacing)
  float irate;    // 1 / samplingRate
  int small;      // <= 2^30 voxels, x-y plane and z extent < 2^24: 32-bit byte offsets
  int brick;      // 0, or B = 4 / 8: a bricked copy of the voxels is in bdata
  int bsmall;     // the bricked copy spans under 4 GiB and a plane of bricks holds
                  // under 2^24 floats: 32-bit byte offsets from 24-bit multiplies
  int nb[3];      // bricks per axis: ceil((n - 1) / B)
  const float *bdata;  // bricks of (B+1)^3 voxels (one-voxel overlap), x fastest
                       // inside a brick, each padded to a multiple of 128 bytes
};

// Floats per stored brick: (B+1)^3 rounded up to 32 (128 bytes).
__host__ __device__ constexpr int brick_floats(int B) { return ((B + 1) * (B + 1) * (B + 1) + 31) / 32 * 32; }

// Piecewise-linear TF: nc colours (rgb) and no opacities over [lo, hi].
struct DTF {
  const float *col;
  const float *op;
  int nc, no;
  float lo, hi;
  float sc, so;   // (nc-1)/(hi-lo), (no-1)/(hi-lo)
  float xc, xo;   // nc-1, no-1 as floats (the clamp bounds of the table coordinate)
  float lo2, hsc, hso;  // 2 lo, sc / 2, so / 2: coordinates of a doubled value
};

struct DVolVis {
  DVolume vol;
  DTF tf;
  int nSlices, nIso, dvr, pad_;
  float slices[4 * GXY_MAX_SLICES];
  float iso[GXY_MAX_ISO];
};

// Binary BVH with both child boxes stored in the parent (64 B per node).
// child index >= 0: inner node; < 0: leaf holding primitive range
// (~c) >> 4 .. count (c & 15) in the sorted primitive arrays.
struct BVHNode {
  float lo0[3], hi0[3];
  float lo1[3], hi1[3];
  int c0, c1;
  int pad_[2];
};

// 4-wide BVH node (128 bytes): the boxes of up to four children stored per
// axis, then the child codes (inner node index >= 0, leaf code < 0 as in
// BVHNode).  An empty slot has an empty box (lo = +inf, hi = -inf).
struct BVH4Node {
  float lox[4], loy[4], loz[4];
  float hix[4], hiy[4], hiz[4];
  int child[4];
  int pad_[4];
};

struct DModel {
  // triangles, in BVH leaf order: 9 floats per triangle (v0 v1 v2)
  const BVHNode *tnodes;
  const float *tverts;
  const int *tgeom;     // geomID per sorted triangle
  const int *tprim;     // primID per sorted triangle
  int ntris;
  int troot_leaf;       // 1 when the whole tree is a single leaf
  // spheres, in BVH leaf order
  const BVHNode *snodes;
  const float *scenter;  // 3 per sphere
  const float *sradius;  // per sphere
  const float *seps;     // per sphere: owning geometry's epsilon
  const int *sgeom;
  const int *sprim;
  int nsph;
  int sroot_leaf;
  // path-line curve pieces (round cones), in BVH leaf order: sphere a and
  // sphere b (x y z radius each) and the curve parameters (u0, u1) they sit at
  const BVHNode *cnodes;
  const float *ca;        // 4 per piece
  const float *cb;        // 4 per piece
  const float *cu;        // 2 per piece
  const int *cgeom;
  const int *cprim;       // curve segment (Embree primID) of each piece
  int ncurve;
  // shading data, indexed by geomID
  int ngeom;
  const struct DGeomShade *shade;
  int depth;            // deepest leaf of either tree (traversal stack bound)
  // the same trees collapsed to 4-wide nodes (null when not built)
  const BVH4Node *tnodes4;
  const BVH4Node *snodes4;
  const BVH4Node *cnodes4;
  int depth4;           // deepest leaf of either 4-wide tree
  int wide;             // traverse the 4-wide trees
};

// Per-geometry shading data (DataDrivenTriangleMesh_postIntersect,
// DataDrivenSpheres_postIntersect, DataDrivenPathLines_postIntersect).
struct DGeomShade {
  int kind;               // 0 triangles, 1 spheres, 2 path lines
  int has_normals, has_data, has_tf;
  const float *vertices;  // triangles: 3 per vertex; path lines: control points, 4 each (x y z radius)
  const uint32_t *indices;  // path lines: first control point of each curve segment
  const float *normals;
  const float *vdata;     // per vertex (triangles) or per sphere
  const float *centers;
  float r0, r1, v0, v1;   // path lines: radius and value range
  DTF tf;
};

struct TraceParams {
  int nvv;
  int integrate;
  int has_model;
  int occlusion;       // 1: non-primary rays end their geometry search at the first hit (no volumes)
  float blo[3], bhi[3];
  float step, eps;
  float istep;         // 1 / step
  float opaque_alpha;  // early termination: alpha above this is opaque (0.999)
  DVolVis vv[GXY_MAX_VOLVIS];
  DModel model;
  unsigned long long *sample_count;    // volume samples are added here, one atomic per wave (may be null)
  int *hit_geom;                       // side-band geomID per ray, may be null
  int *hit_prim;
  // a launch queued before its list's size is known: when dn is set, the
  // launch covers min(n, *dn * dn_mul) rays (n is then the list's capacity)
  const int *dn;
  int dn_mul;
};

// ------------------------------------------------------------------ math
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 scl3(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float mnf(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mxf(float a, float b) { return a > b ? a : b; }
// [0, 1] clamp of a number (v_med3_f32; the march's opacities are never NaN)
__device__ __forceinline__ float clamp01f(float v) { return __builtin_amdgcn_fmed3f(v, 0.f, 1.f); }
__device__ __forceinline__ f3 normalize3(f3 v) { float s = 1.0f / sqrtf(dot3(v, v)); return scl3(s, v); }
__device__ __forceinline__ f3 safe_normalize3(f3 v) {
  float d = dot3(v, v);
  d = d > 1e-12f ? d : 1e-12f;
  float s = 1.0f / sqrtf(d);
  return scl3(s, v);
}

// A pointer read from device memory (rather than passed as a kernel argument)
// is a generic pointer, so its loads would be flat loads with a 64-bit address
// per lane; the round trip through the global address space lets the compiler
// use global loads with a scalar base instead.  Every such pointer here points
// into device memory allocated with hipMalloc.
template <class T>
__device__ __forceinline__ T *gptr(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)p;
}

// Unaligned (4-byte aligned) pair load: one global_load_dwordx2 for two
// x-adjacent voxels.
struct __attribute__((packed, aligned(4))) float2u { float x, y; };

// Trilinear sample (OSPRay StructuredVolume semantics; x, then y, then z).
// The clamp to [0, nextafter(dims-1, 0)] is one v_med3_f32 per axis (equal to
// the compare-select clamp for every non-NaN input; NaN is mapped to 0 after).
// SM = 1: the volume is "small" (at most 2^30 voxels, an x-y plane and the
// z extent below 2^24), so the voxel offset is two full-rate 24-bit
// multiply-adds and the four pair loads are scalar base + one 32-bit byte
// offset (no 64-bit address arithmetic per sample); SM = 0: 64-bit offsets;
// SM = -1: chosen per call from V.small.
// The gathers of one sample (issued) and their interpolation (which waits for
// them), separable so that a march can issue the next sample's loads before it
// composites the current one.
struct VFetch {
  float2u p00, p10, p01, p11;
  float f0, f1, f2;
};

// The four x-pair loads of cell (i0, i1, i2) from the bricked copy (B = 4 or
// 8).  BIG = false when the whole copy spans under 4 GiB and a plane of
// bricks holds under 2^24 floats (DVolume::bsmall): the byte offset is formed
// in 32 bits (24-bit multiply-adds) and the loads take it beside the uniform
// base address; BIG = true forms a 64-bit address.
template <int B, bool BIG>
__device__ __forceinline__ void brick_fetch(const DVolume &V, int i0, int i1, int i2, VFetch &F) {
  constexpr int LB = B == 4 ? 2 : 3, E = B + 1, PL = E * E;
  if (!BIG) {
    // The float offset BF * brick + loc, with brick = a0 + nb0 a1 + nb0 nb1 a2
    // (a = i >> log2 B), loc = b0 + E b1 + E^2 b2 and b = i - B a, is
    //   i0 + (BF - B) a0 + E i1 + (BF nb0 - E B) a1 + E^2 i2 + (BF nb0 nb1 - E^2 B) a2:
    // one chain of 24-bit multiply-adds with no masks (every factor is under
    // 2^24: DVolume::bsmall requires BF nb0 nb1 < 2^24), scaled to bytes at
    // the end (the copy spans under 4 GiB).
    constexpr unsigned BF = brick_floats(B);
    const unsigned k1 = BF * (unsigned)V.nb[0] - E * B;
    const unsigned k2 = BF * (unsigned)V.nb[0] * (unsigned)V.nb[1] - PL * B;
    const unsigned a0 = (unsigned)i0 >> LB, a1 = (unsigned)i1 >> LB, a2 = (unsigned)i2 >> LB;
    unsigned w = __umul24(a0, BF - B) + (unsigned)i0;
    w = __umul24((unsigned)i1, (unsigned)E) + w;
    w = __umul24(a1, k1) + w;
    w = __umul24((unsigned)i2, (unsigned)PL) + w;
    w = __umul24(a2, k2) + w;
    const char *d = (const char *)gptr(V.bdata) + (size_t)(w * 4u);
    F.p00 = *(const float2u *)(d);
    F.p10 = *(const float2u *)(d + 4 * E);
    F.p01 = *(const float2u *)(d + 4 * PL);
    F.p11 = *(const float2u *)(d + 4 * (PL + E));
    return;
  }
  unsigned ty = __umul24((unsigned)(i1 >> LB), (unsigned)V.nb[0]);
  unsigned tz = __umul24((unsigned)(i2 >> LB), (unsigned)V.nb[0] * (unsigned)V.nb[1]);
  // (an empty statement that keeps the two products as they are: the
  // compiler would otherwise merge them with the sums into 64-bit
  // multiply-adds, which issue at a quarter of the rate of the 24-bit
  // multiplies and the three-input add)
  asm("" : "+v"(ty), "+v"(tz));
  const unsigned brick = (unsigned)(i0 >> LB) + ty + tz;
  const unsigned loc = (unsigned)(i0 & (B - 1)) + (unsigned)(i1 & (B - 1)) * E + (unsigned)(i2 & (B - 1)) * PL;
  const float *d = gptr(V.bdata) + ((size_t)brick * brick_floats(B) + loc);
  F.p00 = *(const float2u *)(d);
  F.p10 = *(const float2u *)(d + E);
  F.p01 = *(const float2u *)(d + PL);
  F.p11 = *(const float2u *)(d + PL + E);
}

// Voxel addressing variants (SM): 4 / 8 the bricked copy with 32-bit offsets,
// 12 / 16 the bricked copy (B = 4 / 8) with 64-bit addresses, 1 the linear
// voxels with 32-bit offsets, 0 linear with 64-bit addresses, -1 chosen per
// volume at run time (linear).
__host__ __device__ constexpr bool sm_bricked(int SM) { return SM == 4 || SM == 8 || SM == 12 || SM == 16; }
__host__ __device__ constexpr int sm_brick(int SM) { return (SM == 4 || SM == 12) ? 4 : 8; }
__host__ __device__ constexpr bool sm_big(int SM) { return SM == 12 || SM == 16; }
// the variant for a volume's current layout
__host__ __device__ inline int sm_of(const DVolume &V) {
  return V.brick ? V.brick + (V.bsmall ? 0 : 8) : (V.small ? 1 : 0);
}

// SM = 4 / 8: the bricked copy.  Cell (i0, i1, i2) lies in brick (i >> log2 B)
// at local (i & (B-1)) and its eight corners are inside that brick (the
// overlap voxel), so the four pair loads sit at constant offsets from one
// address: a ray's consecutive samples stay inside one (B+1)^3 block instead
// of touching new cache lines of four x-rows per voxel step.
template <int SM>
__device__ __forceinline__ VFetch vol_fetch(const DVolume &V, f3 p) {
  float l0 = (p.x - V.org[0]) * V.isp[0];
  float l1 = (p.y - V.org[1]) * V.isp[1];
  float l2 = (p.z - V.org[2]) * V.isp[2];
  l0 = __builtin_amdgcn_fmed3f(l0, 0.f, V.ub[0]);
  l1 = __builtin_amdgcn_fmed3f(l1, 0.f, V.ub[1]);
  l2 = __builtin_amdgcn_fmed3f(l2, 0.f, V.ub[2]);
  if (l0 != l0) l0 = 0.f;
  if (l1 != l1) l1 = 0.f;
  if (l2 != l2) l2 = 0.f;
  int i0 = (int)l0, i1 = (int)l1, i2 = (int)l2;
  VFetch F;
  F.f0 = l0 - (float)i0;
  F.f1 = l1 - (float)i1;
  F.f2 = l2 - (float)i2;
  if (sm_bricked(SM)) {
    brick_fetch<sm_brick(SM), sm_big(SM)>(V, i0, i1, i2, F);
    return F;
  }
  const float *vd = gptr(V.data);
  const bool small = SM == 1 || (SM < 0 && V.small);
  if (small) {
    const unsigned sy = (unsigned)V.n[0], sz = (unsigned)V.n[0] * (unsigned)V.n[1];
    const unsigned off = (unsigned)i0 + __umul24((unsigned)i1, sy) + __umul24((unsigned)i2, sz);
    const size_t bo = (size_t)(off * 4u);
    F.p00 = *(const float2u *)((const char *)vd + bo);
    F.p10 = *(const float2u *)((const char *)(vd + sy) + bo);
    F.p01 = *(const float2u *)((const char *)(vd + sz) + bo);
    F.p11 = *(const float2u *)((const char *)(vd + sz + sy) + bo);
  } else {
    const int64_t sy = V.n[0], SZ = (int64_t)V.n[0] * V.n[1];
    const float *d = vd + ((int64_t)i0 + sy * i1 + SZ * i2);
    F.p00 = *(const float2u *)(d);
    F.p10 = *(const float2u *)(d + sy);
    F.p01 = *(const float2u *)(d + SZ);
    F.p11 = *(const float2u *)(d + SZ + sy);
  }
  return F;
}

// The march's fetch: the position is given in voxel coordinates (a ray's
// l(t) = A + t * B, one FMA per axis, A = (org - origin) / spacing and
// B = dir / spacing), clamped into the grid with one v_med3_f32 per axis.  A
// NaN coordinate (which a finite ray never produces) would still address voxel
// 0: v_cvt_i32_f32 maps NaN to 0.
template <int SM>
__device__ __forceinline__ VFetch vol_fetch_vox(const DVolume &V, float l0, float l1, float l2) {
  l0 = __builtin_amdgcn_fmed3f(l0, 0.f, V.ub[0]);
  l1 = __builtin_amdgcn_fmed3f(l1, 0.f, V.ub[1]);
  l2 = __builtin_amdgcn_fmed3f(l2, 0.f, V.ub[2]);
  const int i0 = (int)l0, i1 = (int)l1, i2 = (int)l2;
  VFetch F;
  // l - floor(l) in one v_fract_f32: exact, and equal to l - (float)i for the
  // clamped l >= 0
  F.f0 = __builtin_amdgcn_fractf(l0);
  F.f1 = __builtin_amdgcn_fractf(l1);
  F.f2 = __builtin_amdgcn_fractf(l2);
  if (sm_bricked(SM)) {
    brick_fetch<sm_brick(SM), sm_big(SM)>(V, i0, i1, i2, F);
    return F;
  }
  const float *vd = gptr(V.data);
  const bool small = SM == 1 || (SM < 0 && V.small);
  if (small) {
    const unsigned sy = (unsigned)V.n[0], sz = (unsigned)V.n[0] * (unsigned)V.n[1];
    const unsigned off = (unsigned)i0 + __umul24((unsigned)i1, sy) + __umul24((unsigned)i2, sz);
    const size_t bo = (size_t)(off * 4u);
    F.p00 = *(const float2u *)((const char *)vd + bo);
    F.p10 = *(const float2u *)((const char *)(vd + sy) + bo);
    F.p01 = *(const float2u *)((const char *)(vd + sz) + bo);
    F.p11 = *(const float2u *)((const char *)(vd + sz + sy) + bo);
  } else {
    const int64_t sy = V.n[0], SZ = (int64_t)V.n[0] * V.n[1];
    const float *d = vd + ((int64_t)i0 + sy * i1 + SZ * i2);
    F.p00 = *(const float2u *)(d);
    F.p10 = *(const float2u *)(d + sy);
    F.p01 = *(const float2u *)(d + SZ);
    F.p11 = *(const float2u *)(d + SZ + sy);
  }
  return F;
}

// a + f * (b - a): one subtraction and one FMA
__device__ __forceinline__ float lerp_fma(float a, float b, float f) { return __builtin_fmaf(f, b - a, a); }

// Trilinear interpolation of a fetch with FMA lerps (x, then y, then z).
__device__ __forceinline__ float vol_finish_fma(const VFetch &F) {
  const float a00 = lerp_fma(F.p00.x, F.p00.y, F.f0);
  const float a10 = lerp_fma(F.p10.x, F.p10.y, F.f0);
  const float a01 = lerp_fma(F.p01.x, F.p01.y, F.f0);
  const float a11 = lerp_fma(F.p11.x, F.p11.y, F.f0);
  return lerp_fma(lerp_fma(a00, a10, F.f1), lerp_fma(a01, a11, F.f1), F.f2);
}

// A ray's position in one volume's voxel coordinates, l(t) = A + t * B.
struct VoxRay {
  float a0, a1, a2, b0, b1, b2;
};
__device__ __forceinline__ VoxRay vox_ray(const DVolume &V, f3 org, f3 dir) {
  VoxRay r;
  r.a0 = (org.x - V.org[0]) * V.isp[0];
  r.a1 = (org.y - V.org[1]) * V.isp[1];
  r.a2 = (org.z - V.org[2]) * V.isp[2];
  r.b0 = dir.x * V.isp[0];
  r.b1 = dir.y * V.isp[1];
  r.b2 = dir.z * V.isp[2];
  return r;
}
template <int SM>
__device__ __forceinline__ VFetch vox_fetch(const DVolume &V, const VoxRay &r, float t) {
  return vol_fetch_vox<SM>(V, __builtin_fmaf(t, r.b0, r.a0), __builtin_fmaf(t, r.b1, r.a1),
                           __builtin_fmaf(t, r.b2, r.a2));
}
template <int SM>
__device__ __forceinline__ float vox_sample(const DVolume &V, const VoxRay &r, float t) {
  return vol_finish_fma(vox_fetch<SM>(V, r, t));
}

__device__ __forceinline__ float vol_finish(const VFetch &F) {
  float g0 = 1.0f - F.f0, g1 = 1.0f - F.f1, g2 = 1.0f - F.f2;
  float a00 = g0 * F.p00.x + F.f0 * F.p00.y;
  float a10 = g0 * F.p10.x + F.f0 * F.p10.y;
  float a01 = g0 * F.p01.x + F.f0 * F.p01.y;
  float a11 = g0 * F.p11.x + F.f0 * F.p11.y;
  float b0 = g1 * a00 + F.f1 * a10;
  float b1 = g1 * a01 + F.f1 * a11;
  return g2 * b0 + F.f2 * b1;
}

template <int SM>
__device__ __forceinline__ float vol_sample_t(const DVolume &V, f3 p) {
  return vol_finish(vol_fetch<SM>(V, p));
}

__device__ __forceinline__ float vol_sample(const DVolume &V, f3 p) { return vol_sample_t<-1>(V, p); }

// Forward differences at one grid spacing, divided by the spacing.
__device__ __forceinline__ f3 vol_gradient(const DVolume &V, f3 p) {
  float s = vol_sample(V, p);
  float gx = vol_sample(V, add3(p, mk3(V.sp[0], 0.f, 0.f))) - s;
  float gy = vol_sample(V, add3(p, mk3(0.f, V.sp[1], 0.f))) - s;
  float gz = vol_sample(V, add3(p, mk3(0.f, 0.f, V.sp[2]))) - s;
  return mk3(gx / V.sp[0], gy / V.sp[1], gz / V.sp[2]);
}

// LinearTransferFunction lookups over tables that may live in LDS; the
// division by (hi-lo) is a multiplication by the precomputed (n-1)/(hi-lo).
// The table coordinate x = (v - lo) * (n-1)/(hi-lo) is clamped to [0, n-1]
// (one v_med3_f32), which gives the end entries outside [lo, hi]; entry
// i = min(floor(x), n-2) and its successor are blended with one FMA per
// channel (at x = n-1 the blend reaches the last entry within an ulp).
// Tables hold two entries at least (an empty or one-entry table is stored as
// two equal entries at scale 0: the same value everywhere).  A NaN value
// looks up zero colour and zero opacity, selected at the end rather than
// branched on.
__device__ __forceinline__ f3 tf_color(const float *col, const DTF &T, float v) {
  const bool nan = v != v;
  const float x = nan ? 0.f : __builtin_amdgcn_fmed3f((v - T.lo) * T.sc, 0.f, T.xc);
  int i = (int)x;
  i = i < T.nc - 2 ? i : T.nc - 2;
  const float r = x - (float)i;
  const float *c = col + __umul24((unsigned)i, 3u);
  return mk3(nan ? 0.f : lerp_fma(c[0], c[3], r), nan ? 0.f : lerp_fma(c[1], c[4], r),
             nan ? 0.f : lerp_fma(c[2], c[5], r));
}
__device__ __forceinline__ float tf_opacity(const float *op, const DTF &T, float v) {
  const bool nan = v != v;
  // (a NaN v reads entry 0, v_cvt_i32_f32 mapping NaN to 0, and the result
  // is replaced by zero)
  const float x = __builtin_amdgcn_fmed3f((v - T.lo) * T.so, 0.f, T.xo);
  int i = (int)x;
  i = i < T.no - 2 ? i : T.no - 2;
  const float o = lerp_fma(op[i], op[i + 1], x - (float)i);
  return nan ? 0.f : o;
}
// The opacity and colour lookups of v = s / 2 given s = 2 v and its offset
// d = s - 2 lo: d * (so / 2) rounds exactly as (v - lo) * so (halving and
// doubling are exact), so the march looks both up from one sum and one
// difference, without the halving.
__device__ __forceinline__ float tf_opacity_d2(const float *op, const DTF &T, float d) {
  const bool nan = d != d;
  const float x = __builtin_amdgcn_fmed3f(d * T.hso, 0.f, T.xo);
  int i = (int)x;
  i = i < T.no - 2 ? i : T.no - 2;
  const float o = lerp_fma(op[i], op[i + 1], x - (float)i);
  return nan ? 0.f : o;
}
__device__ __forceinline__ f3 tf_color_num_d2(const float *col, const DTF &T, float d) {
  const float x = __builtin_amdgcn_fmed3f(d * T.hsc, 0.f, T.xc);
  int i = (int)x;
  i = i < T.nc - 2 ? i : T.nc - 2;
  const float r = x - (float)i;
  const float *c = col + __umul24((unsigned)i, 3u);
  return mk3(lerp_fma(c[0], c[3], r), lerp_fma(c[1], c[4], r), lerp_fma(c[2], c[5], r));
}
// The colour lookup of a value known not to be NaN (the march composites a
// sample only when its opacity lookup is positive, which a NaN's is not).
__device__ __forceinline__ f3 tf_color_num(const float *col, const DTF &T, float v) {
  const float x = __builtin_amdgcn_fmed3f((v - T.lo) * T.sc, 0.f, T.xc);
  int i = (int)x;
  i = i < T.nc - 2 ? i : T.nc - 2;
  const float r = x - (float)i;
  const float *c = col + __umul24((unsigned)i, 3u);
  return mk3(lerp_fma(c[0], c[3], r), lerp_fma(c[1], c[4], r), lerp_fma(c[2], c[5], r));
}

// MyIntersectBox (TraceRays.ispc:90-106), rcp as 1/x.
__device__ __forceinline__ void box_clip(f3 org, f3 dir, const float *lo, const float *hi, float &t0, float &t1) {
  float ix = 1.0f / dir.x, iy = 1.0f / dir.y, iz = 1.0f / dir.z;
  float minx = (lo[0] - org.x) * ix, miny = (lo[1] - org.y) * iy, minz = (lo[2] - org.z) * iz;
  float maxx = (hi[0] - org.x) * ix, maxy = (hi[1] - org.y) * iy, maxz = (hi[2] - org.z) * iz;
  t0 = mxf(mnf(minx, maxx), mxf(mnf(miny, maxy), mnf(minz, maxz)));
  t1 = mnf(mxf(minx, maxx), mnf(mxf(miny, maxy), mxf(minz, maxz)));
}

// Wave-wide sum for 64-lane wavefronts.
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exclusive sum of in[0, m) into out by one workgroup of 1024 work-items
// (m <= 16 * 1024): work-item t takes the run [t * per, t * per + per) of
// per = ceil(m / 1024) ints, loads all of it before using any (one memory
// latency instead of one per element), and the 1024 run totals are scanned
// with wave shuffles and one exchange of the 16 wave totals through LDS.
// Returns the sum of all m ints (to every work-item).
__device__ __forceinline__ int scan1024_runs(const int *in, int *out, int m) {
  __shared__ int wtot[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int per = (m + 1023) >> 10;
  const int s0 = t * per;
  int v[16];
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = (k < per && s0 + k < m) ? in[s0 + k] : 0;
  int sum = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) sum += v[k];
  int x = sum;   // inclusive scan of the run totals within the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wtot[w] = x;
  __syncthreads();
  int wb = 0, total = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const int z = wtot[k];
    wb += k < w ? z : 0;
    total += z;
  }
  int base = wb + x - sum;
#pragma unroll
  for (int k = 0; k < 16; k++)
    if (k < per && s0 + k < m) {
      out[s0 + k] = base;
      base += v[k];
    }
  return total;
}
