// light_dev.h — per-parent lighting and secondary-ray generation shared by the
// lighting kernel (spawn.hip) and the batched pass kernels (batch.hip).
#pragma once

#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

__device__ __forceinline__ bool is_parent(const RayList_ispc &R, int i) {
  return R.type[i] == RAY_PRIMARY && (R.term[i] & RAY_SURFACE);
}

// A parent ray's fields as the lighting sequence reads them.
struct ParentRay {
  f3 nrm;
  float r, g, b, o;
  float sr, sg, sb;
  float ox0, oy0, oz0, dx0, dy0, dz0, t;
  int px, py;
};

__device__ __forceinline__ ParentRay parent_load(const RayList_ispc &R, int i) {
  ParentRay P;
  P.nrm = mk3(R.nx[i], R.ny[i], R.nz[i]);
  P.r = R.r[i]; P.g = R.g[i]; P.b = R.b[i]; P.o = R.o[i];
  P.sr = R.sr[i]; P.sg = R.sg[i]; P.sb = R.sb[i];
  P.ox0 = R.ox[i]; P.oy0 = R.oy[i]; P.oz0 = R.oz[i];
  P.dx0 = R.dx[i]; P.dy0 = R.dy[i]; P.dz0 = R.dz[i];
  P.t = R.t[i];
  P.px = R.x[i]; P.py = R.y[i];
  return P;
}

// TraceRays_ambientLighting (TraceRays.ispc:746-759)
__device__ __forceinline__ void light_ambient(const LightParams &L, ParentRay &P) {
  float s = L.Ka * (1.0f - P.o);
  P.r += s * P.sr;
  P.g += s * P.sg;
  P.b += s * P.sb;
}

// TraceRays_generateAORays (TraceRays.ispc:640-730): AO child j of P (P.o
// before diffuseLighting), written to O[off]; its colour is negative under
// reverse lighting (:639-643).
// PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically
// Good Algorithms for Random Number Generation", 2014; pcg32_random_r and
// pcg32_srandom_r): a 64-bit LCG state, output by xorshift-high and a random
// rotation.
struct Pcg32 {
  unsigned long long state, inc;
};
__device__ __forceinline__ unsigned pcg32_next(Pcg32 &g) {
  const unsigned long long old = g.state;
  g.state = old * 6364136223846793005ull + g.inc;
  const unsigned xs = (unsigned)(((old >> 18u) ^ old) >> 27u);
  const unsigned rot = (unsigned)(old >> 59u);
  return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
__device__ __forceinline__ Pcg32 pcg32_seeded(unsigned long long initstate, unsigned long long initseq) {
  Pcg32 g;
  g.state = 0ull;
  g.inc = (initseq << 1u) | 1u;
  pcg32_next(g);
  g.state += initstate;
  pcg32_next(g);
  return g;
}
// The (U, V) pair of AO ray j of the pixel (px, py): the stream selected by the
// pixel, started at (seed, j); each value is the top 24 bits of an output
// over 2^24 (exact in FP32, in [0, 1)).
__device__ __forceinline__ void ao_pcg_uv(unsigned seed, int px, int py, int j, float &u, float &v) {
  Pcg32 g = pcg32_seeded(((unsigned long long)seed << 32) | (unsigned)j,
                         ((unsigned long long)(unsigned)py << 32) | (unsigned)px);
  u = (float)(pcg32_next(g) >> 8) * (1.0f / 16777216.0f);
  v = (float)(pcg32_next(g) >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ void light_ao_child(const LightParams &L, const ParentRay &P, int j, const RayList_ispc &O,
                                               int off) {
  const float Ka = (L.reverse ? -L.Ka : L.Ka) / L.nAO;
  float as = Ka * (1.0f - P.o);
  float ar = as * P.sr, ag = as * P.sg, ab = as * P.sb;
  const f3 nrm = P.nrm;
  f3 b0 = mk3(1.f, 0.f, 0.f);
  if (fabsf(dot3(b0, nrm)) > 0.95f) b0 = mk3(0.f, 1.f, 0.f);
  f3 b1 = normalize3(cross3(b0, nrm));
  b0 = normalize3(cross3(b1, nrm));
  float ox = P.ox0 + P.t * P.dx0, oy = P.oy0 + P.t * P.dy0, oz = P.oz0 + P.t * P.dz0;
  ox = ox + L.eps * nrm.x;
  oy = oy + L.eps * nrm.y;
  oz = oz + L.eps * nrm.z;
  float r0, r1;
  if (L.ao_pcg) {
    ao_pcg_uv(L.ao_seed, P.px, P.py, j, r0, r1);
  } else {
    int rr = ((P.px * 9949 + P.py * 9613 + j * 9151) >> 8) & 0xff;
    r0 = L.U[rr];
    r1 = L.V[rr];
  }
  float w = sqrtf(1.f - r1);
  float x = cosf(6.2831855f * r0) * w;
  float y = sinf(6.2831855f * r0) * w;
  float z = sqrtf(r1) + L.eps;
  f3 rd = add3(add3(scl3(x, b0), scl3(y, b1)), scl3(z, nrm));
  O.ox[off] = ox; O.oy[off] = oy; O.oz[off] = oz;
  O.dx[off] = rd.x; O.dy[off] = rd.y; O.dz[off] = rd.z;
  O.r[off] = ar; O.g[off] = ag; O.b[off] = ab; O.o[off] = 0.f;
  O.t[off] = 0.f; O.tMax[off] = L.aoRadius;
  O.x[off] = P.px; O.y[off] = P.py;
  O.type[off] = RAY_AO; O.term[off] = 0;
}

// TraceRays_diffuseLighting (TraceRays.ispc:868-921)
__device__ __forceinline__ void light_diffuse(const LightParams &L, ParentRay &P) {
  const float Kd = L.Kd / L.nLights;
  const f3 nrm = P.nrm;
  const float o = P.o;
  float tr = 0, tg = 0, tb = 0;
  for (int k = 0; k < L.nLights; k++) {
    f3 Lk = mk3(L.L[3 * k], L.L[3 * k + 1], L.L[3 * k + 2]);
    f3 lv;
    if (L.types[k]) {
      f3 p = mk3(P.ox0 + P.t * P.dx0, P.oy0 + P.t * P.dy0, P.oz0 + P.t * P.dz0);
      lv = safe_normalize3(sub3(Lk, p));
    } else {
      lv = neg3(Lk);
    }
    float d = dot3(nrm, lv);
    if (d > 0) {
      float dff = (1.0f - o) * d;
      tr += dff * P.sr;
      tg += dff * P.sg;
      tb += dff * P.sb;
    }
  }
  P.r = P.r + Kd * (1 - o) * tr;
  P.g = P.g + Kd * (1 - o) * tg;
  P.b = P.b + Kd * (1 - o) * tb;
  P.o = o + Kd * (1 - o) * o;
}

// the parent's alpha after diffuseLighting (the value shadow rays read)
__device__ __forceinline__ float light_diffused_alpha(const LightParams &L, float o) {
  const float Kd = L.Kd / L.nLights;
  return o + Kd * (1 - o) * o;
}

// TraceRays_generateShadowRays (TraceRays.ispc:778-855): shadow child k of P
// for parent alpha o (after diffuseLighting under reverse lighting, which also
// makes its colour negative, :778-782), to O[off].
__device__ __forceinline__ void light_shadow_child(const LightParams &L, const ParentRay &P, float o, int k,
                                                   const RayList_ispc &O, int off) {
  const float Kd = (L.reverse ? -L.Kd : L.Kd) / L.nLights;
  const f3 nrm = P.nrm;
  f3 sp = mk3(P.ox0 + P.t * P.dx0 + L.eps * nrm.x, P.oy0 + P.t * P.dy0 + L.eps * nrm.y,
              P.oz0 + P.t * P.dz0 + L.eps * nrm.z);
  f3 Lk = mk3(L.L[3 * k], L.L[3 * k + 1], L.L[3 * k + 2]);
  f3 lv = L.types[k] ? safe_normalize3(sub3(Lk, sp)) : neg3(Lk);
  lv = safe_normalize3(lv);
  float d = dot3(nrm, lv);
  if (d < 0) d = 0;
  float dff = (1.0f - o) * Kd * d;
  O.ox[off] = sp.x; O.oy[off] = sp.y; O.oz[off] = sp.z;
  O.dx[off] = lv.x; O.dy[off] = lv.y; O.dz[off] = lv.z;
  O.r[off] = dff * P.sr; O.g[off] = dff * P.sg; O.b[off] = dff * P.sb; O.o[off] = 0.f;
  O.t[off] = 0.f; O.tMax[off] = __builtin_inff();
  O.x[off] = P.px; O.y[off] = P.py;
  O.type[off] = RAY_SHADOW; O.term[off] = 0;
}

// One parent of TraceRays::Trace's lighting sequence under reverse lighting
// (TraceRays.cpp:124-131): ambientLighting -> generateAORays -> diffuseLighting
// -> generateShadowRays, applied to ray i of R for the LM_* bits in mode.  AO
// children go to O[ao_base .. ao_base + nAO), shadow children to
// O[shoff .. shoff + nLights) (shoff -1: none).
__device__ __forceinline__ void light_one(const LightParams &L, int mode, const RayList_ispc &R, int i, int ao_base,
                                          int shoff, const RayList_ispc &O) {
  const bool parent = is_parent(R, i);
  if (!parent && shoff == -1) return;
  ParentRay P = parent_load(R, i);
  if ((mode & LM_AMBIENT) && parent) light_ambient(L, P);
  if ((mode & LM_AO) && parent && L.nAO > 0)
    for (int j = 0; j < L.nAO; j++) light_ao_child(L, P, j, O, ao_base + j);
  if ((mode & LM_DIFFUSE) && parent) light_diffuse(L, P);
  if ((mode & LM_SHADOW) && shoff != -1)
    for (int k = 0; k < L.nLights; k++) light_shadow_child(L, P, P.o, k, O, shoff + k);
  if ((mode & (LM_AMBIENT | LM_DIFFUSE)) && parent) {
    R.r[i] = P.r; R.g[i] = P.g; R.b[i] = P.b; R.o[i] = P.o;
  }
}

}  // namespace gxy
