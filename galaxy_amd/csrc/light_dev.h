// light_dev.h — per-parent lighting and secondary-ray generation shared by the
// lighting kernel (spawn.hip) and the batched pass kernels (batch.hip).
#pragma once

#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

__device__ __forceinline__ bool is_parent(const RayList_ispc &R, int i) {
  return R.type[i] == RAY_PRIMARY && (R.term[i] & RAY_SURFACE);
}

// One parent of TraceRays::Trace's lighting sequence under reverse lighting
// (TraceRays.cpp:124-131): ambientLighting -> generateAORays -> diffuseLighting
// -> generateShadowRays (TraceRays.ispc:640-921), applied to ray i of R for the
// LM_* bits in mode.  AO children go to O[ao_base .. ao_base + nAO), shadow
// children to O[shoff .. shoff + nLights) (shoff -1: none).
__device__ __forceinline__ void light_one(const LightParams &L, int mode, const RayList_ispc &R, int i, int ao_base,
                                          int shoff, const RayList_ispc &O) {
  const bool parent = is_parent(R, i);
  if (!parent && shoff == -1) return;
  f3 nrm = mk3(R.nx[i], R.ny[i], R.nz[i]);
  float r = R.r[i], g = R.g[i], b = R.b[i], o = R.o[i];
  const float sr = R.sr[i], sg = R.sg[i], sb = R.sb[i];
  const float ox0 = R.ox[i], oy0 = R.oy[i], oz0 = R.oz[i];
  const float dx0 = R.dx[i], dy0 = R.dy[i], dz0 = R.dz[i];
  const float t = R.t[i];
  const int px = R.x[i], py = R.y[i];

  // TraceRays_ambientLighting (TraceRays.ispc:746-759)
  if ((mode & LM_AMBIENT) && parent) {
    float s = L.Ka * (1.0f - o);
    r += s * sr;
    g += s * sg;
    b += s * sb;
  }
  // TraceRays_generateAORays (TraceRays.ispc:640-730), reverse lighting
  if ((mode & LM_AO) && parent && L.nAO > 0) {
    const float Ka = -L.Ka / L.nAO;
    float as = Ka * (1.0f - o);
    float ar = as * sr, ag = as * sg, ab = as * sb;
    f3 b0 = mk3(1.f, 0.f, 0.f);
    if (fabsf(dot3(b0, nrm)) > 0.95f) b0 = mk3(0.f, 1.f, 0.f);
    f3 b1 = normalize3(cross3(b0, nrm));
    b0 = normalize3(cross3(b1, nrm));
    float ox = ox0 + t * dx0, oy = oy0 + t * dy0, oz = oz0 + t * dz0;
    ox = ox + L.eps * nrm.x;
    oy = oy + L.eps * nrm.y;
    oz = oz + L.eps * nrm.z;
    const int base = ao_base;
    for (int j = 0; j < L.nAO; j++) {
      int off = base + j;
      int rr = ((px * 9949 + py * 9613 + j * 9151) >> 8) & 0xff;
      float r0 = L.U[rr], r1 = L.V[rr];
      float w = sqrtf(1.f - r1);
      float x = cosf(6.2831855f * r0) * w;
      float y = sinf(6.2831855f * r0) * w;
      float z = sqrtf(r1) + L.eps;
      f3 rd = add3(add3(scl3(x, b0), scl3(y, b1)), scl3(z, nrm));
      O.ox[off] = ox; O.oy[off] = oy; O.oz[off] = oz;
      O.dx[off] = rd.x; O.dy[off] = rd.y; O.dz[off] = rd.z;
      O.r[off] = ar; O.g[off] = ag; O.b[off] = ab; O.o[off] = 0.f;
      O.t[off] = 0.f; O.tMax[off] = L.aoRadius;
      O.x[off] = px; O.y[off] = py;
      O.type[off] = RAY_AO; O.term[off] = 0;
    }
  }
  // TraceRays_diffuseLighting (TraceRays.ispc:868-921)
  if ((mode & LM_DIFFUSE) && parent) {
    const float Kd = L.Kd / L.nLights;
    float tr = 0, tg = 0, tb = 0;
    for (int k = 0; k < L.nLights; k++) {
      f3 Lk = mk3(L.L[3 * k], L.L[3 * k + 1], L.L[3 * k + 2]);
      f3 lv;
      if (L.types[k]) {
        f3 p = mk3(ox0 + t * dx0, oy0 + t * dy0, oz0 + t * dz0);
        lv = safe_normalize3(sub3(Lk, p));
      } else {
        lv = neg3(Lk);
      }
      float d = dot3(nrm, lv);
      if (d > 0) {
        float dff = (1.0f - o) * d;
        tr += dff * sr;
        tg += dff * sg;
        tb += dff * sb;
      }
    }
    r = r + Kd * (1 - o) * tr;
    g = g + Kd * (1 - o) * tg;
    b = b + Kd * (1 - o) * tb;
    o = o + Kd * (1 - o) * o;
  }
  // TraceRays_generateShadowRays (TraceRays.ispc:778-855), reverse lighting
  if ((mode & LM_SHADOW) && shoff != -1) {
    const float Kd = -L.Kd / L.nLights;
    int off = shoff;
    for (int k = 0; k < L.nLights; k++) {
      f3 sp = mk3(ox0 + t * dx0 + L.eps * nrm.x, oy0 + t * dy0 + L.eps * nrm.y, oz0 + t * dz0 + L.eps * nrm.z);
      f3 Lk = mk3(L.L[3 * k], L.L[3 * k + 1], L.L[3 * k + 2]);
      f3 lv = L.types[k] ? safe_normalize3(sub3(Lk, sp)) : neg3(Lk);
      lv = safe_normalize3(lv);
      float d = dot3(nrm, lv);
      if (d < 0) d = 0;
      float dff = (1.0f - o) * Kd * d;
      O.ox[off] = sp.x; O.oy[off] = sp.y; O.oz[off] = sp.z;
      O.dx[off] = lv.x; O.dy[off] = lv.y; O.dz[off] = lv.z;
      O.r[off] = dff * sr; O.g[off] = dff * sg; O.b[off] = dff * sb; O.o[off] = 0.f;
      O.t[off] = 0.f; O.tMax[off] = __builtin_inff();
      O.x[off] = px; O.y[off] = py;
      O.type[off] = RAY_SHADOW; O.term[off] = 0;
      off++;
    }
  }
  if ((mode & (LM_AMBIENT | LM_DIFFUSE)) && parent) {
    R.r[i] = r; R.g[i] = g; R.b[i] = b; R.o[i] = o;
  }
}

}  // namespace gxy
