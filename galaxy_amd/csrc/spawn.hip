// spawn.hip — secondary-ray bookkeeping and lighting on gfx950.
//
// TraceRays.cpp:89-117 computes, serially on the host, exclusive offsets of the
// AO block (nAO slots per primary surface ray) followed by the shadow block
// (nLights slots each).  Since every parent contributes 0 or a fixed count,
// both offset arrays follow from one exclusive rank of the "primary & surface"
// flag: a wave ballot + popcount gives the rank inside a wave, LDS combines the
// four waves of a workgroup, and a single-workgroup scan over workgroup totals
// gives the global base.  The lighting kernel then applies, per parent and in
// the reference order, ambientLighting -> generateAORays -> diffuseLighting ->
// generateShadowRays (TraceRays.cpp:124-131, TraceRays.ispc:625-923).
#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

__constant__ float c_randomU[256];
__constant__ float c_randomV[256];

hipError_t upload_uv_tables(const float *U, const float *V) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_randomU), U, 256 * sizeof(float));
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_randomV), V, 256 * sizeof(float));
}

__device__ __forceinline__ bool is_parent(const RayList_ispc &R, int i) {
  return R.type[i] == RAY_PRIMARY && (R.term[i] & RAY_SURFACE);
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Per-workgroup count of flagged rays.
__global__ __launch_bounds__(GXY_BLOCK) void k_flag_count(RayList_ispc R, int n, int *block_counts) {
  __shared__ int wsum[GXY_BLOCK / 64];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = i < n && is_parent(R, i);
  unsigned long long m = __ballot(f);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < GXY_BLOCK / 64; w++) s += wsum[w];
    block_counts[blockIdx.x] = s;
  }
}

// Exclusive scan of nb ints in place (single workgroup, 1024 work-items);
// total written to *total.
__global__ __launch_bounds__(1024) void k_scan_single(int *a, int nb, int *total) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int s0 = t * per, s1 = min(nb, s0 + per);
  int sum = 0;
  for (int k = s0; k < s1; k++) sum += a[k];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int base = t > 0 ? part[t - 1] : 0;
  for (int k = s0; k < s1; k++) {
    int v = a[k];
    a[k] = base;
    base += v;
  }
  if (t == 1023) *total = part[1023];
}

__global__ __launch_bounds__(GXY_BLOCK) void k_flag_rank(RayList_ispc R, int n, const int *block_base, int *rank) {
  __shared__ int wsum[GXY_BLOCK / 64];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = i < n && is_parent(R, i);
  unsigned long long m = __ballot(f);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wsum[w] = __popcll(m);
  __syncthreads();
  int base = block_base[blockIdx.x];
  for (int k = 0; k < w; k++) base += wsum[k];
  // exclusive count of flagged rays before i (defined for every ray)
  if (i < n) rank[i] = base + __popcll(m & lanemask_lt());
}

hipError_t launch_surface_scan(const RayList_ispc &R, int n, int *scratch, int *rank_out, int *total,
                               hipStream_t s) {
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  if (nb == 0) {
    (void)hipMemsetAsync(total, 0, sizeof(int), s);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, scratch);
  hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, scratch, nb, total);
  hipLaunchKernelGGL(k_flag_rank, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, scratch, rank_out);
  return hipGetLastError();
}

__global__ __launch_bounds__(GXY_BLOCK) void k_offsets(RayList_ispc R, int n, const int *rank, const int *total,
                                                       int nAO, int nL, int doShadows, int *ao_off, int *sh_off) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= n) return;
  int r = rank[i];
  int T = *total;
  // every ray gets the running AO count (TraceRays.cpp:90-99); only parents
  // get a shadow slot (TraceRays.cpp:103-115)
  if (ao_off) ao_off[i] = nAO * r;
  if (sh_off) sh_off[i] = (doShadows && is_parent(R, i)) ? nAO * T + nL * r : -1;
}

hipError_t launch_offsets_from_rank(const RayList_ispc &R, int n, const int *rank, const int *total, int nAO,
                                    int nLights, int doShadows, int *ao_off, int *sh_off, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_offsets, dim3((n + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, R, n, rank, total, nAO,
                     nLights, doShadows, ao_off, sh_off);
  return hipGetLastError();
}

// ------------------------------------------------------------------ lighting
__global__ __launch_bounds__(GXY_BLOCK) void k_lighting(LightParams L, int mode, RayList_ispc R, int n,
                                                        const int *ao_off, const int *sh_off, RayList_ispc O) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= n) return;
  const bool parent = is_parent(R, i);
  int shoff = (mode & LM_SHADOW) ? sh_off[i] : -1;
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
    const int base = ao_off[i];
    for (int j = 0; j < L.nAO; j++) {
      int off = base + j;
      int rr = ((px * 9949 + py * 9613 + j * 9151) >> 8) & 0xff;
      float r0 = c_randomU[rr], r1 = c_randomV[rr];
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

hipError_t launch_lighting(const LightParams &L, int mode, const RayList_ispc &R, int n, const int *ao_off,
                           const int *sh_off, const RayList_ispc &O, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lighting, dim3((n + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, L, mode, R, n, ao_off,
                     sh_off, O);
  return hipGetLastError();
}

}  // namespace gxy
