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
#include "trace_dev.h"

namespace gxy {

// A per-item count added to *ctr: the wave's total, one 64-bit atomic per wave
// (every lane of the wave calls it)
__device__ __forceinline__ void wave_add_count(unsigned long long v, unsigned long long *ctr) {
  v = wave_sum_u64(v);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(ctr, v);
}

// WPE > 0 asks the compiler for at least WPE waves per SIMD (fewer VGPRs);
// the autotuner picks the variant.
// SM = 1 when every volume is small (32-bit voxel offsets), -1 otherwise.
// STK > 0: the BVH traversal stack holds STK entries per work-item in LDS
// (instead of 64 in private memory).
// W: BVH walked by geometry variants (2 binary, 4 four-wide, 0 per model).
// MI: isovalues per volume the iso test is unrolled for (1 when no volume
// has more than one, fewer scalar branches per sample).
// BS: work-items per workgroup (64, 128 or 256).  A workgroup's slots are
// refilled only as a whole, so smaller workgroups let a SIMD take new rays as
// soon as one wave's rays are done, at the price of staging the transfer
// function tables once per wave.
// Workgroups reach the eight XCDs (one L2 each) round robin; this renumbering
// hands each XCD runs of C consecutive workgroups' rays instead (neighbouring
// camera tiles read the same bricks), the runs still dealt round robin so the
// XCDs' shares of the image stay alike.  C = 1 is the dispatch order itself;
// the workgroups past the last whole round of 8*C keep their own numbers.
template <int C>
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x;
  if (C <= 1) return b;
  const int full = (int)gridDim.x / (8 * C) * (8 * C);
  if (b >= full) return b;
  const int x = b & 7, j = b >> 3;
  return ((j / C) * 8 + x) * C + j % C;
}

template <int NV, bool GEOM, int WPE, int SM, int STK = 0, int W = 0, int MI = GXY_MAX_ISO, int BS = GXY_BLOCK>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8)))
void k_trace(TraceParams P, RayList_ispc R, int n) {
  constexpr int NVS = NV > 0 ? NV : 1;
  __shared__ float s_col[NVS][3 * GXY_TF_MAX];
  __shared__ float s_op[NVS][GXY_TF_MAX];
  // a launch queued before its list's size was known covers the capacity:
  // workgroups past the size leave at once (before staging the TF tables)
  const int blk = xcd_block<(NV > 0 && !GEOM && STK == 0) ? 8 : 1>();
  if (P.dn) {
    n = min(n, *P.dn * P.dn_mul);
    if (blk * BS >= n) return;
  }
#pragma unroll
  for (int v = 0; v < NV; v++) {
    for (int k = threadIdx.x; k < 3 * P.vv[v].tf.nc; k += BS) s_col[v][k] = P.vv[v].tf.col[k];
    for (int k = threadIdx.x; k < P.vv[v].tf.no; k += BS) s_op[v][k] = P.vv[v].tf.op[k];
  }
  if (NV > 0) __syncthreads();

  const int i = blk * BS + threadIdx.x;
  unsigned long long nsamp = 0;
  if (STK > 0) {
    __shared__ int s_stack[STK > 0 ? STK : 1][BS];
    LdsStack<(STK > 0 ? STK : 1)> stack;
    stack.base = &s_stack[0][threadIdx.x];
    stack.stride = BS;
    if (i < n) nsamp = trace_one<NV, GEOM, LdsStack<(STK > 0 ? STK : 1)>, SM, W, MI>(P, R, i, s_col, s_op, stack);
  } else if (NV > 0 && !GEOM) {
    // volumes only: a wave whose rays are all primary, or all not, takes the
    // march compiled for that kind (lists of one kind are the usual case)
    PrivStack<GXY_BVH_STACK> stack;
    if (i < n) {
      const unsigned long long pm = __builtin_amdgcn_ballot_w64(gptr(R.type)[i] == RAY_PRIMARY);
      const unsigned long long am = __builtin_amdgcn_ballot_w64(true);
      if (pm == am) nsamp = trace_one<NV, GEOM, PrivStack<GXY_BVH_STACK>, SM, W, MI, 1>(P, R, i, s_col, s_op, stack);
      else if (pm == 0) nsamp = trace_one<NV, GEOM, PrivStack<GXY_BVH_STACK>, SM, W, MI, 2>(P, R, i, s_col, s_op, stack);
      else nsamp = trace_one<NV, GEOM, PrivStack<GXY_BVH_STACK>, SM, W, MI>(P, R, i, s_col, s_op, stack);
    }
  } else {
    PrivStack<GXY_BVH_STACK> stack;
    if (i < n) nsamp = trace_one<NV, GEOM, PrivStack<GXY_BVH_STACK>, SM, W, MI>(P, R, i, s_col, s_op, stack);
  }
  if (P.sample_count) wave_add_count(nsamp, P.sample_count);
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
  // the arithmetic of the one-ray-per-item march (voxel-coordinate positions)
#pragma unroll
  for (int v = 0; v < NV; v++) sThis[v] = vox_sample<-1>(P.vv[v].vol, vox_ray(P.vv[v].vol, m.org, m.dir), tThis);
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
#pragma unroll
      for (int k = 0; k < GXY_MAX_ISO; k++) {
        if (k < ni) {
          float isoval = P.vv[major].iso[k];
          if (((isoval >= sl) && (isoval < st)) || ((isoval <= sl) && (isoval > st))) {
            ih = true;
            ivid = major;
            ht = m.tLast + ((isoval - sl) / (st - sl)) * (tThis - m.tLast);
            hs = isoval;
          }
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
#pragma unroll
      for (int v = 0; v < NV; v++) sThis[v] = vox_sample<-1>(P.vv[v].vol, vox_ray(P.vv[v].vol, m.org, m.dir), tThis);
      nsamp += NV;
    }
#pragma unroll
    for (int major = 0; major < NV; major++) {
      if (P.vv[major].dvr) {
        const DTF &T = P.vv[major].tf;
        float sVolume = (sLast[major] + sThis[major]) / 2;
        float sop = tf_opacity(s_op[major], T, sVolume);
        if (sop > 0) {
          if (m.flags & 1) {
            f3 sc = tf_color_num(s_col[major], T, sVolume);
            composite(m.cr, m.cg, m.cb, m.ca, sc, clamp01f(sop * P.vv[major].vol.irate));
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
  if (P.sample_count) wave_add_count(nsamp, P.sample_count);
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
template <int NV, int SM>
__global__ __launch_bounds__(GXY_BLOCK) void k_sampler(TraceParams P, RayList_ispc R, int n) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  unsigned long long nsamp = 0;
  if (i < n) nsamp = sample_one<NV, SM>(P, R, i);
  if (P.sample_count) wave_add_count(nsamp, P.sample_count);
}

// ------------------------------------------------------------- launchers
// one volume with a bricked copy: the variant for its addressing (32-bit
// offsets when the copy spans under 4 GiB) and the occupancy request
template <int SM>
static void launch_trace_sm(const TraceParams &P, const RayList_ispc &R, int n, int wpe, int bs, dim3 g, dim3 b,
                            hipStream_t s) {
  const bool one_iso = P.vv[0].nIso <= 1;
  if (one_iso && wpe == 6 && bs != 256) {
    dim3 gs((n + bs - 1) / bs), bb(bs);
    if (bs == 64) hipLaunchKernelGGL((k_trace<1, false, 6, SM, 0, 0, 1, 64>), gs, bb, 0, s, P, R, n);
    else hipLaunchKernelGGL((k_trace<1, false, 6, SM, 0, 0, 1, 128>), gs, bb, 0, s, P, R, n);
    return;
  }
  if (wpe == 5 && one_iso) hipLaunchKernelGGL((k_trace<1, false, 5, SM, 0, 0, 1>), g, b, 0, s, P, R, n);
  else if (wpe == 5) hipLaunchKernelGGL((k_trace<1, false, 5, SM>), g, b, 0, s, P, R, n);
  else if (wpe == 6 && one_iso) hipLaunchKernelGGL((k_trace<1, false, 6, SM, 0, 0, 1>), g, b, 0, s, P, R, n);
  else if (wpe == 7 && one_iso) hipLaunchKernelGGL((k_trace<1, false, 7, SM, 0, 0, 1>), g, b, 0, s, P, R, n);
  else if (wpe == 8 && one_iso) hipLaunchKernelGGL((k_trace<1, false, 8, SM, 0, 0, 1>), g, b, 0, s, P, R, n);
  else if (wpe >= 6) hipLaunchKernelGGL((k_trace<1, false, 6, SM>), g, b, 0, s, P, R, n);
  else if (one_iso) hipLaunchKernelGGL((k_trace<1, false, 0, SM, 0, 0, 1>), g, b, 0, s, P, R, n);
  else hipLaunchKernelGGL((k_trace<1, false, 0, SM>), g, b, 0, s, P, R, n);
}
static void launch_trace_bricked(const TraceParams &P, const RayList_ispc &R, int n, int wpe, int bs, int sm, dim3 g,
                                 dim3 b, hipStream_t s) {
  switch (sm) {
    case 4: launch_trace_sm<4>(P, R, n, wpe, bs, g, b, s); break;
    case 8: launch_trace_sm<8>(P, R, n, wpe, bs, g, b, s); break;
    case 12: launch_trace_sm<12>(P, R, n, wpe, bs, g, b, s); break;
    default: launch_trace_sm<16>(P, R, n, wpe, bs, g, b, s); break;
  }
}

// the workgroup size launch_trace uses for these parameters
int trace_block_size(const TraceParams &P, int wpe, int bs) {
  return (P.nvv == 1 && !P.has_model && P.vv[0].vol.brick && P.vv[0].nIso <= 1 && wpe == 6 && (bs == 64 || bs == 128))
             ? bs
             : GXY_BLOCK;
}

template <int NV>
static hipError_t launch_trace_nv(const TraceParams &P, const RayList_ispc &R, int n, int wpe, int stack,
                                  hipStream_t s, int gwaves, int bs) {
  dim3 g((n + GXY_BLOCK - 1) / GXY_BLOCK), b(GXY_BLOCK);
  bool small = true;
  for (int v = 0; v < NV; v++) small = small && P.vv[v].vol.small;
  // an LDS stack only when it holds the deepest path of the model
  if (NV == 0 && P.has_model && stack > 0 && (P.model.wide ? P.model.depth4 : P.model.depth) + 2 > stack) stack = 0;
  const bool wide = P.model.wide != 0;
  if (NV == 0 && P.has_model && P.model.ncurve > 0) {
    // path lines: the variants that trace curve pieces too (private stack)
    if (wide) hipLaunchKernelGGL((k_trace<0, true, 0, -1, 0, 20>), g, b, 0, s, P, R, n);
    else if (gwaves == 8) hipLaunchKernelGGL((k_trace<0, true, 8, -1, 0, 18>), g, b, 0, s, P, R, n);
    else hipLaunchKernelGGL((k_trace<0, true, 0, -1, 0, 18>), g, b, 0, s, P, R, n);
  } else if (NV == 0 && P.has_model && wide) {
    if (stack == 16) hipLaunchKernelGGL((k_trace<0, true, 0, -1, 16, 4>), g, b, 0, s, P, R, n);
    else if (stack == 24) hipLaunchKernelGGL((k_trace<0, true, 0, -1, 24, 4>), g, b, 0, s, P, R, n);
    else if (stack == 32) hipLaunchKernelGGL((k_trace<0, true, 0, -1, 32, 4>), g, b, 0, s, P, R, n);
    else hipLaunchKernelGGL((k_trace<0, true, 0, -1, 0, 4>), g, b, 0, s, P, R, n);
  } else if (NV == 0 && P.has_model) {
    if (stack == 16) hipLaunchKernelGGL((k_trace<0, true, 0, -1, 16, 2>), g, b, 0, s, P, R, n);
    else if (stack == 24) hipLaunchKernelGGL((k_trace<0, true, 0, -1, 24, 2>), g, b, 0, s, P, R, n);
    else if (stack == 32) hipLaunchKernelGGL((k_trace<0, true, 0, -1, 32, 2>), g, b, 0, s, P, R, n);
    else if (gwaves == 8) hipLaunchKernelGGL((k_trace<0, true, 8, -1, 0, 2>), g, b, 0, s, P, R, n);
    else hipLaunchKernelGGL((k_trace<0, true, 0, -1, 0, 2>), g, b, 0, s, P, R, n);
  } else if (P.has_model) hipLaunchKernelGGL((k_trace<NV, true, 0, -1>), g, b, 0, s, P, R, n);
  else if (NV == 1 && P.vv[0].vol.brick) launch_trace_bricked(P, R, n, wpe, trace_block_size(P, wpe, bs), sm_of(P.vv[0].vol), g, b, s);
  else if (small && wpe == 5) hipLaunchKernelGGL((k_trace<NV, false, 5, 1>), g, b, 0, s, P, R, n);
  else if (small && wpe >= 6) hipLaunchKernelGGL((k_trace<NV, false, 6, 1>), g, b, 0, s, P, R, n);
  else if (small) hipLaunchKernelGGL((k_trace<NV, false, 0, 1>), g, b, 0, s, P, R, n);
  else if (wpe == 5) hipLaunchKernelGGL((k_trace<NV, false, 5, -1>), g, b, 0, s, P, R, n);
  else if (wpe >= 6) hipLaunchKernelGGL((k_trace<NV, false, 6, -1>), g, b, 0, s, P, R, n);
  else hipLaunchKernelGGL((k_trace<NV, false, 0, -1>), g, b, 0, s, P, R, n);
  return hipGetLastError();
}

hipError_t launch_trace(const TraceParams &P, const RayList_ispc &R, int n, int wpe, int stack, hipStream_t s,
                        int gwaves, int bs) {
  if (n <= 0) return hipSuccess;
  switch (P.nvv) {
    case 0: return launch_trace_nv<0>(P, R, n, wpe, stack, s, gwaves, bs);
    case 1: return launch_trace_nv<1>(P, R, n, wpe, stack, s, gwaves, bs);
    case 2: return launch_trace_nv<2>(P, R, n, wpe, stack, s, gwaves, bs);
    case 3: return launch_trace_nv<3>(P, R, n, wpe, stack, s, gwaves, bs);
    case 4: return launch_trace_nv<4>(P, R, n, wpe, stack, s, gwaves, bs);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_sampler(const TraceParams &P, const RayList_ispc &R, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  dim3 g((n + GXY_BLOCK - 1) / GXY_BLOCK), b(GXY_BLOCK);
  bool small = P.nvv > 0;
  for (int v = 0; v < P.nvv && v < GXY_MAX_VOLVIS; v++) small = small && P.vv[v].vol.small;
  if (P.nvv == 1 && P.vv[0].vol.brick) {
    switch (sm_of(P.vv[0].vol)) {
      case 4: hipLaunchKernelGGL((k_sampler<1, 4>), g, b, 0, s, P, R, n); break;
      case 8: hipLaunchKernelGGL((k_sampler<1, 8>), g, b, 0, s, P, R, n); break;
      case 12: hipLaunchKernelGGL((k_sampler<1, 12>), g, b, 0, s, P, R, n); break;
      default: hipLaunchKernelGGL((k_sampler<1, 16>), g, b, 0, s, P, R, n); break;
    }
    return hipGetLastError();
  }
  switch (P.nvv * 2 + (small ? 1 : 0)) {
    case 2: hipLaunchKernelGGL((k_sampler<1, -1>), g, b, 0, s, P, R, n); break;
    case 3: hipLaunchKernelGGL((k_sampler<1, 1>), g, b, 0, s, P, R, n); break;
    case 4: hipLaunchKernelGGL((k_sampler<2, -1>), g, b, 0, s, P, R, n); break;
    case 5: hipLaunchKernelGGL((k_sampler<2, 1>), g, b, 0, s, P, R, n); break;
    case 6: hipLaunchKernelGGL((k_sampler<3, -1>), g, b, 0, s, P, R, n); break;
    case 7: hipLaunchKernelGGL((k_sampler<3, 1>), g, b, 0, s, P, R, n); break;
    case 8: hipLaunchKernelGGL((k_sampler<4, -1>), g, b, 0, s, P, R, n); break;
    case 9: hipLaunchKernelGGL((k_sampler<4, 1>), g, b, 0, s, P, R, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gxy
