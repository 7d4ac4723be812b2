// batch.hip — batched passes: many uneven ray lists over many resident
// datasets, traced by one persistent launch per pass.
//
// A pass reads a device chunk table (list, source, start, count; at most CH
// rays per chunk, every ray of a chunk belongs to one list and one dataset)
// whose length lives in device memory.  The trace kernel is persistent: each
// wave pops the next chunk from a device work queue (one atomic per chunk),
// stages the dataset's transfer-function tables in its own LDS slice when the
// dataset changes, and runs the per-ray trace (TraceRays_TraceRays,
// SamplerTraceRays) on up to CH rays with its 64 lanes.  The post-pass then
// applies, per chunk, the reference's lighting sequence and spawns secondary
// rays (TraceRays.cpp:68-146), classifies (Renderer.cpp:304-454, with no
// neighbouring partition), accumulates terminated rays into the list's
// framebuffer (Renderer.cpp:456-502) or appends sampler particles
// (Sampler.cpp:51-89), and compacts spawned and kept rays into the next pass's
// ray pool with wave ballots; chunk offsets come from a device-count scan, so
// no pass needs the host, and a fixed sequence of passes can be captured in a
// HIP graph.  All kernels here are launched with fixed grids and read their
// work counts from device memory.
#include "light_dev.h"
#include "route_dev.h"
#include "trace_dev.h"

namespace gxy {

#define BWAVES (GXY_BLOCK / 64)
// isovalue crossings a sampler ray follows within one pass before it is
// re-queued as a KEEP_HERE ray

__device__ __forceinline__ unsigned long long bl_lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ int first_lane_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// LDS writes of one wave become visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Reads of tables that no kernel of the pass writes while it runs, through the
// constant address space: a wave-uniform address then becomes a scalar load
// into SGPRs instead of a vector load into every lane's VGPRs.
// (The address goes through readfirstlane and an integer so that the cast is
// not folded back into the generic pointer it came from.)
template <class T>
__device__ __forceinline__ const T &kconst(const T *p) {
  const unsigned long long a = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  const __attribute__((address_space(4))) T *q =
      (const __attribute__((address_space(4))) T *)(((unsigned long long)hi << 32) | lo);
  return *(const T *)q;
}

// a RayList whose 25 array pointers are known to be global
__device__ __forceinline__ RayList_ispc rl_global(const RayList_ispc &S) {
  RayList_ispc R;
  R.ox = gptr(S.ox); R.oy = gptr(S.oy); R.oz = gptr(S.oz);
  R.dx = gptr(S.dx); R.dy = gptr(S.dy); R.dz = gptr(S.dz);
  R.nx = gptr(S.nx); R.ny = gptr(S.ny); R.nz = gptr(S.nz);
  R.sample = gptr(S.sample);
  R.r = gptr(S.r); R.g = gptr(S.g); R.b = gptr(S.b); R.o = gptr(S.o);
  R.sr = gptr(S.sr); R.sg = gptr(S.sg); R.sb = gptr(S.sb); R.so = gptr(S.so);
  R.t = gptr(S.t); R.tMax = gptr(S.tMax);
  R.x = gptr(S.x); R.y = gptr(S.y);
  R.type = gptr(S.type); R.term = gptr(S.term); R.classification = gptr(S.classification);
  return R;
}

__device__ __forceinline__ RayList_ispc chunk_rays(const BatchArgs &A, const BatchChunk &ch) {
  return rl_global(kconst(A.rls + ch.src));
}

// per-workgroup totals of the BS_* statistics, one slot per workgroup
__device__ __forceinline__ void block_partials(unsigned long long *v, int k0, int k1, unsigned long long *part,
                                               int nblocks, int q = 0) {
  __shared__ unsigned long long ws[BWAVES][BS_COUNT];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = k0; k < k1; k++) ws[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x < k1 - k0) {
    const int k = k0 + threadIdx.x;
    unsigned long long t = 0;
    for (int q = 0; q < BWAVES; q++) t += ws[q][k];
    part[(size_t)k * 4 * nblocks + (size_t)q * nblocks + blockIdx.x] = t;
  }
}

// voxel addressing of a dataset's volume (the SM codes of the trace kernels):
// bricked (8 or 4) with 32-bit offsets when the bricked copy allows them
// (DVolume::bsmall), else with 64-bit addresses (16 or 12), small linear (1),
// large linear (0)
__device__ __forceinline__ int chunk_layout(const TraceParams &P) { return sm_of(P.vv[0].vol); }

// The render rays of one volume chunk by the chunk's layout, one ray per lane
// at a time (MI = 1: the volume has at most one isovalue).
template <int SM, int MI>
__device__ __forceinline__ unsigned long long render_chunk(const TraceParams &P, const RayList_ispc &R,
                                                           const BatchChunk &ch, const float (*col)[3 * GXY_TF_MAX],
                                                           const float (*op)[GXY_TF_MAX],
                                                           PrivStack<GXY_BVH_STACK> &pstack) {
  unsigned long long n = 0;
  for (int k = threadIdx.x & 63; k < ch.count; k += 64)
    n += trace_one<1, false, PrivStack<GXY_BVH_STACK>, SM, 0, MI>(P, R, ch.start + k, col, op, pstack);
  return n;
}
// (bricked volumes with more than one isovalue take the 64-bit variants,
// which serve any bricked copy)
__device__ __forceinline__ unsigned long long render_chunk_any(const TraceParams &P, const RayList_ispc &R,
                                                               const BatchChunk &ch, int lay,
                                                               const float (*col)[3 * GXY_TF_MAX],
                                                               const float (*op)[GXY_TF_MAX],
                                                               PrivStack<GXY_BVH_STACK> &pstack) {
  const bool one = P.vv[0].nIso <= 1;
  if (one && lay == 4) return render_chunk<4, 1>(P, R, ch, col, op, pstack);
  if (one && lay == 12) return render_chunk<12, 1>(P, R, ch, col, op, pstack);
  if (one && lay == 8) return render_chunk<8, 1>(P, R, ch, col, op, pstack);
  if (one && lay == 16) return render_chunk<16, 1>(P, R, ch, col, op, pstack);
  if (lay == 8 || lay == 16) return render_chunk<16, GXY_MAX_ISO>(P, R, ch, col, op, pstack);
  if (lay == 4 || lay == 12) return render_chunk<12, GXY_MAX_ISO>(P, R, ch, col, op, pstack);
  if (lay == 1) return render_chunk<1, GXY_MAX_ISO>(P, R, ch, col, op, pstack);
  return render_chunk<0, GXY_MAX_ISO>(P, R, ch, col, op, pstack);
}

// The sampler rays of one chunk (samples and ray passes into acc).
template <int SM>
__device__ __forceinline__ void sample_chunk_as(const BatchArgs &A, const TraceParams &P, const RayList_ispc &R,
                                                const BatchChunk &ch, float *xyz, int *pc, int cap,
                                                unsigned long long *acc) {
  unsigned segs = 0;
  for (int k = threadIdx.x & 63; k < ch.count; k += 64) {
    acc[BS_SAMPLES] += sample_all<1, SM>(P, R, ch.start + k, xyz, pc, cap, A.max_segments, segs);
    acc[BS_RAYS] += segs;
  }
}
__device__ __forceinline__ void sample_chunk_any(const BatchArgs &A, const TraceParams &P, const RayList_ispc &R,
                                                 const BatchChunk &ch, int lay, float *xyz, int *pc, int cap,
                                                 unsigned long long *acc) {
  if (lay == 4) sample_chunk_as<4>(A, P, R, ch, xyz, pc, cap, acc);
  else if (lay == 8) sample_chunk_as<8>(A, P, R, ch, xyz, pc, cap, acc);
  else if (lay == 16) sample_chunk_as<16>(A, P, R, ch, xyz, pc, cap, acc);
  else if (lay == 12) sample_chunk_as<12>(A, P, R, ch, xyz, pc, cap, acc);
  else if (lay == 1) sample_chunk_as<1>(A, P, R, ch, xyz, pc, cap, acc);
  else sample_chunk_as<0>(A, P, R, ch, xyz, pc, cap, acc);
}

// ------------------------------------------------------------------ pass set-up
// work-queue head, per-list pending counts and the output count cleared
// (kernels, not memsets, so that the pass can be captured in a graph)
__global__ __launch_bounds__(1024) void k_bbegin(BatchArgs A) {
  if (threadIdx.x == 0) {
    *A.head = 0;
    *A.nout = 0;
    *A.nrays_out = 0;
  }
  if (threadIdx.x < 3) {
    A.kcount[threadIdx.x] = 0;
    A.khead[threadIdx.x] = 0;
  }
  for (int l = threadIdx.x; l < A.nlists; l += 1024) A.lpend[l] = 0;
  for (long long k = threadIdx.x; k < (long long)BS_COUNT * 4 * A.nblocks; k += 1024) A.part[k] = 0;
}

// kind of a chunk: 0 volume render, 1 geometry render, 2 sampler
__device__ __forceinline__ int chunk_kind(const BatchArgs &A, const BatchChunk &ch) {
  const BatchList &Ls = A.lists[ch.list];
  if (Ls.mode == BM_SAMPLE) return 2;
  return A.ds[Ls.ds].has_model ? 1 : 0;
}

// per-kind lists of the input chunks (order within a list is irrelevant)
__global__ __launch_bounds__(GXY_BLOCK) void k_bsplit(BatchArgs A) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nin = *A.nin;
  for (int base = (blockIdx.x * BWAVES + w) * 64; base < nin; base += gridDim.x * BWAVES * 64) {
    const int c = base + lane;
    const int kind = c < nin ? chunk_kind(A, A.tin[c]) : -1;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const unsigned long long m = __ballot(kind == q);
      if (!m) continue;
      int b0 = 0;
      if (lane == 0) b0 = atomicAdd(A.kcount + q, __popcll(m));
      b0 = first_lane_i(b0);
      if (kind == q) A.kidx[(size_t)q * A.kcap + b0 + __popcll(m & bl_lanemask_lt())] = c;
    }
  }
}

// appends n host-built chunks (already on the device) to the input table
__global__ __launch_bounds__(256) void k_bappend(BatchChunk *tab, int *ntab, int cap, const BatchChunk *add, int n,
                                                 int *overflow) {
  const int base = *ntab;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
    if (base + k < cap) tab[base + k] = add[k];
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (base + n > cap) *overflow |= 2;
  }
}

__global__ void k_bappend_count(int *ntab, int cap, int n) {
  int v = *ntab + n;
  *ntab = v < cap ? v : cap;
}

__global__ void k_bset_count(int *ntab, int v) { *ntab = v; }

// ------------------------------------------------------------------ trace
template <int STK, int WPE>
__global__ __launch_bounds__(GXY_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8)))
void k_btrace(BatchArgs A) {
  __shared__ float s_col[BWAVES][1][3 * GXY_TF_MAX];
  __shared__ float s_op[BWAVES][1][GXY_TF_MAX];
  __shared__ int s_stack[STK > 0 ? STK : 1][GXY_BLOCK];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nc = *A.nin;
  unsigned long long acc[BS_COUNT];
#pragma unroll
  for (int k = 0; k < BS_COUNT; k++) acc[k] = 0;
  int cur_tf = -1;
  PrivStack<GXY_BVH_STACK> pstack;
  LdsStack<(STK > 0 ? STK : 1)> lstack;
  lstack.base = &s_stack[0][threadIdx.x];
  lstack.stride = GXY_BLOCK;
  for (;;) {
    int c = 0;
    if (lane == 0) c = atomicAdd(A.head, 1);
    c = first_lane_i(c);
    if (c >= nc) break;
    const BatchChunk ch = kconst(A.tin + c);
    const BatchList &Ls = kconst(A.lists + ch.list);
    const int ds = Ls.ds, mode = Ls.mode;
    const TraceParams &P = kconst(A.ds + ds);
    const RayList_ispc R = chunk_rays(A, ch);
    const int lay = chunk_layout(P);
    if (mode == BM_SAMPLE) {
      // every crossing of the ray in this launch (segments counted as ray passes)
      float *xyz = gptr(Ls.xyz);
      int *pc = Ls.pcount;
      const int cap = Ls.pcap;
      sample_chunk_any(A, P, R, ch, lay, xyz, pc, cap, acc);
    } else if (P.has_model) {
      acc[BS_RAYS] += lane < ch.count ? (ch.count - 1 - lane) / 64 + 1 : 0;
      const bool wide = P.model.wide != 0;
      for (int k = lane; k < ch.count; k += 64) {
        if (STK > 0) {
          if (wide) acc[BS_SAMPLES] += trace_one<0, true, decltype(lstack), -1, 4>(P, R, ch.start + k, s_col[w], s_op[w], lstack);
          else acc[BS_SAMPLES] += trace_one<0, true, decltype(lstack), -1, 2>(P, R, ch.start + k, s_col[w], s_op[w], lstack);
        } else {
          if (wide) acc[BS_SAMPLES] += trace_one<0, true, decltype(pstack), -1, 4>(P, R, ch.start + k, s_col[w], s_op[w], pstack);
          else acc[BS_SAMPLES] += trace_one<0, true, decltype(pstack), -1, 2>(P, R, ch.start + k, s_col[w], s_op[w], pstack);
        }
      }
    } else {
      acc[BS_RAYS] += lane < ch.count ? (ch.count - 1 - lane) / 64 + 1 : 0;
      if (ds != cur_tf) {
        const DTF &T = P.vv[0].tf;
        for (int k = lane; k < 3 * T.nc; k += 64) s_col[w][0][k] = T.col[k];
        for (int k = lane; k < T.no; k += 64) s_op[w][0][k] = T.op[k];
        wave_lds_sync();
        cur_tf = ds;
      }
      acc[BS_SAMPLES] += render_chunk_any(P, R, ch, lay, s_col[w], s_op[w], pstack);
    }
  }
  // lane sums of the sample and ray-pass counts
  acc[BS_SAMPLES] = wave_sum_u64(acc[BS_SAMPLES]);
  acc[BS_RAYS] = wave_sum_u64(acc[BS_RAYS]);
  block_partials(acc, BS_RAYS, BS_SAMPLES + 1, A.part, A.nblocks);
}

// ------------------------------------------------------------------ trace, one kernel per kind
// The same per-chunk work as k_btrace for the chunks of one kind (the list
// k_bsplit built), so that each kind gets its own register allocation and
// occupancy; the three kernels of a pass run concurrently on three streams.
template <int KIND, int STK, int WPE>
__global__ __launch_bounds__(GXY_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8)))
void k_btrace_kind(BatchArgs A) {
  constexpr int TFW = KIND == 0 ? BWAVES : 1;
  __shared__ float s_col[TFW][1][3 * GXY_TF_MAX];
  __shared__ float s_op[TFW][1][GXY_TF_MAX];
  __shared__ int s_stack[(KIND == 1 && STK > 0) ? STK : 1][GXY_BLOCK];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nk = A.kcount[KIND];
  const int *idx = A.kidx + (size_t)KIND * A.kcap;
  unsigned long long acc[BS_COUNT];
#pragma unroll
  for (int k = 0; k < BS_COUNT; k++) acc[k] = 0;
  int cur_tf = -1;
  PrivStack<GXY_BVH_STACK> pstack;
  LdsStack<((KIND == 1 && STK > 0) ? STK : 1)> lstack;
  lstack.base = &s_stack[0][threadIdx.x];
  lstack.stride = GXY_BLOCK;
  for (;;) {
    int q = 0;
    if (lane == 0) q = atomicAdd(A.khead + KIND, 1);
    q = first_lane_i(q);
    if (q >= nk) break;
    const int c = first_lane_i(idx[q]);
    const BatchChunk ch = kconst(A.tin + c);
    const BatchList &Ls = kconst(A.lists + ch.list);
    const int ds = Ls.ds;
    const TraceParams &P = kconst(A.ds + ds);
    const RayList_ispc R = chunk_rays(A, ch);
    const int lay = chunk_layout(P);
    if (KIND == 2) {
      float *xyz = gptr(Ls.xyz);
      int *pc = Ls.pcount;
      const int cap = Ls.pcap;
      sample_chunk_any(A, P, R, ch, lay, xyz, pc, cap, acc);
    } else if (KIND == 1) {
      acc[BS_RAYS] += lane < ch.count ? (ch.count - 1 - lane) / 64 + 1 : 0;
      const bool wide = P.model.wide != 0;
      for (int k = lane; k < ch.count; k += 64) {
        if (STK > 0) {
          if (wide) acc[BS_SAMPLES] += trace_one<0, true, decltype(lstack), -1, 4>(P, R, ch.start + k, s_col[0], s_op[0], lstack);
          else acc[BS_SAMPLES] += trace_one<0, true, decltype(lstack), -1, 2>(P, R, ch.start + k, s_col[0], s_op[0], lstack);
        } else {
          if (wide) acc[BS_SAMPLES] += trace_one<0, true, decltype(pstack), -1, 4>(P, R, ch.start + k, s_col[0], s_op[0], pstack);
          else acc[BS_SAMPLES] += trace_one<0, true, decltype(pstack), -1, 2>(P, R, ch.start + k, s_col[0], s_op[0], pstack);
        }
      }
    } else {
      acc[BS_RAYS] += lane < ch.count ? (ch.count - 1 - lane) / 64 + 1 : 0;
      if (ds != cur_tf) {
        const DTF &T = P.vv[0].tf;
        for (int k = lane; k < 3 * T.nc; k += 64) s_col[w % TFW][0][k] = T.col[k];
        for (int k = lane; k < T.no; k += 64) s_op[w % TFW][0][k] = T.op[k];
        wave_lds_sync();
        cur_tf = ds;
      }
      acc[BS_SAMPLES] += render_chunk_any(P, R, ch, lay, s_col[w % TFW], s_op[w % TFW], pstack);
    }
  }
  acc[BS_SAMPLES] = wave_sum_u64(acc[BS_SAMPLES]);
  acc[BS_RAYS] = wave_sum_u64(acc[BS_RAYS]);
  block_partials(acc, BS_RAYS, BS_SAMPLES + 1, A.part, A.nblocks, 1 + KIND);
}

// ------------------------------------------------------------------ post-pass: counts
struct ChunkFlags {
  bool valid, par, keep, term, drop, surf;
  int cls;
};

__device__ __forceinline__ ChunkFlags chunk_flags(const BatchArgs &A, const RayList_ispc &R, int mode, int i,
                                                  bool valid) {
  ChunkFlags f;
  f.valid = valid;
  f.par = f.keep = f.term = f.drop = f.surf = false;
  f.cls = 0;
  if (!valid) return f;
  const int typ = R.type[i], term = R.term[i];
  int c = classify_ray(typ, term, R.classification[i], A.L.reverse);
  // AssignDestinations with no neighbouring partition (Renderer.cpp:423-454)
  if (c == RAY_BOUNDARY)
    c = ((typ == RAY_SHADOW || typ == RAY_AO) && A.L.reverse) ? GXY_DROP_ON_FLOOR : GXY_TERMINATED;
  f.cls = c;
  f.par = mode == BM_RENDER && typ == RAY_PRIMARY && (term & RAY_SURFACE);
  f.keep = c == GXY_KEEP_HERE;
  f.term = c == GXY_TERMINATED;
  f.drop = c == GXY_DROP_ON_FLOOR;
  f.surf = (term & RAY_SURFACE) != 0;
  return f;
}

__global__ __launch_bounds__(GXY_BLOCK) void k_bcount(BatchArgs A) {
  const int lane = threadIdx.x & 63;
  const int nc = *A.nin;
  const int gw = blockIdx.x * BWAVES + (threadIdx.x >> 6), nw = gridDim.x * BWAVES;
  const int nsp = A.L.nAO + (A.L.doShadows ? A.L.nLights : 0);
  for (int c = gw; c < nc; c += nw) {
    const BatchChunk ch = A.tin[c];
    const int mode = A.lists[ch.list].mode;
    const RayList_ispc R = chunk_rays(A, ch);
    int npar = 0, nkeep = 0;
    for (int k0 = 0; k0 < ch.count; k0 += 64) {
      const int k = k0 + lane;
      ChunkFlags f = chunk_flags(A, R, mode, ch.start + k, k < ch.count);
      npar += __popcll(__ballot(f.par));
      nkeep += __popcll(__ballot(f.keep));
    }
    if (lane == 0) {
      const int out = npar * nsp + nkeep;
      A.cpar[c] = npar;
      A.cout[c] = out;
      A.cchk[c] = (out + A.chunk - 1) / A.chunk;
      if (out) atomicAdd(A.lpend + ch.list, out);
    }
  }
}

// ------------------------------------------------------------------ device-count scan
// Exclusive scan of (cout, cchk) over the nin input chunks: NSEG workgroups,
// each owning one contiguous segment.
#define BSCAN_SEG 256

__global__ __launch_bounds__(256) void k_bscan_a(BatchArgs A) {
  __shared__ long long sr[4], sc[4];
  const int n = *A.nin;
  const int seg = (n + BSCAN_SEG - 1) / BSCAN_SEG;
  const int s0 = min(n, blockIdx.x * seg), s1 = min(n, s0 + seg);
  long long r = 0, c = 0;
  for (int k = s0 + threadIdx.x; k < s1; k += 256) {
    r += A.cout[k];
    c += A.cchk[k];
  }
  for (int o = 32; o > 0; o >>= 1) {
    r += __shfl_xor(r, o);
    c += __shfl_xor(c, o);
  }
  if ((threadIdx.x & 63) == 0) {
    sr[threadIdx.x >> 6] = r;
    sc[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    A.segr[blockIdx.x] = sr[0] + sr[1] + sr[2] + sr[3];
    A.segc[blockIdx.x] = sc[0] + sc[1] + sc[2] + sc[3];
  }
}

__global__ __launch_bounds__(BSCAN_SEG) void k_bscan_b(BatchArgs A) {
  __shared__ long long pr[BSCAN_SEG], pc[BSCAN_SEG];
  const int t = threadIdx.x;
  pr[t] = A.segr[t];
  pc[t] = A.segc[t];
  __syncthreads();
  for (int o = 1; o < BSCAN_SEG; o <<= 1) {
    long long vr = t >= o ? pr[t - o] : 0, vc = t >= o ? pc[t - o] : 0;
    __syncthreads();
    pr[t] += vr;
    pc[t] += vc;
    __syncthreads();
  }
  A.segr[t] = t > 0 ? pr[t - 1] : 0;
  A.segc[t] = t > 0 ? pc[t - 1] : 0;
  if (t == BSCAN_SEG - 1) {
    const long long rays = pr[t], chunks = pc[t];
    bool over = rays > A.pool_cap || chunks > A.tcap;
    if (over) *A.overflow |= 1;
    *A.nrays_out = over ? 0 : (int)rays;
    *A.nout = over ? 0 : (int)chunks;
  }
}

__global__ __launch_bounds__(256) void k_bscan_c(BatchArgs A) {
  __shared__ long long wr[4], wc[4];
  __shared__ long long carry_r, carry_c;
  const int n = *A.nin;
  const int seg = (n + BSCAN_SEG - 1) / BSCAN_SEG;
  const int s0 = min(n, blockIdx.x * seg), s1 = min(n, s0 + seg);
  if (threadIdx.x == 0) {
    carry_r = A.segr[blockIdx.x];
    carry_c = A.segc[blockIdx.x];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int b = s0; b < s1; b += 256) {
    const int k = b + threadIdx.x;
    long long r = k < s1 ? A.cout[k] : 0, c = k < s1 ? A.cchk[k] : 0;
    // inclusive wave scan
    long long ir = r, ic = c;
    for (int o = 1; o < 64; o <<= 1) {
      long long vr = __shfl_up(ir, o), vc = __shfl_up(ic, o);
      if (lane >= o) {
        ir += vr;
        ic += vc;
      }
    }
    if (lane == 63) {
      wr[w] = ir;
      wc[w] = ic;
    }
    __syncthreads();
    long long br = carry_r, bc = carry_c;
    for (int q = 0; q < w; q++) {
      br += wr[q];
      bc += wc[q];
    }
    if (k < s1) {
      A.coff[k] = (int)(br + ir - r);
      A.cchkoff[k] = (int)(bc + ic - c);
    }
    __syncthreads();
    if (threadIdx.x == 255) {
      carry_r = br + ir;
      carry_c = bc + ic;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ post-pass: writes
__device__ __forceinline__ void copy_ray(const RayList_ispc &S, int i, const RayList_ispc &D, int j) {
  D.ox[j] = S.ox[i]; D.oy[j] = S.oy[i]; D.oz[j] = S.oz[i];
  D.dx[j] = S.dx[i]; D.dy[j] = S.dy[i]; D.dz[j] = S.dz[i];
  D.nx[j] = S.nx[i]; D.ny[j] = S.ny[i]; D.nz[j] = S.nz[i];
  D.sample[j] = S.sample[i];
  D.r[j] = S.r[i]; D.g[j] = S.g[i]; D.b[j] = S.b[i]; D.o[j] = S.o[i];
  D.sr[j] = S.sr[i]; D.sg[j] = S.sg[i]; D.sb[j] = S.sb[i]; D.so[j] = S.so[i];
  D.t[j] = S.t[i]; D.tMax[j] = S.tMax[i];
  D.x[j] = S.x[i]; D.y[j] = S.y[i];
  D.type[j] = S.type[i]; D.term[j] = S.term[i]; D.classification[j] = S.classification[i];
}

#define BATCH_MAX_CHUNK 1024

// Per chunk (one wave): (1) rank the parents and keep their ray offsets in
// the wave's LDS slice; (2) write every child, one lane per child at
// base + p * (nAO + nLights) + j (AO children first, then shadow children of
// the p-th parent), from the parents' unmodified fields, so that consecutive
// lanes store consecutive children; (3) per ray: the parent's ambient and
// diffuse update, framebuffer accumulation of terminated rays (after that
// update), keepers copied after the children, sampler particles.
__global__ __launch_bounds__(GXY_BLOCK) void k_bwrite(BatchArgs A) {
  __shared__ int s_pl[BWAVES][BATCH_MAX_CHUNK];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nc = *A.nin;
  const bool over = *A.overflow != 0;
  const int gw = blockIdx.x * BWAVES + w, nw = gridDim.x * BWAVES;
  const int nLs = A.L.doShadows ? A.L.nLights : 0;
  const int nAO = A.L.nAO;
  const int nsp = nAO + nLs;
  const RayList_ispc O = rl_global(kconst(A.rls + A.out_pool));
  unsigned long long acc[BS_COUNT];
#pragma unroll
  for (int k = 0; k < BS_COUNT; k++) acc[k] = 0;
  for (int c = gw; c < nc; c += nw) {
    const BatchChunk ch = A.tin[c];
    const BatchList &Ls = A.lists[ch.list];
    const int mode = Ls.mode;
    const RayList_ispc R = chunk_rays(A, ch);
    const int base = A.coff[c], npar = A.cpar[c];
    const int keep_base = base + npar * nsp;
    // (1) parents in rank order
    int par_run = 0;
    if (mode == BM_RENDER && nsp > 0) {
      for (int k0 = 0; k0 < ch.count; k0 += 64) {
        const int k = k0 + lane;
        ChunkFlags f = chunk_flags(A, R, mode, ch.start + k, k < ch.count);
        const unsigned long long mp = __ballot(f.par);
        if (f.par) s_pl[w][par_run + __popcll(mp & bl_lanemask_lt())] = k;
        par_run += __popcll(mp);
      }
      wave_lds_sync();
      // (2) children
      if (!over) {
        for (int q = lane; q < npar * nsp; q += 64) {
          const int p = q / nsp, j = q - p * nsp;
          const ParentRay P = parent_load(R, ch.start + s_pl[w][p]);
          if (j < nAO) light_ao_child(A.L, P, j, O, base + q);
          else light_shadow_child(A.L, P, A.L.reverse ? light_diffused_alpha(A.L, P.o) : P.o, j - nAO, O, base + q);
        }
      }
      wave_lds_sync();   // the slice is rewritten by the next chunk
    }
    // (3) parents' colours, accumulation, keepers, particles
    int keep_run = 0;
    for (int k0 = 0; k0 < ch.count; k0 += 64) {
      const int k = k0 + lane, i = ch.start + k;
      ChunkFlags f = chunk_flags(A, R, mode, i, k < ch.count);
      const unsigned long long mp = __ballot(f.par), mk = __ballot(f.keep);
      const unsigned long long mt = __ballot(f.term), md = __ballot(f.drop);
      const int krank = keep_run + __popcll(mk & bl_lanemask_lt());
      if (!over) {
        if (mode == BM_RENDER) {
          int x = 0, y = 0;
          float cr = 0.f, cg = 0.f, cb = 0.f, co = 0.f;
          if (f.par) {
            // ambientLighting + diffuseLighting (TraceRays.cpp:124-131); without
            // reverse lighting only those whose rays are not cast (:133-141)
            ParentRay P = parent_load(R, i);
            if (A.L.reverse || A.L.nAO <= 0) light_ambient(A.L, P);
            if (A.L.reverse || !(A.L.doShadows && A.L.nLights > 0)) light_diffuse(A.L, P);
            R.r[i] = P.r; R.g[i] = P.g; R.b[i] = P.b; R.o[i] = P.o;
            if (f.term) {
              x = P.px; y = P.py;
              cr = P.r; cg = P.g; cb = P.b; co = P.o;
            }
          } else if (f.term) {
            x = R.x[i];
            y = R.y[i];
            cr = R.r[i];
            cg = R.g[i];
            cb = R.b[i];
            co = R.o[i];
          }
          fb_accumulate_wave(Ls.fb, Ls.width, Ls.height, f.term, x, y, cr, cg, cb, co);
        }
        if (f.keep) copy_ray(R, i, O, keep_base + krank);
        if (mode == BM_SAMPLE) {
          // Sampler::HandleTerminatedRays (Sampler.cpp:71-84): a particle at o + t*d
          const unsigned long long ms = __ballot(f.surf);
          int pbase = 0;
          if (lane == 0 && ms) pbase = atomicAdd(Ls.pcount, __popcll(ms));
          pbase = first_lane_i(pbase);
          if (f.surf) {
            const int off = pbase + __popcll(ms & bl_lanemask_lt());
            if (off < Ls.pcap) {
              const float t = R.t[i];
              Ls.xyz[3 * (int64_t)off + 0] = R.ox[i] + t * R.dx[i];
              Ls.xyz[3 * (int64_t)off + 1] = R.oy[i] + t * R.dy[i];
              Ls.xyz[3 * (int64_t)off + 2] = R.oz[i] + t * R.dz[i];
            }
          }
        }
      }
      acc[BS_SPAWNED] += (unsigned long long)__popcll(mp) * nsp;
      acc[BS_TERMINATED] += __popcll(mt);
      acc[BS_DROPPED] += __popcll(md);
      acc[BS_KEPT] += __popcll(mk);
      keep_run += __popcll(mk);
    }
    if (!over) {
      const int total = npar * nsp + keep_run;
      const int nch = A.cchk[c], cb = A.cchkoff[c];
      for (int q = lane; q < nch; q += 64) {
        BatchChunk o;
        o.src = A.out_pool;
        o.start = base + q * A.chunk;
        o.count = min(A.chunk, total - q * A.chunk);
        o.list = ch.list;
        A.tout[cb + q] = o;
      }
    }
  }
  block_partials(acc, BS_SPAWNED, BS_COUNT, A.part, A.nblocks);
}

// adds the per-workgroup partials of every statistic into A.stats
__global__ __launch_bounds__(1024) void k_breduce(BatchArgs A, int k0, int k1) {
  __shared__ unsigned long long ws[16];
  for (int k = k0; k < k1; k++) {
    unsigned long long v = 0;
    for (int b = threadIdx.x; b < 4 * A.nblocks; b += 1024) v += A.part[(size_t)k * 4 * A.nblocks + b];
    v = wave_sum_u64(v);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int q = 0; q < 16; q++) t += ws[q];
      A.stats[k] += t;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ synthetic ray lists
// Random rays of one list: origins uniform in [lo, hi], isotropic directions
// (z = 1 - 2u, phi = 2 pi u'), t = 0, tMax = inf, pixel x = index % width,
// y = index / width; a PCG hash of (seed, ray index, draw) gives the uniforms.
__device__ __forceinline__ unsigned pcg_hash(unsigned v) {
  unsigned s = v * 747796405u + 2891336453u;
  unsigned w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}

__device__ __forceinline__ float u01(unsigned seed, unsigned i, unsigned k) {
  unsigned h = pcg_hash(seed ^ pcg_hash(i * 8u + k + 0x9e3779b9u * (seed + 1u)));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

__global__ __launch_bounds__(GXY_BLOCK) void k_random_rays(RayList_ispc R, int n, unsigned seed, int width, int type,
                                                           float lx, float ly, float lz, float hx, float hy,
                                                           float hz) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= n) return;
  float ux = u01(seed, i, 0), uy = u01(seed, i, 1), uz = u01(seed, i, 2);
  float zc = 1.0f - 2.0f * u01(seed, i, 3), ph = 6.2831855f * u01(seed, i, 4);
  float rr = sqrtf(fmaxf(0.f, 1.0f - zc * zc));
  R.ox[i] = lx + ux * (hx - lx);
  R.oy[i] = ly + uy * (hy - ly);
  R.oz[i] = lz + uz * (hz - lz);
  R.dx[i] = rr * cosf(ph);
  R.dy[i] = rr * sinf(ph);
  R.dz[i] = zc;
  R.nx[i] = 0.f; R.ny[i] = 0.f; R.nz[i] = 0.f;
  R.sample[i] = 0.f;
  R.r[i] = 0.f; R.g[i] = 0.f; R.b[i] = 0.f; R.o[i] = 0.f;
  R.sr[i] = 0.f; R.sg[i] = 0.f; R.sb[i] = 0.f; R.so[i] = 0.f;
  R.t[i] = 0.f;
  R.tMax[i] = __builtin_inff();
  R.x[i] = width > 0 ? i % width : i;
  R.y[i] = width > 0 ? i / width : 0;
  R.type[i] = type;
  R.term[i] = 0;
  R.classification[i] = 0;
}

__global__ __launch_bounds__(256) void k_fill_f32(float *p, long long n, float v) {
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) p[k] = v;
}
__global__ void k_fill_i32_one(int *p, int v) { *p = v; }

// ------------------------------------------------------------------ launchers
template <int STK, int WPE>
static void btrace_as(const BatchArgs &A, int blocks, hipStream_t s) {
  hipLaunchKernelGGL((k_btrace<STK, WPE>), dim3(blocks), dim3(GXY_BLOCK), 0, s, A);
}

template <int KIND, int STK, int WPE>
static void bkind_as(const BatchArgs &A, int blocks, hipStream_t s) {
  hipLaunchKernelGGL((k_btrace_kind<KIND, STK, WPE>), dim3(blocks), dim3(GXY_BLOCK), 0, s, A);
}

static void launch_split_trace(const BatchArgs &A, const BatchLaunch &K, hipStream_t s, const hipStream_t *ks,
                               const hipEvent_t *ev) {
  const int tb = K.trace_blocks;
  hipLaunchKernelGGL(k_bsplit, dim3(K.post_blocks), dim3(GXY_BLOCK), 0, s, A);
  (void)hipEventRecord(ev[0], s);
  for (int q = 0; q < 3; q++) (void)hipStreamWaitEvent(ks[q], ev[0], 0);
  // volume render: occupancy request 5 or 6 waves per SIMD when asked (LDS TF slices only)
  if (K.wpe == 6) bkind_as<0, 0, 6>(A, tb, ks[0]);
  else if (K.wpe == 5) bkind_as<0, 0, 5>(A, tb, ks[0]);
  else if (K.wpe == 4) bkind_as<0, 0, 4>(A, tb, ks[0]);
  else bkind_as<0, 0, 0>(A, tb, ks[0]);
  if (K.stack == 16) bkind_as<1, 16, 0>(A, tb, ks[1]);
  else if (K.stack == 24) bkind_as<1, 24, 0>(A, tb, ks[1]);
  else if (K.stack == 32) bkind_as<1, 32, 0>(A, tb, ks[1]);
  else bkind_as<1, 0, 0>(A, tb, ks[1]);
  bkind_as<2, 0, 8>(A, tb, ks[2]);
  for (int q = 0; q < 3; q++) {
    (void)hipEventRecord(ev[1 + q], ks[q]);
    (void)hipStreamWaitEvent(s, ev[1 + q], 0);
  }
}

hipError_t launch_batch_pass(const BatchArgs &A, const BatchLaunch &K, hipStream_t s, const hipStream_t *ks,
                             const hipEvent_t *ev) {
  hipLaunchKernelGGL(k_bbegin, dim3(1), dim3(1024), 0, s, A);
  const int tb = K.trace_blocks;
  if (K.split && ks && ev) {
    launch_split_trace(A, K, s, ks, ev);
  } else if (K.stack == 16) {
    if (K.wpe == 4) btrace_as<16, 4>(A, tb, s); else btrace_as<16, 0>(A, tb, s);
  } else if (K.stack == 24) {
    if (K.wpe == 4) btrace_as<24, 4>(A, tb, s); else btrace_as<24, 0>(A, tb, s);
  } else if (K.stack == 32) {
    if (K.wpe == 4) btrace_as<32, 4>(A, tb, s); else btrace_as<32, 0>(A, tb, s);
  } else {
    if (K.wpe == 4) btrace_as<0, 4>(A, tb, s);
    else if (K.wpe == 5) btrace_as<0, 5>(A, tb, s);
    else if (K.wpe == 6) btrace_as<0, 6>(A, tb, s);
    else btrace_as<0, 0>(A, tb, s);
  }
  hipLaunchKernelGGL(k_bcount, dim3(K.post_blocks), dim3(GXY_BLOCK), 0, s, A);
  hipLaunchKernelGGL(k_bscan_a, dim3(BSCAN_SEG), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_bscan_b, dim3(1), dim3(BSCAN_SEG), 0, s, A);
  hipLaunchKernelGGL(k_bscan_c, dim3(BSCAN_SEG), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_bwrite, dim3(K.post_blocks), dim3(GXY_BLOCK), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_batch_reduce(const BatchArgs &A, hipStream_t s) {
  hipLaunchKernelGGL(k_breduce, dim3(1), dim3(1024), 0, s, A, 0, (int)BS_COUNT);
  return hipGetLastError();
}

hipError_t launch_batch_append(BatchChunk *tab, int *ntab, int cap, const BatchChunk *add, int n, int *overflow,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int g = (n + 255) / 256;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_bappend, dim3(g), dim3(256), 0, s, tab, ntab, cap, add, n, overflow);
  hipLaunchKernelGGL(k_bappend_count, dim3(1), dim3(1), 0, s, ntab, cap, n);
  return hipGetLastError();
}

hipError_t launch_set_count(int *p, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32_one, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, long long n, float v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)g), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_random_rays(const RayList_ispc &R, int n, unsigned seed, int width, int type, const float *lo,
                              const float *hi, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_random_rays, dim3((n + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, R, n, seed, width,
                     type, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
  return hipGetLastError();
}

}  // namespace gxy
