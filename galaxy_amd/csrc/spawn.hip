// spawn.hip — secondary-ray bookkeeping and lighting on gfx950.
//
// TraceRays.cpp:89-117 computes, serially on the host, exclusive offsets of the
// AO block (nAO slots per primary surface ray) followed by the shadow block
// (nLights slots each).  Since every parent contributes 0 or a fixed count,
// both offset arrays follow from one exclusive rank of the "primary & surface"
// flag: a wave ballot + popcount gives the rank inside a wave, LDS combines the
// four waves of a workgroup, and a scan over workgroup totals (one workgroup,
// or the device-wide scan below for long lists) gives the global base.  The lighting kernel then applies, per parent and in
// the reference order, ambientLighting -> generateAORays -> diffuseLighting ->
// generateShadowRays (TraceRays.cpp:124-131, TraceRays.ispc:625-923).
#include "light_dev.h"

namespace gxy {

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Per-workgroup count of flagged rays; a zero after the nb counts, so that
// their exclusive sum ends in the total.
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
    if (blockIdx.x == 0) block_counts[gridDim.x] = 0;
  }
}

__global__ __launch_bounds__(GXY_BLOCK) void k_flag_rank(RayList_ispc R, int n, const int *block_base, int *rank,
                                                         int *pidx) {
  __shared__ int wsum[GXY_BLOCK / 64];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = i < n && is_parent(R, i);
  unsigned long long m = __ballot(f);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wsum[w] = __popcll(m);
  __syncthreads();
  int base = block_base[blockIdx.x];
  for (int k = 0; k < w; k++) base += wsum[k];
  // exclusive count of flagged rays before i (defined for every ray); the
  // flagged rays' indices in rank order when pidx is given
  const int rk = base + __popcll(m & lanemask_lt());
  if (i < n && rank) rank[i] = rk;
  if (f && pidx) pidx[rk] = i;
}

// ------------------------------------------------------------------ device-wide scans
// Exclusive sum of m ints (in -> out), reduce-then-scan over tiles of
// SCAN_TILE ints (256 work-items x 8): one launch sums each tile, the tile
// sums are scanned (one workgroup when they fit, else the same scheme one
// level up), and one launch scans each tile from its base.  temp holds
// scan_temp_bytes(m) bytes.  m up to 2^31 - 1.
#define SCAN_ITEMS 8
#define SCAN_TILE (GXY_BLOCK * SCAN_ITEMS)

// Exclusive sum of m ints (in -> out, m <= GXY_SMALL_SCAN) by one workgroup
// of 1024 work-items: one launch where the device-wide scan takes two, for the
// count arrays of lists up to a few million rays.
__global__ __launch_bounds__(1024) void k_scan_small(const int *in, int *out, int m) { scan1024_runs(in, out, m); }

__device__ __forceinline__ int block_sum256(int v, int *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(GXY_BLOCK) void k_scan_tile_sums(const int *in, long long m, int *sums) {
  __shared__ int red[4];
  const long long base = (long long)blockIdx.x * SCAN_TILE;
  int v = 0;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    const long long i = base + (long long)k * GXY_BLOCK + threadIdx.x;
    if (i < m) v += in[i];
  }
  v = block_sum256(v, red);
  if (threadIdx.x == 0) sums[blockIdx.x] = v;
}

// Each work-item scans 8 consecutive ints; the work-items' totals are scanned
// across the workgroup (wave shuffles, then the four wave totals).
__global__ __launch_bounds__(GXY_BLOCK) void k_scan_tiles(const int *in, int *out, long long m, const int *tile_base) {
  __shared__ int wtot[4];
  const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int t = 0;
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < m ? in[base + k] : 0;
    t += v[k];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = t;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wtot[w] = incl;
  __syncthreads();
  int run = tile_base[blockIdx.x] + incl - t;
  for (int k = 0; k < w; k++) run += wtot[k];
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < m) out[base + k] = run;
    run += v[k];
  }
}

static inline size_t scan_pad_ll(long long m) { return (size_t)((m + 63) / 64 * 64); }

size_t scan_temp_bytes(long long m) {
  if (m <= GXY_SMALL_SCAN) return 256;
  const long long nt = (m + SCAN_TILE - 1) / SCAN_TILE;
  return 2 * sizeof(int) * scan_pad_ll(nt) + scan_temp_bytes(nt);
}

hipError_t launch_exclusive_sum(const int *in, int *out, long long m, void *temp, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  if (m <= GXY_SMALL_SCAN) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, in, out, (int)m);
    return hipGetLastError();
  }
  const long long nt = (m + SCAN_TILE - 1) / SCAN_TILE;
  int *sums = (int *)temp, *sbase = sums + scan_pad_ll(nt);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nt), dim3(GXY_BLOCK), 0, s, in, m, sums);
  hipError_t e = launch_exclusive_sum(sums, sbase, nt, sbase + scan_pad_ll(nt), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(GXY_BLOCK), 0, s, in, out, m, (const int *)sbase);
  return hipGetLastError();
}

__global__ void k_set_int(int *dst, const int *src, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = src ? *src : v;
}

// ints of one scan array of m entries, rounded to 64 (256-byte alignment)
static inline size_t scan_pad(size_t m) { return (m + 63) / 64 * 64; }

size_t block_scan_scratch_bytes(int nb) {
  // the device-wide scan's temp storage only where it is used (its size
  // query costs device-property calls on the host)
  return 2 * sizeof(int) * scan_pad((size_t)nb + 1) + (nb + 1 > GXY_SMALL_SCAN ? scan_temp_bytes(nb + 1) : 0);
}

int *block_scan_offsets(int *scratch, int nb) { return scratch + scan_pad((size_t)nb + 1); }

hipError_t launch_block_scan(int *scratch, int nb, hipStream_t s) {
  int *off = block_scan_offsets(scratch, nb);
  if (nb + 1 <= GXY_SMALL_SCAN) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, scratch, off, nb + 1);
    return hipGetLastError();
  }
  return launch_exclusive_sum(scratch, off, nb + 1, scratch + 2 * scan_pad((size_t)nb + 1), s);
}

size_t surface_scan_scratch_bytes(int n) { return block_scan_scratch_bytes((n + GXY_BLOCK - 1) / GXY_BLOCK); }

int *surface_scan_total(int *scratch, int n) {
  const int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  return block_scan_offsets(scratch, nb) + nb;
}

hipError_t launch_surface_scan(const RayList_ispc &R, int n, int *scratch, int *rank_out, int *total,
                               hipStream_t s, int *pidx) {
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  // block counts with a zero after them, and their exclusive sum (the total in its last slot)
  int *off = block_scan_offsets(scratch, nb);
  if (nb == 0) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, off, (const int *)nullptr, 0);
    if (total) hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, total, (const int *)nullptr, 0);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, scratch);
  hipError_t e = launch_block_scan(scratch, nb, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_flag_rank, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, off, rank_out, pidx);
  if (total) hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, total, (const int *)(off + nb), 0);
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

// ------------------------------------------------------------------ allocation-free spawn offsets
// The per-workgroup parent counts and their exclusive sums of the
// allocation-free offsets live in this file-scope array: one copy per GPU,
// held until the process ends (2 * GXY_SMALL_SCAN ints = 128 KiB).  Lists of
// up to GXY_NOALLOC_MAX_RAYS rays fit, so their count array (at most
// GXY_SMALL_SCAN entries with the trailing zero) is scanned by the
// one-workgroup scan, which needs no temporary storage.  Host code never
// passes the array's name to a kernel: it uses the address
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) int scratch[2 * GXY_SMALL_SCAN];

int noalloc_scratch_address(int **ptr) {
  *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    *ptr = nullptr;
    return -1;
  }
  return 0;
}

// Offsets straight from the scanned workgroup counts: each work-item
// recomputes its rank among the parents (ballot + the waves before it), so no
// per-ray rank array is needed.  The parent total is block_base[nb].
__global__ __launch_bounds__(GXY_BLOCK) void k_offsets_direct(RayList_ispc R, int n, const int *block_base, int nb,
                                                              int nAO, int nL, int doShadows, int *ao_off,
                                                              int *sh_off) {
  __shared__ int wsum[GXY_BLOCK / 64];
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  const bool f = i < n && is_parent(R, i);
  const unsigned long long m = __ballot(f);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wsum[w] = __popcll(m);
  __syncthreads();
  if (i >= n) return;
  int r = block_base[blockIdx.x];
  for (int k = 0; k < w; k++) r += wsum[k];
  r += __popcll(m & lanemask_lt());
  const int T = block_base[nb];
  if (ao_off) ao_off[i] = nAO * r;
  if (sh_off) sh_off[i] = (doShadows && f) ? nAO * T + nL * r : -1;
}

// Three launches on s over scr (the array above, by its device address); the
// parent total is left at *total_out (a device address inside scr).  Every
// scratch entry read is written first in the same call (the count kernel
// writes nb counts and the trailing zero; the scan writes nb + 1 sums), so
// nothing needs clearing between calls.
hipError_t launch_spawn_offsets_noalloc(const RayList_ispc &R, int n, int *scr, int nAO, int nL, int doShadows,
                                        int *ao_off, int *sh_off, const int **total_out, hipStream_t s) {
  const int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  if (n <= 0 || nb + 1 > GXY_SMALL_SCAN || !scr) return hipErrorInvalidValue;
  int *counts = scr, *base = scr + GXY_SMALL_SCAN;
  hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, counts);
  hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, (const int *)counts, base, nb + 1);
  hipLaunchKernelGGL(k_offsets_direct, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, (const int *)base, nb, nAO, nL,
                     doShadows, ao_off, sh_off);
  *total_out = base + nb;
  return hipGetLastError();
}

// ------------------------------------------------------------------ lighting
__global__ __launch_bounds__(GXY_BLOCK) void k_lighting(LightParams L, int mode, RayList_ispc R, int n,
                                                        const int *ao_off, const int *sh_off, RayList_ispc O) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= n) return;
  light_one(L, mode, R, i, (mode & LM_AO) ? ao_off[i] : 0, (mode & LM_SHADOW) ? sh_off[i] : -1, O);
}

// The same sequence with one work-item per child, for the offsets of
// TraceRays.cpp:89-117 (AO child j of the p-th parent at nAO*p + j, then its
// shadow child k at nAO*T + nLights*p + k): consecutive work-items write
// consecutive children, so every field store of a wave is one coalesced
// transaction (one work-item per parent writes each field with a stride of
// nAO entries).  pidx: the parents' ray indices in rank order.  The children
// read the parents' unmodified colours, so this runs before k_parents_light.
__global__ __launch_bounds__(GXY_BLOCK) void k_children(LightParams L, int mode, RayList_ispc R, const int *pidx,
                                                        int Th, const int *Td, RayList_ispc O) {
  const int T = Td ? *Td : Th;
  const int nAO = (mode & LM_AO) ? L.nAO : 0, nL = (mode & LM_SHADOW) ? L.nLights : 0;
  const long long nao = (long long)nAO * T, nc = nao + (long long)nL * T;
  // grid-stride: with T on the device the grid is bounded, not sized for every ray being a parent
  const long long stride = (long long)gridDim.x * GXY_BLOCK;
  for (long long t = (long long)blockIdx.x * GXY_BLOCK + threadIdx.x; t < nc; t += stride) {
    if (t < nao) {
      const int p = (int)(t / nAO), j = (int)(t - (long long)p * nAO);
      const ParentRay P = parent_load(R, pidx[p]);
      light_ao_child(L, P, j, O, (int)t);
    } else {
      const long long t2 = t - nao;
      const int p = (int)(t2 / nL), k = (int)(t2 - (long long)p * nL);
      const ParentRay P = parent_load(R, pidx[p]);
      const float o = (mode & LM_DIFFUSE) ? light_diffused_alpha(L, P.o) : P.o;
      light_shadow_child(L, P, o, k, O, (int)t);
    }
  }
}

// ambientLighting and diffuseLighting of the T parents (after k_children)
__global__ __launch_bounds__(GXY_BLOCK) void k_parents_light(LightParams L, int mode, RayList_ispc R, const int *pidx,
                                                             int Th, const int *Td) {
  const int T = Td ? *Td : Th;
  for (int p = blockIdx.x * GXY_BLOCK + threadIdx.x; p < T; p += gridDim.x * GXY_BLOCK) {
    const int i = pidx[p];
    ParentRay P = parent_load(R, i);
    if (mode & LM_AMBIENT) light_ambient(L, P);
    if (mode & LM_DIFFUSE) light_diffuse(L, P);
    R.r[i] = P.r; R.g[i] = P.g; R.b[i] = P.b; R.o[i] = P.o;
  }
}

// k_children and k_parents_light in one pass when each parent has exactly one
// child (one AO ray and no shadow ray, or one shadow ray and no AO ray): the
// work-item of parent p writes child p from the parent's unmodified fields,
// then lights the parent.
__global__ __launch_bounds__(GXY_BLOCK) void k_child_parent(LightParams L, int mode, RayList_ispc R, const int *pidx,
                                                            const int *Td, RayList_ispc O) {
  const int T = *Td;
  const bool ao = (mode & LM_AO) && L.nAO == 1;
  for (int p = blockIdx.x * GXY_BLOCK + threadIdx.x; p < T; p += gridDim.x * GXY_BLOCK) {
    const int i = pidx[p];
    ParentRay P = parent_load(R, i);
    if (ao) {
      light_ao_child(L, P, 0, O, p);
    } else {
      const float o = (mode & LM_DIFFUSE) ? light_diffused_alpha(L, P.o) : P.o;
      light_shadow_child(L, P, o, 0, O, p);
    }
    if (mode & (LM_AMBIENT | LM_DIFFUSE)) {
      if (mode & LM_AMBIENT) light_ambient(L, P);
      if (mode & LM_DIFFUSE) light_diffuse(L, P);
      R.r[i] = P.r; R.g[i] = P.g; R.b[i] = P.b; R.o[i] = P.o;
    }
  }
}

hipError_t launch_lighting_children(const LightParams &L, int mode, const RayList_ispc &R, const int *pidx, int T,
                                    const RayList_ispc &O, hipStream_t s) {
  if (T <= 0) return hipSuccess;
  const long long nc = (long long)T * (((mode & LM_AO) ? L.nAO : 0) + ((mode & LM_SHADOW) ? L.nLights : 0));
  if (nc > 0)
    hipLaunchKernelGGL(k_children, dim3((unsigned)((nc + GXY_BLOCK - 1) / GXY_BLOCK)), dim3(GXY_BLOCK), 0, s, L,
                       mode, R, pidx, T, (const int *)nullptr, O);
  if (mode & (LM_AMBIENT | LM_DIFFUSE))
    hipLaunchKernelGGL(k_parents_light, dim3((T + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, L, mode, R,
                       pidx, T, (const int *)nullptr);
  return hipGetLastError();
}

hipError_t launch_lighting_children_dev(const LightParams &L, int mode, const RayList_ispc &R, const int *pidx,
                                        const int *T, int n, const RayList_ispc &O, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  // grids for the upper bound (n parents), at most 32768 workgroups (128 per
  // CU): the child writes wait on gathers of their parents' fields, and more
  // work-items in flight hide more of that latency (0.727 -> 0.675 ms for the
  // 29 M children of a 4K nine-icosphere frame against 4096 workgroups)
  const long long cap = 32768;
  const int per = ((mode & LM_AO) ? L.nAO : 0) + ((mode & LM_SHADOW) ? L.nLights : 0);
  const long long nc = (long long)n * per;
  if (per == 1) {
    hipLaunchKernelGGL(k_child_parent, dim3((unsigned)std::min(cap, (long long)(n + GXY_BLOCK - 1) / GXY_BLOCK)),
                       dim3(GXY_BLOCK), 0, s, L, mode, R, pidx, T, O);
    return hipGetLastError();
  }
  if (nc > 0)
    hipLaunchKernelGGL(k_children, dim3((unsigned)std::min(cap, (nc + GXY_BLOCK - 1) / GXY_BLOCK)), dim3(GXY_BLOCK),
                       0, s, L, mode, R, pidx, 0, T, O);
  if (mode & (LM_AMBIENT | LM_DIFFUSE))
    hipLaunchKernelGGL(k_parents_light, dim3((unsigned)std::min(cap, (long long)(n + GXY_BLOCK - 1) / GXY_BLOCK)),
                       dim3(GXY_BLOCK), 0, s, L, mode, R, pidx, 0, T);
  return hipGetLastError();
}

hipError_t launch_lighting(const LightParams &L, int mode, const RayList_ispc &R, int n, const int *ao_off,
                           const int *sh_off, const RayList_ispc &O, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lighting, dim3((n + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, L, mode, R, n, ao_off,
                     sh_off, O);
  return hipGetLastError();
}

}  // namespace gxy
