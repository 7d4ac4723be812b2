// bvh.hip — BVH construction on the GPU for the model commit (the place of
// Embree's SAH build at ospCommit(ospModel), Visualization.cpp:219-277).
//
// 1. centroid bounds (wave reduction + ordered-int atomics) and 63-bit Morton
//    codes of the primitive centroids (21 bits per axis);
// 2. LSD radix sort of (code, primitive) pairs, 8 bits per pass: per-tile
//    digit histograms in LDS, one device-wide exclusive scan of the
//    digit-major histogram, and a stable scatter whose in-tile ranks come from
//    wave ballots over the digit bits;
// 3. Karras topology (HPG 2012): one work-item per inner node finds its key
//    range and split with the delta() prefix length (equal codes broken by
//    index);
// 4. bottom-up refit, one launch per depth (depths by pointer jumping over the
//    parent links, then a counting sort of the inner nodes by depth): child
//    boxes, primitive counts, heights, and the SAH cost of every subtree both
//    as built and with the cheaper of leaf or node at each subtree that fits a
//    leaf.  Kernel boundaries order the levels, so no work-item waits on
//    another's stores;
// 5. treelet restructuring (Karras and Aila, "Fast Parallel Construction of
//    High-Quality Bounding Volume Hierarchies", HPG 2013): every node with at
//    least gamma primitives below it roots a treelet of 7 leaves, grown from its
//    two children by opening the inner leaf of largest surface area; the
//    SAH-optimal binary tree over those 7 leaves (dynamic programming over the
//    127 subsets, leaf option included) replaces the treelet when it is
//    cheaper, reusing its inner nodes.  Nodes are processed by height, lowest
//    first (one launch per height; treelets of one height are disjoint), three
//    rounds with gamma = 7, 14, 28;
// 6. layout, top-down in breadth-first frontiers: a subtree whose SAH cost is
//    lower as one leaf (at most GXY_BVH_LEAF primitives) becomes a leaf, inner
//    nodes are renumbered (root 0, the inner children of a node adjacent) and
//    the primitives reordered so that every leaf covers a contiguous range;
// 7. 4-wide collapse, top-down in frontiers: a node takes its two children and
//    opens the inner child of largest surface area until it holds four.
// Node layout: BVHNode (both child boxes, 64 B) and BVH4Node (128 B); child
// codes are inner indices >= 0 or leaf codes ~(first << 4 | count - 1).
#include <cstring>
#include <vector>

#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

// SAH cost of visiting a node (per unit of surface area): a node visit fetches
// 64 B and tests two boxes.  The cost of testing one primitive is the build's
// prim_cost argument (a triangle test, 36 B and about 50 flops, is about 1).
#define BVH_CI 1.2f

__device__ __forceinline__ unsigned int okey(float f) {
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ounkey(unsigned int k) {
  unsigned int u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

__host__ __device__ __forceinline__ int leaf_code(int k) { return ~(k << 4); }
__host__ __device__ __forceinline__ int leaf_code_n(int first, int count) { return ~((first << 4) | (count - 1)); }

__device__ __forceinline__ float half_area(const float lo[3], const float hi[3]) {
  const float dx = fmaxf(hi[0] - lo[0], 0.f), dy = fmaxf(hi[1] - lo[1], 0.f), dz = fmaxf(hi[2] - lo[2], 0.f);
  return dx * dy + dy * dz + dz * dx;
}

// loads and stores of data other work-items of the same launch write
__device__ __forceinline__ float ldf(const float *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ldi(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void stf(float *p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void sti(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ __forceinline__ void slot_box(const BVHNode *nd, int slot, float lo[3], float hi[3]) {
  const float *l = slot ? nd->lo1 : nd->lo0, *h = slot ? nd->hi1 : nd->hi0;
  for (int a = 0; a < 3; a++) { lo[a] = ldf(l + a); hi[a] = ldf(h + a); }
}
__device__ __forceinline__ void set_slot(BVHNode *nd, int slot, int code, const float lo[3], const float hi[3]) {
  float *l = slot ? nd->lo1 : nd->lo0, *h = slot ? nd->hi1 : nd->hi0;
  for (int a = 0; a < 3; a++) { stf(l + a, lo[a]); stf(h + a, hi[a]); }
  sti(slot ? &nd->c1 : &nd->c0, code);
}

// ------------------------------------------------------------------ Morton codes
__global__ void k_bounds_init(unsigned int *b) {
  for (int k = 0; k < 3; k++) { b[k] = 0xffffffffu; b[3 + k] = 0u; }
}

__global__ __launch_bounds__(GXY_BLOCK) void k_centroid_bounds(const float *boxes, int n, unsigned int *b) {
  unsigned int lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  for (int i = blockIdx.x * GXY_BLOCK + threadIdx.x; i < n; i += gridDim.x * GXY_BLOCK) {
    const float *p = boxes + 6 * (int64_t)i;
    for (int k = 0; k < 3; k++) {
      unsigned int c = okey(0.5f * (p[k] + p[3 + k]));
      lo[k] = min(lo[k], c);
      hi[k] = max(hi[k], c);
    }
  }
  for (int k = 0; k < 3; k++) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[k] = min(lo[k], (unsigned int)__shfl_xor((int)lo[k], o, 64));
      hi[k] = max(hi[k], (unsigned int)__shfl_xor((int)hi[k], o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin(b + k, lo[k]);
      atomicMax(b + 3 + k, hi[k]);
    }
  }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

__global__ __launch_bounds__(GXY_BLOCK) void k_morton(const float *boxes, int n, const unsigned int *b,
                                                      uint64_t *keys, int *vals) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= n) return;
  const float *p = boxes + 6 * (int64_t)i;
  uint64_t code = 0;
  for (int k = 0; k < 3; k++) {
    float lo = ounkey(b[k]), hi = ounkey(b[3 + k]);
    float c = 0.5f * (p[k] + p[3 + k]);
    float e = hi - lo;
    float u = e > 0.f ? (c - lo) / e : 0.5f;
    u = u < 0.f ? 0.f : (u > 1.f ? 1.f : u);
    uint64_t q = (uint64_t)(u * 2097151.0f);
    code |= spread21(q) << (2 - k);
  }
  keys[i] = code;
  vals[i] = i;
}

// ------------------------------------------------------------------ radix sort
#define RS_ITEMS 16
#define RS_TILE (GXY_BLOCK * RS_ITEMS)

// digit histogram of one tile, written digit-major (hist[d * ntiles + tile])
__global__ __launch_bounds__(GXY_BLOCK) void k_rs_hist(const uint64_t *keys, int n, int shift, int ntiles, int *hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ITEMS; k++) {
    const int64_t i = base + (int64_t)k * GXY_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(int)(keys[i] >> shift) & 255], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__device__ __forceinline__ unsigned long long rs_lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Stable scatter of one tile: keys are taken in index order (slot by slot,
// wave by wave, lane by lane); a key's destination is its digit's global base
// for this tile + the keys of the same digit before it in the tile.
__global__ __launch_bounds__(GXY_BLOCK) void k_rs_scatter(const uint64_t *kin, const int *vin, int n, int shift,
                                                          int ntiles, const int *off, uint64_t *kout, int *vout) {
  __shared__ int run[256];
  __shared__ int wc[GXY_BLOCK / 64][256];
  const int t = threadIdx.x, w = t >> 6;
  run[t] = off[(int64_t)t * ntiles + blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ITEMS; k++) {
    const int64_t i = base + (int64_t)k * GXY_BLOCK + t;
    const bool valid = i < n;
    uint64_t key = 0;
    int val = 0, d = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
      d = (int)(key >> shift) & 255;
    }
    for (int q = 0; q < GXY_BLOCK / 64; q++) wc[q][t] = 0;
    unsigned long long m = __ballot(valid);
    for (int b = 0; b < 8; b++) {
      const unsigned long long bb = __ballot((d >> b) & 1);
      m &= ((d >> b) & 1) ? bb : ~bb;
    }
    const int rk = __popcll(m & rs_lanemask_lt());
    __syncthreads();
    if (valid && rk == 0) wc[w][d] = __popcll(m);
    __syncthreads();
    if (valid) {
      int pos = run[d] + rk;
      for (int q = 0; q < w; q++) pos += wc[q][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    int add = 0;
    for (int q = 0; q < GXY_BLOCK / 64; q++) add += wc[q][t];
    run[t] += add;
    __syncthreads();
  }
}

// Sorts n (key, value) pairs by the low `bits` bits of the keys (stable);
// the result ends in (k0, v0) or (k1, v1): the return value says which (0/1).
static int radix_sort_pairs(uint64_t *k0, int *v0, uint64_t *k1, int *v1, int n, int bits, int *hist, int *off,
                            void *scan_tmp, hipStream_t s, hipError_t *err) {
  const int ntiles = (n + RS_TILE - 1) / RS_TILE;
  int cur = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    uint64_t *ki = cur ? k1 : k0, *ko = cur ? k0 : k1;
    int *vi = cur ? v1 : v0, *vo = cur ? v0 : v1;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(GXY_BLOCK), 0, s, (const uint64_t *)ki, n, shift, ntiles, hist);
    hipError_t e = launch_exclusive_sum(hist, off, (long long)256 * ntiles, scan_tmp, s);
    if (e != hipSuccess) { *err = e; return cur; }
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(GXY_BLOCK), 0, s, (const uint64_t *)ki, (const int *)vi, n,
                       shift, ntiles, (const int *)off, ko, vo);
    cur ^= 1;
  }
  *err = hipGetLastError();
  return cur;
}

// ------------------------------------------------------------------ Karras topology
__device__ __forceinline__ int delta(const uint64_t *keys, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz((unsigned int)(i ^ j));
  return __clzll(a ^ b);
}

// parent[i] of inner node i, parent[ni + k] of leaf k (ni = n - 1; root -1)
__global__ __launch_bounds__(GXY_BLOCK) void k_karras(const uint64_t *keys, int n, BVHNode *nodes, int *parent) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= n - 1) return;
  const int ni = n - 1;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int g = i + s * d + (d < 0 ? -1 : 0);
  int lo = i < j ? i : j, hi = i < j ? j : i;
  int c0, c1;
  if (lo == g) { c0 = leaf_code(g); parent[ni + g] = i; }
  else { c0 = g; parent[g] = i; }
  if (hi == g + 1) { c1 = leaf_code(g + 1); parent[ni + g + 1] = i; }
  else { c1 = g + 1; parent[g + 1] = i; }
  nodes[i].c0 = c0;
  nodes[i].c1 = c1;
  nodes[i].pad_[0] = nodes[i].pad_[1] = 0;
  if (i == 0) parent[0] = -1;
}

// ------------------------------------------------------------------ refit
// Per inner node: cnt (primitives below), cost (SAH cost as built), costc
// (SAH cost with each subtree of <= GXY_BVH_LEAF primitives made a leaf where
// that is cheaper), leaf (1: costc chose a leaf here), height (1 = both
// children are leaves).  Costs are in units of half surface area.
struct BuildArrays {
  int *cnt;
  float *cost, *costc;
  int *leaf, *height;
  float ct;   // SAH cost of one primitive test
};

__device__ __forceinline__ void child_totals(const BuildArrays &A, int code, const float lo[3], const float hi[3],
                                             int &c, float &cost, float &costc, int &h) {
  if (code >= 0) {
    c = ldi(A.cnt + code);
    cost = ldf(A.cost + code);
    costc = ldf(A.costc + code);
    h = ldi(A.height + code);
  } else {
    c = 1;
    cost = costc = A.ct * half_area(lo, hi);
    h = 0;
  }
}

// Depth of every inner node by pointer jumping over the parent links: after
// round r, anc is the ancestor 2^r levels up (-1 past the root) and dep the
// distance to it (the full depth once anc is -1).  Eight rounds give exact
// depths below 256; deeper nodes keep a depth >= 256.
__global__ __launch_bounds__(GXY_BLOCK) void k_depth_init(const int *parent, int ni, int *anc, int *dep) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= ni) return;
  const int a = parent[i];
  anc[i] = a;
  dep[i] = a >= 0 ? 1 : 0;
}
__global__ __launch_bounds__(GXY_BLOCK) void k_depth_jump(const int *anc, const int *dep, int ni, int *anc2,
                                                          int *dep2) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= ni) return;
  const int a = anc[i];
  dep2[i] = a >= 0 ? dep[i] + dep[a] : dep[i];
  anc2[i] = a >= 0 ? anc[a] : -1;
}

// Counting sort of node indices by a small key (depth or height, clamped to
// GXY_BVH_MAXH - 1).  Equal keys are common within a wave (neighbouring nodes
// sit at the same height), so a wave adds each distinct key once: the lowest
// active lane's key, all lanes holding it, one atomic.  k_level_hist counts;
// k_level_scatter reserves each key's range for its LEVEL_TILE nodes with one
// global atomic per key and places them (order within a key is arbitrary).
#define LEVEL_TILE (GXY_BLOCK * 8)
__device__ __forceinline__ int level_key(const int *key, int64_t i, int n) {
  return i < n ? min(max(key[i], 0), GXY_BVH_MAXH - 1) : -1;
}
__global__ __launch_bounds__(GXY_BLOCK) void k_level_hist(const int *key, int n, int *hist) {
  __shared__ int h[GXY_BVH_MAXH];
  for (int t = threadIdx.x; t < GXY_BVH_MAXH; t += GXY_BLOCK) h[t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * LEVEL_TILE;
  for (int r = 0; r < LEVEL_TILE / GXY_BLOCK; r++) {
    const int k = level_key(key, base + (int64_t)r * GXY_BLOCK + threadIdx.x, n);
    unsigned long long act = __ballot(k >= 0);
    while (act) {
      const int kk = __shfl(k, __ffsll((long long)act) - 1, 64);
      const unsigned long long same = __ballot(k == kk) & act;
      if (k == kk && __popcll(same & rs_lanemask_lt()) == 0) atomicAdd(&h[kk], __popcll(same));
      act &= ~same;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < GXY_BVH_MAXH; t += GXY_BLOCK)
    if (h[t]) atomicAdd(hist + t, h[t]);
}
__global__ __launch_bounds__(GXY_BLOCK) void k_level_scatter(const int *key, int n, int *cursor, int *list) {
  __shared__ int h[GXY_BVH_MAXH];
  for (int t = threadIdx.x; t < GXY_BVH_MAXH; t += GXY_BLOCK) h[t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * LEVEL_TILE;
  for (int r = 0; r < LEVEL_TILE / GXY_BLOCK; r++) {
    const int k = level_key(key, base + (int64_t)r * GXY_BLOCK + threadIdx.x, n);
    unsigned long long act = __ballot(k >= 0);
    while (act) {
      const int kk = __shfl(k, __ffsll((long long)act) - 1, 64);
      const unsigned long long same = __ballot(k == kk) & act;
      if (k == kk && __popcll(same & rs_lanemask_lt()) == 0) atomicAdd(&h[kk], __popcll(same));
      act &= ~same;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < GXY_BVH_MAXH; t += GXY_BLOCK)
    if (h[t]) h[t] = atomicAdd(cursor + t, h[t]);   // this tile's first position for key t
  __syncthreads();
  for (int r = 0; r < LEVEL_TILE / GXY_BLOCK; r++) {
    const int64_t i = base + (int64_t)r * GXY_BLOCK + threadIdx.x;
    const int k = level_key(key, i, n);
    unsigned long long act = __ballot(k >= 0);
    while (act) {
      const int kk = __shfl(k, __ffsll((long long)act) - 1, 64);
      const unsigned long long same = __ballot(k == kk) & act;
      int first = 0;
      if (k == kk && __popcll(same & rs_lanemask_lt()) == 0) first = atomicAdd(&h[kk], __popcll(same));
      first = __shfl(first, __ffsll((long long)same) - 1, 64);
      if (k == kk) list[first + __popcll(same & rs_lanemask_lt())] = (int)i;
      act &= ~same;
    }
  }
}

// Refit of the inner nodes of one depth (every deeper node done by earlier
// launches): the children's boxes go into this node's slots, and the totals
// of its subtree are formed from the children's.
__global__ __launch_bounds__(GXY_BLOCK) void k_refit_level(const float *boxes, const int *order, BVHNode *nodes,
                                                           const int *list, int count, BuildArrays A) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= count) return;
  const int p = list[i];
  BVHNode *nd = nodes + p;
  float l[2][3], h[2][3], s[2], q[2];
  int c[2], g[2];
  for (int slot = 0; slot < 2; slot++) {
    const int code = slot ? nd->c1 : nd->c0;
    if (code >= 0) {
      const BVHNode *ch = nodes + code;
      for (int a = 0; a < 3; a++) {
        l[slot][a] = fminf(ch->lo0[a], ch->lo1[a]);
        h[slot][a] = fmaxf(ch->hi0[a], ch->hi1[a]);
      }
    } else {
      const float *pb = boxes + 6 * (int64_t)order[(~code) >> 4];
      for (int a = 0; a < 3; a++) { l[slot][a] = pb[a]; h[slot][a] = pb[3 + a]; }
    }
    child_totals(A, code, l[slot], h[slot], c[slot], s[slot], q[slot], g[slot]);
  }
  float lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    nd->lo0[a] = l[0][a]; nd->hi0[a] = h[0][a];
    nd->lo1[a] = l[1][a]; nd->hi1[a] = h[1][a];
    lo[a] = fminf(l[0][a], l[1][a]);
    hi[a] = fmaxf(h[0][a], h[1][a]);
  }
  const float ar = half_area(lo, hi);
  const int cc = c[0] + c[1];
  const float cost = BVH_CI * ar + s[0] + s[1];
  const float inner = BVH_CI * ar + q[0] + q[1], asleaf = A.ct * ar * (float)cc;
  const bool lf = p != 0 && cc <= GXY_BVH_LEAF && asleaf <= inner;
  A.cnt[p] = cc;
  A.cost[p] = cost;
  A.costc[p] = lf ? asleaf : inner;
  A.leaf[p] = lf ? 1 : 0;
  A.height[p] = max(g[0], g[1]) + 1;
}

// ------------------------------------------------------------------ treelets
#define TL 7                 // treelet leaves
#define TLS (1 << TL)        // subsets
#define TB 64                // work-items per workgroup of the treelet kernel

// Restructures the treelet rooted at p (the costs below it final).  LDS
// columns: cost and best split of every subset (entry s of this work-item at
// s * TB).  Returns the treelet root's cost as restructured (or as it was).
__device__ float treelet(BVHNode *nodes, int *parent, const BuildArrays &A, int ni, int p, float cur, float *sc,
                         unsigned char *sp) {
  int code[TL];
  float bl[TL][3], bh[TL][3];
  int cntl[TL];
  float cst[TL];
  int inner[TL - 1];
  // grow the treelet: open the inner leaf of largest area until there are 7
  int nl = 2, ni_ = 1;
  inner[0] = p;
  {
    float l[3], h[3];
    slot_box(nodes + p, 0, l, h);
    code[0] = ldi(&nodes[p].c0);
    for (int a = 0; a < 3; a++) { bl[0][a] = l[a]; bh[0][a] = h[a]; }
    slot_box(nodes + p, 1, l, h);
    code[1] = ldi(&nodes[p].c1);
    for (int a = 0; a < 3; a++) { bl[1][a] = l[a]; bh[1][a] = h[a]; }
  }
#pragma unroll
  for (int grow = 2; grow < TL; grow++) {
    int best = -1;
    float ba = -1.f;
#pragma unroll
    for (int j = 0; j < TL; j++)
      if (j < nl && code[j] >= 0) {
        const float a = half_area(bl[j], bh[j]);
        if (a > ba) { ba = a; best = j; }
      }
    if (best < 0) break;
    int q = 0;
#pragma unroll
    for (int j = 0; j < TL; j++)
      if (j == best) q = code[j];
    float l0[3], h0[3], l1[3], h1[3];
    slot_box(nodes + q, 0, l0, h0);
    slot_box(nodes + q, 1, l1, h1);
    const int q0 = ldi(&nodes[q].c0), q1 = ldi(&nodes[q].c1);
#pragma unroll
    for (int j = 0; j < TL - 1; j++)
      if (j == ni_) inner[j] = q;
    ni_++;
#pragma unroll
    for (int j = 0; j < TL; j++) {
      if (j == best) {
        code[j] = q0;
        for (int a = 0; a < 3; a++) { bl[j][a] = l0[a]; bh[j][a] = h0[a]; }
      }
      if (j == nl) {
        code[j] = q1;
        for (int a = 0; a < 3; a++) { bl[j][a] = l1[a]; bh[j][a] = h1[a]; }
      }
    }
    nl++;
  }
  if (nl < TL) return cur;
#pragma unroll
  for (int j = 0; j < TL; j++) {
    int g;
    float s, q;
    child_totals(A, code[j], bl[j], bh[j], cntl[j], s, q, g);
    cst[j] = q;
  }
  // optimal split of every subset, singletons first (a subset's proper
  // subsets are smaller numbers)
  for (int s = 1; s < TLS; s++) {
    float l[3] = {3e38f, 3e38f, 3e38f}, h[3] = {-3e38f, -3e38f, -3e38f};
    int c = 0, single = -1;
#pragma unroll
    for (int j = 0; j < TL; j++)
      if ((s >> j) & 1) {
        for (int a = 0; a < 3; a++) { l[a] = fminf(l[a], bl[j][a]); h[a] = fmaxf(h[a], bh[j][a]); }
        c += cntl[j];
        single = j;
      }
    if ((s & (s - 1)) == 0) {
      float v = 0.f;
#pragma unroll
      for (int j = 0; j < TL; j++)
        if (j == single) v = cst[j];
      sc[s * TB] = v;
      sp[s * TB] = 0;
      continue;
    }
    const int low = s & -s, rest = s ^ low;
    float best = 3e38f;
    int bp = low;
    for (int q = rest;; q = (q - 1) & rest) {
      const int pset = q | low;
      if (pset != s) {
        const float v = sc[pset * TB] + sc[(s ^ pset) * TB];
        if (v < best) { best = v; bp = pset; }
      }
      if (q == 0) break;
    }
    const float ar = half_area(l, h);
    float v = BVH_CI * ar + best;
    if (c <= GXY_BVH_LEAF) v = fminf(v, A.ct * ar * (float)c);
    sc[s * TB] = v;
    sp[s * TB] = (unsigned char)bp;
  }
  const float best = sc[(TLS - 1) * TB];
  if (!(best < cur * 0.99999f)) return cur;
  // rebuild: subsets in a queue, inner node k of the treelet for entry k
  int qs[TL - 1] = {TLS - 1, 0, 0, 0, 0, 0};
  int nq = 1;
#pragma unroll
  for (int k = 0; k < TL - 1; k++) {
    int s = 0, node = 0;
#pragma unroll
    for (int j = 0; j < TL - 1; j++)
      if (j == k) { s = qs[j]; node = inner[j]; }
    if (s == 0 || (s & (s - 1)) == 0) continue;   // (a full binary tree of 7 leaves fills all six)
    const int left = sp[s * TB], right = s ^ left;
    int ccount = 0;
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const int sub = side ? right : left;
      float l[3] = {3e38f, 3e38f, 3e38f}, h[3] = {-3e38f, -3e38f, -3e38f};
      int c = 0, single = -1;
#pragma unroll
      for (int j = 0; j < TL; j++)
        if ((sub >> j) & 1) {
          for (int a = 0; a < 3; a++) { l[a] = fminf(l[a], bl[j][a]); h[a] = fmaxf(h[a], bh[j][a]); }
          c += cntl[j];
          single = j;
        }
      ccount += c;
      int ccode;
      if ((sub & (sub - 1)) == 0) {
        ccode = 0;
#pragma unroll
        for (int j = 0; j < TL; j++)
          if (j == single) ccode = code[j];
      } else {
        ccode = 0;
#pragma unroll
        for (int j = 0; j < TL - 1; j++)
          if (j == nq) { ccode = inner[j]; qs[j] = sub; }
        nq++;
      }
      set_slot(nodes + node, side, ccode, l, h);
      if (ccode >= 0) sti(parent + ccode, node);
      else sti(parent + ni + ((~ccode) >> 4), node);
    }
    sti(A.cnt + node, ccount);
    stf(A.cost + node, sc[s * TB]);
    stf(A.costc + node, sc[s * TB]);
    sti(A.leaf + node, 0);
  }
  return best;
}

// The inner nodes of one height: refresh the node's totals from its children
// (final: lower heights ran before), and restructure its treelet when it has
// at least gamma primitives below it.
__global__ __launch_bounds__(TB) void k_treelet_level(const int *list, int count, BVHNode *nodes, int *parent, int ni,
                                                      BuildArrays A, int gamma) {
  __shared__ float s_cost[TLS * TB];
  __shared__ unsigned char s_split[TLS * TB];
  const int i = blockIdx.x * TB + threadIdx.x;
  if (i >= count) return;
  const int p = list[i];
  BVHNode *nd = nodes + p;
  float l0[3], h0[3], l1[3], h1[3];
  slot_box(nd, 0, l0, h0);
  slot_box(nd, 1, l1, h1);
  int n0, n1, g0, g1;
  float s0, s1, q0, q1;
  child_totals(A, ldi(&nd->c0), l0, h0, n0, s0, q0, g0);
  child_totals(A, ldi(&nd->c1), l1, h1, n1, s1, q1, g1);
  float lo[3], hi[3];
  for (int a = 0; a < 3; a++) { lo[a] = fminf(l0[a], l1[a]); hi[a] = fmaxf(h0[a], h1[a]); }
  const float ar = half_area(lo, hi);
  const int c = n0 + n1;
  float cost = BVH_CI * ar + q0 + q1;
  if (c <= GXY_BVH_LEAF && p != 0) cost = fminf(cost, A.ct * ar * (float)c);
  if (c >= gamma) cost = treelet(nodes, parent, A, ni, p, cost, s_cost + threadIdx.x, s_split + threadIdx.x);
  sti(A.cnt + p, c);
  stf(A.costc + p, cost);
}

// ------------------------------------------------------------------ layout
// A frontier entry: old inner node, new node index, first primitive.
struct LayEnt {
  int old, idx, first;
};

// Primitives of the subtree under code in depth-first order (at most
// GXY_BVH_LEAF of them, so at most GXY_BVH_LEAF - 1 inner nodes).
__device__ int subtree_prims(const BVHNode *nodes, int code, int *out) {
  int st[GXY_BVH_LEAF];
  int sp = 0, n = 0;
  st[sp++] = code;
  while (sp > 0) {
    const int c = st[--sp];
    if (c < 0) {
      if (n < GXY_BVH_LEAF) out[n] = (~c) >> 4;
      n++;
    } else if (sp + 2 <= GXY_BVH_LEAF) {
      st[sp++] = nodes[c].c1;
      st[sp++] = nodes[c].c0;
    } else {
      return GXY_BVH_LEAF + 1;   // malformed: more than a leaf holds
    }
  }
  return n;
}

__global__ __launch_bounds__(GXY_BLOCK) void k_layout(const LayEnt *fr, int count, const BVHNode *nodes, int n,
                                                      const int *cnt, const int *leaf, const int *order,
                                                      BVHNode *out, int *neworder, int *ctr, LayEnt *next,
                                                      int *bad, int cap) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= count) return;
  const LayEnt e = fr[i];
  if (e.idx < 0 || e.idx >= cap) {   // more inner nodes than the tree has: malformed
    atomicAdd(bad, 1);
    return;
  }
  const BVHNode nd = nodes[e.old];
  const int cc[2] = {nd.c0, nd.c1};
  int ninner = 0;
  for (int s = 0; s < 2; s++) ninner += (cc[s] >= 0 && !leaf[cc[s]]) ? 1 : 0;
  int base = 0, nslot = 0;
  if (ninner) {
    base = atomicAdd(ctr, ninner);
    nslot = atomicAdd(ctr + 1, ninner);
    if (nslot + ninner > cap) {
      atomicAdd(bad, 1);
      return;
    }
  }
  BVHNode o = nd;
  int first = e.first, k = 0;
  for (int s = 0; s < 2; s++) {
    const int c = cc[s];
    int code;
    if (c < 0) {
      if (first < n) neworder[first] = order[(~c) >> 4];
      else atomicAdd(bad, 1);
      code = leaf_code_n(first, 1);
      first += 1;
    } else if (leaf[c]) {
      int prims[GXY_BVH_LEAF];
      const int m = subtree_prims(nodes, c, prims);
      if (m < 1 || m > GXY_BVH_LEAF || first + m > n) {
        atomicAdd(bad, 1);
        code = leaf_code_n(first, 1);
      } else {
        for (int j = 0; j < m; j++) neworder[first + j] = order[prims[j]];
        code = leaf_code_n(first, m);
        first += m;
      }
    } else {
      const int idx = base + k;
      next[nslot + k] = {c, idx, first};
      k++;
      code = idx;
      first += cnt[c];
    }
    if (s == 0) o.c0 = code;
    else o.c1 = code;
  }
  o.pad_[0] = o.pad_[1] = 0;
  out[e.idx] = o;
}

// ------------------------------------------------------------------ 4-wide collapse
struct WideEnt {
  int bin, idx;
};

__global__ __launch_bounds__(GXY_BLOCK) void k_wide(const WideEnt *fr, int count, const BVHNode *nodes,
                                                    BVH4Node *out, int *ctr, WideEnt *next, int cap, int *bad) {
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= count) return;
  const WideEnt e = fr[i];
  if (e.idx < 0 || e.idx >= cap || e.bin < 0 || e.bin >= cap) {
    atomicAdd(bad, 1);
    return;
  }
  int code[4];
  float lo[4][3], hi[4][3];
  const BVHNode b = nodes[e.bin];
  code[0] = b.c0;
  code[1] = b.c1;
  for (int a = 0; a < 3; a++) { lo[0][a] = b.lo0[a]; hi[0][a] = b.hi0[a]; lo[1][a] = b.lo1[a]; hi[1][a] = b.hi1[a]; }
  int m = 2;
#pragma unroll
  for (int grow = 2; grow < 4; grow++) {
    int best = -1;
    float ba = -1.f;
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (j < m && code[j] >= 0) {
        const float a = half_area(lo[j], hi[j]);
        if (a > ba) { ba = a; best = j; }
      }
    if (best < 0) break;
    int q = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (j == best) q = code[j];
    const BVHNode c = nodes[q];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (j == best) {
        code[j] = c.c0;
        for (int a = 0; a < 3; a++) { lo[j][a] = c.lo0[a]; hi[j][a] = c.hi0[a]; }
      }
      if (j == m) {
        code[j] = c.c1;
        for (int a = 0; a < 3; a++) { lo[j][a] = c.lo1[a]; hi[j][a] = c.hi1[a]; }
      }
    }
    m++;
  }
  int ninner = 0;
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (j < m && code[j] >= 0) ninner++;
  int base = 0, nslot = 0;
  if (ninner) {
    base = atomicAdd(ctr, ninner);
    nslot = atomicAdd(ctr + 1, ninner);
    if (nslot + ninner > cap) {
      atomicAdd(bad, 1);
      return;
    }
  }
  BVH4Node o;
  int k = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (j < m) {
      o.lox[j] = lo[j][0]; o.loy[j] = lo[j][1]; o.loz[j] = lo[j][2];
      o.hix[j] = hi[j][0]; o.hiy[j] = hi[j][1]; o.hiz[j] = hi[j][2];
      if (code[j] >= 0) {
        next[nslot + k] = {code[j], base + k};
        o.child[j] = base + k;
        k++;
      } else {
        o.child[j] = code[j];
      }
    } else {
      o.lox[j] = o.loy[j] = o.loz[j] = __builtin_inff();
      o.hix[j] = o.hiy[j] = o.hiz[j] = -__builtin_inff();
      o.child[j] = 0x7fffffff;
    }
    o.pad_[j] = 0;
  }
  out[e.idx] = o;
}

// ------------------------------------------------------------------ SAH of the final tree
__global__ __launch_bounds__(GXY_BLOCK) void k_sah(const BVHNode *nodes, int nn, float ct, double *acc) {
  __shared__ double red[GXY_BLOCK / 64];
  const int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  double v = 0.0;
  if (i < nn) {
    const BVHNode b = nodes[i];
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) { lo[a] = fminf(b.lo0[a], b.lo1[a]); hi[a] = fmaxf(b.hi0[a], b.hi1[a]); }
    v = BVH_CI * (double)half_area(lo, hi);
    const int cc[2] = {b.c0, b.c1};
    for (int s = 0; s < 2; s++)
      if (cc[s] < 0)
        v += ct * (double)((((~cc[s]) & 15) + 1)) * (double)half_area(s ? b.lo1 : b.lo0, s ? b.hi1 : b.hi0);
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < GXY_BLOCK / 64; w++) s += red[w];
    atomicAdd(acc, s);
  }
}

// ------------------------------------------------------------------ driver
namespace {
struct Pool {
  alloc_fn_t alloc;
  free_fn_t dfree;
  std::vector<void *> held;
  void *get(size_t bytes) {
    void *p = alloc(bytes ? bytes : 1);
    if (p) held.push_back(p);
    return p;
  }
  void release(void *p) {
    for (auto &q : held)
      if (q == p) { dfree(p); q = nullptr; }
  }
  ~Pool() {
    for (void *p : held)
      if (p) dfree(p);
  }
};
}  // namespace

#define BVH_CHK(x)                      \
  do {                                  \
    const hipError_t e_ = (x);          \
    if (e_ != hipSuccess) return e_;    \
  } while (0)

static hipError_t read_ints(const int *dev, int *host, int n, hipStream_t s) {
  BVH_CHK(hipMemcpyAsync(host, dev, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  return hipStreamSynchronize(s);
}

// Node indices sorted by key (see k_level_hist): cnt[k] and start[k] of each
// key on the host, the list on the device.  hh holds 2 * GXY_BVH_MAXH ints.
static hipError_t sort_by_level(const int *key, int n, int *hh, int *cnt, int *start, int *list, hipStream_t s) {
  const dim3 grid((n + LEVEL_TILE - 1) / LEVEL_TILE);
  BVH_CHK(hipMemsetAsync(hh, 0, sizeof(int) * GXY_BVH_MAXH, s));
  hipLaunchKernelGGL(k_level_hist, grid, dim3(GXY_BLOCK), 0, s, key, n, hh);
  BVH_CHK(read_ints(hh, cnt, GXY_BVH_MAXH, s));
  start[0] = 0;
  for (int k = 0; k < GXY_BVH_MAXH; k++) start[k + 1] = start[k] + cnt[k];
  if (cnt[GXY_BVH_MAXH - 1] > 0) return hipErrorInvalidValue;   // deeper than the level table
  BVH_CHK(hipMemcpyAsync(hh + GXY_BVH_MAXH, start, sizeof(int) * GXY_BVH_MAXH, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_level_scatter, grid, dim3(GXY_BLOCK), 0, s, key, n, hh + GXY_BVH_MAXH, list);
  return hipGetLastError();
}

// Bottom-up refit of the ni inner nodes, one launch per depth, deepest
// first; t0..t3 are ni ints each of scratch, list ni ints.
static hipError_t refit_levels(const float *boxes, const int *order, int ni, BVHNode *nodes, const int *parent,
                               const BuildArrays &A, int *t0, int *t1, int *t2, int *t3, int *hh, int *list,
                               hipStream_t s) {
  const dim3 grid((ni + GXY_BLOCK - 1) / GXY_BLOCK);
  hipLaunchKernelGGL(k_depth_init, grid, dim3(GXY_BLOCK), 0, s, parent, ni, t0, t1);
  int *anc = t0, *dep = t1, *anc2 = t2, *dep2 = t3;
  for (int r = 0; r < 8; r++) {
    hipLaunchKernelGGL(k_depth_jump, grid, dim3(GXY_BLOCK), 0, s, (const int *)anc, (const int *)dep, ni, anc2, dep2);
    int *t = anc; anc = anc2; anc2 = t;
    t = dep; dep = dep2; dep2 = t;
  }
  int cnt[GXY_BVH_MAXH], start[GXY_BVH_MAXH + 1];
  BVH_CHK(sort_by_level(dep, ni, hh, cnt, start, list, s));
  for (int d = GXY_BVH_MAXH - 2; d >= 0; d--)
    if (cnt[d] > 0)
      hipLaunchKernelGGL(k_refit_level, dim3((cnt[d] + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, boxes,
                         order, nodes, (const int *)(list + start[d]), cnt[d], A);
  return hipGetLastError();
}

hipError_t build_lbvh(const float *boxes_dev, int n, float prim_cost, alloc_fn_t alloc, free_fn_t dfree,
                      BVHBuildResult *out, hipStream_t s) {
  memset(out, 0, sizeof *out);
  if (n <= 0) return hipSuccess;
  Pool P{alloc, dfree, {}};
  if (n == 1) {
    BVHNode h;
    float b[6];
    BVH_CHK(hipMemcpyAsync(b, boxes_dev, 6 * sizeof(float), hipMemcpyDeviceToHost, s));
    BVH_CHK(hipStreamSynchronize(s));
    for (int a = 0; a < 3; a++) {
      h.lo0[a] = b[a]; h.hi0[a] = b[3 + a];
      h.lo1[a] = 3.0e38f; h.hi1[a] = -3.0e38f;
    }
    h.c0 = leaf_code(0);
    h.c1 = leaf_code(0);
    h.pad_[0] = h.pad_[1] = 0;
    BVH4Node w;
    for (int j = 0; j < 4; j++) {
      w.lox[j] = j ? INFINITY : b[0]; w.loy[j] = j ? INFINITY : b[1]; w.loz[j] = j ? INFINITY : b[2];
      w.hix[j] = j ? -INFINITY : b[3]; w.hiy[j] = j ? -INFINITY : b[4]; w.hiz[j] = j ? -INFINITY : b[5];
      w.child[j] = j ? 0x7fffffff : leaf_code(0);
      w.pad_[j] = 0;
    }
    out->nodes = (BVHNode *)alloc(sizeof(BVHNode));
    out->order = (int *)alloc(sizeof(int));
    out->nodes4 = (BVH4Node *)alloc(sizeof(BVH4Node));
    if (!out->nodes || !out->order || !out->nodes4) return hipErrorOutOfMemory;
    int zero = 0;
    BVH_CHK(hipMemcpyAsync(out->nodes, &h, sizeof h, hipMemcpyHostToDevice, s));
    BVH_CHK(hipMemcpyAsync(out->order, &zero, sizeof zero, hipMemcpyHostToDevice, s));
    BVH_CHK(hipMemcpyAsync(out->nodes4, &w, sizeof w, hipMemcpyHostToDevice, s));
    BVH_CHK(hipStreamSynchronize(s));
    out->nnodes = 1;
    out->depth = 1;
    out->nnodes4 = 1;
    out->depth4 = 1;
    out->leaves = 1;
    out->sah = BVH_CI + prim_cost;   // the root, then its one primitive, over the same surface
    return hipGetLastError();
  }
  const int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK, ni = n - 1;
  const int ntiles = (n + RS_TILE - 1) / RS_TILE;
  unsigned int *bnd = (unsigned int *)P.get(8 * sizeof(unsigned int));
  uint64_t *k0 = (uint64_t *)P.get(sizeof(uint64_t) * n), *k1 = (uint64_t *)P.get(sizeof(uint64_t) * n);
  int *v0 = (int *)P.get(sizeof(int) * n), *v1 = (int *)P.get(sizeof(int) * n);
  int *hist = (int *)P.get(sizeof(int) * 256 * (size_t)ntiles);
  int *hoff = (int *)P.get(sizeof(int) * 256 * (size_t)ntiles);
  void *stmp = P.get(scan_temp_bytes((long long)256 * ntiles));
  if (!bnd || !k0 || !k1 || !v0 || !v1 || !hist || !hoff || !stmp) return hipErrorOutOfMemory;
  // 1-2: Morton codes, sorted
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, s, bnd);
  hipLaunchKernelGGL(k_centroid_bounds, dim3(nb < 4096 ? nb : 4096), dim3(GXY_BLOCK), 0, s, boxes_dev, n, bnd);
  hipLaunchKernelGGL(k_morton, dim3(nb), dim3(GXY_BLOCK), 0, s, boxes_dev, n, (const unsigned int *)bnd, k0, v0);
  hipError_t e = hipSuccess;
  const int res = radix_sort_pairs(k0, v0, k1, v1, n, 63, hist, hoff, stmp, s, &e);
  BVH_CHK(e);
  uint64_t *keys = res ? k1 : k0;
  int *order = res ? v1 : v0;
  P.release(hist);
  P.release(hoff);
  P.release(stmp);
  // 3: topology
  BVHNode *nodes = (BVHNode *)P.get(sizeof(BVHNode) * (size_t)ni);
  int *parent = (int *)P.get(sizeof(int) * (size_t)(ni + n));
  int *t0 = (int *)P.get(sizeof(int) * (size_t)ni), *t1 = (int *)P.get(sizeof(int) * (size_t)ni);
  int *t2 = (int *)P.get(sizeof(int) * (size_t)ni), *t3 = (int *)P.get(sizeof(int) * (size_t)ni);
  BuildArrays A;
  A.ct = prim_cost;
  A.cnt = (int *)P.get(sizeof(int) * (size_t)ni);
  A.cost = (float *)P.get(sizeof(float) * (size_t)ni);
  A.costc = (float *)P.get(sizeof(float) * (size_t)ni);
  A.leaf = (int *)P.get(sizeof(int) * (size_t)ni);
  A.height = (int *)P.get(sizeof(int) * (size_t)ni);
  int *hh = (int *)P.get(sizeof(int) * 2 * GXY_BVH_MAXH);
  int *hlist = (int *)P.get(sizeof(int) * (size_t)ni);
  if (!nodes || !parent || !t0 || !t1 || !t2 || !t3 || !A.cnt || !A.cost || !A.costc || !A.leaf || !A.height || !hh || !hlist)
    return hipErrorOutOfMemory;
  hipLaunchKernelGGL(k_karras, dim3((ni + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, (const uint64_t *)keys,
                     n, nodes, parent);
  P.release(k0);
  P.release(k1);
  // 4-5: refit, then treelet rounds (each followed by a refit for heights and totals)
  int maxh = 0;
  for (int round = 0; round <= 3; round++) {
    BVH_CHK(refit_levels(boxes_dev, (const int *)order, ni, nodes, (const int *)parent, A, t0, t1, t2, t3, hh, hlist,
                         s));
    if (round == 3 || ni < TL) break;
    // inner nodes by height
    int hcount[GXY_BVH_MAXH], hstart[GXY_BVH_MAXH + 1];
    BVH_CHK(sort_by_level((const int *)A.height, ni, hh, hcount, hstart, hlist, s));
    const int gamma = TL << round;
    for (int h = 1; h < GXY_BVH_MAXH; h++) {
      if (hcount[h] == 0) continue;
      maxh = h > maxh ? h : maxh;
      hipLaunchKernelGGL(k_treelet_level, dim3((hcount[h] + TB - 1) / TB), dim3(TB), 0, s,
                         (const int *)(hlist + hstart[h]), hcount[h], nodes, parent, ni, A, gamma);
    }
    BVH_CHK(hipGetLastError());
  }
  // 6: layout (leaf collapse, renumbering, primitive order)
  BVHNode *fin = (BVHNode *)P.get(sizeof(BVHNode) * (size_t)ni);
  int *neworder = (int *)P.get(sizeof(int) * (size_t)n);
  LayEnt *fa = (LayEnt *)P.get(sizeof(LayEnt) * (size_t)ni), *fb = (LayEnt *)P.get(sizeof(LayEnt) * (size_t)ni);
  int *ctr = (int *)P.get(sizeof(int) * 8);
  if (!fin || !neworder || !fa || !fb || !ctr) return hipErrorOutOfMemory;
  {
    const LayEnt root = {0, 0, 0};
    const int init[8] = {1, 0, 0, 0, 0, 0, 0, 0};   // nodes used, next-frontier entries, malformed leaves
    BVH_CHK(hipMemcpyAsync(fa, &root, sizeof root, hipMemcpyHostToDevice, s));
    BVH_CHK(hipMemcpyAsync(ctr, init, sizeof init, hipMemcpyHostToDevice, s));
  }
  int count = 1, levels = 0, nn = 1;
  while (count > 0) {
    if (++levels > GXY_BVH_MAXH) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_layout, dim3((count + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, (const LayEnt *)fa,
                       count, (const BVHNode *)nodes, n, (const int *)A.cnt, (const int *)A.leaf,
                       (const int *)order, fin, neworder, ctr, fb, ctr + 2, ni);
    int c3[3];
    BVH_CHK(read_ints(ctr, c3, 3, s));
    if (c3[2] != 0) return hipErrorInvalidValue;
    nn = c3[0];
    count = c3[1];
    BVH_CHK(hipMemsetAsync(ctr + 1, 0, sizeof(int), s));
    LayEnt *t = fa;
    fa = fb;
    fb = t;
  }
  P.release(nodes);
  P.release(parent);
  P.release(t0);
  P.release(t1);
  P.release(t2);
  P.release(t3);
  // 7: 4-wide collapse
  BVH4Node *w4 = (BVH4Node *)P.get(sizeof(BVH4Node) * (size_t)nn);
  WideEnt *wa = (WideEnt *)P.get(sizeof(WideEnt) * (size_t)nn), *wb = (WideEnt *)P.get(sizeof(WideEnt) * (size_t)nn);
  if (!w4 || !wa || !wb) return hipErrorOutOfMemory;
  {
    const WideEnt root = {0, 0};
    const int init[3] = {1, 0, 0};
    BVH_CHK(hipMemcpyAsync(wa, &root, sizeof root, hipMemcpyHostToDevice, s));
    BVH_CHK(hipMemcpyAsync(ctr, init, sizeof init, hipMemcpyHostToDevice, s));
  }
  int wcount = 1, wlevels = 0, nn4 = 1;
  while (wcount > 0) {
    if (++wlevels > GXY_BVH_MAXH) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_wide, dim3((wcount + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s,
                       (const WideEnt *)wa, wcount, (const BVHNode *)fin, w4, ctr, wb, nn, ctr + 2);
    int c2[3];
    BVH_CHK(read_ints(ctr, c2, 3, s));
    if (c2[2] != 0) return hipErrorInvalidValue;
    nn4 = c2[0];
    wcount = c2[1];
    BVH_CHK(hipMemsetAsync(ctr + 1, 0, sizeof(int), s));
    WideEnt *t = wa;
    wa = wb;
    wb = t;
  }
  // SAH cost of the final binary tree, relative to the root's surface
  double *acc = (double *)P.get(sizeof(double));
  if (!acc) return hipErrorOutOfMemory;
  BVH_CHK(hipMemsetAsync(acc, 0, sizeof(double), s));
  hipLaunchKernelGGL(k_sah, dim3((nn + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, (const BVHNode *)fin, nn,
                     prim_cost, acc);
  double sah = 0.0;
  BVHNode root;
  BVH_CHK(hipMemcpyAsync(&sah, acc, sizeof sah, hipMemcpyDeviceToHost, s));
  BVH_CHK(hipMemcpyAsync(&root, fin, sizeof root, hipMemcpyDeviceToHost, s));
  BVH_CHK(hipStreamSynchronize(s));
  float rl[3], rh[3];
  for (int a = 0; a < 3; a++) { rl[a] = fminf(root.lo0[a], root.lo1[a]); rh[a] = fmaxf(root.hi0[a], root.hi1[a]); }
  const float rx = rh[0] - rl[0], ry = rh[1] - rl[1], rz = rh[2] - rl[2];
  const double ra = (double)rx * ry + (double)ry * rz + (double)rz * rx;
  // hand the results over (the pool frees everything else)
  for (auto &q : P.held)
    if (q == fin || q == neworder || q == w4) q = nullptr;
  out->nodes = fin;
  out->order = neworder;
  out->nnodes = nn;
  out->depth = levels + 1;
  out->nodes4 = w4;
  out->nnodes4 = nn4;
  out->depth4 = wlevels;
  out->sah = ra > 0.0 ? (float)(sah / ra) : 0.f;
  out->leaves = 0;
  out->heights = maxh;
  return hipGetLastError();
}

}  // namespace gxy
