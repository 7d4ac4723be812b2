// bvh.hip — linear BVH construction on the GPU (Karras, "Maximizing Parallelism
// in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012).
//
// 1. centroid bounds (wave reduction + ordered-int atomics);
// 2. 63-bit Morton code of each primitive centroid (21 bits per axis);
// 3. device radix sort of (code, primitive) pairs (rocPRIM via hipCUB);
// 4. one work-item per internal node finds its key range and split with the
//    delta() prefix length (ties between equal codes broken by index);
// 5. leaves climb to the root with an arrival counter per node; the second
//    arrival writes the union box.  Each node stores both child boxes (64 B),
//    which is the layout the traversal loads with four 16-byte reads.
#include <hipcub/hipcub.hpp>

#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

__device__ __forceinline__ unsigned int okey(float f) {
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ounkey(unsigned int k) {
  unsigned int u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

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

__device__ __forceinline__ int delta(const uint64_t *keys, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz((unsigned int)(i ^ j));
  return __clzll(a ^ b);
}

__host__ __device__ __forceinline__ int leaf_code(int k) { return ~(k << 4); }

__global__ __launch_bounds__(GXY_BLOCK) void k_karras(const uint64_t *keys, int n, BVHNode *nodes, int *nparent,
                                                      int *lparent) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= n - 1) return;
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
  if (lo == g) { c0 = leaf_code(g); lparent[g] = i; }
  else { c0 = g; nparent[g] = i; }
  if (hi == g + 1) { c1 = leaf_code(g + 1); lparent[g + 1] = i; }
  else { c1 = g + 1; nparent[g + 1] = i; }
  nodes[i].c0 = c0;
  nodes[i].c1 = c1;
  if (i == 0) nparent[0] = -1;
}

__global__ __launch_bounds__(GXY_BLOCK) void k_refit(const float *boxes, const int *order, int n, BVHNode *nodes,
                                                     const int *nparent, const int *lparent, int *flags,
                                                     int *maxdepth) {
  int k = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (k >= n) return;
  const float *pb = boxes + 6 * (int64_t)order[k];
  float lo[3] = {pb[0], pb[1], pb[2]}, hi[3] = {pb[3], pb[4], pb[5]};
  int code = leaf_code(k);
  int p = lparent[k];
  // depth of this leaf (number of ancestors)
  int depth = 0;
  for (int q = p; q >= 0 && depth < 1024; q = nparent[q]) depth++;
  atomicMax(maxdepth, depth);
  int climb = 0;
  while (p >= 0 && climb++ < 1024) {
    BVHNode *nd = nodes + p;
    int slot = nd->c0 == code ? 0 : 1;
    float *dst = slot == 0 ? nd->lo0 : nd->lo1;
    float *dsth = slot == 0 ? nd->hi0 : nd->hi1;
    for (int a = 0; a < 3; a++) {
      __hip_atomic_store(dst + a, lo[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(dsth + a, hi[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    int prev = atomicAdd(flags + p, 1);
    if (prev == 0) return;  // the sibling subtree finishes this node
    __threadfence();
    const float *olo = slot == 0 ? nd->lo1 : nd->lo0;
    const float *ohi = slot == 0 ? nd->hi1 : nd->hi0;
    for (int a = 0; a < 3; a++) {
      float ol = __hip_atomic_load(olo + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      float oh = __hip_atomic_load(ohi + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      lo[a] = lo[a] < ol ? lo[a] : ol;
      hi[a] = hi[a] > oh ? hi[a] : oh;
    }
    code = p;
    p = nparent[p];
  }
}

hipError_t build_lbvh(const float *boxes_dev, int n, alloc_fn_t alloc, free_fn_t dfree, BVHBuildResult *out,
                      hipStream_t s) {
  out->nodes = nullptr;
  out->order = nullptr;
  out->nnodes = 0;
  out->depth = 0;
  out->sah = 0.f;
  if (n <= 0) return hipSuccess;
  if (n == 1) {
    BVHNode h;
    float b[6];
    hipMemcpyAsync(b, boxes_dev, 6 * sizeof(float), hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    for (int a = 0; a < 3; a++) {
      h.lo0[a] = b[a]; h.hi0[a] = b[3 + a];
      h.lo1[a] = 3.0e38f; h.hi1[a] = -3.0e38f;
    }
    h.c0 = leaf_code(0);
    h.c1 = leaf_code(0);
    h.pad_[0] = h.pad_[1] = 0;
    out->nodes = (BVHNode *)alloc(sizeof(BVHNode));
    out->order = (int *)alloc(sizeof(int));
    int zero = 0;
    hipMemcpyAsync(out->nodes, &h, sizeof h, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(out->order, &zero, sizeof zero, hipMemcpyHostToDevice, s);
    hipStreamSynchronize(s);
    out->nnodes = 1;
    out->depth = 1;
    return hipGetLastError();
  }
  const int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  unsigned int *bnd = (unsigned int *)alloc(8 * sizeof(unsigned int));
  uint64_t *k0 = (uint64_t *)alloc(sizeof(uint64_t) * n), *k1 = (uint64_t *)alloc(sizeof(uint64_t) * n);
  int *v0 = (int *)alloc(sizeof(int) * n), *v1 = (int *)alloc(sizeof(int) * n);
  BVHNode *nodes = (BVHNode *)alloc(sizeof(BVHNode) * (size_t)(n - 1));
  int *nparent = (int *)alloc(sizeof(int) * (size_t)(n - 1));
  int *lparent = (int *)alloc(sizeof(int) * (size_t)n);
  int *flags = (int *)alloc(sizeof(int) * (size_t)(n - 1) + sizeof(int));
  int *maxdepth = flags + (n - 1);
  if (!bnd || !k0 || !k1 || !v0 || !v1 || !nodes || !nparent || !lparent || !flags) return hipErrorOutOfMemory;
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, s, bnd);
  hipLaunchKernelGGL(k_centroid_bounds, dim3(nb < 4096 ? nb : 4096), dim3(GXY_BLOCK), 0, s, boxes_dev, n, bnd);
  hipLaunchKernelGGL(k_morton, dim3(nb), dim3(GXY_BLOCK), 0, s, boxes_dev, n, bnd, k0, v0);
  size_t tmp = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, k0, k1, v0, v1, n, 0, 63, s);
  void *dtmp = alloc(tmp > 0 ? tmp : 1);
  if (!dtmp) return hipErrorOutOfMemory;
  hipcub::DeviceRadixSort::SortPairs(dtmp, tmp, k0, k1, v0, v1, n, 0, 63, s);
  hipMemsetAsync(flags, 0, sizeof(int) * (size_t)n, s);
  hipLaunchKernelGGL(k_karras, dim3((n - 1 + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, k1, n, nodes, nparent,
                     lparent);
  hipLaunchKernelGGL(k_refit, dim3(nb), dim3(GXY_BLOCK), 0, s, boxes_dev, v1, n, nodes, nparent, lparent, flags,
                     maxdepth);
  int depth = 0;
  hipMemcpyAsync(&depth, maxdepth, sizeof(int), hipMemcpyDeviceToHost, s);
  hipError_t e = hipStreamSynchronize(s);
  dfree(dtmp);
  dfree(bnd);
  dfree(k0);
  dfree(k1);
  dfree(v0);
  dfree(nparent);
  dfree(lparent);
  dfree(flags);
  if (e != hipSuccess) return e;
  out->nodes = nodes;
  out->order = v1;
  out->nnodes = n - 1;
  out->depth = depth + 1;
  return hipGetLastError();
}

}  // namespace gxy
