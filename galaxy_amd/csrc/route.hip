// route.hip — ray classification, routing, framebuffer accumulation, stable
// partition into per-destination lists, primary-ray generation, and ray copies.
//
// Renderer::Classify / AssignDestinations / HandleTerminatedRays
// (Renderer.cpp:304-502) are scalar host loops in the reference; here one
// kernel applies all three per ray (one work-item per ray, SoA loads), with
// FP32 atomic adds into the framebuffer (Rendering::AddLocalPixels,
// Rendering.cpp:125-153).  processRays_task's per-destination copy
// (Renderer.cpp:561-634) becomes a stable multi-way partition: per-wave ballots
// rank each ray within its category, per-workgroup totals are scanned per
// category, and a scatter moves the 25 fields (100 B) of each surviving ray.
#include <cstring>

#include "route_dev.h"

namespace gxy {

__global__ __launch_bounds__(GXY_BLOCK) void k_route(RouteParams P, RayList_ispc R, int n) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool acc = false;
  int ax = 0, ay = 0;
  float ar = 0.f, ag = 0.f, ab = 0.f, ao = 0.f;
  if (i < n) {
    const int typ = R.type[i];
    int c = R.classification[i];
    if (P.mode & RM_CLASSIFY) c = classify_ray(typ, R.term[i], c, P.reverse);
    if ((P.mode & RM_ASSIGN) && c == RAY_BOUNDARY) {
      int f = exit_face(P.blo, P.bhi, R.ox[i], R.oy[i], R.oz[i], R.dx[i], R.dy[i], R.dz[i]);
      if (P.neighbors[f] >= 0) c = P.neighbors[f];
      else if (typ == RAY_SHADOW || typ == RAY_AO) c = P.reverse ? GXY_DROP_ON_FLOOR : GXY_TERMINATED;
      else c = GXY_TERMINATED;
    }
    if (P.mode & (RM_CLASSIFY | RM_ASSIGN)) R.classification[i] = c;
    if ((P.mode & RM_ACCUMULATE) && c == GXY_TERMINATED) {
      acc = true;
      ax = R.x[i];
      ay = R.y[i];
      ar = R.r[i];
      ag = R.g[i];
      ab = R.b[i];
      ao = R.o[i];
    }
  }
  if (P.mode & RM_ACCUMULATE) fb_accumulate_wave(P.fb, P.width, P.height, acc, ax, ay, ar, ag, ab, ao, P.unique != 0);
}

hipError_t launch_route(const RouteParams &P, const RayList_ispc &R, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_route, dim3((n + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, P, R, n);
  return hipGetLastError();
}

// ------------------------------------------------------------------ partition
#define GXY_MAX_CAT 32

__device__ __forceinline__ unsigned long long lanemask_lt64() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// classification -> category map, passed by value: t[class + 4]
#define GXY_CAT_TABLE 64
struct CatTable {
  int n;
  int t[GXY_CAT_TABLE];
};

__device__ __forceinline__ int category(const RayList_ispc &R, int i, const CatTable &T) {
  int c = R.classification[i] + 4;
  return (c >= 0 && c < T.n) ? T.t[c] : -1;
}

// counts[k * nb + b] = rays of category k in workgroup b
__global__ __launch_bounds__(GXY_BLOCK) void k_part_count(RayList_ispc R, int n, CatTable table, int K,
                                                          int nb, int *counts) {
  __shared__ int wc[GXY_BLOCK / 64][GXY_MAX_CAT];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  int cat = i < n ? category(R, i, table) : -1;
  int w = threadIdx.x >> 6;
  for (int k = 0; k < K; k++) {
    unsigned long long m = __ballot(cat == k);
    if ((threadIdx.x & 63) == 0) wc[w][k] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < K) {
    int s = 0;
    for (int q = 0; q < GXY_BLOCK / 64; q++) s += wc[q][threadIdx.x];
    counts[threadIdx.x * nb + blockIdx.x] = s;
  }
  // a zero after the K*nb counts: their exclusive sum then ends in the total
  if (blockIdx.x == 0 && threadIdx.x == 0) counts[K * nb] = 0;
}

// k_route followed by k_part_count's counting, from the classification in
// registers (one pass over the rays instead of two)
__global__ __launch_bounds__(GXY_BLOCK) void k_route_count(RouteParams P, RayList_ispc R, int n, CatTable table,
                                                           int K, int nb, int *counts) {
  __shared__ int wc[GXY_BLOCK / 64][GXY_MAX_CAT];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool acc = false;
  int ax = 0, ay = 0;
  float ar = 0.f, ag = 0.f, ab = 0.f, ao = 0.f;
  int cat = -1;
  if (i < n) {
    const int typ = R.type[i];
    int c = R.classification[i];
    if (P.mode & RM_CLASSIFY) c = classify_ray(typ, R.term[i], c, P.reverse);
    if ((P.mode & RM_ASSIGN) && c == RAY_BOUNDARY) {
      int f = exit_face(P.blo, P.bhi, R.ox[i], R.oy[i], R.oz[i], R.dx[i], R.dy[i], R.dz[i]);
      if (P.neighbors[f] >= 0) c = P.neighbors[f];
      else if (typ == RAY_SHADOW || typ == RAY_AO) c = P.reverse ? GXY_DROP_ON_FLOOR : GXY_TERMINATED;
      else c = GXY_TERMINATED;
    }
    if (P.mode & (RM_CLASSIFY | RM_ASSIGN)) R.classification[i] = c;
    if ((P.mode & RM_ACCUMULATE) && c == GXY_TERMINATED) {
      acc = true;
      ax = R.x[i];
      ay = R.y[i];
      ar = R.r[i];
      ag = R.g[i];
      ab = R.b[i];
      ao = R.o[i];
    }
    const int ci = c + 4;
    cat = (ci >= 0 && ci < table.n) ? table.t[ci] : -1;
  }
  if (P.mode & RM_ACCUMULATE) fb_accumulate_wave(P.fb, P.width, P.height, acc, ax, ay, ar, ag, ab, ao, P.unique != 0);
  const int w = threadIdx.x >> 6;
  for (int k = 0; k < K; k++) {
    unsigned long long m = __ballot(cat == k);
    if ((threadIdx.x & 63) == 0) wc[w][k] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < K) {
    int s = 0;
    for (int q = 0; q < GXY_BLOCK / 64; q++) s += wc[q][threadIdx.x];
    counts[threadIdx.x * nb + blockIdx.x] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) counts[K * nb] = 0;
}

// one workgroup per category: exclusive sum of its nb counts (relative to the
// category's start) and the category's total, for nb <= GXY_SMALL_SCAN
__global__ __launch_bounds__(1024) void k_part_scan_small(const int *counts, int nb, int *off, int *tot) {
  const int64_t base0 = (int64_t)blockIdx.x * nb;
  const int total = scan1024_runs(counts + base0, off + base0, nb);
  if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

// category totals from the exclusive sum over the category-major counts
__global__ void k_part_totals(const int *off, int nb, int K, int *tot) {
  const int k = threadIdx.x;
  if (k < K) tot[k] = off[(k + 1) * nb] - off[k * nb];
}

// destination lists are passed by value in the kernel arguments
#define GXY_MAX_SCATTER 12
struct ListTable {
  RayList_ispc l[GXY_MAX_SCATTER];
};

__device__ __forceinline__ void copy_one(const RayList_ispc &S, int i, const RayList_ispc &D, int j) {
  float *const *sf = &S.ox;
  float *const *df = &D.ox;
#pragma unroll
  for (int k = 0; k < 20; k++) df[k][j] = sf[k][i];
  int *const *si = &S.x;
  int *const *di = &D.x;
#pragma unroll
  for (int k = 0; k < 5; k++) di[k][j] = si[k][i];
}

__global__ __launch_bounds__(GXY_BLOCK) void k_part_scatter(RayList_ispc R, int n, CatTable table, int K,
                                                            int nb, const int *base, ListTable lists) {
  __shared__ int wc[GXY_BLOCK / 64][GXY_MAX_CAT];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  int cat = i < n ? category(R, i, table) : -1;
  int w = threadIdx.x >> 6;
  int rank = 0;
  for (int k = 0; k < K; k++) {
    unsigned long long m = __ballot(cat == k);
    if ((threadIdx.x & 63) == 0) wc[w][k] = __popcll(m);
    if (cat == k) rank = __popcll(m & lanemask_lt64());
  }
  __syncthreads();
  if (cat >= 0 && cat < K && cat < GXY_MAX_SCATTER) {
    int off = base[cat * nb + blockIdx.x] - base[cat * nb] + rank;
    for (int q = 0; q < w; q++) off += wc[q][cat];
    const RayList_ispc &D = lists.l[cat];
    if (D.ox) copy_one(R, i, D, off);
  }
}

static CatTable make_table(const int *cat_table, int ntable) {
  CatTable T;
  T.n = ntable;
  for (int k = 0; k < GXY_CAT_TABLE; k++) T.t[k] = k < ntable ? cat_table[k] : -1;
  return T;
}

// scratch layout: [K*nb+1 counts][their exclusive sum][scan temp], each array
// rounded to 64 ints
static inline size_t part_pad(int n, int K) {
  const size_t m = (size_t)K * ((n + GXY_BLOCK - 1) / GXY_BLOCK) + 1;
  return (m + 63) / 64 * 64;
}

size_t partition_scratch_bytes(int n, int K) {
  const int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  // temp storage of the device-wide scan only on the path that runs it
  return 2 * sizeof(int) * part_pad(n, K) + (nb > GXY_SMALL_SCAN ? scan_temp_bytes(K * nb + 1) : 0);
}

__global__ void k_zero_ints(int *p, int m) {
  if ((int)threadIdx.x < m) p[threadIdx.x] = 0;
}

hipError_t launch_route_count(const RouteParams &P, const RayList_ispc &R, int n, const int *cat_table, int ntable,
                              int K, int *scratch, hipStream_t s) {
  if (K > GXY_MAX_CAT || ntable > GXY_CAT_TABLE) return hipErrorInvalidValue;
  const int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_route_count, dim3(nb), dim3(GXY_BLOCK), 0, s, P, R, n, make_table(cat_table, ntable), K, nb,
                     scratch);
  return hipGetLastError();
}

hipError_t launch_partition(const RayList_ispc &R, int n, const int *cat_table, int ntable, int K,
                            const RayList_ispc *lists, int *scratch, int *counts, hipStream_t s, bool counted) {
  if (K > GXY_MAX_CAT || ntable > GXY_CAT_TABLE) return hipErrorInvalidValue;
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  if (nb == 0) {
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, counts, K);
    return hipGetLastError();
  }
  CatTable T = make_table(cat_table, ntable);
  const size_t pad = part_pad(n, K);
  int *off = scratch + pad;
  if (!counted) hipLaunchKernelGGL(k_part_count, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, T, K, nb, scratch);
  if (nb <= GXY_SMALL_SCAN) {
    // offsets relative to each category's start (the scatter subtracts off[cat * nb] = 0)
    hipLaunchKernelGGL(k_part_scan_small, dim3(K), dim3(1024), 0, s, scratch, nb, off, counts);
  } else {
    hipError_t e = launch_exclusive_sum(scratch, off, K * nb + 1, scratch + 2 * pad, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_part_totals, dim3(1), dim3(64), 0, s, off, nb, K, counts);
  }
  if (lists) return launch_partition_scatter(R, n, cat_table, ntable, K, lists, scratch, s);
  return hipGetLastError();
}

hipError_t launch_partition_scatter(const RayList_ispc &R, int n, const int *cat_table, int ntable, int K,
                                    const RayList_ispc *lists, const int *scratch, hipStream_t s) {
  if (K > GXY_MAX_SCATTER || ntable > GXY_CAT_TABLE) return hipErrorInvalidValue;
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  if (nb == 0) return hipSuccess;
  CatTable T = make_table(cat_table, ntable);
  ListTable L;
  memset(&L, 0, sizeof L);
  for (int k = 0; k < K; k++) L.l[k] = lists[k];
  hipLaunchKernelGGL(k_part_scatter, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, T, K, nb, scratch + part_pad(n, K), L);
  return hipGetLastError();
}

// ------------------------------------------------------------------ sampler particles
// Sampler::HandleTerminatedRays (Sampler.cpp:71-84): one particle per
// RAY_SURFACE ray at o + t*d, appended in ray order.
__global__ __launch_bounds__(GXY_BLOCK) void k_sampler_count(RayList_ispc R, int n, int *bc) {
  __shared__ int ws[GXY_BLOCK / 64];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = i < n && (R.term[i] & RAY_SURFACE);
  unsigned long long m = __ballot(f);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    bc[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
    if (blockIdx.x == 0) bc[gridDim.x] = 0;
  }
}

__global__ __launch_bounds__(GXY_BLOCK) void k_sampler_write(RayList_ispc R, int n, const int *bbase, const int *count,
                                                             float *xyz, int capacity) {
  __shared__ int ws[GXY_BLOCK / 64];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = i < n && (R.term[i] & RAY_SURFACE);
  unsigned long long m = __ballot(f);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) ws[w] = __popcll(m);
  __syncthreads();
  if (!f) return;
  int off = *count + bbase[blockIdx.x] + __popcll(m & lanemask_lt64());
  for (int q = 0; q < w; q++) off += ws[q];
  if (off < capacity) {
    float t = R.t[i];
    xyz[3 * (int64_t)off + 0] = R.ox[i] + t * R.dx[i];
    xyz[3 * (int64_t)off + 1] = R.oy[i] + t * R.dy[i];
    xyz[3 * (int64_t)off + 2] = R.oz[i] + t * R.dz[i];
  }
}

__global__ void k_add_count(int *count, const int *add) { *count += *add; }

// The values the host waits for after a list, written straight into mapped
// host memory: out[0..na) = a, out[16] = *b, out[18..19] = *c (64-bit), and
// out[32] = seq once the others are visible to the host.
__global__ void k_readback(const int *a, int na, const int *b, const unsigned long long *c, int *out, int seq) {
  const int t = threadIdx.x;
  if (t < na) out[t] = a[t];
  if (t == 16 && b) out[16] = *b;
  if (t == 18 && c) *(unsigned long long *)(out + 18) = *c;
  __threadfence_system();
  __syncthreads();
  if (t == 0) {
    *(volatile int *)(out + 32) = seq;
    __threadfence_system();
  }
}

hipError_t launch_readback(const int *a, int na, const int *b, const unsigned long long *c, int *out_dev, int seq,
                           hipStream_t s) {
  if (na < 0 || na > 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(64), 0, s, a, na, b, c, out_dev, seq);
  return hipGetLastError();
}
__global__ void k_copy_total(const int *src, int *dst) { *dst = *src; }

hipError_t launch_accumulate_sampler(const RayList_ispc &R, int n, float *xyz, int capacity, int *count,
                                     int *scratch, hipStream_t s) {
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  if (nb == 0) return hipSuccess;
  int *off = block_scan_offsets(scratch, nb);
  hipLaunchKernelGGL(k_sampler_count, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, scratch);
  hipError_t e = launch_block_scan(scratch, nb, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sampler_write, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, off, count, xyz, capacity);
  hipLaunchKernelGGL(k_add_count, dim3(1), dim3(1), 0, s, count, off + nb);
  return hipGetLastError();
}

// ------------------------------------------------------------------ pixel messages
// HandleTerminatedRays' remote path (Renderer.cpp:490-499): the pixels of the
// TERMINATED rays in ray order, after the SendPixelsMsg header.
__global__ __launch_bounds__(GXY_BLOCK) void k_pix_count(RayList_ispc R, int n, int *bc) {
  __shared__ int ws[GXY_BLOCK / 64];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = i < n && R.classification[i] == GXY_TERMINATED;
  unsigned long long m = __ballot(f);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    bc[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
    if (blockIdx.x == 0) bc[gridDim.x] = 0;
  }
}

__global__ __launch_bounds__(GXY_BLOCK) void k_pix_write(RayList_ispc R, int n, const int *bbase, gxy_pixel *out) {
  __shared__ int ws[GXY_BLOCK / 64];
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = i < n && R.classification[i] == GXY_TERMINATED;
  unsigned long long m = __ballot(f);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) ws[w] = __popcll(m);
  __syncthreads();
  if (!f) return;
  int off = bbase[blockIdx.x] + __popcll(m & lanemask_lt64());
  for (int q = 0; q < w; q++) off += ws[q];
  gxy_pixel p;
  p.x = R.x[i];
  p.y = R.y[i];
  p.r = R.r[i];
  p.g = R.g[i];
  p.b = R.b[i];
  p.o = R.o[i];
  out[off] = p;
}

__global__ void k_pix_header(gxy_pixels_header h, gxy_pixels_header *dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = h;
}

hipError_t launch_pixels_count(const RayList_ispc &R, int n, int *scratch, int *total, hipStream_t s) {
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  if (nb == 0) {
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, total, 1);
    return hipGetLastError();
  }
  int *off = block_scan_offsets(scratch, nb);
  hipLaunchKernelGGL(k_pix_count, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, scratch);
  hipError_t e = launch_block_scan(scratch, nb, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_copy_total, dim3(1), dim3(1), 0, s, off + nb, total);
  return hipGetLastError();
}

hipError_t launch_pixels_write(const RayList_ispc &R, int n, int *scratch, const gxy_pixels_header &h, void *msg,
                               hipStream_t s) {
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  gxy_pixel *px = (gxy_pixel *)((char *)msg + sizeof(gxy_pixels_header));
  if (nb > 0)
    hipLaunchKernelGGL(k_pix_write, dim3(nb), dim3(GXY_BLOCK), 0, s, R, n, block_scan_offsets(scratch, nb), px);
  hipLaunchKernelGGL(k_pix_header, dim3(1), dim3(64), 0, s, h, (gxy_pixels_header *)msg);
  return hipGetLastError();
}

// Rendering.cpp:104-121 (ACCUMULATE_PIXEL with frame stamps): the one
// work-item that raises a stale stamp clears the pixel; the additions follow
// in a second launch.
__global__ __launch_bounds__(GXY_BLOCK) void k_pix_stamp(const gxy_pixel *p, int count, int frame, float *fb,
                                                         int width, int height, int *kbuf) {
  int j = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (j >= count) return;
  const int x = p[j].x, y = p[j].y;
  if (x < 0 || y < 0 || x >= width || y >= height) return;
  const int64_t off = (int64_t)y * width + x;
  if (kbuf[off] < frame && atomicMax(&kbuf[off], frame) < frame) {
    fb[4 * off + 0] = 0.f;
    fb[4 * off + 1] = 0.f;
    fb[4 * off + 2] = 0.f;
    fb[4 * off + 3] = 0.f;
  }
}

__global__ __launch_bounds__(GXY_BLOCK) void k_pix_add(const gxy_pixel *p, int count, float *fb, int width,
                                                       int height, int *inside) {
  int j = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool in = false;
  if (j < count) {
    const gxy_pixel q = p[j];
    in = q.x >= 0 && q.y >= 0 && q.x < width && q.y < height;
    if (in) {
      float *d = fb + 4 * ((int64_t)q.y * width + q.x);
      atomicAdd(d + 0, q.r);
      atomicAdd(d + 1, q.g);
      atomicAdd(d + 2, q.b);
      atomicAdd(d + 3, q.o);
    }
  }
  unsigned long long m = __ballot(in);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(inside, __popcll(m));
}

hipError_t launch_pixels_apply(const void *msg, int count, int frame, float *fb, int width, int height, int *kbuf,
                               int *inside, hipStream_t s) {
  const gxy_pixel *px = (const gxy_pixel *)((const char *)msg + sizeof(gxy_pixels_header));
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, inside, 1);
  if (count <= 0) return hipGetLastError();
  const int nb = (count + GXY_BLOCK - 1) / GXY_BLOCK;
  if (kbuf) hipLaunchKernelGGL(k_pix_stamp, dim3(nb), dim3(GXY_BLOCK), 0, s, px, count, frame, fb, width, height, kbuf);
  hipLaunchKernelGGL(k_pix_add, dim3(nb), dim3(GXY_BLOCK), 0, s, px, count, fb, width, height, inside);
  return hipGetLastError();
}

// ------------------------------------------------------------------ camera
// Box::intersect (Box.cpp:107-150)
__device__ __forceinline__ bool box_intersect(const float *lo, const float *hi, f3 org, f3 dir, float &tmin,
                                              float &tmax) {
  tmin = (lo[0] - org.x) / dir.x;
  tmax = (hi[0] - org.x) / dir.x;
  if (tmin > tmax) { float t = tmax; tmax = tmin; tmin = t; }
  if (tmax < 0) return false;
  float tymin = (lo[1] - org.y) / dir.y, tymax = (hi[1] - org.y) / dir.y;
  if (tymin > tymax) { float t = tymax; tymax = tymin; tymin = t; }
  if (tymax < 0) return false;
  if ((tmin > tymax) || (tymin > tmax)) return false;
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  float tzmin = (lo[2] - org.z) / dir.z, tzmax = (hi[2] - org.z) / dir.z;
  if (tzmin > tzmax) { float t = tzmin; tzmin = tzmax; tzmax = t; }
  if (tzmax < 0) return false;
  if ((tmin > tzmax) || (tzmin > tmax)) return false;
  if (tzmin > tmin) tmin = tzmin;
  if (tzmax < tmax) tmax = tzmax;
  if (tmin < 0) tmin = 0;
  return true;
}

// pixel of the q-th ray slot: row major, or 8x8 tiles (tile-major, row major
// inside a tile) for locality of neighbouring rays in a wavefront
__device__ __forceinline__ void pixel_of(const CameraParams &C, int q, int &x, int &y) {
  if (C.order == 0) {
    x = q % C.width;
    y = q / C.width;
  } else {
    // 8x8 tiles in bands of two tile rows, a band's tiles column by column,
    // so four consecutive tiles (one 256-ray workgroup) make a 16x16 square;
    // a last band of one tile row runs along its row
    const int tw = (C.width + 7) / 8, th = (C.height + 7) / 8;
    const int tile = q / 64, r = q % 64, band = tile / (2 * tw), k = tile % (2 * tw);
    int tx, ty;
    if (band * 2 + 1 < th) { tx = k / 2; ty = band * 2 + (k & 1); }
    else { tx = k; ty = band * 2; }
    x = tx * 8 + (r % 8);
    y = ty * 8 + (r / 8);
  }
}

// dtypes.h normalize: d = len(a); if (d != 0) { d = 1.0/d (double); a *= d }
__device__ __forceinline__ f3 dt_normalize(f3 a) {
  float d = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  if (d != 0) {
    d = (float)(1.0 / (double)d);
    a.x *= d; a.y *= d; a.z *= d;
  }
  return a;
}

__device__ __forceinline__ bool camera_ray(const CameraParams &C, int x, int y, f3 &vorigin, f3 &vray) {
  float fx = ((float)x - C.off_x) * C.scaling;
  float fy = ((float)y - C.off_y) * C.scaling;
  f3 xy = mk3(C.center[0] + fx * C.vr[0] + fy * C.vu[0], C.center[1] + fx * C.vr[1] + fy * C.vu[1],
              C.center[2] + fx * C.vr[2] + fy * C.vu[2]);
  f3 vdir = mk3(C.vdir[0], C.vdir[1], C.vdir[2]);
  if (C.is_ortho) {
    vorigin = sub3(xy, vdir);
    vray = vdir;
  } else {
    vorigin = mk3(C.veye[0], C.veye[1], C.veye[2]);
    vray = dt_normalize(sub3(xy, vorigin));
  }
  float gmin = 0, gmax = 0, lmin = 0, lmax = 0;
  bool hit = box_intersect(C.glo, C.ghi, vorigin, vray, gmin, gmax);
  if (hit) hit = box_intersect(C.llo, C.lhi, vorigin, vray, lmin, lmax);
  float d = fabsf(lmin) - fabsf(gmin);
  return hit && (lmax >= 0) && ((double)d < 0.000001) && ((double)d > -0.000001);
}

__global__ __launch_bounds__(GXY_BLOCK) void k_camera_count(CameraParams C, int npix, int *bc) {
  __shared__ int ws[GXY_BLOCK / 64];
  int q = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = false;
  if (q < npix) {
    int x, y;
    pixel_of(C, q, x, y);
    f3 o, d;
    f = x < C.width && y < C.height && camera_ray(C, x, y, o, d);
  }
  unsigned long long m = __ballot(f);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    bc[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
    if (blockIdx.x == 0) bc[gridDim.x] = 0;
  }
}

__global__ __launch_bounds__(GXY_BLOCK) void k_camera_write(CameraParams C, int npix, const int *bbase,
                                                            RayList_ispc R, int capacity) {
  __shared__ int ws[GXY_BLOCK / 64];
  int q = blockIdx.x * GXY_BLOCK + threadIdx.x;
  bool f = false;
  int x = 0, y = 0;
  f3 o = mk3(0, 0, 0), d = o;
  if (q < npix) {
    pixel_of(C, q, x, y);
    f = x < C.width && y < C.height && camera_ray(C, x, y, o, d);
  }
  unsigned long long m = __ballot(f);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) ws[w] = __popcll(m);
  __syncthreads();
  if (!f) return;
  int off = bbase[blockIdx.x] + __popcll(m & lanemask_lt64());
  for (int k = 0; k < w; k++) off += ws[k];
  if (off >= capacity) return;
  R.x[off] = x; R.y[off] = y;
  R.ox[off] = o.x; R.oy[off] = o.y; R.oz[off] = o.z;
  R.dx[off] = d.x; R.dy[off] = d.y; R.dz[off] = d.z;
  R.r[off] = 0.f; R.g[off] = 0.f; R.b[off] = 0.f; R.o[off] = 0.f;
  R.t[off] = 0.f; R.tMax[off] = 3.402823466e+38f;
  R.type[off] = RAY_PRIMARY; R.term[off] = 0; R.classification[off] = 0;
}

hipError_t launch_camera(const CameraParams &C, const RayList_ispc &R, int capacity, int *scratch, int *total,
                         hipStream_t s) {
  int npix;
  if (C.order == 0) npix = C.width * C.height;
  else npix = ((C.width + 7) / 8) * ((C.height + 7) / 8) * 64;
  int nb = (npix + GXY_BLOCK - 1) / GXY_BLOCK;
  if (nb <= 0) nb = 0;
  int *off = block_scan_offsets(scratch, nb);
  if (nb == 0) {
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, off, 1);
    if (total) hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, total, 1);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_camera_count, dim3(nb), dim3(GXY_BLOCK), 0, s, C, npix, scratch);
  hipError_t e = launch_block_scan(scratch, nb, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_camera_write, dim3(nb), dim3(GXY_BLOCK), 0, s, C, npix, off, R, capacity);
  if (total) hipLaunchKernelGGL(k_copy_total, dim3(1), dim3(1), 0, s, off + nb, total);
  return hipGetLastError();
}

// ------------------------------------------------------------------ copies
__global__ __launch_bounds__(GXY_BLOCK) void k_copy_rays(RayList_ispc S, int start, int count, RayList_ispc D,
                                                         int dst_start) {
  int i = blockIdx.x * GXY_BLOCK + threadIdx.x;
  if (i >= count) return;
  copy_one(S, start + i, D, dst_start + i);
}

hipError_t launch_copy_rays(const RayList_ispc &src, int start, int count, const RayList_ispc &dst, int dst_start,
                            hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_rays, dim3((count + GXY_BLOCK - 1) / GXY_BLOCK), dim3(GXY_BLOCK), 0, s, src, start, count,
                     dst, dst_start);
  return hipGetLastError();
}

// ------------------------------------------------------------------ RayList header
// the 64-byte block header written by a kernel (stream ordered, no host wait)
struct HdrWords {
  unsigned w[16];
};

__global__ void k_write_header(unsigned *dst, HdrWords h) {
  if (threadIdx.x < 16) dst[threadIdx.x] = h.w[threadIdx.x];
}

hipError_t launch_write_header(void *block, const void *hdr64, hipStream_t s) {
  HdrWords h;
  memcpy(&h, hdr64, sizeof h);
  hipLaunchKernelGGL(k_write_header, dim3(1), dim3(64), 0, s, (unsigned *)block, h);
  return hipGetLastError();
}

}  // namespace gxy
