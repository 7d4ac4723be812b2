// gxy_internal.h — launchers shared between the kernel files and the runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "gxy_device.h"

namespace gxy {

// lighting / spawn modes (bits), applied per parent in this order
enum { LM_AMBIENT = 1, LM_AO = 2, LM_DIFFUSE = 4, LM_SHADOW = 8 };

struct LightParams {
  int nAO, nLights, doShadows;
  int reverse;         // Galaxy's reverse lighting model (1, its default) or the direct one (0)
  float aoRadius, Ka, Kd, eps;
  float L[3 * 16];
  int types[16];
  const float *U, *V;  // AO direction tables (256 entries each, device memory)
  int ao_pcg;          // 1: AO directions from PCG32 streams per (pixel, AO index, seed) instead of the tables
  unsigned ao_seed;
};

// route modes (bits)
enum { RM_CLASSIFY = 1, RM_ASSIGN = 2, RM_ACCUMULATE = 4 };

struct RouteParams {
  int reverse;
  int neighbors[6];
  float blo[3], bhi[3];
  float *fb;
  int width, height;
  int mode;
  int unique;   // 1: no two rays of the list carry the same pixel (plain add, no atomics)
};

struct CameraParams {
  float veye[3], vdir[3], vr[3], vu[3], center[3];
  float scaling, off_x, off_y;
  int is_ortho;
  int width, height, order;
  float llo[3], lhi[3], glo[3], ghi[3];
};

// wpe: occupancy variant of the volume-only kernel (0 none, 5 or 6 waves/SIMD);
// stack: BVH traversal stack of geometry-only launches (0: 64 entries in
// private memory; 16, 24, 32: entries in LDS, used when the tree fits)
// gwaves: 8 asks for 8 waves per SIMD (64 VGPRs) in the binary private-stack geometry walks (with or without
// path lines), 0 the compiler's choice
// bs: work-items per workgroup of the one-bricked-volume, one-isovalue march
// at 6 waves per SIMD (64, 128 or 256; GXY_BLOCK elsewhere):
// trace_block_size(P, wpe, bs) is the size a launch uses
hipError_t launch_trace(const TraceParams &P, const RayList_ispc &R, int n, int wpe, int stack, hipStream_t s,
                        int gwaves = 0, int bs = GXY_BLOCK);
int trace_block_size(const TraceParams &P, int wpe, int bs);
hipError_t launch_sampler(const TraceParams &P, const RayList_ispc &R, int n, hipStream_t s);
// Persistent-lane trace for volume-only visualizations (1..4 volumes): each
// wave owns `chunk` consecutive rays; idle lanes are refilled when at least
// `refill` of the 64 are idle.
hipError_t launch_trace_pl(const TraceParams &P, const RayList_ispc &R, int n, int chunk, int refill,
                           hipStream_t s);
int trace_pl_blocks(int n, int chunk);

// Exclusive sum of m ints by the device-wide scan; temp holds scan_temp_bytes(m).
size_t scan_temp_bytes(long long m);
hipError_t launch_exclusive_sum(const int *in, int *out, long long m, void *temp, hipStream_t s);
// Offsets from per-workgroup counts: scratch holds block_scan_scratch_bytes(nb)
// bytes, the count kernel writes nb counts and a zero after them at scratch,
// and launch_block_scan writes their exclusive sum (nb + 1 entries, the total
// last) at block_scan_offsets(scratch, nb).
// (one workgroup for count arrays of at most GXY_SMALL_SCAN entries, the
// device-wide scan above that)
#define GXY_SMALL_SCAN 16384
size_t block_scan_scratch_bytes(int nb);
int *block_scan_offsets(int *scratch, int nb);
hipError_t launch_block_scan(int *scratch, int nb, hipStream_t s);
// Exclusive offsets of primary surface rays (flag scan).  scratch holds
// surface_scan_scratch_bytes(n) bytes; total (device int, may be null)
// receives the count, which also stays at surface_scan_total(scratch, n).
size_t surface_scan_scratch_bytes(int n);
int *surface_scan_total(int *scratch, int n);
hipError_t launch_surface_scan(const RayList_ispc &R, int n, int *scratch, int *rank_out, int *total,
                               hipStream_t s, int *pidx = nullptr);
// Lighting with one work-item per spawned child (children at the offsets of
// TraceRays.cpp:89-117 for the T parents whose ray indices pidx lists in rank
// order), then the parents' ambient + diffuse update.
hipError_t launch_lighting_children(const LightParams &L, int mode, const RayList_ispc &R, const int *pidx, int T,
                                    const RayList_ispc &O, hipStream_t s);
// The same with the parent count T on the device (at most n parents): the
// grids cover the upper bound and work-items beyond T exit.
hipError_t launch_lighting_children_dev(const LightParams &L, int mode, const RayList_ispc &R, const int *pidx,
                                        const int *T, int n, const RayList_ispc &O, hipStream_t s);
// Offsets as TraceRays.cpp:89-117 from the ranks: ao = nAO*rank, shadow =
// nAO*total + nLights*rank (or -1).
hipError_t launch_offsets_from_rank(const RayList_ispc &R, int n, const int *rank, const int *total, int nAO,
                                    int nLights, int doShadows, int *ao_off, int *sh_off, hipStream_t s);
// The same offsets with every temporary in the static scratch array of
// spawn.hip (no allocation).  noalloc_scratch_address gives that array's
// address on the current device (0, or -1 with *ptr null); lists of at most
// GXY_NOALLOC_MAX_RAYS rays; *total_out receives the device address of the
// parent total.
static_assert(GXY_NOALLOC_MAX_RAYS == (GXY_SMALL_SCAN - 1) * GXY_BLOCK, "static scratch limit");
int noalloc_scratch_address(int **ptr);
hipError_t launch_spawn_offsets_noalloc(const RayList_ispc &R, int n, int *scr, int nAO, int nL, int doShadows,
                                        int *ao_off, int *sh_off, const int **total_out, hipStream_t s);
hipError_t launch_lighting(const LightParams &L, int mode, const RayList_ispc &R, int n, const int *ao_off,
                           const int *sh_off, const RayList_ispc &O, hipStream_t s);

hipError_t launch_route(const RouteParams &P, const RayList_ispc &R, int n, hipStream_t s);
// Stable partition by classification into K categories: the category of a ray
// is cat_table[class + 4] (host table, ntable <= 64 entries, -1 = not copied).
// counts: K ints (device).  scratch: partition_scratch_bytes(n, K) bytes, kept
// intact between the count phase (lists == nullptr) and launch_partition_scatter,
// which takes the same n and K (lists: K entries, those with a null ox not copied).
// counted: the per-workgroup counts are already in scratch (launch_route_count:
// the route step and the count phase in one pass over the rays).
size_t partition_scratch_bytes(int n, int K);
hipError_t launch_partition(const RayList_ispc &R, int n, const int *cat_table, int ntable, int K,
                            const RayList_ispc *lists, int *scratch, int *counts, hipStream_t s,
                            bool counted = false);
hipError_t launch_route_count(const RouteParams &P, const RayList_ispc &R, int n, const int *cat_table, int ntable,
                              int K, int *scratch, hipStream_t s);
hipError_t launch_partition_scatter(const RayList_ispc &R, int n, const int *cat_table, int ntable, int K,
                                    const RayList_ispc *lists, const int *scratch, hipStream_t s);
// scratch: block_scan_scratch_bytes(ceil(n/256)) bytes (camera: of the pixel
// order's ray slots; total may be null: the count stays at
// block_scan_offsets(scratch, nb)[nb])
hipError_t launch_accumulate_sampler(const RayList_ispc &R, int n, float *xyz, int capacity, int *count,
                                     int *scratch, hipStream_t s);
hipError_t launch_camera(const CameraParams &C, const RayList_ispc &R, int capacity, int *scratch, int *total,
                         hipStream_t s);
// one kernel copying up to 16 ints (a), one int (b, optional) and one 64-bit
// counter (c, optional) into the mapped host area out_dev (a at 0, b at 16,
// c at 18), then seq at 32 once those are visible
hipError_t launch_readback(const int *a, int na, const int *b, const unsigned long long *c, int *out_dev, int seq,
                           hipStream_t s);
// SendPixelsMsg contents: count of TERMINATED rays (device int), then the
// header and pixels into msg (scratch: block_scan_scratch_bytes(ceil(n/256))
// bytes, kept between the two calls); applying a message of count pixels
// (inside: device int, pixels within the framebuffer)
hipError_t launch_pixels_count(const RayList_ispc &R, int n, int *scratch, int *total, hipStream_t s);
hipError_t launch_pixels_write(const RayList_ispc &R, int n, int *scratch, const gxy_pixels_header &h, void *msg,
                               hipStream_t s);
hipError_t launch_pixels_apply(const void *msg, int count, int frame, float *fb, int width, int height, int *kbuf,
                               int *inside, hipStream_t s);
// writes the 64-byte RayList header hdr64 (host memory, copied into the launch) to block
hipError_t launch_write_header(void *block, const void *hdr64, hipStream_t s);
hipError_t launch_copy_rays(const RayList_ispc &src, int start, int count, const RayList_ispc &dst, int dst_start,
                            hipStream_t s);

// bricked copy of a linear volume (B = 4 or 8; out holds nb0*nb1*nb2*brick_floats(B))
hipError_t launch_brickify(const float *data, const int n[3], int B, const int nb[3], float *out, hipStream_t s);
hipError_t launch_noise(float *data, int nx, int ny, int nz, const int *perm_dev, int octaves, float base_freq,
                        float *minmax_scratch, hipStream_t s);

// ---------------------------------------------------------------- batched passes
enum { BM_RENDER = 0, BM_SAMPLE = 1 };
// statistics slots
enum { BS_RAYS = 0, BS_SAMPLES, BS_SPAWNED, BS_TERMINATED, BS_DROPPED, BS_KEPT, BS_COUNT };

struct BatchList {
  RayList_ispc R;   // the list's own (primary) rays, device
  float *fb;        // BM_RENDER: RGBA accumulation, width * height * 4
  float *xyz;       // BM_SAMPLE: particles, 3 floats each
  int *pcount;      // BM_SAMPLE: particle count (device)
  int width, height;
  int pcap;         // particle capacity
  int ds;           // dataset index
  int mode;         // BM_RENDER / BM_SAMPLE
  int n;
};

// rays [start, start + count) of one list, read from ray list src of the
// RayList table (0, 1: the two ray pools; 2 + l: list l's own rays)
struct BatchChunk {
  int src, start, count, list;
};

struct BatchArgs {
  const BatchList *lists;
  int nlists;
  int chunk;                  // rays per chunk (CH)
  const TraceParams *ds;      // per dataset
  const RayList_ispc *rls;    // RayList table: pools 0 and 1, then every list's own rays
  int pool_cap;
  int out_pool;               // pool (0 or 1) written by this pass
  const BatchChunk *tin;
  const int *nin;
  BatchChunk *tout;
  int *nout;
  int tcap;
  int *nrays_out;
  int *head;                  // work-queue head of the trace kernel
  int *cpar, *cout, *cchk;    // per input chunk: parents, rays out, chunks out
  int *coff, *cchkoff;        // their exclusive offsets
  long long *segr, *segc;     // scan segments (256 each)
  int *lpend;                 // per list: rays in the next pass
  unsigned long long *part;   // per-workgroup partials, BS_COUNT * nblocks
  int nblocks;
  unsigned long long *stats;  // BS_COUNT totals
  int *overflow;              // bit 0: pool or table full; bit 1: append overflow
  int *kidx;                  // per kind (volume render, geometry, sampler): input chunk indices, kcap each
  int *kcount;                // 3 counts
  int *khead;                 // 3 work-queue heads
  int kcap;
  int max_segments;           // sampler rays: isovalue crossings followed per pass (the rest re-queued)
  LightParams L;
};

struct BatchLaunch {
  int trace_blocks;  // persistent trace grid (workgroups)
  int post_blocks;   // count / write grid
  int stack;         // traversal stack: 0 private (64 entries), or 16/24/32 entries in LDS
  int wpe;           // occupancy request (waves per SIMD), 0 none
  int split;         // 0: one fused trace kernel; 1: one kernel per ray kind on three streams
};

// Partial sums: A.part holds BS_COUNT x (4 * nblocks) slots; nblocks equals
// trace_blocks and post_blocks.  ks / ev: three streams and four events for the
// split trace (may be null when split is 0).
hipError_t launch_batch_pass(const BatchArgs &A, const BatchLaunch &K, hipStream_t s, const hipStream_t *ks,
                             const hipEvent_t *ev);
hipError_t launch_batch_reduce(const BatchArgs &A, hipStream_t s);
hipError_t launch_batch_append(BatchChunk *tab, int *ntab, int cap, const BatchChunk *add, int n, int *overflow,
                               hipStream_t s);
hipError_t launch_set_count(int *p, int v, hipStream_t s);
hipError_t launch_fill_f32(float *p, long long n, float v, hipStream_t s);
hipError_t launch_random_rays(const RayList_ispc &R, int n, unsigned seed, int width, int type, const float *lo,
                              const float *hi, hipStream_t s);

// LBVH build (bvh.hip).  Boxes per primitive (lo, hi) on device; returns nodes
// (device, allocated through alloc_fn) and the sorted primitive order.
typedef void *(*alloc_fn_t)(size_t);
typedef void (*free_fn_t)(void *);
// Leaves of up to GXY_BVH_LEAF primitives; trees of height below GXY_BVH_MAXH.
#define GXY_BVH_LEAF 8
#define GXY_BVH_MAXH 256
struct BVHBuildResult {
  BVHNode *nodes;     // binary tree, root 0
  int *order;         // primitive of each leaf slot (leaf ranges index it)
  int nnodes;
  int depth;          // levels of the binary tree, leaves included
  float sah;          // SAH cost of the binary tree over its root's surface area
  BVH4Node *nodes4;   // the same tree collapsed to 4-wide nodes, root 0
  int nnodes4;
  int depth4;         // levels of 4-wide inner nodes
  int leaves;
  int heights;        // height of the tree before the last treelet round
};
// prim_cost: SAH cost of one primitive test against one node visit's 1.2
hipError_t build_lbvh(const float *boxes_dev, int n, float prim_cost, alloc_fn_t alloc, free_fn_t dfree,
                      BVHBuildResult *out, hipStream_t s);

}  // namespace gxy
