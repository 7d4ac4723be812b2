// runtime.cpp — host side of libgxy_hip.so: device selection, stream, caching
// device allocator, scene objects resident in HBM, device RayLists with the
// reference block layout, the C-ABI entry points of include/gxy_hip.h, and the
// ray-list loop of Renderer/processRays_task run against device lists.
//
// Host semantics cited per entry point; kernels live in trace.hip, spawn.hip,
// route.hip, bvh.hip and noise.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "gxy_internal.h"

#define GXY_EXPORT extern "C" __attribute__((visibility("default")))

namespace gxy {

// ------------------------------------------------------------------ errors
static thread_local std::string t_err;
static int fail(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  if (getenv("GXY_VERBOSE")) fprintf(stderr, "gxy: %s\n", buf);
  return -1;
}
#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail("%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

// ------------------------------------------------------------------ runtime state
struct Timing {
  double ms = 0;
  int launches = 0;
};
struct PendingEv {
  std::string name;
  hipEvent_t a, b;
};

// An execution context: one HIP stream (default priority), its device
// counters ([0] samples), and a cache of freed device blocks that are reused
// only on this stream (stream order makes the reuse safe without a sync).
struct Ctx {
  hipStream_t stream = nullptr;
  unsigned long long *counters = nullptr;
  int *hread = nullptr, *dread = nullptr;   // mapped host readback area (host / device address)
  int rb_seq = 0;                            // sequence number of the last readback
  std::multimap<size_t, void *> free_blocks;
  size_t bytes_cached = 0;
};

struct Runtime {
  std::recursive_mutex mu;
  int device = -1;
  Ctx def;                      // the library's default context
  hipStream_t user_stream = nullptr;
  std::map<void *, size_t> live;  // live block -> size class
  size_t bytes_live = 0;
  // timing
  int timing = 0;                 // 0 off, 1 every kernel, 2 trace kernels only
  std::vector<PendingEv> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, Timing> timings;
  // registries for handle validation
  std::set<void *> volumes, tfs, models, raylists, tracers, lightings, contexts;
  std::atomic<int> raylist_id{0};
  float opaque_alpha = 0.999f;
  int reverse_lighting = 1;       // CMakeLists.txt:66 GXY_REVERSE_LIGHTING (default ON)
  // SAH cost of one triangle, sphere and round-cone test (a node visit costs
  // 1.2): the trace times of the mesh and path-line frames fall as the cost
  // rises to these values and are flat beyond them
  float bvh_cost[3] = {2.0f, 1.0f, 8.0f};
  std::vector<void *> outbound[6];   // RayListObj* per exit face, last frames
  std::vector<std::pair<int *, int *>> free_reads;   // readback areas of destroyed contexts (kept mapped)
  // volume-only trace kernel: 0 one ray per work-item, 1 persistent lanes
  int trace_kind = 0, trace_chunk = 256, trace_refill = 32, trace_wpe = 0;
  int ao_pcg = 0;          // AO directions: 0 the UV table (reference), 1 PCG32 streams
  unsigned ao_seed = 0;
  int geom_stack = 0;           // BVH traversal stack of geometry-only traces (0 private, or LDS entries)
  int occlusion = 1;            // rank loops: shadow/AO rays of volume-free scenes stop at their first hit
  int bvh_width = 2;            // traverse the binary trees (2) or the 4-wide ones (4)
  int geom_waves = 0;           // binary private-stack walks: 0 the compiler's occupancy, 8 waves per SIMD
  float *uv = nullptr;          // AO direction tables U[256], V[256] (device)
  // device address of the static scratch array of the allocation-free spawn
  // offsets, per device (taken once, with that device current, at set-up)
  int *noalloc_scratch[64] = {};
};
static Runtime G;
static thread_local Ctx *t_ctx = nullptr;

static Ctx &CX() { return t_ctx ? *t_ctx : G.def; }
static hipStream_t S() {
  if (t_ctx) return t_ctx->stream;
  return G.user_stream ? G.user_stream : G.def.stream;
}

static size_t size_class(size_t b) {
  if (b < 256) return 256;
  if (b <= (1u << 20)) {
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
  }
  // eighth-octave classes above 1 MiB, rounded to 64 KiB
  size_t c = 1u << 20;
  while (c * 2 <= b) c <<= 1;
  size_t step = c / 8;
  size_t r = ((b + step - 1) / step) * step;
  return (r + 65535) & ~(size_t)65535;
}

static void release_cache(Ctx &c) {
  for (auto &kv : c.free_blocks) hipFree(kv.second);
  c.free_blocks.clear();
  c.bytes_cached = 0;
}

void *dalloc(size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  Ctx &cx = CX();
  size_t c = size_class(bytes ? bytes : 1);
  auto it = cx.free_blocks.find(c);
  void *p = nullptr;
  if (it != cx.free_blocks.end()) {
    p = it->second;
    cx.free_blocks.erase(it);
    cx.bytes_cached -= c;
  } else {
    if (hipMalloc(&p, c) != hipSuccess) {
      (void)hipGetLastError();
      hipStreamSynchronize(S());
      release_cache(cx);
      if (hipMalloc(&p, c) != hipSuccess) {
        (void)hipGetLastError();
        fail("device allocation of %zu bytes failed (%zu live)", c, G.bytes_live);
        return nullptr;
      }
    }
  }
  G.live[p] = c;
  G.bytes_live += c;
  return p;
}

// Frees into the current context's cache; the block is reused on this stream
// only, after the work queued before the free (stream order).
static void dfree_stream_ordered(void *p) {
  if (!p) return;
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  auto it = G.live.find(p);
  if (it == G.live.end()) return;
  size_t c = it->second;
  G.live.erase(it);
  G.bytes_live -= c;
  Ctx &cx = CX();
  cx.free_blocks.insert({c, p});
  cx.bytes_cached += c;
}

// Frees a block that other streams may have used: waits for the device first.
void dfree(void *p) {
  if (!p) return;
  if (G.contexts.empty()) hipStreamSynchronize(S());
  else hipDeviceSynchronize();   // another context's stream may still read it
  dfree_stream_ordered(p);
}

// ------------------------------------------------------------------ readback
// A context's readback area: 64 ints of pinned, mapped, coherent host memory
// that kernels write directly, so a list's counts reach the host with one
// kernel and one stream wait.  Areas outlive their contexts (reused, never
// unmapped while the process runs).
static int *readback_area(Ctx &c) {
  if (c.hread) return c.dread;
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  if (!G.free_reads.empty()) {
    c.hread = G.free_reads.back().first;
    c.dread = G.free_reads.back().second;
    G.free_reads.pop_back();
    memset(c.hread, 0, 64 * sizeof(int));   // no sequence number of its last owner
    c.rb_seq = 0;
    return c.dread;
  }
  void *h = nullptr, *d = nullptr;
  if (hipHostMalloc(&h, 64 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    (void)hipGetLastError();
    fail("pinned readback allocation failed");
    return nullptr;
  }
  if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
    (void)hipGetLastError();
    hipHostFree(h);
    fail("no device address for the readback area");
    return nullptr;
  }
  c.hread = (int *)h;
  c.dread = (int *)d;
  memset(c.hread, 0, 64 * sizeof(int));
  c.rb_seq = 0;
  return c.dread;
}

struct Readback {
  int ints[16];
  int one;
  unsigned long long u64;
};

// Returns up to 16 device ints, one device int and one device 64-bit counter
// (each optional) once the stream has reached them.  The host polls the
// sequence number the readback kernel writes last (the kernel runs after all
// work queued before it); every 256 polls it asks the stream, so an error or
// a kernel that never wrote ends the wait through hipStreamSynchronize.
// readback_start / readback_wait split it, so that work queued between them
// runs while the host waits (one readback in flight per context).
static int readback_start(const int *ints, int nints, const int *one, const unsigned long long *u64, int &seq) {
  Ctx &cx = CX();
  int *dev = readback_area(cx);
  if (!dev) return -1;
  seq = ++cx.rb_seq;
  HIPCHK(launch_readback(ints, nints, one, u64, dev, seq, S()));
  return 0;
}
static int readback_wait(int seq, int nints, bool one, bool u64, Readback &out) {
  Ctx &cx = CX();
  const volatile int *h = cx.hread;
  for (unsigned spins = 1; h[32] != seq; spins++) {
    if ((spins & 255) == 0) {
      const hipError_t q = hipStreamQuery(S());
      if (q == hipErrorNotReady) continue;
      (void)hipGetLastError();
      HIPCHK(hipStreamSynchronize(S()));
      if (h[32] != seq) return fail("readback %d did not arrive", seq);
      break;
    }
    __builtin_ia32_pause();
  }
  for (int k = 0; k < nints; k++) out.ints[k] = h[k];
  out.one = one ? h[16] : 0;
  out.u64 = u64 ? ((unsigned long long)(unsigned)h[19] << 32) | (unsigned)h[18] : 0;
  return 0;
}
static int readback(const int *ints, int nints, const int *one, const unsigned long long *u64, Readback &out) {
  int seq = 0;
  if (readback_start(ints, nints, one, u64, seq) != 0) return -1;
  return readback_wait(seq, nints, one != nullptr, u64 != nullptr, out);
}

// ------------------------------------------------------------------ timing
static hipEvent_t timing_event() {
  if (!G.free_events.empty()) {
    hipEvent_t e = G.free_events.back();
    G.free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct KTimer {
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  explicit KTimer(const char *n) : name(n) {
    if (G.timing == 1 || (G.timing == 2 && strcmp(n, "trace") == 0)) {
      std::lock_guard<std::recursive_mutex> lk(G.mu);
      a = timing_event();
      b = timing_event();
      if (!a || !b) {
        if (a) G.free_events.push_back(a);
        if (b) G.free_events.push_back(b);
        a = b = nullptr;
        return;
      }
      hipEventRecord(a, S());
    }
  }
  ~KTimer() {
    if (a) {
      hipEventRecord(b, S());
      std::lock_guard<std::recursive_mutex> lk(G.mu);
      G.pending.push_back({name, a, b});
    }
  }
};

static void resolve_timings() {
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  for (auto &p : G.pending) {
    hipEventSynchronize(p.b);
    float ms = 0;
    hipEventElapsedTime(&ms, p.a, p.b);
    G.timings[p.name].ms += ms;
    G.timings[p.name].launches++;
    G.free_events.push_back(p.a);
    G.free_events.push_back(p.b);
  }
  G.pending.clear();
}

// ------------------------------------------------------------------ UV tables
// UV.ih:21-58: base-2 / base-3 radical inverses accumulated in float with the
// digit weight multiplied by 1/b each step, rounded to 6 significant digits.
static void make_uv(float *U, float *V) {
  for (int b = 2; b <= 3; b++)
    for (int i = 0; i < 256; i++) {
      float inv = 1.0f / (float)b, f = inv, r = 0.f;
      for (int j = i; j > 0; j /= b) {
        r = r + (float)(j % b) * f;
        f = f * inv;
      }
      char buf[64];
      snprintf(buf, sizeof buf, "%.6g", (double)r);
      (b == 2 ? U : V)[i] = strtof(buf, nullptr);
    }
}

static int ensure_init() {
  if (G.device >= 0) return 0;
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  if (G.device >= 0) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return fail("no HIP device available (libgxy_hip requires an AMD GPU)");
  }
  int dev = 0;
  if (getenv("GXY_DEVICE")) dev = atoi(getenv("GXY_DEVICE"));
  HIPCHK(hipSetDevice(dev));
  HIPCHK(hipStreamCreateWithFlags(&G.def.stream, hipStreamNonBlocking));
  if (dev >= 0 && dev < 64 && !G.noalloc_scratch[dev] && noalloc_scratch_address(&G.noalloc_scratch[dev]) != 0) {
    (void)hipGetLastError();   // gxy_spawn_offsets_noalloc reports it; nothing else needs the array
    G.noalloc_scratch[dev] = nullptr;
  }
  G.device = dev;
  // the default context's readback area now, so that no later call allocates it
  if (!readback_area(G.def)) return -1;
  float UV[512];
  make_uv(UV, UV + 256);
  G.uv = (float *)dalloc(sizeof UV);
  if (!G.uv) return -1;
  HIPCHK(hipMemcpy(G.uv, UV, sizeof UV, hipMemcpyHostToDevice));
  G.def.counters = (unsigned long long *)dalloc(8 * sizeof(unsigned long long));
  if (!G.def.counters) return -1;
  HIPCHK(hipMemsetAsync(G.def.counters, 0, 8 * sizeof(unsigned long long), S()));
  return 0;
}
#define INIT_OR(ret) \
  do {               \
    if (ensure_init() != 0) return ret; \
  } while (0)

// ------------------------------------------------------------------ scene objects
struct VolumeObj {
  DVolume d;
  float *data = nullptr;
  float *bdata = nullptr;   // bricked copy (d.brick > 0)
  size_t count = 0;
};

// (re)builds the bricked copy from the linear voxels
static int volume_brickify(VolumeObj *v) {
  if (v->d.brick == 0) return 0;
  const int B = v->d.brick;
  HIPCHK(launch_brickify(v->data, v->d.n, B, v->d.nb, v->bdata, S()));
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}

struct TFObj {
  DTF d;
  float *col = nullptr, *op = nullptr;
  std::vector<float> hcol, hop;
};

struct MeshObj {
  int geomID;
  int nv, nt;
  float *V = nullptr, *N = nullptr, *D = nullptr;
  uint32_t *I = nullptr;
  TFObj *tf = nullptr;
};

struct SpheresObj {
  int geomID;
  int n;
  float *C = nullptr, *D = nullptr;
  float r0, r1, v0, v1, eps;
  TFObj *tf = nullptr;
};

// Path lines as round Bezier curves (DataDrivenPathLines.cpp:46-160): the
// control points (x y z radius) of every curve segment and, on the host, the
// round-cone pieces each segment is traced as (curve_pieces of them).
struct PathObj {
  int geomID;
  int nseg = 0;
  float *CP = nullptr;         // device: control points, 4 floats each
  uint32_t *IC = nullptr;      // device: first control point of each segment
  std::vector<float> pa, pb;   // pieces: sphere a and sphere b, 4 floats each
  std::vector<float> pu;       // pieces: (u0, u1)
  std::vector<int> pseg;       // pieces: curve segment
  float r0, r1, v0, v1;
  TFObj *tf = nullptr;
};

struct ModelObj {
  std::vector<MeshObj> meshes;
  std::vector<SpheresObj> spheres;
  std::vector<PathObj> paths;
  int ngeom = 0;
  bool committed = false;
  DModel d;
  std::vector<void *> owned;   // device buffers built at commit
  DGeomShade *shade = nullptr;
  int *hit_geom = nullptr, *hit_prim = nullptr;
  int tnodes = 0, snodes = 0, tdepth = 0, sdepth = 0, tdepth4 = 0, sdepth4 = 0;
  int cnodes = 0, cdepth = 0, cdepth4 = 0;
  float tsah = 0.f, ssah = 0.f, csah = 0.f;   // SAH cost of each tree over its root's surface
  void free_built() {
    for (void *p : owned) dfree(p);
    owned.clear();
    committed = false;
  }
};

struct RayListObj {
  void *block = nullptr;
  size_t bytes = 0;
  // what the rank loop knows of the lists it makes itself (0 for others):
  // kinds 1 every ray primary, 2 none primary; unique: each pixel at most once
  int kinds = 0;
  int unique = 0;
  bool hdr_dirty = false;   // the device copy of hdr not written yet (lists only the loop sees)
  bool pretraced = false;   // its rays' trace is already queued (the rank loop's next pass skips it)
  gxy_raylist_hdr hdr;
  RayList_ispc ispc;
};

static void set_pointers(RayListObj *r) {
  char *base = (char *)r->block + 64;
  size_t nn = (size_t)r->hdr.aligned_size;
  float **fp = &r->ispc.ox;
  for (int k = 0; k < 20; k++) fp[k] = (float *)(base + (size_t)k * nn * 4);
  int **ip = &r->ispc.x;
  for (int k = 0; k < 5; k++) ip[k] = (int *)(base + (size_t)(20 + k) * nn * 4);
}

// write_header false: the 64-byte block header is written only when the
// block leaves the library (lists the rank loop makes and consumes itself)
static RayListObj *raylist_new(int nrays, int frame, int type, bool write_header = true) {
  if (nrays < 0) {
    fail("negative ray count");
    return nullptr;
  }
  RayListObj *r = new RayListObj;
  memset(&r->hdr, 0, sizeof r->hdr);
  int nn = (nrays + 15) & ~15;
  r->hdr.size = nrays;
  r->hdr.frame = frame;
  r->hdr.aligned_size = nn;
  r->hdr.type = type;
  r->hdr.id = G.raylist_id++;
  r->bytes = 64 + (size_t)nn * 100;
  r->block = dalloc(r->bytes);
  if (!r->block) {
    delete r;
    return nullptr;
  }
  set_pointers(r);
  if (!write_header) {
    r->hdr_dirty = true;
  } else if (launch_write_header(r->block, &r->hdr, S()) != hipSuccess) {
    fail("RayList header write failed");
    dfree(r->block);
    delete r;
    return nullptr;
  }
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.raylists.insert(r);
  return r;
}

// writes the block header of a list whose header was left to be written
static int raylist_sync_header(RayListObj *r) {
  if (!r->hdr_dirty) return 0;
  HIPCHK(launch_write_header(r->block, &r->hdr, S()));
  r->hdr_dirty = false;
  return 0;
}

static void raylist_delete(RayListObj *r, bool sync = true) {
  if (!r) return;
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    G.raylists.erase(r);
  }
  if (sync) dfree(r->block);
  else dfree_stream_ordered(r->block);
  delete r;
}

template <class T>
static T *lookup(std::set<void *> &reg, void *h) {
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  return reg.count(h) ? (T *)h : nullptr;
}

// ------------------------------------------------------------------ param builders
static int make_trace_params(const Visualization_ispc *vis, float global_eps, TraceParams &P) {
  memset(&P, 0, sizeof P);
  if (!vis) return fail("null Visualization_ispc");
  if (vis->nVolumeVis < 0 || vis->nVolumeVis > GXY_MAX_VOLVIS)
    return fail("%d volume visualizations; at most %d per launch", vis->nVolumeVis, GXY_MAX_VOLVIS);
  P.nvv = vis->nVolumeVis;
  P.step = -1;
  for (int m = 0; m < vis->nVolumeVis; m++) {
    const VolumeVis_ispc *vv = vis->volumeVis[m];
    if (!vv) return fail("null VolumeVis_ispc %d", m);
    VolumeObj *V = lookup<VolumeObj>(G.volumes, vv->vis.vis.data);
    TFObj *T = lookup<TFObj>(G.tfs, vv->vis.transferFunction);
    if (!V) return fail("VolumeVis %d: data is not a gxy volume handle", m);
    if (!T) return fail("VolumeVis %d: transferFunction is not a gxy TF handle", m);
    if (vv->nSlices > GXY_MAX_SLICES) return fail("%d slices; at most %d", vv->nSlices, GXY_MAX_SLICES);
    if (vv->nIsovalues > GXY_MAX_ISO) return fail("%d isovalues; at most %d", vv->nIsovalues, GXY_MAX_ISO);
    DVolVis &D = P.vv[m];
    D.vol = V->d;
    D.tf = T->d;
    D.nSlices = vv->nSlices;
    for (int k = 0; k < 4 * vv->nSlices; k++) D.slices[k] = vv->slices[k];
    D.nIso = vv->nIsovalues;
    for (int k = 0; k < vv->nIsovalues; k++) D.iso[k] = vv->isovalues[k];
    D.dvr = vv->volume_render ? 1 : 0;
    if (D.dvr || D.nIso > 0) P.integrate = 1;
    // TraceRays.ispc:342-361
    float s = V->d.step * V->d.rate;
    if (P.step < 0 || P.step > s) {
      P.step = s;
      P.eps = global_eps * s;
    }
  }
  for (int k = 0; k < 3; k++) {
    P.blo[k] = vis->local_bb.lower[k];
    P.bhi[k] = vis->local_bb.upper[k];
  }
  if (P.step != P.step) return fail("the volume step is NaN (check the volumes' sampling rates)");
  P.istep = P.step > 0 ? 1.0f / P.step : 0.f;
  if (vis->model) {
    ModelObj *M = lookup<ModelObj>(G.models, vis->model);
    if (!M) return fail("Visualization model is not a gxy model handle");
    if (!M->committed) return fail("model used before gxy_model_commit");
    P.has_model = (M->d.ntris > 0 || M->d.nsph > 0 || M->d.ncurve > 0) ? 1 : 0;
    P.model = M->d;
    P.model.wide = (G.bvh_width == 4 && (M->d.ntris == 0 || M->d.tnodes4) && (M->d.nsph == 0 || M->d.snodes4)) ? 1 : 0;
    P.hit_geom = M->hit_geom;
    P.hit_prim = M->hit_prim;
  }
  P.sample_count = nullptr;
  P.opaque_alpha = G.opaque_alpha;
  return 0;
}

static int make_light_params(const Lighting_ispc *l, float eps, LightParams &L) {
  memset(&L, 0, sizeof L);
  if (!l) return fail("null Lighting_ispc");
  if (l->nLights > 16) return fail("%u lights; at most 16", l->nLights);
  L.nAO = l->n_ao_rays;
  L.nLights = (int)l->nLights;
  L.doShadows = l->do_shadows;
  L.aoRadius = l->ao_radius;
  L.Ka = l->Ka;
  L.Kd = l->Kd;
  L.eps = eps;
  L.reverse = G.reverse_lighting;
  L.U = G.uv;
  L.V = G.uv + 256;
  L.ao_pcg = G.ao_pcg;
  L.ao_seed = G.ao_seed;
  for (int k = 0; k < L.nLights; k++) {
    for (int a = 0; a < 3; a++) L.L[3 * k + a] = l->lights[3 * k + a];
    L.types[k] = l->types ? l->types[k] : 2;
  }
  return 0;
}

}  // namespace gxy

using namespace gxy;

// ====================================================================== C-ABI
GXY_EXPORT int gxy_init(int device) {
  if (G.device >= 0) return 0;
  char buf[16];
  snprintf(buf, sizeof buf, "%d", device);
  if (device > 0) setenv("GXY_DEVICE", buf, 1);
  return ensure_init();
}
GXY_EXPORT const char *gxy_last_error(void) { return t_err.c_str(); }
GXY_EXPORT int gxy_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}
GXY_EXPORT int gxy_synchronize(void) {
  INIT_OR(-1);
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}
GXY_EXPORT void *gxy_context_create(void) {
  INIT_OR(nullptr);
  Ctx *c = new Ctx;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fail("hipStreamCreate failed");
    delete c;
    return nullptr;
  }
  Ctx *prev = t_ctx;
  t_ctx = c;
  c->counters = (unsigned long long *)dalloc(8 * sizeof(unsigned long long));
  // the readback area now, so that no later call on this context allocates it
  if (!c->counters || !readback_area(*c)) {
    dfree_stream_ordered(c->counters);
    {
      std::lock_guard<std::recursive_mutex> lk(G.mu);
      release_cache(*c);
      if (c->hread) G.free_reads.push_back({c->hread, c->dread});
    }
    t_ctx = prev;
    hipStreamDestroy(c->stream);
    delete c;
    return nullptr;
  }
  t_ctx = prev;
  hipMemsetAsync(c->counters, 0, 8 * sizeof(unsigned long long), c->stream);
  hipStreamSynchronize(c->stream);
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.contexts.insert(c);
  return c;
}
GXY_EXPORT int gxy_context_bind(void *ctx) {
  if (ctx == nullptr) {
    t_ctx = nullptr;
    return 0;
  }
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  if (!G.contexts.count(ctx)) return fail("not a context handle");
  t_ctx = (Ctx *)ctx;
  return 0;
}
GXY_EXPORT int gxy_context_destroy(void *ctx) {
  Ctx *c;
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    if (!G.contexts.count(ctx)) return fail("not a context handle");
    G.contexts.erase(ctx);
    c = (Ctx *)ctx;
  }
  hipStreamSynchronize(c->stream);
  Ctx *prev = t_ctx;
  t_ctx = c;
  dfree_stream_ordered(c->counters);
  t_ctx = prev;
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    release_cache(*c);
    if (c->hread) G.free_reads.push_back({c->hread, c->dread});
  }
  hipStreamDestroy(c->stream);
  if (t_ctx == c) t_ctx = nullptr;
  delete c;
  return 0;
}

GXY_EXPORT int gxy_set_stream(void *s) {
  INIT_OR(-1);
  G.user_stream = (hipStream_t)s;
  return 0;
}
GXY_EXPORT void *gxy_get_stream(void) {
  INIT_OR(nullptr);
  return (void *)S();
}
GXY_EXPORT void *gxy_device_alloc(size_t bytes) {
  INIT_OR(nullptr);
  return dalloc(bytes);
}
GXY_EXPORT int gxy_device_free(void *p) {
  INIT_OR(-1);
  dfree(p);
  return 0;
}
GXY_EXPORT int gxy_copy_to_device(void *dst, const void *src, size_t bytes) {
  INIT_OR(-1);
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S()));
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}
GXY_EXPORT int gxy_copy_to_host(void *dst, const void *src, size_t bytes) {
  INIT_OR(-1);
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}
GXY_EXPORT int gxy_copy_device(void *dst, const void *src, size_t bytes) {
  INIT_OR(-1);
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S()));
  return 0;
}
GXY_EXPORT int gxy_memset(void *dst, int value, size_t bytes) {
  INIT_OR(-1);
  HIPCHK(hipMemsetAsync(dst, value, bytes, S()));
  return 0;
}
GXY_EXPORT size_t gxy_device_bytes_in_use(void) { return G.bytes_live; }
GXY_EXPORT int gxy_set_opacity_threshold(float alpha) {
  if (!(alpha > 0.f && alpha <= 1.f)) return fail("opacity threshold %g outside (0, 1]", alpha);
  G.opaque_alpha = alpha;
  return 0;
}
GXY_EXPORT float gxy_get_opacity_threshold(void) { return G.opaque_alpha; }
GXY_EXPORT int gxy_set_trace_config(int kind, int chunk, int refill) {
  // kind 0: one ray per work-item; 10 + w: the same with a w-waves-per-SIMD
  // occupancy hint (w = 5, 6; 7 and 8 for one bricked volume with at most one
  // isovalue, 6 otherwise), chunk 64 or 128 being the workgroup size of w = 6
  // on one bricked volume with at most one isovalue; 1: persistent lanes
  int wpe = 0;
  if (kind >= 10) {
    wpe = kind - 10;
    if (wpe < 5 || wpe > 8) return fail("trace occupancy variant %d: 15, 16, 17 or 18", kind);
    kind = 0;
  }
  if (kind < 0 || kind > 1) return fail("trace kernel kind %d: 0, 1, 15, 16, 17 or 18", kind);
  if (chunk < 64 || chunk % 64 || refill < 1 || refill > 64)
    return fail("trace config chunk %d (multiple of 64) refill %d (1..64)", chunk, refill);
  G.trace_kind = kind;
  G.trace_wpe = wpe;
  G.trace_chunk = chunk;
  G.trace_refill = refill;
  return 0;
}
GXY_EXPORT int gxy_set_geometry_stack(int entries) {
  if (entries != 0 && entries != 16 && entries != 24 && entries != 32)
    return fail("geometry stack %d: 0 (private, 64 entries), 16, 24 or 32 (LDS)", entries);
  G.geom_stack = entries;
  return 0;
}
GXY_EXPORT int gxy_get_geometry_stack(void) { return G.geom_stack; }
GXY_EXPORT int gxy_set_occlusion_rays(int on) {
  INIT_OR(-1);
  G.occlusion = on ? 1 : 0;
  return 0;
}
GXY_EXPORT int gxy_get_occlusion_rays(void) { return G.occlusion; }
GXY_EXPORT int gxy_set_bvh_width(int width) {
  if (width != 2 && width != 4) return fail("BVH width %d: 2 or 4", width);
  G.bvh_width = width;
  return 0;
}
GXY_EXPORT int gxy_get_bvh_width(void) { return G.bvh_width; }
GXY_EXPORT int gxy_set_geometry_waves(int waves) {
  if (waves != 0 && waves != 8) return fail("geometry waves %d: 0 or 8", waves);
  G.geom_waves = waves;
  return 0;
}
GXY_EXPORT int gxy_get_geometry_waves(void) { return G.geom_waves; }
GXY_EXPORT int gxy_get_trace_config(int *kind, int *chunk, int *refill) {
  *kind = G.trace_kind == 0 && G.trace_wpe ? 10 + G.trace_wpe : G.trace_kind;
  *chunk = G.trace_chunk;
  *refill = G.trace_refill;
  return 0;
}
GXY_EXPORT size_t gxy_device_bytes_total(void) {
  if (ensure_init() != 0) return 0;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
  return tot;
}

GXY_EXPORT int gxy_timing_enable(int on) {
  INIT_OR(-1);
  if (on < 0 || on > 2) return fail("timing mode %d: 0, 1 or 2", on);
  G.timing = on;
  return 0;
}
GXY_EXPORT int gxy_timing_reset(void) {
  INIT_OR(-1);
  resolve_timings();
  G.timings.clear();
  return 0;
}
GXY_EXPORT int gxy_timing_get(const char *kernel, double *total_ms, int *launches) {
  INIT_OR(-1);
  resolve_timings();
  auto it = G.timings.find(kernel);
  *total_ms = it == G.timings.end() ? 0.0 : it->second.ms;
  *launches = it == G.timings.end() ? 0 : it->second.launches;
  return 0;
}

// ------------------------------------------------------------------ RayLists
GXY_EXPORT void *gxy_raylist_create(int nrays, int frame, int type, RayList_ispc *ispc) {
  INIT_OR(nullptr);
  RayListObj *r = raylist_new(nrays, frame, type);
  if (r && ispc) *ispc = r->ispc;
  return r;
}
GXY_EXPORT int gxy_raylist_destroy(void *rl) {
  RayListObj *r = lookup<RayListObj>(G.raylists, rl);
  if (!r) return fail("not a raylist handle");
  raylist_delete(r);
  return 0;
}
GXY_EXPORT int gxy_raylist_size(void *rl) {
  RayListObj *r = lookup<RayListObj>(G.raylists, rl);
  return r ? r->hdr.size : fail("not a raylist handle");
}
GXY_EXPORT int gxy_raylist_aligned_size(void *rl) {
  RayListObj *r = lookup<RayListObj>(G.raylists, rl);
  return r ? r->hdr.aligned_size : fail("not a raylist handle");
}
GXY_EXPORT void *gxy_raylist_split_part(void *rl, int start, int count, RayList_ispc *ispc) {
  RayListObj *r = lookup<RayListObj>(G.raylists, rl);
  if (!r) {
    fail("not a raylist handle");
    return nullptr;
  }
  if (start < 0 || count < 0 || start + count > r->hdr.size) {
    fail("split range [%d, %d) outside list of %d", start, start + count, r->hdr.size);
    return nullptr;
  }
  RayListObj *p = raylist_new(count, r->hdr.frame, r->hdr.type);
  if (!p) return nullptr;
  if (launch_copy_rays(r->ispc, start, count, p->ispc, 0, S()) != hipSuccess) {
    fail("copy kernel failed");
    raylist_delete(p);
    return nullptr;
  }
  if (ispc) *ispc = p->ispc;
  return p;
}
GXY_EXPORT int gxy_raylist_truncate(void *rl, int n, RayList_ispc *ispc) {
  RayListObj *r = lookup<RayListObj>(G.raylists, rl);
  if (!r) return fail("not a raylist handle");
  if (n < 0 || n > r->hdr.size) return fail("truncate to %d of %d rays", n, r->hdr.size);
  int nn = (n + 15) & ~15;
  // Rays.cpp:245-286: shrink the block only when the aligned size drops
  if (nn < r->hdr.aligned_size) {
    RayListObj tmp = *r;
    r->hdr.size = n;
    r->hdr.aligned_size = nn;
    r->bytes = 64 + (size_t)nn * 100;
    r->block = dalloc(r->bytes);
    if (!r->block) return -1;
    set_pointers(r);
    if (launch_copy_rays(tmp.ispc, 0, n, r->ispc, 0, S()) != hipSuccess) return fail("copy kernel failed");
    HIPCHK(hipMemcpyAsync(r->block, &r->hdr, sizeof r->hdr, hipMemcpyHostToDevice, S()));
    dfree(tmp.block);
  } else {
    r->hdr.size = n;
    HIPCHK(hipMemcpyAsync(r->block, &r->hdr, sizeof r->hdr, hipMemcpyHostToDevice, S()));
    HIPCHK(hipStreamSynchronize(S()));
  }
  if (ispc) *ispc = r->ispc;
  return 0;
}
GXY_EXPORT size_t gxy_raylist_bytes(void *rl) {
  RayListObj *r = lookup<RayListObj>(G.raylists, rl);
  return r ? r->bytes : 0;
}
GXY_EXPORT int gxy_raylist_download(void *rl, void *host_block) {
  RayListObj *r = lookup<RayListObj>(G.raylists, rl);
  if (!r) return fail("not a raylist handle");
  if (raylist_sync_header(r) != 0) return -1;
  HIPCHK(hipMemcpyAsync(host_block, r->block, r->bytes, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}
GXY_EXPORT void *gxy_raylist_upload(const void *host_block, RayList_ispc *ispc) {
  INIT_OR(nullptr);
  const gxy_raylist_hdr *h = (const gxy_raylist_hdr *)host_block;
  if (h->aligned_size != ((h->size + 15) & ~15) || h->size < 0) {
    fail("raylist block header inconsistent (size %d aligned %d)", h->size, h->aligned_size);
    return nullptr;
  }
  RayListObj *r = raylist_new(h->size, h->frame, h->type);
  if (!r) return nullptr;
  r->hdr = *h;
  if (hipMemcpyAsync(r->block, host_block, r->bytes, hipMemcpyHostToDevice, S()) != hipSuccess ||
      hipStreamSynchronize(S()) != hipSuccess) {
    fail("upload failed");
    raylist_delete(r);
    return nullptr;
  }
  if (ispc) *ispc = r->ispc;
  return r;
}

// ------------------------------------------------------------------ volumes
GXY_EXPORT void *gxy_volume_create(const int dims[3], const float origin[3], const float spacing[3],
                                   const float *voxels, float samplingRate) {
  INIT_OR(nullptr);
  for (int k = 0; k < 3; k++)
    if (dims[k] < 2) {
      fail("volume dimension %d is %d; at least 2 voxels per axis", k, dims[k]);
      return nullptr;
    }
  VolumeObj *v = new VolumeObj;
  memset(&v->d, 0, sizeof v->d);
  v->count = (size_t)dims[0] * dims[1] * dims[2];
  v->data = (float *)dalloc(v->count * 4);
  if (!v->data) {
    delete v;
    return nullptr;
  }
  for (int k = 0; k < 3; k++) {
    v->d.n[k] = dims[k];
    v->d.org[k] = origin[k];
    v->d.sp[k] = spacing[k];
    v->d.isp[k] = 1.0f / spacing[k];
    v->d.ub[k] = std::nextafter((float)(dims[k] - 1), 0.0f);
  }
  v->d.data = v->data;
  v->d.rate = samplingRate;
  v->d.irate = 1.0f / samplingRate;
  // 32-bit byte offsets (at most 4 (2^30 - 1)) and 24-bit index products
  v->d.small = (v->count <= (size_t)1 << 30 && (size_t)v->d.n[0] * v->d.n[1] < ((size_t)1 << 24) &&
                v->d.n[2] < (1 << 24)) ? 1 : 0;
  v->d.step = std::min(spacing[0], std::min(spacing[1], spacing[2]));
  if (voxels) {
    if (hipMemcpyAsync(v->data, voxels, v->count * 4, hipMemcpyHostToDevice, S()) != hipSuccess ||
        hipStreamSynchronize(S()) != hipSuccess) {
      fail("volume upload failed");
      dfree(v->data);
      delete v;
      return nullptr;
    }
  }
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.volumes.insert(v);
  return v;
}
GXY_EXPORT int gxy_volume_destroy(void *vol) {
  VolumeObj *v = lookup<VolumeObj>(G.volumes, vol);
  if (!v) return fail("not a volume handle");
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    G.volumes.erase(v);
  }
  dfree(v->data);
  dfree(v->bdata);
  delete v;
  return 0;
}
GXY_EXPORT int gxy_volume_set_layout(void *vol, int brick) {
  INIT_OR(-1);
  VolumeObj *v = lookup<VolumeObj>(G.volumes, vol);
  if (!v) return fail("not a volume handle");
  if (brick != 0 && brick != 4 && brick != 8) return fail("volume layout %d: 0 (linear), 4 or 8 (bricks)", brick);
  if (brick == v->d.brick) return 0;
  dfree(v->bdata);
  v->bdata = nullptr;
  v->d.bdata = nullptr;
  v->d.brick = 0;
  v->d.bsmall = 0;
  if (brick == 0) return 0;
  int nb[3];
  size_t nbr = 1;
  for (int k = 0; k < 3; k++) {
    nb[k] = (v->d.n[k] - 1 + brick - 1) / brick;
    nbr *= (size_t)nb[k];
  }
  if ((size_t)nb[0] * nb[1] >= ((size_t)1 << 24) || nb[1] >= (1 << 24) || nb[2] >= (1 << 24))
    return fail("volume too large for the bricked layout");
  v->bdata = (float *)dalloc(nbr * brick_floats(brick) * sizeof(float));
  if (!v->bdata) return -1;
  for (int k = 0; k < 3; k++) v->d.nb[k] = nb[k];
  v->d.bdata = v->bdata;
  v->d.brick = brick;
  // a cell's last load ends (PL + E + 2) floats past its offset, inside its own brick
  v->d.bsmall = nbr * brick_floats(brick) * sizeof(float) < ((size_t)1 << 32) &&
                        (size_t)brick_floats(brick) * nb[0] * nb[1] < ((size_t)1 << 24)
                    ? 1
                    : 0;
  return volume_brickify(v);
}
GXY_EXPORT int gxy_volume_get_layout(void *vol) {
  VolumeObj *v = lookup<VolumeObj>(G.volumes, vol);
  if (!v) return fail("not a volume handle");
  return v->d.brick;
}
GXY_EXPORT float *gxy_volume_device_data(void *vol) {
  VolumeObj *v = lookup<VolumeObj>(G.volumes, vol);
  return v ? v->data : nullptr;
}
GXY_EXPORT int gxy_volume_set_sampling_rate(void *vol, float rate) {
  VolumeObj *v = lookup<VolumeObj>(G.volumes, vol);
  if (!v) return fail("not a volume handle");
  v->d.rate = rate;
  v->d.irate = 1.0f / rate;
  return 0;
}
GXY_EXPORT int gxy_volume_fill_noise(void *vol, uint32_t seed, int octaves, float base_frequency) {
  VolumeObj *v = lookup<VolumeObj>(G.volumes, vol);
  if (!v) return fail("not a volume handle");
  // seeded permutation: Fisher-Yates driven by a PCG32 stream
  int perm[256];
  for (int i = 0; i < 256; i++) perm[i] = i;
  uint64_t state = 0x853c49e6748fea9bull ^ ((uint64_t)seed << 1), inc = ((uint64_t)seed << 1) | 1u;
  auto pcg = [&]() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  };
  for (int i = 255; i > 0; i--) {
    int j = (int)(pcg() % (uint32_t)(i + 1));
    std::swap(perm[i], perm[j]);
  }
  int *dperm = (int *)dalloc(256 * sizeof(int) + 16);
  float *mm = (float *)(dperm + 256);
  if (!dperm) return -1;
  HIPCHK(hipMemcpyAsync(dperm, perm, sizeof perm, hipMemcpyHostToDevice, S()));
  HIPCHK(launch_noise(v->data, v->d.n[0], v->d.n[1], v->d.n[2], dperm, octaves, base_frequency, mm, S()));
  HIPCHK(hipStreamSynchronize(S()));
  dfree(dperm);
  return volume_brickify(v);
}
GXY_EXPORT int gxy_volume_download(void *vol, float *host) {
  VolumeObj *v = lookup<VolumeObj>(G.volumes, vol);
  if (!v) return fail("not a volume handle");
  HIPCHK(hipMemcpyAsync(host, v->data, v->count * 4, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}

// ------------------------------------------------------------------ TFs
static void *tf_make(std::vector<float> col, std::vector<float> op, float lo, float hi) {
  if (col.size() / 3 > GXY_TF_MAX || op.size() > GXY_TF_MAX) {
    fail("transfer function tables are limited to %d entries", GXY_TF_MAX);
    return nullptr;
  }
  TFObj *t = new TFObj;
  t->hcol = std::move(col);
  t->hop = std::move(op);
  // The device lookups read two entries at least: no colours is white
  // everywhere, no opacities opaque, and a one-entry table is that entry
  // everywhere (two equal entries at scale 0), so the lookups need no cases.
  std::vector<float> dcol = t->hcol, dop = t->hop;
  if (dcol.empty()) dcol = {1.f, 1.f, 1.f};
  if (dcol.size() == 3) dcol.insert(dcol.end(), {dcol[0], dcol[1], dcol[2]});
  if (dop.empty()) dop = {1.f};
  if (dop.size() == 1) dop.push_back(dop[0]);
  t->col = (float *)dalloc(dcol.size() * 4);
  t->op = (float *)dalloc(dop.size() * 4);
  if (!t->col || !t->op) {
    dfree(t->col);
    dfree(t->op);
    delete t;
    return nullptr;
  }
  hipMemcpyAsync(t->col, dcol.data(), dcol.size() * 4, hipMemcpyHostToDevice, S());
  hipMemcpyAsync(t->op, dop.data(), dop.size() * 4, hipMemcpyHostToDevice, S());
  hipStreamSynchronize(S());
  t->d.col = t->col;
  t->d.op = t->op;
  t->d.nc = (int)dcol.size() / 3;
  t->d.no = (int)dop.size();
  t->d.lo = lo;
  t->d.hi = hi;
  t->d.sc = t->hcol.size() / 3 > 1 ? (float)(t->d.nc - 1) / (hi - lo) : 0.f;
  t->d.so = t->hop.size() > 1 ? (float)(t->d.no - 1) / (hi - lo) : 0.f;
  t->d.xc = (float)(t->d.nc - 1);
  t->d.xo = (float)(t->d.no - 1);
  t->d.lo2 = 2.f * lo;
  t->d.hsc = 0.5f * t->d.sc;
  t->d.hso = 0.5f * t->d.so;
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.tfs.insert(t);
  return t;
}

// MappedVis::local_commit (MappedVis.cpp:286-333)
GXY_EXPORT void *gxy_tf_create_from_maps(const float *cmap, int nc, const float *omap, int no,
                                         const float *range) {
  INIT_OR(nullptr);
  if (nc < 2 || no < 2) {
    fail("colormap and opacitymap need at least two points");
    return nullptr;
  }
  std::vector<float> col(256 * 3), op(256);
  int i0 = 0, i1 = 1;
  float xmin = cmap[0], xmax = cmap[4 * (nc - 1)];
  for (int i = 0; i < 256; i++) {
    float x = (float)(xmin + (i / (255.0)) * (xmax - xmin));
    if (x > xmax) x = xmax;
    while (i1 < nc - 1 && cmap[4 * i1] < x) i0++, i1++;
    float d = (x - cmap[4 * i0]) / (cmap[4 * i1] - cmap[4 * i0]);
    col[3 * i + 0] = cmap[4 * i0 + 1] + d * (cmap[4 * i1 + 1] - cmap[4 * i0 + 1]);
    col[3 * i + 1] = cmap[4 * i0 + 2] + d * (cmap[4 * i1 + 2] - cmap[4 * i0 + 2]);
    col[3 * i + 2] = cmap[4 * i0 + 3] + d * (cmap[4 * i1 + 3] - cmap[4 * i0 + 3]);
  }
  i0 = 0, i1 = 1;
  xmin = omap[0], xmax = omap[2 * (no - 1)];
  for (int i = 0; i < 256; i++) {
    float x = (float)(xmin + (i / (255.0)) * (xmax - xmin));
    if (x > xmax) x = xmax;
    while (i1 < no - 1 && omap[2 * i1] < x) i0++, i1++;
    float d = (x - omap[2 * i0]) / (omap[2 * i1] - omap[2 * i0]);
    op[i] = omap[2 * i0 + 1] + d * (omap[2 * i1 + 1] - omap[2 * i0 + 1]);
  }
  float lo = range ? range[0] : cmap[0], hi = range ? range[1] : cmap[4 * (nc - 1)];
  return tf_make(std::move(col), std::move(op), lo, hi);
}
GXY_EXPORT void *gxy_tf_create(const float *colors, int nc, const float *opac, int no, float lo, float hi) {
  INIT_OR(nullptr);
  return tf_make(std::vector<float>(colors, colors + 3 * nc), std::vector<float>(opac, opac + no), lo, hi);
}
GXY_EXPORT int gxy_tf_destroy(void *tf) {
  TFObj *t = lookup<TFObj>(G.tfs, tf);
  if (!t) return fail("not a TF handle");
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    G.tfs.erase(t);
  }
  dfree(t->col);
  dfree(t->op);
  delete t;
  return 0;
}
GXY_EXPORT int gxy_tf_get_tables(void *tf, float *c, float *o, float *range) {
  TFObj *t = lookup<TFObj>(G.tfs, tf);
  if (!t) return fail("not a TF handle");
  memcpy(c, t->hcol.data(), t->hcol.size() * 4);
  memcpy(o, t->hop.data(), t->hop.size() * 4);
  range[0] = t->d.lo;
  range[1] = t->d.hi;
  return 0;
}

// ------------------------------------------------------------------ models
template <class T>
static T *upload(const T *h, size_t n) {
  if (!h || n == 0) return nullptr;
  T *d = (T *)dalloc(n * sizeof(T));
  if (!d) return nullptr;
  hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, S());
  return d;
}

GXY_EXPORT void *gxy_model_create(void) {
  INIT_OR(nullptr);
  ModelObj *m = new ModelObj;
  memset(&m->d, 0, sizeof m->d);
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.models.insert(m);
  return m;
}
GXY_EXPORT int gxy_model_destroy(void *model) {
  ModelObj *M = lookup<ModelObj>(G.models, model);
  if (!M) return fail("not a model handle");
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    G.models.erase(M);
  }
  M->free_built();
  for (auto &m : M->meshes) { dfree(m.V); dfree(m.I); dfree(m.N); dfree(m.D); }
  for (auto &s : M->spheres) { dfree(s.C); dfree(s.D); }
  for (auto &p : M->paths) { dfree(p.CP); dfree(p.IC); }
  delete M;
  return 0;
}
GXY_EXPORT int gxy_model_add_triangles(void *model, const float *V, int nv, const uint32_t *I, int nt,
                                       const float *N, const float *D, void *tf) {
  ModelObj *M = lookup<ModelObj>(G.models, model);
  if (!M) return fail("not a model handle");
  if (nv <= 0 || nt <= 0 || !V || !I) return fail("empty triangle mesh");
  for (int64_t k = 0; k < 3 * (int64_t)nt; k++)
    if (I[k] >= (uint32_t)nv) return fail("triangle index %u out of range (%d vertices)", I[k], nv);
  MeshObj m;
  m.geomID = M->ngeom++;
  m.nv = nv;
  m.nt = nt;
  m.V = upload(V, 3 * (size_t)nv);
  m.I = upload(I, 3 * (size_t)nt);
  m.N = upload(N, N ? 3 * (size_t)nv : 0);
  m.D = upload(D, D ? (size_t)nv : 0);
  m.tf = tf ? lookup<TFObj>(G.tfs, tf) : nullptr;
  if (tf && !m.tf) return fail("mesh TF is not a TF handle");
  HIPCHK(hipStreamSynchronize(S()));
  M->meshes.push_back(m);
  M->free_built();
  return m.geomID;
}
GXY_EXPORT int gxy_model_add_spheres(void *model, const float *C, const float *D, int n, float r0, float r1,
                                     float v0, float v1, void *tf) {
  ModelObj *M = lookup<ModelObj>(G.models, model);
  if (!M) return fail("not a model handle");
  if (n <= 0 || !C) return fail("empty sphere set");
  SpheresObj s;
  s.geomID = M->ngeom++;
  s.n = n;
  s.C = upload(C, 3 * (size_t)n);
  std::vector<float> zero;
  if (!D) zero.assign(n, 0.f);
  s.D = upload(D ? D : zero.data(), (size_t)n);
  s.r0 = r0; s.r1 = r1; s.v0 = v0; s.v1 = v1;
  // DataDrivenSpheresGeometry_set (DataDrivenSpheres.ispc:211-220)
  float e = std::log(r0);
  if (e < 0.f) e = -1.f / e;
  if (e > r0 / 100.0) e = (float)(r0 / 100.0);
  s.eps = e;
  s.tf = tf ? lookup<TFObj>(G.tfs, tf) : nullptr;
  if (tf && !s.tf) return fail("sphere TF is not a TF handle");
  HIPCHK(hipStreamSynchronize(S()));
  M->spheres.push_back(s);
  M->free_built();
  return s.geomID;
}

// MAP_RADIUS of DataDrivenPathLines.cpp:11-19
static float path_radius(float d, float r0, float r1, float v0, float v1) {
  if (v0 == v1) return r0;
  const float R = (d - v0) / (v1 - v0);
  return (R <= 0.0f) ? r0 : (R >= 1.0f) ? r1 : r0 + R * (r1 - r0);
}

// Round cones per curve segment (control points c, x y z radius each).  With
// n equal steps of u, the pieces' centres and radii are the linear
// interpolants of p(u) and r(u) on each step, which differ from them by at
// most (1/n)^2 / 8 times max|p''| and max|r''|; a ball moved by d and resized
// by e stays within d + e of the original, so the union of the pieces (the
// traced solid) and the union of the spheres (p(u), r(u)) (the swept-sphere
// solid) lie within (1/n)^2 / 8 (max|p''| + max|r''|) of each other.  For a
// cubic, p'' is linear in u: max|p''| <= 6 max(|c0 - 2c1 + c2|, |c1 - 2c2 + c3|),
// likewise for r, and r(u) >= the least control radius.  The count is the
// least n that brings the bound to GXY_CURVE_TOL times that radius (one piece
// for a straight segment of constant radius), at most GXY_CURVE_MAX_PIECES.
static int curve_pieces(const float *c) {
  double dp = 0.0, dr = 0.0, rmin = c[3];
  for (int k = 0; k < 2; k++) {
    double d[4];
    for (int a = 0; a < 4; a++)
      d[a] = (double)c[4 * k + a] - 2.0 * (double)c[4 * (k + 1) + a] + (double)c[4 * (k + 2) + a];
    const double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    dp = l > dp ? l : dp;
    dr = std::fabs(d[3]) > dr ? std::fabs(d[3]) : dr;
  }
  for (int k = 1; k < 4; k++) rmin = (double)c[4 * k + 3] < rmin ? (double)c[4 * k + 3] : rmin;
  const double m = 6.0 * (dp + dr);
  if (!(m > 0.0)) return 1;
  if (!(rmin > 0.0)) return GXY_CURVE_MAX_PIECES;
  const double n = std::ceil(std::sqrt(m / (8.0 * GXY_CURVE_TOL * rmin)));
  return n < 1.0 ? 1 : (n > (double)GXY_CURVE_MAX_PIECES ? GXY_CURVE_MAX_PIECES : (int)n);
}

GXY_EXPORT int gxy_model_add_pathlines(void *model, const float *V, const float *D, int nv, const int *I, int nseg,
                                       float r0, float r1, float v0, float v1, void *tf) {
  ModelObj *M = lookup<ModelObj>(G.models, model);
  if (!M) return fail("not a model handle");
  if (nv <= 0 || !V || !D || nseg <= 0 || !I) return fail("empty path-line set");
  for (int i = 0; i < nseg; i++)
    if (I[i] < 0 || I[i] + 1 >= nv) return fail("path-line segment %d starts at vertex %d of %d", i, I[i], nv);
  PathObj P;
  P.r0 = r0; P.r1 = r1; P.v0 = v0; P.v1 = v1;
  P.tf = tf ? lookup<TFObj>(G.tfs, tf) : nullptr;
  if (tf && !P.tf) return fail("path-line TF is not a TF handle");
  // control points: a segment continuing the previous one starts with the
  // previous segment's outgoing tangent, one followed by a continuation ends
  // with a tangent through the next vertex split in the ratio of the two
  // segment lengths; curve ends repeat their end point
  // (DataDrivenPathLines.cpp:97-157)
  std::vector<float> cp;
  std::vector<uint32_t> ic;
  auto v3 = [&](int i, int k) { return V[3 * (size_t)i + k]; };
  auto push = [&](float x, float y, float z, float r) { cp.insert(cp.end(), {x, y, z, r}); };
  bool middle = false;
  float tan[3] = {0.f, 0.f, 0.f};
  for (int i = 0; i < nseg; i++) {
    const int idx = I[i];
    const float st[3] = {v3(idx, 0), v3(idx, 1), v3(idx, 2)}, en[3] = {v3(idx + 1, 0), v3(idx + 1, 1), v3(idx + 1, 2)};
    const float dx = st[0] - en[0], dy = st[1] - en[1], dz = st[2] - en[2];
    const float lseg = std::sqrt(dx * dx + dy * dy + dz * dz);
    const float rs = path_radius(D[idx], r0, r1, v0, v1), re = path_radius(D[idx + 1], r0, r1, v0, v1);
    ic.push_back((uint32_t)(cp.size() / 4));
    push(st[0], st[1], st[2], rs);
    if (middle) push(st[0] + tan[0], st[1] + tan[1], st[2] + tan[2], (1.f - 1.f / 3) * rs + (1.f / 3) * re);
    else push(st[0], st[1], st[2], rs);
    middle = i + 1 < nseg && I[i + 1] == idx + 1;
    if (middle) {
      if (idx + 2 >= nv) return fail("path-line vertex %d out of range", idx + 2);
      const float nx[3] = {v3(idx + 2, 0), v3(idx + 2, 1), v3(idx + 2, 2)};
      float dl[3], bl2 = 0.f;
      for (int k = 0; k < 3; k++) {
        dl[k] = (1.f / 3) * (nx[k] - st[k]);
        bl2 += (nx[k] - en[k]) * (nx[k] - en[k]);
      }
      const float bl = std::sqrt(bl2);
      // two zero-length segments in a row: the reference's 0/0 taken as 1/2
      const float r = lseg + bl > 0.f ? lseg / (lseg + bl) : 0.5f;
      push(en[0] - r * dl[0], en[1] - r * dl[1], en[2] - r * dl[2], (1.f - 2.f / 3) * rs + (2.f / 3) * re);
      for (int k = 0; k < 3; k++) tan[k] = (1.f - r) * dl[k];
    } else {
      push(en[0], en[1], en[2], re);
    }
    push(en[0], en[1], en[2], re);
  }
  // pieces: the curve sampled at u = j / n (Bernstein weights), n from curve_pieces
  for (int i = 0; i < nseg; i++) {
    const float *c = &cp[4 * (size_t)ic[i]];
    const int np = curve_pieces(c);
    float prev[4] = {0, 0, 0, 0}, prevu = 0.f;
    for (int j = 0; j <= np; j++) {
      const float u = (float)j / (float)np, q = 1.0f - u;
      const float w0 = q * q * q, w1 = 3.0f * q * q * u, w2 = 3.0f * q * u * u, w3 = u * u * u;
      float x[4];
      for (int k = 0; k < 4; k++) x[k] = w0 * c[k] + w1 * c[4 + k] + w2 * c[8 + k] + w3 * c[12 + k];
      if (j > 0) {
        P.pa.insert(P.pa.end(), prev, prev + 4);
        P.pb.insert(P.pb.end(), x, x + 4);
        P.pu.insert(P.pu.end(), {prevu, u});
        P.pseg.push_back(i);
      }
      for (int k = 0; k < 4; k++) prev[k] = x[k];
      prevu = u;
    }
  }
  P.geomID = M->ngeom++;
  P.nseg = nseg;
  P.CP = upload(cp.data(), cp.size());
  P.IC = upload(ic.data(), ic.size());
  if (!P.CP || !P.IC || hipStreamSynchronize(S()) != hipSuccess) {
    dfree(P.CP);
    dfree(P.IC);
    return fail("path lines: device upload failed");
  }
  M->paths.push_back(std::move(P));
  M->free_built();
  return M->paths.back().geomID;
}

namespace gxy {
__global__ void k_tri_boxes(const float *V, const uint32_t *I, int nt, float *boxes) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nt) return;
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  for (int c = 0; c < 3; c++) {
    const float *v = V + 3 * (int64_t)I[3 * (int64_t)p + c];
    for (int a = 0; a < 3; a++) {
      lo[a] = v[a] < lo[a] ? v[a] : lo[a];
      hi[a] = v[a] > hi[a] ? v[a] : hi[a];
    }
  }
  for (int a = 0; a < 3; a++) { boxes[6 * (int64_t)p + a] = lo[a]; boxes[6 * (int64_t)p + 3 + a] = hi[a]; }
}

__device__ float sphere_radius_dev(float r0, float r1, float v0, float v1, float dv) {
  if (v0 != v1) {
    float d = (dv - v0) / (v1 - v0);
    if (d > 1) return r1;
    if (d < 0) return r0;
    return r0 + d * (r1 - r0);
  }
  return r0;
}

__global__ void k_sph_boxes(const float *C, const float *D, int n, float r0, float r1, float v0, float v1,
                            float *boxes, float *radius) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float r = sphere_radius_dev(r0, r1, v0, v1, D[p]);
  radius[p] = r;
  for (int a = 0; a < 3; a++) {
    boxes[6 * (int64_t)p + a] = C[3 * (int64_t)p + a] - r;
    boxes[6 * (int64_t)p + 3 + a] = C[3 * (int64_t)p + a] + r;
  }
}

// sorted triangle k <- primitive order[k] of the flat list (geom, prim)
__global__ void k_gather_tris(const int *order, const int *fgeom, const int *fprim, int n, const DGeomShade *shade,
                              float *tverts, int *tgeom, int *tprim) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int src = order[k];
  int g = fgeom[src], p = fprim[src];
  const DGeomShade &S = shade[g];
  for (int c = 0; c < 3; c++) {
    uint32_t vi = S.indices[3 * (int64_t)p + c];
    for (int a = 0; a < 3; a++) tverts[9 * (int64_t)k + 3 * c + a] = S.vertices[3 * (int64_t)vi + a];
  }
  tgeom[k] = g;
  tprim[k] = p;
}

__global__ void k_gather_sph(const int *order, const int *fgeom, const int *fprim, const float *fC,
                             const float *frad, const float *feps, int n, float *sc, float *sr, float *se, int *sg,
                             int *sp) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int src = order[k];
  for (int a = 0; a < 3; a++) sc[3 * (int64_t)k + a] = fC[3 * (int64_t)src + a];
  sr[k] = frad[src];
  se[k] = feps[src];
  sg[k] = fgeom[src];
  sp[k] = fprim[src];
}
}  // namespace gxy

// Structural check of a built tree: every child index in range, every inner
// node and every leaf reached exactly once from the root.  Run for trees of up
// to GXY_BVH_VALIDATE_MAX nodes (default 4M) before the tree is used.
// Structural check of a 4-wide tree built on the device: child indices in
// range, every node reached once from the root, every primitive in exactly one
// leaf.  Run for trees of up to GXY_BVH_VALIDATE_MAX nodes (default 4M).
static int validate_bvh4(const BVH4Node *d4, int nn4, int nprim, const char *what) {
  size_t lim = getenv("GXY_BVH_VALIDATE_MAX") ? (size_t)atol(getenv("GXY_BVH_VALIDATE_MAX")) : (size_t)(4u << 20);
  if ((size_t)nn4 > lim) return 0;
  std::vector<BVH4Node> n4(nn4);
  HIPCHK(hipMemcpyAsync(n4.data(), d4, sizeof(BVH4Node) * nn4, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  std::vector<char> seen_n(nn4, 0);
  std::vector<int> seen_p(nprim, 0);
  std::vector<int> st{0};
  while (!st.empty()) {
    const int n = st.back();
    st.pop_back();
    if (n < 0 || n >= nn4 || seen_n[n]) return fail("%s 4-wide BVH: node %d invalid or shared", what, n);
    seen_n[n] = 1;
    for (int k = 0; k < 4; k++) {
      if (!(n4[n].lox[k] <= n4[n].hix[k])) continue;   // empty slot
      const int c = n4[n].child[k];
      if (c >= 0) {
        st.push_back(c);
      } else {
        const int code = ~c, a = code >> 4, cnt = (code & 15) + 1;
        if (a < 0 || a + cnt > nprim) return fail("%s 4-wide BVH: leaf range %d+%d outside %d", what, a, cnt, nprim);
        for (int q = a; q < a + cnt; q++) seen_p[q]++;
      }
    }
  }
  for (int q = 0; q < nprim; q++)
    if (seen_p[q] != 1) return fail("%s 4-wide BVH: primitive %d in %d leaves", what, q, seen_p[q]);
  for (int k = 0; k < nn4; k++)
    if (!seen_n[k]) return fail("%s 4-wide BVH: node %d unreachable", what, k);
  return 0;
}

static int validate_bvh(const BVHNode *dnodes, int nnodes, int nprim, const char *what) {
  size_t lim = getenv("GXY_BVH_VALIDATE_MAX") ? (size_t)atol(getenv("GXY_BVH_VALIDATE_MAX")) : (size_t)(4u << 20);
  if ((size_t)nnodes > lim) return 0;
  std::vector<BVHNode> h(nnodes);
  HIPCHK(hipMemcpyAsync(h.data(), dnodes, sizeof(BVHNode) * nnodes, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  std::vector<char> seen_node(nnodes, 0), seen_prim(nprim, 0);
  std::vector<int> st;
  st.push_back(0);
  while (!st.empty()) {
    int c = st.back();
    st.pop_back();
    if (c >= 0) {
      if (c >= nnodes || seen_node[c]) return fail("%s BVH: inner node %d invalid or shared", what, c);
      seen_node[c] = 1;
      st.push_back(h[c].c0);
      if (!(nnodes == 1 && h[c].c1 == h[c].c0)) st.push_back(h[c].c1);
    } else {
      int code = ~c, start = code >> 4, cnt = (code & 15) + 1;
      if (start < 0 || start + cnt > nprim) return fail("%s BVH: leaf range %d+%d outside %d", what, start, cnt, nprim);
      for (int k = start; k < start + cnt; k++) {
        if (seen_prim[k]) return fail("%s BVH: primitive %d in two leaves", what, k);
        seen_prim[k] = 1;
      }
    }
  }
  for (int k = 0; k < nnodes; k++)
    if (!seen_node[k]) return fail("%s BVH: node %d unreachable", what, k);
  for (int k = 0; k < nprim; k++)
    if (!seen_prim[k]) return fail("%s BVH: primitive %d unreachable", what, k);
  return 0;
}

GXY_EXPORT int gxy_model_commit(void *model) {
  ModelObj *M = lookup<ModelObj>(G.models, model);
  if (!M) return fail("not a model handle");
  M->free_built();
  memset(&M->d, 0, sizeof M->d);
  // shading table indexed by geomID
  std::vector<DGeomShade> sh(std::max(1, M->ngeom));
  memset(sh.data(), 0, sh.size() * sizeof(DGeomShade));
  int64_t ntri = 0, nsph = 0;
  for (auto &m : M->meshes) {
    DGeomShade &s = sh[m.geomID];
    s.kind = 0;
    s.vertices = m.V;
    s.indices = m.I;
    s.normals = m.N;
    s.has_normals = m.N != nullptr;
    s.vdata = m.D;
    s.has_data = m.D != nullptr;
    s.has_tf = m.tf != nullptr;
    if (m.tf) s.tf = m.tf->d;
    ntri += m.nt;
  }
  int64_t ncur = 0;
  for (auto &p : M->paths) {
    DGeomShade &s = sh[p.geomID];
    s.kind = 2;
    s.vertices = p.CP;
    s.indices = p.IC;
    s.r0 = p.r0; s.r1 = p.r1; s.v0 = p.v0; s.v1 = p.v1;
    s.has_tf = p.tf != nullptr;
    if (p.tf) s.tf = p.tf->d;
    ncur += (int64_t)p.pseg.size();
  }
  for (auto &p : M->spheres) {
    DGeomShade &s = sh[p.geomID];
    s.kind = 1;
    s.centers = p.C;
    s.vdata = p.D;
    s.has_data = 1;
    s.has_tf = p.tf != nullptr;
    if (p.tf) s.tf = p.tf->d;
    nsph += p.n;
  }
  if (ntri > (1 << 27) || nsph > (1 << 27) || ncur > (1 << 27)) return fail("more than 2^27 primitives of one kind");
  DGeomShade *dsh = upload(sh.data(), sh.size());
  M->owned.push_back(dsh);
  M->d.shade = dsh;
  M->d.ngeom = M->ngeom;
  hipStream_t s = S();
  if (ntri > 0) {
    std::vector<int> fg(ntri), fp(ntri);
    float *boxes = (float *)dalloc(6 * 4 * (size_t)ntri);
    if (!boxes) return -1;
    int64_t off = 0;
    for (auto &m : M->meshes) {
      for (int p = 0; p < m.nt; p++) { fg[off + p] = m.geomID; fp[off + p] = p; }
      hipLaunchKernelGGL(k_tri_boxes, dim3((m.nt + 255) / 256), dim3(256), 0, s, m.V, m.I, m.nt, boxes + 6 * off);
      off += m.nt;
    }
    int *dfg = upload(fg.data(), fg.size()), *dfp = upload(fp.data(), fp.size());
    BVHBuildResult B;
    HIPCHK(build_lbvh(boxes, (int)ntri, G.bvh_cost[0], dalloc, dfree, &B, s));
    float *tverts = (float *)dalloc(9 * 4 * (size_t)ntri);
    int *tg = (int *)dalloc(4 * (size_t)ntri), *tp = (int *)dalloc(4 * (size_t)ntri);
    if (!tverts || !tg || !tp || !B.nodes) return fail("model commit: out of device memory");
    hipLaunchKernelGGL(k_gather_tris, dim3((ntri + 255) / 256), dim3(256), 0, s, B.order, dfg, dfp, (int)ntri, dsh,
                       tverts, tg, tp);
    HIPCHK(hipStreamSynchronize(s));
    dfree(boxes);
    dfree(dfg);
    dfree(dfp);
    dfree(B.order);
    M->owned.insert(M->owned.end(), {(void *)B.nodes, (void *)tverts, (void *)tg, (void *)tp});
    M->d.tnodes = B.nodes;
    M->d.tverts = tverts;
    M->d.tgeom = tg;
    M->d.tprim = tp;
    M->d.ntris = (int)ntri;
    M->tnodes = B.nnodes;
    M->tdepth = B.depth;
    if (B.depth > 60) return fail("triangle BVH depth %d exceeds the traversal stack", B.depth);
    if (validate_bvh(B.nodes, B.nnodes, (int)ntri, "triangle") != 0) return -1;
    M->owned.push_back(B.nodes4);
    if (validate_bvh4(B.nodes4, B.nnodes4, (int)ntri, "triangle") != 0) return -1;
    M->d.tnodes4 = B.nodes4;
    M->tdepth4 = B.depth4;
    M->tsah = B.sah;
  }
  if (nsph > 0) {
    std::vector<int> fg(nsph), fp(nsph);
    std::vector<float> feps(nsph);
    float *boxes = (float *)dalloc(6 * 4 * (size_t)nsph);
    float *frad = (float *)dalloc(4 * (size_t)nsph);
    float *fC = (float *)dalloc(12 * (size_t)nsph);
    if (!boxes || !frad || !fC) return -1;
    int64_t off = 0;
    for (auto &p : M->spheres) {
      for (int k = 0; k < p.n; k++) { fg[off + k] = p.geomID; fp[off + k] = k; feps[off + k] = p.eps; }
      hipLaunchKernelGGL(k_sph_boxes, dim3((p.n + 255) / 256), dim3(256), 0, s, p.C, p.D, p.n, p.r0, p.r1, p.v0,
                         p.v1, boxes + 6 * off, frad + off);
      hipMemcpyAsync(fC + 3 * off, p.C, 12 * (size_t)p.n, hipMemcpyDeviceToDevice, s);
      off += p.n;
    }
    int *dfg = upload(fg.data(), fg.size()), *dfp = upload(fp.data(), fp.size());
    float *dfe = upload(feps.data(), feps.size());
    BVHBuildResult B;
    HIPCHK(build_lbvh(boxes, (int)nsph, G.bvh_cost[1], dalloc, dfree, &B, s));
    float *sc = (float *)dalloc(12 * (size_t)nsph), *sr = (float *)dalloc(4 * (size_t)nsph);
    float *se = (float *)dalloc(4 * (size_t)nsph);
    int *sg = (int *)dalloc(4 * (size_t)nsph), *sp = (int *)dalloc(4 * (size_t)nsph);
    if (!sc || !sr || !se || !sg || !sp || !B.nodes) return fail("model commit: out of device memory");
    hipLaunchKernelGGL(k_gather_sph, dim3((nsph + 255) / 256), dim3(256), 0, s, B.order, dfg, dfp, fC, frad, dfe,
                       (int)nsph, sc, sr, se, sg, sp);
    HIPCHK(hipStreamSynchronize(s));
    for (void *p : {(void *)boxes, (void *)frad, (void *)fC, (void *)dfg, (void *)dfp, (void *)dfe, (void *)B.order})
      dfree(p);
    M->owned.insert(M->owned.end(), {(void *)B.nodes, (void *)sc, (void *)sr, (void *)se, (void *)sg, (void *)sp});
    M->d.snodes = B.nodes;
    M->d.scenter = sc;
    M->d.sradius = sr;
    M->d.seps = se;
    M->d.sgeom = sg;
    M->d.sprim = sp;
    M->d.nsph = (int)nsph;
    M->snodes = B.nnodes;
    M->sdepth = B.depth;
    if (B.depth > 60) return fail("sphere BVH depth %d exceeds the traversal stack", B.depth);
    if (validate_bvh(B.nodes, B.nnodes, (int)nsph, "sphere") != 0) return -1;
    M->owned.push_back(B.nodes4);
    if (validate_bvh4(B.nodes4, B.nnodes4, (int)nsph, "sphere") != 0) return -1;
    M->d.snodes4 = B.nodes4;
    M->sdepth4 = B.depth4;
    M->ssah = B.sah;
  }
  if (ncur > 0) {
    // round-cone pieces: boxes on the host, tree on the device, pieces
    // gathered into leaf order on the host
    std::vector<float> fa, fb, fu, boxes;
    std::vector<int> fg, fp;
    for (auto &p : M->paths)
      for (size_t k = 0; k < p.pseg.size(); k++) {
        const float *a = &p.pa[4 * k], *b = &p.pb[4 * k];
        for (int c = 0; c < 3; c++) boxes.push_back(std::min(a[c] - a[3], b[c] - b[3]));
        for (int c = 0; c < 3; c++) boxes.push_back(std::max(a[c] + a[3], b[c] + b[3]));
        fa.insert(fa.end(), a, a + 4);
        fb.insert(fb.end(), b, b + 4);
        fu.insert(fu.end(), {p.pu[2 * k], p.pu[2 * k + 1]});
        fg.push_back(p.geomID);
        fp.push_back(p.pseg[k]);
      }
    float *dbox = upload(boxes.data(), boxes.size());
    if (!dbox) return -1;
    BVHBuildResult B;
    HIPCHK(build_lbvh(dbox, (int)ncur, G.bvh_cost[2], dalloc, dfree, &B, s));
    if (!B.nodes) return fail("model commit: out of device memory");
    std::vector<int> order(ncur);
    const hipError_t e1 = hipMemcpyAsync(order.data(), B.order, 4 * (size_t)ncur, hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    dfree(dbox);
    dfree(B.order);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      dfree(B.nodes);
      return fail("curve BVH: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    }
    std::vector<float> sa(4 * ncur), sb(4 * ncur), su(2 * ncur);
    std::vector<int> sg(ncur), sp(ncur);
    for (int64_t k = 0; k < ncur; k++) {
      const int o = order[k];
      if (o < 0 || o >= ncur) {
        dfree(B.nodes);
        return fail("curve BVH order entry %d out of range", o);
      }
      for (int c = 0; c < 4; c++) { sa[4 * k + c] = fa[4 * (size_t)o + c]; sb[4 * k + c] = fb[4 * (size_t)o + c]; }
      su[2 * k] = fu[2 * (size_t)o];
      su[2 * k + 1] = fu[2 * (size_t)o + 1];
      sg[k] = fg[o];
      sp[k] = fp[o];
    }
    float *da = upload(sa.data(), sa.size()), *db = upload(sb.data(), sb.size()), *du = upload(su.data(), su.size());
    int *dg = upload(sg.data(), sg.size()), *dp = upload(sp.data(), sp.size());
    // owned from here on, so that the model frees them on any failure below
    M->owned.insert(M->owned.end(), {(void *)B.nodes, (void *)da, (void *)db, (void *)du, (void *)dg, (void *)dp});
    if (!da || !db || !du || !dg || !dp) return fail("model commit: out of device memory");
    M->d.cnodes = B.nodes;
    M->d.ca = da;
    M->d.cb = db;
    M->d.cu = du;
    M->d.cgeom = dg;
    M->d.cprim = dp;
    M->d.ncurve = (int)ncur;
    M->cnodes = B.nnodes;
    M->cdepth = B.depth;
    if (B.depth > 60) return fail("curve BVH depth %d exceeds the traversal stack", B.depth);
    if (validate_bvh(B.nodes, B.nnodes, (int)ncur, "curve") != 0) return -1;
    M->owned.push_back(B.nodes4);
    if (validate_bvh4(B.nodes4, B.nnodes4, (int)ncur, "curve") != 0) return -1;
    M->d.cnodes4 = B.nodes4;
    M->cdepth4 = B.depth4;
    M->csah = B.sah;
    HIPCHK(hipStreamSynchronize(s));
  }
  M->d.depth = std::max(std::max(M->tdepth, M->sdepth), M->cdepth);
  M->d.depth4 = std::max(std::max(M->tdepth4, M->sdepth4), M->cdepth4);
  M->committed = true;
  return 0;
}
GXY_EXPORT int gxy_set_bvh_prim_cost(int kind, float cost) {
  INIT_OR(-1);
  if (kind < 0 || kind > 2) return fail("BVH primitive kind %d: 0 triangles, 1 spheres, 2 path-line pieces", kind);
  if (!(cost > 0.f && cost < 1000.f)) return fail("BVH primitive cost %g out of (0, 1000)", cost);
  G.bvh_cost[kind] = cost;
  return 0;
}
GXY_EXPORT float gxy_get_bvh_prim_cost(int kind) { return (kind >= 0 && kind <= 2) ? G.bvh_cost[kind] : -1.f; }

GXY_EXPORT int gxy_model_bvh_stats(void *model, int *nodes, int *leaves, int *depth, float *sah) {
  ModelObj *M = lookup<ModelObj>(G.models, model);
  if (!M) return fail("not a model handle");
  if (nodes) *nodes = M->tnodes + M->snodes + M->cnodes;
  if (leaves) *leaves = M->d.ntris + M->d.nsph + M->d.ncurve;
  if (depth) *depth = std::max(std::max(M->tdepth, M->sdepth), M->cdepth);
  // the SAH costs of the triangle, sphere and curve trees, weighted by their primitives
  if (sah) {
    const double np = (double)M->d.ntris + M->d.nsph + M->d.ncurve;
    *sah = np > 0 ? (float)((M->tsah * (double)M->d.ntris + M->ssah * (double)M->d.nsph +
                             M->csah * (double)M->d.ncurve) / np)
                  : 0.f;
  }
  return 0;
}
GXY_EXPORT int gxy_model_set_hit_output(void *model, int *g, int *p) {
  ModelObj *M = lookup<ModelObj>(G.models, model);
  if (!M) return fail("not a model handle");
  M->hit_geom = g;
  M->hit_prim = p;
  return 0;
}

// Trace (or sampler) launch; its volume samples are added to the context's
// sample counter (counters[0]) by the kernel's waves.
static int trace_launch(TraceParams P, const RayList_ispc &R, int n, bool sampler) {
  if (n <= 0) return 0;
  const bool pl = !sampler && G.trace_kind == 1 && P.nvv >= 1 && !P.has_model;
  P.sample_count = CX().counters;
  KTimer kt(sampler ? "sampler" : "trace");
  if (pl) HIPCHK(launch_trace_pl(P, R, n, G.trace_chunk, G.trace_refill, S()));
  else HIPCHK(sampler ? launch_sampler(P, R, n, S())
                      : launch_trace(P, R, n, G.trace_wpe, G.geom_stack, S(), G.geom_waves, G.trace_chunk));
  return 0;
}

// ------------------------------------------------------------------ TraceRays_* (reference ABI)
struct TraceRaysObj {
  int debug[100];
};

GXY_EXPORT void *TraceRays_allocate(void) {
  if (ensure_init() != 0) return nullptr;
  TraceRaysObj *t = new TraceRaysObj;
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.tracers.insert(t);
  return t;
}
GXY_EXPORT void TraceRays_initialize(void *self) {
  TraceRaysObj *t = lookup<TraceRaysObj>(G.tracers, self);
  if (t) for (int i = 0; i < 100; i++) t->debug[i] = -1;
}
GXY_EXPORT void TraceRays_destroy(void *self) {
  TraceRaysObj *t = lookup<TraceRaysObj>(G.tracers, self);
  if (!t) return;
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.tracers.erase(t);
  delete t;
}

// The reference's entry points return nothing; their errors are reported by
// gxy_last_error(), which each of them clears first, so that an earlier
// failure elsewhere is not taken for theirs.
GXY_EXPORT void *TraceRays_TraceRays(void *self, void *vis, int nRaysIn, void *raysIn, float epsilon) {
  (void)self;
  t_err.clear();
  if (ensure_init() != 0) return nullptr;
  TraceParams P;
  if (make_trace_params((const Visualization_ispc *)vis, epsilon, P) != 0) return nullptr;
  trace_launch(P, *(const RayList_ispc *)raysIn, nRaysIn, false);
  return nullptr;
}

static void lighting_entry(void *lights, int n, void *raysIn, int *off, void *raysOut, float eps, int mode,
                           const char *name) {
  t_err.clear();
  if (ensure_init() != 0) return;
  LightParams L;
  if (make_light_params((const Lighting_ispc *)lights, eps, L) != 0) return;
  RayList_ispc O;
  memset(&O, 0, sizeof O);
  if (raysOut) O = *(const RayList_ispc *)raysOut;
  KTimer kt(name);
  hipError_t e = launch_lighting(L, mode, *(const RayList_ispc *)raysIn, n, (mode & LM_AO) ? off : nullptr,
                                 (mode & LM_SHADOW) ? off : nullptr, O, S());
  if (e != hipSuccess) fail("lighting launch: %s", hipGetErrorString(e));
}
GXY_EXPORT void TraceRays_generateAORays(void *self, void *lights, int n, void *raysIn, int *offsets, void *raysOut,
                                         float eps) {
  (void)self;
  lighting_entry(lights, n, raysIn, offsets, raysOut, eps, LM_AO, "ao");
}
GXY_EXPORT void TraceRays_generateShadowRays(void *self, void *lights, int n, void *raysIn, int *offsets,
                                             void *raysOut, float eps) {
  (void)self;
  lighting_entry(lights, n, raysIn, offsets, raysOut, eps, LM_SHADOW, "shadow");
}
GXY_EXPORT void TraceRays_ambientLighting(void *self, void *lights, int n, void *raysIn) {
  (void)self;
  lighting_entry(lights, n, raysIn, nullptr, nullptr, 0.f, LM_AMBIENT, "ambient");
}
GXY_EXPORT void TraceRays_diffuseLighting(void *self, void *lights, int n, void *raysIn) {
  (void)self;
  lighting_entry(lights, n, raysIn, nullptr, nullptr, 0.f, LM_DIFFUSE, "diffuse");
}

GXY_EXPORT void *SamplerTraceRays_SamplerTraceRays(void *self, void *vis, int n, void *raysIn) {
  (void)self;
  t_err.clear();
  if (ensure_init() != 0) return nullptr;
  const Visualization_ispc *V = (const Visualization_ispc *)vis;
  if (!V || V->nVolumeVis < 1) return nullptr;  // SamplerTraceRays.ispc:139
  TraceParams P;
  if (make_trace_params(V, 0.f, P) != 0) return nullptr;
  // SamplerTraceRays.ispc:139-150: step = min samplingStep*samplingRate
  for (int k = 0; k < P.nvv; k++)
    if (P.vv[k].nIso == 0) P.vv[k].iso[0] = 0.f;  // IsoSamplerVis_initialize default
  trace_launch(P, *(const RayList_ispc *)raysIn, n, true);
  return nullptr;
}

// ------------------------------------------------------------------ Lighting_* (Lighting.ispc:23-134)
struct LightingObj {
  Lighting_ispc l;
  std::vector<float> L;
  std::vector<int> T;
};
GXY_EXPORT void *Lighting_allocate(void) {
  LightingObj *o = new LightingObj;
  memset(&o->l, 0, sizeof o->l);
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.lightings.insert(o);
  return &o->l;  // Lighting_ispc is the first member
}
static LightingObj *lobj(void *self) {
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  return G.lightings.count(self) ? (LightingObj *)self : nullptr;
}
GXY_EXPORT void Lighting_initialize(void *self) {
  LightingObj *o = lobj(self);
  if (!o) return;
  o->l.n_ao_rays = 0;
  o->l.do_shadows = 0;
  o->l.ao_radius = 1.0f;
  o->l.Ka = 0.5f;
  o->l.Kd = 0.5f;
  o->L = {1.f, 1.f, 1.f};
  o->T = {2};
  o->l.lights = o->L.data();
  o->l.types = o->T.data();
  o->l.nLights = 1;
}
GXY_EXPORT void Lighting_destroy(void *self) {
  LightingObj *o = lobj(self);
  if (!o) return;
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  G.lightings.erase(o);
  delete o;
}
GXY_EXPORT void Lighting_SetLights(void *self, int n, float *l, int *t) {
  LightingObj *o = lobj(self);
  if (!o) return;
  o->L.assign(l, l + 3 * n);
  if (t) o->T.assign(t, t + n);
  else if ((int)o->T.size() != n) o->T.assign(n, 0);
  o->l.lights = o->L.data();
  o->l.types = o->T.data();
  o->l.nLights = n;
}
GXY_EXPORT void Lighting_GetLights(void *self, int *n, float **l, int **t) {
  LightingObj *o = lobj(self);
  if (!o) return;
  *n = (int)o->l.nLights;
  *l = o->l.lights;
  *t = o->l.types;
}
GXY_EXPORT void Lighting_SetK(void *self, float ka, float kd) {
  LightingObj *o = lobj(self);
  if (o) { o->l.Ka = ka; o->l.Kd = kd; }
}
GXY_EXPORT void Lighting_GetK(void *self, float *ka, float *kd) {
  LightingObj *o = lobj(self);
  if (o) { *ka = o->l.Ka; *kd = o->l.Kd; }
}
GXY_EXPORT void Lighting_SetAO(void *self, int n, float r) {
  LightingObj *o = lobj(self);
  if (o) { o->l.n_ao_rays = n; o->l.ao_radius = r; }
}
GXY_EXPORT void Lighting_GetAO(void *self, int *n, float *r) {
  LightingObj *o = lobj(self);
  if (o) { *n = o->l.n_ao_rays; *r = o->l.ao_radius; }
}
GXY_EXPORT void Lighting_SetShadowFlag(void *self, int b) {
  LightingObj *o = lobj(self);
  if (o) o->l.do_shadows = b;
}
GXY_EXPORT void Lighting_GetShadowFlag(void *self, int *b) {
  LightingObj *o = lobj(self);
  if (o) *b = o->l.do_shadows;
}

// ------------------------------------------------------------------ pass pieces
namespace gxy {
struct Scratch {
  std::vector<void *> blocks;
  void *get(size_t bytes) {
    void *p = dalloc(bytes);
    if (p) blocks.push_back(p);
    return p;
  }
  ~Scratch() {
    for (void *p : blocks) dfree_stream_ordered(p);
  }
};

struct OwnedList {
  RayListObj *p = nullptr;
  RayListObj *release() {
    RayListObj *r = p;
    p = nullptr;
    return r;
  }
  ~OwnedList() { raylist_delete(p, false); }
};

// TraceRays.cpp:89-117 on the device; n_ao / n_shadow through one readback
static int spawn_offsets(int n, const RayList_ispc &R, int nAO, int nL, int doSh, int *ao_off, int *sh_off,
                         int *T_out) {
  Scratch sc;
  int *scratch = (int *)sc.get(surface_scan_scratch_bytes(n));
  int *rank = (int *)sc.get(sizeof(int) * std::max(1, n));
  if (!scratch || !rank) return -1;
  int *total = surface_scan_total(scratch, n);
  {
    KTimer kt("spawn_scan");
    HIPCHK(launch_surface_scan(R, n, scratch, rank, nullptr, S()));
    HIPCHK(launch_offsets_from_rank(R, n, rank, total, nAO, nL, doSh, ao_off, sh_off, S()));
  }
  Readback rb;
  if (readback(nullptr, 0, total, nullptr, rb) != 0) return -1;
  *T_out = rb.one;
  return 0;
}
}  // namespace gxy

GXY_EXPORT int gxy_spawn_offsets(int n, RayList_ispc *R, int nAO, int nL, int doSh, int *ao_off, int *sh_off,
                                 int *n_ao, int *n_sh) {
  INIT_OR(-1);
  int T = 0;
  if (n < 0) return fail("negative ray count");
  if (n == 0) {
    *n_ao = *n_sh = 0;
    return 0;
  }
  if (spawn_offsets(n, *R, nAO, nL, doSh, ao_off, sh_off, &T) != 0) return -1;
  *n_ao = nAO * T;
  *n_sh = doSh ? nL * T : 0;
  return 0;
}

// gxy_spawn_offsets without any allocation: the scan's temporaries are the
// static scratch array of spawn.hip, by the device address taken at set-up;
// the parent total comes back through the context's readback area (allocated
// when the library or the context was set up).  Calls that overlap on one GPU
// share that array and race.
GXY_EXPORT int gxy_spawn_offsets_noalloc(int n, RayList_ispc *R, int nAO, int nL, int doSh, int *ao_off,
                                         int *sh_off, int *n_ao, int *n_sh) {
  INIT_OR(-1);
  if (n < 0) return fail("negative ray count");
  if (n > GXY_NOALLOC_MAX_RAYS)
    return fail("gxy_spawn_offsets_noalloc: %d rays exceed the static scratch limit of %d rays", n,
                (int)GXY_NOALLOC_MAX_RAYS);
  int *scr = (G.device >= 0 && G.device < 64) ? G.noalloc_scratch[G.device] : nullptr;
  if (!scr) return fail("gxy_spawn_offsets_noalloc: no device address for the static scratch array");
  if (n == 0) {
    *n_ao = *n_sh = 0;
    return 0;
  }
  const int *total = nullptr;
  HIPCHK(launch_spawn_offsets_noalloc(*R, n, scr, nAO, nL, doSh, ao_off, sh_off, &total, S()));
  Readback rb;
  if (readback(nullptr, 0, total, nullptr, rb) != 0) return -1;
  *n_ao = nAO * rb.one;
  *n_sh = doSh ? nL * rb.one : 0;
  return 0;
}

static RouteParams route_params(int mode, int reverse, const gxy_box3f *box, const int nb[6], float *fb, int w, int h,
                                int unique) {
  RouteParams P;
  memset(&P, 0, sizeof P);
  P.unique = unique;
  P.reverse = reverse;
  for (int k = 0; k < 6; k++) P.neighbors[k] = nb ? nb[k] : -1;
  if (box)
    for (int k = 0; k < 3; k++) { P.blo[k] = box->lower[k]; P.bhi[k] = box->upper[k]; }
  P.fb = fb;
  P.width = w;
  P.height = h;
  P.mode = mode;
  return P;
}
static int route(int n, const RayList_ispc &R, int mode, int reverse, const gxy_box3f *box, const int nb[6],
                 float *fb, int w, int h, int unique = 0) {
  const RouteParams P = route_params(mode, reverse, box, nb, fb, w, h, unique);
  KTimer kt("route");
  HIPCHK(launch_route(P, R, n, S()));
  return 0;
}

// The lighting model is one library setting (gxy_set_reverse_lighting), as
// Galaxy's is one build option: a flag that disagrees with it is an error,
// since reverse-signed children classified by the direct rules (or the
// converse) give results no Galaxy build produces.
static int check_reverse(int reverse, const char *what) {
  if ((reverse != 0) != (G.reverse_lighting != 0))
    return fail("%s: reverse_lighting=%d but the library's lighting model is %s (gxy_set_reverse_lighting)", what,
                reverse, G.reverse_lighting ? "reverse" : "direct");
  return 0;
}
GXY_EXPORT int gxy_set_reverse_lighting(int on) {
  INIT_OR(-1);
  G.reverse_lighting = on ? 1 : 0;
  return 0;
}
GXY_EXPORT int gxy_get_reverse_lighting(void) { return G.reverse_lighting; }
GXY_EXPORT int gxy_set_ao_directions(int mode, unsigned seed) {
  INIT_OR(-1);
  if (mode != 0 && mode != 1) return fail("AO directions %d: 0 the UV table, 1 PCG32 streams", mode);
  G.ao_pcg = mode;
  G.ao_seed = seed;
  return 0;
}
GXY_EXPORT int gxy_get_ao_directions(unsigned *seed) {
  if (seed) *seed = G.ao_seed;
  return G.ao_pcg;
}

GXY_EXPORT int gxy_classify(int n, RayList_ispc *R, int reverse) {
  INIT_OR(-1);
  if (check_reverse(reverse, "gxy_classify") != 0) return -1;
  return route(n, *R, RM_CLASSIFY, reverse, nullptr, nullptr, nullptr, 0, 0);
}
GXY_EXPORT int gxy_assign_destinations(int n, RayList_ispc *R, const gxy_box3f *box, const int nb[6], int reverse) {
  INIT_OR(-1);
  if (check_reverse(reverse, "gxy_assign_destinations") != 0) return -1;
  return route(n, *R, RM_ASSIGN, reverse, box, nb, nullptr, 0, 0);
}
GXY_EXPORT int gxy_accumulate_terminated(int n, RayList_ispc *R, float *fb, int w, int h) {
  INIT_OR(-1);
  return route(n, *R, RM_ACCUMULATE, 1, nullptr, nullptr, fb, w, h);
}

// SendPixelsMsg (Renderer.h:246-305) from HandleTerminatedRays' remote path
// (Renderer.cpp:490-499): count, check the capacity, then header + pixels
GXY_EXPORT int gxy_pixels_pack(int n, RayList_ispc *R, int64_t rkey, int64_t rskey, int frame, int source,
                               void *msg, size_t capacity_bytes, int *count_out) {
  INIT_OR(-1);
  if (n < 0 || !R) return fail("gxy_pixels_pack: %d rays", n);
  if (!msg || ((uintptr_t)msg & 7)) return fail("gxy_pixels_pack: message buffer null or not 8-byte aligned");
  if (capacity_bytes < sizeof(gxy_pixels_header)) return fail("gxy_pixels_pack: %zu bytes hold no header", capacity_bytes);
  Scratch sc;
  const int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  int *scratch = (int *)sc.get(block_scan_scratch_bytes(nb));
  int *total = (int *)sc.get(sizeof(int));
  if (!scratch || !total) return -1;
  {
    KTimer kt("pixels");
    HIPCHK(launch_pixels_count(*R, n, scratch, total, S()));
  }
  Readback rb;
  if (readback(nullptr, 0, total, nullptr, rb) != 0) return -1;
  const int count = rb.one;
  const size_t need = sizeof(gxy_pixels_header) + sizeof(gxy_pixel) * (size_t)count;
  if (need > capacity_bytes) return fail("gxy_pixels_pack: %d pixels need %zu bytes, capacity %zu", count, need,
                                         capacity_bytes);
  gxy_pixels_header h;
  memset(&h, 0, sizeof h);
  h.rkey = rkey;
  h.rskey = rskey;
  h.count = count;
  h.frame = frame;
  h.source = source;
  {
    KTimer kt("pixels");
    HIPCHK(launch_pixels_write(*R, n, scratch, h, msg, S()));
  }
  HIPCHK(hipStreamSynchronize(S()));
  if (count_out) *count_out = count;
  return 0;
}

// SendPixelsMsg::Action + Rendering::AddLocalPixels (Renderer.h:313-336,
// Rendering.cpp:126-153): the header is read from the device message
GXY_EXPORT int gxy_pixels_apply(const void *msg, size_t bytes, float *fb, int w, int h, int *kbuf, int *fb_frame,
                                int *applied) {
  INIT_OR(-1);
  if (applied) *applied = 0;
  if (!msg || ((uintptr_t)msg & 7)) return fail("gxy_pixels_apply: message null or not 8-byte aligned");
  if (bytes < sizeof(gxy_pixels_header)) return fail("gxy_pixels_apply: %zu bytes hold no header", bytes);
  if (!fb || !fb_frame || w <= 0 || h <= 0) return fail("gxy_pixels_apply: framebuffer %dx%d", w, h);
  Readback rb;
  // count, frame, source: the header's ints 4..6
  if (readback((const int *)msg + 4, 3, nullptr, nullptr, rb) != 0) return -1;
  const int count = rb.ints[0], frame = rb.ints[1];
  if (count < 0 || sizeof(gxy_pixels_header) + sizeof(gxy_pixel) * (size_t)count > bytes)
    return fail("gxy_pixels_apply: header count %d does not fit %zu bytes", count, bytes);
  if (frame < *fb_frame) return 0;   // an older frame's pixels are dropped
  *fb_frame = frame;
  Scratch sc;
  int *inside = (int *)sc.get(sizeof(int));
  if (!inside) return -1;
  {
    KTimer kt("pixels");
    HIPCHK(launch_pixels_apply(msg, count, frame, fb, w, h, kbuf, inside, S()));
  }
  if (readback(nullptr, 0, inside, nullptr, rb) != 0) return -1;
  if (applied) *applied = rb.one;
  return 0;
}

GXY_EXPORT int gxy_partition_rays(int n, RayList_ispc *R, int ndest, RayList_ispc *const *lists,
                                  RayList_ispc *keepers, int *counts) {
  INIT_OR(-1);
  if (ndest < 0 || ndest + 4 > 64) return fail("ndest %d outside [0, 60]", ndest);
  int K = ndest + 1;
  std::vector<int> table(ndest + 4, -1);
  table[GXY_KEEP_HERE + 4] = ndest;
  for (int d = 0; d < ndest; d++) table[d + 4] = d;
  Scratch sc;
  int *scratch = (int *)sc.get(partition_scratch_bytes(n, K));
  int *dcounts = (int *)sc.get(sizeof(int) * K);
  if (!scratch || !dcounts) return -1;
  {
    KTimer kt("partition");
    HIPCHK(launch_partition(*R, n, table.data(), (int)table.size(), K, nullptr, scratch, dcounts, S()));
  }
  HIPCHK(hipMemcpyAsync(counts, dcounts, sizeof(int) * K, hipMemcpyDeviceToHost, S()));
  bool any = keepers != nullptr;
  for (int d = 0; d < ndest && lists; d++) any |= lists[d] != nullptr;
  if (any) {
    if (K > 12) return fail("list copies support at most 11 destinations");
    std::vector<RayList_ispc> L(K);
    memset(L.data(), 0, sizeof(RayList_ispc) * K);
    for (int d = 0; d < ndest; d++)
      if (lists && lists[d]) L[d] = *lists[d];
    if (keepers) L[ndest] = *keepers;
    KTimer kt("partition");
    HIPCHK(launch_partition_scatter(*R, n, table.data(), (int)table.size(), K, L.data(), scratch, S()));
  }
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}

// Camera::generate_initial_rays host part (Camera.cpp:528-582): image-plane
// frame in float with the dtypes.h normalise; per-pixel rays on the device.
static void dt_norm(float *a) {
  float d = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (d != 0) {
    d = (float)(1.0 / d);
    a[0] *= d; a[1] *= d; a[2] *= d;
  }
}
static void crossf(const float *a, const float *b, float *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static void camera_params(const float eye[3], const float cdir[3], int is_dir, const float up[3], float aov, int w,
                          int h, const gxy_box3f *lb, const gxy_box3f *gb, int order, CameraParams &C) {
  memset(&C, 0, sizeof C);
  float vdir[3], vu[3] = {up[0], up[1], up[2]}, center[3];
  for (int k = 0; k < 3; k++) vdir[k] = is_dir ? cdir[k] : cdir[k] - eye[k];
  if (aov == 0.0f) {
    for (int k = 0; k < 3; k++) center[k] = eye[k] + vdir[k];
    dt_norm(vdir);
  } else {
    float d = (float)(1.0 / std::tan(2 * 3.1415926 * (aov / 2.0) / 360.0));
    dt_norm(vdir);
    for (int k = 0; k < 3; k++) center[k] = eye[k] + d * vdir[k];
  }
  float vr[3];
  crossf(vdir, vu, vr);
  dt_norm(vr);
  crossf(vr, vdir, vu);
  dt_norm(vu);
  float pixel_scaling = (float)(((w < h ? w : h) - 1.0) / 2.0);
  for (int k = 0; k < 3; k++) {
    C.veye[k] = eye[k];
    C.vdir[k] = vdir[k];
    C.vr[k] = vr[k];
    C.vu[k] = vu[k];
    C.center[k] = center[k];
    C.llo[k] = lb->lower[k];
    C.lhi[k] = lb->upper[k];
    C.glo[k] = gb->lower[k];
    C.ghi[k] = gb->upper[k];
  }
  C.scaling = (float)(1.0 / pixel_scaling);
  C.off_x = (float)((w - 1) / 2.0);
  C.off_y = (float)((h - 1) / 2.0);
  C.is_ortho = aov == 0.0f;
  C.width = w;
  C.height = h;
  C.order = order;
}

// counter_out (may be null): the context's sample counter as of the camera
// rays, read in the same wait
// between(total): work queued after the readback of the ray count and before
// the host waits for it, given the count's device address (may be null)
static int camera_generate(const float eye[3], const float cdir[3], int is_dir, const float up[3], float aov,
                           int w, int h, const gxy_box3f *lb, const gxy_box3f *gb, const RayList_ispc &R,
                           int capacity, int order, int *n_out, unsigned long long *counter_out = nullptr,
                           const std::function<int(const int *)> &between = nullptr) {
  CameraParams C;
  camera_params(eye, cdir, is_dir, up, aov, w, h, lb, gb, order, C);
  int npix = order == 0 ? w * h : ((w + 7) / 8) * ((h + 7) / 8) * 64;
  Scratch sc;
  int nb = (npix + GXY_BLOCK - 1) / GXY_BLOCK;
  int *scratch = (int *)sc.get(block_scan_scratch_bytes(nb));
  if (!scratch) return -1;
  int *total = block_scan_offsets(scratch, nb) + nb;
  {
    KTimer kt("camera");
    HIPCHK(launch_camera(C, R, capacity, scratch, nullptr, S()));
  }
  Readback rb;
  int seq = 0;
  if (readback_start(nullptr, 0, total, counter_out ? CX().counters : nullptr, seq) != 0) return -1;
  if (between && between(total) != 0) return -1;
  if (readback_wait(seq, 0, true, counter_out != nullptr, rb) != 0) return -1;
  const int T = rb.one;
  if (counter_out) *counter_out = rb.u64;
  if (T > capacity) return fail("camera produced %d rays for a list of %d", T, capacity);
  *n_out = T;
  return 0;
}

GXY_EXPORT int gxy_camera_generate(const float eye[3], const float cdir[3], int is_dir, const float up[3], float aov,
                                   int w, int h, const gxy_box3f *lb, const gxy_box3f *gb, RayList_ispc *R,
                                   int capacity, int order, int *n_out) {
  INIT_OR(-1);
  return camera_generate(eye, cdir, is_dir, up, aov, w, h, lb, gb, *R, capacity, order, n_out);
}

GXY_EXPORT int gxy_sampler_collect(int n, RayList_ispc *R, float *xyz, int capacity, int *count) {
  INIT_OR(-1);
  Scratch sc;
  int nb = (n + GXY_BLOCK - 1) / GXY_BLOCK;
  int *scratch = (int *)sc.get(block_scan_scratch_bytes(nb));
  if (!scratch) return -1;
  KTimer kt("sampler_collect");
  HIPCHK(launch_accumulate_sampler(*R, n, xyz, capacity, count, scratch, S()));
  return 0;
}

// ------------------------------------------------------------------ ray-list loop
namespace gxy {
struct QItem {
  RayListObj *rl;
  int prio;
};

struct LoopCtx {
  TraceParams P;
  gxy_box3f box;
  LightParams L;
  bool have_lights = false;
  int nb[6];
  float eps;
  int maxlist;
  float *fb;
  int w, h;
  gxy_frame_stats st;
  std::vector<QItem> q;
  bool sampler = false;
  float *xyz = nullptr;
  int capacity = 0;
  int *dcount = nullptr;
  unsigned long long samples_read = 0;   // the sample counter as of the last host wait
  bool have_base = false;                // the counter was read at the loop's start (not cleared)
  unsigned long long samples_base = 0;
  bool have_samples = false;
};

// RayQManager + ThreadPool order (RayQManager.cpp:138-194, Threading.cpp:176-233):
// the highest priority first (primary lists 3, secondary 2), LIFO within it.
static QItem pop_next(std::vector<QItem> &q) {
  int best = -1;
  for (int i = 0; i < (int)q.size(); i++)
    if (best < 0 || q[i].prio >= q[best].prio) best = i;
  QItem it = q[best];
  q.erase(q.begin() + best);
  return it;
}

// Renderer::Trace's split (Renderer.cpp:675-690) + enqueue
static int enqueue_split(LoopCtx &C, RayListObj *out, int prio) {
  int n = out->hdr.size;
  if (n <= C.maxlist) {
    C.q.push_back({out, prio});
    return 0;
  }
  for (int s = 0; s < n; s += C.maxlist) {
    int c = std::min(C.maxlist, n - s);
    RayListObj *p = raylist_new(c, out->hdr.frame, out->hdr.type, false);
    if (!p) return -1;
    p->kinds = out->kinds;
    p->unique = out->unique;
    HIPCHK(launch_copy_rays(out->ispc, s, c, p->ispc, 0, S()));
    C.q.push_back({p, prio});
  }
  raylist_delete(out, false);
  return 0;
}

// processRays_task::work (Renderer.cpp:511-651) for one list and its keepers
// Queues the trace of list L before its size is known: the launch covers its
// capacity ncap and the kernel reads the size, *dn * mul, from the device.
// Only the one-ray-per-item kernels read it, so only they are queued ahead;
// and not when the model's per-ray hit side-band is on (a list traced ahead
// would write it out of the loop's list order).
static bool can_pretrace(const LoopCtx &C, long long ncap) {
  return !C.sampler && G.trace_kind != 1 && !C.P.hit_geom && ncap > 0 && ncap <= C.maxlist;
}
static int pretrace(LoopCtx &C, RayListObj *L, int ncap, const int *dn, int mul) {
  TraceParams P = C.P;
  P.dn = dn;
  P.dn_mul = mul;
  if (trace_launch(P, L->ispc, ncap, false) != 0) return -1;
  L->pretraced = true;
  return 0;
}

static int process_list(LoopCtx &C, RayListObj *L, int prio, bool owned) {
  while (L && L->hdr.size > 0) {
    const int n = L->hdr.size;
    const RayList_ispc &R = L->ispc;
    if (L->pretraced) L->pretraced = false;
    else if (trace_launch(C.P, R, n, C.sampler) != 0) return -1;
    C.st.ray_passes += n;
    C.st.passes++;
    C.st.lists++;
    Scratch spawn_sc;
    int *scr_spawn = nullptr, *dT = nullptr, nsp = 0;
    OwnedList out;   // the children, freed on an early return
    // a list without primary rays has no parents: no spawn stage
    if (!C.sampler && L->kinds != 2) {
      // TraceRays::Trace (TraceRays.cpp:68-146): the parents (primary surface
      // rays) ranked by one flag scan; the AO block then the shadow block of
      // their children, written one work-item per child
      // the parent count stays on the device until the partition counts are
      // read (one host wait per list): the children list is allocated for
      // every ray being a parent and its size set afterwards
      scr_spawn = (int *)spawn_sc.get(surface_scan_scratch_bytes(n));
      int *pidx = (int *)spawn_sc.get(sizeof(int) * std::max(1, n));
      if (!scr_spawn || !pidx) return -1;
      int *total = surface_scan_total(scr_spawn, n);
      {
        KTimer kt("spawn_scan");
        HIPCHK(launch_surface_scan(R, n, scr_spawn, nullptr, nullptr, S(), pidx));
      }
      nsp = C.L.nAO + (C.L.doShadows ? C.L.nLights : 0);
      if ((long long)n * nsp > INT32_MAX - 16) return fail("%lld secondary rays exceed one RayList", (long long)n * nsp);
      if (nsp > 0) {
        out.p = raylist_new(n * nsp, L->hdr.frame, 1, false);
        if (!out.p) return -1;
        // AO and shadow rays only; one child per parent keeps the parents' pixels unique
        out.p->kinds = 2;
        out.p->unique = (L->unique && nsp == 1) ? 1 : 0;
      }
      // TraceRays.cpp:124-131 (reverse lighting: every step) or :133-141 (the
      // cast rays carry ambient / diffuse instead of the parent)
      const bool ao = C.L.nAO > 0, sh = C.L.doShadows && C.L.nLights > 0;
      const int mode = C.L.reverse ? (LM_AMBIENT | LM_DIFFUSE | (ao ? LM_AO : 0) | (sh ? LM_SHADOW : 0))
                                   : ((ao ? LM_AO : LM_AMBIENT) | (sh ? LM_SHADOW : LM_DIFFUSE));
      RayList_ispc O;
      memset(&O, 0, sizeof O);
      if (out.p) O = out.p->ispc;
      {
        KTimer kt("lighting");
        HIPCHK(launch_lighting_children_dev(C.L, mode, R, pidx, total, n, O, S()));
      }
      dT = total;
    }
    // split by destination: keepers + one list per neighbour face
    // (Renderer.cpp:561-634); terminated and dropped rays are counted only
    int table[64];
    for (int k = 0; k < 64; k++) table[k] = -1;
    table[GXY_KEEP_HERE + 4] = 0;
    table[GXY_TERMINATED + 4] = 7;
    table[GXY_DROP_ON_FLOOR + 4] = 8;
    int ntable = 4;
    for (int f = 0; f < 6; f++)
      if (C.nb[f] >= 0 && C.nb[f] + 4 < 64) {
        table[C.nb[f] + 4] = 1 + f;
        ntable = std::max(ntable, C.nb[f] + 5);
      }
    const int K = 9;
    Scratch sc;
    int *scr = (int *)sc.get(partition_scratch_bytes(n, K));
    int *dcounts = (int *)sc.get(sizeof(int) * K);
    if (!scr || !dcounts) return -1;
    // Classify, AssignDestinations, HandleTerminatedRays (Renderer.cpp:550-556),
    // with the partition's count phase in the same pass over the rays
    {
      const RouteParams RP =
          route_params(C.sampler ? (RM_CLASSIFY | RM_ASSIGN) : (RM_CLASSIFY | RM_ASSIGN | RM_ACCUMULATE),
                       G.reverse_lighting, &C.box, C.nb, C.fb, C.w, C.h, L->unique);
      KTimer kt("route");
      HIPCHK(launch_route_count(RP, R, n, table, ntable, K, scr, S()));
    }
    if (C.sampler) {
      // Sampler::HandleTerminatedRays (Sampler.cpp:51-89)
      Scratch ssc;
      int nbk = (n + GXY_BLOCK - 1) / GXY_BLOCK;
      int *sscr = (int *)ssc.get(block_scan_scratch_bytes(nbk));
      if (!sscr) return -1;
      KTimer kt("sampler_collect");
      HIPCHK(launch_accumulate_sampler(R, n, C.xyz, C.capacity, C.dcount, sscr, S()));
    }
    int counts[K];
    {
      KTimer kt("partition");
      HIPCHK(launch_partition(R, n, table, ntable, K, nullptr, scr, dcounts, S(), true));
    }
    // one wait per list: the partition counts, the parent count and the
    // sample counter (after the loop's last list no trace follows).  The
    // children's trace is queued behind the readback when they will form one
    // list (the loop traces them next, or after lists whose rays they do not
    // touch), so the GPU traces while the host waits and decides.
    Readback rb;
    int seq = 0;
    if (readback_start(dcounts, K, dT, CX().counters, seq) != 0) return -1;
    if (out.p && dT && can_pretrace(C, (long long)n * nsp) && pretrace(C, out.p, n * nsp, dT, nsp) != 0) return -1;
    if (readback_wait(seq, K, dT != nullptr, true, rb) != 0) return -1;
    for (int k = 0; k < K; k++) counts[k] = rb.ints[k];
    const int T = rb.one;
    C.samples_read = rb.u64;
    C.have_samples = true;
    if (out.p && T > 0) {
      // the children of the T parents (TraceRays.cpp:89-117 layout); the
      // arrays stay laid out for the capacity (hdr.aligned_size)
      RayListObj *o = out.release();
      o->hdr.size = T * nsp;
      C.st.secondary_rays += (long long)T * nsp;
      if (enqueue_split(C, o, 2) != 0) return -1;
    }
    C.st.terminated += counts[7];
    C.st.dropped += counts[8];
    RayListObj *keep = nullptr;
    // categories 7 and 8 (terminated, dropped) are counted, not copied
    RayList_ispc lists[K];
    memset(lists, 0, sizeof lists);
    bool any = false;
    if (counts[0] > 0) {
      keep = raylist_new(counts[0], L->hdr.frame, L->hdr.type, false);
      if (!keep) return -1;
      keep->kinds = L->kinds;
      keep->unique = L->unique;
      lists[0] = keep->ispc;
      C.st.kept += counts[0];
      any = true;
    }
    // outbound queue: one list per face with a neighbour (SendRays,
    // Renderer.cpp:620-634); kept for gxy_outbound_take
    for (int f = 0; f < 6; f++) {
      C.st.outbound[f] += counts[1 + f];
      if (counts[1 + f] > 0) {
        RayListObj *ob = raylist_new(counts[1 + f], L->hdr.frame, L->hdr.type);
        if (!ob) return -1;
        lists[1 + f] = ob->ispc;
        std::lock_guard<std::recursive_mutex> lk(G.mu);
        G.outbound[f].push_back(ob);
        any = true;
      }
    }
    if (any) {
      KTimer kt("partition");
      HIPCHK(launch_partition_scatter(R, n, table, ntable, K, lists, scr, S()));
    }
    if (owned) raylist_delete(L, false);
    L = keep;
    owned = true;
  }
  if (L && owned) raylist_delete(L, false);
  (void)prio;
  return 0;
}

static int run_loop(LoopCtx &C, int n, const RayList_ispc &rays, RayListObj *adopt = nullptr) {
  if (!C.have_base) HIPCHK(hipMemsetAsync(CX().counters, 0, 8 * sizeof(unsigned long long), S()));
  if (adopt && n <= C.maxlist) {
    // the caller's list is processed in place (it is the loop's to consume)
    C.q.push_back({adopt, 3});
  } else {
    // initial lists, at most maxlist rays each (primary priority)
    for (int s = 0; s < n; s += C.maxlist) {
      int c = std::min(C.maxlist, n - s);
      RayListObj *p = raylist_new(c, 0, 0, false);
      if (!p) return -1;
      HIPCHK(launch_copy_rays(rays, s, c, p->ispc, 0, S()));
      C.q.push_back({p, 3});
    }
    if (adopt) raylist_delete(adopt, false);
  }
  while (!C.q.empty()) {
    QItem it = pop_next(C.q);
    if (process_list(C, it.rl, it.prio, true) != 0) {
      for (auto &x : C.q) raylist_delete(x.rl);
      C.q.clear();
      return -1;
    }
  }
  if (!C.have_samples) {
    Readback rb;
    if (readback(nullptr, 0, nullptr, CX().counters, rb) != 0) return -1;
    C.samples_read = rb.u64;
  }
  C.st.samples = (long long)(C.samples_read - (C.have_base ? C.samples_base : 0));
  return 0;
}

static int loop_setup(LoopCtx &C, void *vis, void *lights, const int nb[6], float eps, int maxlist) {
  memset(&C.st, 0, sizeof C.st);
  if (make_trace_params((const Visualization_ispc *)vis, eps, C.P) != 0) return -1;
  // the loop reads only the classification of shadow and AO rays, never their t
  C.P.occlusion = (G.occlusion && C.P.nvv == 0) ? 1 : 0;
  C.box = ((const Visualization_ispc *)vis)->local_bb;
  if (lights) {
    if (make_light_params((const Lighting_ispc *)lights, 0.001f, C.L) != 0) return -1;
    C.L.eps = eps;
    C.have_lights = true;
  }
  for (int k = 0; k < 6; k++) C.nb[k] = nb ? nb[k] : -1;
  C.eps = eps;
  C.maxlist = maxlist > 0 ? maxlist : 1000000;
  return 0;
}
}  // namespace gxy

GXY_EXPORT int gxy_outbound_count(int face) {
  if (face < 0 || face > 5) return fail("face %d outside 0..5", face);
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  return (int)G.outbound[face].size();
}
GXY_EXPORT void *gxy_outbound_take(int face, RayList_ispc *ispc, int *n) {
  if (face < 0 || face > 5) {
    fail("face %d outside 0..5", face);
    return nullptr;
  }
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  if (G.outbound[face].empty()) return nullptr;
  RayListObj *r = (RayListObj *)G.outbound[face].front();
  G.outbound[face].erase(G.outbound[face].begin());
  if (ispc) *ispc = r->ispc;
  if (n) *n = r->hdr.size;
  return r;
}
GXY_EXPORT int gxy_outbound_clear(void) {
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  for (int f = 0; f < 6; f++) {
    for (void *r : G.outbound[f]) raylist_delete((RayListObj *)r);
    G.outbound[f].clear();
  }
  return 0;
}

GXY_EXPORT int gxy_process_raylist(void *vis, void *lights, int n, RayList_ispc *rays, const int nb[6], float eps,
                                   int maxlist, float *fb, int w, int h, gxy_frame_stats *stats) {
  INIT_OR(-1);
  LoopCtx C;
  if (loop_setup(C, vis, lights, nb, eps, maxlist) != 0) return -1;
  if (!C.have_lights) return fail("gxy_process_raylist needs a Lighting_ispc");
  C.fb = fb;
  C.w = w;
  C.h = h;
  int r = run_loop(C, n, *rays);
  if (stats) *stats = C.st;
  return r;
}

GXY_EXPORT int gxy_render_frame(void *vis, void *lights, const float eye[3], const float center[3],
                                const float up[3], float aov, int w, int h, const int nb[6], float eps, int maxlist,
                                float *fb, gxy_frame_stats *stats) {
  INIT_OR(-1);
  const Visualization_ispc *V = (const Visualization_ispc *)vis;
  if (!V) return fail("null visualization");
  RayListObj *prim = raylist_new(w * h, 0, 0, false);
  if (!prim) return -1;
  prim->kinds = 1;    // camera rays: all primary, one per pixel
  prim->unique = 1;
  LoopCtx C;
  if (loop_setup(C, vis, lights, nb, eps, maxlist) != 0 || !C.have_lights) {
    raylist_delete(prim);
    return C.have_lights ? -1 : fail("gxy_render_frame needs a Lighting_ispc");
  }
  int np = 0;
  unsigned long long base = 0;
  // the camera rays' trace is queued behind the readback of their count (the
  // list is processed in place when it fits one list)
  auto ahead = [&](const int *dcount) {
    return can_pretrace(C, (long long)w * h) ? pretrace(C, prim, w * h, dcount, 1) : 0;
  };
  if (camera_generate(eye, center, 0, up, aov, w, h, &V->local_bb, &V->global_bb, prim->ispc, w * h, 1, &np,
                      &base, ahead) != 0) {
    raylist_delete(prim);
    return -1;
  }
  // the sample counter is not cleared: the frame's samples are counted from its
  // value read before the camera rays' trace
  C.have_base = true;
  C.samples_base = base;
  C.fb = fb;
  C.w = w;
  C.h = h;
  // Truncate (Rays.cpp:245-286) keeps the block; the loop consumes the list
  prim->hdr.size = np;
  int r = run_loop(C, np, prim->ispc, prim);
  C.st.primary_rays = np;
  if (stats) *stats = C.st;
  return r;
}

GXY_EXPORT int gxy_sample_raylist(void *vis, int n, RayList_ispc *rays, int maxlist, float *xyz, int capacity,
                                  int *count, gxy_frame_stats *stats) {
  INIT_OR(-1);
  LoopCtx C;
  if (loop_setup(C, vis, nullptr, nullptr, 0.f, maxlist) != 0) return -1;
  if (C.P.nvv < 1) return fail("sampler needs at least one volume");
  C.sampler = true;
  C.xyz = xyz;
  C.capacity = capacity;
  Scratch sc;
  C.dcount = (int *)sc.get(sizeof(int));
  if (!C.dcount) return -1;
  HIPCHK(hipMemsetAsync(C.dcount, 0, sizeof(int), S()));
  int r = run_loop(C, n, *rays);
  int c = 0;
  HIPCHK(hipMemcpyAsync(&c, C.dcount, sizeof(int), hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  *count = c;
  if (stats) *stats = C.st;
  return r;
}

// ====================================================================== batched form
namespace gxy {

struct BatchObj {
  // launch configuration
  int chunk = 256, stack = 0, wpe = 0, bpc = 4, split = 0;
  int cus = 0;
  hipStream_t ks[3] = {nullptr, nullptr, nullptr};   // streams of the per-kind trace kernels
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int *kidx = nullptr;                                // per-kind chunk lists, 3 x kcap
  int kcap = 0;
  // datasets and lighting
  std::vector<TraceParams> ds;
  std::vector<int> ds_depth;
  LightParams L;
  bool have_L = false;
  // lists (slot index = list id); rls[2 + l] = list l's rays
  std::vector<BatchList> lists;
  std::vector<char> live;
  std::vector<unsigned> seed;
  std::vector<std::array<float, 6>> seed_box;
  std::vector<char> has_seed;
  struct Waiting {
    int list, start, end;            // primary rays [start, end) of the list not yet in a step
  };
  std::vector<Waiting> fresh;        // in the order the lists were added
  // device state
  TraceParams *d_ds = nullptr;
  int d_ds_cap = 0;
  BatchList *d_lists = nullptr;
  RayList_ispc *d_rls = nullptr;
  int d_lists_cap = 0;
  RayListObj *pool[2] = {nullptr, nullptr};
  BatchChunk *tab[2] = {nullptr, nullptr};
  int tcap[2] = {0, 0};
  BatchChunk *d_init = nullptr, *d_fresh = nullptr;
  int init_cap = 0, fresh_cap = 0, ninit = 0;
  int *pc = nullptr;                 // per-chunk arrays: cpar, cout, cchk, coff, cchkoff (pc_cap each)
  int pc_cap = 0;
  long long *seg = nullptr;          // 2 * 256
  int *lpend = nullptr;
  int lpend_cap = 0;
  int lpend_n = 0;    // lists whose lpend entries the passes reset (the last pass arguments built)
  unsigned long long *part = nullptr;
  int part_blocks = 0;
  unsigned long long *stats = nullptr;  // BS_COUNT
  int *ints = nullptr;               // 0,1: table counts; 2 nrays_out; 3 head; 4 overflow
  int pass = 0;
  long long queued = 0;              // rays queued for the next pass (host view)
  long long step_rays = 0;           // primary rays of added lists per step (0 = all)
  int max_segments = 4096;           // sampler rays: crossings followed per pass
  int passes_run = 0;
  bool dirty_tables = true;          // list / dataset tables need upload
  bool dirty_graph = true;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gsig;
  std::vector<void *> owned;
  void free_all() {
    if (gexec) hipGraphExecDestroy(gexec);
    gexec = nullptr;
    for (int k = 0; k < 2; k++)
      if (pool[k]) raylist_delete(pool[k]);
    pool[0] = pool[1] = nullptr;
    void *bufs[] = {d_ds, d_lists, d_rls, tab[0], tab[1], d_init, d_fresh, pc, seg, lpend, part, stats, ints, kidx};
    for (void *p : bufs) dfree(p);
    for (int q = 0; q < 3; q++)
      if (ks[q]) hipStreamDestroy(ks[q]);
    for (int q = 0; q < 4; q++)
      if (ev[q]) hipEventDestroy(ev[q]);
  }
};

static std::set<void *> g_batches;

static int nsp_of(const LightParams &L) { return L.nAO + (L.doShadows ? L.nLights : 0); }

// grows a device buffer of `elem`-byte entries to at least n entries (keeps
// the first `keep` entries)
static int grow(void **p, int *cap, long long n, size_t elem, long long keep = 0) {
  if (n <= *cap) return 0;
  if (n > INT32_MAX) return fail("batch table of %lld entries exceeds 2^31", n);
  long long c = std::max<long long>(n, (long long)*cap * 3 / 2);
  if (c > INT32_MAX) c = n;
  void *q = dalloc((size_t)c * elem);
  if (!q) return -1;
  if (keep > 0 && *p) HIPCHK(hipMemcpyAsync(q, *p, (size_t)keep * elem, hipMemcpyDeviceToDevice, S()));
  if (*p) dfree(*p);
  *p = q;
  *cap = (int)c;
  return 0;
}

static int batch_blocks(BatchObj *B) { return std::max(1, B->cus * B->bpc); }

static int batch_upload_tables(BatchObj *B) {
  if (!B->dirty_tables) return 0;
  const int nl = (int)B->lists.size(), nd = (int)B->ds.size();
  if (grow((void **)&B->d_ds, &B->d_ds_cap, std::max(1, nd), sizeof(TraceParams)) != 0) return -1;
  int lc = B->d_lists_cap;
  if (nl + 2 > lc) {
    int c1 = lc, c2 = lc;
    if (grow((void **)&B->d_lists, &c1, nl + 2, sizeof(BatchList)) != 0) return -1;
    if (grow((void **)&B->d_rls, &c2, nl + 2, sizeof(RayList_ispc)) != 0) return -1;
    B->d_lists_cap = std::min(c1, c2);
  }
  if (nl > B->lpend_cap && grow((void **)&B->lpend, &B->lpend_cap, nl, sizeof(int)) != 0) return -1;
  if (nd) HIPCHK(hipMemcpyAsync(B->d_ds, B->ds.data(), nd * sizeof(TraceParams), hipMemcpyHostToDevice, S()));
  if (nl) HIPCHK(hipMemcpyAsync(B->d_lists, B->lists.data(), nl * sizeof(BatchList), hipMemcpyHostToDevice, S()));
  std::vector<RayList_ispc> rl(nl + 2);
  memset(rl.data(), 0, rl.size() * sizeof(RayList_ispc));
  for (int k = 0; k < 2; k++)
    if (B->pool[k]) rl[k] = B->pool[k]->ispc;
  for (int l = 0; l < nl; l++) rl[2 + l] = B->lists[l].R;
  HIPCHK(hipMemcpyAsync(B->d_rls, rl.data(), rl.size() * sizeof(RayList_ispc), hipMemcpyHostToDevice, S()));
  HIPCHK(hipStreamSynchronize(S()));   // host vectors may change after return
  B->dirty_tables = false;
  return 0;
}

// pool k holds at least n rays
static int batch_pool(BatchObj *B, int k, long long n) {
  if (B->pool[k] && B->pool[k]->hdr.aligned_size >= n) return 0;
  if (n > INT32_MAX - 16) return fail("ray pool of %lld rays exceeds 2^31", n);
  long long c = std::max<long long>(n, B->pool[k] ? (long long)B->pool[k]->hdr.aligned_size * 3 / 2 : 0);
  if (c > INT32_MAX - 16) c = n;
  RayListObj *r = raylist_new((int)c, 0, 1, false);
  if (!r) return -1;
  if (B->pool[k]) raylist_delete(B->pool[k]);
  B->pool[k] = r;
  B->dirty_tables = true;
  B->dirty_graph = true;
  return 0;
}

static void batch_chunks_of(BatchObj *B, int l, std::vector<BatchChunk> &out) {
  const int n = B->lists[l].n;
  for (int s = 0; s < n; s += B->chunk) {
    BatchChunk c;
    c.src = 2 + l;
    c.start = s;
    c.count = std::min(B->chunk, n - s);
    c.list = l;
    out.push_back(c);
  }
}

static BatchArgs batch_args(BatchObj *B, int pass) {
  BatchArgs A;
  memset(&A, 0, sizeof A);
  const int in = (pass + 1) & 1, out = pass & 1;
  A.lists = B->d_lists;
  A.nlists = (int)B->lists.size();
  B->lpend_n = A.nlists;
  A.chunk = B->chunk;
  A.ds = B->d_ds;
  A.rls = B->d_rls;
  A.pool_cap = B->pool[out] ? B->pool[out]->hdr.aligned_size : 0;
  A.out_pool = out;
  A.tin = B->tab[in];
  A.nin = B->ints + in;
  A.tout = B->tab[out];
  A.nout = B->ints + out;
  A.tcap = std::min(B->tcap[out], B->pc_cap);
  A.nrays_out = B->ints + 2;
  A.head = B->ints + 3;
  A.overflow = B->ints + 4;
  A.cpar = B->pc;
  A.cout = B->pc + (size_t)B->pc_cap;
  A.cchk = B->pc + 2 * (size_t)B->pc_cap;
  A.coff = B->pc + 3 * (size_t)B->pc_cap;
  A.cchkoff = B->pc + 4 * (size_t)B->pc_cap;
  A.segr = B->seg;
  A.segc = B->seg + 256;
  A.lpend = B->lpend;
  A.part = B->part;
  A.nblocks = B->part_blocks;
  A.stats = B->stats;
  A.kidx = B->kidx;
  A.kcount = B->ints + 8;
  A.khead = B->ints + 11;
  A.kcap = B->kcap;
  A.max_segments = B->max_segments;
  A.L = B->L;
  return A;
}

static BatchLaunch batch_launch(BatchObj *B) {
  BatchLaunch K;
  K.trace_blocks = batch_blocks(B);
  K.post_blocks = B->part_blocks;   // equal to trace_blocks: every partial slot is written
  K.stack = B->stack;
  K.wpe = B->wpe;
  K.split = B->split;
  return K;
}

// capacity for a pass whose input holds in_rays rays (at most prim of them
// primary, the only rays that spawn) in in_chunks chunks
static int batch_reserve_pass(BatchObj *B, int pass, long long in_rays, long long in_chunks, long long prim) {
  const int in = (pass + 1) & 1, out = pass & 1;
  const long long out_rays = in_rays + std::min(in_rays, prim) * nsp_of(B->L) + 16;
  if (batch_pool(B, out, out_rays) != 0) return -1;
  const long long out_chunks = out_rays / B->chunk + in_chunks + 16;
  if (grow((void **)&B->tab[out], &B->tcap[out], out_chunks, sizeof(BatchChunk)) != 0) return -1;
  (void)in;
  const long long need_pc = std::max(in_chunks, out_chunks) + 16;
  if (need_pc > B->pc_cap) {
    int c = B->pc_cap;
    void *p = B->pc;
    long long want = std::max<long long>(need_pc, (long long)c * 3 / 2);
    if (want > INT32_MAX / 5) return fail("batch chunk arrays too large");
    void *q = dalloc((size_t)want * 5 * sizeof(int));
    if (!q) return -1;
    if (p) dfree(p);
    B->pc = (int *)q;
    B->pc_cap = (int)want;
    dfree(B->kidx);
    B->kidx = (int *)dalloc((size_t)want * 3 * sizeof(int));
    if (!B->kidx) return -1;
    B->kcap = (int)want;
    B->dirty_graph = true;
  }
  return 0;
}

static int batch_ensure_base(BatchObj *B) {
  if (!B->ints) {
    B->ints = (int *)dalloc(16 * sizeof(int));
    B->seg = (long long *)dalloc(512 * sizeof(long long));
    B->stats = (unsigned long long *)dalloc(BS_COUNT * sizeof(unsigned long long));
    if (!B->ints || !B->seg || !B->stats) return -1;
    HIPCHK(hipMemsetAsync(B->ints, 0, 16 * sizeof(int), S()));
    for (int q = 0; q < 3; q++) HIPCHK(hipStreamCreateWithFlags(&B->ks[q], hipStreamNonBlocking));
    for (int q = 0; q < 4; q++) HIPCHK(hipEventCreateWithFlags(&B->ev[q], hipEventDisableTiming));
    HIPCHK(hipMemsetAsync(B->stats, 0, BS_COUNT * sizeof(unsigned long long), S()));
  }
  const int nb = batch_blocks(B);
  if (nb != B->part_blocks) {
    dfree(B->part);
    B->part = (unsigned long long *)dalloc((size_t)nb * 4 * BS_COUNT * sizeof(unsigned long long));
    if (!B->part) return -1;
    B->part_blocks = nb;
    B->dirty_graph = true;
  }
  return 0;
}

static int batch_check(BatchObj *B) {
  if (!B->have_L) return fail("gxy_batch_set_lighting not called");
  if (B->stack > 0)
    for (size_t d = 0; d < B->ds.size(); d++)
      if (B->ds_depth[d] + 2 > B->stack)
        return fail("dataset %zu: BVH depth %d needs a traversal stack of at least %d entries (configured %d)", d,
                    B->ds_depth[d], B->ds_depth[d] + 2, B->stack);
  return 0;
}

// the set-up of a run: statistics and outputs cleared, seeded lists
// regenerated, every list's primary chunks queued as the input of pass 0
// everything a captured run bakes into its kernel arguments
static std::vector<uint64_t> batch_graph_sig(BatchObj *B, int graph_passes) {
  std::vector<uint64_t> v = {(uint64_t)graph_passes, (uint64_t)B->split, (uint64_t)(uintptr_t)B->kidx,
                             (uint64_t)(uintptr_t)B->tab[0], (uint64_t)(uintptr_t)B->tab[1],
                             (uint64_t)B->tcap[0], (uint64_t)B->tcap[1], (uint64_t)(uintptr_t)B->pc,
                             (uint64_t)B->pc_cap, (uint64_t)(uintptr_t)B->d_init, (uint64_t)B->ninit,
                             (uint64_t)(uintptr_t)B->d_ds, (uint64_t)(uintptr_t)B->d_lists,
                             (uint64_t)(uintptr_t)B->d_rls, (uint64_t)(uintptr_t)B->part, (uint64_t)B->part_blocks,
                             (uint64_t)(uintptr_t)B->lpend, (uint64_t)B->chunk, (uint64_t)B->stack,
                             (uint64_t)B->wpe, (uint64_t)B->bpc, (uint64_t)B->lists.size(),
                             (uint64_t)B->max_segments};
  for (int k = 0; k < 2; k++) {
    v.push_back((uint64_t)(uintptr_t)B->pool[k]);
    v.push_back(B->pool[k] ? (uint64_t)(uintptr_t)B->pool[k]->block : 0);
  }
  const unsigned char *lp = (const unsigned char *)&B->L;
  for (size_t k = 0; k < sizeof B->L; k += 8) {
    uint64_t w = 0;
    memcpy(&w, lp + k, std::min<size_t>(8, sizeof B->L - k));
    v.push_back(w);
  }
  for (size_t l = 0; l < B->lists.size(); l++) {
    const BatchList &Ls = B->lists[l];
    v.push_back(B->live[l]);
    v.push_back((uint64_t)(uintptr_t)Ls.fb);
    v.push_back((uint64_t)(uintptr_t)Ls.pcount);
    v.push_back((uint64_t)(uintptr_t)Ls.R.ox);
    v.push_back(((uint64_t)Ls.n << 32) | (uint32_t)Ls.width);
    v.push_back(((uint64_t)Ls.height << 32) | (uint32_t)Ls.mode);
    v.push_back(B->has_seed[l] ? (((uint64_t)1 << 32) | B->seed[l]) : 0);
    for (int k = 0; k < 6; k++) {
      uint32_t w;
      memcpy(&w, &B->seed_box[l][k], 4);
      v.push_back(w);
    }
  }
  return v;
}

static int batch_setup_kernels(BatchObj *B) {
  HIPCHK(launch_fill_f32((float *)B->stats, 2 * BS_COUNT, 0.f, S()));
  HIPCHK(launch_set_count(B->ints + 4, 0, S()));
  for (size_t l = 0; l < B->lists.size(); l++) {
    if (!B->live[l]) continue;
    const BatchList &Ls = B->lists[l];
    if (Ls.mode == BM_RENDER) HIPCHK(launch_fill_f32(Ls.fb, 4LL * Ls.width * Ls.height, 0.f, S()));
    else HIPCHK(launch_set_count(Ls.pcount, 0, S()));
    if (B->has_seed[l]) {
      const float *bx = B->seed_box[l].data();
      HIPCHK(launch_random_rays(Ls.R, Ls.n, B->seed[l], Ls.width, RAY_PRIMARY, bx, bx + 3, S()));
    }
  }
  HIPCHK(launch_set_count(B->ints + 1, 0, S()));
  HIPCHK(launch_set_count(B->ints + 0, 0, S()));
  HIPCHK(launch_batch_append(B->tab[1], B->ints + 1, B->tcap[1], B->d_init, B->ninit, B->ints + 4, S()));
  return 0;
}

static int batch_pass_kernels(BatchObj *B, int pass) {
  BatchArgs A = batch_args(B, pass);
  HIPCHK(launch_batch_pass(A, batch_launch(B), S(), B->ks, B->ev));
  HIPCHK(launch_batch_reduce(A, S()));
  return 0;
}

static int batch_read_counts(BatchObj *B, int pass, long long *rays, int *chunks, int *overflow) {
  int v[8];
  HIPCHK(hipMemcpyAsync(v, B->ints, sizeof v, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  *chunks = v[pass & 1];
  *rays = v[2];
  *overflow = v[4];
  return 0;
}

static int batch_stats_read(BatchObj *B, gxy_batch_stats *st) {
  if (!st) return 0;
  unsigned long long v[BS_COUNT];
  HIPCHK(hipMemcpyAsync(v, B->stats, sizeof v, hipMemcpyDeviceToHost, S()));
  int ov = 0;
  HIPCHK(hipMemcpyAsync(&ov, B->ints + 4, sizeof ov, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  st->rays = v[BS_RAYS];
  st->samples = v[BS_SAMPLES];
  st->spawned = v[BS_SPAWNED];
  st->terminated = v[BS_TERMINATED];
  st->dropped = v[BS_DROPPED];
  st->kept = v[BS_KEPT];
  st->passes = B->passes_run;
  st->overflow = ov;
  return 0;
}

}  // namespace gxy

GXY_EXPORT void *gxy_batch_create(int chunk) {
  INIT_OR(nullptr);
  if (chunk < 64 || chunk > 1024 || chunk % 64) {
    fail("chunk %d: must be a multiple of 64 in [64, 1024]", chunk);
    return nullptr;
  }
  BatchObj *B = new BatchObj;
  B->chunk = chunk;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, G.device) != hipSuccess) {
    delete B;
    fail("hipGetDeviceProperties failed");
    return nullptr;
  }
  B->cus = prop.multiProcessorCount;
  memset(&B->L, 0, sizeof B->L);
  if (batch_ensure_base(B) != 0) {
    B->free_all();
    delete B;
    return nullptr;
  }
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  g_batches.insert(B);
  return B;
}

GXY_EXPORT int gxy_batch_destroy(void *b) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    g_batches.erase(B);
  }
  HIPCHK(hipStreamSynchronize(S()));
  B->free_all();
  delete B;
  return 0;
}

static int batch_dataset_params(BatchObj *B, const Visualization_ispc *vis, float eps, TraceParams &P, int &depth);

GXY_EXPORT int gxy_batch_update_dataset(void *b, int ds, const Visualization_ispc *vis, float eps) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (ds < 0 || ds >= (int)B->ds.size()) return fail("dataset %d out of range", ds);
  TraceParams P;
  int depth = 0;
  if (batch_dataset_params(B, vis, eps, P, depth) != 0) return -1;
  if ((P.nvv == 1) != (B->ds[ds].nvv == 1)) return fail("dataset %d changes kind", ds);
  B->ds[ds] = P;
  B->ds_depth[ds] = depth;
  B->dirty_tables = B->dirty_graph = true;
  return 0;
}

static int batch_dataset_params(BatchObj *B, const Visualization_ispc *vis, float eps, TraceParams &P, int &depth) {
  (void)B;
  if (make_trace_params(vis, eps, P) != 0) return -1;
  const bool volume = P.nvv == 1 && !P.has_model, geometry = P.nvv == 0 && P.has_model;
  if (!volume && !geometry)
    return fail("a batch dataset is one volume without geometry or geometry without volumes (got %d volumes%s)",
                P.nvv, P.has_model ? " and a model" : "");
  if (geometry && P.model.ncurve > 0)
    return fail("a batch dataset with path lines: path lines are traced by the per-list loop only");
  P.occlusion = (G.occlusion && geometry) ? 1 : 0;
  P.hit_geom = nullptr;
  P.hit_prim = nullptr;
  P.sample_count = nullptr;
  depth = 0;
  if (geometry) {
    ModelObj *M = lookup<ModelObj>(G.models, vis->model);
    depth = M ? (P.model.wide ? M->d.depth4 : std::max(M->tdepth, M->sdepth)) : 64;
  }
  return 0;
}

GXY_EXPORT int gxy_batch_add_dataset(void *b, const Visualization_ispc *vis, float eps) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  TraceParams P;
  int depth = 0;
  if (batch_dataset_params(B, vis, eps, P, depth) != 0) return -1;
  B->ds.push_back(P);
  B->ds_depth.push_back(depth);
  B->dirty_tables = B->dirty_graph = true;
  return (int)B->ds.size() - 1;
}

GXY_EXPORT int gxy_batch_set_lighting(void *b, const Lighting_ispc *lights, float eps) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (make_light_params(lights, eps, B->L) != 0) return -1;
  B->have_L = true;
  B->dirty_graph = true;
  return 0;
}

GXY_EXPORT int gxy_batch_add_list(void *b, int ds, int mode, int n, const RayList_ispc *rays, float *fb, int width,
                                  int height, float *xyz, int capacity, int *count) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (ds < 0 || ds >= (int)B->ds.size()) return fail("dataset %d out of range", ds);
  if (n < 0 || !rays) return fail("bad ray list");
  if (mode != BM_RENDER && mode != BM_SAMPLE) return fail("mode %d", mode);
  if (mode == BM_RENDER && (!fb || width <= 0 || height <= 0)) return fail("render list needs a framebuffer");
  if (mode == BM_SAMPLE && (!xyz || !count || capacity < 0)) return fail("sample list needs a particle buffer");
  if (mode == BM_SAMPLE && B->ds[ds].nvv != 1) return fail("sample list over a dataset without a volume");
  BatchList Ls;
  memset(&Ls, 0, sizeof Ls);
  Ls.R = *rays;
  Ls.fb = fb;
  Ls.xyz = xyz;
  Ls.pcount = count;
  Ls.width = width;
  Ls.height = height;
  Ls.pcap = capacity;
  Ls.ds = ds;
  Ls.mode = mode;
  Ls.n = n;
  int l = -1;
  for (size_t k = 0; k < B->live.size(); k++)
    if (!B->live[k]) {
      l = (int)k;
      break;
    }
  if (l < 0) {
    l = (int)B->lists.size();
    B->lists.push_back(Ls);
    B->live.push_back(1);
    B->seed.push_back(0);
    B->seed_box.push_back({0, 0, 0, 1, 1, 1});
    B->has_seed.push_back(0);
  } else {
    B->lists[l] = Ls;
    B->live[l] = 1;
    B->has_seed[l] = 0;
  }
  if (mode == BM_RENDER) HIPCHK(launch_fill_f32(fb, 4LL * width * height, 0.f, S()));
  else HIPCHK(launch_set_count(count, 0, S()));
  if (n > 0) B->fresh.push_back({l, 0, n});
  B->dirty_tables = B->dirty_graph = true;
  return l;
}

GXY_EXPORT int gxy_batch_set_list_seed(void *b, int l, unsigned seed, const float lo[3], const float hi[3]) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (l < 0 || l >= (int)B->lists.size() || !B->live[l]) return fail("no list %d", l);
  B->seed[l] = seed;
  for (int k = 0; k < 3; k++) {
    B->seed_box[l][k] = lo ? lo[k] : 0.f;
    B->seed_box[l][3 + k] = hi ? hi[k] : 1.f;
  }
  B->has_seed[l] = 1;
  B->dirty_graph = true;
  return 0;
}

GXY_EXPORT int gxy_batch_remove_list(void *b, int l) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (l < 0 || l >= (int)B->lists.size() || !B->live[l]) return fail("no list %d", l);
  for (const BatchObj::Waiting &w : B->fresh)
    if (w.list == l) return fail("list %d has not been traced yet", l);
  B->live[l] = 0;
  B->lists[l].n = 0;
  B->dirty_tables = B->dirty_graph = true;
  return 0;
}

GXY_EXPORT int gxy_batch_set_config(void *b, int chunk, int stack, int wpe, int bpc) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (chunk < 64 || chunk > 1024 || chunk % 64) return fail("chunk %d: a multiple of 64 in [64, 1024]", chunk);
  if (stack != 0 && stack != 16 && stack != 24 && stack != 32) return fail("stack %d: 0, 16, 24 or 32", stack);
  if (wpe != 0 && wpe != 4 && wpe != 5 && wpe != 6) return fail("wpe %d: 0, 4, 5 or 6", wpe);
  if (bpc < 1 || bpc > 16) return fail("blocks per CU %d", bpc);
  if (chunk != B->chunk && B->queued > 0) return fail("chunk size changes only with no rays queued");
  B->chunk = chunk;
  B->stack = stack;
  B->wpe = wpe;
  B->bpc = bpc;
  B->dirty_graph = true;
  return batch_ensure_base(B);
}

GXY_EXPORT int gxy_batch_set_split(void *b, int split) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (split != 0 && split != 1) return fail("split %d: 0 or 1", split);
  B->split = split;
  B->dirty_graph = true;
  return 0;
}
GXY_EXPORT int gxy_batch_set_step_rays(void *b, long long rays) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (rays < 0) return fail("step budget %lld", rays);
  B->step_rays = rays;
  return 0;
}
GXY_EXPORT int gxy_batch_set_max_segments(void *b, int n) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (n < 1 || n > 4096) return fail("max segments %d: 1..4096", n);
  B->max_segments = n;
  B->dirty_graph = true;
  return 0;
}
GXY_EXPORT long long gxy_batch_get_step_rays(void *b) {
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  return B->step_rays;
}
GXY_EXPORT int gxy_batch_get_split(void *b) {
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  return B->split;
}
GXY_EXPORT int gxy_batch_get_config(void *b, int *chunk, int *stack, int *wpe, int *bpc) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (chunk) *chunk = B->chunk;
  if (stack) *stack = B->stack;
  if (wpe) *wpe = B->wpe;
  if (bpc) *bpc = B->bpc;
  return 0;
}

GXY_EXPORT int gxy_batch_run(void *b, int graph_passes, gxy_batch_stats *st) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (batch_check(B) != 0) return -1;
  if (B->queued > 0) return fail("gxy_batch_run with rays still queued by gxy_batch_step");
  B->fresh.clear();   // the run queues every list itself
  // capacities for the worst case of pass 0 (every primary ray spawning)
  std::vector<BatchChunk> init;
  long long nprim = 0;
  for (size_t l = 0; l < B->lists.size(); l++)
    if (B->live[l]) {
      batch_chunks_of(B, (int)l, init);
      nprim += B->lists[l].n;
    }
  if (batch_ensure_base(B) != 0) return -1;
  const long long nch = (long long)init.size();
  if (grow((void **)&B->tab[1], &B->tcap[1], nch + 16, sizeof(BatchChunk)) != 0) return -1;
  // pass 0 reads the primary chunks; later passes read at most what pass 0 wrote
  if (batch_reserve_pass(B, 0, nprim, nch, nprim) != 0) return -1;
  if (batch_reserve_pass(B, 1, nprim * (1 + nsp_of(B->L)), (long long)B->tcap[0], 0) != 0) return -1;
  if ((int)nch > B->init_cap || B->dirty_graph) {
    if (grow((void **)&B->d_init, &B->init_cap, std::max<long long>(1, nch), sizeof(BatchChunk)) != 0) return -1;
    if (nch) HIPCHK(hipMemcpyAsync(B->d_init, init.data(), nch * sizeof(BatchChunk), hipMemcpyHostToDevice, S()));
    B->ninit = (int)nch;
    HIPCHK(hipStreamSynchronize(S()));
  }
  if (batch_upload_tables(B) != 0) return -1;
  B->passes_run = 0;
  int pass = 0;
  if (graph_passes > 0) {
    std::vector<uint64_t> sig = batch_graph_sig(B, graph_passes);
    if (!B->gexec || sig != B->gsig) {
      if (B->gexec) HIPCHK(hipGraphExecDestroy(B->gexec));
      B->gexec = nullptr;
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(S(), hipStreamCaptureModeThreadLocal));
      int rc = batch_setup_kernels(B);
      for (int p = 0; rc == 0 && p < graph_passes; p++) rc = batch_pass_kernels(B, p);
      hipError_t e = hipStreamEndCapture(S(), &g);
      if (rc != 0) return -1;
      HIPCHK(e);
      e = hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      HIPCHK(e);
      B->gsig = sig;
      B->dirty_graph = false;
    }
    {
      KTimer kt("batch_graph");
      HIPCHK(hipGraphLaunch(B->gexec, S()));
    }
    pass = graph_passes;
    B->passes_run = graph_passes;
  } else {
    if (batch_setup_kernels(B) != 0) return -1;
  }
  for (;;) {
    long long rays = 0;
    int chunks = 0, ov = 0;
    if (pass > 0) {
      if (batch_read_counts(B, pass - 1, &rays, &chunks, &ov) != 0) return -1;
      if (ov) return fail("batch pass %d overflowed its ray pool or chunk table (flags %d)", pass - 1, ov);
      if (chunks == 0) break;
      if (batch_reserve_pass(B, pass, rays, chunks, nprim) != 0) return -1;
      if (B->dirty_tables && batch_upload_tables(B) != 0) return -1;
    }
    {
      KTimer kt("batch_pass");
      if (batch_pass_kernels(B, pass) != 0) return -1;
    }
    pass++;
    B->passes_run++;
    if (pass > 4096) return fail("batch did not drain after 4096 passes");
  }
  B->pass = 0;
  B->queued = 0;
  return batch_stats_read(B, st);
}

GXY_EXPORT int gxy_batch_step(void *b, long long *next_rays) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  if (batch_check(B) != 0) return -1;
  const int p = B->pass, in = (p + 1) & 1;
  int nin = 0;
  HIPCHK(hipMemcpyAsync(&nin, B->ints + in, sizeof nin, hipMemcpyDeviceToHost, S()));
  HIPCHK(hipStreamSynchronize(S()));
  if (B->queued == 0) nin = 0;
  // the primary chunks joining this step: all, or up to the step budget with
  // the lists that have the fewest waiting rays first (each list's chunks stay
  // in order, so its remaining rays are a suffix)
  std::vector<BatchChunk> take;
  {
    // waiting ranges by rays left, fewest first (stable: ties in the order added)
    std::vector<int> order(B->fresh.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
    if (B->step_rays > 0)
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return B->fresh[a].end - B->fresh[a].start < B->fresh[b].end - B->fresh[b].start;
      });
    long long budget = B->step_rays > 0 ? B->step_rays : LLONG_MAX;
    // whole chunks: a list's share may exceed its limit by less than a chunk
    auto take_from = [&](BatchObj::Waiting &w, long long limit) {
      long long got = 0;
      while (w.start < w.end && budget > 0 && got < limit) {
        BatchChunk c;
        c.src = 2 + w.list;
        c.start = w.start;
        c.count = std::min(B->chunk, w.end - w.start);
        c.list = w.list;
        take.push_back(c);
        w.start += c.count;
        budget -= c.count;
        got += c.count;
      }
    };
    // the oldest waiting list keeps a quarter of every step, so that a stream
    // of small lists cannot hold a large one back indefinitely
    if (B->step_rays > 0 && !B->fresh.empty()) take_from(B->fresh[0], B->step_rays / 4);
    for (int k : order) {
      if (budget <= 0) break;
      take_from(B->fresh[k], LLONG_MAX);
    }
    size_t o = 0;
    for (size_t k = 0; k < B->fresh.size(); k++)
      if (B->fresh[k].start < B->fresh[k].end) B->fresh[o++] = B->fresh[k];
    B->fresh.resize(o);
  }
  long long in_rays = B->queued;
  for (const BatchChunk &c : take) in_rays += c.count;
  const long long in_chunks = (long long)nin + (long long)take.size();
  if (grow((void **)&B->tab[in], &B->tcap[in], in_chunks + 16, sizeof(BatchChunk), nin) != 0) return -1;
  long long prim = 0;
  for (size_t l = 0; l < B->lists.size(); l++)
    if (B->live[l] && B->lists[l].mode == BM_RENDER) prim += B->lists[l].n;
  if (batch_reserve_pass(B, p, in_rays, in_chunks, prim) != 0) return -1;
  if (batch_upload_tables(B) != 0) return -1;
  if (B->queued == 0) HIPCHK(launch_set_count(B->ints + in, 0, S()));
  if (!take.empty()) {
    const int nf = (int)take.size();
    if (grow((void **)&B->d_fresh, &B->fresh_cap, nf, sizeof(BatchChunk)) != 0) return -1;
    HIPCHK(hipMemcpyAsync(B->d_fresh, take.data(), nf * sizeof(BatchChunk), hipMemcpyHostToDevice, S()));
    HIPCHK(launch_batch_append(B->tab[in], B->ints + in, B->tcap[in], B->d_fresh, nf, B->ints + 4, S()));
    HIPCHK(hipStreamSynchronize(S()));
  }
  {
    KTimer kt("batch_pass");
    if (batch_pass_kernels(B, p) != 0) return -1;
  }
  long long rays = 0;
  int chunks = 0, ov = 0;
  if (batch_read_counts(B, p, &rays, &chunks, &ov) != 0) return -1;
  if (ov) return fail("batch pass overflowed its ray pool or chunk table (flags %d)", ov);
  B->pass = p + 1;
  B->passes_run++;
  B->queued = rays;
  long long waiting = 0;
  for (const BatchObj::Waiting &w : B->fresh) waiting += w.end - w.start;
  if (next_rays) *next_rays = rays + waiting;
  return 0;
}

GXY_EXPORT int gxy_batch_pending(void *b, int *per_list, int n) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  const int nl = std::min(n, (int)B->lists.size());
  if (nl <= 0) return 0;
  // lpend holds the counts of the lists the last passes covered; lists added
  // since then have only waiting rays
  const int nc = B->passes_run == 0 ? 0 : std::min(nl, std::min(B->lpend_n, B->lpend_cap));
  for (int l = nc; l < nl; l++) per_list[l] = 0;
  if (nc > 0) {
    HIPCHK(hipMemcpyAsync(per_list, B->lpend, nc * sizeof(int), hipMemcpyDeviceToHost, S()));
    HIPCHK(hipStreamSynchronize(S()));
  }
  for (const BatchObj::Waiting &w : B->fresh)
    if (w.list < nl) per_list[w.list] += w.end - w.start;
  return 0;
}

GXY_EXPORT int gxy_batch_get_stats(void *b, gxy_batch_stats *st) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  return batch_stats_read(B, st);
}

GXY_EXPORT long long gxy_batch_bytes(void *b) {
  INIT_OR(-1);
  BatchObj *B = lookup<BatchObj>(g_batches, b);
  if (!B) return fail("not a batch handle");
  // the device blocks the batch holds, at their allocated (size-class) sizes
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  auto held = [](const void *p) -> long long {
    auto it = G.live.find((void *)p);
    return it == G.live.end() ? 0 : (long long)it->second;
  };
  long long t = 0;
  for (int k = 0; k < 2; k++)
    if (B->pool[k]) t += held(B->pool[k]->block);
  const void *bufs[] = {B->d_ds, B->d_lists, B->d_rls, B->tab[0], B->tab[1], B->d_init, B->d_fresh,
                        B->pc,   B->seg,     B->lpend, B->part,   B->stats,  B->ints,    B->kidx};
  for (const void *p : bufs) t += held(p);
  return t;
}

GXY_EXPORT int gxy_rays_random(const RayList_ispc *rays, int n, unsigned seed, int width, int type, const float lo[3],
                               const float hi[3]) {
  INIT_OR(-1);
  if (!rays || n < 0) return fail("bad ray list");
  const float dlo[3] = {0, 0, 0}, dhi[3] = {1, 1, 1};
  HIPCHK(launch_random_rays(*rays, n, seed, width, type, lo ? lo : dlo, hi ? hi : dhi, S()));
  return 0;
}

// ====================================================================== served form
namespace gxy {

// Request states (gxy_request_result.status)
enum { RQ_DONE = 0, RQ_FAILED = -1, RQ_REJECTED = -2, RQ_CANCELLED = -3, RQ_QUEUED = 1, RQ_RUNNING = 2 };

struct Flight {
  int req = -1, list = -1;
  RayListObj *rays = nullptr;
  float *fb = nullptr, *xyz = nullptr;
  int *cnt = nullptr;
  long long bytes = 0;
  int n = 0;
};

// One request of a server: the caller's description (its out buffer must stay
// valid until the request is final), its result, and whether a cancellation
// is asked for.
struct ReqRecord {
  gxy_request req;
  gxy_request_result res;
  bool cancel = false;
};

struct ServerObj {
  long long budget = 0;
  int slots = 1;
  void *batch = nullptr;
  std::vector<const Visualization_ispc *> vis;
  std::vector<float> eps;
  const Lighting_ispc *lights = nullptr;
  float leps = 0.001f;
  int nsp = 0;
  // the requests, the queue of those not yet admitted (ids in arrival order)
  // and the admitted ones; guarded by mu.  A record's request is not changed
  // after it is added (the deque keeps references to records valid), so the
  // service thread reads it without the lock.
  std::mutex mu;
  std::condition_variable cv;        // a request arrived or a cancellation was asked for
  std::condition_variable idle_cv;   // a step ended or the queue changed
  std::deque<ReqRecord> recs;
  std::deque<int> queue;
  std::vector<Flight> fl;
  int admitting = 0;                   // requests taken off the queue, not yet in fl (admission under way)
  long long base_bytes = 0;          // device bytes in use when serving started
  long long peak = 0;
  int max_in_flight = 0;
  std::chrono::steady_clock::time_point t0;
  // the live service thread (gxy_server_start)
  std::thread th;
  bool running = false, stop = false;
  int loop_error = 0;
  std::string loop_msg;
};

static std::set<void *> g_servers;

static double since(const ServerObj *S_) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - S_->t0).count();
}

static long long request_bytes(const ServerObj *S_, const gxy_request &r) {
  if (r.kind == 0) {
    const long long px = (long long)r.width * r.height;
    return px * (100 + 16) + 2 * px * (1 + S_->nsp) * 100;
  }
  return r.nrays * 100 * 3 + r.capacity * 12;
}

// a request the server cannot take at all (before any device work)
static const char *request_invalid(const ServerObj *S_, const gxy_request &r) {
  if (r.kind != 0 && r.kind != 1) return "request kind must be 0 (render) or 1 (sample)";
  if (r.dataset < 0 || r.dataset >= (int)S_->vis.size()) return "request dataset out of range";
  if (r.kind == 0 && (r.width <= 0 || r.height <= 0 || (long long)r.width * r.height > INT32_MAX / 2))
    return "render request size out of range";
  if (r.kind == 1 && (r.nrays <= 0 || r.nrays > INT32_MAX / 2 || r.capacity < 0))
    return "sample request size out of range";
  if (request_bytes(S_, r) > S_->budget) return "request needs more device memory than the server's budget";
  return nullptr;
}

static void flight_free(Flight &f) {
  if (f.rays) raylist_delete(f.rays);
  dfree(f.fb);
  dfree(f.xyz);
  dfree(f.cnt);
  f.rays = nullptr;
  f.fb = f.xyz = nullptr;
  f.cnt = nullptr;
}

// the request's rays and outputs on the device
static int flight_make(ServerObj *S_, const gxy_request &r, Flight &f) {
  const Visualization_ispc *v = S_->vis[r.dataset];
  if (r.kind == 0) {
    const int px = r.width * r.height;
    f.rays = raylist_new(px, 0, 0, false);
    f.fb = (float *)dalloc((size_t)px * 16);
    if (!f.rays || !f.fb) return -1;
    int nout = 0;
    if (camera_generate(r.eye, r.center, 0, r.up, r.aov, r.width, r.height, &v->local_bb, &v->global_bb,
                        f.rays->ispc, px, 1, &nout) != 0)
      return -1;
    f.n = nout;
    HIPCHK(launch_fill_f32(f.fb, 4LL * px, 0.f, S()));
  } else {
    f.n = (int)r.nrays;
    f.rays = raylist_new(f.n, 0, 0, false);
    f.xyz = (float *)dalloc((size_t)std::max<long long>(1, r.capacity) * 12);
    f.cnt = (int *)dalloc(sizeof(int));
    if (!f.rays || !f.xyz || !f.cnt) return -1;
    HIPCHK(launch_random_rays(f.rays->ispc, f.n, r.seed, f.n, RAY_PRIMARY, v->local_bb.lower, v->local_bb.upper, S()));
    HIPCHK(launch_set_count(f.cnt, 0, S()));
  }
  return 0;
}

// results to the host buffer of the request
static int flight_finish(const gxy_request &r, Flight &f, gxy_request_result &res) {
  if (r.kind == 0) {
    res.rays = f.n;
    if (r.out) HIPCHK(hipMemcpyAsync(r.out, f.fb, (size_t)r.width * r.height * 16, hipMemcpyDeviceToHost, S()));
  } else {
    int c = 0;
    HIPCHK(hipMemcpyAsync(&c, f.cnt, sizeof c, hipMemcpyDeviceToHost, S()));
    HIPCHK(hipStreamSynchronize(S()));
    res.particles = c;
    res.rays = f.n;
    const long long k = std::min<long long>(c, r.capacity);
    if (r.out && k > 0) HIPCHK(hipMemcpyAsync(r.out, f.xyz, (size_t)k * 12, hipMemcpyDeviceToHost, S()));
  }
  HIPCHK(hipStreamSynchronize(S()));
  return 0;
}

// Ends an admitted request: its batch list removed, its buffers freed, its
// status set (under the server lock).
static void flight_end(ServerObj *S_, size_t k, int status) {
  Flight f = S_->fl[k];
  if (f.list >= 0) gxy_batch_remove_list(S_->batch, f.list);
  flight_free(f);
  std::lock_guard<std::mutex> lk(S_->mu);
  ReqRecord &rr = S_->recs[f.req];
  rr.res.status = status;
  rr.res.finished = since(S_);
  S_->fl.erase(S_->fl.begin() + k);
  S_->idle_cv.notify_all();
}

// Every admitted request ended as failed (after an error of the batch), so
// that no list of a failed step is stepped again and no buffer leaks.
static void serve_abort(ServerObj *S_) {
  while (!S_->fl.empty()) flight_end(S_, S_->fl.size() - 1, RQ_FAILED);
}

// Admission: queued requests whose arrival has come, in order, while a slot is
// free and the device bytes held since serving started plus the request's
// estimate fit the budget (an estimate over the budget alone was rejected at
// submission).  A request whose set-up fails is failed alone.
// nothing queued, being admitted or in flight (S_->mu held)
static bool server_idle(const ServerObj *S_) { return S_->queue.empty() && S_->admitting == 0 && S_->fl.empty(); }

static void serve_admit(ServerObj *S_) {
  BatchObj *B = (BatchObj *)S_->batch;
  for (;;) {
    int id;
    {
      std::lock_guard<std::mutex> lk(S_->mu);
      if (S_->queue.empty() || (int)S_->fl.size() >= S_->slots) return;
      id = S_->queue.front();
      ReqRecord &rr = S_->recs[id];
      if (rr.req.arrival > since(S_)) return;
      const long long est = request_bytes(S_, rr.req);
      if (!S_->fl.empty() && (long long)G.bytes_live - S_->base_bytes + est > S_->budget) return;   // wait for memory
      S_->queue.pop_front();
      if (rr.cancel) {
        rr.res.status = RQ_CANCELLED;
        rr.res.finished = since(S_);
        S_->idle_cv.notify_all();
        continue;
      }
      rr.res.status = RQ_RUNNING;
      rr.res.admitted = since(S_);
      S_->admitting++;
    }
    const gxy_request r = S_->recs[id].req;
    Flight f;
    f.req = id;
    f.bytes = request_bytes(S_, r);
    bool ok = flight_make(S_, r, f) == 0;
    if (ok) {
      f.list = r.kind == 0 ? gxy_batch_add_list(B, r.dataset, BM_RENDER, f.n, &f.rays->ispc, f.fb, r.width, r.height,
                                                nullptr, 0, nullptr)
                           : gxy_batch_add_list(B, r.dataset, BM_SAMPLE, f.n, &f.rays->ispc, nullptr, 0, 0, f.xyz,
                                                (int)std::min<long long>(r.capacity, INT32_MAX), f.cnt);
      ok = f.list >= 0;
    }
    std::lock_guard<std::mutex> lk(S_->mu);
    S_->admitting--;
    S_->fl.push_back(f);
    if (!ok) {
      S_->fl.pop_back();
      flight_free(f);
      ReqRecord &rr = S_->recs[id];
      rr.res.status = RQ_FAILED;
      rr.res.finished = since(S_);
      S_->idle_cv.notify_all();
    }
    S_->max_in_flight = std::max(S_->max_in_flight, (int)S_->fl.size());
  }
}

// Cancellations asked for since the last step: admitted requests leave the
// batch now (their partial outputs are dropped).
static void serve_cancel(ServerObj *S_) {
  for (size_t k = 0; k < S_->fl.size();) {
    bool c;
    {
      std::lock_guard<std::mutex> lk(S_->mu);
      c = S_->recs[S_->fl[k].req].cancel;
    }
    if (c) flight_end(S_, k, RQ_CANCELLED);
    else k++;
  }
}

// One continuous-batching step over the admitted requests, then every request
// with nothing left in the batch is finished.  Returns -1 on a batch error
// (every admitted request is then failed and released).
static int serve_step(ServerObj *S_) {
  BatchObj *B = (BatchObj *)S_->batch;
  if (S_->fl.empty()) return 0;
  long long q = 0;
  if (gxy_batch_step(B, &q) != 0) {
    serve_abort(S_);
    return -1;
  }
  S_->peak = std::max(S_->peak, (long long)G.bytes_live - S_->base_bytes);
  std::vector<int> pend(B->lists.size(), 0);
  if (gxy_batch_pending(B, pend.data(), (int)pend.size()) != 0) {
    serve_abort(S_);
    return -1;
  }
  for (size_t k = 0; k < S_->fl.size();) {
    Flight &f = S_->fl[k];
    if (pend[f.list] == 0) {
      gxy_request_result tmp;
      {
        std::lock_guard<std::mutex> lk(S_->mu);
        tmp = S_->recs[f.req].res;
      }
      const int rc = flight_finish(S_->recs[f.req].req, f, tmp);
      {
        std::lock_guard<std::mutex> lk(S_->mu);
        gxy_request_result &res = S_->recs[f.req].res;
        res.rays = tmp.rays;
        res.particles = tmp.particles;
      }
      flight_end(S_, k, rc == 0 ? RQ_DONE : RQ_FAILED);
    } else {
      k++;
    }
  }
  return 0;
}

static int server_ready(ServerObj *S_) {
  if (!S_->lights) return fail("gxy_server_set_lighting not called");
  return 0;
}

// A new request record (status queued, or rejected when it cannot be taken).
static int record_request(ServerObj *S_, const gxy_request &r, double arrival) {
  ReqRecord rr;
  rr.req = r;
  rr.req.arrival = arrival;
  memset(&rr.res, 0, sizeof rr.res);
  rr.res.status = RQ_QUEUED;
  const int id = (int)S_->recs.size();
  if (request_invalid(S_, r)) {
    rr.res.status = RQ_REJECTED;
    rr.res.admitted = rr.res.finished = arrival;
    S_->recs.push_back(rr);
  } else {
    S_->recs.push_back(rr);
    S_->queue.push_back(id);
  }
  return id;
}

// The live service: its own execution context (stream), admission, cancel and
// step until asked to stop; a batch error fails the admitted requests and the
// ones still queued, and is reported by gxy_server_drain / gxy_server_stop.
static void service_loop(ServerObj *S_) {
  void *ctx = gxy_context_create();
  if (!ctx || gxy_context_bind(ctx) != 0) {
    std::lock_guard<std::mutex> lk(S_->mu);
    S_->loop_error = -1;
    S_->loop_msg = "service thread: no execution context";
    S_->running = false;
    S_->idle_cv.notify_all();
    return;
  }
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S_->mu);
      if (S_->stop && server_idle(S_)) break;
      if (S_->fl.empty()) {
        // nothing admitted: sleep until a request arrives (or its arrival time)
        bool wait = S_->queue.empty() || S_->recs[S_->queue.front()].req.arrival > since(S_);
        if (wait && !S_->stop) S_->cv.wait_for(lk, std::chrono::milliseconds(2));
      }
    }
    serve_cancel(S_);
    serve_admit(S_);
    if (serve_step(S_) != 0) {
      std::lock_guard<std::mutex> lk(S_->mu);
      S_->loop_error = -1;
      S_->loop_msg = t_err;
      for (int id : S_->queue) {
        S_->recs[id].res.status = RQ_FAILED;
        S_->recs[id].res.finished = since(S_);
      }
      S_->queue.clear();
      S_->idle_cv.notify_all();
    }
  }
  gxy_context_bind(nullptr);
  gxy_context_destroy(ctx);
  std::lock_guard<std::mutex> lk(S_->mu);
  S_->running = false;
  S_->idle_cv.notify_all();
}

}  // namespace gxy

GXY_EXPORT void *gxy_server_create(long long budget, int slots, int chunk) {
  INIT_OR(nullptr);
  if (budget <= 0 || slots <= 0) {
    fail("server needs a positive HBM budget and slot count");
    return nullptr;
  }
  void *b = gxy_batch_create(chunk);
  if (!b) return nullptr;
  ServerObj *S_ = new ServerObj;
  S_->budget = budget;
  S_->slots = slots;
  S_->batch = b;
  S_->t0 = std::chrono::steady_clock::now();
  std::lock_guard<std::recursive_mutex> lk(G.mu);
  g_servers.insert(S_);
  return S_;
}

// Stops the service thread: with cancel, every request still queued or
// admitted is cancelled first; otherwise the thread finishes them.
static int server_stop(ServerObj *S_, bool cancel) {
  {
    std::lock_guard<std::mutex> lk(S_->mu);
    if (!S_->th.joinable()) return 0;
    if (cancel)
      for (auto &rr : S_->recs)
        if (rr.res.status == RQ_QUEUED || rr.res.status == RQ_RUNNING) rr.cancel = true;
    S_->stop = true;
    S_->cv.notify_all();
  }
  S_->th.join();
  std::lock_guard<std::mutex> lk(S_->mu);
  S_->stop = false;
  return S_->loop_error;
}

GXY_EXPORT int gxy_server_destroy(void *s) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  server_stop(S_, true);
  {
    std::lock_guard<std::recursive_mutex> lk(G.mu);
    g_servers.erase(S_);
  }
  int rc = gxy_batch_destroy(S_->batch);
  delete S_;
  return rc;
}

GXY_EXPORT void *gxy_server_batch(void *s) {
  INIT_OR(nullptr);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) {
    fail("not a server handle");
    return nullptr;
  }
  return S_->batch;
}

GXY_EXPORT int gxy_server_add_dataset(void *s, const Visualization_ispc *vis, float eps) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  if (S_->th.joinable()) return fail("add datasets before gxy_server_start");
  int k = gxy_batch_add_dataset(S_->batch, vis, eps);
  if (k < 0) return -1;
  S_->vis.push_back(vis);
  S_->eps.push_back(eps);
  return k;
}

GXY_EXPORT int gxy_server_set_lighting(void *s, const Lighting_ispc *lights, float eps) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  if (S_->th.joinable()) return fail("set the lighting before gxy_server_start");
  if (gxy_batch_set_lighting(S_->batch, lights, eps) != 0) return -1;
  S_->lights = lights;
  S_->leps = eps;
  S_->nsp = (int)lights->n_ao_rays + (lights->do_shadows ? (int)lights->nLights : 0);
  return 0;
}

GXY_EXPORT int gxy_server_start(void *s) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  if (server_ready(S_) != 0) return -1;
  std::lock_guard<std::mutex> lk(S_->mu);
  if (S_->th.joinable()) return fail("server already started");
  HIPCHK(hipDeviceSynchronize());   // datasets uploaded on other streams are complete
  S_->t0 = std::chrono::steady_clock::now();
  S_->base_bytes = (long long)G.bytes_live;
  S_->loop_error = 0;
  S_->loop_msg.clear();
  S_->running = true;
  S_->th = std::thread(service_loop, S_);
  return 0;
}

GXY_EXPORT int gxy_server_submit(void *s, const gxy_request *req) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  if (!req) return fail("null request");
  std::lock_guard<std::mutex> lk(S_->mu);
  if (!S_->th.joinable() || !S_->running) return fail("server not started");
  const int id = record_request(S_, *req, since(S_));
  S_->cv.notify_all();
  return id;
}

GXY_EXPORT int gxy_server_poll(void *s, int id, gxy_request_result *res) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  std::lock_guard<std::mutex> lk(S_->mu);
  if (id < 0 || id >= (int)S_->recs.size()) return fail("request id %d unknown", id);
  if (res) *res = S_->recs[id].res;
  return S_->recs[id].res.status;
}

GXY_EXPORT int gxy_server_cancel(void *s, int id) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  std::lock_guard<std::mutex> lk(S_->mu);
  if (id < 0 || id >= (int)S_->recs.size()) return fail("request id %d unknown", id);
  ReqRecord &rr = S_->recs[id];
  if (rr.res.status != RQ_QUEUED && rr.res.status != RQ_RUNNING) return 0;   // already final
  rr.cancel = true;
  S_->cv.notify_all();
  return 0;
}

GXY_EXPORT int gxy_server_drain(void *s, double timeout_s) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  std::unique_lock<std::mutex> lk(S_->mu);
  if (!S_->th.joinable()) return fail("server not started");
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  const bool drained = S_->idle_cv.wait_until(lk, until, [&] {
    return server_idle(S_) || !S_->running;
  });
  if (S_->loop_error) return fail("server loop failed: %s", S_->loop_msg.c_str());
  if (drained) return 0;
  // the watchdog: cancel what is left and report the timeout
  int left = (int)S_->queue.size() + (int)S_->fl.size();
  for (auto &rr : S_->recs)
    if (rr.res.status == RQ_QUEUED || rr.res.status == RQ_RUNNING) rr.cancel = true;
  S_->cv.notify_all();
  const bool emptied = S_->idle_cv.wait_for(lk, std::chrono::seconds(60), [&] {
    return server_idle(S_) || !S_->running;
  });
  return fail("drain timed out after %.3f s with %d requests left (%s)", timeout_s, left,
              emptied ? "cancelled" : "cancellation still pending");
}

GXY_EXPORT int gxy_server_stop(void *s) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  const int rc = server_stop(S_, false);
  if (rc != 0) return fail("server loop failed: %s", S_->loop_msg.c_str());
  return 0;
}

GXY_EXPORT int gxy_server_run(void *s, const gxy_request *reqs, int n, int mode, gxy_request_result *res,
                              gxy_serve_stats *st) {
  INIT_OR(-1);
  ServerObj *S_ = lookup<ServerObj>(g_servers, s);
  if (!S_) return fail("not a server handle");
  if (server_ready(S_) != 0) return -1;
  if (S_->th.joinable()) return fail("the service thread is running (gxy_server_stop first)");
  for (int k = 1; k < n; k++)
    if (reqs[k].arrival < reqs[k - 1].arrival) return fail("requests must be sorted by arrival");
  gxy_serve_stats out;
  memset(&out, 0, sizeof out);
  S_->recs.clear();
  S_->queue.clear();
  S_->t0 = std::chrono::steady_clock::now();
  S_->base_bytes = (long long)G.bytes_live;
  S_->peak = 0;
  S_->max_in_flight = 0;
  for (int k = 0; k < n; k++) record_request(S_, reqs[k], reqs[k].arrival);
  auto wait_until = [&](double t) {
    for (;;) {
      double d = t - since(S_);
      if (d <= 0) return;
      std::this_thread::sleep_for(std::chrono::microseconds(d > 0.002 ? 1000 : 50));
    }
  };
  int rc = 0;
  if (mode == 1) {
    // one request at a time through the per-list loop
    const int nb[6] = {-1, -1, -1, -1, -1, -1};
    for (int k = 0; k < n; k++) {
      ReqRecord &rr = S_->recs[k];
      if (rr.res.status != RQ_QUEUED) continue;
      const gxy_request &r = rr.req;
      wait_until(r.arrival);
      rr.res.admitted = since(S_);
      Flight f;
      int frc = flight_make(S_, r, f);
      gxy_frame_stats fs;
      memset(&fs, 0, sizeof fs);
      if (frc == 0) {
        if (r.kind == 0) {
          frc = gxy_process_raylist((void *)S_->vis[r.dataset], (void *)S_->lights, f.n, &f.rays->ispc, nb,
                                    S_->eps[r.dataset], 1 << 30, f.fb, r.width, r.height, &fs);
        } else {
          int c = 0;
          frc = gxy_sample_raylist((void *)S_->vis[r.dataset], f.n, &f.rays->ispc, 1 << 30, f.xyz, (int)r.capacity,
                                   &c, &fs);
          if (frc == 0 && hipMemcpyAsync(f.cnt, &c, sizeof c, hipMemcpyHostToDevice, S()) != hipSuccess) frc = -1;
        }
      }
      S_->peak = std::max(S_->peak, (long long)G.bytes_live - S_->base_bytes);
      if (frc == 0) frc = flight_finish(r, f, rr.res);
      flight_free(f);
      rr.res.status = frc == 0 ? RQ_DONE : RQ_FAILED;
      rr.res.finished = since(S_);
      if (frc == 0) {
        out.rays += fs.ray_passes;
        out.samples += fs.samples;
        out.passes += fs.passes;
      }
    }
    S_->max_in_flight = n > 0 ? 1 : 0;
  } else {
    // continuous batching: requests join the batch as they arrive
    BatchObj *B = (BatchObj *)S_->batch;
    gxy_batch_stats b0;
    if (gxy_batch_get_stats(B, &b0) != 0) return -1;
    const int passes0 = B->passes_run;
    for (;;) {
      serve_admit(S_);
      if (S_->fl.empty()) {
        if (S_->queue.empty()) break;
        wait_until(S_->recs[S_->queue.front()].req.arrival);
        continue;
      }
      if (serve_step(S_) != 0) {
        rc = -1;
        for (int id : S_->queue) S_->recs[id].res.status = RQ_FAILED;
        S_->queue.clear();
        break;
      }
    }
    gxy_batch_stats b1;
    if (gxy_batch_get_stats(B, &b1) == 0) {
      out.rays = b1.rays - b0.rays;
      out.samples = b1.samples - b0.samples;
      out.passes = B->passes_run - passes0;
    }
  }
  double first = n > 0 ? reqs[0].arrival : 0, last = 0;
  for (int k = 0; k < n; k++) {
    res[k] = S_->recs[k].res;
    last = std::max(last, res[k].finished);
  }
  out.wall = last - first;
  out.peak_bytes = S_->peak;
  out.max_in_flight = S_->max_in_flight;
  if (st) *st = out;
  return rc;
}
