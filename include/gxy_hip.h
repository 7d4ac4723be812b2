/*
 * gxy_hip.h — C-ABI of the MI355X ray-processing path (libgxy_hip.so).
 *
 * The first block mirrors the ISPC exports that Galaxy's C++ renderer binds
 * (generated headers TraceRays_ispc.h / Lighting_ispc.h / SamplerTraceRays_ispc.h,
 * all C linkage).  The argument structs keep the reference layouts byte for byte;
 * the difference is where the memory lives: every array a RayList_ispc points at
 * is DEVICE memory (HBM), and the opaque OSPRay pointers of the reference
 * (Model*, Volume*, TransferFunction*) are replaced by handles returned from the
 * gxy_scene_* / gxy_volume_* / gxy_tf_* / gxy_model_* calls below, whose data is
 * resident on the GPU.
 *
 * Errors: the reference has no return codes (it prints and carries on).  Every
 * entry point here that can fail returns 0 on success and a negative value on
 * failure, and gxy_last_error() returns the message.  Entry points that mirror a
 * reference `void` export keep the `void`/`void*` return and record the error.
 *
 * Streams: all work is enqueued on the library's stream for the calling thread
 * (gxy_set_stream overrides it).  Nothing blocks unless the entry point returns
 * host-visible counts, which it documents.
 */
#ifndef GXY_HIP_H
#define GXY_HIP_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ray flags: src/renderer/RayFlags.h:28-36 ---------------------------- */
#define RAY_PRIMARY            0x001
#define RAY_SHADOW             0x002
#define RAY_AO                 0x004
#define RAY_EMPTY              0x008
#define RAY_SURFACE            0x001
#define RAY_OPAQUE             0x002
#define RAY_BOUNDARY           0x004
#define RAY_TIMEOUT            0x008
#define RAY_EXTERNAL_BOUNDARY  0x010

/* ---- classification sentinels: src/renderer/Renderer.cpp:78-82 ----------- */
#define GXY_TERMINATED     (-1)
#define GXY_DROP_ON_FLOOR  (-2)
#define GXY_KEEP_HERE      (-3)
#define GXY_UNDETERMINED   (-4)
#define GXY_NO_NEIGHBOR    (-1)

/* ---- RayList_ispc: src/renderer/Rays.ih:21-48 (25 SoA arrays, 100 B/ray) --
 * Arrays are laid out inside one block exactly as RayList::setup_ispc_pointers
 * does (src/renderer/Rays.cpp:173-204): a 64-byte header, then 20 float arrays
 * and 5 int arrays, each aligned_size = ceil(n/16)*16 elements long. */
typedef struct RayList_ispc {
  float *ox, *oy, *oz;
  float *dx, *dy, *dz;
  float *nx, *ny, *nz;
  float *sample;
  float *r, *g, *b, *o;
  float *sr, *sg, *sb, *so;
  float *t, *tMax;
  int   *x, *y, *type, *term, *classification;
} RayList_ispc;

/* RayList block header (src/renderer/Rays.h:49-59): three 8-byte Keys, then
 * size, frame, aligned_size, id, type; padded to 64 bytes. */
typedef struct gxy_raylist_hdr {
  int64_t rendererKey, renderingKey, renderingSetKey;
  int32_t size, frame, aligned_size, id, type;
  int32_t pad_[5];
} gxy_raylist_hdr;

/* Renderer::SendPixelsMsg contents (src/renderer/Renderer.h:246-262,
 * src/renderer/Pixel.h:33-37): this 32-byte header (two 8-byte Keys, count,
 * frame, sending rank, tail padding to 8-byte alignment), then count pixels of
 * 24 bytes. */
typedef struct gxy_pixels_header {
  int64_t rkey, rskey;
  int32_t count, frame, source;
  int32_t pad_;
} gxy_pixels_header;
typedef struct gxy_pixel {
  int32_t x, y;
  float r, g, b, o;
} gxy_pixel;

/* ---- Lighting_ispc: src/renderer/Lighting.ih:23-32 ------------------------ */
typedef struct Lighting_ispc {
  int      n_ao_rays;
  int      do_shadows;
  float    ao_radius;
  float    Ka, Kd;
  int     *types;    /* host array, nLights entries: 0 dir, 1 camera-rel, 2 point */
  float   *lights;   /* host array, nLights * 3 floats (vec3f)                     */
  uint32_t nLights;
} Lighting_ispc;

/* ---- visualization structs: Vis.ih:23-26, MappedVis.ih:26-30,
 *      VolumeVis.ih:26-37, Visualization.ih:30-47 ---------------------------- */
typedef struct Vis_ispc { void *data; } Vis_ispc;              /* data: gxy volume handle */
typedef struct MappedVis_ispc { Vis_ispc vis; void *transferFunction; } MappedVis_ispc; /* gxy TF handle */
typedef struct VolumeVis_ispc {
  MappedVis_ispc vis;
  int    nSlices;
  float *slices;      /* host array of nSlices vec4f planes (nx, ny, nz, d) */
  int    nIsovalues;
  float *isovalues;   /* host array */
  bool   volume_render;
} VolumeVis_ispc;
typedef struct gxy_box3f { float lower[3], upper[3]; } gxy_box3f;
typedef struct Visualization_ispc {
  void            *model;        /* gxy model handle (triangles + spheres), or NULL */
  int32_t          nVolumeVis;
  VolumeVis_ispc **volumeVis;
  int32_t          nMappedVis;
  MappedVis_ispc **mappedVis;
  gxy_box3f        global_bb;
  gxy_box3f        local_bb;
} Visualization_ispc;

/* ==== reference entry points (src/renderer/TraceRays.ispc) ================= */

/* TraceRays.ispc:58-77 */
void *TraceRays_allocate(void);
void  TraceRays_initialize(void *self);
void  TraceRays_destroy(void *self);

/* TraceRays.ispc:326-623.  Traces nRaysIn device-resident rays in place:
 * box clip, slices, geometry (BVH), volume march with TF compositing and
 * isosurfaces; writes r g b o, t, term and, on a surface hit, sr sg sb so and
 * nx ny nz.  Returns NULL like the reference; errors are reported by
 * gxy_last_error(), which this and the other TraceRays_* / SamplerTraceRays_*
 * entry points clear on entry (empty after a call that succeeded). */
void *TraceRays_TraceRays(void *self, void *vis, int nRaysIn, void *raysIn, float epsilon);

/* TraceRays.ispc:625-733.  offsets: device int array, one per input ray. */
void TraceRays_generateAORays(void *self, void *lights, int nRaysIn, void *raysIn,
                              int *offsets, void *raysOut, float epsilon);
/* TraceRays.ispc:763-857 */
void TraceRays_generateShadowRays(void *self, void *lights, int nRaysIn, void *raysIn,
                                  int *offsets, void *raysOut, float epsilon);
/* TraceRays.ispc:735-761 */
void TraceRays_ambientLighting(void *self, void *lights, int nRaysIn, void *raysIn);
/* TraceRays.ispc:859-923 */
void TraceRays_diffuseLighting(void *self, void *lights, int nRaysIn, void *raysIn);

/* src/sampler/SamplerTraceRays.ispc:128-222 (iso-threshold sampler; the
 * isovalue of each VolumeVis is its isovalues[0]). */
void *SamplerTraceRays_SamplerTraceRays(void *self, void *vis, int nRaysIn, void *raysIn);

/* src/renderer/Lighting.ispc:23-134 */
void *Lighting_allocate(void);
void  Lighting_initialize(void *self);
void  Lighting_destroy(void *self);
void  Lighting_SetLights(void *self, int n, float *l, int *t);
void  Lighting_GetLights(void *self, int *n, float **l, int **t);
void  Lighting_SetK(void *self, float ka, float kd);
void  Lighting_GetK(void *self, float *ka, float *kd);
void  Lighting_SetAO(void *self, int n, float r);
void  Lighting_GetAO(void *self, int *n, float *r);
void  Lighting_SetShadowFlag(void *self, int b);
void  Lighting_GetShadowFlag(void *self, int *b);

/* ==== runtime, memory and errors =========================================== */
int         gxy_init(int device);           /* selects the device; idempotent */
const char *gxy_last_error(void);
int         gxy_device_count(void);
int         gxy_synchronize(void);
int         gxy_set_stream(void *hip_stream); /* NULL restores the library stream */
/* Execution contexts: each has its own HIP stream (default priority), device
 * counters and block cache.  A context bound to a host thread carries all of
 * that thread's work, so several threads can run frames concurrently on one
 * GPU (the serving loop does this).  NULL unbinds (back to the default). */
void       *gxy_context_create(void);
int         gxy_context_bind(void *ctx);
int         gxy_context_destroy(void *ctx);
void       *gxy_get_stream(void);
void       *gxy_device_alloc(size_t bytes);   /* cached pool; 256-byte aligned */
int         gxy_device_free(void *p);
int         gxy_copy_to_device(void *dst, const void *src, size_t bytes);
int         gxy_copy_to_host(void *dst, const void *src, size_t bytes);
int         gxy_copy_device(void *dst, const void *src, size_t bytes);
int         gxy_memset(void *dst, int value, size_t bytes);
size_t      gxy_device_bytes_in_use(void);
/* Early ray termination: TraceRays.ispc:466,550,581 treat a ray as opaque when
 * min(r,g,b) >= 1 or alpha > 0.999.  The alpha threshold is a process-wide
 * setting here (default 0.999, the reference value). */
int         gxy_set_opacity_threshold(float alpha);
float       gxy_get_opacity_threshold(void);
/* Launch configuration of the volume-only trace (no reference counterpart):
 * kind 0 = one ray per work-item; 15 / 16 / 17 / 18 = the same kernel compiled
 * for at least 5 / 6 / 7 / 8 waves per SIMD (17 and 18 for one bricked volume
 * with at most one isovalue, 16 otherwise), where `chunk` 64 or 128 selects that
 * many work-items per workgroup for kind 16 on one bricked volume with at most
 * one isovalue (256 otherwise); kind 1 = persistent lanes, each wave owning
 * `chunk` consecutive rays (multiple of 64) and refilling idle lanes once
 * `refill` (1..64) of them are idle.  Results are identical for every setting;
 * galaxy_amd/autotune.py picks the fastest. */
int         gxy_set_trace_config(int kind, int chunk, int refill);
/* BVH traversal stack of geometry-only traces (no reference counterpart):
 * 0 = 64 entries per ray in private memory; 16 / 24 / 32 = that many entries
 * in LDS, used for models whose tree depth + 2 fits (others keep the private
 * stack).  Results are identical for every setting. */
int         gxy_set_geometry_stack(int entries);
int         gxy_get_geometry_stack(void);
/* Occlusion rays (default 1): in the rank loops (gxy_render_frame,
 * gxy_process_raylist, the batched and served forms) a shadow or AO ray of a
 * scene without volumes ends its geometry search at its first hit: Classify
 * terminates it for any hit and its colour does not depend on which, and the
 * loops never read its t.  TraceRays_TraceRays always finds the closest hit.
 * 0: closest hit everywhere (same frames and counts). */
int gxy_set_occlusion_rays(int on);
int gxy_get_occlusion_rays(void);
/* BVH traversed by geometry traces (no reference counterpart): 2 (default) =
 * the binary LBVH, 4 = the 4-wide trees every model commit also builds
 * (128-byte nodes, about half the depth).  Hits are identical either way.
 * The binary walk is faster on scenes of several objects and on path lines
 * (1.3-1.8x on MI355X), the 4-wide one by about 9% on one closed mesh. */
int         gxy_set_bvh_width(int width);
int         gxy_get_bvh_width(void);
/* Occupancy of the binary walks with the private stack, path lines included
 * (no reference counterpart): 0 (default) = the compiler's register
 * allocation (70 VGPRs, 7 waves per SIMD; 76 and 6 with path lines), 8 = 8
 * waves per SIMD (64 VGPRs, some spills; 4 % faster on the 11.8 M-triangle
 * nine-icosphere frame, 5 % on the path-line frame, 5 % slower on a
 * 20 K-triangle mesh).  Results are identical. */
int         gxy_set_geometry_waves(int waves);
int         gxy_get_geometry_waves(void);
int         gxy_get_trace_config(int *kind, int *chunk, int *refill);
size_t      gxy_device_bytes_total(void);

/* Kernel timing: launches record HIP events on their stream when enabled
 * (on = 1: every kernel; 2: the trace kernels only, "trace"; 0: off).
 * gxy_timing_get returns total ms and launch count for the named kernel. */
int gxy_timing_enable(int on);
int gxy_timing_reset(void);
int gxy_timing_get(const char *kernel, double *total_ms, int *launches);

/* ==== device RayLists (Rays.h:43-227, Rays.cpp:48-286) ====================== */
/* Allocates one device block with the reference layout and fills *ispc with the
 * 25 device array pointers.  Returns an opaque handle (NULL on failure). */
void *gxy_raylist_create(int nrays, int frame, int type, RayList_ispc *ispc);
int   gxy_raylist_destroy(void *rl);
int   gxy_raylist_size(void *rl);
int   gxy_raylist_aligned_size(void *rl);
/* RayList::Split (Rays.cpp:206-243): copies rays [start, start+count) into a new
 * list; RayList::Truncate (Rays.cpp:245-286). */
void *gxy_raylist_split_part(void *rl, int start, int count, RayList_ispc *ispc);
int   gxy_raylist_truncate(void *rl, int n, RayList_ispc *ispc);
/* Whole-block host<->device copies for serialization (the wire format is the
 * block itself: header + arrays). */
size_t gxy_raylist_bytes(void *rl);
int    gxy_raylist_download(void *rl, void *host_block);
void  *gxy_raylist_upload(const void *host_block, RayList_ispc *ispc);

/* ==== scene objects ========================================================== */
/* Structured volume (OsprayVolume.cpp:25-52 over Volume.cpp:174-392).  dims,
 * origin and spacing describe the ghosted brick; voxels is a HOST float array
 * (x fastest) copied to HBM, or NULL to allocate without data.
 * samplingRate as OSPRay's "samplingRate" (default 1). samplingStep is the
 * minimum grid spacing. */
void *gxy_volume_create(const int dims[3], const float origin[3], const float spacing[3],
                        const float *voxels, float samplingRate);
int   gxy_volume_destroy(void *vol);
float *gxy_volume_device_data(void *vol);
int   gxy_volume_set_sampling_rate(void *vol, float rate);
/* Fills the volume with seeded 4-octave gradient noise normalised to [0,1]. */
int   gxy_volume_fill_noise(void *vol, uint32_t seed, int octaves, float base_frequency);
int   gxy_volume_download(void *vol, float *host);
/* Voxel layout the trace kernels read (no reference counterpart): 0 = the
 * linear brick only; 4 / 8 = also a bricked copy (B^3 cells per brick stored
 * with a one-voxel overlap, (B+1)^3 voxels padded to 128 bytes: 1.4x the
 * voxels for B = 8, 2x for B = 4) that single-volume traces and samplers read.
 * The copy is rebuilt from the linear voxels by this call and by
 * gxy_volume_fill_noise; call it again after writing through
 * gxy_volume_device_data.  Samples are identical in every layout. */
int   gxy_volume_set_layout(void *vol, int brick);
int   gxy_volume_get_layout(void *vol);

/* Transfer function: the 256-entry resample of MappedVis::local_commit
 * (MappedVis.cpp:277-338).  colormap: n x [d, r, g, b]; opacitymap: m x [d, o];
 * data_range: NULL or [lo, hi]. */
void *gxy_tf_create_from_maps(const float *colormap, int ncolors,
                              const float *opacitymap, int nopacities,
                              const float *data_range);
/* Direct tables (colors: nc x 3, opacities: no), value range [lo, hi]. */
void *gxy_tf_create(const float *colors, int nc, const float *opacities, int no, float lo, float hi);
int   gxy_tf_destroy(void *tf);
int   gxy_tf_get_tables(void *tf, float *colors256x3, float *opac256, float *range2);

/* Geometry model (OSPModel + Embree scene): triangle meshes
 * (DataDrivenTriangleMesh.cpp:43-180) and data-driven spheres
 * (DataDrivenSpheres.ispc:178-278).  Arrays are host arrays; the model copies
 * them to HBM.  gxy_model_commit builds the device BVHs. */
void *gxy_model_create(void);
int   gxy_model_destroy(void *model);
int   gxy_model_add_triangles(void *model, const float *vertices, int nvertices,
                              const uint32_t *indices, int ntriangles,
                              const float *normals, const float *vertex_data, void *tf);
int   gxy_model_add_spheres(void *model, const float *centers, const float *data, int n,
                            float radius0, float radius1, float value0, float value1, void *tf);
/* Data-driven path lines (DataDrivenPathLines.cpp:46-160,
 * DataDrivenPathLines.ispc:212-269): segment i joins vertex indices[i] and
 * indices[i] + 1; per-vertex data gives the radius (radius0..radius1 over
 * value0..value1, MAP_RADIUS); consecutive segments form round cubic Bezier
 * curves with the reference's control points; the hit colour is TF(value)
 * of the radius at the curve parameter.  Each curve segment is traced as 8
 * round cones between curve samples (the reference intersects the curves in
 * Embree, which is not available: that surface is approximated).  Returns
 * the geomID; primID of a hit = segment index. */
int   gxy_model_add_pathlines(void *model, const float *vertices, const float *data, int nvertices,
                              const int *indices, int nsegments, float radius0, float radius1,
                              float value0, float value1, void *tf);
int   gxy_model_commit(void *model);
/* BVH statistics after commit: binary inner nodes, primitives, levels of the
 * deepest tree (leaves included), and the SAH cost of the trees (each over its
 * root's surface area: node visits 1.2, primitive tests their set cost),
 * averaged with the primitive counts as weights. */
int   gxy_model_bvh_stats(void *model, int *nodes, int *leaves, int *depth, float *sah);
/* SAH cost of one primitive test, per kind (0 triangles, 1 spheres, 2
 * path-line pieces; defaults 2, 1, 8) against 1.2 for a node visit, used by
 * the next commits: it sets how readily the build makes leaves of several
 * primitives (up to 8).  Any value gives the same hits. */
int   gxy_set_bvh_prim_cost(int kind, float cost);
float gxy_get_bvh_prim_cost(int kind);
/* Side-band hit record of the last TraceRays_TraceRays on this model:
 * per ray (geomID, primID); device arrays of n ints each, or NULL to disable. */
int   gxy_model_set_hit_output(void *model, int *geom_ids, int *prim_ids);

/* ==== the pass around TraceRays (TraceRays.cpp:68-146, Renderer.cpp:304-691) */

/* TraceRays.cpp:89-117 on the device: exclusive offsets, AO block first, then
 * the shadow block.  ao_offsets / shadow_offsets are device int arrays of
 * nRays; n_ao / n_shadow receive the counts (this call synchronises). */
int gxy_spawn_offsets(int nRays, RayList_ispc *raysIn, int nAO, int nLights, int do_shadows,
                      int *ao_offsets, int *shadow_offsets, int *n_ao, int *n_shadow);
/* No counterpart in the reference.  gxy_spawn_offsets for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): same arguments and results, but every temporary lives in
 * one static device array per GPU (128 KiB, held until the process ends) and
 * the call allocates nothing once gxy_init (or the first call) has set the
 * library up.  At most GXY_NOALLOC_MAX_RAYS rays; larger lists fail with an
 * error.  Calls that overlap on one GPU share the array and race.
 * Synchronises, like gxy_spawn_offsets. */
#define GXY_NOALLOC_MAX_RAYS 4194048   /* 16383 workgroups of 256 rays */
int gxy_spawn_offsets_noalloc(int nRays, RayList_ispc *raysIn, int nAO, int nLights, int do_shadows,
                              int *ao_offsets, int *shadow_offsets, int *n_ao, int *n_shadow);

/* Galaxy's lighting model, a build option there (GXY_REVERSE_LIGHTING,
 * CMakeLists.txt:66, default ON) and one library setting here: 1 = reverse
 * lighting (the default), 0 = the direct model of TraceRays.cpp:133-141,
 * TraceRays.ispc:639-643,778-782, Renderer.cpp:365-447.  The TraceRays_*
 * lighting entry points, the frame loops and the batched form follow it (a
 * batch takes it when its lighting is set); gxy_classify and
 * gxy_assign_destinations fail when their reverse_lighting argument disagrees
 * with it. */
int gxy_set_reverse_lighting(int on);
int gxy_get_reverse_lighting(void);
/* AO ray directions (no counterpart in the reference's ABI; its
 * generateAORays, TraceRays.ispc:625-733, always uses the UV table): mode 0
 * (default) = the reference's 256-entry UV table indexed by its pixel hash
 * ((x*9949 + y*9613 + j*9151) >> 8) & 0xff; mode 1 = (U, V) drawn from a PCG32
 * stream (O'Neill 2014, pcg32_srandom_r(initstate = seed << 32 | j,
 * initseq = y << 32 | x)), the top 24 bits of two outputs over 2^24, for
 * workloads configured with seeded random AO directions.  The direction
 * formula (cosine-weighted, not renormalised) is the reference's either way.
 * Applies to every entry point that spawns AO rays. */
int gxy_set_ao_directions(int mode, unsigned seed);
int gxy_get_ao_directions(unsigned *seed);

/* Renderer::Classify (Renderer.cpp:304-421), reverse lighting on when
 * reverse_lighting != 0 (the reference default, CMakeLists.txt:66); it must
 * equal the library setting above. */
int gxy_classify(int nRays, RayList_ispc *rays, int reverse_lighting);
/* Renderer::AssignDestinations (Renderer.cpp:423-454) + Box::exit_face
 * (Box.cpp:84-97); neighbors: 6 ranks or -1. */
int gxy_assign_destinations(int nRays, RayList_ispc *rays, const gxy_box3f *local_box,
                            const int neighbors[6], int reverse_lighting);
/* HandleTerminatedRays + Rendering::AddLocalPixels (Renderer.cpp:456-502,
 * Rendering.cpp:125-153): fb[(y*width+x)*4 + c] += rgba for every TERMINATED
 * ray.  fb is a device float array of width*height*4. */
int gxy_accumulate_terminated(int nRays, RayList_ispc *rays, float *fb, int width, int height);
/* HandleTerminatedRays' remote path (Renderer.cpp:490-499) with
 * SendPixelsMsg::SendPixelsMsg and StashPixel (Renderer.h:267-305): writes the
 * message contents into msg (device memory, 8-byte aligned, capacity_bytes):
 * the header {rkey, rskey, count, frame, source}, then the pixels of the
 * TERMINATED rays in ray order.  *count_out = pixels.  Fails with nothing
 * written when 32 + 24*count exceeds capacity_bytes.  Synchronises. */
int gxy_pixels_pack(int nRays, RayList_ispc *rays, int64_t rkey, int64_t rskey, int frame, int source,
                    void *msg, size_t capacity_bytes, int *count_out);
/* SendPixelsMsg::Action + Rendering::AddLocalPixels (Renderer.h:313-336,
 * Rendering.cpp:104-153) for a message in device memory (bytes long): when
 * its frame is not older than *fb_frame, *fb_frame takes it and every pixel
 * is added to fb (width*height*4 floats, FP32 atomics); with kbuffer
 * (width*height device ints of per-pixel frame stamps, may be NULL) a pixel
 * whose stamp is older than the frame is cleared and stamped first.  Pixels
 * outside the framebuffer are skipped (the reference does not check them).
 * *applied = pixels added (0 for an older frame).  Synchronises. */
int gxy_pixels_apply(const void *msg, size_t bytes, float *fb, int width, int height, int *kbuffer,
                     int *fb_frame, int *applied);
/* processRays_task::work's split (Renderer.cpp:561-634): stable partition of the
 * rays by classification into ndest per-destination device lists plus keepers.
 * counts[0..ndest-1] = rays for destination rank d, counts[ndest] = keepers.
 * lists[d] may be NULL when only counts are wanted.  Synchronises. */
int gxy_partition_rays(int nRays, RayList_ispc *rays, int ndest,
                       RayList_ispc *const *lists, RayList_ispc *keepers, int *counts);

/* Camera::generate_initial_rays + SpawnRays (Camera.cpp:379-493, 528-829):
 * writes up to width*height primary rays (pixels whose ray meets the boxes)
 * into rays; returns the count through n_out (synchronises).  order: 0 = row
 * major as the reference, 1 = 8x8 pixel tiles (same ray set, better locality:
 * bands of two tile rows, each band's tiles column by column, so every four
 * consecutive tiles cover a 16x16 square; a last single tile row in order). */
int gxy_camera_generate(const float eye[3], const float center_or_dir[3], int is_dir,
                        const float up[3], float aov, int width, int height,
                        const gxy_box3f *local_box, const gxy_box3f *global_box,
                        RayList_ispc *rays, int capacity, int order, int *n_out);

/* Sampler::HandleTerminatedRays (Sampler.cpp:51-89): appends the position
 * o + t*d of every ray with term & RAY_SURFACE to the device particle array
 * xyz (capacity entries, 3 floats each) at *count (device int, atomically
 * advanced).  Order of particles within one call follows ray order. */
int gxy_sampler_collect(int nRays, RayList_ispc *rays, float *xyz, int capacity, int *count);

/* ==== one whole frame: the host loop of Renderer/RayQManager on the device == */
typedef struct gxy_frame_stats {
  long long primary_rays;      /* originated */
  long long secondary_rays;    /* spawned AO + shadow */
  long long ray_passes;        /* every TraceRays_TraceRays ray */
  long long terminated;        /* accumulated into the framebuffer */
  long long dropped;
  long long kept;              /* KEEP_HERE re-traces */
  long long outbound[6];       /* rays that left through face f to a neighbour */
  long long samples;           /* volume sample points taken (all volumes) */
  int       passes;            /* TraceRays launches */
  int       lists;             /* RayLists processed */
} gxy_frame_stats;

/* Renders one camera view of vis under lights into fb (device, w*h*4 floats,
 * accumulated into; clear it first).  Runs camera rays, then processes ray
 * lists until the queue drains: trace, spawn, classify, route, accumulate,
 * keepers and secondaries re-enqueued (priority: primary lists before
 * secondary ones, LIFO within a level, as Threading.cpp:176-233).  Lists larger
 * than max_rays_per_list are split (Renderer.cpp:675-690).  neighbors: 6 ranks
 * or -1; outbound rays are counted per face and their lists queued for
 * gxy_outbound_take.  epsilon: Renderer epsilon (0.001 default). */
int gxy_render_frame(void *vis, void *lights, const float eye[3], const float center[3],
                     const float up[3], float aov, int width, int height,
                     const int neighbors[6], float epsilon, int max_rays_per_list,
                     float *fb, gxy_frame_stats *stats);

/* Outbound queue: the rays each pass routed to a neighbour (exit face f with
 * neighbors[f] >= 0), one device RayList per pass and face, in pass order.
 * gxy_outbound_take hands ownership of the next list to the caller (free with
 * gxy_raylist_destroy); gxy_outbound_clear drops them all. */
int   gxy_outbound_count(int face);
void *gxy_outbound_take(int face, RayList_ispc *ispc, int *n);
int   gxy_outbound_clear(void);

/* Same loop starting from caller-supplied device rays (count n, not consumed). */
int gxy_process_raylist(void *vis, void *lights, int n, RayList_ispc *rays,
                        const int neighbors[6], float epsilon, int max_rays_per_list,
                        float *fb, int width, int height, gxy_frame_stats *stats);

/* Sampler loop (Sampler::Trace + KEEP_HERE re-trace): collects every isovalue
 * crossing of the n rays into xyz (device, capacity*3 floats); *count receives
 * the number of particles (synchronises). */
int gxy_sample_raylist(void *vis, int n, RayList_ispc *rays, int max_rays_per_list,
                       float *xyz, int capacity, int *count, gxy_frame_stats *stats);


/* ---- Batched form -------------------------------------------------------
 * No counterpart in the reference, which traces one RayList per thread-pool
 * task (Renderer.cpp:659-691).  Many uneven ray lists over many resident
 * datasets are traced pass by pass: each pass is one persistent trace launch
 * whose waves pop chunks of at most `chunk` rays of one list from a device
 * work queue, followed by a post-pass that applies, per list, exactly the
 * reference's per-list steps (lighting and secondary spawn of
 * TraceRays::Trace, Classify / AssignDestinations with no neighbours,
 * HandleTerminatedRays or Sampler::HandleTerminatedRays, KEEP_HERE re-trace)
 * and compacts spawned and kept rays into the next pass.  Every per-ray result
 * equals that of gxy_process_raylist / gxy_sample_raylist on the list alone;
 * framebuffer sums may differ in float addition order. */
typedef struct gxy_batch_stats {
  unsigned long long rays;        /* ray passes traced */
  unsigned long long samples;     /* volume sample points */
  unsigned long long spawned;     /* shadow + AO rays generated */
  unsigned long long terminated;
  unsigned long long dropped;
  unsigned long long kept;        /* rays re-queued as KEEP_HERE (sampler rays follow up to 4096
                                     crossings within one pass before they are re-queued) */
  int passes;                     /* passes run */
  int overflow;                   /* nonzero: a pass ran out of pool or table space (results incomplete) */
} gxy_batch_stats;

enum { GXY_BATCH_RENDER = 0, GXY_BATCH_SAMPLE = 1 };

/* chunk: rays per work-queue entry (64..1024, multiple of 64). */
void *gxy_batch_create(int chunk);
int   gxy_batch_destroy(void *batch);
/* A resident dataset: vis holds either one VolumeVis and no model, or a
 * model and no VolumeVis (handles as for TraceRays_TraceRays).  Returns its
 * index.  epsilon: Renderer epsilon of the march step (0.001). */
int   gxy_batch_add_dataset(void *batch, const Visualization_ispc *vis, float epsilon);
/* Re-reads dataset ds from vis (after a change of its objects, e.g. a new
 * volume layout); the kind (volume or geometry) must stay the same. */
int   gxy_batch_update_dataset(void *batch, int ds, const Visualization_ispc *vis, float epsilon);
/* Lighting shared by every render list; epsilon offsets spawned rays. */
int   gxy_batch_set_lighting(void *batch, const Lighting_ispc *lights, float epsilon);
/* A list of n device rays (consumed: traced in place) over dataset ds.
 * GXY_BATCH_RENDER: terminated rays accumulate into fb (device, width*height*4
 * floats, cleared when the list is added and at each gxy_batch_run).
 * GXY_BATCH_SAMPLE: every isovalue crossing appends a particle to xyz (device,
 * capacity*3 floats); *count (device int) is the particle total.  Returns the
 * list index (indices of removed lists are reused). */
int   gxy_batch_add_list(void *batch, int ds, int mode, int n, const RayList_ispc *rays, float *fb, int width,
                         int height, float *xyz, int capacity, int *count);
/* Regenerate the list's rays at the start of every gxy_batch_run with
 * gxy_rays_random(seed, width, lo, hi). */
int   gxy_batch_set_list_seed(void *batch, int list, unsigned seed, const float lo[3], const float hi[3]);
int   gxy_batch_remove_list(void *batch, int list);
/* Launch configuration: chunk (rays per queue entry, 64..1024), stack (BVH traversal
 * stack: 0 = 64 entries in private memory, 16/24/32 = entries in LDS; must
 * exceed the deepest tree by 2), wpe (waves per SIMD requested: 0, 4, 5 or 6; with an
 * LDS stack the fused kernel ignores 5 and 6, split mode applies it to the volume
 * kernel; in split mode the sampler kernel always asks for 8),
 * blocks_per_cu (persistent workgroups per compute unit, 1..16). */
int   gxy_batch_set_config(void *batch, int chunk, int stack, int wpe, int blocks_per_cu);
int   gxy_batch_get_config(void *batch, int *chunk, int *stack, int *wpe, int *blocks_per_cu);
/* split 0: one persistent trace kernel takes every chunk; 1: three persistent
 * kernels (volume render, geometry, sampler chunks), each with its own
 * register budget, run concurrently on three streams. */
int   gxy_batch_set_split(void *batch, int split);
int   gxy_batch_get_split(void *batch);
/* Runs every list from its own rays to completion (clears framebuffers and
 * particle counts first).  graph_passes > 0: the set-up and that many passes
 * are captured once in a HIP graph and replayed on later runs (re-captured
 * after any change of lists, datasets or configuration); passes still needed
 * afterwards run one by one. */
int   gxy_batch_run(void *batch, int graph_passes, gxy_batch_stats *stats);
/* Continuous form: one pass over the rays now queued, including the primary
 * rays of lists added since the last pass (up to the step budget, below).
 * *next_rays receives the number of rays queued for the following pass,
 * primary rays still waiting for a step included (synchronises). */
int   gxy_batch_step(void *batch, long long *next_rays);
/* Step budget: at most `rays` primary rays of added lists join one step
 * (0 = all of them, the default; a list's last chunk may overshoot).  The
 * oldest waiting list gets a quarter of the budget, then the lists with the
 * fewest primary rays left go first; a list's remaining rays join later
 * steps, so a step stays short and a list added meanwhile starts at the next
 * step.  Results do not depend on the budget. */
int   gxy_batch_set_step_rays(void *batch, long long rays);
/* Sample lists: isovalue crossings one ray follows within a pass (1..4096,
 * default 4096); a ray with more continues in the next pass.  Results do not
 * depend on it; a small value shortens the longest rays of a pass. */
int   gxy_batch_set_max_segments(void *batch, int segments);
long long gxy_batch_get_step_rays(void *batch);
/* per_list[l] = rays of list l queued for the next pass, with its primary
 * rays still waiting for a step (l < n). */
int   gxy_batch_pending(void *batch, int *per_list, int n);
int   gxy_batch_get_stats(void *batch, gxy_batch_stats *stats);
/* Device bytes held by the batch's ray pools and tables. */
long long gxy_batch_bytes(void *batch);

/* Synthetic rays: origins uniform in [lo, hi], isotropic directions, t = 0,
 * tMax = inf, colour 0, pixel (i % width, i / width), from a PCG hash of
 * (seed, i).  type: RAY_PRIMARY etc. */
int gxy_rays_random(const RayList_ispc *rays, int n, unsigned seed, int width, int type, const float lo[3],
                    const float hi[3]);


/* ---- Served form ---------------------------------------------------------
 * No counterpart in the reference (whose multiserver renders one client
 * command at a time).  A request loop over resident datasets: requests are
 * admitted in arrival order as slots and device memory (an HBM budget) allow,
 * their rays join the batch of in-flight rays at the next pass boundary
 * (continuous batching over gxy_batch_step), and each request completes when
 * none of its rays remain queued.  The same stream can be served one request
 * at a time through the per-list loop (gxy_process_raylist /
 * gxy_sample_raylist) for comparison. */
typedef struct gxy_request {
  int kind;                 /* 0 render (camera rays, lighting, shadows/AO), 1 sample (isovalue crossings) */
  int dataset;              /* server dataset index */
  int width, height;        /* render: image size */
  long long nrays;          /* sample: random rays through the dataset's box */
  unsigned seed;            /* sample: ray seed */
  double arrival;           /* seconds after the start of the stream */
  float eye[3], center[3], up[3], aov;   /* render: camera */
  float *out;               /* optional host buffer: render width*height*4 floats, sample capacity*3 */
  long long capacity;       /* sample: particle capacity */
} gxy_request;

typedef struct gxy_request_result {
  double admitted, finished;   /* seconds after the start of the stream */
  long long particles;         /* sample: isovalue crossings found */
  long long rays;              /* render: primary rays generated; sample: rays */
  int status;                  /* 0 done, -1 failed, -2 rejected (invalid, or over the HBM
                                  budget by itself), -3 cancelled; while pending:
                                  1 queued, 2 admitted */
  int pad_;
} gxy_request_result;

typedef struct gxy_serve_stats {
  double wall;                 /* seconds from the first arrival to the last completion */
  unsigned long long rays;     /* ray passes traced */
  unsigned long long samples;  /* volume samples */
  long long peak_bytes;        /* device bytes in use, maximum over the run */
  int passes;                  /* batch passes (continuous) or trace launches (one at a time) */
  int max_in_flight;
} gxy_serve_stats;

enum { GXY_SERVE_CONTINUOUS = 0, GXY_SERVE_ONE_AT_A_TIME = 1 };

/* hbm_budget: device bytes the server may hold (requests + ray pools on top of
 * what is already allocated); max_slots: requests in flight at once. */
void *gxy_server_create(long long hbm_budget, int max_slots, int chunk);
int   gxy_server_destroy(void *server);
/* vis: one VolumeVis or one model (kept by the caller while the server runs). */
int   gxy_server_add_dataset(void *server, const Visualization_ispc *vis, float epsilon);
int   gxy_server_set_lighting(void *server, const Lighting_ispc *lights, float epsilon);
/* The batch the server drives (for gxy_batch_set_config). */
void *gxy_server_batch(void *server);
/* Serves n requests (sorted by arrival) in real time; res[k] receives request
 * k's admission and completion times. */
int   gxy_server_run(void *server, const gxy_request *reqs, int n, int mode, gxy_request_result *res,
                     gxy_serve_stats *stats);

/* The live served form (the request loop of the multiserver viewer,
 * ViewerClientServer.cpp:63-250, on the GPU): gxy_server_start runs a service
 * thread on its own stream that admits submitted requests as slots and the
 * HBM budget allow and steps every admitted request together (continuous
 * batching), until gxy_server_stop.  gxy_server_submit queues a copy of the
 * request (its out buffer must stay valid until the request is final; its
 * arrival is the time of submission) and returns its id; a request that is
 * invalid or over the budget by itself is rejected at once (status -2) while
 * the others go on.  gxy_server_poll returns the status (and fills res).
 * gxy_server_cancel cancels a queued or admitted request (admitted ones leave
 * the batch at the next step boundary).  gxy_server_drain waits until nothing
 * is queued or admitted; after timeout_s seconds it cancels what is left and
 * fails (the queue-drain watchdog).  A batch error fails the admitted and
 * queued requests, releases their buffers, and is reported by the next
 * gxy_server_drain or gxy_server_stop.  gxy_server_run needs the service
 * thread stopped. */
int   gxy_server_start(void *server);
int   gxy_server_submit(void *server, const gxy_request *req);
int   gxy_server_poll(void *server, int id, gxy_request_result *res);
int   gxy_server_cancel(void *server, int id);
int   gxy_server_drain(void *server, double timeout_s);
int   gxy_server_stop(void *server);

/* ==== the viewer's command front end (ViewerClientServer::handle,
 * multiserver/ViewerClientServer.cpp:63-252) on the live served form ====== */
/* A viewer takes the command lines of Galaxy's viewer (permute_pixels,
 * max_rays_per_packet, commit, json, window, render, renderMany N) and writes
 * the reference's reply ("ok" or its error text) into reply; returns 1 when
 * the command is one of these, 0 otherwise.  json Datasets name volumes
 * registered with gxy_viewer_add_volume (by file name, else name); json
 * Visualization (VolumeVis operators: colormap, opacitymap, data range,
 * isovalues, slices / plane, volume rendering; Lighting), Camera and Renderer
 * epsilon are kept; render / renderMany submit render requests to a live
 * server (gxy_server_start) holding the visualization, rebuilt when it or the
 * camera-resolved lighting changes, and wait for them.  Creating a viewer and
 * answering commands need no GPU until a frame is rendered.  permute_pixels
 * and max_rays_per_packet are kept and reported (the served form's camera
 * rays are tiled and its lists are stepped by its own step budget). */
void *gxy_viewer_create(long long hbm_budget, int slots);   /* 0: 64 GiB, 64 slots */
int   gxy_viewer_destroy(void *viewer);
int   gxy_viewer_add_volume(void *viewer, const char *key, void *volume, const gxy_box3f *local_box,
                            const gxy_box3f *global_box);
int   gxy_viewer_handle(void *viewer, const char *line, char *reply, size_t capacity);
/* The last frame (width*height*4 floats into out, when not NULL); returns 1
 * when there is one. */
int   gxy_viewer_frame(void *viewer, float *out, int *width, int *height, int *frames);
int   gxy_viewer_settings(void *viewer, int *permute, int *max_rays_per_packet, int *width, int *height);

#ifdef __cplusplus
}
#endif
#endif /* GXY_HIP_H */
