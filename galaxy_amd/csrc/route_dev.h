// route_dev.h — per-ray classification rules shared by the routing kernel
// (route.hip) and the batched pass kernels (batch.hip).
#pragma once

#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

// Renderer::Classify (Renderer.cpp:304-421) for one ray: returns the new
// classification (RAY_BOUNDARY marks a ray still to be routed); rays of other
// types keep c.
__device__ __forceinline__ int classify_ray(int typ, int term, int c, int reverse) {
  if (typ == RAY_PRIMARY) {
    if (term & RAY_BOUNDARY) c = RAY_BOUNDARY;
    else if ((term & RAY_OPAQUE) | (term & RAY_TIMEOUT)) c = GXY_TERMINATED;
    else c = GXY_KEEP_HERE;
  } else if (typ == RAY_SHADOW) {
    if ((term & RAY_OPAQUE) | (term & RAY_SURFACE)) c = reverse ? GXY_TERMINATED : GXY_DROP_ON_FLOOR;
    else if (term & RAY_BOUNDARY) c = RAY_BOUNDARY;
    else c = GXY_DROP_ON_FLOOR;
  } else if (typ == RAY_AO) {
    if ((term & RAY_OPAQUE) | (term & RAY_SURFACE)) c = reverse ? GXY_TERMINATED : GXY_DROP_ON_FLOOR;
    else if (term & RAY_BOUNDARY) c = RAY_BOUNDARY;
    else if (term & RAY_TIMEOUT) c = reverse ? GXY_DROP_ON_FLOOR : GXY_TERMINATED;
    else c = GXY_DROP_ON_FLOOR;
  }
  return c;
}

// Box::exit_face (Box.cpp:84-97), using the ray origin as the reference does.
__device__ __forceinline__ int exit_face(const float *lo, const float *hi, float x, float y, float z, float dx,
                                         float dy, float dz) {
  const float FM = 3.402823466e+38f;
  float tx = (dx > 0.0001f) ? ((hi[0] - x) / dx) : (dx < -0.0001f) ? ((lo[0] - x) / dx) : FM;
  float ty = (dy > 0.0001f) ? ((hi[1] - y) / dy) : (dy < -0.0001f) ? ((lo[1] - y) / dy) : FM;
  float tz = (dz > 0.0001f) ? ((hi[2] - z) / dz) : (dz < -0.0001f) ? ((lo[2] - z) / dz) : FM;
  if (tx < 0) tx = FM;
  if (ty < 0) ty = FM;
  if (tz < 0) tz = FM;
  if (tx < ty && tx < tz) return (dx < 0) ? 0 : 1;
  else if (ty < tz) return (dy < 0) ? 2 : 3;
  else return (dz < 0) ? 4 : 5;
}

// Rendering::AddLocalPixels (Rendering.cpp:125-153) for one wave: every lane
// with f set adds (r, g, b, a) to pixel (x, y) of fb.  A run of adjacent lanes
// with the same pixel (the AO and shadow children of one parent are adjacent)
// is summed first with a segmented suffix scan over shuffles that stops at the
// end of each run, and only the first lane of the run issues the four FP32
// atomics, so the children of one pixel do not serialise on one address.
// Lanes without a pixel (key -1) form runs of their own that are never
// written; equal pixels in separate runs are added separately.  All 64 lanes
// must call it.
// unique: no other ray of the launch adds to the same pixel, so a plain
// read-add-write gives the atomic's sum (one addition per pixel).
__device__ __forceinline__ void fb_accumulate_wave(float *fb, int W, int H, bool f, int x, int y, float r, float g,
                                                   float b, float a, bool unique = false) {
  const int lane = threadIdx.x & 63;
  const int key = (f && x >= 0 && y >= 0 && x < W && y < H) ? y * W + x : -1;
  if (unique) {
    if (key >= 0) {
      float *p = fb + 4 * (int64_t)key;
      p[0] = p[0] + r;
      p[1] = p[1] + g;
      p[2] = p[2] + b;
      p[3] = p[3] + a;
    }
    return;
  }
  if (!__any(key >= 0)) return;
  // end: the window [lane, lane + off) summed so far reaches the end of the run
  const int nkey = __shfl_down(key, 1, 64);
  int end = (lane == 63 || nkey != key) ? 1 : 0;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const float nr = __shfl_down(r, off, 64), ng = __shfl_down(g, off, 64);
    const float nb = __shfl_down(b, off, 64), na = __shfl_down(a, off, 64);
    const int nend = __shfl_down(end, off, 64);
    if (!end && lane + off < 64) {
      r += nr;
      g += ng;
      b += nb;
      a += na;
      end = nend;
    }
  }
  const int pk = __shfl_up(key, 1, 64);
  if (key >= 0 && (lane == 0 || pk != key)) {
    float *p = fb + 4 * (int64_t)key;
    atomicAdd(p + 0, r);
    atomicAdd(p + 1, g);
    atomicAdd(p + 2, b);
    atomicAdd(p + 3, a);
  }
}

}  // namespace gxy
