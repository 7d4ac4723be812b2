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

}  // namespace gxy
