// noise.hip — seeded 4-octave gradient-noise volumes (synthetic input of the
// volume-render workload): improved Perlin noise (K. Perlin, "Improving
// Noise", SIGGRAPH 2002) with a seeded permutation, summed over octaves with
// amplitude 1/2^o and frequency base*2^o cells across the volume, then scaled
// to exactly [0, 1] by the field's own minimum and maximum.
#include "gxy_device.h"
#include "gxy_internal.h"

namespace gxy {

__device__ __forceinline__ float fade(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }
__device__ __forceinline__ float lerpf(float t, float a, float b) { return a + t * (b - a); }
__device__ __forceinline__ float grad(int h, float x, float y, float z) {
  h &= 15;
  float u = h < 8 ? x : y;
  float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
  return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}

__device__ float pnoise(const int *P, float x, float y, float z) {
  float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
  x -= fx; y -= fy; z -= fz;
  float u = fade(x), v = fade(y), w = fade(z);
  int A = P[X] + Y, AA = P[A & 511] + Z, AB = P[(A + 1) & 511] + Z;
  int B = P[X + 1] + Y, BA = P[B & 511] + Z, BB = P[(B + 1) & 511] + Z;
  return lerpf(w,
               lerpf(v, lerpf(u, grad(P[AA & 511], x, y, z), grad(P[BA & 511], x - 1, y, z)),
                     lerpf(u, grad(P[AB & 511], x, y - 1, z), grad(P[BB & 511], x - 1, y - 1, z))),
               lerpf(v, lerpf(u, grad(P[(AA + 1) & 511], x, y, z - 1), grad(P[(BA + 1) & 511], x - 1, y, z - 1)),
                     lerpf(u, grad(P[(AB + 1) & 511], x, y - 1, z - 1),
                           grad(P[(BB + 1) & 511], x - 1, y - 1, z - 1))));
}

// order-preserving float <-> uint mapping for atomic min / max
__device__ __forceinline__ unsigned int fkey(float f) {
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(unsigned int k) {
  unsigned int u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

__global__ __launch_bounds__(GXY_BLOCK) void k_noise(float *d, int nx, int ny, int nz, const int *perm, int octaves,
                                                     float base, unsigned int *mm) {
  __shared__ int P[512];
  for (int k = threadIdx.x; k < 512; k += GXY_BLOCK) P[k] = perm[k & 255];
  __syncthreads();
  int64_t total = (int64_t)nx * ny * nz;
  unsigned int kmin = 0xffffffffu, kmax = 0u;
  for (int64_t v = (int64_t)blockIdx.x * GXY_BLOCK + threadIdx.x; v < total; v += (int64_t)gridDim.x * GXY_BLOCK) {
    int x = (int)(v % nx);
    int y = (int)((v / nx) % ny);
    int z = (int)(v / ((int64_t)nx * ny));
    float px = (x + 0.5f) / nx, py = (y + 0.5f) / ny, pz = (z + 0.5f) / nz;
    float s = 0.f, amp = 1.f, f = base;
    for (int o = 0; o < octaves; o++) {
      s += amp * pnoise(P, px * f, py * f, pz * f);
      amp *= 0.5f;
      f *= 2.f;
    }
    d[v] = s;
    unsigned int k = fkey(s);
    kmin = min(kmin, k);
    kmax = max(kmax, k);
  }
  for (int o = 32; o > 0; o >>= 1) {
    kmin = min(kmin, (unsigned int)__shfl_xor((int)kmin, o, 64));
    kmax = max(kmax, (unsigned int)__shfl_xor((int)kmax, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(mm, kmin);
    atomicMax(mm + 1, kmax);
  }
}

__global__ __launch_bounds__(GXY_BLOCK) void k_normalise(float *d, int64_t total, const unsigned int *mm) {
  float lo = funkey(mm[0]), hi = funkey(mm[1]);
  float inv = hi > lo ? 1.0f / (hi - lo) : 0.f;
  for (int64_t v = (int64_t)blockIdx.x * GXY_BLOCK + threadIdx.x; v < total; v += (int64_t)gridDim.x * GXY_BLOCK) {
    float t = (d[v] - lo) * inv;
    d[v] = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
  }
}

__global__ void k_mm_init(unsigned int *mm) {
  mm[0] = 0xffffffffu;
  mm[1] = 0u;
}

hipError_t launch_noise(float *data, int nx, int ny, int nz, const int *perm_dev, int octaves, float base_freq,
                        float *minmax_scratch, hipStream_t s) {
  unsigned int *mm = (unsigned int *)minmax_scratch;
  int64_t total = (int64_t)nx * ny * nz;
  int64_t want = (total + GXY_BLOCK - 1) / GXY_BLOCK;
  int nb = (int)(want < 65536 ? want : 65536);
  hipLaunchKernelGGL(k_mm_init, dim3(1), dim3(1), 0, s, mm);
  hipLaunchKernelGGL(k_noise, dim3(nb), dim3(GXY_BLOCK), 0, s, data, nx, ny, nz, perm_dev, octaves, base_freq, mm);
  hipLaunchKernelGGL(k_normalise, dim3(nb), dim3(GXY_BLOCK), 0, s, data, total, mm);
  return hipGetLastError();
}

// ------------------------------------------------------------------ bricked copy
// out[b * BF + (lz*(B+1) + ly)*(B+1) + lx] = voxel (B*bx + lx, B*by + ly,
// B*bz + lz) clamped to the volume (clamped voxels beyond the last cell are
// never read by a sample); padding entries are 0.
template <int B>
__global__ __launch_bounds__(256) void k_brickify(const float *data, int nx, int ny, int nz, int nbx, int nby,
                                                  long long total, float *out) {
  constexpr int E = B + 1, BF = brick_floats(B);
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < total; k += (long long)gridDim.x * 256) {
    const long long b = k / BF;
    const int r = (int)(k - b * BF);
    if (r >= E * E * E) {
      out[k] = 0.f;
      continue;
    }
    const int lx = r % E, ly = (r / E) % E, lz = r / (E * E);
    const int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long long)nbx * nby));
    const int x = min(B * bx + lx, nx - 1), y = min(B * by + ly, ny - 1), z = min(B * bz + lz, nz - 1);
    out[k] = data[(long long)x + (long long)nx * y + (long long)nx * ny * z];
  }
}

hipError_t launch_brickify(const float *data, const int n[3], int B, const int nb[3], float *out, hipStream_t s) {
  const long long total = (long long)nb[0] * nb[1] * nb[2] * brick_floats(B);
  long long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (B == 4)
    hipLaunchKernelGGL(k_brickify<4>, dim3((unsigned)g), dim3(256), 0, s, data, n[0], n[1], n[2], nb[0], nb[1], total,
                       out);
  else if (B == 8)
    hipLaunchKernelGGL(k_brickify<8>, dim3((unsigned)g), dim3(256), 0, s, data, n[0], n[1], n[2], nb[0], nb[1], total,
                       out);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace gxy
