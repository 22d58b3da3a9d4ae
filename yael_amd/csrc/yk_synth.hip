// yk_synth.hip -- synthetic vector sets generated in HBM, for benchmarks of
// sizes the host would take long to generate and copy (1e7 x 128 floats).
//
// Counter-based: every value is a function of (seed, row, column) only, so a
// set is reproducible whatever the launch shape.  Uniforms come from the
// splitmix64 finaliser (Steele, Lea, Flood 2014) of the counter; normal
// deviates from the Box-Muller transform of two uniforms.
#include <hip/hip_runtime.h>
#include <math.h>

#include "yk.h"
#include "yk_dev.h"

__device__ __forceinline__ uint64_t yk_mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// uniform in (0, 1]: 24 high bits of the hash
__device__ __forceinline__ float yk_u01(uint64_t h) {
  return ((float)(uint32_t)(h >> 40) + 1.0f) * 0x1p-24f;
}

__device__ __forceinline__ float yk_normal(uint64_t seed, uint64_t ctr) {
  const uint64_t h1 = yk_mix64(seed ^ yk_mix64(2 * ctr));
  const uint64_t h2 = yk_mix64(seed ^ yk_mix64(2 * ctr + 1));
  const float u1 = yk_u01(h1), u2 = yk_u01(h2);
  return sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
}

// kind 0: uniform [lo, hi); 1: mixture (centre of a hashed component +
// sigma * N(0,1), clipped to [lo, hi], optionally rounded); 2: |N(0,1)|.
// Pad columns [d, ld) are written as zero.
__global__ void yk_synth_kernel(float *out, long n, int d, int ld, int kind, float lo, float hi,
                                const float *__restrict__ cen, int ncomp, float sigma, int rnd,
                                uint64_t seed) {
  const long total = n * ld;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const long i = t / ld;
    const int j = (int)(t - i * ld);
    float x = 0.f;
    if (j < d) {
      const uint64_t ctr = (uint64_t)i * d + j;
      if (kind == 0) {
        x = lo + (hi - lo) * (yk_u01(yk_mix64(seed ^ yk_mix64(ctr))) - 0x1p-24f);
      } else if (kind == 1) {
        const int c = (int)(yk_mix64(~seed ^ yk_mix64((uint64_t)i)) % (uint64_t)ncomp);
        x = cen[(long)c * d + j] + sigma * yk_normal(seed, ctr);
        x = fminf(fmaxf(x, lo), hi);
        if (rnd) x = rintf(x);
      } else {
        x = fabsf(yk_normal(seed, ctr));
      }
    }
    out[t] = x;
  }
}

// rows scaled to unit L2 norm, one wave per row
__global__ void yk_unit_rows_kernel(float *x, long n, int d, int ld) {
  const long row = blockIdx.x * (long)(blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  float *p = x + row * ld;
  float s = 0.f;
  for (int j = lane; j < d; j += 64) s += p[j] * p[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  const float inv = s > 0.f ? 1.0f / sqrtf(s) : 0.f;
  for (int j = lane; j < d; j += 64) p[j] *= inv;
}

extern "C" int yk_synth(float *out, long n, int d, int ld, int kind, float lo, float hi,
                        const float *centres, int ncomp, float sigma, int round_values,
                        unsigned long long seed, int unit_rows) {
  YK_TRY(yk_init());
  if (n <= 0) return YK_OK;
  if (d < 1 || ld < d || kind < 0 || kind > 2 || (kind == 1 && (!centres || ncomp < 1)))
    return yk_set_error(YK_EINVAL, "yk_synth: bad arguments");
  hipLaunchKernelGGL(yk_synth_kernel, dim3(8192), dim3(256), 0, yk_stream(), out, n, d, ld, kind,
                     lo, hi, centres, ncomp, sigma, round_values, (uint64_t)seed);
  YK_HIP(hipGetLastError());
  if (unit_rows) {
    hipLaunchKernelGGL(yk_unit_rows_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0,
                       yk_stream(), out, n, d, ld);
    YK_HIP(hipGetLastError());
  }
  return YK_OK;
}
