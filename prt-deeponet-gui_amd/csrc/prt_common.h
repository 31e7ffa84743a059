// Shared helpers of the PRT-LB HIP library: error reporting and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/prtlb.h"

namespace prt {

// Thread-local message of the last failure (read by prt_last_error()).
void set_error(const char *fmt, ...);

inline int hip_fail(hipError_t e, const char *what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return PRT_E_HIP;
}

// Check the launch that was just issued.
inline int launch_status(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? PRT_OK : hip_fail(e, what);
}

// Division of a non-negative 31-bit index by a fixed divisor d >= 1 with a
// multiply-high, an add and a shift (Granlund & Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, round-up variant):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (mulhi(n, m) + n) >> l for 0 <= n < 2^31.
struct FastDiv {
  uint32_t m;
  uint32_t l;
  int d;
};

inline FastDiv make_fastdiv(int d) {
  FastDiv f;
  f.d = d < 1 ? 1 : d;
  f.l = 0;
  while ((1ull << f.l) < (unsigned long long)f.d) ++f.l;
  f.m = (uint32_t)((((1ull << f.l) - (unsigned long long)f.d) << 32) / (unsigned long long)f.d + 1ull);
  return f;
}

__device__ inline int fastdiv(int n, const FastDiv &f) {
  const uint32_t u = (uint32_t)n;
  return (int)((__umulhi(u, f.m) + u) >> f.l);
}

// Workgroups are dispatched to the device's 8 XCDs round-robin (linear
// workgroup b to XCD b % 8), and each XCD has its own L2.  xcd_block(b, nb)
// renumbers the nb workgroups so that XCD x runs one contiguous range of
// logical blocks: data shared by neighbouring blocks (stencil neighbours,
// cache lines split between blocks) then meets in one L2 instead of being
// fetched by several.  A bijection of [0, nb), so results do not depend on it.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb / kXcds, r = nb % kXcds, x = b % kXcds, slot = b / kXcds;
  return x * q + (x < r ? x : r) + slot;
}

// Block-wide FP64 sum of one value per thread (blockDim.x * blockDim.y <=
// 1024, a multiple of 64).  Wavefront reduction with shuffles, then LDS.
__device__ inline double block_sum(double v, double *lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  const int nw = (blockDim.x * blockDim.y) >> 6;
  if ((t & 63) == 0) lds[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0)
    for (int k = 0; k < nw; ++k) s += lds[k];
  return s;  // valid in thread 0
}

// The pore voxels (g == pore_code) face-connected to a pore voxel on x = 0,
// as reach[n] = 1, by repeated face dilation inside the pore (prt_geom.hip).
// changed: one device int of scratch.  Host code; synchronises the stream.
int inlet_reach(const uint8_t *g, int X, int Y, int Z, uint8_t pore_code, uint8_t *reach, int *changed,
                int max_sweeps, hipStream_t st);

}  // namespace prt
