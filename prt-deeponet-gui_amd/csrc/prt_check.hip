// prt_check.hip -- the data passes of check_physics on the device.
//
// check_physics (3D/tools/prtlb_3d.py:1632-1843) validates the stored
// transport snapshots out (n_t, ns, *grid) with seven rules.  Every rule
// reduces the snapshots to a handful of numbers: non-finite counts, per
// (snapshot, chemical) pore maxima and minima, per-snapshot maxima of the
// donor on the mid and outlet x planes, the checkerboard counts of the last
// donor field along x, and second moments between chemical 0 and each other
// chemical.  These kernels make those numbers in one pass each over the
// device-resident snapshots; every kernel writes one partial per block (no
// atomics), so the host's final fold is deterministic.  The rules and their
// messages stay in the host package.
#include <hip/hip_runtime.h>

#include <cmath>

#include "prt_common.h"

namespace prt {

constexpr int kCheckBlock = 256;

template <typename T, typename Op>
__device__ inline T block_fold(T v, T *lds, Op op) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int w = kCheckBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) lds[threadIdx.x] = op(lds[threadIdx.x], lds[threadIdx.x + w]);
    __syncthreads();
  }
  return lds[0];
}

struct FMax {
  __device__ float operator()(float a, float b) const { return a > b ? a : b; }
};
struct FMin {
  __device__ float operator()(float a, float b) const { return a < b ? a : b; }
};
struct DAdd {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct LAdd {
  __device__ long long operator()(long long a, long long b) const { return a + b; }
};

// counts[2*b] = NaNs, counts[2*b+1] = infinities among x[b*chunk ...].
__global__ __launch_bounds__(kCheckBlock) void nonfinite_kernel(const float *__restrict__ x, long long n,
                                                               long long chunk, long long *__restrict__ counts) {
  __shared__ long long lds[kCheckBlock];
  const long long lo = (long long)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  long long nan = 0, inf = 0;
  for (long long i = lo + threadIdx.x; i < hi; i += kCheckBlock) {
    const float v = x[i];
    nan += v != v;
    inf += (v == INFINITY || v == -INFINITY);
  }
  const long long a = block_fold(nan, lds, LAdd());
  __syncthreads();
  const long long b = block_fold(inf, lds, LAdd());
  if (threadIdx.x == 0) {
    counts[2 * blockIdx.x] = a;
    counts[2 * blockIdx.x + 1] = b;
  }
}

// Pore maximum and minimum of out[t, j] over voxel chunk blockIdx.x, for
// (t, j) = blockIdx.y: mx/mn[blockIdx.y * gridDim.x + blockIdx.x].  A chunk
// without pore voxels gives -inf / +inf.
__global__ __launch_bounds__(kCheckBlock) void pore_range_kernel(const float *__restrict__ out,
                                                                const uint8_t *__restrict__ pore, long long N,
                                                                long long chunk, float *__restrict__ mx,
                                                                float *__restrict__ mn) {
  __shared__ float lds[kCheckBlock];
  const float *f = out + (size_t)blockIdx.y * N;
  const long long lo = (long long)blockIdx.x * chunk, hi = lo + chunk < N ? lo + chunk : N;
  float a = -INFINITY, b = INFINITY;
  for (long long i = lo + threadIdx.x; i < hi; i += kCheckBlock)
    if (pore[i]) {
      const float v = f[i];
      a = v > a ? v : a;
      b = v < b ? v : b;
    }
  const float A = block_fold(a, lds, FMax());
  __syncthreads();
  const float B = block_fold(b, lds, FMin());
  if (threadIdx.x == 0) {
    mx[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = A;
    mn[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = B;
  }
}

// Maximum of the donor (chemical 0) of snapshot t = blockIdx.y over every
// voxel of x plane xs[blockIdx.z] (two planes): part[(z * n_t + t) * gx + bx].
__global__ __launch_bounds__(kCheckBlock) void plane_max_kernel(const float *__restrict__ out, int ns, long long N,
                                                               long long plane, int x_mid, int x_out,
                                                               long long chunk, float *__restrict__ part) {
  __shared__ float lds[kCheckBlock];
  const int xs = blockIdx.z == 0 ? x_mid : x_out;
  const float *f = out + (size_t)blockIdx.y * ns * N + (size_t)xs * plane;
  const long long lo = (long long)blockIdx.x * chunk, hi = lo + chunk < plane ? lo + chunk : plane;
  float a = -INFINITY;
  for (long long i = lo + threadIdx.x; i < hi; i += kCheckBlock) a = f[i] > a ? f[i] : a;
  const float A = block_fold(a, lds, FMax());
  if (threadIdx.x == 0) part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = A;
}

// Checkerboard counts of f (the last donor field) along x, the reference's
// d = where(m[1:] & m[:-1], f[1:] - f[:-1], 0), pair = m[2:] & m[1:-1] &
// m[:-2], flip = sign(d[1:]) * sign(d[:-1]) < 0 within pair: counts[2b] =
// pairs, counts[2b+1] = flips over the middle voxels of chunk b.
__global__ __launch_bounds__(kCheckBlock) void checker_kernel(const float *__restrict__ f,
                                                             const uint8_t *__restrict__ pore, int X,
                                                             long long plane, long long chunk,
                                                             long long *__restrict__ counts) {
  __shared__ long long lds[kCheckBlock];
  const long long M = (long long)(X > 2 ? X - 2 : 0) * plane;  // voxels with x in 1 .. X-2
  const long long lo = (long long)blockIdx.x * chunk, hi = lo + chunk < M ? lo + chunk : M;
  long long pairs = 0, flips = 0;
  for (long long k = lo + threadIdx.x; k < hi; k += kCheckBlock) {
    const long long i = k + plane;  // x >= 1
    const bool m0 = pore[i - plane], m1 = pore[i], m2 = pore[i + plane];
    if (!(m0 && m1 && m2)) continue;
    ++pairs;
    const float d_hi = f[i + plane] - f[i], d_lo = f[i] - f[i - plane];
    const int s_hi = (d_hi > 0.0f) - (d_hi < 0.0f), s_lo = (d_lo > 0.0f) - (d_lo < 0.0f);
    flips += (s_hi * s_lo) < 0;
  }
  const long long a = block_fold(pairs, lds, LAdd());
  __syncthreads();
  const long long b = block_fold(flips, lds, LAdd());
  if (threadIdx.x == 0) {
    counts[2 * blockIdx.x] = a;
    counts[2 * blockIdx.x + 1] = b;
  }
}

// FP64 moments of (a0, aj) over every (snapshot, pore voxel), chemical
// j = blockIdx.y + 1, about the means mean[0] (chemical 0) and mean[j]
// (zero when mean is null): mom[((j-1) * gx + bx) * 5 + {0..4}] = sums of
// a0, aj, (a0-m0)^2, (aj-mj)^2, (a0-m0)(aj-mj).  Two calls (sums, then
// centred sums) give the two-pass standard deviations and correlation.
__global__ __launch_bounds__(kCheckBlock) void moments_kernel(const float *__restrict__ out, int n_t, int ns,
                                                             const uint8_t *__restrict__ pore, long long N,
                                                             long long chunk, const double *__restrict__ mean,
                                                             double *__restrict__ mom) {
  __shared__ double lds[kCheckBlock];
  const int j = blockIdx.y + 1;
  const double m0 = mean ? mean[0] : 0.0, mj = mean ? mean[j] : 0.0;
  const long long M = (long long)n_t * N;
  const long long lo = (long long)blockIdx.x * chunk, hi = lo + chunk < M ? lo + chunk : M;
  double s[5] = {0, 0, 0, 0, 0};
  for (long long k = lo + threadIdx.x; k < hi; k += kCheckBlock) {
    const long long t = k / N, i = k - t * N;
    if (!pore[i]) continue;
    const double a = out[((size_t)t * ns) * N + i], b = out[((size_t)t * ns + j) * N + i];
    const double da = a - m0, db = b - mj;
    s[0] += a;
    s[1] += b;
    s[2] += da * da;
    s[3] += db * db;
    s[4] += da * db;
  }
  for (int q = 0; q < 5; ++q) {
    const double v = block_fold(s[q], lds, DAdd());
    if (threadIdx.x == 0) mom[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 5 + q] = v;
    __syncthreads();
  }
}

// max |aj - sc a0| over every (snapshot, pore voxel), in float32 as the
// reference evaluates it: part[bx].
__global__ __launch_bounds__(kCheckBlock) void copy_residual_kernel(const float *__restrict__ out, int n_t, int ns,
                                                                   int j, float sc, const uint8_t *__restrict__ pore,
                                                                   long long N, long long chunk,
                                                                   float *__restrict__ part) {
  __shared__ float lds[kCheckBlock];
  const long long M = (long long)n_t * N;
  const long long lo = (long long)blockIdx.x * chunk, hi = lo + chunk < M ? lo + chunk : M;
  float r = 0.0f;
  for (long long k = lo + threadIdx.x; k < hi; k += kCheckBlock) {
    const long long t = k / N, i = k - t * N;
    if (!pore[i]) continue;
    const float a = out[((size_t)t * ns) * N + i], b = out[((size_t)t * ns + j) * N + i];
    const float d = fabsf(b - sc * a);  // file compiled with -ffp-contract=off
    r = d > r ? d : r;
  }
  const float R = block_fold(r, lds, FMax());
  if (threadIdx.x == 0) part[blockIdx.x] = R;
}

static long long chunk_for(long long n, long long blocks) { return (n + blocks - 1) / blocks; }

}  // namespace prt

using namespace prt;

extern "C" int prt_check_blocks(int64_t n) {
  // blocks per pass: enough to fill the device, each at least 4096 items
  long long b = (n + 4095) / 4096;
  if (b > 2048) b = 2048;
  return (int)(b < 1 ? 1 : b);
}

extern "C" int prt_check_nonfinite(const float *x, int64_t n, int blocks, int64_t *counts, void *stream) {
  if (!x || !counts || n < 1 || blocks < 1) {
    set_error("prt_check_nonfinite: bad arguments");
    return PRT_E_ARG;
  }
  nonfinite_kernel<<<blocks, kCheckBlock, 0, (hipStream_t)stream>>>(x, n, chunk_for(n, blocks),
                                                                   (long long *)counts);
  return launch_status("prt_check_nonfinite launch");
}

extern "C" int prt_check_pore_range(const float *out, int n_t, int ns, int64_t n_vox, const uint8_t *pore,
                                    int blocks, float *mx, float *mn, void *stream) {
  if (!out || !pore || !mx || !mn || n_t < 1 || ns < 1 || n_vox < 1 || blocks < 1 || (long long)n_t * ns > 65535) {
    set_error("prt_check_pore_range: bad arguments");
    return PRT_E_ARG;
  }
  const dim3 g((unsigned)blocks, (unsigned)(n_t * ns));
  pore_range_kernel<<<g, kCheckBlock, 0, (hipStream_t)stream>>>(out, pore, n_vox, chunk_for(n_vox, blocks), mx,
                                                                mn);
  return launch_status("prt_check_pore_range launch");
}

extern "C" int prt_check_plane_max(const float *out, int n_t, int ns, int nx, int64_t plane, int blocks,
                                   float *part, void *stream) {
  if (!out || !part || n_t < 1 || ns < 1 || nx < 1 || plane < 1 || blocks < 1 || n_t > 65535) {
    set_error("prt_check_plane_max: bad arguments");
    return PRT_E_ARG;
  }
  const dim3 g((unsigned)blocks, (unsigned)n_t, 2);
  plane_max_kernel<<<g, kCheckBlock, 0, (hipStream_t)stream>>>(out, ns, (long long)nx * plane, plane, nx / 2,
                                                               nx - 1, chunk_for(plane, blocks), part);
  return launch_status("prt_check_plane_max launch");
}

extern "C" int prt_check_checker(const float *f, const uint8_t *pore, int nx, int64_t plane, int blocks,
                                 int64_t *counts, void *stream) {
  if (!f || !pore || !counts || nx < 1 || plane < 1 || blocks < 1) {
    set_error("prt_check_checker: bad arguments");
    return PRT_E_ARG;
  }
  const long long M = (long long)(nx > 2 ? nx - 2 : 0) * plane;
  checker_kernel<<<blocks, kCheckBlock, 0, (hipStream_t)stream>>>(f, pore, nx, plane,
                                                                  M > 0 ? chunk_for(M, blocks) : 0,
                                                                  (long long *)counts);
  return launch_status("prt_check_checker launch");
}

extern "C" int prt_check_moments(const float *out, int n_t, int ns, const uint8_t *pore, int64_t n_vox, int blocks,
                                 const double *mean, double *mom, void *stream) {
  if (!out || !pore || !mom || n_t < 1 || ns < 2 || n_vox < 1 || blocks < 1) {
    set_error("prt_check_moments: bad arguments");
    return PRT_E_ARG;
  }
  const dim3 g((unsigned)blocks, (unsigned)(ns - 1));
  moments_kernel<<<g, kCheckBlock, 0, (hipStream_t)stream>>>(out, n_t, ns, pore, n_vox,
                                                             chunk_for((long long)n_t * n_vox, blocks), mean, mom);
  return launch_status("prt_check_moments launch");
}

extern "C" int prt_check_copy_residual(const float *out, int n_t, int ns, int j, float sc, const uint8_t *pore,
                                       int64_t n_vox, int blocks, float *part, void *stream) {
  if (!out || !pore || !part || n_t < 1 || j < 1 || j >= ns || n_vox < 1 || blocks < 1) {
    set_error("prt_check_copy_residual: bad arguments");
    return PRT_E_ARG;
  }
  copy_residual_kernel<<<blocks, kCheckBlock, 0, (hipStream_t)stream>>>(
      out, n_t, ns, j, sc, pore, n_vox, chunk_for((long long)n_t * n_vox, blocks), part);
  return launch_status("prt_check_copy_residual launch");
}
