// prt_features.hip -- the harmonic pressure prior and its gradient on the GPU
// (3D/tools/harmonic_pressure.py:91-197), the per-rock flow features the
// velocity operator is trained on.
//
// Harmonic pressure.  Laplace's equation on the pore voxels face-connected to
// the inlet: P = 1 on x = 0, P = 0 on x = nx - 1, and at every other such
// voxel sum over its connected face neighbours (P_nb - P) = 0, a wall simply
// dropping out of the sum (the reference's sparse rows, :110-146).  The
// reference factorises that system (scipy spsolve); here the Dirichlet
// voxels are moved to the right-hand side, which leaves the graph Laplacian
// of the interior voxels plus their Dirichlet degree: symmetric positive
// definite, because every interior component touches the inlet plane through
// a connected voxel.  It is solved matrix-free over the box by
// Jacobi-preconditioned conjugate gradients in FP64 to a relative residual
// (default 1e-12), and rounded to float32 as the reference stores it.
//
// Every vector is a box array (zero off the unknowns), every dot product is
// per-block partials summed by one block in a fixed order, and the CG
// scalars stay on the device, so an iteration is five launches and the host
// only reads the residual every 32 iterations.
//
// Gradient (:148-197): per axis a central difference where both neighbours
// are pore, one-sided where only one is, zero where neither is, all in
// float32; then, normalised, everything is divided by the largest |component|
// over the pore.
//
// Flow descriptors (3D/tools/flow_features.py), bit-identical to the
// reference's float32 maps:
//   - UPRM, the widest path from the inlet (:205-257): the reference pops a
//     max-first heap; here every pore voxel repeatedly takes the largest
//     min(value of a pore face neighbour, own wall distance) until nothing
//     changes.  Values only grow and each is realised by a path, so the
//     fixed point is the same widest-path value, whatever the order.
//   - dw2 (:260-280): the squared wall distance scaled to [0, 1], float32.
//   - MIS, the local thickness (:128-150): the reference sweeps radii
//     r_k = rmax - k r_step down and paints, with a ball of radius r_k,
//     around every voxel at least r_k - 1e-6 from a wall.  A voxel c is a
//     centre from its first admissible radius rc(c) = r_k0 on, so a voxel x
//     ends up with the largest rc(c) over the voxels c with |x - c|^2 <=
//     rc(c)^2 (smaller radii cover less): one gather over the ball of
//     radius ceil(rmax), with the reference's comparisons in FP64.
//   - Painting into the buffers (:153-177): a buffer voxel takes the largest
//     interior value v whose ball (|x - c|^2 <= v * v in float32) reaches it.
#include <hip/hip_runtime.h>

#include <cmath>

#include "prt_common.h"

#pragma clang fp contract(off)

namespace prt {

constexpr int kFeatBlock = 256;
constexpr int kSumThreads = 1024;

static unsigned fblocks(long long n) { return (unsigned)((n + kFeatBlock - 1) / kFeatBlock); }

__device__ inline void unflat(long long n, int Y, int Z, int &x, int &y, int &z) {
  const long long yz = (long long)Y * Z;
  x = (int)(n / yz);
  const int r = (int)(n - (long long)x * yz);
  y = r / Z;
  z = r - y * Z;
}

// Degree (live face neighbours) of each unknown, 0 elsewhere; the
// right-hand side (its live neighbours on x = 0, each P = 1); x = 0.
__global__ __launch_bounds__(kFeatBlock) void hp_setup_kernel(const uint8_t *live, int X, int Y, int Z,
                                                              uint8_t *deg, double *xv, double *r) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z, yz = (long long)Y * Z;
  if (n >= N) return;
  int x, y, z;
  unflat(n, Y, Z, x, y, z);
  int d = 0;
  double b = 0.0;
  if (live[n] && x > 0 && x < X - 1) {
    d += live[n - yz];
    d += live[n + yz];
    if (y > 0) d += live[n - Z];
    if (y < Y - 1) d += live[n + Z];
    if (z > 0) d += live[n - 1];
    if (z < Z - 1) d += live[n + 1];
    if (x == 1 && live[n - yz]) b = 1.0;
  }
  deg[n] = (uint8_t)d;
  xv[n] = 0.0;
  r[n] = d ? b : 0.0;
}

// A applied to v at unknown n: deg v(n) - sum of v over unknown neighbours.
__device__ inline double hp_apply(const uint8_t *deg, const double *v, long long n, int x, int y, int z, int X, int Y,
                                  int Z) {
  const long long yz = (long long)Y * Z;
  double s = 0.0;
  if (x > 1 && deg[n - yz]) s += v[n - yz];
  if (x < X - 2 && deg[n + yz]) s += v[n + yz];
  if (y > 0 && deg[n - Z]) s += v[n - Z];
  if (y < Y - 1 && deg[n + Z]) s += v[n + Z];
  if (z > 0 && deg[n - 1]) s += v[n - 1];
  if (z < Z - 1 && deg[n + 1]) s += v[n + 1];
  return (double)deg[n] * v[n] - s;
}

// p = r / deg (the preconditioned residual) and partial sums of r.z, r.r.
__global__ __launch_bounds__(kFeatBlock) void hp_start_kernel(const uint8_t *deg, const double *r, long long N,
                                                              double *p, double *part) {
  __shared__ double lds[16];
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (n < N) {
    const double z = deg[n] ? r[n] / (double)deg[n] : 0.0;
    p[n] = z;
    rz = r[n] * z;
    rr = r[n] * r[n];
  }
  const double a = block_sum(rz, lds);
  __syncthreads();
  const double b = block_sum(rr, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// q = A p and partial sums of p.q.
__global__ __launch_bounds__(kFeatBlock) void hp_apply_kernel(const uint8_t *deg, const double *p, int X, int Y, int Z,
                                                              double *q, double *part) {
  __shared__ double lds[16];
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  double pq = 0.0;
  if (n < N) {
    double v = 0.0;
    if (deg[n]) {
      int x, y, z;
      unflat(n, Y, Z, x, y, z);
      v = hp_apply(deg, p, n, x, y, z, X, Y, Z);
      pq = p[n] * v;
    }
    q[n] = v;
  }
  const double a = block_sum(pq, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// Fixed-order sum of `width` interleaved partial streams over nb blocks by
// one block; mode 0: s[0] = rz, s[1] = rr (start); mode 1: alpha =
// rz / pq into s[2]; mode 2: beta = rz_new / rz, rz = rz_new, rr into s[1].
// Scalars: s[0] rz, s[1] rr, s[2] alpha, s[3] beta.
__global__ __launch_bounds__(kSumThreads) void hp_scalar_kernel(const double *part, long long nb, int width,
                                                                int mode, double *s) {
  __shared__ double acc[kSumThreads];
  double tot[2] = {0.0, 0.0};
  for (int w = 0; w < width; ++w) {
    double v = 0.0;
    for (long long k = threadIdx.x; k < nb; k += kSumThreads) v += part[k * width + w];
    acc[threadIdx.x] = v;
    __syncthreads();
    for (int h = kSumThreads / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) acc[threadIdx.x] += acc[threadIdx.x + h];
      __syncthreads();
    }
    tot[w] = acc[0];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    s[0] = tot[0];
    s[1] = tot[1];
  } else if (mode == 1) {
    s[2] = tot[0] != 0.0 ? s[0] / tot[0] : 0.0;
  } else {
    s[3] = s[0] != 0.0 ? tot[0] / s[0] : 0.0;
    s[0] = tot[0];
    s[1] = tot[1];
  }
}

// x += alpha p, r -= alpha q; partial sums of r.z (z = r / deg) and r.r.
__global__ __launch_bounds__(kFeatBlock) void hp_update_kernel(const uint8_t *deg, const double *p, const double *q,
                                                               const double *s, long long N, double *xv, double *r,
                                                               double *part) {
  __shared__ double lds[16];
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (n < N && deg[n]) {
    const double alpha = s[2];
    xv[n] += alpha * p[n];
    const double rn = r[n] - alpha * q[n];
    r[n] = rn;
    rz = rn * (rn / (double)deg[n]);
    rr = rn * rn;
  }
  const double a = block_sum(rz, lds);
  __syncthreads();
  const double b = block_sum(rr, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// p = r / deg + beta p.
__global__ __launch_bounds__(kFeatBlock) void hp_direction_kernel(const uint8_t *deg, const double *r, const double *s,
                                                                  long long N, double *p) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N && deg[n]) p[n] = r[n] / (double)deg[n] + s[3] * p[n];
}

// The float32 field: 1 on live x = 0, 0 on live x = nx - 1, the solution on
// the unknowns, 0 everywhere else (solid, unreached pore, isolated voxels).
__global__ __launch_bounds__(kFeatBlock) void hp_field_kernel(const uint8_t *live, const uint8_t *deg,
                                                              const double *xv, int X, int Y, int Z, float *out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z, yz = (long long)Y * Z;
  if (n >= N) return;
  const int x = (int)(n / yz);
  float v = 0.0f;
  if (live[n]) {
    if (x == 0) v = 1.0f;
    else if (x == X - 1) v = 0.0f;
    else if (deg[n]) v = (float)xv[n];
  }
  out[n] = v;
}

// Pore voxels the inlet does not reach.
__global__ __launch_bounds__(kFeatBlock) void hp_dropped_kernel(const uint8_t *pore, const uint8_t *live, long long N,
                                                                int *count) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N && pore[n] && !live[n]) atomicAdd(count, 1);
}

// Gradient components (nd of them) of P at every voxel, float32, and
// per-block maxima of |component| over the pore.
__global__ __launch_bounds__(kFeatBlock) void hg_kernel(const uint8_t *pore, const float *P, int nd, int X, int Y,
                                                        int Z, float *out, float *bmax) {
  __shared__ float lds[kFeatBlock];
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  float m = 0.0f;
  if (n < N) {
    int c[3];
    unflat(n, Y, Z, c[0], c[1], c[2]);
    const int ext[3] = {X, Y, Z};
    const long long stride[3] = {(long long)Y * Z, Z, 1};
    const bool here = pore[n] != 0;
    for (int a = 0; a < nd; ++a) {
      // a 2D grid is (X, 1, Z): its axes are x and z
      const int ax = (nd == 2 && a == 1) ? 2 : a;
      float g = 0.0f;
      if (here && ext[ax] >= 2) {
        const bool okp = c[ax] < ext[ax] - 1 && pore[n + stride[ax]];
        const bool okm = c[ax] > 0 && pore[n - stride[ax]];
        const float p0 = P[n];
        if (okp && okm) g = (P[n + stride[ax]] - P[n - stride[ax]]) * 0.5f;
        else if (okp) g = P[n + stride[ax]] - p0;
        else if (okm) g = p0 - P[n - stride[ax]];
      }
      out[(size_t)a * N + n] = g;
      if (here) m = fmaxf(m, fabsf(g));
    }
  }
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int h = kFeatBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) lds[threadIdx.x] = fmaxf(lds[threadIdx.x], lds[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blockIdx.x] = lds[0];
}

__global__ __launch_bounds__(kSumThreads) void max_kernel(const float *v, long long n, float *out) {
  __shared__ float lds[kSumThreads];
  float m = 0.0f;
  for (long long k = threadIdx.x; k < n; k += kSumThreads) m = fmaxf(m, v[k]);
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int h = kSumThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) lds[threadIdx.x] = fmaxf(lds[threadIdx.x], lds[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

__global__ __launch_bounds__(kFeatBlock) void div_kernel(float *v, long long n, const float *m) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const float d = *m;
  if (k < n && d > 0.0f) v[k] = v[k] / d;
}

// ------------------------------------------------------------ UPRM
__global__ __launch_bounds__(kFeatBlock) void uprm_init_kernel(const uint8_t *pore, const float *edt, long long N,
                                                               long long plane, float *best) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) best[n] = (n < plane && pore[n]) ? edt[n] : 0.0f;
}

// One relaxation sweep in place: best(v) = max(best(v), max over pore face
// neighbours u of min(best(u), edt(v))).
__global__ __launch_bounds__(kFeatBlock) void uprm_relax_kernel(const uint8_t *pore, const float *edt, int X, int Y,
                                                                int Z, float *best, int *changed) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z, yz = (long long)Y * Z;
  if (n >= N || !pore[n]) return;
  int x, y, z;
  unflat(n, Y, Z, x, y, z);
  float m = 0.0f;
  if (x > 0 && pore[n - yz]) m = fmaxf(m, best[n - yz]);
  if (x < X - 1 && pore[n + yz]) m = fmaxf(m, best[n + yz]);
  if (y > 0 && pore[n - Z]) m = fmaxf(m, best[n - Z]);
  if (y < Y - 1 && pore[n + Z]) m = fmaxf(m, best[n + Z]);
  if (z > 0 && pore[n - 1]) m = fmaxf(m, best[n - 1]);
  if (z < Z - 1 && pore[n + 1]) m = fmaxf(m, best[n + 1]);
  const float cand = fminf(m, edt[n]);
  if (cand > best[n]) {
    best[n] = cand;
    *changed = 1;
  }
}

// ------------------------------------------------------------ dw2
__global__ __launch_bounds__(kFeatBlock) void dw2_kernel(const uint8_t *pore, const float *edt, long long N, float lo,
                                                         float span, float *out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float v = 0.0f;
  if (pore[n]) {
    const float f2 = edt[n] * edt[n];
    v = (f2 - lo) / span;
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);  // np.clip (no NaN here)
  }
  out[n] = v;
}

// ------------------------------------------------------------ MIS
// rc(c): the first radius of the sweep rmax, rmax - step, ... (K of them)
// at which voxel c is a centre (edt >= r - 1e-6, compared in FP64), or 0.
__global__ __launch_bounds__(kFeatBlock) void lt_centre_kernel(const uint8_t *pore, const float *edt, long long N,
                                                               double rmax, double step, int K, double *rc) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double r = 0.0;
  if (pore[n]) {
    const double e = (double)edt[n];
    int k = (int)floor((rmax - e) / step) - 1;
    k = k < 0 ? 0 : k;
    while (k < K && !(e >= (rmax + (double)k * -step) - 1e-6)) ++k;
    if (k < K) r = rmax + (double)k * -step;
  }
  rc[n] = r;
}

// out(x) = float32 of the largest rc(c) with |x - c|^2 <= rc(c)^2 over the
// ball of radius R around a pore voxel x, 0 off the pore.
__global__ __launch_bounds__(kFeatBlock) void lt_gather_kernel(const uint8_t *pore, const double *rc, int X, int Y,
                                                               int Z, int R, float *out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  if (n >= N) return;
  float v = 0.0f;
  if (pore[n]) {
    int x, y, z;
    unflat(n, Y, Z, x, y, z);
    double best = 0.0;
    const int x0 = x - R < 0 ? 0 : x - R, x1 = x + R > X - 1 ? X - 1 : x + R;
    const int y0 = y - R < 0 ? 0 : y - R, y1 = y + R > Y - 1 ? Y - 1 : y + R;
    const int z0 = z - R < 0 ? 0 : z - R, z1 = z + R > Z - 1 ? Z - 1 : z + R;
    for (int a = x0; a <= x1; ++a)
      for (int b = y0; b <= y1; ++b) {
        const long long row = ((long long)a * Y + b) * Z;
        const int dab = (a - x) * (a - x) + (b - y) * (b - y);
        for (int c = z0; c <= z1; ++c) {
          const double r = rc[row + c];
          const int d2 = dab + (c - z) * (c - z);
          if (r > best && (double)d2 <= r * r) best = r;
        }
      }
    v = (float)best;
  }
  out[n] = v;
}

// Buffer voxels (x < buf or x >= X - buf) of `full`: the largest interior
// value v (x in [buf, X - buf)) with |x - c|^2 <= v * v, v * v in float32.
__global__ __launch_bounds__(kFeatBlock) void paint_kernel(int X, int Y, int Z, int buf, int R, float *full) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  if (n >= N) return;
  int x, y, z;
  unflat(n, Y, Z, x, y, z);
  if (x >= buf && x < X - buf) return;
  float best = 0.0f;
  const int lo = buf, hi = X - buf - 1;  // the interior slab
  int x0 = x - R < lo ? lo : x - R, x1 = x + R > hi ? hi : x + R;
  const int y0 = y - R < 0 ? 0 : y - R, y1 = y + R > Y - 1 ? Y - 1 : y + R;
  const int z0 = z - R < 0 ? 0 : z - R, z1 = z + R > Z - 1 ? Z - 1 : z + R;
  for (int a = x0; a <= x1; ++a)
    for (int b = y0; b <= y1; ++b) {
      const long long row = ((long long)a * Y + b) * Z;
      const int dab = (a - x) * (a - x) + (b - y) * (b - y);
      for (int c = z0; c <= z1; ++c) {
        const float v = full[row + c];
        const int d2 = dab + (c - z) * (c - z);
        if (v > best && (double)d2 <= (double)(v * v)) best = v;
      }
    }
  full[n] = best;
}

// MIS finish: unreached pore in the buffers takes vmin; off the pore, 0.
__global__ __launch_bounds__(kFeatBlock) void mis_finish_kernel(const uint8_t *pore, int X, long long plane, int buf,
                                                                float vmin, float *full) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)X * plane) return;
  const int x = (int)(n / plane);
  const bool ends = x < buf || x >= X - buf;
  if (!pore[n]) full[n] = 0.0f;
  else if (ends && full[n] <= 0.0f) full[n] = vmin;
}

static int feat_dims(int nx, int ny, int nz, int &X, int &Y, int &Z) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("grid dimensions must be positive (%d, %d, %d)", nx, ny, nz);
    return PRT_E_ARG;
  }
  if ((long long)nx * ny * nz >= (1LL << 31)) {
    set_error("grid too large");
    return PRT_E_RANGE;
  }
  X = nx;
  Y = ny;
  Z = nz;
  return PRT_OK;
}

}  // namespace prt

using namespace prt;

extern "C" int64_t prt_harmonic_work_size(int nx, int ny, int nz) {
  // bytes: four FP64 box vectors, the degree bytes, the reach bytes, the
  // block partials (two per block) and the four scalars plus a flag
  const long long N = (long long)nx * ny * nz;
  if (N < 1) return -1;
  const long long nb = (N + kFeatBlock - 1) / kFeatBlock;
  return 32 * N + 2 * ((N + 7) / 8 * 8) + 16 * nb + 64;
}

extern "C" int prt_harmonic_pressure(const uint8_t *pore, int nx, int ny, int nz, void *work, int64_t work_bytes,
                                     double tol, int max_iter, float *out, int *dropped_host, int *iters_host,
                                     void *stream) {
  int X, Y, Z;
  int rc = feat_dims(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!pore || !work || !out || !(tol > 0.0) || max_iter < 1 ||
      work_bytes < prt_harmonic_work_size(nx, ny, nz)) {
    set_error("prt_harmonic_pressure: bad arguments or work buffer too small");
    return PRT_E_ARG;
  }
  const long long N = (long long)X * Y * Z, nb = (N + kFeatBlock - 1) / kFeatBlock;
  char *w = static_cast<char *>(work);
  double *xv = reinterpret_cast<double *>(w);
  double *r = xv + N, *p = r + N, *q = p + N;
  uint8_t *deg = reinterpret_cast<uint8_t *>(q + N);
  uint8_t *live = deg + (N + 7) / 8 * 8;
  double *part = reinterpret_cast<double *>(live + (N + 7) / 8 * 8);
  double *s = part + 2 * nb;
  int *flag = reinterpret_cast<int *>(s + 4);
  hipStream_t st = (hipStream_t)stream;
  // the pore face-connected to the inlet (3D/tools/harmonic_pressure.py:76-88)
  if ((rc = inlet_reach(pore, X, Y, Z, 1, live, flag, 1 << 30, st))) return rc;
  const unsigned g = fblocks(N);
  hp_setup_kernel<<<g, kFeatBlock, 0, st>>>(live, X, Y, Z, deg, xv, r);
  hp_start_kernel<<<g, kFeatBlock, 0, st>>>(deg, r, N, p, part);
  hp_scalar_kernel<<<1, kSumThreads, 0, st>>>(part, nb, 2, 0, s);
  if ((rc = launch_status("prt_harmonic_pressure setup"))) return rc;
  double h[2];
  hipError_t e = hipMemcpyAsync(h, s, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "prt_harmonic_pressure copy");
  const double stop = tol * tol * h[1];  // |r|^2 <= tol^2 |b|^2
  int it = 0;
  bool done = h[1] == 0.0;
  while (!done && it < max_iter) {
    const int chunk = (max_iter - it) < 32 ? (max_iter - it) : 32;
    for (int k = 0; k < chunk; ++k) {
      hp_apply_kernel<<<g, kFeatBlock, 0, st>>>(deg, p, X, Y, Z, q, part);
      hp_scalar_kernel<<<1, kSumThreads, 0, st>>>(part, nb, 1, 1, s);
      hp_update_kernel<<<g, kFeatBlock, 0, st>>>(deg, p, q, s, N, xv, r, part);
      hp_scalar_kernel<<<1, kSumThreads, 0, st>>>(part, nb, 2, 2, s);
      hp_direction_kernel<<<g, kFeatBlock, 0, st>>>(deg, r, s, N, p);
    }
    if ((rc = launch_status("prt_harmonic_pressure iteration"))) return rc;
    it += chunk;
    e = hipMemcpyAsync(h, s, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "prt_harmonic_pressure copy");
    done = h[1] <= stop;
  }
  if (!done) {
    set_error("prt_harmonic_pressure: not converged to %g in %d iterations", tol, max_iter);
    return PRT_E_RANGE;
  }
  hp_field_kernel<<<g, kFeatBlock, 0, st>>>(live, deg, xv, X, Y, Z, out);
  if ((rc = launch_status("prt_harmonic_pressure field"))) return rc;
  if (iters_host) *iters_host = it;
  if (dropped_host) {
    // pore voxels the inlet does not reach: reported, and left at 0
    e = hipMemsetAsync(flag, 0, sizeof(int), st);
    if (e != hipSuccess) return hip_fail(e, "prt_harmonic_pressure memset");
    hp_dropped_kernel<<<g, kFeatBlock, 0, st>>>(pore, live, N, flag);
    if ((rc = launch_status("prt_harmonic_pressure count"))) return rc;
    e = hipMemcpyAsync(dropped_host, flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "prt_harmonic_pressure count copy");
  }
  return PRT_OK;
}

extern "C" int prt_harmonic_gradient(const uint8_t *pore, int nx, int ny, int nz, int nd, const float *P,
                                     int normalise, float *out, float *scratch, void *stream) {
  int X, Y, Z;
  int rc = feat_dims(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!pore || !P || !out || !scratch || (nd != 2 && nd != 3) || (nd == 2 && ny != 1)) {
    set_error("prt_harmonic_gradient: bad arguments (a 2D grid is passed as (nx, 1, ny))");
    return PRT_E_ARG;
  }
  const long long N = (long long)X * Y * Z;
  const unsigned g = fblocks(N);
  hipStream_t st = (hipStream_t)stream;
  float *bmax = scratch, *m = scratch + g;
  hg_kernel<<<g, kFeatBlock, 0, st>>>(pore, P, nd, X, Y, Z, out, bmax);
  if (normalise) {
    max_kernel<<<1, kSumThreads, 0, st>>>(bmax, g, m);
    div_kernel<<<fblocks((long long)nd * N), kFeatBlock, 0, st>>>(out, (long long)nd * N, m);
  }
  return launch_status("prt_harmonic_gradient launch");
}

extern "C" int prt_uprm(const uint8_t *pore, int nx, int ny, int nz, const float *edt, float *best, int *scratch1,
                        int max_sweeps, int *sweeps_host, void *stream) {
  int X, Y, Z;
  int rc = feat_dims(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!pore || !edt || !best || !scratch1 || max_sweeps < 1) {
    set_error("prt_uprm: bad arguments");
    return PRT_E_ARG;
  }
  const long long N = (long long)X * Y * Z;
  hipStream_t st = (hipStream_t)stream;
  uprm_init_kernel<<<fblocks(N), kFeatBlock, 0, st>>>(pore, edt, N, (long long)Y * Z, best);
  if ((rc = launch_status("prt_uprm init"))) return rc;
  int done = 0;
  for (;;) {
    hipError_t e = hipMemsetAsync(scratch1, 0, sizeof(int), st);
    if (e != hipSuccess) return hip_fail(e, "prt_uprm memset");
    for (int k = 0; k < 16; ++k) uprm_relax_kernel<<<fblocks(N), kFeatBlock, 0, st>>>(pore, edt, X, Y, Z, best, scratch1);
    if ((rc = launch_status("prt_uprm sweep"))) return rc;
    done += 16;
    int h = 0;
    e = hipMemcpyAsync(&h, scratch1, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "prt_uprm flag copy");
    if (!h) break;
    if (done >= max_sweeps) {
      set_error("prt_uprm: not converged in %d sweeps", max_sweeps);
      return PRT_E_RANGE;
    }
  }
  if (sweeps_host) *sweeps_host = done;
  return PRT_OK;
}

extern "C" int prt_dw2(const uint8_t *pore, int64_t n, const float *edt, float lo, float span, float *out,
                       void *stream) {
  if (!pore || !edt || !out || n < 1 || !(span > 0.0f)) {
    set_error("prt_dw2: bad arguments");
    return PRT_E_ARG;
  }
  dw2_kernel<<<fblocks(n), kFeatBlock, 0, (hipStream_t)stream>>>(pore, edt, n, lo, span, out);
  return launch_status("prt_dw2 launch");
}

extern "C" int prt_local_thickness(const uint8_t *pore, int nx, int ny, int nz, const float *edt, double rmax,
                                   double r_step, double *rc_work, float *out, void *stream) {
  int X, Y, Z;
  int rc = feat_dims(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!pore || !edt || !rc_work || !out || !(r_step > 0.0) || rmax < 0.0 || rmax > 1e6) {
    set_error("prt_local_thickness: bad arguments");
    return PRT_E_ARG;
  }
  const long long N = (long long)X * Y * Z;
  hipStream_t st = (hipStream_t)stream;
  // the sweep's radii: numpy.arange(rmax, 0, -r_step)
  const int K = rmax > 0.0 ? (int)ceil(rmax / r_step) : 0;
  const int R = (int)ceil(rmax);
  lt_centre_kernel<<<fblocks(N), kFeatBlock, 0, st>>>(pore, edt, N, rmax, r_step, K, rc_work);
  lt_gather_kernel<<<fblocks(N), kFeatBlock, 0, st>>>(pore, rc_work, X, Y, Z, R, out);
  return launch_status("prt_local_thickness launch");
}

extern "C" int prt_paint_buffers(const uint8_t *pore, int nx, int ny, int nz, int buf, float vmax, float vmin,
                                 float *full, void *stream) {
  int X, Y, Z;
  int rc = feat_dims(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!pore || !full || buf < 1 || 2 * buf >= X || vmax < 0.0f || vmax > 1e6f) {
    set_error("prt_paint_buffers: bad arguments");
    return PRT_E_ARG;
  }
  const long long N = (long long)X * Y * Z;
  hipStream_t st = (hipStream_t)stream;
  paint_kernel<<<fblocks(N), kFeatBlock, 0, st>>>(X, Y, Z, buf, (int)ceilf(vmax), full);
  mis_finish_kernel<<<fblocks(N), kFeatBlock, 0, st>>>(pore, X, (long long)Y * Z, buf, vmin, full);
  return launch_status("prt_paint_buffers launch");
}
