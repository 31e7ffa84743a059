// prt_geom.hip -- pore geometry on the GPU: seeded grain packs, inlet
// connectivity and biofilm coating.
//
// Grain pack.  The host draws the candidate grains (centres, radii) from the
// seeded generator in order; a grain covers a voxel when its periodic
// minimum-image distance is at most the radius.  Each voxel finds the FIRST
// candidate covering it (candidates are binned into cells, each cell's list
// in ascending index order, so the first hit ends the search); a histogram
// of first-cover indices gives the solid count after every prefix of
// candidates, and the pack is the shortest prefix that brings the pore
// fraction to the target -- exactly what adding grains one at a time until
// the porosity is reached produces, without the serial loop.  Distances are
// formed with the host generator's operations and order (contraction off),
// so the mask is bit-identical.
//
// Connectivity (3D/tools/prtlb_3d.py:530-564): pore voxels face-connected to
// a pore voxel on x = 0 are found by repeated face dilation restricted to
// the pore until nothing changes; the rest of the pore becomes SOLID.
// Coating (:567-593): `layers` face dilations of the non-pore, no wrap.
//
// Geodesic distance from the inlet (3D/tools/build_practice_dataset.py:
// 53-72): the reference's relaxation, pass for pass -- per sweep, for each
// axis and each shift direction in turn, every pore voxel takes
// min(d, d(upstream neighbour) + 1) with solid and out-of-grid neighbours
// read as 1e9, all in float32; sweeps stop after the first one that changes
// nothing, or after 4 nx sweeps.  Each pass is a Jacobi update from the
// previous pass's array, as the reference's whole-array roll is.
//
// Euclidean distance transform of the pore (scipy.ndimage.
// distance_transform_edt(g == PORE), used at 3D/tools/build_dataset_3d.py:659):
// exact squared distances to the nearest non-pore voxel by three separable
// minimisations (z, then y, then x) in integers, then sqrt in FP64 rounded
// to float32, which is what the reference stores.
#include <hip/hip_runtime.h>

#include <cmath>

#include "prt_common.h"

namespace prt {

__global__ __launch_bounds__(256) void first_cover_kernel(int nd, int X, int Y, int Z, const double *centres,
                                                          const double *radii, const int32_t *cell_start,
                                                          const int32_t *cell_items, int cell, int cx, int cy,
                                                          int cz, int K, int32_t *first) {
#pragma clang fp contract(off)
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  if (n >= N) return;
  const int x = (int)(n / ((long long)Y * Z));
  const int r = (int)(n - (long long)x * Y * Z);
  const int y = r / Z, z = r - (r / Z) * Z;
  // a 2D grid is (X, 1, Z): its two axes are x and z
  const int c = ((x / cell) * cy + (y / cell)) * cz + (z / cell);
  int best = K;
  for (int t = cell_start[c]; t < cell_start[c + 1]; ++t) {
    const int k = cell_items[t];
    const double *ce = centres + (size_t)k * nd;
    double d2;
    if (nd == 3) {
      double d = fabs((double)x - ce[0]);
      d = fmin(d, (double)X - d);
      d2 = d * d;
      d = fabs((double)y - ce[1]);
      d = fmin(d, (double)Y - d);
      d2 = d2 + d * d;
      d = fabs((double)z - ce[2]);
      d = fmin(d, (double)Z - d);
      d2 = d2 + d * d;
    } else {
      double d = fabs((double)x - ce[0]);
      d = fmin(d, (double)X - d);
      d2 = d * d;
      d = fabs((double)z - ce[1]);
      d = fmin(d, (double)Z - d);
      d2 = d2 + d * d;
    }
    const double rr = radii[k];
    if (d2 <= rr * rr) {
      best = k;
      break;
    }
  }
  first[n] = best;
}

__global__ __launch_bounds__(256) void cover_hist_kernel(const int32_t *first, long long N, int32_t *hist) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) atomicAdd(&hist[first[n]], 1);
}

__global__ __launch_bounds__(256) void pack_mask_kernel(const int32_t *first, long long N, int kstar, uint8_t solid_code,
                                                        uint8_t pore_code, uint8_t *g) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) g[n] = first[n] < kstar ? solid_code : pore_code;
}

// reach = pore voxels on x = 0
__global__ __launch_bounds__(256) void reach_init_kernel(const uint8_t *g, uint8_t pore, int X, long long plane,
                                                         uint8_t *reach) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)X * plane) return;
  reach[n] = (n < plane && g[n] == pore) ? 1 : 0;
}

// One face dilation of `reach` inside the pore (no wrap); *changed set when
// any voxel joins.  Red-black independent: reads and writes the same array,
// which only ever turns 0 -> 1, so any interleaving converges to the same
// fixed point.
__global__ __launch_bounds__(256) void reach_grow_kernel(const uint8_t *g, uint8_t pore, int X, int Y, int Z,
                                                         uint8_t *reach, int *changed) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  if (n >= N || reach[n] || g[n] != pore) return;
  const long long yz = (long long)Y * Z;
  const int x = (int)(n / yz);
  const int r = (int)(n - x * yz);
  const int y = r / Z, z = r - (r / Z) * Z;
  const bool hit = (x > 0 && reach[n - yz]) || (x < X - 1 && reach[n + yz]) || (y > 0 && reach[n - Z]) ||
                   (y < Y - 1 && reach[n + Z]) || (z > 0 && reach[n - 1]) || (z < Z - 1 && reach[n + 1]);
  if (hit) {
    reach[n] = 1;
    *changed = 1;
  }
}

__global__ __launch_bounds__(256) void reach_apply_kernel(uint8_t *g, uint8_t pore, uint8_t solid, const uint8_t *reach,
                                                          long long N, int *removed) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N && g[n] == pore && !reach[n]) {
    g[n] = solid;
    atomicAdd(removed, 1);
  }
}

// dst = src dilated by the face cross (no wrap, outside = false)
__global__ __launch_bounds__(256) void dilate_kernel(const uint8_t *src, int X, int Y, int Z, uint8_t *dst) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  if (n >= N) return;
  const long long yz = (long long)Y * Z;
  const int x = (int)(n / yz);
  const int r = (int)(n - x * yz);
  const int y = r / Z, z = r - (r / Z) * Z;
  dst[n] = src[n] || (x > 0 && src[n - yz]) || (x < X - 1 && src[n + yz]) || (y > 0 && src[n - Z]) ||
           (y < Y - 1 && src[n + Z]) || (z > 0 && src[n - 1]) || (z < Z - 1 && src[n + 1]);
}

__global__ __launch_bounds__(256) void nonpore_kernel(const uint8_t *g, uint8_t pore, long long N, uint8_t *out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) out[n] = g[n] != pore;
}

__global__ __launch_bounds__(256) void coat_apply_kernel(uint8_t *g, uint8_t pore, uint8_t code, const uint8_t *grown,
                                                         long long N, int *marked) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N && grown[n] && g[n] == pore) {
    g[n] = code;
    atomicAdd(marked, 1);
  }
}

// ---- geodesic relaxation
constexpr float kGeoInf = 1e9f;

__global__ __launch_bounds__(256) void geo_init_kernel(const uint8_t *g, uint8_t pore, long long plane, long long N,
                                                       float *d) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  d[n] = g[n] == pore ? (n < plane ? 0.0f : kGeoInf) : __builtin_nanf("");
}

// One (axis, shift) pass: dout = min(din, din[upstream] + 1) on pore voxels,
// the upstream voxel being the one at -sh along the axis; *changed is set
// when a value drops.
__global__ __launch_bounds__(256) void geo_pass_kernel(const uint8_t *g, uint8_t pore, int X, int Y, int Z, int ax,
                                                       int sh, const float *din, float *dout, int *changed) {
#pragma clang fp contract(off)
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  if (n >= N) return;
  if (g[n] != pore) {
    dout[n] = __builtin_nanf("");
    return;
  }
  const long long yz = (long long)Y * Z;
  const int x = (int)(n / yz);
  const int r = (int)(n - x * yz);
  const int y = r / Z, z = r - (r / Z) * Z;
  const int p = ax == 0 ? x : (ax == 1 ? y : z);
  const int ext = ax == 0 ? X : (ax == 1 ? Y : Z);
  const long long stride = ax == 0 ? yz : (ax == 1 ? (long long)Z : 1);
  float nb = kGeoInf;
  const int q = p - sh;
  if (q >= 0 && q < ext) {
    const float v = din[n - (long long)sh * stride];
    nb = v != v ? kGeoInf : v;
  }
  const float d = din[n];
  const float cand = nb + 1.0f;
  const float out = cand < d ? cand : d;
  dout[n] = out;
  if (out != d) *changed = 1;
}

__global__ __launch_bounds__(256) void geo_finish_kernel(const float *d, long long N, float *out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) out[n] = d[n] != d[n] ? 0.0f : d[n];
}

// ---- exact Euclidean distance transform (squared, int32)
constexpr int kEdtInf = 1 << 30;

// D[n] = min over z' of the row: non-pore at z' ? (z - z')^2 : inf.
__global__ __launch_bounds__(256) void edt_z_kernel(const uint8_t *g, uint8_t pore, int Z, long long N, int *D) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const long long row = n / Z;
  const int z = (int)(n - row * Z);
  const uint8_t *gr = g + row * Z;
  int best = kEdtInf;
  for (int q = 0; q < Z; ++q)
    if (gr[q] != pore) {
      const int d = z - q;
      best = d * d < best ? d * d : best;
    }
  D[n] = best;
}

// Dout[n] = min over the axis line through n of Din[line k] + (p - k)^2.
__global__ __launch_bounds__(256) void edt_line_kernel(const int *Din, int X, int Y, int Z, int ax, int *Dout) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)X * Y * Z;
  if (n >= N) return;
  const long long yz = (long long)Y * Z;
  const int x = (int)(n / yz);
  const int r = (int)(n - x * yz);
  const int y = r / Z;
  const int p = ax == 0 ? x : y;
  const int ext = ax == 0 ? X : Y;
  const long long stride = ax == 0 ? yz : (long long)Z;
  const int *line = Din + (n - (long long)p * stride);
  int best = kEdtInf;
  for (int k = 0; k < ext; ++k) {
    const int v = line[(long long)k * stride];
    if (v >= kEdtInf) continue;
    const int d = p - k;
    const int c = v + d * d;
    best = c < best ? c : best;
  }
  Dout[n] = best;
}

__global__ __launch_bounds__(256) void edt_sqrt_kernel(const int *D, long long N, float *out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) out[n] = (float)sqrt((double)D[n]);
}

static unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

int inlet_reach(const uint8_t *g, int X, int Y, int Z, uint8_t pore_code, uint8_t *reach, int *changed,
                int max_sweeps, hipStream_t st) {
  const long long N = (long long)X * Y * Z, plane = (long long)Y * Z;
  reach_init_kernel<<<nblk(N), 256, 0, st>>>(g, pore_code, X, plane, reach);
  int rc = launch_status("reach init");
  if (rc) return rc;
  // sweeps in chunks of 16 between host checks of the change flag; each
  // sweep only ever sets voxels, so the fixed point is the reachable set
  for (int done = 0; done < max_sweeps;) {
    hipError_t e = hipMemsetAsync(changed, 0, sizeof(int), st);
    if (e != hipSuccess) return hip_fail(e, "inlet reach memset");
    for (int k = 0; k < 16; ++k) reach_grow_kernel<<<nblk(N), 256, 0, st>>>(g, pore_code, X, Y, Z, reach, changed);
    if ((rc = launch_status("reach grow"))) return rc;
    done += 16;
    int h = 0;
    e = hipMemcpyAsync(&h, changed, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "inlet reach flag copy");
    if (!h) break;
    if (done >= max_sweeps) {
      set_error("inlet reach: not converged in %d sweeps", max_sweeps);
      return PRT_E_RANGE;
    }
  }
  return PRT_OK;
}

static int dims3(int nx, int ny, int nz, int &X, int &Y, int &Z) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("grid dimensions must be positive");
    return PRT_E_ARG;
  }
  X = nx;
  Y = nz == 1 ? 1 : ny;
  Z = nz == 1 ? ny : nz;
  return PRT_OK;
}

}  // namespace prt

using namespace prt;

extern "C" int prt_grain_first_cover(int nx, int ny, int nz, const double *centres, const double *radii, int n_grains,
                                     const int32_t *cell_start, const int32_t *cell_items, int cell_edge,
                                     int32_t *first, void *stream) {
  int X, Y, Z;
  int rc = dims3(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!centres || !radii || !cell_start || !cell_items || !first || n_grains < 0 || cell_edge < 1) {
    set_error("prt_grain_first_cover: bad arguments");
    return PRT_E_ARG;
  }
  const int nd = nz == 1 ? 2 : 3;
  const int cx = (X + cell_edge - 1) / cell_edge, cy = (Y + cell_edge - 1) / cell_edge,
            cz = (Z + cell_edge - 1) / cell_edge;
  const long long N = (long long)X * Y * Z;
  first_cover_kernel<<<nblk(N), 256, 0, (hipStream_t)stream>>>(nd, X, Y, Z, centres, radii, cell_start, cell_items,
                                                              cell_edge, cx, cy, cz, n_grains, first);
  return launch_status("prt_grain_first_cover launch");
}

extern "C" int prt_grain_histogram(const int32_t *first, int64_t n_voxels, int32_t *hist, void *stream) {
  if (!first || !hist || n_voxels < 1) {
    set_error("prt_grain_histogram: bad arguments");
    return PRT_E_ARG;
  }
  cover_hist_kernel<<<nblk(n_voxels), 256, 0, (hipStream_t)stream>>>(first, n_voxels, hist);
  return launch_status("prt_grain_histogram launch");
}

extern "C" int prt_grain_mask(const int32_t *first, int64_t n_voxels, int n_grains_used, int solid_code,
                              int pore_code, uint8_t *g, void *stream) {
  if (!first || !g || n_voxels < 1) {
    set_error("prt_grain_mask: bad arguments");
    return PRT_E_ARG;
  }
  pack_mask_kernel<<<nblk(n_voxels), 256, 0, (hipStream_t)stream>>>(first, n_voxels, n_grains_used,
                                                                   (uint8_t)solid_code, (uint8_t)pore_code, g);
  return launch_status("prt_grain_mask launch");
}

extern "C" int prt_keep_inlet_connected(uint8_t *g, int nx, int ny, int nz, int pore_code, int solid_code,
                                        uint8_t *reach, int *scratch2, int max_sweeps, int *removed_host,
                                        void *stream) {
  int X, Y, Z;
  int rc = dims3(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!g || !reach || !scratch2 || !removed_host || max_sweeps < 1) {
    set_error("prt_keep_inlet_connected: bad arguments");
    return PRT_E_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const long long N = (long long)X * Y * Z;
  int *changed = scratch2, *removed = scratch2 + 1;
  if ((rc = inlet_reach(g, X, Y, Z, (uint8_t)pore_code, reach, changed, max_sweeps, st))) return rc;
  hipError_t e = hipMemsetAsync(removed, 0, sizeof(int), st);
  if (e != hipSuccess) return hip_fail(e, "prt_keep_inlet_connected memset");
  reach_apply_kernel<<<nblk(N), 256, 0, st>>>(g, (uint8_t)pore_code, (uint8_t)solid_code, reach, N, removed);
  if ((rc = launch_status("reach apply"))) return rc;
  e = hipMemcpyAsync(removed_host, removed, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "prt_keep_inlet_connected count copy");
  return PRT_OK;
}

extern "C" int prt_coat_with_biofilm(uint8_t *g, int nx, int ny, int nz, int pore_code, int code, int layers,
                                     uint8_t *work_a, uint8_t *work_b, int *scratch1, int *marked_host,
                                     void *stream) {
  int X, Y, Z;
  int rc = dims3(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!g || !work_a || !work_b || !scratch1 || !marked_host || layers < 0) {
    set_error("prt_coat_with_biofilm: bad arguments");
    return PRT_E_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const long long N = (long long)X * Y * Z;
  *marked_host = 0;
  if (layers == 0) return PRT_OK;
  nonpore_kernel<<<nblk(N), 256, 0, st>>>(g, (uint8_t)pore_code, N, work_a);
  uint8_t *src = work_a, *dst = work_b;
  for (int l = 0; l < layers; ++l) {
    dilate_kernel<<<nblk(N), 256, 0, st>>>(src, X, Y, Z, dst);
    uint8_t *t = src;
    src = dst;
    dst = t;
  }
  if ((rc = launch_status("coat dilate"))) return rc;
  hipError_t e = hipMemsetAsync(scratch1, 0, sizeof(int), st);
  if (e != hipSuccess) return hip_fail(e, "prt_coat_with_biofilm memset");
  coat_apply_kernel<<<nblk(N), 256, 0, st>>>(g, (uint8_t)pore_code, (uint8_t)code, src, N, scratch1);
  if ((rc = launch_status("coat apply"))) return rc;
  e = hipMemcpyAsync(marked_host, scratch1, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "prt_coat_with_biofilm count copy");
  return PRT_OK;
}

extern "C" int prt_geodesic(const uint8_t *g, int nx, int ny, int nz, int pore_code, float *d_a, float *d_b,
                            int *scratch1, int max_sweeps, float *out, int *sweeps_host, void *stream) {
  int X, Y, Z;
  int rc = dims3(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!g || !d_a || !d_b || !scratch1 || !out || max_sweeps < 0) {
    set_error("prt_geodesic: bad arguments");
    return PRT_E_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const long long N = (long long)X * Y * Z, plane = (long long)Y * Z;
  geo_init_kernel<<<nblk(N), 256, 0, st>>>(g, (uint8_t)pore_code, plane, N, d_a);
  if ((rc = launch_status("geodesic init"))) return rc;
  float *src = d_a, *dst = d_b;
  int done = 0;
  // sweeps in chunks between host checks; a chunk that changes nothing
  // means the relaxation had stopped, and sweeps past that are identities
  while (done < max_sweeps) {
    const int chunk = max_sweeps - done < 8 ? max_sweeps - done : 8;
    hipError_t e = hipMemsetAsync(scratch1, 0, sizeof(int), st);
    if (e != hipSuccess) return hip_fail(e, "prt_geodesic memset");
    for (int k = 0; k < chunk; ++k)
      for (int ax = 0; ax < 3; ++ax)
        for (int sh = 1; sh >= -1; sh -= 2) {
          geo_pass_kernel<<<nblk(N), 256, 0, st>>>(g, (uint8_t)pore_code, X, Y, Z, ax, sh, src, dst, scratch1);
          float *t = src;
          src = dst;
          dst = t;
        }
    if ((rc = launch_status("geodesic pass"))) return rc;
    done += chunk;
    int h = 0;
    e = hipMemcpyAsync(&h, scratch1, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "prt_geodesic flag copy");
    if (!h) break;
  }
  geo_finish_kernel<<<nblk(N), 256, 0, st>>>(src, N, out);
  if ((rc = launch_status("geodesic finish"))) return rc;
  if (sweeps_host) *sweeps_host = done;
  return PRT_OK;
}

extern "C" int prt_edt(const uint8_t *g, int nx, int ny, int nz, int pore_code, int32_t *work_a, int32_t *work_b,
                       float *out, void *stream) {
  int X, Y, Z;
  int rc = dims3(nx, ny, nz, X, Y, Z);
  if (rc) return rc;
  if (!g || !work_a || !work_b || !out || X > 16384 || Y > 16384 || Z > 16384) {
    set_error("prt_edt: bad arguments (or an edge above 16384)");
    return PRT_E_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const long long N = (long long)X * Y * Z;
  edt_z_kernel<<<nblk(N), 256, 0, st>>>(g, (uint8_t)pore_code, Z, N, work_a);
  edt_line_kernel<<<nblk(N), 256, 0, st>>>(work_a, X, Y, Z, 1, work_b);
  edt_line_kernel<<<nblk(N), 256, 0, st>>>(work_b, X, Y, Z, 0, work_a);
  edt_sqrt_kernel<<<nblk(N), 256, 0, st>>>(work_a, N, out);
  return launch_status("prt_edt launch");
}
