// prt_lbm.hip -- fused TRT lattice-Boltzmann Stokes kernels for gfx950.
//
// One thread per node, pull scheme over an A/B pair of population buffers
// stored by z-row (each row holds its Q populations one after the other,
// each a contiguous run along z; prt_lbm_core.h):
// every launch gathers the Q populations that stream INTO a node (periodic on
// every axis), collides them (TRT with the magic parameter, first-order Guo
// force with omega-), applies the solid-node reversal and writes the node's
// Q post-collision populations once.  Each population is read once and
// written once per iteration: 2*Q*sizeof(pop) bytes per node update, the
// whole of the algorithmic traffic (HBM-bound, 1.3 flop/B in FP64; no MFMA).
//
// The reference's iteration (3D/tools/prtlb_3d.py:427-445) is
//   collide at every node -> solid nodes take f_opp(i) -> np.roll stream.
// Storing post-collision populations and pulling at the start of the next
// launch is the same sequence cut at a different place: f^t(x) =
// fp^{t-1}(x - e_i) is gathered, fp^t(x) is written.
//
// A 2D D2Q9 grid (nx, ny) is indexed as the 3D grid (nx, 1, ny): same bytes
// (C order), one kernel.  The D2Q9 velocity (e_x, e_y) becomes (e_x, 0, e_y).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "prt_lbm_core.h"

namespace prt {

template <int Q, int ST, bool SKIP, bool RED, int NPT, bool NTL = false, bool XCD = false>
__global__ __launch_bounds__(512) void lbm_step_kernel(LbmArgs a) {
  __shared__ double lds[16];
  uint32_t bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  if constexpr (XCD) {
    // each XCD takes a contiguous run of (z, y, x) blocks: a slab of planes
    const uint32_t L = xcd_block(bx + gridDim.x * (by + gridDim.y * bz), gridDim.x * gridDim.y * gridDim.z);
    bx = L % gridDim.x;
    by = (L / gridDim.x) % gridDim.y;
    bz = L / (gridDim.x * gridDim.y);
  }
  const int z = bx * blockDim.x + threadIdx.x;
  const int y0 = (by * blockDim.y + threadIdx.y) * NPT;
  const int x = bz + a.xoff;
  double contrib = 0.0;
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int y = y0 + k;
    if (z < a.Z && y < a.Y) {
      const int n = (x * a.Y + y) * a.Z + z;
      const bool solid = a.solid[n] != 0;
      if (!(SKIP && solid)) {
        const Nbr nb = make_nbr(x, y, z, a.X, a.Y, a.Z);
        double ux;
        const size_t self = (size_t)nb.own(Q, (int)a.zs);
        if constexpr (ST == PRT_STORE_FP64) {
          double f[Q], fp[Q];
          gather<Q, ST, SKIP, double, NTL>(a, nb, f);
          collide_ref<Q>(a, solid, f, fp, ux);
          scatter<Q, ST>(a, self, a.zs, fp);
        } else {
          float f[Q], fp[Q];
          gather<Q, ST, SKIP, float, NTL>(a, nb, f);
          collide_dev<Q>(a, solid, f, fp, ux);
          scatter<Q, ST>(a, self, a.zs, fp);
        }
        if (RED && !solid) contrib += fabs(ux);
      }
    }
  }
  if (RED) {
    const double s = block_sum(contrib, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0)
      a.partials[bx + gridDim.x * (by + gridDim.y * bz)] = s;
  }
}

// Macroscopic field of the final state (3D/tools/prtlb_3d.py:451-455).
template <int Q, int ST, bool SKIP>
__global__ __launch_bounds__(256) void lbm_velocity_kernel(LbmArgs a, double *v64, float *v32) {
#pragma clang fp contract(off)
  __shared__ double lds[16];
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int x = blockIdx.z + a.xoff;
  double contrib = 0.0;
  if (z < a.Z && y < a.Y) {
    const int n = (x * a.Y + y) * a.Z + z;
    const bool solid = a.solid[n] != 0;
    const Nbr nb = make_nbr(x, y, z, a.X, a.Y, a.Z);
    double u[3];
    if constexpr (ST == PRT_STORE_FP64) {
      double f[Q];
      gather<Q, ST, SKIP>(a, nb, f);
      double rho = f[0];
#pragma unroll
      for (int i = 1; i < Q; ++i) rho += f[i];
      double j[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        if (Lat<Q>::cx[i]) j[0] += (double)Lat<Q>::cx[i] * f[i];
        if (Lat<Q>::cy[i]) j[1] += (double)Lat<Q>::cy[i] * f[i];
        if (Lat<Q>::cz[i]) j[2] += (double)Lat<Q>::cz[i] * f[i];
      }
      u[0] = (j[0] + a.half_g) / rho;
      u[1] = j[1] / rho;
      u[2] = j[2] / rho;
    } else {
      float f[Q];
      gather<Q, ST, SKIP>(a, nb, f);
      double dr = 0.0, j[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        dr += (double)f[i];
        if (Lat<Q>::cx[i]) j[0] += (double)Lat<Q>::cx[i] * (double)f[i];
        if (Lat<Q>::cy[i]) j[1] += (double)Lat<Q>::cy[i] * (double)f[i];
        if (Lat<Q>::cz[i]) j[2] += (double)Lat<Q>::cz[i] * (double)f[i];
      }
      const double rho = 1.0 + dr;
      u[0] = (j[0] + a.half_g) / rho;
      u[1] = j[1] / rho;
      u[2] = j[2] / rho;
    }
    if (solid) u[0] = u[1] = u[2] = 0.0;
    // a 2D grid carries its second velocity component in the z slot
    const double o0 = u[0], o1 = (Q == 9) ? u[2] : u[1], o2 = (Q == 9) ? 0.0 : u[2];
    if (v64) {
      v64[n] = o0;
      v64[a.N + n] = o1;
      v64[2 * a.N + n] = o2;
    }
    if (v32) {
      v32[n] = (float)o0;
      v32[a.N + n] = (float)o1;
      v32[2 * a.N + n] = (float)o2;
    }
    contrib = o0;
  }
  if (a.partials) {
    const double s = block_sum(contrib, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0)
      a.partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = s;
  }
}

// Rest state over a whole z-row buffer (row pitch zs, `total` elements),
// run padding included.
template <int Q, int ST>
__global__ void fill_rest_kernel(void *f, long long zs, long long total) {
  using T = typename Pop<ST>::T;
  T *p = static_cast<T *>(f);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
    const int i = (int)((k / zs) % Q);
    if constexpr (ST == PRT_STORE_FP64)
      p[k] = weight<Q>(i);
    else
      p[k] = (T)0.0f;
  }
}

// Halo message of one x-plane of a slab: the populations with e_x = dir
// (5 of D3Q19, 3 of D2Q9, in lattice order), each as the plane's Y row runs
// of zs elements, message element (k*Y + y)*zs + z <-> population element
// ((plane*Y + y)*Q + i_k)*zs + z.  T is a bit container of the storage
// width: the copy moves bits.
template <int Q>
__device__ inline int halo_pop(int dir, int k) {
  int seen = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i)
    if (Lat<Q>::cx[i] == dir) {
      if (seen == k) return i;
      ++seen;
    }
  return 0;
}

template <int Q, typename T, bool PACK>
__global__ __launch_bounds__(256) void halo_kernel(T *f, int plane, int dir, int Y, long long zs, T *buf) {
  constexpr int ND = Q == 19 ? 5 : 3;
  const long long total = (long long)ND * Y * zs;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += stride) {
    const long long z = m % zs, r = m / zs;
    const int y = (int)(r % Y), k = (int)(r / Y);
    const long long e = (((long long)plane * Y + y) * Q + halo_pop<Q>(dir, k)) * zs + z;
    if constexpr (PACK) buf[m] = f[e];
    else f[e] = buf[m];
  }
}

// Deterministic sum: one block, fixed per-thread stride order, fixed combine.
__global__ __launch_bounds__(1024) void sum_f64_kernel(const double *x, long long n, double *out) {
  __shared__ double part[1024];
  double s = 0.0;
  for (long long k = threadIdx.x; k < n; k += blockDim.x) s += x[k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// One block per x-plane: sum of v64[0][x, ...] in a fixed order.
__global__ __launch_bounds__(256) void plane_flux_kernel(const double *v, int plane_n, double *out) {
  __shared__ double part[256];
  const double *p = v + (size_t)blockIdx.x * plane_n;
  double s = 0.0;
  for (int k = threadIdx.x; k < plane_n; k += blockDim.x) s += p[k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

// ------------------------------------------------------------------ host side
struct Geom {
  int X, Y, Z;
  long long N;
  dim3 block, grid;
  int npt;
  bool ntl;
  bool xcd;
  int xoff;  // PRT_LBM_SLAB: planes 1 .. X-2 are launched, 0 and X-1 are halo
};

static int make_geom(int q, int nx, int ny, int nz, Geom &g, int flags = 0) {
  if (q != 9 && q != 19) {
    set_error("q must be 9 (D2Q9) or 19 (D3Q19), got %d", q);
    return PRT_E_ARG;
  }
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("grid dimensions must be positive (%d, %d, %d)", nx, ny, nz);
    return PRT_E_ARG;
  }
  if (q == 9 && nz != 1) {
    set_error("D2Q9 takes a 2D grid: pass nz = 1 (got %d)", nz);
    return PRT_E_ARG;
  }
  if (q == 9) {
    g.X = nx; g.Y = 1; g.Z = ny;
  } else {
    g.X = nx; g.Y = ny; g.Z = nz;
  }
  g.N = (long long)nx * ny * nz;
  // every population element index (z-row layout, the widest row pitch of
  // the three storages) stays below 2^31
  if ((long long)g.X * g.Y * q * pop_row_pitch(g.Z, 2) >= (1LL << 31) || g.X > 65535) {
    set_error("grid of %lld nodes exceeds the 32-bit node index", g.N);
    return PRT_E_RANGE;
  }
  // block shape: (bz, by) threads, npt nodes per thread along y.  Default:
  // the smallest power of two >= Z (64..256) along z, 256 threads.
  int bz = 64;
  while (bz < g.Z && bz < 256) bz <<= 1;
  int by = 256 / bz;
  g.npt = 1;
  g.ntl = (flags >> 18) & 1;
  g.xcd = (flags >> 19) & 1;
  const int tz = (flags >> 8) & 15, ty = (flags >> 12) & 15, tn = (flags >> 16) & 3;
  if (tz) {
    bz = 1 << tz;
    by = 1 << ty;
    g.npt = tn + 1;
    if (bz < 1 || bz * by > 512 || bz * by < 64 || (bz * by) % 64 || g.npt > 2) {
      set_error("bad block tuning: %d x %d threads, %d nodes per thread", bz, by, g.npt);
      return PRT_E_ARG;
    }
  }
  g.block = dim3(bz, by, 1);
  const int rows = (g.Y + g.npt - 1) / g.npt;
  g.xoff = (flags & PRT_LBM_SLAB) ? 1 : 0;
  if (g.xoff && g.X < 3) {
    set_error("a slab needs at least one plane between its two halo planes (nx = %d)", g.X);
    return PRT_E_ARG;
  }
  g.grid = dim3((g.Z + bz - 1) / bz, (rows + by - 1) / by, g.X - 2 * g.xoff);
  return PRT_OK;
}

long long pop_row_pitch(int Z, int elem_bytes) {
  const long long line = 128 / elem_bytes;
  return ((long long)Z + line - 1) / line * line;
}

long long pop_elems(int q, int X, int Y, int Z, int elem_bytes) {
  return (long long)X * Y * q * pop_row_pitch(Z, elem_bytes);
}

long long pop_stride(long long n) { return (n + kListTile - 1) / kListTile * kListTile; }

static int elem_size(int storage) { return storage == PRT_STORE_FP64 ? 8 : (storage == PRT_STORE_FP32 ? 4 : 2); }

static void fill_args(LbmArgs &a, int q, const Geom &g, const uint8_t *solid, const void *src, void *dst,
                      double omp, double omm, double force, double s16, double *partials, int elem_bytes) {
  std::memset(&a, 0, sizeof(a));
  a.solid = solid; a.src = src; a.dst = dst; a.partials = partials;
  a.X = g.X; a.Y = g.Y; a.Z = g.Z; a.N = g.N;
  a.zs = pop_row_pitch(g.Z, elem_bytes);
  a.omp = omp; a.omm = omm; a.g = force; a.half_g = 0.5 * force;
  a.omh = 1 - 0.5 * omm;
  (void)q;
  a.s16 = (float)s16;
  a.inv_s16 = (float)(1.0 / s16);
  a.xoff = g.xoff;
}

template <int Q, int ST, int NPT>
static void launch_step_n(const Geom &g, const LbmArgs &a, bool skip, bool red, hipStream_t st) {
  if (skip) {
    if (red) lbm_step_kernel<Q, ST, true, true, NPT><<<g.grid, g.block, 0, st>>>(a);
    else lbm_step_kernel<Q, ST, true, false, NPT><<<g.grid, g.block, 0, st>>>(a);
  } else {
    if (red) lbm_step_kernel<Q, ST, false, true, NPT><<<g.grid, g.block, 0, st>>>(a);
    else lbm_step_kernel<Q, ST, false, false, NPT><<<g.grid, g.block, 0, st>>>(a);
  }
}

template <int Q, int ST>
static void launch_step_t(const Geom &g, const LbmArgs &a, bool skip, bool red, hipStream_t st) {
  if (g.xcd && !skip && g.npt == 1) {
    // XCD-contiguous block order (dense mode, one node per thread)
    if (g.ntl) {
      if (red) lbm_step_kernel<Q, ST, false, true, 1, true, true><<<g.grid, g.block, 0, st>>>(a);
      else lbm_step_kernel<Q, ST, false, false, 1, true, true><<<g.grid, g.block, 0, st>>>(a);
    } else {
      if (red) lbm_step_kernel<Q, ST, false, true, 1, false, true><<<g.grid, g.block, 0, st>>>(a);
      else lbm_step_kernel<Q, ST, false, false, 1, false, true><<<g.grid, g.block, 0, st>>>(a);
    }
  } else if (g.ntl && !skip && g.npt == 1) {
    // nontemporal population loads (dense mode, one node per thread)
    if (red) lbm_step_kernel<Q, ST, false, true, 1, true><<<g.grid, g.block, 0, st>>>(a);
    else lbm_step_kernel<Q, ST, false, false, 1, true><<<g.grid, g.block, 0, st>>>(a);
  } else if (g.npt == 2) {
    launch_step_n<Q, ST, 2>(g, a, skip, red, st);
  } else {
    launch_step_n<Q, ST, 1>(g, a, skip, red, st);
  }
}

static int launch_step(int q, int storage, const Geom &g, const LbmArgs &a, bool skip, bool red,
                       hipStream_t st) {
  switch (q * 4 + storage) {
    case 19 * 4 + PRT_STORE_FP64: launch_step_t<19, PRT_STORE_FP64>(g, a, skip, red, st); break;
    case 19 * 4 + PRT_STORE_FP32: launch_step_t<19, PRT_STORE_FP32>(g, a, skip, red, st); break;
    case 19 * 4 + PRT_STORE_FP16: launch_step_t<19, PRT_STORE_FP16>(g, a, skip, red, st); break;
    case 9 * 4 + PRT_STORE_FP64: launch_step_t<9, PRT_STORE_FP64>(g, a, skip, red, st); break;
    case 9 * 4 + PRT_STORE_FP32: launch_step_t<9, PRT_STORE_FP32>(g, a, skip, red, st); break;
    case 9 * 4 + PRT_STORE_FP16: launch_step_t<9, PRT_STORE_FP16>(g, a, skip, red, st); break;
    default:
      set_error("unknown storage %d", storage);
      return PRT_E_ARG;
  }
  return launch_status("prt_lbm_step launch");
}

static int check_omega(double omp, double omm, double s16, int storage) {
  if (!(omp > 0.0 && omp < 2.0) || !(omm > 0.0 && omm < 2.0)) {
    set_error("relaxation rates must lie in (0, 2): omega+ %g, omega- %g (tau must exceed 0.5)", omp, omm);
    return PRT_E_ARG;
  }
  if (storage == PRT_STORE_FP16 && !(s16 > 0.0)) {
    set_error("fp16_scale must be positive");
    return PRT_E_ARG;
  }
  return PRT_OK;
}

}  // namespace prt

using namespace prt;

extern "C" int prt_lbm_fill_rest(int q, int storage, int nx, int ny, int nz, void *f, void *stream) {
  Geom g;
  if (!f || storage < 0 || storage > 2) {
    set_error("prt_lbm_fill_rest: bad arguments");
    return PRT_E_ARG;
  }
  const int rc = make_geom(q, nx, ny, nz, g);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int blocks = 2048;
  const int eb = elem_size(storage);
  const long long zs = pop_row_pitch(g.Z, eb), tot = pop_elems(q, g.X, g.Y, g.Z, eb);
  switch (q * 4 + storage) {
    case 19 * 4 + PRT_STORE_FP64: fill_rest_kernel<19, PRT_STORE_FP64><<<blocks, 256, 0, st>>>(f, zs, tot); break;
    case 19 * 4 + PRT_STORE_FP32: fill_rest_kernel<19, PRT_STORE_FP32><<<blocks, 256, 0, st>>>(f, zs, tot); break;
    case 19 * 4 + PRT_STORE_FP16: fill_rest_kernel<19, PRT_STORE_FP16><<<blocks, 256, 0, st>>>(f, zs, tot); break;
    case 9 * 4 + PRT_STORE_FP64: fill_rest_kernel<9, PRT_STORE_FP64><<<blocks, 256, 0, st>>>(f, zs, tot); break;
    case 9 * 4 + PRT_STORE_FP32: fill_rest_kernel<9, PRT_STORE_FP32><<<blocks, 256, 0, st>>>(f, zs, tot); break;
    case 9 * 4 + PRT_STORE_FP16: fill_rest_kernel<9, PRT_STORE_FP16><<<blocks, 256, 0, st>>>(f, zs, tot); break;
    default:
      set_error("unknown storage %d", storage);
      return PRT_E_ARG;
  }
  return launch_status("prt_lbm_fill_rest launch");
}

extern "C" int64_t prt_lbm_launch_blocks(int q, int nx, int ny, int nz, int flags) {
  // partials a launch with these flags writes: one per block, indices
  // [0, blocks); the velocity kernel launches the flags = 0 shape
  Geom g;
  if (make_geom(q, nx, ny, nz, g, flags) != PRT_OK) return -1;
  return (int64_t)g.grid.x * g.grid.y * g.grid.z;
}

extern "C" int64_t prt_lbm_partials_count(int q, int nx, int ny, int nz) {
  // the largest block count any tuning of this grid launches (64 threads)
  Geom g;
  if (make_geom(q, nx, ny, nz, g) != PRT_OK) return -1;
  const long long zb = (g.Z + 63) / 64;
  return (int64_t)(zb * g.Y * g.X);
}

template <int Q>
static void lattice_copy(int32_t *c, double *w, int32_t *opp) {
  for (int i = 0; i < Q; ++i) {
    // kernels keep D2Q9 as (x, -, y) in their (x, y, z) slots
    c[3 * i] = Lat<Q>::cx[i];
    c[3 * i + 1] = Q == 9 ? Lat<Q>::cz[i] : Lat<Q>::cy[i];
    c[3 * i + 2] = Q == 9 ? 0 : Lat<Q>::cz[i];
    w[i] = weight<Q>(i);
    opp[i] = Lat<Q>::opp[i];
  }
}

extern "C" int prt_lbm_lattice(int q, int32_t *c, double *w, int32_t *opp) {
  if ((q != 9 && q != 19) || !c || !w || !opp) {
    set_error("prt_lbm_lattice: q must be 9 or 19 and the arrays non-null");
    return PRT_E_ARG;
  }
  if (q == 9) lattice_copy<9>(c, w, opp);
  else lattice_copy<19>(c, w, opp);
  return PRT_OK;
}

extern "C" int64_t prt_lbm_pop_elems(int q, int storage, int nx, int ny, int nz) {
  Geom g;
  if (storage < 0 || storage > 2) {
    set_error("prt_lbm_pop_elems: bad storage %d", storage);
    return -1;
  }
  if (make_geom(q, nx, ny, nz, g) != PRT_OK) return -1;
  return pop_elems(q, g.X, g.Y, g.Z, elem_size(storage));
}

extern "C" int64_t prt_lbm_pop_stride(int64_t n_nodes, int storage) {
  if (n_nodes < 1 || storage < 0 || storage > 2) {
    set_error("prt_lbm_pop_stride: bad arguments");
    return -1;
  }
  return pop_stride(n_nodes);
}

extern "C" int prt_lbm_step(int q, int storage, int nx, int ny, int nz, const uint8_t *solid, const void *f_src,
                            void *f_dst, double omega_p, double omega_m, double force, double fp16_scale,
                            int flags, double *partials, void *stream) {
  Geom g;
  int rc = make_geom(q, nx, ny, nz, g, flags);
  if (rc) return rc;
  if ((rc = check_omega(omega_p, omega_m, fp16_scale, storage))) return rc;
  if (!solid || !f_src || !f_dst || f_src == f_dst) {
    set_error("prt_lbm_step: null or aliased population buffers");
    return PRT_E_ARG;
  }
  LbmArgs a;
  fill_args(a, q, g, solid, f_src, f_dst, omega_p, omega_m, force, fp16_scale, partials, elem_size(storage));
  return launch_step(q, storage, g, a, (flags & PRT_LBM_SKIP_SOLID) != 0, partials != nullptr,
                     (hipStream_t)stream);
}

extern "C" int prt_lbm_run(int q, int storage, int nx, int ny, int nz, const uint8_t *solid, void *f_a, void *f_b,
                           int steps, double omega_p, double omega_m, double force, double fp16_scale, int flags,
                           double *partials, void *stream) {
  Geom g;
  int rc = make_geom(q, nx, ny, nz, g, flags);
  if (rc) return rc;
  if ((rc = check_omega(omega_p, omega_m, fp16_scale, storage))) return rc;
  if (!solid || !f_a || !f_b || f_a == f_b || steps < 0) {
    set_error("prt_lbm_run: null or aliased population buffers, or negative steps");
    return PRT_E_ARG;
  }
  LbmArgs a;
  fill_args(a, q, g, solid, f_a, f_b, omega_p, omega_m, force, fp16_scale, nullptr, elem_size(storage));
  for (int s = 0; s < steps; ++s) {
    a.src = (s & 1) ? f_b : f_a;
    a.dst = (s & 1) ? f_a : f_b;
    const bool red = partials && s == steps - 1;
    a.partials = red ? partials : nullptr;
    rc = launch_step(q, storage, g, a, (flags & PRT_LBM_SKIP_SOLID) != 0, red, (hipStream_t)stream);
    if (rc) return rc;
  }
  return PRT_OK;
}

extern "C" int prt_sum_f64(const double *x, int64_t n, double *out, void *stream) {
  if (!x || !out || n < 0) {
    set_error("prt_sum_f64: bad arguments");
    return PRT_E_ARG;
  }
  sum_f64_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(x, n, out);
  return launch_status("prt_sum_f64 launch");
}

// ---- the allocation-free convergence tick (prt_lbm_run_sum_static)
//
// The per-block |u_x| partials of the tick's last iteration go to this
// library's own device array, one copy per GPU held until the process
// ends, instead of a caller buffer.  In host code the array's name stands
// for a host-side symbol, not a device address: the device address of each
// GPU's copy is looked up once per device (hipGetSymbolAddress, that device
// current) and only that address is handed to kernels.
constexpr long long kStaticPartials = 1LL << 20;   // doubles: 8 MiB per GPU
static __device__ __align__(256) double scratch[kStaticPartials];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices];

// The current device's copy of scratch, or nullptr (error text set).
static double *scratch_on_current_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("prt_lbm_run_sum_static: no usable current device (%s)", hipGetErrorString(e));
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("prt_lbm_run_sum_static: hipGetSymbolAddress failed on device %d (%s)", dev,
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  return g_scratch_dev[dev];
}

extern "C" int64_t prt_lbm_static_partials_max(void) { return kStaticPartials; }

extern "C" int prt_lbm_run_sum_static(int q, int storage, int nx, int ny, int nz, const uint8_t *solid,
                                      void *f_a, void *f_b, int steps, double omega_p, double omega_m,
                                      double force, double fp16_scale, int flags, double *sum_out,
                                      void *stream) {
  Geom g;
  int rc = make_geom(q, nx, ny, nz, g, flags);
  if (rc) return rc;
  if (steps < 1 || !sum_out) {
    set_error("prt_lbm_run_sum_static: needs at least one step and a device output");
    return PRT_E_ARG;
  }
  const long long nblk = (long long)g.grid.x * g.grid.y * g.grid.z;
  if (nblk > kStaticPartials) {
    set_error("prt_lbm_run_sum_static: the launch has %lld blocks, more than the %lld partials of the "
              "static scratch; use prt_lbm_run with a partials buffer",
              nblk, kStaticPartials);
    return PRT_E_RANGE;
  }
  double *part = scratch_on_current_device();
  if (!part) return PRT_E_HIP;
  // every block of the last iteration writes its partial: nothing to clear
  rc = prt_lbm_run(q, storage, nx, ny, nz, solid, f_a, f_b, steps, omega_p, omega_m, force, fp16_scale, flags,
                   part, stream);
  if (rc) return rc;
  sum_f64_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(part, nblk, sum_out);
  return launch_status("prt_lbm_run_sum_static sum launch");
}

extern "C" int prt_lbm_velocity(int q, int storage, int nx, int ny, int nz, const uint8_t *solid,
                                const void *f_last, const void *f_prev, int flags, double force,
                                double fp16_scale, double *v64, float *v32, double *partials, void *stream) {
  Geom g;
  int rc = make_geom(q, nx, ny, nz, g, flags & PRT_LBM_SLAB);
  if (rc) return rc;
  const bool skip = (flags & PRT_LBM_SKIP_SOLID) != 0;
  if (!solid || !f_last || (skip && (!f_prev || f_prev == f_last)) ||
      (storage == PRT_STORE_FP16 && !(fp16_scale > 0))) {
    set_error("prt_lbm_velocity: bad arguments");
    return PRT_E_ARG;
  }
  LbmArgs a;
  // in the fluid-only mode the populations a solid neighbour would hand
  // back are a fluid node's own two-step-old slots: the other buffer
  fill_args(a, q, g, solid, f_last, const_cast<void *>(f_prev), 1.0, 1.0, force,
            fp16_scale > 0 ? fp16_scale : 1.0, partials, elem_size(storage));
  hipStream_t st = (hipStream_t)stream;
#define PRT_VK(QQ, SS)                                                                              \
  do {                                                                                              \
    if (skip) lbm_velocity_kernel<QQ, SS, true><<<g.grid, g.block, 0, st>>>(a, v64, v32);            \
    else lbm_velocity_kernel<QQ, SS, false><<<g.grid, g.block, 0, st>>>(a, v64, v32);                \
  } while (0)
  switch (q * 4 + storage) {
    case 19 * 4 + PRT_STORE_FP64: PRT_VK(19, PRT_STORE_FP64); break;
    case 19 * 4 + PRT_STORE_FP32: PRT_VK(19, PRT_STORE_FP32); break;
    case 19 * 4 + PRT_STORE_FP16: PRT_VK(19, PRT_STORE_FP16); break;
    case 9 * 4 + PRT_STORE_FP64: PRT_VK(9, PRT_STORE_FP64); break;
    case 9 * 4 + PRT_STORE_FP32: PRT_VK(9, PRT_STORE_FP32); break;
    case 9 * 4 + PRT_STORE_FP16: PRT_VK(9, PRT_STORE_FP16); break;
    default:
      set_error("unknown storage %d", storage);
      return PRT_E_ARG;
  }
#undef PRT_VK
  return launch_status("prt_lbm_velocity launch");
}

static int halo_copy(const char *who, bool pack, int q, int storage, int nx, int ny, int nz, void *f, int plane,
                     int dir, void *buf, void *stream) {
  Geom g;
  int rc = make_geom(q, nx, ny, nz, g);
  if (rc) return rc;
  if (!f || !buf || storage < 0 || storage > 2 || plane < 0 || plane >= g.X || (dir != 1 && dir != -1)) {
    set_error("%s: bad arguments (plane %d of %d, dir %d)", who, plane, g.X, dir);
    return PRT_E_ARG;
  }
  const int eb = elem_size(storage);
  const long long zs = pop_row_pitch(g.Z, eb);
  const long long total = (long long)(q == 19 ? 5 : 3) * g.Y * zs;
  const int blocks = (int)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
  hipStream_t st = (hipStream_t)stream;
#define PRT_HK(QQ, TT)                                                                                      \
  do {                                                                                                      \
    if (pack) halo_kernel<QQ, TT, true><<<blocks, 256, 0, st>>>((TT *)f, plane, dir, g.Y, zs, (TT *)buf);    \
    else halo_kernel<QQ, TT, false><<<blocks, 256, 0, st>>>((TT *)f, plane, dir, g.Y, zs, (TT *)buf);        \
  } while (0)
  if (q == 19) {
    if (eb == 8) PRT_HK(19, uint64_t);
    else if (eb == 4) PRT_HK(19, uint32_t);
    else PRT_HK(19, uint16_t);
  } else {
    if (eb == 8) PRT_HK(9, uint64_t);
    else if (eb == 4) PRT_HK(9, uint32_t);
    else PRT_HK(9, uint16_t);
  }
#undef PRT_HK
  return launch_status(who);
}

extern "C" int64_t prt_lbm_halo_elems(int q, int storage, int nx, int ny, int nz) {
  Geom g;
  if (storage < 0 || storage > 2 || make_geom(q, nx, ny, nz, g) != PRT_OK) {
    set_error("prt_lbm_halo_elems: bad arguments");
    return -1;
  }
  return (int64_t)(q == 19 ? 5 : 3) * g.Y * pop_row_pitch(g.Z, elem_size(storage));
}

extern "C" int prt_lbm_halo_pack(int q, int storage, int nx, int ny, int nz, const void *f, int plane, int dir,
                                 void *buf, void *stream) {
  return halo_copy("prt_lbm_halo_pack", true, q, storage, nx, ny, nz, const_cast<void *>(f), plane, dir, buf,
                   stream);
}

extern "C" int prt_lbm_halo_unpack(int q, int storage, int nx, int ny, int nz, void *f, int plane, int dir,
                                   const void *buf, void *stream) {
  return halo_copy("prt_lbm_halo_unpack", false, q, storage, nx, ny, nz, f, plane, dir, const_cast<void *>(buf),
                   stream);
}

extern "C" int prt_plane_flux(const double *v64, int nx, int ny, int nz, double *plane, void *stream) {
  if (!v64 || !plane || nx < 1 || ny < 1 || nz < 1) {
    set_error("prt_plane_flux: bad arguments");
    return PRT_E_ARG;
  }
  plane_flux_kernel<<<nx, 256, 0, (hipStream_t)stream>>>(v64, ny * nz, plane);
  return launch_status("prt_plane_flux launch");
}
