// prt_adr.hip -- fused finite-volume advection-diffusion-reaction step.
//
// Two launches of the same per-voxel update (prt_adr_core.h):
//  - dense: one thread per box voxel; everything the reference builds once
//    per solve as full-grid face tables (open/far faces, face velocities,
//    harmonic face diffusivities, div u, surface weights;
//    3D/tools/prtlb_3d.py:1100-1158) is recomputed in registers from one
//    flag byte per voxel, the cell-centred velocity and a per-species 2x2
//    diffusivity table;
//  - pore list (the default): one lane per pore voxel, its geometry from a
//    pore table built once per solve (aligned per-voxel records), the
//    radius-2 neighbours from L1/L2.
//
// The float32 instantiation reproduces the reference's float32 NumPy
// arithmetic bit for bit: the same operations in the same order, no
// contraction (this file is compiled with -ffp-contract=off and every
// function also pins it), IEEE division, NumPy's NaN-propagating
// minimum/maximum, and edge-replicated (clamped) neighbours.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "prt_adr_core.h"

namespace prt {

template <typename R, int ND, int NSM>
__global__ __launch_bounds__(256) void adr_step_kernel(AdrArgs<R> a) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int x = blockIdx.z;
  if (z >= a.Z || y >= a.Y) return;
  R cnew[NSM];
  adr_voxel<R, ND, NSM>(a, (uint32_t)((x * a.Y + y) * a.Z + z), x, y, z, cnew);
}

// The pore-list launch: one thread per list entry, so no lane is spent on
// solid; the geometry comes from the pore table.  Same update code as the
// dense launch, so the same bits.
template <typename R, int ND, int NSM, bool NTT>
__global__ __launch_bounds__(256) void adr_list_kernel(AdrArgs<R> a, const void *__restrict__ tab, uint32_t ta,
                                                      uint32_t npore) {
  const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (k >= npore) return;
  R cnew[NSM];
  adr_entry<R, ND, NSM, NTT>(a, tab, ta, k, cnew);
}

// Pore table of a list: the record of every entry, or kPoreWordNone (and
// zero faces) for an entry that is not a pore voxel.
template <typename R, int ND>
__global__ __launch_bounds__(256) void adr_table_kernel(AdrArgs<R> a, const int32_t *__restrict__ pidx,
                                                       uint32_t npore, uint32_t ta, void *__restrict__ tab) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npore) return;
  const int n = pidx[k];
  VoxGeom<R, ND> g;
  g.w = kPoreWordNone;
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) g.u_lo[ax] = g.u_hi[ax] = R(0);
  if (n >= 0 && (long long)n < a.N) {
    const uint8_t fl = a.flags[n];
    if (fl & 1) {
      int x, y, z;
      adr_coords(a, n, x, y, z);
      geom_dense<R, ND>(a, (uint32_t)n, fl, x, y, z, g);
    }
  }
  tab_store<R, ND>(tab, ta, k, (uint32_t)n, g.w, g);
}

// ------------------------------------------------------------------ host side
template <typename R>
static int build_args(const prt_adr_params *p, int nx, int ny, int nz, const uint8_t *flags, const R *u,
                      const R *cin, R *cout, AdrArgs<R> &a, dim3 &grid, dim3 &block) {
  if (!flags || !u || !cin || !cout || cin == cout) {
    set_error("prt_adr_step: null or aliased buffers");
    return PRT_E_ARG;
  }
  const int rc = adr_args_from_params<R>(p, nx, ny, nz, a, grid, block);
  if (rc) return rc;
  if (p->immobile_list) {
    set_error("prt_adr_step: immobile_list needs the pore-list launch");
    return PRT_E_ARG;
  }
  a.flags = flags; a.u = u; a.cin = cin; a.cout = cout;
  return PRT_OK;
}

template <typename R, int ND>
static void launch_nd(const AdrArgs<R> &a, dim3 g, dim3 b, hipStream_t st) {
  if (a.ns == 1) adr_step_kernel<R, ND, 1><<<g, b, 0, st>>>(a);
  else if (a.ns == 2) adr_step_kernel<R, ND, 2><<<g, b, 0, st>>>(a);
  else if (a.ns <= 4) adr_step_kernel<R, ND, 4><<<g, b, 0, st>>>(a);
  else adr_step_kernel<R, ND, 8><<<g, b, 0, st>>>(a);
}

template <typename R>
static int launch(const prt_adr_params *p, const AdrArgs<R> &a, dim3 g, dim3 b, hipStream_t st) {
  if (p->nd == 2) launch_nd<R, 2>(a, g, b, st);
  else launch_nd<R, 3>(a, g, b, st);
  return launch_status("prt_adr_step launch");
}

template <typename R, int ND, bool NTT>
static void launch_list_ndt(const AdrArgs<R> &a, const void *tab, uint32_t npore, hipStream_t st) {
  const dim3 g((npore + 255) / 256), b(256);
  const uint32_t ta = (uint32_t)tab_entries(npore);
  if (a.ns == 1) adr_list_kernel<R, ND, 1, NTT><<<g, b, 0, st>>>(a, tab, ta, npore);
  else if (a.ns == 2) adr_list_kernel<R, ND, 2, NTT><<<g, b, 0, st>>>(a, tab, ta, npore);
  else if (a.ns <= 4) adr_list_kernel<R, ND, 4, NTT><<<g, b, 0, st>>>(a, tab, ta, npore);
  else adr_list_kernel<R, ND, 8, NTT><<<g, b, 0, st>>>(a, tab, ta, npore);
}

template <typename R, int ND>
static void launch_list_nd(const AdrArgs<R> &a, const void *tab, uint32_t npore, hipStream_t st) {
  if (npore >= kNtTableEntries) launch_list_ndt<R, ND, true>(a, tab, npore, st);
  else launch_list_ndt<R, ND, false>(a, tab, npore, st);
}

// Arguments of a list launch: parameters, grid and table.
template <typename R>
static int list_args(const char *who, const prt_adr_params *p, int nx, int ny, int nz, const void *tab,
                     int64_t n_pore, const R *cin, R *cout, AdrArgs<R> &a) {
  dim3 g, b;
  const int rc = adr_args_from_params<R>(p, nx, ny, nz, a, g, b);
  if (rc) return rc;
  if (!cin || !cout || cin == cout || n_pore < 0 || n_pore > a.N || (n_pore > 0 && !tab)) {
    set_error("%s: null or aliased buffers, or a bad pore count", who);
    return PRT_E_ARG;
  }
  if ((unsigned long long)tab_entries(n_pore) * 16ull >= (1ull << 32)) {
    set_error("%s: pore list too long for the 32-bit table offsets", who);
    return PRT_E_RANGE;
  }
  if (p->write_solid) {
    set_error("%s: write_solid needs the dense launch", who);
    return PRT_E_ARG;
  }
  a.cin = cin;
  a.cout = cout;
  return PRT_OK;
}

template <typename R>
static int table(int nd, int nx, int ny, int nz, const uint8_t *flags, const R *u, const int32_t *pidx,
                 int64_t n_pore, void *tab, hipStream_t st) {
  AdrArgs<R> a;
  const int rc = adr_grid<R>(nd, nx, ny, nz, a);
  if (rc) return rc;
  if (n_pore < 0 || n_pore > a.N || (n_pore > 0 && (!flags || !u || !pidx || !tab))) {
    set_error("prt_adr_pore_table: null buffers or a bad pore count");
    return PRT_E_ARG;
  }
  if ((unsigned long long)tab_entries(n_pore) * 16ull >= (1ull << 32)) {
    set_error("prt_adr_pore_table: pore list too long for the 32-bit table offsets");
    return PRT_E_RANGE;
  }
  if (n_pore == 0) return PRT_OK;
  a.flags = flags;
  a.u = u;
  const uint32_t ta = (uint32_t)tab_entries(n_pore);
  const dim3 g((unsigned)((n_pore + 255) / 256)), b(256);
  if (nd == 2) adr_table_kernel<R, 2><<<g, b, 0, st>>>(a, pidx, (uint32_t)n_pore, ta, tab);
  else adr_table_kernel<R, 3><<<g, b, 0, st>>>(a, pidx, (uint32_t)n_pore, ta, tab);
  return launch_status("prt_adr_pore_table launch");
}

}  // namespace prt

using namespace prt;

extern "C" int prt_adr_step(const prt_adr_params *p, int nx, int ny, int nz, const uint8_t *flags, const float *u,
                            const float *c_in, float *c_out, void *stream) {
  AdrArgs<float> a;
  dim3 g, b;
  int rc = build_args<float>(p, nx, ny, nz, flags, u, c_in, c_out, a, g, b);
  if (rc) return rc;
  return launch<float>(p, a, g, b, (hipStream_t)stream);
}

extern "C" int prt_adr_step_f64(const prt_adr_params *p, int nx, int ny, int nz, const uint8_t *flags,
                                const double *u, const double *c_in, double *c_out, void *stream) {
  AdrArgs<double> a;
  dim3 g, b;
  int rc = build_args<double>(p, nx, ny, nz, flags, u, c_in, c_out, a, g, b);
  if (rc) return rc;
  return launch<double>(p, a, g, b, (hipStream_t)stream);
}

extern "C" int prt_adr_run(const prt_adr_params *p, int nx, int ny, int nz, const uint8_t *flags, const float *u,
                           float *c_a, float *c_b, int steps, void *stream) {
  AdrArgs<float> a;
  dim3 g, b;
  int rc = build_args<float>(p, nx, ny, nz, flags, u, c_a, c_b, a, g, b);
  if (rc) return rc;
  if (steps < 0) {
    set_error("prt_adr_run: negative step count");
    return PRT_E_ARG;
  }
  for (int s = 0; s < steps; ++s) {
    a.cin = (s & 1) ? c_b : c_a;
    a.cout = (s & 1) ? c_a : c_b;
    rc = launch<float>(p, a, g, b, (hipStream_t)stream);
    if (rc) return rc;
  }
  return PRT_OK;
}

extern "C" int64_t prt_adr_table_bytes(int nd, int fp64, int64_t n_pore) {
  if ((nd != 2 && nd != 3) || n_pore < 0) {
    set_error("prt_adr_table_bytes: nd must be 2 or 3 and n_pore >= 0");
    return -1;
  }
  return tab_bytes(nd, fp64 != 0, n_pore);
}

extern "C" int prt_adr_pore_table(int nd, int nx, int ny, int nz, const uint8_t *flags, const float *u,
                                  const int32_t *pore_idx, int64_t n_pore, void *table_out, void *stream) {
  return table<float>(nd, nx, ny, nz, flags, u, pore_idx, n_pore, table_out, (hipStream_t)stream);
}

extern "C" int prt_adr_pore_table_f64(int nd, int nx, int ny, int nz, const uint8_t *flags, const double *u,
                                      const int32_t *pore_idx, int64_t n_pore, void *table_out, void *stream) {
  return table<double>(nd, nx, ny, nz, flags, u, pore_idx, n_pore, table_out, (hipStream_t)stream);
}

extern "C" int prt_adr_run_list(const prt_adr_params *p, int nx, int ny, int nz, const void *pore_table,
                                int64_t n_pore, float *c_a, float *c_b, int steps, void *stream) {
  AdrArgs<float> a;
  int rc = list_args<float>("prt_adr_run_list", p, nx, ny, nz, pore_table, n_pore, c_a, c_b, a);
  if (rc) return rc;
  if (steps < 0) {
    set_error("prt_adr_run_list: negative step count");
    return PRT_E_ARG;
  }
  if (n_pore == 0) return PRT_OK;
  for (int s = 0; s < steps; ++s) {
    a.cin = (s & 1) ? c_b : c_a;
    a.cout = (s & 1) ? c_a : c_b;
    if (p->nd == 2) launch_list_nd<float, 2>(a, pore_table, (uint32_t)n_pore, (hipStream_t)stream);
    else launch_list_nd<float, 3>(a, pore_table, (uint32_t)n_pore, (hipStream_t)stream);
    rc = launch_status("prt_adr_run_list launch");
    if (rc) return rc;
  }
  return PRT_OK;
}

extern "C" int prt_adr_step_list_f64(const prt_adr_params *p, int nx, int ny, int nz, const void *pore_table,
                                     int64_t n_pore, const double *c_in, double *c_out, void *stream) {
  AdrArgs<double> a;
  const int rc = list_args<double>("prt_adr_step_list_f64", p, nx, ny, nz, pore_table, n_pore, c_in, c_out, a);
  if (rc) return rc;
  if (n_pore == 0) return PRT_OK;
  if (p->nd == 2) launch_list_nd<double, 2>(a, pore_table, (uint32_t)n_pore, (hipStream_t)stream);
  else launch_list_nd<double, 3>(a, pore_table, (uint32_t)n_pore, (hipStream_t)stream);
  return launch_status("prt_adr_step_list_f64 launch");
}
