// prt_adr.hip -- fused finite-volume advection-diffusion-reaction step.
//
// One thread per voxel.  Everything the reference builds once per solve as
// full-grid face tables (open/far faces, face velocities, harmonic face
// diffusivities, div u, surface weights; 3D/tools/prtlb_3d.py:1100-1158) is
// recomputed in registers from one flag byte per voxel, the cell-centred
// velocity and a per-species 2x2 diffusivity table, so a step streams only
//   C in (4 B/species) + C out (4 B/species) + u (4 B/axis) + flags (1 B)
// per pore voxel; the radius-2 neighbours come from L1/L2.
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
  adr_voxel<R, ND, NSM>(a, x, y, z, cnew);
}

// One thread per PORE voxel (pidx: the pore voxels' flat indices): no lane
// of a wave is spent on solid.  Same per-voxel code, so the same bits.
template <typename R, int ND, int NSM>
__global__ __launch_bounds__(256) void adr_list_kernel(AdrArgs<R> a, const int32_t *__restrict__ pidx,
                                                      int npore) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npore) return;
  const int n = pidx[k];
  if ((unsigned long long)n >= (unsigned long long)a.N) return;
  const int yz = a.Y * a.Z;
  const int x = n / yz;
  const int r = n - x * yz;
  const int y = r / a.Z;
  const int z = r - y * a.Z;
  R cnew[NSM];
  adr_voxel<R, ND, NSM>(a, x, y, z, cnew);
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

template <typename R, int ND>
static void launch_list_nd(const AdrArgs<R> &a, const int32_t *pidx, int npore, hipStream_t st) {
  const dim3 g((npore + 255) / 256), b(256);
  if (a.ns == 1) adr_list_kernel<R, ND, 1><<<g, b, 0, st>>>(a, pidx, npore);
  else if (a.ns == 2) adr_list_kernel<R, ND, 2><<<g, b, 0, st>>>(a, pidx, npore);
  else if (a.ns <= 4) adr_list_kernel<R, ND, 4><<<g, b, 0, st>>>(a, pidx, npore);
  else adr_list_kernel<R, ND, 8><<<g, b, 0, st>>>(a, pidx, npore);
}

template <typename R>
static int launch_list(const prt_adr_params *p, const AdrArgs<R> &a, const int32_t *pidx, int npore,
                       hipStream_t st) {
  if (npore <= 0) return PRT_OK;
  if (p->nd == 2) launch_list_nd<R, 2>(a, pidx, npore, st);
  else launch_list_nd<R, 3>(a, pidx, npore, st);
  return launch_status("prt_adr_list launch");
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

extern "C" int prt_adr_run_list(const prt_adr_params *p, int nx, int ny, int nz, const uint8_t *flags,
                                const float *u, const int32_t *pore_idx, int64_t n_pore, float *c_a, float *c_b,
                                int steps, void *stream) {
  AdrArgs<float> a;
  dim3 g, b;
  int rc = build_args<float>(p, nx, ny, nz, flags, u, c_a, c_b, a, g, b);
  if (rc) return rc;
  if (steps < 0 || n_pore < 0 || n_pore > a.N || (n_pore > 0 && !pore_idx)) {
    set_error("prt_adr_run_list: bad step count or pore list");
    return PRT_E_ARG;
  }
  if (p->write_solid) {
    set_error("prt_adr_run_list: write_solid needs the dense launch (prt_adr_run)");
    return PRT_E_ARG;
  }
  for (int s = 0; s < steps; ++s) {
    a.cin = (s & 1) ? c_b : c_a;
    a.cout = (s & 1) ? c_a : c_b;
    rc = launch_list<float>(p, a, pore_idx, (int)n_pore, (hipStream_t)stream);
    if (rc) return rc;
  }
  return PRT_OK;
}

extern "C" int prt_adr_step_list_f64(const prt_adr_params *p, int nx, int ny, int nz, const uint8_t *flags,
                                     const double *u, const int32_t *pore_idx, int64_t n_pore,
                                     const double *c_in, double *c_out, void *stream) {
  AdrArgs<double> a;
  dim3 g, b;
  int rc = build_args<double>(p, nx, ny, nz, flags, u, c_in, c_out, a, g, b);
  if (rc) return rc;
  if (n_pore < 0 || n_pore > a.N || (n_pore > 0 && !pore_idx)) {
    set_error("prt_adr_step_list_f64: bad pore list");
    return PRT_E_ARG;
  }
  if (p->write_solid) {
    set_error("prt_adr_step_list_f64: write_solid needs the dense launch (prt_adr_step_f64)");
    return PRT_E_ARG;
  }
  return launch_list<double>(p, a, pore_idx, (int)n_pore, (hipStream_t)stream);
}
