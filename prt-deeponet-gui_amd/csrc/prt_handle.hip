// prt_handle.hip -- solver handles over the stateless entry points, for
// hosts that prefer an object to threading buffers through every call
// (a C/C++ or FFI caller of the Stokes and transport stages).
//
// A handle keeps the caller's device pointers, the grid, the relaxation
// rates and the stream, and which population buffer is newest.  It owns no
// device memory (only its small host struct), so creating and destroying
// one never touches the device allocator; it is bound to one stream and is
// not shared between threads.  The arithmetic is the stateless functions'
// own, so results are the same bits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <new>

#include "prt_common.h"

// omega- of the TRT model from omega+ and the magic parameter,
// 3D/tools/prtlb_3d.py:130-139: 1 / (magic / (1/omega+ - 1/2) + 1/2).
extern "C" double prt_trt_minus(double omega_p, double magic) {
  const double t = 1.0 / omega_p - 0.5;
  if (!(t > 0.0)) {
    prt::set_error("prt_trt_minus: tau must exceed 1/2 (omega+ %g)", omega_p);
    return NAN;
  }
  return 1.0 / (magic / t + 0.5);
}

struct prt_lbm_solver {
  int q, storage, nx, ny, nz, flags;
  const uint8_t *solid;
  void *f[2];
  int latest;               // index of the buffer holding the newest state
  double *partials, *scalar;
  double omega_p, omega_m, force, fp16_scale;
  long long n, n_fluid;
  void *stream;
};

struct prt_adr_solver {
  prt_adr_params params;
  int nx, ny, nz;
  const void *table;
  long long n_pore;
  float *c[2];
  int latest;
  void *stream;
};

extern "C" int prt_lbm_solver_create(int q, int storage, int nx, int ny, int nz, const uint8_t *solid,
                                     long long n_fluid, void *f_a, void *f_b, double *partials, double *scalar,
                                     double tau, double magic, double force, double fp16_scale, int flags,
                                     void *stream, prt_lbm_solver **out) {
  if (!out || !solid || !f_a || !f_b || f_a == f_b || !partials || !scalar || nx < 1 || ny < 1 || nz < 1 ||
      n_fluid < 0 || (q != 9 && q != 19) || storage < 0 || storage > 2) {
    prt::set_error("prt_lbm_solver_create: bad arguments");
    return PRT_E_ARG;
  }
  const double omp = 1.0 / tau;
  const double omm = prt_trt_minus(omp, magic);
  if (!(omm == omm)) return PRT_E_ARG;
  prt_lbm_solver *h = new (std::nothrow) prt_lbm_solver();
  if (!h) {
    prt::set_error("prt_lbm_solver_create: out of host memory");
    return PRT_E_ARG;
  }
  h->q = q; h->storage = storage; h->nx = nx; h->ny = ny; h->nz = nz; h->flags = flags;
  h->solid = solid; h->f[0] = f_a; h->f[1] = f_b; h->latest = 0;
  h->partials = partials; h->scalar = scalar;
  h->omega_p = omp; h->omega_m = omm; h->force = force; h->fp16_scale = fp16_scale > 0 ? fp16_scale : 1024.0;
  h->n = (long long)nx * ny * nz; h->n_fluid = n_fluid;
  h->stream = stream;
  *out = h;
  return PRT_OK;
}

extern "C" int prt_lbm_solver_init(prt_lbm_solver *h) {
  if (!h) return PRT_E_ARG;
  int rc = prt_lbm_fill_rest(h->q, h->storage, h->nx, h->ny, h->nz, h->f[0], h->stream);
  if (!rc) rc = prt_lbm_fill_rest(h->q, h->storage, h->nx, h->ny, h->nz, h->f[1], h->stream);
  h->latest = 0;
  return rc;
}

extern "C" int prt_lbm_solver_step_n(prt_lbm_solver *h, int n) {
  if (!h || n < 0) {
    prt::set_error("prt_lbm_solver_step_n: bad arguments");
    return PRT_E_ARG;
  }
  if (n == 0) return PRT_OK;
  const int rc = prt_lbm_run(h->q, h->storage, h->nx, h->ny, h->nz, h->solid, h->f[h->latest],
                             h->f[1 - h->latest], n, h->omega_p, h->omega_m, h->force, h->fp16_scale,
                             h->flags, h->partials, h->stream);
  if (!rc && (n & 1)) h->latest = 1 - h->latest;
  return rc;
}

extern "C" int prt_lbm_solver_reduce(prt_lbm_solver *h, int what, double *out_host) {
  if (!h || !out_host || (what != PRT_REDUCE_MEAN_ABS_UX && what != PRT_REDUCE_SUM_UX)) {
    prt::set_error("prt_lbm_solver_reduce: bad arguments");
    return PRT_E_ARG;
  }
  // sum exactly the partials the producing launch wrote
  const long long np = prt_lbm_launch_blocks(h->q, h->nx, h->ny, h->nz, what == PRT_REDUCE_SUM_UX ? 0 : h->flags);
  int rc;
  if (what == PRT_REDUCE_SUM_UX) {
    // the velocity kernel's per-block sums of u_x over every node
    rc = prt_lbm_velocity(h->q, h->storage, h->nx, h->ny, h->nz, h->solid, h->f[h->latest],
                          h->f[1 - h->latest], h->flags, h->force, h->fp16_scale, nullptr, nullptr,
                          h->partials, h->stream);
    if (rc) return rc;
  }
  // PRT_REDUCE_MEAN_ABS_UX reads the partials the last step_n left
  rc = prt_sum_f64(h->partials, np, h->scalar, h->stream);
  if (rc) return rc;
  double s = 0.0;
  hipError_t e = hipMemcpyAsync(&s, h->scalar, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)h->stream);
  if (e != hipSuccess) return prt::hip_fail(e, "prt_lbm_solver_reduce copy");
  if (what == PRT_REDUCE_MEAN_ABS_UX) s = h->n_fluid ? s / (double)h->n_fluid : 0.0;
  *out_host = s;
  return PRT_OK;
}

extern "C" int prt_lbm_solver_velocity(prt_lbm_solver *h, double *v64, float *v32) {
  if (!h || (!v64 && !v32)) {
    prt::set_error("prt_lbm_solver_velocity: bad arguments");
    return PRT_E_ARG;
  }
  return prt_lbm_velocity(h->q, h->storage, h->nx, h->ny, h->nz, h->solid, h->f[h->latest], h->f[1 - h->latest],
                          h->flags, h->force, h->fp16_scale, v64, v32, nullptr, h->stream);
}

extern "C" void *prt_lbm_solver_newest(prt_lbm_solver *h) { return h ? h->f[h->latest] : nullptr; }

extern "C" int prt_lbm_solver_destroy(prt_lbm_solver *h) {
  delete h;
  return PRT_OK;
}

extern "C" int prt_adr_solver_create(const prt_adr_params *p, int nx, int ny, int nz, const void *pore_table,
                                     long long n_pore, float *c_a,
                                     float *c_b, void *stream, prt_adr_solver **out) {
  if (!out || !p || !c_a || !c_b || c_a == c_b || n_pore < 0 || (n_pore > 0 && !pore_table)) {
    prt::set_error("prt_adr_solver_create: bad arguments");
    return PRT_E_ARG;
  }
  prt_adr_solver *h = new (std::nothrow) prt_adr_solver();
  if (!h) {
    prt::set_error("prt_adr_solver_create: out of host memory");
    return PRT_E_ARG;
  }
  h->params = *p;
  h->nx = nx; h->ny = ny; h->nz = nz;
  h->table = pore_table; h->n_pore = n_pore;
  h->c[0] = c_a; h->c[1] = c_b; h->latest = 0;
  h->stream = stream;
  // validate once through a zero-step run (argument checks only)
  const int rc = prt_adr_run_list(&h->params, nx, ny, nz, pore_table, n_pore, c_a, c_b, 0, stream);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return PRT_OK;
}

extern "C" int prt_adr_solver_step_n(prt_adr_solver *h, int n) {
  if (!h || n < 0) {
    prt::set_error("prt_adr_solver_step_n: bad arguments");
    return PRT_E_ARG;
  }
  const int rc = prt_adr_run_list(&h->params, h->nx, h->ny, h->nz, h->table, h->n_pore,
                                  h->c[h->latest], h->c[1 - h->latest], n, h->stream);
  if (!rc && (n & 1)) h->latest = 1 - h->latest;
  return rc;
}

extern "C" float *prt_adr_solver_newest(prt_adr_solver *h) { return h ? h->c[h->latest] : nullptr; }

extern "C" int prt_adr_solver_snapshot(prt_adr_solver *h, float *dst) {
  if (!h || !dst) {
    prt::set_error("prt_adr_solver_snapshot: bad arguments");
    return PRT_E_ARG;
  }
  const size_t bytes = (size_t)h->params.ns * (size_t)h->nx * h->ny * h->nz * sizeof(float);
  const hipError_t e = hipMemcpyAsync(dst, h->c[h->latest], bytes, hipMemcpyDeviceToDevice, (hipStream_t)h->stream);
  return e == hipSuccess ? PRT_OK : prt::hip_fail(e, "prt_adr_solver_snapshot copy");
}

extern "C" int prt_adr_solver_destroy(prt_adr_solver *h) {
  delete h;
  return PRT_OK;
}
