// prt_sparse.hip -- the TRT lattice-Boltzmann step over a fluid-node list.
//
// The dense kernels keep every box node's populations; their fluid-only mode
// skips solid nodes but still pulls whole cache lines of the dense arrays,
// so it moves nearly the dense bytes.  Here only the fluid nodes are stored,
// compacted in C order (entry k = the k-th fluid node), with a link table
// giving, per direction, the entry of the node that streams into each entry
// (periodic wrap as the dense grid) or -1 where that node is solid.  A step
// then moves 2 Q s + 4 (Q - 1) bytes per FLUID node (FP32: 224 B; 188 B
// with the 16-bit link table), against 2 Q s per BOX node dense (FP32:
// 152 B, i.e. 152 / porosity per fluid node).
//
// The values are the dense fluid-only kernel's, bit for bit: the same gather
// rule (a solid upstream node hands back the entry's own two-step-old
// opposite slot, 3D/tools/prtlb_3d.py:442-443 reversal), the same collision
// code (prt_lbm_core.h) and the same scatter.
#include <hip/hip_runtime.h>

#include <cstring>

#include "prt_lbm_core.h"

namespace prt {

constexpr int kSparseBlock = 256;

// map[n] = -1 for every box node.
__global__ __launch_bounds__(256) void map_clear_kernel(int32_t *map, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) map[i] = -1;
}

// map[fluid_idx[k]] = k.
__global__ __launch_bounds__(256) void map_fill_kernel(const int32_t *fluid_idx, long long nf, long long n,
                                                       int32_t *map) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nf) {
    const int32_t v = fluid_idx[k];
    if (v >= 0 && v < n) map[v] = (int32_t)k;
  }
}

// link[(i - 1) * lk + k] = map[upstream node of entry k along e_i].
template <int Q>
__global__ __launch_bounds__(256) void link_build_kernel(const int32_t *fluid_idx, long long nf, int X, int Y, int Z,
                                                         const int32_t *map, long long lk, int32_t *link) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  const int n = fluid_idx[k];
  const int yz = Y * Z;
  const int x = n / yz, r = n - x * yz, y = r / Z, z = r - y * Z;
  const Nbr nb = make_nbr(x, y, z, X, Y, Z);
#pragma unroll
  for (int i = 1; i < Q; ++i) link[(size_t)(i - 1) * lk + k] = map[nb.at(Lat<Q>::cx[i], Lat<Q>::cy[i], Lat<Q>::cz[i])];
}

// link[(i - 1) * lk + k] as a 16-bit step modulo nf (prt_lbm_core.h);
// *overflow = 1 if some link does not fit.
__global__ __launch_bounds__(256) void link_pack16_kernel(const int32_t *__restrict__ link, long long nf, long long lk,
                                                          int rows, int16_t *__restrict__ out, int32_t *overflow) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  bool bad = false;
  for (int r = 0; r < rows; ++r) {
    const int32_t v = link[(size_t)r * lk + k];
    int16_t e = (int16_t)kLink16Solid;
    if (v >= 0) {
      long long d = (long long)v - k;
      if (d > nf / 2) d -= nf;
      if (d < -(nf / 2)) d += nf;
      if (d < -32767 || d > 32767 || v >= nf) bad = true;
      else e = (int16_t)d;
    }
    out[(size_t)r * lk + k] = e;
  }
  if (bad) atomicOr(overflow, 1);
}

__global__ void flag_clear_kernel(int32_t *p) { *p = 0; }

template <int Q, int ST, bool RED, typename L>
__global__ __launch_bounds__(kSparseBlock) void lbm_sparse_kernel(LbmArgs a, const L *__restrict__ link,
                                                                 long long lk) {
  __shared__ double lds[16];
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int k = (int)blk * kSparseBlock + threadIdx.x;
  double contrib = 0.0;
  if (k < a.N) {
    double ux;
    if constexpr (ST == PRT_STORE_FP64) {
      double f[Q], fp[Q];
      gather_sparse<Q, ST>(a, link, lk, k, f);
      collide_ref<Q>(a, false, f, fp, ux);
      scatter<Q, ST>(a, (size_t)list_elem(k, 0, Q), kListTile, fp);
    } else {
      float f[Q], fp[Q];
      gather_sparse<Q, ST>(a, link, lk, k, f);
      collide_dev<Q>(a, false, f, fp, ux);
      scatter<Q, ST>(a, (size_t)list_elem(k, 0, Q), kListTile, fp);
    }
    if (RED) contrib = fabs(ux);
  }
  if (RED) {
    const double t = block_sum(contrib, lds);
    if (threadIdx.x == 0) a.partials[blk] = t;
  }
}

// Velocity of every fluid entry written at its box node (the caller zeroes
// the solid nodes), moments as the dense velocity kernel; partials[b] = the
// block's sum of u_x (FP64).
template <int Q, int ST, typename L>
__global__ __launch_bounds__(kSparseBlock) void sparse_velocity_kernel(LbmArgs a, const L *__restrict__ link,
                                                                      long long lk, const int32_t *fluid_idx,
                                                                      long long nbox, double *v64, float *v32) {
#pragma clang fp contract(off)
  __shared__ double lds[16];
  const int k = blockIdx.x * kSparseBlock + threadIdx.x;
  double contrib = 0.0;
  if (k < a.N) {
    double u[3];
    if constexpr (ST == PRT_STORE_FP64) {
      double f[Q];
      gather_sparse<Q, ST>(a, link, lk, k, f);
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
      gather_sparse<Q, ST>(a, link, lk, k, f);
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
    // a 2D grid carries its second velocity component in the z slot
    const double o0 = u[0], o1 = (Q == 9) ? u[2] : u[1], o2 = (Q == 9) ? 0.0 : u[2];
    const long long n = fluid_idx[k];
    if (v64) {
      v64[n] = o0;
      v64[nbox + n] = o1;
      v64[2 * nbox + n] = o2;
    }
    if (v32) {
      v32[n] = (float)o0;
      v32[nbox + n] = (float)o1;
      v32[2 * nbox + n] = (float)o2;
    }
    contrib = o0;
  }
  if (a.partials) {
    const double s = block_sum(contrib, lds);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
  }
}

// Rest state over a whole tiled list buffer of Q * ld elements.
template <int Q, int ST>
__global__ void sparse_fill_rest_kernel(void *f, long long ld) {
  using T = typename Pop<ST>::T;
  T *p = static_cast<T *>(f);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < ld * Q; k += stride) {
    const int i = (int)((k / kListTile) % Q);
    if constexpr (ST == PRT_STORE_FP64)
      p[k] = weight<Q>(i);
    else
      p[k] = (T)0.0f;
  }
}

static int sp_args(const char *who, int q, int storage, long long nf, const void *src, void *dst, double omp,
                   double omm, double force, double s16, LbmArgs &a) {
  if ((q != 9 && q != 19) || storage < 0 || storage > 2 || nf < 1 || pop_stride(nf) * q >= (1LL << 31) || !src || !dst ||
      src == dst || !(omp > 0 && omp < 2) || !(omm > 0 && omm < 2) ||
      (storage == PRT_STORE_FP16 && !(s16 > 0))) {
    set_error("%s: bad arguments", who);
    return PRT_E_ARG;
  }
  std::memset(&a, 0, sizeof(a));
  a.src = src;
  a.dst = dst;
  a.N = nf;
  a.X = a.Y = a.Z = 0;
  a.omp = omp;
  a.omm = omm;
  a.g = force;
  a.half_g = 0.5 * force;
  a.omh = 1 - 0.5 * omm;
  a.s16 = (float)(s16 > 0 ? s16 : 1.0);
  a.inv_s16 = (float)(1.0 / (s16 > 0 ? s16 : 1.0));
  return PRT_OK;
}

long long link_stride(long long nf) { return (nf + 63) / 64 * 64; }

}  // namespace prt

using namespace prt;

extern "C" int64_t prt_lbm_link_stride(int64_t n_fluid) { return link_stride(n_fluid); }

extern "C" int prt_lbm_link_table(int q, int nx, int ny, int nz, const int32_t *fluid_idx, int64_t n_fluid,
                                  int32_t *map_work, int32_t *link, void *stream) {
  if ((q != 9 && q != 19) || nx < 1 || ny < 1 || nz < 1 || (q == 9 && nz != 1) || !fluid_idx || !map_work ||
      !link || n_fluid < 1) {
    set_error("prt_lbm_link_table: bad arguments");
    return PRT_E_ARG;
  }
  const int X = nx, Y = q == 9 ? 1 : ny, Z = q == 9 ? ny : nz;
  const long long n = (long long)nx * ny * nz;
  if (n >= (1LL << 31) || n_fluid > n) {
    set_error("prt_lbm_link_table: grid or list too large");
    return PRT_E_RANGE;
  }
  hipStream_t st = (hipStream_t)stream;
  const long long lk = link_stride(n_fluid);
  map_clear_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(map_work, n);
  map_fill_kernel<<<(unsigned)((n_fluid + 255) / 256), 256, 0, st>>>(fluid_idx, n_fluid, n, map_work);
  const unsigned g = (unsigned)((n_fluid + 255) / 256);
  if (q == 19) link_build_kernel<19><<<g, 256, 0, st>>>(fluid_idx, n_fluid, X, Y, Z, map_work, lk, link);
  else link_build_kernel<9><<<g, 256, 0, st>>>(fluid_idx, n_fluid, X, Y, Z, map_work, lk, link);
  return launch_status("prt_lbm_link_table launch");
}

extern "C" int prt_lbm_sparse_fill_rest(int q, int storage, int64_t n_fluid, void *f, void *stream) {
  if (!f || n_fluid < 1 || (q != 9 && q != 19) || storage < 0 || storage > 2) {
    set_error("prt_lbm_sparse_fill_rest: bad arguments");
    return PRT_E_ARG;
  }
  const long long ld = pop_stride(n_fluid);
  hipStream_t st = (hipStream_t)stream;
  switch (q * 4 + storage) {
    case 19 * 4 + PRT_STORE_FP64: sparse_fill_rest_kernel<19, PRT_STORE_FP64><<<2048, 256, 0, st>>>(f, ld); break;
    case 19 * 4 + PRT_STORE_FP32: sparse_fill_rest_kernel<19, PRT_STORE_FP32><<<2048, 256, 0, st>>>(f, ld); break;
    case 19 * 4 + PRT_STORE_FP16: sparse_fill_rest_kernel<19, PRT_STORE_FP16><<<2048, 256, 0, st>>>(f, ld); break;
    case 9 * 4 + PRT_STORE_FP64: sparse_fill_rest_kernel<9, PRT_STORE_FP64><<<2048, 256, 0, st>>>(f, ld); break;
    case 9 * 4 + PRT_STORE_FP32: sparse_fill_rest_kernel<9, PRT_STORE_FP32><<<2048, 256, 0, st>>>(f, ld); break;
    default: sparse_fill_rest_kernel<9, PRT_STORE_FP16><<<2048, 256, 0, st>>>(f, ld); break;
  }
  return launch_status("prt_lbm_sparse_fill_rest launch");
}

extern "C" int prt_lbm_sparse_partials_count(int64_t n_fluid) {
  return (int)((n_fluid + kSparseBlock - 1) / kSparseBlock);
}

extern "C" int prt_lbm_link_pack16(int q, int64_t n_fluid, const int32_t *link, int16_t *link16, int32_t *overflow,
                                   void *stream) {
  if ((q != 9 && q != 19) || n_fluid < 1 || !link || !link16 || !overflow) {
    set_error("prt_lbm_link_pack16: bad arguments");
    return PRT_E_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  flag_clear_kernel<<<1, 1, 0, st>>>(overflow);
  link_pack16_kernel<<<(unsigned)((n_fluid + 255) / 256), 256, 0, st>>>(link, n_fluid, link_stride(n_fluid), q - 1,
                                                                        link16, overflow);
  return launch_status("prt_lbm_link_pack16 launch");
}

extern "C" int prt_lbm_sparse_run(int q, int storage, int64_t n_fluid, const void *link, int link_bits, void *f_a,
                                  void *f_b, int steps, double omega_p, double omega_m, double force,
                                  double fp16_scale, double *partials, void *stream) {
  LbmArgs a;
  int rc = sp_args("prt_lbm_sparse_run", q, storage, n_fluid, f_a, f_b, omega_p, omega_m, force, fp16_scale, a);
  if (rc) return rc;
  if (!link || steps < 0 || (link_bits != 32 && link_bits != 16)) {
    set_error("prt_lbm_sparse_run: null link table, link width not 16 or 32, or negative steps");
    return PRT_E_ARG;
  }
  const int32_t *l32 = static_cast<const int32_t *>(link);
  const int16_t *l16 = static_cast<const int16_t *>(link);
  const long long lk = link_stride(n_fluid);
  const dim3 grid((unsigned)((n_fluid + kSparseBlock - 1) / kSparseBlock)), block(kSparseBlock);
  hipStream_t st = (hipStream_t)stream;
  for (int s = 0; s < steps; ++s) {
    a.src = (s & 1) ? f_b : f_a;
    a.dst = (s & 1) ? f_a : f_b;
    const bool red = partials && s == steps - 1;
    a.partials = red ? partials : nullptr;
#define PRT_SK(QQ, SS)                                                                            \
  do {                                                                                            \
    if (link_bits == 16) {                                                                        \
      if (red) lbm_sparse_kernel<QQ, SS, true><<<grid, block, 0, st>>>(a, l16, lk);               \
      else lbm_sparse_kernel<QQ, SS, false><<<grid, block, 0, st>>>(a, l16, lk);                  \
    } else {                                                                                      \
      if (red) lbm_sparse_kernel<QQ, SS, true><<<grid, block, 0, st>>>(a, l32, lk);               \
      else lbm_sparse_kernel<QQ, SS, false><<<grid, block, 0, st>>>(a, l32, lk);                  \
    }                                                                                             \
  } while (0)
    switch (q * 4 + storage) {
      case 19 * 4 + PRT_STORE_FP64: PRT_SK(19, PRT_STORE_FP64); break;
      case 19 * 4 + PRT_STORE_FP32: PRT_SK(19, PRT_STORE_FP32); break;
      case 19 * 4 + PRT_STORE_FP16: PRT_SK(19, PRT_STORE_FP16); break;
      case 9 * 4 + PRT_STORE_FP64: PRT_SK(9, PRT_STORE_FP64); break;
      case 9 * 4 + PRT_STORE_FP32: PRT_SK(9, PRT_STORE_FP32); break;
      default: PRT_SK(9, PRT_STORE_FP16); break;
    }
#undef PRT_SK
    rc = launch_status("prt_lbm_sparse_run launch");
    if (rc) return rc;
  }
  return PRT_OK;
}

extern "C" int prt_lbm_sparse_velocity(int q, int storage, int64_t n_box, int64_t n_fluid, const int32_t *fluid_idx,
                                       const void *link, int link_bits, const void *f_last, const void *f_prev,
                                       double force, double fp16_scale, double *v64, float *v32, double *partials,
                                       void *stream) {
  LbmArgs a;
  int rc = sp_args("prt_lbm_sparse_velocity", q, storage, n_fluid, f_last, const_cast<void *>(f_prev), 1.0, 1.0,
                   force, fp16_scale, a);
  if (rc) return rc;
  if (!fluid_idx || !link || n_box < n_fluid || (!v64 && !v32) || (link_bits != 32 && link_bits != 16)) {
    set_error("prt_lbm_sparse_velocity: bad arguments");
    return PRT_E_ARG;
  }
  const int32_t *l32 = static_cast<const int32_t *>(link);
  const int16_t *l16 = static_cast<const int16_t *>(link);
  a.partials = partials;
  const long long lk = link_stride(n_fluid);
  const dim3 grid((unsigned)((n_fluid + kSparseBlock - 1) / kSparseBlock)), block(kSparseBlock);
  hipStream_t st = (hipStream_t)stream;
#define PRT_SV(QQ, SS)                                                                                      \
  do {                                                                                                      \
    if (link_bits == 16)                                                                                    \
      sparse_velocity_kernel<QQ, SS><<<grid, block, 0, st>>>(a, l16, lk, fluid_idx, n_box, v64, v32);       \
    else                                                                                                    \
      sparse_velocity_kernel<QQ, SS><<<grid, block, 0, st>>>(a, l32, lk, fluid_idx, n_box, v64, v32);       \
  } while (0)
  switch (q * 4 + storage) {
    case 19 * 4 + PRT_STORE_FP64: PRT_SV(19, PRT_STORE_FP64); break;
    case 19 * 4 + PRT_STORE_FP32: PRT_SV(19, PRT_STORE_FP32); break;
    case 19 * 4 + PRT_STORE_FP16: PRT_SV(19, PRT_STORE_FP16); break;
    case 9 * 4 + PRT_STORE_FP64: PRT_SV(9, PRT_STORE_FP64); break;
    case 9 * 4 + PRT_STORE_FP32: PRT_SV(9, PRT_STORE_FP32); break;
    default: PRT_SV(9, PRT_STORE_FP16); break;
  }
#undef PRT_SV
  return launch_status("prt_lbm_sparse_velocity launch");
}
