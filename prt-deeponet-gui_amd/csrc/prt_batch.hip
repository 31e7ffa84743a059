// prt_batch.hip -- many geometries of uneven size in one launch.
//
// A batch is a set of segments, one per geometry, each with its own dense
// box, solid mask, A/B population (or concentration) buffers and physical
// parameters.  One launch covers every active segment: a block-to-segment
// table maps each 256-thread block to its segment and to 256 consecutive
// nodes of that segment's box, so blocks never straddle two geometries and
// no lane idles except in each segment's last block.  The per-node update is
// the very same device code as the single-geometry kernels (prt_lbm_core.h,
// prt_adr_core.h): a geometry gives bit-identical results alone, in a batch,
// or in any mix -- which is what the serving loop relies on.
//
// Segments enter and leave between launches (continuous batching): each
// segment carries a phase bit saying which of its buffers holds the newest
// state at launch parity 0, so a segment admitted at any step joins the
// alternation without copying.  Per-segment reductions (mean |u_x| for the
// flow convergence, pore sums of each chemical) are per-block partials
// reduced per segment in a fixed order.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>

#include "prt_adr_core.h"
#include "prt_lbm_core.h"

namespace prt {

struct SegLbm {
  LbmArgs a;        // src/dst chosen per launch from fa/fb and the phase
  void *fa, *fb;
  const void *link;     // fluid-list member: link table (a.N = list length)
  long long lk;
  int phase, blk0, nblk, link16;  // link16: 16-bit link table
};

struct SegAdr {
  AdrArgs<float> a;
  float *ca, *cb;
  const void *tab;       // pore table of the segment's list
  uint32_t ta;           // its entries per section (tab_entries)
  int npts;              // list entries
  int phase, blk0, nblk, nd;
};

// Linear node index -> (x, y, z) of a (X, Y, Z) box.
__device__ inline void unflatten(int n, int Y, int Z, int &x, int &y, int &z) {
  const int yz = Y * Z;
  x = n / yz;
  const int r = n - x * yz;
  y = r / Z;
  z = r - y * Z;
}

template <int Q, int ST, bool SKIP, bool RED>
__global__ __launch_bounds__(256) void lbm_batch_kernel(const SegLbm *__restrict__ segs,
                                                        const int *__restrict__ blk_seg, int parity,
                                                        double *partials) {
  __shared__ double lds[16];
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int s = __builtin_amdgcn_readfirstlane(blk_seg[blk]);
  const SegLbm &d = segs[s];
  LbmArgs a = d.a;
  const bool newest_a = ((parity ^ d.phase) & 1) == 0;
  a.src = newest_a ? d.fa : d.fb;
  a.dst = newest_a ? d.fb : d.fa;
  const int n = (int)(blk - d.blk0) * 256 + threadIdx.x;
  double contrib = 0.0;
  if (n < a.N) {
    int x, y, z;
    unflatten(n, a.Y, a.Z, x, y, z);
    const bool solid = a.solid[n] != 0;
    if (!(SKIP && solid)) {
      const Nbr nb = make_nbr(x, y, z, a.X, a.Y, a.Z);
      const size_t self = (size_t)nb.own(Q, (int)a.zs);
      double ux;
      if constexpr (ST == PRT_STORE_FP64) {
        double f[Q], fp[Q];
        gather<Q, ST, SKIP>(a, nb, f);
        collide_ref<Q>(a, solid, f, fp, ux);
        scatter<Q, ST>(a, self, a.zs, fp);
      } else {
        float f[Q], fp[Q];
        gather<Q, ST, SKIP>(a, nb, f);
        collide_dev<Q>(a, solid, f, fp, ux);
        scatter<Q, ST>(a, self, a.zs, fp);
      }
      if (RED && !solid) contrib = fabs(ux);
    }
  }
  if (RED) {
    const double t = block_sum(contrib, lds);
    if (threadIdx.x == 0) partials[blk] = t;
  }
}

// Fluid-list members: entry k of member s's list, the single-geometry
// sparse step's code (gather_sparse, the same collision and scatter).
template <int Q, int ST, bool RED>
__global__ __launch_bounds__(256) void lbm_batch_sparse_kernel(const SegLbm *__restrict__ segs,
                                                               const int *__restrict__ blk_seg, int parity,
                                                               double *partials) {
  __shared__ double lds[16];
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int s = __builtin_amdgcn_readfirstlane(blk_seg[blk]);
  const SegLbm &d = segs[s];
  LbmArgs a = d.a;
  const bool newest_a = ((parity ^ d.phase) & 1) == 0;
  a.src = newest_a ? d.fa : d.fb;
  a.dst = newest_a ? d.fb : d.fa;
  const int k = (int)(blk - d.blk0) * 256 + threadIdx.x;
  double contrib = 0.0;
  if (k < a.N) {
    double ux;
    using R = typename std::conditional<ST == PRT_STORE_FP64, double, float>::type;
    R f[Q], fp[Q];
    if (d.link16) gather_sparse<Q, ST>(a, static_cast<const int16_t *>(d.link), d.lk, k, f);
    else gather_sparse<Q, ST>(a, static_cast<const int32_t *>(d.link), d.lk, k, f);
    if constexpr (ST == PRT_STORE_FP64) collide_ref<Q>(a, false, f, fp, ux);
    else collide_dev<Q>(a, false, f, fp, ux);
    scatter<Q, ST>(a, (size_t)list_elem(k, 0, Q), kListTile, fp);
    if (RED) contrib = fabs(ux);
  }
  if (RED) {
    const double t = block_sum(contrib, lds);
    if (threadIdx.x == 0) partials[blk] = t;
  }
}

// Transport members: block b of member s covers 256 consecutive entries of
// its pore list.  With one or two chemicals the kernel is held to the
// registers of eight waves per SIMD (left alone the compiler takes 68 vector
// registers, seven waves; held, 60 and some scalars spilled to lanes): the
// 64-rock campaign step measured 558 -> 551 us.
template <int NSM, bool RED, bool NTT>
__global__ __launch_bounds__(256, NSM <= 2 ? 8 : 1) void adr_batch_kernel(const SegAdr *__restrict__ segs,
                                                                         const int *__restrict__ blk_seg,
                                                                         int parity, double *partials) {
  __shared__ double lds[16];
  // round-robin block order here: measured faster than the XCD-contiguous
  // order on the 64-rock campaign mix
  const uint32_t blk = blockIdx.x;
  const int s = __builtin_amdgcn_readfirstlane(blk_seg[blk]);
  const SegAdr &d = segs[s];
  AdrArgs<float> a = d.a;
  const bool newest_a = ((parity ^ d.phase) & 1) == 0;
  a.cin = newest_a ? d.ca : d.cb;
  a.cout = newest_a ? d.cb : d.ca;
  const uint32_t k = (blk - (uint32_t)d.blk0) * 256 + threadIdx.x;
  float cnew[NSM];
  bool pore = false;
  if (k < (uint32_t)d.npts)
    pore = d.nd == 3 ? adr_entry<float, 3, NSM, NTT>(a, d.tab, d.ta, k, cnew)
                     : adr_entry<float, 2, NSM, NTT>(a, d.tab, d.ta, k, cnew);
  if (RED) {
    // per-block pore sums of every chemical (FP64), NSM slots per block
#pragma unroll
    for (int j = 0; j < NSM; ++j) {
      const double t = block_sum((pore && j < a.ns) ? (double)cnew[j] : 0.0, lds);
      if (threadIdx.x == 0) partials[(size_t)blk * NSM + j] = t;
      __syncthreads();
    }
  }
}

// out[s * width + j] = sum over segment s's blocks of partials[b * width + j].
__global__ __launch_bounds__(256) void seg_sum_kernel(const int *blk0, const int *nblk, int width,
                                                      const double *partials, double *out) {
  __shared__ double part[256];
  const int s = blockIdx.x;
  for (int j = 0; j < width; ++j) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk[s]; b += 256) acc += partials[(size_t)(blk0[s] + b) * width + j];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)s * width + j] = part[0];
    __syncthreads();
  }
}

static int blocks_of(long long n) { return (int)((n + 255) / 256); }

}  // namespace prt

using namespace prt;

extern "C" int64_t prt_lbm_batch_desc_size(void) { return (int64_t)sizeof(SegLbm); }
extern "C" int64_t prt_adr_batch_desc_size(void) { return (int64_t)sizeof(SegAdr); }

extern "C" int prt_lbm_batch_plan(int q, int storage, const prt_lbm_segment *segs, int nseg, void *desc_host,
                                  int32_t *blk_seg_host, int64_t blk_capacity, int64_t *n_blocks) {
  if (!segs || nseg < 1 || !desc_host || !n_blocks || (q != 9 && q != 19) || storage < 0 || storage > 2) {
    set_error("prt_lbm_batch_plan: bad arguments");
    return PRT_E_ARG;
  }
  SegLbm *d = static_cast<SegLbm *>(desc_host);
  long long total = 0;
  for (int s = 0; s < nseg; ++s) {
    const prt_lbm_segment &g = segs[s];
    const bool sparse = g.link != nullptr;
    if (g.nx < 1 || g.ny < 1 || g.nz < 1 || (q == 9 && g.nz != 1) || (!sparse && !g.solid) || !g.f_a ||
        !g.f_b || g.f_a == g.f_b || (sparse && (g.n_fluid < 1 || g.n_fluid > (long long)g.nx * g.ny * g.nz)) ||
        (s > 0 && sparse != (segs[0].link != nullptr)) ||
        (sparse && g.link_bits != 0 && g.link_bits != 16 && g.link_bits != 32)) {
      set_error("prt_lbm_batch_plan: segment %d has a bad grid or buffers", s);
      return PRT_E_ARG;
    }
    const double omp = g.omega_p, omm = g.omega_m;
    if (!(omp > 0 && omp < 2) || !(omm > 0 && omm < 2)) {
      set_error("prt_lbm_batch_plan: segment %d relaxation rates out of (0, 2)", s);
      return PRT_E_ARG;
    }
    std::memset(&d[s], 0, sizeof(SegLbm));
    LbmArgs &a = d[s].a;
    a.solid = g.solid;
    if (q == 9) {
      a.X = g.nx; a.Y = 1; a.Z = g.ny;
    } else {
      a.X = g.nx; a.Y = g.ny; a.Z = g.nz;
    }
    a.N = sparse ? (long long)g.n_fluid : (long long)g.nx * g.ny * g.nz;
    const int eb = storage == PRT_STORE_FP64 ? 8 : (storage == PRT_STORE_FP32 ? 4 : 2);
    a.zs = sparse ? 0 : pop_row_pitch(a.Z, eb);
    if ((sparse ? pop_stride(a.N) * q : pop_elems(q, a.X, a.Y, a.Z, eb)) >= (1LL << 31)) {
      set_error("prt_lbm_batch_plan: segment %d too large for the 32-bit index", s);
      return PRT_E_RANGE;
    }
    a.omp = omp; a.omm = omm; a.g = g.force; a.half_g = 0.5 * g.force;
    a.omh = 1 - 0.5 * omm;
    const double s16 = g.fp16_scale > 0 ? g.fp16_scale : 1.0;
    a.s16 = (float)s16;
    a.inv_s16 = (float)(1.0 / s16);
    d[s].fa = g.f_a;
    d[s].fb = g.f_b;
    d[s].link = g.link;
    d[s].link16 = sparse && g.link_bits == 16;
    d[s].lk = sparse ? link_stride(a.N) : 0;
    d[s].phase = g.phase & 1;
    d[s].blk0 = (int)total;
    d[s].nblk = blocks_of(a.N);
    total += d[s].nblk;
  }
  *n_blocks = total;
  if (!blk_seg_host) return PRT_OK;  // size query
  if (blk_capacity < total) {
    set_error("prt_lbm_batch_plan: block table holds %lld entries, %lld needed", (long long)blk_capacity, total);
    return PRT_E_ARG;
  }
  for (int s = 0; s < nseg; ++s)
    for (int b = 0; b < d[s].nblk; ++b) blk_seg_host[d[s].blk0 + b] = s;
  return PRT_OK;
}

extern "C" int prt_lbm_batch_step(int q, int storage, const void *desc_dev, const int32_t *blk_seg_dev,
                                  int64_t n_blocks, int parity, int flags, double *partials, void *stream) {
  if (!desc_dev || !blk_seg_dev || n_blocks < 1 || n_blocks > 0x7fffffffLL) {
    set_error("prt_lbm_batch_step: bad arguments");
    return PRT_E_ARG;
  }
  const SegLbm *d = static_cast<const SegLbm *>(desc_dev);
  hipStream_t st = (hipStream_t)stream;
  const bool skip = (flags & PRT_LBM_SKIP_SOLID) != 0, red = partials != nullptr;
  const bool sparse = (flags & PRT_LBM_SPARSE) != 0;
  const dim3 grid((unsigned)n_blocks), block(256);
#define PRT_LB(QQ, SS)                                                                       \
  do {                                                                                       \
    if (sparse && red) lbm_batch_sparse_kernel<QQ, SS, true><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials); \
    else if (sparse) lbm_batch_sparse_kernel<QQ, SS, false><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials); \
    else                                                                                     \
    if (skip && red) lbm_batch_kernel<QQ, SS, true, true><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials);  \
    else if (skip) lbm_batch_kernel<QQ, SS, true, false><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials);   \
    else if (red) lbm_batch_kernel<QQ, SS, false, true><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials);    \
    else lbm_batch_kernel<QQ, SS, false, false><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials);            \
  } while (0)
  switch (q * 4 + storage) {
    case 19 * 4 + PRT_STORE_FP64: PRT_LB(19, PRT_STORE_FP64); break;
    case 19 * 4 + PRT_STORE_FP32: PRT_LB(19, PRT_STORE_FP32); break;
    case 19 * 4 + PRT_STORE_FP16: PRT_LB(19, PRT_STORE_FP16); break;
    case 9 * 4 + PRT_STORE_FP64: PRT_LB(9, PRT_STORE_FP64); break;
    case 9 * 4 + PRT_STORE_FP32: PRT_LB(9, PRT_STORE_FP32); break;
    case 9 * 4 + PRT_STORE_FP16: PRT_LB(9, PRT_STORE_FP16); break;
    default:
      set_error("prt_lbm_batch_step: unknown q/storage %d/%d", q, storage);
      return PRT_E_ARG;
  }
#undef PRT_LB
  return launch_status("prt_lbm_batch_step launch");
}

extern "C" int prt_adr_batch_plan(const prt_adr_segment *segs, int nseg, void *desc_host, int32_t *blk_seg_host,
                                  int64_t blk_capacity, int64_t *n_blocks, int *max_species) {
  if (!segs || nseg < 1 || !desc_host || !n_blocks) {
    set_error("prt_adr_batch_plan: bad arguments");
    return PRT_E_ARG;
  }
  SegAdr *d = static_cast<SegAdr *>(desc_host);
  long long total = 0;
  int nsm = 1;
  for (int s = 0; s < nseg; ++s) {
    const prt_adr_segment &g = segs[s];
    if (!g.params || !g.c_a || !g.c_b || g.c_a == g.c_b || (g.n_pore > 0 && !g.pore_table)) {
      set_error("prt_adr_batch_plan: segment %d has null or aliased buffers", s);
      return PRT_E_ARG;
    }
    std::memset(&d[s], 0, sizeof(SegAdr));
    dim3 gr, bl;
    const int rc = adr_args_from_params<float>(g.params, g.nx, g.ny, g.nz, d[s].a, gr, bl);
    if (rc) return rc;
    if (g.n_pore < 0 || g.n_pore > d[s].a.N || (unsigned long long)tab_entries(g.n_pore) * 16ull >= (1ull << 32) ||
        g.params->write_solid) {
      set_error("prt_adr_batch_plan: segment %d: pore count out of range or write_solid set", s);
      return PRT_E_ARG;
    }
    d[s].ca = g.c_a;
    d[s].cb = g.c_b;
    d[s].tab = g.pore_table;
    d[s].ta = (uint32_t)tab_entries(g.n_pore);
    d[s].npts = (int)g.n_pore;
    d[s].phase = g.phase & 1;
    d[s].nd = g.params->nd;
    d[s].blk0 = (int)total;
    d[s].nblk = d[s].npts > 0 ? blocks_of(d[s].npts) : 0;
    total += d[s].nblk;
    nsm = g.params->ns > nsm ? g.params->ns : nsm;
  }
  *n_blocks = total;
  if (max_species) *max_species = nsm;
  if (!blk_seg_host) return PRT_OK;
  if (blk_capacity < total) {
    set_error("prt_adr_batch_plan: block table holds %lld entries, %lld needed", (long long)blk_capacity, total);
    return PRT_E_ARG;
  }
  for (int s = 0; s < nseg; ++s)
    for (int b = 0; b < d[s].nblk; ++b) blk_seg_host[d[s].blk0 + b] = s;
  return PRT_OK;
}

extern "C" int prt_adr_batch_step(const void *desc_dev, const int32_t *blk_seg_dev, int64_t n_blocks,
                                  int max_species, int parity, double *partials, void *stream) {
  if (!desc_dev || !blk_seg_dev || n_blocks < 1 || n_blocks > 0x7fffffffLL || max_species < 1 ||
      max_species > 8) {
    set_error("prt_adr_batch_step: bad arguments");
    return PRT_E_ARG;
  }
  const SegAdr *d = static_cast<const SegAdr *>(desc_dev);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)n_blocks), block(256);
  const bool red = partials != nullptr;
  // the members' tables together decide the load kind, as one list's does
  const bool ntt = n_blocks * 256 >= (int64_t)kNtTableEntries;
#define PRT_AB3(NN, R)                                                                       \
  do {                                                                                       \
    if (ntt) adr_batch_kernel<NN, R, true><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials); \
    else adr_batch_kernel<NN, R, false><<<grid, block, 0, st>>>(d, blk_seg_dev, parity, partials);    \
  } while (0)
#define PRT_AB(NN)          \
  do {                      \
    if (red) PRT_AB3(NN, true);  \
    else PRT_AB3(NN, false);     \
  } while (0)
  if (max_species == 1) PRT_AB(1);
  else if (max_species == 2) PRT_AB(2);
  else if (max_species <= 4) PRT_AB(4);
  else PRT_AB(8);
#undef PRT_AB
#undef PRT_AB3
  return launch_status("prt_adr_batch_step launch");
}

extern "C" int prt_seg_sums(const int32_t *blk0_dev, const int32_t *nblk_dev, int nseg, int width,
                            const double *partials, double *out, void *stream) {
  if (!blk0_dev || !nblk_dev || nseg < 1 || width < 1 || !partials || !out) {
    set_error("prt_seg_sums: bad arguments");
    return PRT_E_ARG;
  }
  seg_sum_kernel<<<nseg, 256, 0, (hipStream_t)stream>>>(blk0_dev, nblk_dev, width, partials, out);
  return launch_status("prt_seg_sums launch");
}
