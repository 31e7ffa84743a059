// prt_batch.hip -- many geometries of uneven size in one launch: the
// batched Stokes step (the batched transport step is prt_batch_adr.hip).
//
// A batch is a set of segments, one per geometry, each with its own box or
// fluid list, A/B population buffers and physical parameters.  One launch
// covers every active segment: a block-to-segment table maps each block to
// its segment and to one tile of consecutive nodes (256 or 512, the launch
// shape's) of that segment, so blocks never straddle two geometries and no
// lane idles except in each segment's last block.  The per-node update is
// the very same device code as the single-geometry kernels (prt_lbm_core.h,
// prt_adr_core.h): a geometry gives bit-identical results alone, in a batch,
// in any mix and under any launch shape -- which is what the serving loop
// and the launch-shape tuner rely on.
//
// Segments enter and leave between launches (continuous batching): each
// segment carries a phase bit saying which of its buffers holds the newest
// state at launch parity 0, so a segment admitted at any step joins the
// alternation without copying.  Per-segment reductions (mean |u_x| for the
// flow convergence, pore sums of each chemical) are partials per 256-node
// tile reduced per segment in a fixed order.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>

#include "prt_batch.h"

namespace prt {

// Linear node index -> (x, y, z) of a (X, Y, Z) box.
__device__ inline void unflatten(int n, int Y, int Z, int &x, int &y, int &z) {
  const int yz = Y * Z;
  x = n / yz;
  const int r = n - x * yz;
  y = r / Z;
  z = r - y * Z;
}

// Dense members (every box node, or the fluid ones with SKIP): one node
// per thread, 256-node tiles.
template <int Q, int ST, bool SKIP, bool RED>
__global__ __launch_bounds__(256) void lbm_batch_kernel(const SegLbm *__restrict__ segs,
                                                        const int *__restrict__ blk_seg, int parity, int xcd,
                                                        double *partials) {
  __shared__ double lds[4];
  const uint32_t blk = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int s = __builtin_amdgcn_readfirstlane(blk_seg[blk]);
  const SegLbm &d = segs[s];
  if (d.tile != 256) return;  // planned for another shape: update nothing
  LbmArgs a = d.a;
  const bool newest_a = ((parity ^ d.phase) & 1) == 0;
  a.src = newest_a ? d.fa : d.fb;
  a.dst = newest_a ? d.fb : d.fa;
  const int base = (int)(blk - d.blk0) * 256;
  const int n = base + threadIdx.x;
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
    group_to_lds(contrib, lds, threadIdx.x >> 6);
    tile_partials<256>(lds, 1, base, a.N, partials + d.rt0 + (blk - d.blk0));
  }
}

// Fluid-list members: entries base + e*T + thread (e < E) of member s's
// list, each with the single-geometry sparse step's code (gather_sparse,
// the same collision and scatter).
template <int Q, int ST, bool RED, int T, int E>
__global__ __launch_bounds__(T) void lbm_batch_sparse_kernel(const SegLbm *__restrict__ segs,
                                                             const int *__restrict__ blk_seg, int parity,
                                                             int xcd, double *partials) {
  constexpr int TILE = T * E;
  __shared__ double lds[TILE / 64];
  const uint32_t blk = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int s = __builtin_amdgcn_readfirstlane(blk_seg[blk]);
  const SegLbm &d = segs[s];
  if (d.tile != TILE) return;  // planned for another shape: update nothing
  LbmArgs a = d.a;
  const bool newest_a = ((parity ^ d.phase) & 1) == 0;
  a.src = newest_a ? d.fa : d.fb;
  a.dst = newest_a ? d.fb : d.fa;
  const int base = (int)(blk - d.blk0) * TILE;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = base + e * T + (int)threadIdx.x;
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
    if (RED) group_to_lds(contrib, lds, (e * T + (int)threadIdx.x) >> 6);
  }
  if (RED) tile_partials<TILE>(lds, 1, base, a.N, partials + d.rt0 + (blk - d.blk0) * (TILE / 256));
}

// out[s * width + j] = sum over segment s's tiles of partials[b * width + j].
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

template <int Q, int ST, bool RED>
static void launch_sparse(const BatchShape &sh, dim3 grid, hipStream_t st, const SegLbm *d, const int32_t *bs,
                          int parity, double *partials) {
#define PRT_LS(TT, EE)                                                                             \
  if (sh.t == TT && sh.e == EE) {                                                                  \
    lbm_batch_sparse_kernel<Q, ST, RED, TT, EE><<<grid, dim3(TT), 0, st>>>(d, bs, parity, sh.xcd, partials); \
    return;                                                                                        \
  }
  PRT_LS(256, 1) PRT_LS(128, 2) PRT_LS(64, 4) PRT_LS(256, 2) PRT_LS(128, 4)
#undef PRT_LS
}

template <int Q, int ST>
static void launch_lbm(bool sparse, bool skip, bool red, const BatchShape &sh, dim3 grid, hipStream_t st,
                       const SegLbm *d, const int32_t *bs, int parity, double *partials) {
  if (sparse) {
    if (red) launch_sparse<Q, ST, true>(sh, grid, st, d, bs, parity, partials);
    else launch_sparse<Q, ST, false>(sh, grid, st, d, bs, parity, partials);
    return;
  }
  const dim3 block(256);
  if (skip && red) lbm_batch_kernel<Q, ST, true, true><<<grid, block, 0, st>>>(d, bs, parity, sh.xcd, partials);
  else if (skip) lbm_batch_kernel<Q, ST, true, false><<<grid, block, 0, st>>>(d, bs, parity, sh.xcd, partials);
  else if (red) lbm_batch_kernel<Q, ST, false, true><<<grid, block, 0, st>>>(d, bs, parity, sh.xcd, partials);
  else lbm_batch_kernel<Q, ST, false, false><<<grid, block, 0, st>>>(d, bs, parity, sh.xcd, partials);
}

}  // namespace prt

using namespace prt;

extern "C" int64_t prt_lbm_batch_desc_size(void) { return (int64_t)sizeof(SegLbm); }

extern "C" int prt_batch_shape_tile(int shape) {
  BatchShape s;
  return decode_shape(shape, s) ? s.tile() : 0;
}

extern "C" int prt_lbm_batch_plan_shaped(int q, int storage, const prt_lbm_segment *segs, int nseg, int shape,
                                         void *desc_host, int32_t *blk_seg_host, int64_t blk_capacity,
                                         int64_t *n_blocks, int64_t *n_tiles) {
  BatchShape sh;
  if (!segs || nseg < 1 || !desc_host || !n_blocks || (q != 9 && q != 19) || storage < 0 || storage > 2 ||
      !decode_shape(shape, sh)) {
    set_error("prt_lbm_batch_plan: bad arguments");
    return PRT_E_ARG;
  }
  const int tile = sh.tile();
  SegLbm *d = static_cast<SegLbm *>(desc_host);
  long long total = 0, tiles = 0;
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
    if (!sparse && (sh.t != 256 || sh.e != 1)) {
      set_error("prt_lbm_batch_plan: dense members take 256 threads with one node each");
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
    // entries index up to a whole tile past the list (k < N guards them)
    if ((sparse ? pop_stride(a.N) * q : pop_elems(q, a.X, a.Y, a.Z, eb)) >= (1LL << 31) ||
        a.N + tile >= (1LL << 31)) {
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
    d[s].tile = tile;
    d[s].blk0 = (int)total;
    d[s].nblk = blocks_of(a.N, tile);
    d[s].rt0 = (int)tiles;
    total += d[s].nblk;
    tiles += blocks_of(a.N, 256);
    if (total > 0x7fffffffLL || tiles > 0x7fffffffLL) {
      set_error("prt_lbm_batch_plan: more than 2^31 blocks");
      return PRT_E_RANGE;
    }
  }
  *n_blocks = total;
  if (n_tiles) *n_tiles = tiles;
  if (!blk_seg_host) return PRT_OK;  // size query
  if (blk_capacity < total) {
    set_error("prt_lbm_batch_plan: block table holds %lld entries, %lld needed", (long long)blk_capacity, total);
    return PRT_E_ARG;
  }
  for (int s = 0; s < nseg; ++s)
    for (int b = 0; b < d[s].nblk; ++b) blk_seg_host[d[s].blk0 + b] = s;
  return PRT_OK;
}

extern "C" int prt_lbm_batch_plan(int q, int storage, const prt_lbm_segment *segs, int nseg, void *desc_host,
                                  int32_t *blk_seg_host, int64_t blk_capacity, int64_t *n_blocks) {
  return prt_lbm_batch_plan_shaped(q, storage, segs, nseg, PRT_BATCH_SHAPE_LBM_DEFAULT, desc_host, blk_seg_host,
                                   blk_capacity, n_blocks, nullptr);
}

extern "C" int prt_lbm_batch_step_shaped(int q, int storage, const void *desc_dev, const int32_t *blk_seg_dev,
                                         int64_t n_blocks, int parity, int flags, int shape, double *partials,
                                         void *stream) {
  BatchShape sh;
  if (!desc_dev || !blk_seg_dev || n_blocks < 1 || n_blocks > 0x7fffffffLL || !decode_shape(shape, sh)) {
    set_error("prt_lbm_batch_step: bad arguments");
    return PRT_E_ARG;
  }
  const SegLbm *d = static_cast<const SegLbm *>(desc_dev);
  hipStream_t st = (hipStream_t)stream;
  const bool skip = (flags & PRT_LBM_SKIP_SOLID) != 0, red = partials != nullptr;
  const bool sparse = (flags & PRT_LBM_SPARSE) != 0;
  if (!sparse && (sh.t != 256 || sh.e != 1)) {
    set_error("prt_lbm_batch_step: dense members take 256 threads with one node each");
    return PRT_E_ARG;
  }
  const dim3 grid((unsigned)n_blocks);
  switch (q * 4 + storage) {
    case 19 * 4 + PRT_STORE_FP64: launch_lbm<19, PRT_STORE_FP64>(sparse, skip, red, sh, grid, st, d, blk_seg_dev, parity, partials); break;
    case 19 * 4 + PRT_STORE_FP32: launch_lbm<19, PRT_STORE_FP32>(sparse, skip, red, sh, grid, st, d, blk_seg_dev, parity, partials); break;
    case 19 * 4 + PRT_STORE_FP16: launch_lbm<19, PRT_STORE_FP16>(sparse, skip, red, sh, grid, st, d, blk_seg_dev, parity, partials); break;
    case 9 * 4 + PRT_STORE_FP64: launch_lbm<9, PRT_STORE_FP64>(sparse, skip, red, sh, grid, st, d, blk_seg_dev, parity, partials); break;
    case 9 * 4 + PRT_STORE_FP32: launch_lbm<9, PRT_STORE_FP32>(sparse, skip, red, sh, grid, st, d, blk_seg_dev, parity, partials); break;
    case 9 * 4 + PRT_STORE_FP16: launch_lbm<9, PRT_STORE_FP16>(sparse, skip, red, sh, grid, st, d, blk_seg_dev, parity, partials); break;
    default:
      set_error("prt_lbm_batch_step: unknown q/storage %d/%d", q, storage);
      return PRT_E_ARG;
  }
  return launch_status("prt_lbm_batch_step launch");
}

extern "C" int prt_lbm_batch_step(int q, int storage, const void *desc_dev, const int32_t *blk_seg_dev,
                                  int64_t n_blocks, int parity, int flags, double *partials, void *stream) {
  return prt_lbm_batch_step_shaped(q, storage, desc_dev, blk_seg_dev, n_blocks, parity, flags,
                                   PRT_BATCH_SHAPE_LBM_DEFAULT, partials, stream);
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
