// prt_batch_adr.hip -- the batched transport step: every member's pore
// list advanced by one launch (the batching scheme is described in
// prt_batch.hip).  The per-entry update is adr_entry (prt_adr_core.h), the
// single-geometry pore-list kernel's own code, so a member's field is the
// same bits alone, in any batch and under any launch shape.
#include <hip/hip_runtime.h>

#include <cstring>

#include "prt_batch.h"

namespace prt {

// Registers: with one or two chemicals and one entry per thread the kernel
// is held to eight waves per SIMD (left alone the compiler takes 68 vector
// registers, seven waves; held, 60 and some scalars spilled to lanes): the
// 64-rock campaign step measured 558 -> 551 us.  Shapes with several
// entries per thread are left to the compiler.
template <int NSM, int T, int E>
struct AdrOcc {
  static constexpr int min_blocks = (E == 1 && NSM <= 2) ? 2048 / T : 1;
};

// Block b of member s covers entries base + e*T + thread (e < E) of its
// pore list, base = (b - blk0) * T * E.
template <int NSM, bool RED, bool NTT, int T, int E>
__global__ __launch_bounds__(T, (AdrOcc<NSM, T, E>::min_blocks)) void adr_batch_kernel(
    const SegAdr *__restrict__ segs, const int *__restrict__ blk_seg, int parity, int xcd, double *partials) {
  constexpr int TILE = T * E;
  __shared__ double lds[RED ? (TILE / 64) * NSM : 1];
  const uint32_t blk = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int s = __builtin_amdgcn_readfirstlane(blk_seg[blk]);
  const SegAdr &d = segs[s];
  if (d.tile != TILE) return;  // planned for another shape: update nothing
  AdrArgs<float> a = d.a;
  const bool newest_a = ((parity ^ d.phase) & 1) == 0;
  a.cin = newest_a ? d.ca : d.cb;
  a.cout = newest_a ? d.cb : d.ca;
  const uint32_t base = (blk - (uint32_t)d.blk0) * TILE;
  if constexpr (E == 1) {
    const uint32_t k = base + threadIdx.x;
    float cnew[NSM];
    bool pore = false;
    if (k < (uint32_t)d.npts)
      pore = d.nd == 3 ? adr_entry<float, 3, NSM, NTT>(a, d.tab, d.ta, k, cnew)
                       : adr_entry<float, 2, NSM, NTT>(a, d.tab, d.ta, k, cnew);
    if (RED) {
      // pore sums of every chemical (FP64), NSM slots per 64-entry group
      const int g = (int)threadIdx.x >> 6;
#pragma unroll
      for (int j = 0; j < NSM; ++j) group_to_lds((pore && j < a.ns) ? (double)cnew[j] : 0.0, lds, g * NSM + j);
    }
  } else {
    // entries in pairs: both entries' loads in flight together
    auto pair = [&](int e) {
      const uint32_t k[2] = {base + e * T + threadIdx.x, base + (e + 1) * T + threadIdx.x};
      if (k[0] < (uint32_t)d.npts) {
        float cnew[2][NSM];
        bool pore[2];
        if (d.nd == 3) adr_entry2<float, 3, NSM, NTT>(a, d.tab, d.ta, (uint32_t)d.npts, k, cnew, pore);
        else adr_entry2<float, 2, NSM, NTT>(a, d.tab, d.ta, (uint32_t)d.npts, k, cnew, pore);
      }
    };
    pair(0);
    if constexpr (E == 4) pair(2);
  }
  if (RED)
    tile_partials<TILE>(lds, NSM, base, d.npts, partials + (size_t)(d.rt0 + (blk - d.blk0) * (TILE / 256)) * NSM);
}

template <int NSM, bool RED, bool NTT>
static void launch_adr(const BatchShape &sh, dim3 grid, hipStream_t st, const SegAdr *d, const int32_t *bs,
                       int parity, double *partials) {
#define PRT_AS(TT, EE)                                                                            \
  if (sh.t == TT && sh.e == EE) {                                                                 \
    adr_batch_kernel<NSM, RED, NTT, TT, EE><<<grid, dim3(TT), 0, st>>>(d, bs, parity, sh.xcd, partials); \
    return;                                                                                       \
  }
  if constexpr (RED) {
    PRT_AS(256, 1)
  } else {
    PRT_AS(256, 1) PRT_AS(128, 2) PRT_AS(64, 4) PRT_AS(256, 2) PRT_AS(128, 4)
  }
#undef PRT_AS
}

template <int NSM>
static void launch_adr_n(bool red, bool ntt, const BatchShape &sh, dim3 grid, hipStream_t st, const SegAdr *d,
                         const int32_t *bs, int parity, double *partials) {
  if (red && ntt) launch_adr<NSM, true, true>(sh, grid, st, d, bs, parity, partials);
  else if (red) launch_adr<NSM, true, false>(sh, grid, st, d, bs, parity, partials);
  else if (ntt) launch_adr<NSM, false, true>(sh, grid, st, d, bs, parity, partials);
  else launch_adr<NSM, false, false>(sh, grid, st, d, bs, parity, partials);
}

}  // namespace prt

using namespace prt;

extern "C" int64_t prt_adr_batch_desc_size(void) { return (int64_t)sizeof(SegAdr); }

extern "C" int prt_adr_batch_plan_shaped(const prt_adr_segment *segs, int nseg, int shape, void *desc_host,
                                         int32_t *blk_seg_host, int64_t blk_capacity, int64_t *n_blocks,
                                         int64_t *n_tiles, int *max_species) {
  BatchShape sh;
  if (!segs || nseg < 1 || !desc_host || !n_blocks || !decode_shape(shape, sh)) {
    set_error("prt_adr_batch_plan: bad arguments");
    return PRT_E_ARG;
  }
  const int tile = sh.tile();
  SegAdr *d = static_cast<SegAdr *>(desc_host);
  long long total = 0, tiles = 0;
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
    d[s].tile = tile;
    d[s].blk0 = (int)total;
    d[s].nblk = d[s].npts > 0 ? blocks_of(d[s].npts, tile) : 0;
    d[s].rt0 = (int)tiles;
    total += d[s].nblk;
    tiles += d[s].npts > 0 ? blocks_of(d[s].npts, 256) : 0;
    if (total > 0x7fffffffLL || tiles * 8 > 0x7fffffffLL) {
      set_error("prt_adr_batch_plan: too many blocks");
      return PRT_E_RANGE;
    }
    nsm = g.params->ns > nsm ? g.params->ns : nsm;
  }
  *n_blocks = total;
  if (n_tiles) *n_tiles = tiles;
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

extern "C" int prt_adr_batch_plan(const prt_adr_segment *segs, int nseg, void *desc_host, int32_t *blk_seg_host,
                                  int64_t blk_capacity, int64_t *n_blocks, int *max_species) {
  return prt_adr_batch_plan_shaped(segs, nseg, PRT_BATCH_SHAPE_ADR_DEFAULT, desc_host, blk_seg_host, blk_capacity,
                                   n_blocks, nullptr, max_species);
}

extern "C" int prt_adr_batch_step_shaped(const void *desc_dev, const int32_t *blk_seg_dev, int64_t n_blocks,
                                         int max_species, int parity, int shape, double *partials, void *stream) {
  BatchShape sh;
  if (!desc_dev || !blk_seg_dev || n_blocks < 1 || n_blocks > 0x7fffffffLL || max_species < 1 ||
      max_species > 8 || !decode_shape(shape, sh)) {
    set_error("prt_adr_batch_step: bad arguments");
    return PRT_E_ARG;
  }
  const bool red = partials != nullptr;
  if (red && (sh.t != 256 || sh.e != 1)) {
    set_error("prt_adr_batch_step: partial sums need 256 threads with one entry each");
    return PRT_E_ARG;
  }
  const SegAdr *d = static_cast<const SegAdr *>(desc_dev);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)n_blocks);
  // the members' tables together decide the load kind, as one list's does
  const bool ntt = n_blocks * sh.tile() >= (int64_t)kNtTableEntries;
  if (max_species == 1) launch_adr_n<1>(red, ntt, sh, grid, st, d, blk_seg_dev, parity, partials);
  else if (max_species == 2) launch_adr_n<2>(red, ntt, sh, grid, st, d, blk_seg_dev, parity, partials);
  else if (max_species <= 4) launch_adr_n<4>(red, ntt, sh, grid, st, d, blk_seg_dev, parity, partials);
  else launch_adr_n<8>(red, ntt, sh, grid, st, d, blk_seg_dev, parity, partials);
  return launch_status("prt_adr_batch_step launch");
}

extern "C" int prt_adr_batch_step(const void *desc_dev, const int32_t *blk_seg_dev, int64_t n_blocks,
                                  int max_species, int parity, double *partials, void *stream) {
  return prt_adr_batch_step_shaped(desc_dev, blk_seg_dev, n_blocks, max_species, parity,
                                   PRT_BATCH_SHAPE_ADR_DEFAULT, partials, stream);
}
