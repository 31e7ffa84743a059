// prt_batch.h -- what the batched Stokes kernels (prt_batch.hip) and the
// batched transport kernels (prt_batch_adr.hip) share: the per-member
// descriptors, the launch shapes a tuner picks from, and the per-tile sums.
#pragma once

#include <hip/hip_runtime.h>

#include "prt_adr_core.h"
#include "prt_lbm_core.h"

namespace prt {

// Per-member device descriptors.  blk0 / nblk: the member's blocks in the
// launch (tile list entries each); rt0: its first partial-sum slot (one
// per 256 entries, whatever the tile).
struct SegLbm {
  LbmArgs a;        // src/dst chosen per launch from fa/fb and the phase
  void *fa, *fb;
  const void *link;     // fluid-list member: link table (a.N = list length)
  long long lk;
  int phase, blk0, nblk, link16;  // link16: 16-bit link table
  int tile, rt0;
};

struct SegAdr {
  AdrArgs<float> a;
  float *ca, *cb;
  const void *tab;       // pore table of the segment's list
  uint32_t ta;           // its entries per section (tab_entries)
  int npts;              // list entries
  int phase, blk0, nblk, nd;
  int tile, rt0;
};

// Launch shape of a batched step (PRT_BATCH_SHAPE in prtlb.h): T threads
// per block, E consecutive runs of T list entries per thread (entry
// base + e*T + thread, so every load of a wave stays as contiguous as with
// one entry per thread), and the block order.  A block covers T*E = 256 or
// 512 entries.
struct BatchShape {
  int t, e, xcd;
  int tile() const { return t * e; }
};

inline bool decode_shape(int code, BatchShape &s) {
  const int tl = code & 15, el = (code >> 4) & 15, x = code >> 8;
  if (x < 0 || x > 1 || tl < 6 || tl > 8 || el > 2 || tl + el < 8 || tl + el > 9) return false;
  s.t = 1 << tl;
  s.e = 1 << el;
  s.xcd = x;
  return true;
}

// Partial sums go per 256-entry tile in one order whatever the launch
// shape: each 64-entry group is summed by a wavefront butterfly into slot g
// of LDS, and a tile's partial is 0 + g0 + g1 + g2 + g3 -- block_sum's own
// order for a 256-thread block, so every shape gives the same bits.
__device__ inline void group_to_lds(double v, double *lds, int slot) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) lds[slot] = v;
}

// After every group of the block is in LDS (width values per group):
// out[r * width + j] for each 256-entry tile r of the block that starts
// below n.
template <int TILE>
__device__ inline void tile_partials(const double *lds, int width, long long base, long long n, double *out) {
  __syncthreads();
  constexpr int kTiles = TILE / 256;
  if ((int)threadIdx.x < kTiles * width) {
    const int r = (int)threadIdx.x / width, j = (int)threadIdx.x % width;
    if (base + 256LL * r < n) {
      double s = 0.0;
#pragma unroll
      for (int g = 0; g < 4; ++g) s += lds[(4 * r + g) * width + j];
      out[r * width + j] = s;
    }
  }
}

inline int blocks_of(long long n, int tile) { return (int)((n + tile - 1) / tile); }

}  // namespace prt
