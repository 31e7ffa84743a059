// prt_lbm_core.h -- device building blocks of the TRT lattice-Boltzmann
// kernels: lattice tables, population storage, the periodic pull gather,
// the post-collision scatter and the two collision forms.  Shared by the
// single-geometry kernels (prt_lbm.hip) and the batched multi-geometry
// kernels (prt_batch.hip) so both evaluate bit-identical arithmetic.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "prt_common.h"

namespace prt {

template <int Q>
struct Lat;

// D3Q19, ordering of 3D/tools/prtlb_3d.py:370-375 (required for population
// level parity with the reference's state).
template <>
struct Lat<19> {
  static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
  static constexpr int cls[19] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
};
// D2Q9 of 3D/tools/prtlb_2d.py:391-394, stored as (x, -, y) -> (x, y=0, z).
template <>
struct Lat<9> {
  static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int cz[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr int cls[9] = {0, 1, 1, 1, 1, 2, 2, 2, 2};
};

template <int Q>
__host__ __device__ constexpr double weight(int i) {
  return Q == 19 ? (Lat<19>::cls[i] == 0 ? 1.0 / 3.0 : Lat<19>::cls[i] == 1 ? 1.0 / 18.0 : 1.0 / 36.0)
                 : (Lat<9>::cls[i] == 0 ? 4.0 / 9.0 : Lat<9>::cls[i] == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
}

struct LbmArgs {
  const uint8_t *solid;
  const void *src;
  void *dst;
  double *partials;
  int X, Y, Z;
  long long N;
  long long zs;       // dense grid: row pitch, elements per population per z-row (>= Z)
  double omp, omm, g, half_g;
  double omh;         // 1 - omm/2, formed on the host as the reference does
  float s16, inv_s16; // FP16 storage scale s and 1/s
  int xoff;           // first plane a dense launch updates (1 for a slab, else 0)
};

// Dense grids store populations by z-row: the Q populations of row
// (x, y) follow one another, each a contiguous run of zs elements,
//
//   element (i, x, y, z) at ((x*Y + y)*Q + i)*zs + z,
//
// with zs = Z rounded up to a whole number of 128-byte lines.  A block
// that updates whole rows then writes one contiguous run of Q*zs elements
// per row, and reads each population from a contiguous run of its source
// row, instead of touching Q separate arrays across the whole buffer: far
// fewer DRAM pages open at once.  Every population run starts on a cache
// line, so no line is shared by two runs.
long long pop_row_pitch(int Z, int elem_bytes);
long long pop_elems(int q, int X, int Y, int Z, int elem_bytes);

// Fluid lists store populations by tile of kListTile consecutive entries,
// the same idea as the z-rows of a dense grid: the Q populations of a tile
// follow one another, each a run of kListTile elements,
//
//   element (i, k) at ((k / kListTile)*Q + i)*kListTile + k % kListTile,
//
// so a 256-entry block writes one contiguous run of Q*kListTile elements.
// A buffer holds Q * pop_stride(n) elements (n rounded up to a tile).
constexpr int kListTile = 256;
long long pop_stride(long long n);

__device__ inline int list_elem(int k, int i, int q) {
  return ((k / kListTile) * q + i) * kListTile + k % kListTile;
}

// Element stride of the link-table rows of an n_fluid fluid list.
long long link_stride(long long n_fluid);

// Guo forcing term w_i (1 - omm/2) 3 g e_ix with the reference's association
// (3D/tools/prtlb_3d.py:439-440); exact constants, contraction off.
template <int Q>
__device__ inline double guo_term(const LbmArgs &a, int i) {
#pragma clang fp contract(off)
  double t = weight<Q>(i) * a.omh;
  t = t * 3.0;
  t = t * a.g;
  return t * (double)Lat<Q>::cx[i];
}

// FP16 storage: h_i = (f_i - w_i) * s / w_i; 1/w_i is an exact constant.
template <int Q>
__device__ inline float ld16(const LbmArgs &a, int i) {
  return (float)weight<Q>(i) * a.inv_s16;
}
template <int Q>
__device__ inline float st16(const LbmArgs &a, int i) {
  return a.s16 * (float)(1.0 / weight<Q>(i));
}

// ---------------------------------------------------------------- storage
template <int ST>
struct Pop;
template <>
struct Pop<PRT_STORE_FP64> {
  using T = double;
};
template <>
struct Pop<PRT_STORE_FP32> {
  using T = float;
};
template <>
struct Pop<PRT_STORE_FP16> {
  using T = __half;
};

// Gather index of the node that streams into (x, y, z) along e_i.
struct Nbr {
  int xm, xp, ym, yp, zm, zp, x, y, z, Y, Z;
  __device__ inline int at(int ex, int ey, int ez) const {
    const int sx = ex > 0 ? xm : (ex < 0 ? xp : x);
    const int sy = ey > 0 ? ym : (ey < 0 ? yp : y);
    const int sz = ez > 0 ? zm : (ez < 0 ? zp : z);
    return (sx * Y + sy) * Z + sz;
  }
  // element of population i of that node in the z-row layout (q, zs as
  // in LbmArgs); below 2^31 by the grid check
  __device__ inline int elem(int ex, int ey, int ez, int i, int q, int zs) const {
    const int sx = ex > 0 ? xm : (ex < 0 ? xp : x);
    const int sy = ey > 0 ? ym : (ey < 0 ? yp : y);
    const int sz = ez > 0 ? zm : (ez < 0 ? zp : z);
    return ((sx * Y + sy) * q + i) * zs + sz;
  }
  // element of population 0 of (x, y, z) itself
  __device__ inline int own(int q, int zs) const { return (x * Y + y) * q * zs + z; }
};

__device__ inline Nbr make_nbr(int x, int y, int z, int X, int Y, int Z) {
  Nbr b;
  b.x = x; b.y = y; b.z = z; b.Y = Y; b.Z = Z;
  b.xm = x == 0 ? X - 1 : x - 1;
  b.xp = x == X - 1 ? 0 : x + 1;
  b.ym = y == 0 ? Y - 1 : y - 1;
  b.yp = y == Y - 1 ? 0 : y + 1;
  b.zm = z == 0 ? Z - 1 : z - 1;
  b.zp = z == Z - 1 ? 0 : z + 1;
  return b;
}

// Load the Q populations streaming into the node of nb (z-row layout;
// SKIP: from the node's own two-step-old opposite slot where the upstream
// node is solid).
template <int Q, int ST, bool SKIP, typename R, bool NTL = false>
__device__ inline void gather(const LbmArgs &a, const Nbr &nb, R *f) {
  using T = typename Pop<ST>::T;
  const T *src = static_cast<const T *>(a.src);
  const T *own = static_cast<const T *>(a.dst);
  const int zs = (int)a.zs;
  const int self = nb.own(Q, zs);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int e = (i == 0) ? self : nb.elem(Lat<Q>::cx[i], Lat<Q>::cy[i], Lat<Q>::cz[i], i, Q, zs);
    T v;
    if (SKIP && i != 0 && a.solid[nb.at(Lat<Q>::cx[i], Lat<Q>::cy[i], Lat<Q>::cz[i])])
      v = own[(size_t)self + (size_t)Lat<Q>::opp[i] * zs];
    else if constexpr (NTL && ST != PRT_STORE_FP16)
      v = __builtin_nontemporal_load(src + (size_t)e);   // read once: do not keep
    else
      v = src[(size_t)e];
    if constexpr (ST == PRT_STORE_FP16)
      f[i] = __half2float(v) * ld16<Q>(a, i);
    else
      f[i] = (R)v;
  }
}

// Sparse (fluid-list) storage: only the fluid nodes' populations, tiled as
// above (list_elem); link[(i - 1) * lk + k] is the list entry
// of the node that streams into entry k along e_i, or -1 where that node is
// solid -- then, as in gather<SKIP>, the entry's own two-step-old opposite
// slot (still in the write buffer).  Same values as the dense fluid-only
// gather, so the same bits after the collision.
//
// Link tables come in two widths.  int32: the entry itself, -1 for solid.
// int16 (kLink16Solid for solid): the entry minus k taken modulo the list
// length into [-32767, 32767], which every member fits whose fluid nodes
// per x-plane stay below 32k (the periodic wrap along x is a small step
// modulo the list); it halves the table's traffic, 4 (Q - 1) -> 2 (Q - 1)
// bytes per fluid node.
constexpr int kLink16Solid = -32768;

__device__ inline int link_entry(const int32_t *__restrict__ link, size_t at, int, int) { return link[at]; }
__device__ inline int link_entry(const int16_t *__restrict__ link, size_t at, int k, int nf) {
  const int d = link[at];
  if (d == kLink16Solid) return -1;
  const int s = k + d;
  return s < 0 ? s + nf : (s >= nf ? s - nf : s);
}

template <int Q, int ST, typename R, typename L>
__device__ inline void gather_sparse(const LbmArgs &a, const L *__restrict__ link, long long lk, int k, R *f) {
  using T = typename Pop<ST>::T;
  const T *src = static_cast<const T *>(a.src);
  const T *own = static_cast<const T *>(a.dst);
  int s[Q];
  s[0] = k;
#pragma unroll
  for (int i = 1; i < Q; ++i) s[i] = link_entry(link, (size_t)(i - 1) * lk + k, k, (int)a.N);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const T *p = (i == 0 || s[i] >= 0) ? src + list_elem(s[i], i, Q) : own + list_elem(k, Lat<Q>::opp[i], Q);
    const T v = *p;
    if constexpr (ST == PRT_STORE_FP16)
      f[i] = __half2float(v) * ld16<Q>(a, i);
    else
      f[i] = (R)v;
  }
}

// FP16 storage rounds stochastically: the 13 bits of float32 significand
// that FP16 drops get a dither added before they are cut, so an increment
// smaller than half an FP16 unit still moves the stored value with the
// matching probability instead of being rounded away at every step (with
// round-to-nearest the shifted populations of slow pore throats stall).
// The dither comes from the node's own post-collision populations (the low
// significand bits of the rest population, or of population 1 for the rest
// population itself), mixed per direction: the same state gives the same
// bits, with no counter or generator state, so every launch path agrees.
//
// Odd 24-bit multipliers, one per direction, that turn the node's seed into
// each population's dither: bits 19-31 of the 24 x 24-bit product of the
// seed's low 24 bits (significand bits) and the multiplier, one full-rate
// multiply (a 32-bit integer multiply issues at a quarter of the rate).
constexpr uint32_t kDitherMul[19] = {
    0x9E3779u, 0x85EBCBu, 0xC2B2AFu, 0x27D4EBu, 0x165667u, 0xD3A265u, 0xFD7047u,
    0xB55A4Fu, 0x8F1BBDu, 0xCA62C1u, 0x5A8279u, 0x6ED9EBu, 0x3C6EF3u, 0xA54FF5u,
    0x510E53u, 0x9B0569u, 0x1F83D9u, 0x5BE0CDu, 0xCBBB9Du};
template <int Q>
__device__ inline uint32_t f16_dither_seed(const float *fp, int i) {
  return __float_as_uint(fp[i == 0 ? 1 : 0]);
}
// x rounded to FP16 with the dither: the dither is added to the 13 bits
// FP16 drops and they are cut, which leaves at most FP16's 11 significand
// bits, so the conversion is exact in FP16's normal range.  Infinities and
// quiet NaNs keep their bits (their low 13 bits are zero, and the add does
// not reach the exponent).
__device__ inline unsigned short f16_round_dithered(float x, uint32_t seed, int i) {
  // __umul24 returns a signed int: the shift must see the product unsigned
  const uint32_t d = static_cast<uint32_t>(__umul24(seed, kDitherMul[i])) >> 19;
  return __half_as_ushort(__float2half_rn(__uint_as_float((__float_as_uint(x) + d) & ~0x1fffu)));
}

// Store the Q post-collision populations: population i at element
// base + i * stride (dense: the node's own element and zs; fluid list:
// list_elem(k, 0, Q) and kListTile).
template <int Q, int ST, typename R>
__device__ inline void scatter(const LbmArgs &a, size_t base, long long stride, const R *fp) {
  using T = typename Pop<ST>::T;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if constexpr (ST == PRT_STORE_FP16) {
      unsigned short *dst = static_cast<unsigned short *>(a.dst);
      const unsigned short v = f16_round_dithered(fp[i] * st16<Q>(a, i), f16_dither_seed<Q>(fp, i), i);
      __builtin_nontemporal_store(v, dst + base + (size_t)i * stride);
    } else {
      T *dst = static_cast<T *>(a.dst);
      __builtin_nontemporal_store((T)fp[i], dst + base + (size_t)i * stride);
    }
  }
}

// FP64 collision in the reference's own evaluation order (contraction off):
// rho = sum f, j = einsum(e, f), u = (j + g/2 e_x)/rho masked, feq, TRT with
// f+/- = (f +/- f_opp)/2, then + Guo.  3D/tools/prtlb_3d.py:142-154, 427-440.
template <int Q>
__device__ inline void collide_ref(const LbmArgs &a, bool solid, const double *f,
                                   double *fp, double &ux_out) {
#pragma clang fp contract(off)
  double rho = f[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) rho += f[i];
  double jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (Lat<Q>::cx[i]) jx += (double)Lat<Q>::cx[i] * f[i];
    if (Lat<Q>::cy[i]) jy += (double)Lat<Q>::cy[i] * f[i];
    if (Lat<Q>::cz[i]) jz += (double)Lat<Q>::cz[i] * f[i];
  }
  double ux = (jx + a.half_g) / rho;
  double uy = jy / rho;
  double uz = jz / rho;
  if (solid) ux = uy = uz = 0.0;
  ux_out = ux;
  const double usq = (ux * ux + uy * uy) + uz * uz;
  double feq[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const double eu = ((double)Lat<Q>::cx[i] * ux + (double)Lat<Q>::cy[i] * uy) +
                      (double)Lat<Q>::cz[i] * uz;
    feq[i] = (weight<Q>(i) * rho) * (((1.0 + 3.0 * eu) + (4.5 * eu) * eu) - 1.5 * usq);
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int o = Lat<Q>::opp[i];
    const double f_p = 0.5 * (f[i] + f[o]);
    const double f_m = 0.5 * (f[i] - f[o]);
    const double e_p = 0.5 * (feq[i] + feq[o]);
    const double e_m = 0.5 * (feq[i] - feq[o]);
    double v = (f[i] - a.omp * (f_p - e_p)) - a.omm * (f_m - e_m);
    v = v + guo_term<Q>(a, i);
    fp[i] = solid ? f[o] : v;
  }
}

// FP32 collision on deviations g_i = f_i - w_i (the shifted storages).
// Contraction is off so every kernel that inlines this (single-geometry,
// batched) evaluates exactly the same operations: a geometry gives the same
// bits alone or in a batch.
template <int Q>
__device__ inline void collide_dev(const LbmArgs &a, bool solid, const float *gq,
                                   float *out, double &ux_out) {
#pragma clang fp contract(off)
  float dr = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    dr += gq[i];
    if (Lat<Q>::cx[i]) jx += Lat<Q>::cx[i] > 0 ? gq[i] : -gq[i];
    if (Lat<Q>::cy[i]) jy += Lat<Q>::cy[i] > 0 ? gq[i] : -gq[i];
    if (Lat<Q>::cz[i]) jz += Lat<Q>::cz[i] > 0 ? gq[i] : -gq[i];
  }
  const float rho = 1.f + dr;
  const float irho = 1.f / rho;
  float ux = (jx + (float)a.half_g) * irho, uy = jy * irho, uz = jz * irho;
  if (solid) ux = uy = uz = 0.f;
  ux_out = ux;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float omp = (float)a.omp, omm = (float)a.omm;
  {
    const float eq0 = (float)weight<Q>(0) * (dr - 1.5f * rho * usq);
    out[0] = solid ? gq[0] : gq[0] - omp * (gq[0] - eq0);
  }
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    const int o = Lat<Q>::opp[i];
    if (o < i) continue;
    const float w = (float)weight<Q>(i);
    const float eu = (float)Lat<Q>::cx[i] * ux + (float)Lat<Q>::cy[i] * uy + (float)Lat<Q>::cz[i] * uz;
    const float sym = w * (dr + rho * (4.5f * eu * eu - 1.5f * usq));
    const float asym = w * rho * 3.f * eu;
    const float gp = 0.5f * (gq[i] + gq[o]);
    const float gm = 0.5f * (gq[i] - gq[o]);
    const float A = omp * (gp - sym);
    const float B = omm * (gm - asym);
    const float F = (float)guo_term<Q>(a, i);   // F_opp = -F
    const float vi = gq[i] - A - B + F;
    const float vo = gq[o] - A + B - F;
    out[i] = solid ? gq[o] : vi;
    out[o] = solid ? gq[i] : vo;
  }
}

}  // namespace prt
