// prt_adr_core.h -- one voxel of the fused finite-volume transport step,
// shared by the single-geometry kernel (prt_adr.hip) and the batched
// multi-geometry kernel (prt_batch.hip), so both evaluate identical
// arithmetic.  See prt_adr.hip for the scheme and its float32 contract.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "prt_common.h"

namespace prt {

template <typename R>
struct AdrArgs {
  const uint8_t *flags;
  const R *u;
  const R *cin;
  R *cout;
  int X, Y, Z;
  long long N;
  int ns, biotic_on, coupled, abio_index, abio_target, abio_product, surface_only;
  int sorption, sorb_c, sorb_s, write_solid;
  int imm_list;  // immobile chemicals stored in pore-list order (list launches)
  int mad24;     // every axis stride in bytes fits a signed 24-bit operand
  R dt, k, ka, kd, ks_d, ks_a, alpha_dt, neg_alpha, c2, yield_, c3_init, k_sorb, k_desorb;
  int alpha_pos;
  FastDiv div_yz, div_z;                 // flat index -> (x, y, z)
  int mobile[8], pinned[8];
  R fed[8];
  R dface[8][4];
};

// np.minimum / np.maximum: the first argument wins ties and NaN propagates.
template <typename R>
__device__ inline R npmin(R a, R b) {
  return (a <= b || a != a) ? a : b;
}
template <typename R>
__device__ inline R npmax(R a, R b) {
  return (a >= b || a != a) ? a : b;
}
// np.sign: +-1 with the sign of d, and d itself for +-0 and NaN (one class
// test and a sign copy instead of two compares).
template <typename R>
__device__ inline R npsign(R d) {
  const bool keep = __builtin_amdgcn_class(d, 0x63);  // NaN, -0 or +0
  return keep ? d : __builtin_copysign(R(1), d);
}

// C_face = up + 1/2 minmod(down - up, up - upup), 3D/tools/prtlb_3d.py:700-705.
// The magnitude is the hardware minimum: it differs from np.minimum only
// when d1 or d2 is NaN, and then the sign factor is NaN and so is the slope.
template <typename R>
__device__ inline R muscl(R up, R upup, R down) {
#pragma clang fp contract(off)
  const R d1 = down - up;
  const R d2 = up - upup;
  const R slope = (R(0.5) * (npsign(d1) + npsign(d2))) * fmin(fabs(d1), fabs(d2));
  return up + R(0.5) * slope;
}

// base[i] through a 32-bit byte offset from a wave-uniform base: one scalar
// base register plus one vector offset, no 64-bit address arithmetic.  The
// argument check keeps every offset below 2^32.
template <typename T>
__device__ inline T ld32(const T *base, uint32_t i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T));
}
// The same for data read once per step and never again by a neighbour (the
// pore list and table of a large grid): a nontemporal load, so it does not
// displace the concentration lines the stencil re-reads.
template <typename T>
__device__ inline T ldnt32(const T *base, uint32_t i) {
  return __builtin_nontemporal_load(
      reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ inline void st32(T *base, uint32_t i, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T)) = v;
}

// Pore-table word of one pore voxel (one uint32):
//   bits 4a .. 4a+3, axis a = 0 .. ND-1: reach on the low and the high side
//     (2 bits each), which places the neighbour values the scheme reads:
//     the second one on a side r strides away and the first min(r, 1);
//   bits 12+4a .. 12+4a+3: low face open, high face open, low neighbour
//     biofilm, high neighbour biofilm (a face is open when the neighbour
//     across it is pore; a face on the box edge is open, towards the voxel
//     itself, as the edge padding makes it);
//   bit 24: the voxel itself is biofilm;
//   bits 25-26: min(x, 3), for the inlet pin;
//   bits 27+2a, 28+2a, axes a < ND-1: far low face open, far high face open.
// Along the axes a < ND-1 the reach is the distance to the box edge, capped
// at 2 (the edge clamp), and the face bits pick among the values read.
// Along the last axis, whose neighbours share the voxel's cache lines, the
// reach is cut back to the values the scheme uses, so no select is needed:
// the first value is the neighbour across the face when that face is open
// and not on the box edge, else the voxel itself; the second is the next
// one out when that is pore and inside the box, else the first.
// 0xFFFFFFFF (reach 3) marks a list entry that is not a pore voxel.
constexpr uint32_t kPoreWordNone = 0xFFFFFFFFu;
constexpr int kWordFace = 12, kWordBiofilm = 24, kWordX = 25, kWordFar = 27;

// Geometry of one pore voxel for one step: the byte offsets of the
// neighbour values in an array of R, the word above and the face
// velocities.  Built from the flag and velocity grids (dense launch, table
// build) or read from the pore table (list launch); both give the same
// values, so the update is bit-identical.
template <typename R, int ND>
struct VoxGeom {
  uint32_t noff[ND][4];  // [axis][m2, m1, p1, p2], bytes
  uint32_t w;
  R u_lo[ND], u_hi[ND];
};

template <typename R>
__device__ inline int axis_stride(const AdrArgs<R> &a, int ax3) {
  return ax3 == 0 ? a.Y * a.Z : (ax3 == 1 ? a.Z : 1);
}

// base at a byte offset (a wave-uniform base plus one vector offset).
template <typename T>
__device__ inline T ldb(const T *base, uint32_t off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// Neighbour offsets of voxel n from reach codes rc (bits 4a .. 4a+3): the
// second value on a side sits r strides away and the first min(r, 1).
// With a.mad24 each offset is one 24-bit multiply-add of the reach and the
// stride in bytes; otherwise the strides are picked with selects.
template <typename R, int ND>
__device__ inline void geom_index(const AdrArgs<R> &a, uint32_t n, uint32_t rc, VoxGeom<R, ND> &g) {
  const uint32_t nb = n * (uint32_t)sizeof(R);
  // 3D axes x, y, z; a 2D grid is (X, 1, Z) with axes x, z
  auto sbytes = [&](int ax) { return axis_stride(a, (ND == 2 && ax == 1) ? 2 : ax) * (int)sizeof(R); };
  if (a.mad24) {
#pragma unroll
    for (int ax = 0; ax < ND; ++ax) {
      const int sb = sbytes(ax);
      const int rl = (rc >> (4 * ax)) & 3u, rh = (rc >> (4 * ax + 2)) & 3u;
      g.noff[ax][0] = nb + (uint32_t)__mul24(rl, -sb);
      g.noff[ax][1] = nb + (uint32_t)__mul24(min(rl, 1), -sb);
      g.noff[ax][2] = nb + (uint32_t)__mul24(min(rh, 1), sb);
      g.noff[ax][3] = nb + (uint32_t)__mul24(rh, sb);
    }
  } else {
#pragma unroll
    for (int ax = 0; ax < ND; ++ax) {
      const int sb = sbytes(ax);
      const int rl = (rc >> (4 * ax)) & 3u, rh = (rc >> (4 * ax + 2)) & 3u;
      const int om1 = rl ? sb : 0, op1 = rh ? sb : 0;
      g.noff[ax][0] = nb - (uint32_t)(rl == 2 ? 2 * sb : om1);
      g.noff[ax][1] = nb - (uint32_t)om1;
      g.noff[ax][2] = nb + (uint32_t)op1;
      g.noff[ax][3] = nb + (uint32_t)(rh == 2 ? 2 * sb : op1);
    }
  }
}

// Geometry from the grids: pore voxel n = (x*Y + y)*Z + z with flag byte fl.
// The neighbours are first taken with the edge clamp alone (reach = the
// distance to the box edge, capped at 2), which fixes the face bits and
// velocities; along the last axis the reach is then cut back where a face or
// the face beyond it is closed.
template <typename R, int ND>
__device__ inline void geom_dense(const AdrArgs<R> &a, uint32_t n, uint8_t fl, int x, int y, int z,
                                  VoxGeom<R, ND> &g) {
#pragma clang fp contract(off)
  uint32_t edge = 0;
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) {
    const int ax3 = (ND == 2 && ax == 1) ? 2 : ax;
    const int ext = ax3 == 0 ? a.X : (ax3 == 1 ? a.Y : a.Z);
    const int p = ax3 == 0 ? x : (ax3 == 1 ? y : z);
    edge |= ((uint32_t)min(p, 2) | (uint32_t)min(ext - 1 - p, 2) << 2) << (4 * ax);
  }
  geom_index<R, ND>(a, n, edge, g);
  uint32_t w = ((uint32_t)(fl >> 1) & 1u) << kWordBiofilm | (uint32_t)min(x, 3) << kWordX;
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) {
    uint8_t nf[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) nf[k] = ld32(a.flags, g.noff[ax][k] / (uint32_t)sizeof(R));
    const R *ua = a.u + (size_t)ax * a.N;
    const R um = ldb(ua, g.noff[ax][1]), uc = ld32(ua, n), up = ldb(ua, g.noff[ax][2]);
    const bool o_lo = nf[1] & 1, o_hi = nf[2] & 1;
    const bool f_lo = o_lo && (nf[0] & 1);
    const bool f_hi = o_hi && (nf[3] & 1);
    g.u_lo[ax] = o_lo ? R(0.5) * (uc + um) : R(0);
    g.u_hi[ax] = o_hi ? R(0.5) * (uc + up) : R(0);
    const uint32_t el = (edge >> (4 * ax)) & 3u, eh = (edge >> (4 * ax + 2)) & 3u;
    if (ax == ND - 1) {
      const uint32_t rl = o_lo ? (f_lo ? el : min(el, 1u)) : 0u;
      const uint32_t rh = o_hi ? (f_hi ? eh : min(eh, 1u)) : 0u;
      w |= (rl | rh << 2) << (4 * ax);
    } else {
      w |= (el | eh << 2) << (4 * ax);
      w |= ((uint32_t)f_lo | (uint32_t)f_hi << 1) << (kWordFar + 2 * ax);
    }
    w |= ((uint32_t)o_lo | (uint32_t)o_hi << 1 | (uint32_t)((nf[1] >> 1) & 1) << 2 |
          (uint32_t)((nf[2] >> 1) & 1) << 3)
         << (kWordFace + 4 * ax);
  }
  g.w = w;
  // only the last axis changes; its stride is one element
  const uint32_t nb = n * (uint32_t)sizeof(R), sb = (uint32_t)sizeof(R);
  const uint32_t rl = (w >> (4 * (ND - 1))) & 3u, rh = (w >> (4 * (ND - 1) + 2)) & 3u;
  g.noff[ND - 1][0] = nb - rl * sb;
  g.noff[ND - 1][1] = nb - min(rl, 1u) * sb;
  g.noff[ND - 1][2] = nb + min(rh, 1u) * sb;
  g.noff[ND - 1][3] = nb + rh * sb;
}

// All ones when bit b of w is set, else zero (one sign-extending bit
// extract).
__device__ inline uint32_t bitmask(uint32_t w, int b) {
  return (uint32_t)((int32_t)(w << (31 - b)) >> 31);
}
// m ? x : y for a mask m from bitmask: a bitwise select of the two values'
// bits, so it is exact for every value (NaN included) and needs no compare.
// In float32 it is written as the one bitfield-insert instruction: left to
// itself the compiler turns the bit operations back into a bit test, a
// compare and a conditional move per select.
template <typename R>
__device__ inline R bsel(uint32_t m, R x, R y) {
  if constexpr (sizeof(R) == 4) {
    const uint32_t xi = __builtin_bit_cast(uint32_t, x), yi = __builtin_bit_cast(uint32_t, y);
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(xi), "v"(yi));
    return __builtin_bit_cast(R, r);
  } else {
    const uint64_t m2 = (uint64_t)(int64_t)(int32_t)m;
    const uint64_t xi = __builtin_bit_cast(uint64_t, x), yi = __builtin_bit_cast(uint64_t, y);
    return __builtin_bit_cast(R, (xi & m2) | (yi & ~m2));
  }
}

template <typename R>
using vec4 = R __attribute__((ext_vector_type(4)));
template <typename R>
using vec2 = R __attribute__((ext_vector_type(2)));

// Pore table of a list of n entries: one device buffer of sections, each
// holding ta entries (n rounded up to 64), every entry of a section one
// aligned 8- or 16-byte record, so a wave reads each section as one
// contiguous run.
//   float32: section 0  {pore index, word, u_lo0, u_hi0}        16 B
//            section 1  {u_lo1, u_hi1, u_lo2, u_hi2} (3D)      16 B
//                       {u_lo1, u_hi1}              (2D)        8 B
//   FP64:    section 0  {pore index, word}                       8 B
//            section a+1 {u_lo_a, u_hi_a}, a < nd                16 B
// (u_lo_a, u_hi_a: the velocities on the low and high face of the voxel
// along axis a, 3D/tools/prtlb_3d.py:1100-1135.)
inline long long tab_entries(long long n) { return (n + 63) / 64 * 64; }
inline long long tab_bytes(int nd, bool fp64, long long n) {
  const long long ta = tab_entries(n);
  if (!fp64) return ta * (16 + (nd == 3 ? 16 : 8));
  return ta * (8 + 16LL * nd);
}

// Pore lists (one list, or all members of a batch together) at least this
// long stream their table with nontemporal loads: the table (32 B per pore
// voxel in float32) then exceeds what the 256 MB last-level cache keeps
// between steps next to the concentration arrays.
constexpr uint32_t kNtTableEntries = 1u << 22;

// Load of a table record: nontemporal (NTT) for lists too large to stay in
// the last-level cache between steps, ordinary otherwise.
template <bool NTT, typename T>
__device__ inline T ldtab(const T *base, uint32_t i) {
  if constexpr (NTT) return ldnt32(base, i);
  else return ld32(base, i);
}

// The head of entry k: its voxel index and word (and, in float32, the face
// velocities of axis 0, which share the record).
template <typename R, int ND, bool NTT>
__device__ inline void tab_head(const void *tab, uint32_t k, uint32_t &n, uint32_t &w, VoxGeom<R, ND> &g) {
  if constexpr (sizeof(R) == 4) {
    using U4 = uint32_t __attribute__((ext_vector_type(4)));
    const U4 h = ldtab<NTT>(static_cast<const U4 *>(tab), k);
    // components to scalars first: a bit cast of a vector component reads
    // the start of the vector
    const uint32_t lo = h.z, hi = h.w;
    n = h.x;
    w = h.y;
    g.u_lo[0] = __builtin_bit_cast(R, lo);
    g.u_hi[0] = __builtin_bit_cast(R, hi);
  } else {
    using U2 = uint32_t __attribute__((ext_vector_type(2)));
    const U2 h = ldtab<NTT>(static_cast<const U2 *>(tab), k);
    n = h.x;
    w = h.y;
  }
}

// The face velocities of entry k not in its head.
template <typename R, int ND, bool NTT>
__device__ inline void tab_faces(const void *tab, uint32_t ta, uint32_t k, VoxGeom<R, ND> &g) {
  const char *t = static_cast<const char *>(tab);
  if constexpr (sizeof(R) == 4) {
    const char *s1 = t + (size_t)ta * 16;
    if constexpr (ND == 3) {
      const vec4<R> f = ldtab<NTT>(reinterpret_cast<const vec4<R> *>(s1), k);
      g.u_lo[1] = f.x; g.u_hi[1] = f.y; g.u_lo[2] = f.z; g.u_hi[2] = f.w;
    } else {
      const vec2<R> f = ldtab<NTT>(reinterpret_cast<const vec2<R> *>(s1), k);
      g.u_lo[1] = f.x; g.u_hi[1] = f.y;
    }
  } else {
#pragma unroll
    for (int ax = 0; ax < ND; ++ax) {
      const vec2<R> f = ldtab<NTT>(reinterpret_cast<const vec2<R> *>(t + (size_t)ta * (8 + 16 * ax)), k);
      g.u_lo[ax] = f.x;
      g.u_hi[ax] = f.y;
    }
  }
}

// Store of entry k's record (the table build).
template <typename R, int ND>
__device__ inline void tab_store(void *tab, uint32_t ta, uint32_t k, uint32_t n, uint32_t w, const VoxGeom<R, ND> &g) {
  char *t = static_cast<char *>(tab);
  if constexpr (sizeof(R) == 4) {
    using U4 = uint32_t __attribute__((ext_vector_type(4)));
    const R lo = g.u_lo[0], hi = g.u_hi[0];
    reinterpret_cast<U4 *>(t)[k] = U4{n, w, __builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi)};
    char *s1 = t + (size_t)ta * 16;
    if constexpr (ND == 3)
      reinterpret_cast<vec4<R> *>(s1)[k] = vec4<R>{g.u_lo[1], g.u_hi[1], g.u_lo[2], g.u_hi[2]};
    else
      reinterpret_cast<vec2<R> *>(s1)[k] = vec2<R>{g.u_lo[1], g.u_hi[1]};
  } else {
    using U2 = uint32_t __attribute__((ext_vector_type(2)));
    reinterpret_cast<U2 *>(t)[k] = U2{n, w};
#pragma unroll
    for (int ax = 0; ax < ND; ++ax)
      reinterpret_cast<vec2<R> *>(t + (size_t)ta * (8 + 16 * ax))[k] = vec2<R>{g.u_lo[ax], g.u_hi[ax]};
  }
}

// The step at one pore voxel n (list entry k) with geometry g, given the
// values its stencil reads: c[j] at the voxel and cn[j][axis][m2, m1, p1,
// p2] as the reach codes place them (mobile chemicals; an immobile one is
// only read at the voxel).  Writes the new values to cout and returns them
// in cnew.  An immobile chemical is only ever read and written at the voxel
// itself, so with a.imm_list its row holds the pore voxels in list order
// (entry k at k, rows of mobile chemicals stay box arrays for the stencil):
// its load and store are then contiguous across a wave instead of scattered
// over the box.
template <typename R, int ND, int NSM>
__device__ inline void adr_apply(const AdrArgs<R> &a, uint32_t n, uint32_t k, const VoxGeom<R, ND> &g,
                                 R (&c)[NSM], R (&cn)[NSM][ND][4], R *cnew) {
#pragma clang fp contract(off)
  const int ns = a.ns;
  const uint32_t w = g.w;
  const int bf_self = (w >> kWordBiofilm) & 1;
  R tr[NSM];
#pragma unroll
  for (int j = 0; j < NSM; ++j) tr[j] = R(0);
  // the inlet pin (3D/tools/prtlb_3d.py:1398-1400) applied as the values are
  // read: a voxel on x = 0 that is read at all is a pore voxel.  Only voxels
  // with x <= 2 can see x = 0; along x the values read sit at x - reach,
  // x - min(reach, 1), x + min(reach, 1) and x + reach.
  const uint32_t xpos = (w >> kWordX) & 3u;
  const bool x_is0 = xpos == 0;
  if (xpos <= 2) {
    const uint32_t el = w & 3u, eh = (w >> 2) & 3u;
    bool onx0[ND][4];
#pragma unroll
    for (int ax = 0; ax < ND; ++ax)
#pragma unroll
      for (int k = 0; k < 4; ++k) onx0[ax][k] = x_is0;
    onx0[0][0] = xpos == el;
    onx0[0][1] = xpos == min(el, 1u);
    onx0[0][2] = x_is0 && eh == 0;
    onx0[0][3] = x_is0 && eh == 0;
#pragma unroll
    for (int j = 0; j < NSM; ++j) {
      if (j >= ns || !a.pinned[j]) continue;
      if (x_is0) c[j] = a.fed[j];
      if (!a.mobile[j]) continue;
#pragma unroll
      for (int ax = 0; ax < ND; ++ax)
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (onx0[ax][k]) cn[j][ax][k] = a.fed[j];
    }
  }

  // diffusion coefficients of a face, from the 2x2 table (own and
  // neighbour biofilm), picked with selects: a runtime index into the
  // kernel argument block would copy the whole block to scratch memory
  R t0[NSM], t1[NSM];
#pragma unroll
  for (int j = 0; j < NSM; ++j) {
    t0[j] = bf_self ? a.dface[j][2] : a.dface[j][0];
    t1[j] = bf_self ? a.dface[j][3] : a.dface[j][1];
  }
  R divu = R(0);
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) {
    // face bits as all-ones/zero masks: each select below is one bitwise
    // select instead of a bit test, a compare and a conditional move
    const int fb = kWordFace + 4 * ax, ff = kWordFar + 2 * ax;
    const uint32_t m_lo = bitmask(w, fb), m_hi = bitmask(w, fb + 1);
    const uint32_t m_bm1 = bitmask(w, fb + 2), m_bp1 = bitmask(w, fb + 3);
    const bool direct = ax == ND - 1;
    const uint32_t m_flo = direct ? 0u : bitmask(w, ff), m_fhi = direct ? 0u : bitmask(w, ff + 1);
    const R u_lo = g.u_lo[ax], u_hi = g.u_hi[ax];
    divu = (ax == 0) ? R(0) + (u_hi - u_lo) : divu + (u_hi - u_lo);
#pragma unroll
    for (int j = 0; j < NSM; ++j) {
      if (j >= ns || !a.mobile[j]) continue;
      // along the last axis the reach codes already point a closed face's
      // values at the voxel (Cm1 = C) and a closed far face's at the first
      // neighbour (Cm2 = Cm1); along the others the face bits pick
      const R C = c[j];
      const R Cm1 = direct ? cn[j][ax][1] : bsel(m_lo, cn[j][ax][1], C);
      const R Cp1 = direct ? cn[j][ax][2] : bsel(m_hi, cn[j][ax][2], C);
      const R Cm2 = direct ? cn[j][ax][0] : bsel(m_flo, cn[j][ax][0], Cm1);
      const R Cp2 = direct ? cn[j][ax][3] : bsel(m_fhi, cn[j][ax][3], Cp1);
      // a closed face's difference is +0 and its coefficient is not zeroed:
      // t * +0 = 0 * +0 for a finite t >= 0
      const R d_lo = bsel(m_bm1, t1[j], t0[j]);
      const R d_hi = bsel(m_bp1, t1[j], t0[j]);
      const R dif = d_hi * (Cp1 - C) + d_lo * (Cm1 - C);
      tr[j] = (ax == 0) ? R(0) + dif : tr[j] + dif;
      // upwind face values: the upwind side picks the three muscl operands,
      // so one evaluation per face (u >= 0 is false for NaN, as in np.where)
      const bool uh = u_hi >= R(0), ul = u_lo >= R(0);
      const R fh = muscl(uh ? C : Cp1, uh ? Cm1 : Cp2, uh ? Cp1 : C);
      const R flo = muscl(ul ? Cm1 : C, ul ? Cm2 : Cp1, ul ? C : Cm1);
      tr[j] = tr[j] - (u_hi * fh - u_lo * flo);
    }
  }
#pragma unroll
  for (int j = 0; j < NSM; ++j)
    if (j < ns && a.mobile[j]) tr[j] = tr[j] + c[j] * divu;

  // closed faces over 2*nd, edge-replicated neighbours (:1000-1018)
  uint32_t open_bits = 0;
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) open_bits |= 3u << (kWordFace + 4 * ax);
  const int closed = 2 * ND - __builtin_popcount(w & open_bits);

  // ---- reactions, 3D/tools/prtlb_3d.py:1224-1367
  R r[NSM];
#pragma unroll
  for (int j = 0; j < NSM; ++j) r[j] = R(0);
  R Rb = R(0);
  if (a.biotic_on) {
    const R c0 = c[0];
    Rb = a.ks_d > R(0) ? (a.k * c0) / (a.ks_d + c0) : a.k * c0;
    if (ns > 1) {
      const R c1 = c[NSM > 1 ? 1 : 0];
      Rb = Rb * (a.ks_a <= R(0) ? c1 : c1 / (a.ks_a + c1));
    }
    if (a.coupled) Rb = (Rb * c[NSM > 3 ? 3 : 0]) / a.c3_init;
    R lim = c0 / a.dt;
    if (ns > 1 && a.alpha_pos) lim = npmin(lim, c[NSM > 1 ? 1 : 0] / a.alpha_dt);
    Rb = npmin(Rb, R(0.5) * lim);
  }
  R Ra = R(0);
  if (a.abio_index >= 0) {
    R cj = R(0);
#pragma unroll
    for (int j = 0; j < NSM; ++j)
      if (j == a.abio_index) cj = c[j];
    R surf = R(1);
    if (a.surface_only) {
      surf = R(closed) / R(2 * ND);
      Ra = (a.ka * cj) * surf;
    } else {
      Ra = a.ka * cj;
    }
    Ra = npmin(Ra, (R(0.5) * cj) / a.dt);
  }
  r[0] = -Rb;
  if (NSM > 1 && ns > 1) r[NSM > 1 ? 1 : 0] = a.neg_alpha * Rb;
  if (NSM > 2 && ns > 2) r[NSM > 2 ? 2 : 0] = a.c2 * Rb;
  if (NSM > 3 && ns > 3)
    r[NSM > 3 ? 3 : 0] = a.biotic_on ? a.yield_ * Rb - a.kd * c[NSM > 3 ? 3 : 0] : R(0);
  if (a.abio_target >= 0) {
#pragma unroll
    for (int j = 0; j < NSM; ++j)
      if (j == a.abio_target) r[j] = r[j] - Ra;
    if (a.abio_product >= 0)
#pragma unroll
      for (int j = 0; j < NSM; ++j)
        if (j == a.abio_product) r[j] = r[j] + Ra;
  }
  if (a.sorption) {
    // reversible sorption C <-> S (build-defined): forward ka C (x surface
    // weight when surface_only), back kdes S, each capped at half of what
    // its source holds per step.  No counterpart in the reference.
    R cs = R(0), ss = R(0);
#pragma unroll
    for (int j = 0; j < NSM; ++j) {
      if (j == a.sorb_c) cs = c[j];
      if (j == a.sorb_s) ss = c[j];
    }
    R surf = R(1);
    if (a.surface_only) surf = R(closed) / R(2 * ND);
    R fw = (a.k_sorb * cs) * surf;
    fw = npmin(fw, (R(0.5) * cs) / a.dt);
    R bk = a.k_desorb * ss;
    bk = npmin(bk, (R(0.5) * ss) / a.dt);
    const R net = fw - bk;
#pragma unroll
    for (int j = 0; j < NSM; ++j) {
      if (j == a.sorb_c) r[j] = r[j] - net;
      if (j == a.sorb_s) r[j] = r[j] + net;
    }
  }

  // ---- C + dt (T + R), clip at zero (the voxel is in the pore)
#pragma unroll
  for (int j = 0; j < NSM; ++j) {
    if (j >= ns) continue;
    const R t = a.mobile[j] ? tr[j] : R(0);
    const R v = npmax(c[j] + a.dt * (t + r[j]), R(0));
    st32(a.cout + (size_t)j * a.N, (!a.mobile[j] && a.imm_list) ? k : n, v);
    cnew[j] = v;
  }
}

// Every value the stencil of voxel n (list entry k) may need, loaded
// unconditionally through the offsets of g (the reach codes keep every
// offset inside the box).
template <typename R, int ND, int NSM>
__device__ inline void adr_gather(const AdrArgs<R> &a, uint32_t n, uint32_t k, const VoxGeom<R, ND> &g,
                                  R (&c)[NSM], R (&cn)[NSM][ND][4]) {
#pragma unroll
  for (int j = 0; j < NSM; ++j) {
    c[j] = R(0);
    if (j >= a.ns) continue;
    const R *cj = a.cin + (size_t)j * a.N;
    if (!a.mobile[j]) {
      c[j] = ld32(cj, a.imm_list ? k : n);
      continue;
    }
    c[j] = ld32(cj, n);
#pragma unroll
    for (int ax = 0; ax < ND; ++ax)
#pragma unroll
      for (int m = 0; m < 4; ++m) cn[j][ax][m] = ldb(cj, g.noff[ax][m]);
  }
}

// One voxel (flat index n = (x*Y + y)*Z + z) from the grids.  Returns false
// (and writes nothing unless write_solid) outside the pore.
template <typename R, int ND, int NSM>
__device__ inline bool adr_voxel(const AdrArgs<R> &a, uint32_t n, int x, int y, int z, R *cnew) {
  const uint8_t fl = ld32(a.flags, n);
  if (!(fl & 1)) {
    if (a.write_solid)
#pragma unroll
      for (int j = 0; j < NSM; ++j)
        if (j < a.ns) st32(a.cout + (size_t)j * a.N, n, R(0));
    return false;
  }
  VoxGeom<R, ND> g;
  geom_dense<R, ND>(a, n, fl, x, y, z, g);
  R c[NSM], cn[NSM][ND][4];
  adr_gather<R, ND, NSM>(a, n, n, g, c, cn);
  adr_apply<R, ND, NSM>(a, n, n, g, c, cn, cnew);
  return true;
}

// List entry k (k below the list length) from the pore table: the whole
// record is requested at once, then every value the stencil reads in one
// round.  Returns false for an entry marked not pore (kPoreWordNone).
template <typename R, int ND, int NSM, bool NTT>
__device__ inline bool adr_entry(const AdrArgs<R> &a, const void *tab, uint32_t ta, uint32_t k, R *cnew) {
  VoxGeom<R, ND> g;
  uint32_t n, w;
  tab_head<R, ND, NTT>(tab, k, n, w, g);
  tab_faces<R, ND, NTT>(tab, ta, k, g);
  if (w == kPoreWordNone || n >= (uint32_t)a.N) return false;
  g.w = w;
  geom_index<R, ND>(a, n, w, g);
  R c[NSM], cn[NSM][ND][4];
  adr_gather<R, ND, NSM>(a, n, k, g, c, cn);
  adr_apply<R, ND, NSM>(a, n, k, g, c, cn, cnew);
  return true;
}

// Two list entries k[0], k[1] of an npts-entry list (npts >= 1) in one
// pass: both records, then every value both stencils read, all requested
// before either update, so the second entry's loads are in flight while
// the first is computed (cin and cout are distinct buffers, so no load of
// one entry depends on a store of the other).  Every load is unconditional:
// an entry past the list reads the last record, and one that is not a pore
// voxel reads voxel 0 with zero reach; both are then skipped.  pore[e] says
// which entries were updated.  The same updates as adr_entry, entry by
// entry.
template <typename R, int ND, int NSM, bool NTT>
__device__ inline void adr_entry2(const AdrArgs<R> &a, const void *tab, uint32_t ta, uint32_t npts,
                                  const uint32_t (&k)[2], R (&cnew)[2][NSM], bool (&pore)[2]) {
  VoxGeom<R, ND> g[2];
  uint32_t n[2], w[2], kk[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    kk[e] = min(k[e], npts - 1);
    tab_head<R, ND, NTT>(tab, kk[e], n[e], w[e], g[e]);
    tab_faces<R, ND, NTT>(tab, ta, kk[e], g[e]);
  }
  R c[2][NSM], cn[2][NSM][ND][4];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    pore[e] = k[e] < npts && w[e] != kPoreWordNone && n[e] < (uint32_t)a.N;
    const uint32_t ne = pore[e] ? n[e] : 0u, we = pore[e] ? w[e] : 0u;
    g[e].w = we;
    geom_index<R, ND>(a, ne, we, g[e]);
    adr_gather<R, ND, NSM>(a, ne, pore[e] ? kk[e] : 0u, g[e], c[e], cn[e]);
  }
#pragma unroll
  for (int e = 0; e < 2; ++e)
    if (pore[e]) adr_apply<R, ND, NSM>(a, n[e], kk[e], g[e], c[e], cn[e], cnew[e]);
}

// Flat voxel index -> (x, y, z).
template <typename R>
__device__ inline void adr_coords(const AdrArgs<R> &a, int n, int &x, int &y, int &z) {
  x = fastdiv(n, a.div_yz);
  const int r = n - x * a.div_yz.d;
  y = fastdiv(r, a.div_z);
  z = r - y * a.Z;
}

// Validate a grid (nd = 2: nx x ny with nz = 1, stored as (nx, 1, ny)) and
// set the grid fields of a, zeroing the rest.  Host code.
template <typename R>
inline int adr_grid(int nd, int nx, int ny, int nz, AdrArgs<R> &a) {
  if (nd != 2 && nd != 3) {
    set_error("prt_adr_step: nd must be 2 or 3 (got %d)", nd);
    return PRT_E_ARG;
  }
  if (nx < 1 || ny < 1 || nz < 1 || (nd == 2 && nz != 1)) {
    set_error("prt_adr_step: bad grid (%d, %d, %d) for nd = %d", nx, ny, nz, nd);
    return PRT_E_ARG;
  }
  std::memset(&a, 0, sizeof(a));
  if (nd == 2) {
    a.X = nx; a.Y = 1; a.Z = ny;
  } else {
    a.X = nx; a.Y = ny; a.Z = nz;
  }
  a.N = (long long)nx * ny * nz;
  if (a.N >= (1LL << 31) || a.N * (long long)sizeof(R) >= (1LL << 32) || a.X > 65535) {
    set_error("prt_adr_step: grid too large for the 32-bit voxel index");
    return PRT_E_RANGE;
  }
  a.div_yz = make_fastdiv(a.Y * a.Z);
  a.mad24 = (long long)a.Y * a.Z * (long long)sizeof(R) < (1LL << 23);
  a.div_z = make_fastdiv(a.Z);
  return PRT_OK;
}

// Validate a parameter block and grid and convert them to kernel arguments
// (pointers are set by the caller).  Host code.
template <typename R>
inline int adr_args_from_params(const prt_adr_params *p, int nx, int ny, int nz, AdrArgs<R> &a, dim3 &grid,
                                dim3 &block) {
  if (!p) {
    set_error("prt_adr_step: null parameter block");
    return PRT_E_ARG;
  }

  if (p->ns < 1 || p->ns > 8) {
    set_error("prt_adr_step: 1..8 chemicals supported (got %d)", p->ns);
    return PRT_E_ARG;
  }

  if (!(p->dt > 0)) {
    set_error("prt_adr_step: dt must be positive");
    return PRT_E_ARG;
  }
  auto idx_ok = [&](int j) { return j >= -1 && j < p->ns; };
  if (!idx_ok(p->abio_index) || !idx_ok(p->abio_target) || !idx_ok(p->abio_product) ||
      (p->sorption && (p->sorb_c < 0 || p->sorb_c >= p->ns || p->sorb_s < 0 || p->sorb_s >= p->ns)) ||
      (p->coupled && p->ns < 4)) {
    set_error("prt_adr_step: reaction species index out of range");
    return PRT_E_ARG;
  }
  const int grc = adr_grid<R>(p->nd, nx, ny, nz, a);
  if (grc) return grc;
  if (a.N * p->ns >= (1LL << 31)) {
    set_error("prt_adr_step: grid too large for the 32-bit voxel index");
    return PRT_E_RANGE;
  }
  a.ns = p->ns; a.biotic_on = p->biotic_on; a.coupled = p->coupled;
  a.abio_index = p->abio_index; a.abio_target = p->abio_target; a.abio_product = p->abio_product;
  a.surface_only = p->surface_only; a.sorption = p->sorption; a.sorb_c = p->sorb_c; a.sorb_s = p->sorb_s;
  a.write_solid = p->write_solid;
  a.imm_list = p->immobile_list;
  a.dt = (R)p->dt; a.k = (R)p->k; a.ka = (R)p->ka; a.kd = (R)p->kd;
  a.ks_d = (R)p->ks_donor; a.ks_a = (R)p->ks_acceptor; a.alpha_dt = (R)p->alpha_dt;
  a.neg_alpha = (R)p->neg_alpha; a.c2 = (R)p->c2; a.yield_ = (R)p->yield_; a.c3_init = (R)p->c3_init;
  a.k_sorb = (R)p->k_sorb; a.k_desorb = (R)p->k_desorb;
  a.alpha_pos = p->alpha_pos > 0;
  for (int j = 0; j < 8; ++j) {
    a.mobile[j] = p->mobile[j];
    a.pinned[j] = p->pinned[j];
    a.fed[j] = (R)p->fed[j];
    for (int q = 0; q < 4; ++q) {
      // the step multiplies a closed face's zero difference by its
      // coefficient, which stays zero only for a finite one
      if (!(p->dface[j][q] >= 0) || !std::isfinite((R)p->dface[j][q])) {
        set_error("prt_adr_step: face diffusion coefficients must be finite and non-negative");
        return PRT_E_ARG;
      }
      a.dface[j][q] = (R)p->dface[j][q];
    }
  }
  int bz = 64;
  while (bz < a.Z && bz < 256) bz <<= 1;
  const int by = 256 / bz;
  block = dim3(bz, by, 1);
  grid = dim3((a.Z + bz - 1) / bz, (a.Y + by - 1) / by, a.X);
  return PRT_OK;
}

}  // namespace prt
