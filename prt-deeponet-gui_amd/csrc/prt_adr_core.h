// prt_adr_core.h -- one voxel of the fused finite-volume transport step,
// shared by the single-geometry kernel (prt_adr.hip) and the batched
// multi-geometry kernel (prt_batch.hip), so both evaluate identical
// arithmetic.  See prt_adr.hip for the scheme and its float32 contract.
#pragma once

#include <hip/hip_runtime.h>

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
  R dt, k, ka, kd, ks_d, ks_a, alpha_dt, neg_alpha, c2, yield_, c3_init, k_sorb, k_desorb;
  int alpha_pos;
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
template <typename R>
__device__ inline R npsign(R d) {
  return d > R(0) ? R(1) : (d < R(0) ? R(-1) : d);
}

// C_face = up + 1/2 minmod(down - up, up - upup), 3D/tools/prtlb_3d.py:700-705
template <typename R>
__device__ inline R muscl(R up, R upup, R down) {
#pragma clang fp contract(off)
  const R d1 = down - up;
  const R d2 = up - upup;
  const R slope = (R(0.5) * (npsign(d1) + npsign(d2))) * npmin(fabs(d1), fabs(d2));
  return up + R(0.5) * slope;
}

// One voxel of the step.  Returns false (and writes nothing unless
// write_solid) outside the pore; otherwise writes the new values to cout and
// also returns them in cnew.
template <typename R, int ND, int NSM>
__device__ inline bool adr_voxel(const AdrArgs<R> &a, int x, int y, int z, R *cnew) {
#pragma clang fp contract(off)
  const int n = (x * a.Y + y) * a.Z + z;
  const uint8_t fl = a.flags[n];
  const int ns = a.ns;
  if (!(fl & 1)) {
    if (a.write_solid)
#pragma unroll
      for (int j = 0; j < NSM; ++j)
        if (j < ns) a.cout[(size_t)j * a.N + n] = R(0);
    return false;
  }
  const int bf_self = (fl >> 1) & 1;

  // Every neighbour value the stencil may need is loaded up front and
  // unconditionally (indices are clamped, so always in bounds); the face
  // rules then pick among loaded values.  Loads that depended on neighbour
  // flags would chain several memory round trips per axis.
  int nidx[ND][4], nx0[ND][4];           // [axis][m2, m1, p1, p2]
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) {
    // axis geometry: 3D axes x, y, z; a 2D grid is (X, 1, Z) with axes x, z
    const int ax3 = (ND == 2 && ax == 1) ? 2 : ax;
    const int stride = ax3 == 0 ? a.Y * a.Z : (ax3 == 1 ? a.Z : 1);
    const int ext = ax3 == 0 ? a.X : (ax3 == 1 ? a.Y : a.Z);
    const int p = ax3 == 0 ? x : (ax3 == 1 ? y : z);
    const int pm1 = p > 0 ? p - 1 : 0, pm2 = p > 1 ? p - 2 : 0;
    const int pp1 = p < ext - 1 ? p + 1 : ext - 1, pp2 = p < ext - 2 ? p + 2 : ext - 1;
    nidx[ax][0] = n + (pm2 - p) * stride;
    nidx[ax][1] = n + (pm1 - p) * stride;
    nidx[ax][2] = n + (pp1 - p) * stride;
    nidx[ax][3] = n + (pp2 - p) * stride;
    nx0[ax][0] = ax3 == 0 ? pm2 : x;
    nx0[ax][1] = ax3 == 0 ? pm1 : x;
    nx0[ax][2] = ax3 == 0 ? pp1 : x;
    nx0[ax][3] = ax3 == 0 ? pp2 : x;
  }
  uint8_t nf[ND][4];
  R um[ND], uc[ND], up[ND];
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) {
#pragma unroll
    for (int k = 0; k < 4; ++k) nf[ax][k] = a.flags[nidx[ax][k]];
    const R *ua = a.u + (size_t)ax * a.N;
    um[ax] = ua[nidx[ax][1]];
    uc[ax] = ua[n];
    up[ax] = ua[nidx[ax][2]];
  }
  // the inlet pin (3D/tools/prtlb_3d.py:1398-1400) applied as the value is
  // read: a voxel on x = 0 that is read at all is a pore voxel
  R c[NSM], tr[NSM], cn[NSM][ND][4];
#pragma unroll
  for (int j = 0; j < NSM; ++j) {
    tr[j] = R(0);
    c[j] = R(0);
    if (j >= ns) continue;
    const R *cj = a.cin + (size_t)j * a.N;
    const R v = cj[n];
    c[j] = (x == 0 && a.pinned[j]) ? a.fed[j] : v;
    if (!a.mobile[j]) continue;
#pragma unroll
    for (int ax = 0; ax < ND; ++ax)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const R w = cj[nidx[ax][k]];
        cn[j][ax][k] = (nx0[ax][k] == 0 && a.pinned[j]) ? a.fed[j] : w;
      }
  }

  R divu = R(0);
#pragma unroll
  for (int ax = 0; ax < ND; ++ax) {
    const uint8_t fm1 = nf[ax][1], fp1 = nf[ax][2];
    const bool o_lo = fm1 & 1, o_hi = fp1 & 1;
    const bool f_lo = o_lo && (nf[ax][0] & 1);
    const bool f_hi = o_hi && (nf[ax][3] & 1);
    const R u_lo = o_lo ? R(0.5) * (uc[ax] + um[ax]) : R(0);
    const R u_hi = o_hi ? R(0.5) * (uc[ax] + up[ax]) : R(0);
    divu = (ax == 0) ? R(0) + (u_hi - u_lo) : divu + (u_hi - u_lo);
    const int bm1 = (fm1 >> 1) & 1, bp1 = (fp1 >> 1) & 1;
#pragma unroll
    for (int j = 0; j < NSM; ++j) {
      if (j >= ns || !a.mobile[j]) continue;
      const R C = c[j];
      const R Cm1 = o_lo ? cn[j][ax][1] : C;
      const R Cp1 = o_hi ? cn[j][ax][2] : C;
      const R Cm2 = f_lo ? cn[j][ax][0] : Cm1;
      const R Cp2 = f_hi ? cn[j][ax][3] : Cp1;
      // 2x2 table picked with selects: a runtime index into the kernel
      // argument block would copy the whole block to scratch memory
      const R t0 = bf_self ? a.dface[j][2] : a.dface[j][0];
      const R t1 = bf_self ? a.dface[j][3] : a.dface[j][1];
      const R d_lo = o_lo ? (bm1 ? t1 : t0) : R(0);
      const R d_hi = o_hi ? (bp1 ? t1 : t0) : R(0);
      const R dif = d_hi * (Cp1 - C) + d_lo * (Cm1 - C);
      tr[j] = (ax == 0) ? R(0) + dif : tr[j] + dif;
      const R fh = u_hi >= R(0) ? muscl(C, Cm1, Cp1) : muscl(Cp1, Cp2, C);
      const R flo = u_lo >= R(0) ? muscl(Cm1, Cm2, C) : muscl(C, Cp1, Cm1);
      tr[j] = tr[j] - (u_hi * fh - u_lo * flo);
    }
  }
#pragma unroll
  for (int j = 0; j < NSM; ++j)
    if (j < ns && a.mobile[j]) tr[j] = tr[j] + c[j] * divu;

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
      // closed faces over 2*nd, edge-replicated neighbours (:1000-1018)
      int closed = 0;
#pragma unroll
      for (int ax = 0; ax < ND; ++ax) closed += !(nf[ax][1] & 1) + !(nf[ax][2] & 1);
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
    if (a.surface_only) {
      int closed = 0;
#pragma unroll
      for (int ax = 0; ax < ND; ++ax) closed += !(nf[ax][1] & 1) + !(nf[ax][2] & 1);
      surf = R(closed) / R(2 * ND);
    }
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
    a.cout[(size_t)j * a.N + n] = v;
    cnew[j] = v;
  }
  return true;
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
  if (p->nd != 2 && p->nd != 3) {
    set_error("prt_adr_step: nd must be 2 or 3 (got %d)", p->nd);
    return PRT_E_ARG;
  }
  if (p->ns < 1 || p->ns > 8) {
    set_error("prt_adr_step: 1..8 chemicals supported (got %d)", p->ns);
    return PRT_E_ARG;
  }
  if (nx < 1 || ny < 1 || nz < 1 || (p->nd == 2 && nz != 1)) {
    set_error("prt_adr_step: bad grid (%d, %d, %d) for nd = %d", nx, ny, nz, p->nd);
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
  std::memset(&a, 0, sizeof(a));
  if (p->nd == 2) {
    a.X = nx; a.Y = 1; a.Z = ny;
  } else {
    a.X = nx; a.Y = ny; a.Z = nz;
  }
  a.N = (long long)nx * ny * nz;
  if (a.N * p->ns >= (1LL << 31) || a.X > 65535) {
    set_error("prt_adr_step: grid too large for the 32-bit voxel index");
    return PRT_E_RANGE;
  }
  a.ns = p->ns; a.biotic_on = p->biotic_on; a.coupled = p->coupled;
  a.abio_index = p->abio_index; a.abio_target = p->abio_target; a.abio_product = p->abio_product;
  a.surface_only = p->surface_only; a.sorption = p->sorption; a.sorb_c = p->sorb_c; a.sorb_s = p->sorb_s;
  a.write_solid = p->write_solid;
  a.dt = (R)p->dt; a.k = (R)p->k; a.ka = (R)p->ka; a.kd = (R)p->kd;
  a.ks_d = (R)p->ks_donor; a.ks_a = (R)p->ks_acceptor; a.alpha_dt = (R)p->alpha_dt;
  a.neg_alpha = (R)p->neg_alpha; a.c2 = (R)p->c2; a.yield_ = (R)p->yield_; a.c3_init = (R)p->c3_init;
  a.k_sorb = (R)p->k_sorb; a.k_desorb = (R)p->k_desorb;
  a.alpha_pos = p->alpha_pos > 0;
  for (int j = 0; j < 8; ++j) {
    a.mobile[j] = p->mobile[j];
    a.pinned[j] = p->pinned[j];
    a.fed[j] = (R)p->fed[j];
    for (int q = 0; q < 4; ++q) a.dface[j][q] = (R)p->dface[j][q];
  }
  int bz = 64;
  while (bz < a.Z && bz < 256) bz <<= 1;
  const int by = 256 / bz;
  block = dim3(bz, by, 1);
  grid = dim3((a.Z + bz - 1) / bz, (a.Y + by - 1) / by, a.X);
  return PRT_OK;
}

}  // namespace prt
