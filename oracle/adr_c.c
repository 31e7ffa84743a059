/* adr_c.c -- the oracle's float32 transport step in C.  TEST INFRASTRUCTURE
 * ONLY: it checks the HIP transport kernels over the long runs of the
 * benchmark workloads, where the vectorised NumPy oracle would take hours.
 *
 * One call is one oracle_adr_step: oracle/prtlb_oracle.py adr_step (pin the
 * fed inlet in place, transport the mobile rows, C + dt (T + R), clip at
 * zero in the pore, zero outside), with adr_transport and adr_reaction
 * written out per voxel in NumPy's float32 evaluation order (the reference's
 * 3D/tools/prtlb_3d.py:1182-1208, :1224-1367, :1392-1413).  It reads the
 * oracle's own set-up arrays (adr_setup: the clamped neighbours as the face
 * masks select them, face velocities and diffusivities, div u, surface
 * weights) taken over the pore voxels, so the set-up stays the NumPy
 * oracle's.
 *
 * NumPy semantics kept: every operation in float32 with no contraction
 * (-ffp-contract=off, no fast math), IEEE division, np.sign giving +0 for a
 * zero and NaN for NaN, np.minimum / np.maximum propagating NaN and giving
 * the second argument on ties (as NumPy's float32 loops do), Python-float
 * scalars taken as float32 (NEP 50; the caller passes them rounded).
 */
#include <stdint.h>
#include <string.h>

/* np.sign: +-1, +0 for a zero, the NaN itself for NaN (the comparisons
 * are branch-free; NaN is the rare case) */
static inline float np_sign(float d) {
  const float s = (float)((d > 0.0f) - (d < 0.0f));
  return d != d ? d : s;
}
static inline float np_min(float a, float b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? a : b;
}
static inline float np_max(float a, float b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}
static inline float np_abs(float a) {
  union { float f; uint32_t u; } v = {a};
  v.u &= 0x7fffffffu;
  return v.f;
}
/* prtlb_oracle.muscl */
static inline float muscl(float up, float upup, float down) {
  const float d1 = down - up;
  const float d2 = up - upup;
  const float slope = (0.5f * (np_sign(d1) + np_sign(d2))) * np_min(np_abs(d1), np_abs(d2));
  return up + 0.5f * slope;
}

/* Reaction constants, each already the float32 value the oracle's
 * expression forms (Python floats rounded as NEP 50 rounds them). */
typedef struct {
  int biotic_on;   /* ch.biotic_on */
  int has_ksd;     /* ch.ks_donor > 0 */
  int has_ksa;     /* ch.ks_acceptor > 0 */
  int coupled;     /* ch.biomass_coupled and ns > 3 and c_init[3] > 0 */
  int alpha_pos;   /* ch.alpha > 0 */
  int abio_j;      /* abiotic_index (may be >= ns) */
  int abio_on;     /* ka > 0 and ns > j */
  int abio_surf;   /* abiotic_surface_only */
  int abio_mk;     /* product index, -1 for none */
  int sorption, sorb_c, sorb_s;
  float k, ka, kd, dt, ks_d, ks_a, c3_init, alpha_dt, neg_alpha, c2, yield_, k_sorb, k_des;
} adr_rx;

/* c has ns rows of N voxels (the previous state; the inlet pin is written
 * into it, as the oracle pins in place); c_out receives the new state at
 * the P pore voxels (its other voxels are left as they are: the caller
 * keeps them zero, as np.where(mf, ..., 0) makes them).  rec holds one
 * record of W = 1 + nd (6 + 2 nmov) + 3 four-byte words per pore voxel:
 *   [0]                      the voxel's flat index (int32);
 *   per axis a, at 1 + a (6 + 2 nmov):
 *     4 int32: the voxels whose values the oracle's selects take for Cm2,
 *              Cm1, Cp1, Cp2 (np.where(o_lo, C[m1], C) is C[o_lo ? m1 : n],
 *              and so on);
 *     u_lo, u_hi, then d_lo and d_hi of each transported row (float32);
 *   then div u, the surface weight, the sorption surface weight (float32).
 * mov: the nmov transported rows, in order. */
void oracle_adr_step(int64_t N, int64_t P, int ns, int nd, int nmov, const int32_t *mov,
                     const uint32_t *rec, const int64_t *inlet, int64_t n_inlet, const float *fed,
                     const int32_t *pin, const adr_rx *rx, float *c, float *c_out) {
  const int W = 1 + nd * (6 + 2 * nmov) + 3;
  /* the inlet pin, in place (prtlb_oracle.adr_step) */
  for (int q = 0; q < nmov; ++q) {
    const int j = mov[q];
    if (!pin[j]) continue;
    float *cj = c + (int64_t)j * N;
    for (int64_t i = 0; i < n_inlet; ++i) cj[inlet[i]] = fed[j];
  }
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; ++p) {
    const uint32_t *r0 = rec + p * (int64_t)W;
    const int64_t n = (int32_t)r0[0];
    const float *tail = (const float *)(r0 + 1 + nd * (6 + 2 * nmov));
    float tr[8];
    for (int j = 0; j < ns; ++j) tr[j] = 0.0f;
    /* adr_transport over the mobile rows */
    for (int q = 0; q < nmov; ++q) {
      const int j = mov[q];
      const float *cj = c + (int64_t)j * N;
      const float C = cj[n];
      float out = 0.0f;
      for (int ax = 0; ax < nd; ++ax) {
        const uint32_t *ra = r0 + 1 + ax * (6 + 2 * nmov);
        const float *fa = (const float *)(ra + 4);
        const float Cm2 = cj[(int32_t)ra[0]], Cm1 = cj[(int32_t)ra[1]];
        const float Cp1 = cj[(int32_t)ra[2]], Cp2 = cj[(int32_t)ra[3]];
        const float ul = fa[0], uh = fa[1], dl = fa[2 + 2 * q], dh = fa[3 + 2 * q];
        const float dif = dh * (Cp1 - C) + dl * (Cm1 - C);
        out = out + dif;
        /* np.where(u >= 0, muscl(a...), muscl(b...)) takes one of two
         * evaluations; the upwind side's operands are picked first and
         * the one evaluation made */
        const int ph = uh >= 0.0f, pl = ul >= 0.0f;
        const float fh = muscl(ph ? C : Cp1, ph ? Cm1 : Cp2, ph ? Cp1 : C);
        const float fl = muscl(pl ? Cm1 : C, pl ? Cm2 : Cp1, pl ? C : Cm1);
        out = out - (uh * fh - ul * fl);
      }
      out = out + C * tail[0];
      tr[j] = out;
    }
    /* adr_reaction */
    float cv[8], r[8];
    for (int j = 0; j < ns; ++j) {
      cv[j] = c[(int64_t)j * N + n];
      r[j] = 0.0f;
    }
    const float dt = rx->dt;
    float R = 0.0f;
    if (rx->biotic_on) {
      R = rx->has_ksd ? (rx->k * cv[0]) / (rx->ks_d + cv[0]) : rx->k * cv[0];
      if (ns > 1) R = R * (rx->has_ksa ? cv[1] / (rx->ks_a + cv[1]) : cv[1]);
      if (rx->coupled) R = (R * cv[3]) / rx->c3_init;
    }
    float Ra = 0.0f;
    if (rx->abio_on) {
      const float cj = cv[rx->abio_j];
      Ra = rx->abio_surf ? (rx->ka * cj) * tail[1] : rx->ka * cj;
    }
    if (rx->biotic_on) {
      float lim = cv[0] / dt;
      if (ns > 1 && rx->alpha_pos) lim = np_min(lim, cv[1] / rx->alpha_dt);
      R = np_min(R, 0.5f * lim);
    }
    if (rx->abio_on) Ra = np_min(Ra, (0.5f * cv[rx->abio_j]) / dt);
    r[0] = -R;
    if (ns > 1) r[1] = rx->neg_alpha * R;
    if (ns > 2) r[2] = rx->c2 * R;
    if (ns > 3) r[3] = rx->biotic_on ? rx->yield_ * R - rx->kd * cv[3] : 0.0f;
    if (rx->abio_j < ns) {
      r[rx->abio_j] = r[rx->abio_j] - Ra;
      if (rx->abio_mk >= 0) r[rx->abio_mk] = r[rx->abio_mk] + Ra;
    }
    if (rx->sorption) {
      const int cc = rx->sorb_c, ss = rx->sorb_s;
      float fw = (rx->k_sorb * cv[cc]) * tail[2];
      fw = np_min(fw, (0.5f * cv[cc]) / dt);
      float bk = rx->k_des * cv[ss];
      bk = np_min(bk, (0.5f * cv[ss]) / dt);
      const float net = fw - bk;
      r[cc] = r[cc] - net;
      r[ss] = r[ss] + net;
    }
    /* C + dt (T + R), clipped at zero */
    for (int j = 0; j < ns; ++j) {
      const float v = cv[j] + dt * (tr[j] + r[j]);
      c_out[(int64_t)j * N + n] = np_max(v, 0.0f);
    }
  }
}
