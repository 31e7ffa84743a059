/* stokes_c.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the oracle's FP64 Stokes iteration
 * (oracle/prtlb_oracle.py:stokes), which itself restates the reference's
 * stokes_d2q9 / stokes_d3q19 (3D/tools/prtlb_2d.py:348-486,
 * 3D/tools/prtlb_3d.py:348-456).  It exists so that field-by-field checks of
 * the HIP solver can run at full workload sizes (a 128^3 D3Q19 box), where
 * the vectorised NumPy form needs seconds per iteration.  It is checked
 * against the NumPy oracle and the reference's golden vectors in
 * tests/test_oracle_c.py; nothing in the product links it.
 *
 * Evaluation order per node (3D/tools/prtlb_3d.py:427-445), FP64, built
 * with -ffp-contract=off:
 *   rho = sum_i f_i (f_0 first, then f_1, ...)
 *   j_a = sum_i e_ia f_i
 *   u_x = (j_x + g/2) / rho, u_y = j_y / rho, u_z = j_z / rho; 0 in solid
 *   usq = (u_x^2 + u_y^2) + u_z^2
 *   eu_i = (e_ix u_x + e_iy u_y) + e_iz u_z
 *   feq_i = (w_i rho) ((1 + 3 eu) + (4.5 eu) eu - 1.5 usq)
 *   fp_i = (f_i - w+ (f+_i - feq+_i)) - w- (f-_i - feq-_i),  f+- = (f_i +- f_opp)/2
 *   fp_i += ((w_i (1 - w-/2)) 3 g) e_ix
 *   solid: fp_i = f_opp(i) (the arriving populations reversed)
 *   stream: f_i(x) = fp_i(x - e_i), periodic on every axis
 *
 * Layout: population i of node n = (x*Y + y)*Z + z at f[i*N + n]; a 2D grid
 * (nx, ny) is the 3D grid (nx, 1, ny) with D2Q9's (e_x, e_y) as (e_x, 0, e_y).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const int D3_E[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                                {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
                                {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                                {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
static const int D3_OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
/* D2Q9 (3D/tools/prtlb_2d.py:391-394) as (e_x, 0, e_y) */
static const int D2_E[9][3] = {{0, 0, 0}, {1, 0, 0},  {0, 0, 1},   {-1, 0, 0}, {0, 0, -1},
                               {1, 0, 1}, {-1, 0, 1}, {-1, 0, -1}, {1, 0, -1}};
static const int D2_OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

typedef struct {
  int q;
  int e[19][3];
  int opp[19];
  double w[19];
} lattice_t;

static void make_lattice(int q, lattice_t *L) {
  L->q = q;
  for (int i = 0; i < q; ++i) {
    for (int a = 0; a < 3; ++a) L->e[i][a] = q == 19 ? D3_E[i][a] : D2_E[i][a];
    L->opp[i] = q == 19 ? D3_OPP[i] : D2_OPP[i];
    const int m = abs(L->e[i][0]) + abs(L->e[i][1]) + abs(L->e[i][2]);
    if (q == 19)
      L->w[i] = m == 0 ? 1.0 / 3.0 : (m == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
    else
      L->w[i] = m == 0 ? 4.0 / 9.0 : (m == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
  }
}

/* Moments of one node: rho and the masked velocity. */
static void moments(const lattice_t *L, const double *f, long long N, long long n, int solid,
                    double half_g, double *rho_o, double u[3]) {
  const int q = L->q;
  double rho = f[n];
  for (int i = 1; i < q; ++i) rho += f[(long long)i * N + n];
  double j[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < q; ++i) {
    const double fi = f[(long long)i * N + n];
    for (int a = 0; a < 3; ++a)
      if (L->e[i][a]) j[a] += (double)L->e[i][a] * fi;
  }
  u[0] = (j[0] + half_g) / rho;
  u[1] = j[1] / rho;
  u[2] = j[2] / rho;
  if (solid) u[0] = u[1] = u[2] = 0.0;
  *rho_o = rho;
}

/* Number of populations in one state: q * X * Y * Z doubles. */
long long oracle_stokes_elems(int q, int X, int Y, int Z) { return (long long)q * X * Y * Z; }

/* Fill f with the rest state f_i = w_i at every node (solid included). */
int oracle_stokes_init(int q, int X, int Y, int Z, double *f) {
  if ((q != 9 && q != 19) || X < 1 || Y < 1 || Z < 1 || !f) return -1;
  lattice_t L;
  make_lattice(q, &L);
  const long long N = (long long)X * Y * Z;
  for (int i = 0; i < q; ++i)
    for (long long n = 0; n < N; ++n) f[(long long)i * N + n] = L.w[i];
  return 0;
}

/* Advance `steps` iterations in place (f holds q*N doubles; tmp as many).
 * sum_abs_ux (may be NULL) receives, for the last iteration, the sum of
 * |u_x| over fluid nodes of its pre-collision moments (summed per x-plane in
 * node order, then the planes in order).  Returns 0, or -1 on bad
 * arguments. */
int oracle_stokes_run(int q, int X, int Y, int Z, const uint8_t *solid, double *f, double *tmp,
                      int steps, double om_p, double om_m, double g, double *sum_abs_ux) {
  if ((q != 9 && q != 19) || X < 1 || Y < 1 || Z < 1 || !solid || !f || !tmp || steps < 0) return -1;
  lattice_t L;
  make_lattice(q, &L);
  const long long N = (long long)X * Y * Z;
  const double half_g = 0.5 * g;
  double guo[19];
  for (int i = 0; i < q; ++i) {
    double t = L.w[i] * (1.0 - 0.5 * om_m);
    t = t * 3.0;
    t = t * g;
    guo[i] = t * (double)L.e[i][0];
  }
  double s_ux = 0.0;
  double *plane = (double *)calloc((size_t)X, sizeof(double));
  if (!plane) return -1;
  const long long YZ = (long long)Y * Z;
  for (int it = 0; it < steps; ++it) {
    const int last = it == steps - 1;
    /* collide every node into tmp; |u_x| summed per x-plane in node order,
       then the planes in order, so the result does not depend on threads */
#pragma omp parallel for schedule(static) if (N >= (1LL << 18))
    for (int x = 0; x < X; ++x) {
      double acc = 0.0;
      for (long long n = x * YZ; n < (x + 1) * YZ; ++n) {
        const int sd = solid[n] != 0;
        double rho, u[3];
        moments(&L, f, N, n, sd, half_g, &rho, u);
        if (last && !sd) acc += u[0] < 0 ? -u[0] : u[0];
        const double usq = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
        double fl[19], feq[19];
        for (int i = 0; i < q; ++i) {
          fl[i] = f[(long long)i * N + n];
          const double eu = ((double)L.e[i][0] * u[0] + (double)L.e[i][1] * u[1]) + (double)L.e[i][2] * u[2];
          feq[i] = (L.w[i] * rho) * (((1.0 + 3.0 * eu) + (4.5 * eu) * eu) - 1.5 * usq);
        }
        for (int i = 0; i < q; ++i) {
          const int o = L.opp[i];
          double v;
          if (sd) {
            v = fl[o];
          } else {
            const double fp_ = 0.5 * (fl[i] + fl[o]);
            const double fm_ = 0.5 * (fl[i] - fl[o]);
            const double ep_ = 0.5 * (feq[i] + feq[o]);
            const double em_ = 0.5 * (feq[i] - feq[o]);
            v = (fl[i] - om_p * (fp_ - ep_)) - om_m * (fm_ - em_);
            v = v + guo[i];
          }
          tmp[(long long)i * N + n] = v;
        }
      }
      plane[x] = acc;
    }
    if (last) {
      s_ux = 0.0;
      for (int x = 0; x < X; ++x) s_ux += plane[x];
    }
    /* pull-stream tmp into f, periodic */
#pragma omp parallel for schedule(static) if (N >= (1LL << 18))
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          const long long n = ((long long)x * Y + y) * Z + z;
          for (int i = 0; i < q; ++i) {
            int sx = x - L.e[i][0], sy = y - L.e[i][1], sz = z - L.e[i][2];
            sx = sx < 0 ? sx + X : (sx >= X ? sx - X : sx);
            sy = sy < 0 ? sy + Y : (sy >= Y ? sy - Y : sy);
            sz = sz < 0 ? sz + Z : (sz >= Z ? sz - Z : sz);
            f[(long long)i * N + n] = tmp[(long long)i * N + ((long long)sx * Y + sy) * Z + sz];
          }
        }
  }
  free(plane);
  if (sum_abs_ux) *sum_abs_ux = s_ux;
  return 0;
}

/* The velocity of the current state as the reference returns it before its
 * float32 cast: u from the moments (u_x with g/2), zero in solid; v holds
 * 3*N doubles (component-major).  sum_ux (may be NULL) receives the sum of
 * u_x over every node, in node order. */
int oracle_stokes_velocity(int q, int X, int Y, int Z, const uint8_t *solid, const double *f, double g,
                           double *v, double *sum_ux) {
  if ((q != 9 && q != 19) || X < 1 || Y < 1 || Z < 1 || !solid || !f || !v) return -1;
  lattice_t L;
  make_lattice(q, &L);
  const long long N = (long long)X * Y * Z;
  double s = 0.0;
  for (long long n = 0; n < N; ++n) {
    double rho, u[3];
    moments(&L, f, N, n, solid[n] != 0, 0.5 * g, &rho, u);
    for (int a = 0; a < 3; ++a) v[(long long)a * N + n] = u[a];
    s += u[0];
  }
  if (sum_ux) *sum_ux = s;
  return 0;
}
