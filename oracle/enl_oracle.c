/*
 * enl_oracle.c -- CPU restatement of Enlightning's per-patch update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels
 * in enlightning_amd/csrc: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load the library built from it.  It is
 * never the product path.
 *
 * It restates, in plain C with the same operation order, the same
 * single-precision-folded literals and no FMA contraction (build with
 * -ffp-contract=off), the reference's Fortran kernels:
 *
 *   rhs           fortran/rhs_hybrid_high_llf.m4:28-433   (default build)
 *                 fortran/rhs_hybrid_high.m4:28-415
 *                 fortran/rhs_hybrid_low_llf.m4:28-399
 *                 fortran/rhs_hybrid_low.m4:28-381
 *                 fortran/rhs_weno_high_llf.m4:27-391
 *                 fortran/rhs_weno_high.m4:27-373
 *                 fortran/rhs_weno_low_llf.m4:27-357
 *                 fortran/rhs_weno_low.m4:27-339
 *                 fortran/rhs_drp.m4:26-241
 *   runge_kutta   fortran/runge_kutta.m4:26-69
 *   update_state  fortran/update_state.m4:26-110
 *   tag_cells     fortran/tag_cells.m4:27-112
 *
 * plus the host-side pieces of the patch loop the kernels are called from:
 *
 *   physical BC   enlightning.C:790-1013 (setPhysicalBoundaryConditions)
 *   singleStep    enlightning.C:459-583 (rhs -> runge_kutta -> update_state)
 *   MOL RK step   SAMRAI MethodOfLinesIntegrator stage loop (not vendored;
 *                 inferred from enlightning.C:465-490 contexts)
 *
 * Array layout follows SAMRAI CellData: column-major, component slowest.
 * Element (i,j,k) of a g-ghost array of patch [if0..il0]x[if1..il1] sits at
 * (i-if0+g) + Nx*((j-if1+g) + Ny*k), Nx = il0-if0+1+2g, Ny = il1-if1+1+2g
 * (enlightning.C:663-666, 856).  Interior-only arrays use g = 0.
 *
 * Literals: Fortran default-REAL literals such as 13.0/12.0 or 0.9 are folded
 * in single precision and then widened; the F() macro reproduces that.
 *
 * The Fortran reference itself is not built here (it needs m4 and SAMRAI's
 * pdat_m4arrdim2d.i, neither of which exists in this image); this file is
 * pinned instead by the closed-form, invariant and cross-variant checks in
 * tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "enl_oracle.h"

/* Fortran REAL literal -> double, as flang/gfortran fold it. */
#define F(x) ((double)(float)(x))

/* Every single-precision-folded literal of the Fortran kernels, in one table
 * (SURVEY Appendix A order).  The functions below read their constants from
 * here and orc_literals() exports this same table, so the pin test of the
 * table checks the values the kernels actually use. */
enum {
  L_13_12, L_0_1, L_0_6, L_0_3, L_EP5, L_1_6, L_1_3, L_2_3, L_4_3,
  L_A1, L_A2, L_A3, L_0_78, L_0_9, L_0_9999, L_EP3, L_COUNT
};
static const double LIT[L_COUNT] = {
  F(13.0f / 12.0f), /* wenop/wenom beta, rhs_hybrid_high_llf.m4:375-380 */
  F(0.1f), F(0.6f), F(0.3f), /* WENO5 ideal weights, :382-384, 417-419 */
  F(10.0e-6f),      /* WENO5 epsilon, :373, 408 */
  F(1.0f / 6.0f),   /* WENO5 candidates, :392-394, 427-429 */
  F(1.0f / 3.0f), F(2.0f / 3.0f), /* WENO3 weights (rhs_hybrid_low_llf.m4:360-361,
                                     386-387) and viscous phi (rhs_hybrid_high_llf.m4:174-175) */
  F(4.0f / 3.0f),   /* viscous phi */
  F(0.79926643f), F(-0.18941314f), F(0.02651995f), /* DRP a_1..a_3, :117-119 */
  F(0.78f),         /* c_vn, :217 */
  F(0.9f),          /* sensor / tag epsilon, :115, tag_cells.m4:65 */
  F(0.9999f),       /* sensor / tag threshold, :247, tag_cells.m4:72 */
  F(10e-6f),        /* WENO3 epsilon, rhs_hybrid_low_llf.m4:355, 381 */
};

static double sq(double x) { return x * x; }

/* Strided view of one cell-centred array.  base points at element (i0,j0,0). */
typedef struct {
  double *base;
  int i0, j0;
  long sy, sk;
} view_t;

static inline double at(const view_t *v, int i, int j, int k) {
  return v->base[(long)(i - v->i0) + (long)(j - v->j0) * v->sy + (long)k * v->sk];
}
static inline double *ref(const view_t *v, int i, int j, int k) {
  return &v->base[(long)(i - v->i0) + (long)(j - v->j0) * v->sy + (long)k * v->sk];
}

/* View of a full ghosted patch array (SAMRAI getPointer() layout). */
static view_t patch_view(const double *p, int if0, int il0, int if1, int il1, int g0, int g1) {
  view_t v;
  long nx = il0 - if0 + 1 + 2 * g0, ny = il1 - if1 + 1 + 2 * g1;
  v.base = (double *)p;
  v.i0 = if0 - g0;
  v.j0 = if1 - g1;
  v.sy = nx;
  v.sk = nx * ny;
  return v;
}

/* ------------------------------------------------------------------------ */
/* WENO reconstructions.                                                     */
/* ------------------------------------------------------------------------ */

/* rhs_hybrid_high_llf.m4:365-398 (identical copy in every *_high* file). */
double orc_wenom5(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
  const double ep = LIT[L_EP5];
  double beta0 = LIT[L_13_12] * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  double beta1 = LIT[L_13_12] * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  double beta2 = LIT[L_13_12] * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  double zeta0 = LIT[L_0_1] / sq(ep + beta0);
  double zeta1 = LIT[L_0_6] / sq(ep + beta1);
  double zeta2 = LIT[L_0_3] / sq(ep + beta2);
  double sigma = zeta0 + zeta1 + zeta2;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma, omega2 = zeta2 / sigma;
  double f0 = LIT[L_1_6] * (11 * fg0 - 7 * fgp1 + 2 * fgp2);
  double f1 = LIT[L_1_6] * (2 * fgm1 + 5 * fg0 - fgp1);
  double f2 = LIT[L_1_6] * (-fgm2 + 5 * fgm1 + 2 * fg0);
  return omega0 * f0 + omega1 * f1 + omega2 * f2;
}

/* rhs_hybrid_high_llf.m4:400-433. */
double orc_wenop5(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
  const double ep = LIT[L_EP5];
  double beta0 = LIT[L_13_12] * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  double beta1 = LIT[L_13_12] * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  double beta2 = LIT[L_13_12] * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  double zeta0 = LIT[L_0_3] / sq(ep + beta0);
  double zeta1 = LIT[L_0_6] / sq(ep + beta1);
  double zeta2 = LIT[L_0_1] / sq(ep + beta2);
  double sigma = zeta0 + zeta1 + zeta2;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma, omega2 = zeta2 / sigma;
  double f0 = LIT[L_1_6] * (2 * fg0 + 5 * fgp1 - fgp2);
  double f1 = LIT[L_1_6] * (-fgm1 + 5 * fg0 + 2 * fgp1);
  double f2 = LIT[L_1_6] * (2 * fgm2 - 7 * fgm1 + 11 * fg0);
  return omega0 * f0 + omega1 * f1 + omega2 * f2;
}

/* rhs_hybrid_low_llf.m4:349-373 (ep literal 10e-6 there, 10.0e-6 in
 * rhs_weno_low*.m4:313,339 -- both fold to the same float). */
double orc_wenom3(double fgm1, double fg0, double fgp1) {
  const double ep = LIT[L_EP3];
  double beta0 = sq(fgp1 - fg0);
  double beta1 = sq(fg0 - fgm1);
  double zeta0 = LIT[L_1_3] / sq(ep + beta0);
  double zeta1 = LIT[L_2_3] / sq(ep + beta1);
  double sigma = zeta0 + zeta1;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma;
  double f0 = 0.5 * (3 * fg0 - fgp1);
  double f1 = 0.5 * (fgm1 + fg0);
  return omega0 * f0 + omega1 * f1;
}

/* rhs_hybrid_low_llf.m4:375-399. */
double orc_wenop3(double fgm1, double fg0, double fgp1) {
  const double ep = LIT[L_EP3];
  double beta0 = sq(fgp1 - fg0);
  double beta1 = sq(fg0 - fgm1);
  double zeta0 = LIT[L_2_3] / sq(ep + beta0);
  double zeta1 = LIT[L_1_3] / sq(ep + beta1);
  double sigma = zeta0 + zeta1;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma;
  double f0 = 0.5 * (fg0 + fgp1);
  double f1 = 0.5 * (-fgm1 + 3 * fg0);
  return omega0 * f0 + omega1 * f1;
}

/* Smoothness ratio shared by the hybrid sensor (rhs_hybrid_high_llf.m4:244-254)
 * and tag_cells (tag_cells.m4:67-106). */
static inline double smooth_ratio(double tmp1, double tmp2, double ep) {
  return (2 * fabs(tmp1 * tmp2) + ep) / (sq(tmp1) + sq(tmp2) + ep);
}

/* ep = 0.9*tol*(xi**2)/(1-0.9*(tol**2)) with xi = 1
 * (rhs_hybrid_high_llf.m4:114-115, tag_cells.m4:64-65). */
double orc_sensor_eps(double tol) {
  double xi = 1;
  return LIT[L_0_9] * tol * sq(xi) / (1 - LIT[L_0_9] * sq(tol));
}

/* ------------------------------------------------------------------------ */
/* rhs: one cell.                                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
  int hybrid;   /* level-0 density sensor picks DRP or WENO */
  int weno;     /* WENO branch exists */
  int drp;      /* DRP branch exists */
  int order;    /* 5 or 3 */
  int llf;      /* per-cell LLF alpha, else fixed alpha */
  int reach;    /* flux stencil half width: 3, or 2 for rhs_weno_low* */
  int llf_reach;/* LLF window half width: 3 for *_high_llf, 2 for *_low_llf */
} scheme_t;

static scheme_t scheme_of(int variant) {
  scheme_t s;
  memset(&s, 0, sizeof s);
  switch (variant) {
    case ENL_RHS_HYBRID_HIGH_LLF: s.hybrid = 1; s.weno = 1; s.drp = 1; s.order = 5; s.llf = 1; s.reach = 3; s.llf_reach = 3; break;
    case ENL_RHS_HYBRID_HIGH:     s.hybrid = 1; s.weno = 1; s.drp = 1; s.order = 5; s.llf = 0; s.reach = 3; break;
    case ENL_RHS_HYBRID_LOW_LLF:  s.hybrid = 1; s.weno = 1; s.drp = 1; s.order = 3; s.llf = 1; s.reach = 3; s.llf_reach = 2; break;
    case ENL_RHS_HYBRID_LOW:      s.hybrid = 1; s.weno = 1; s.drp = 1; s.order = 3; s.llf = 0; s.reach = 3; break;
    case ENL_RHS_WENO_HIGH_LLF:   s.weno = 1; s.order = 5; s.llf = 1; s.reach = 3; s.llf_reach = 3; break;
    case ENL_RHS_WENO_HIGH:       s.weno = 1; s.order = 5; s.llf = 0; s.reach = 3; break;
    case ENL_RHS_WENO_LOW_LLF:    s.weno = 1; s.order = 3; s.llf = 1; s.reach = 2; s.llf_reach = 2; break;
    case ENL_RHS_WENO_LOW:        s.weno = 1; s.order = 3; s.llf = 0; s.reach = 2; break;
    case ENL_RHS_DRP:             s.drp = 1; s.order = 0; s.reach = 3; break;
    default: s.order = -1; break;
  }
  return s;
}

int orc_rhs_ghost_need(int variant) {
  scheme_t s = scheme_of(variant);
  if (s.order < 0) return -1;
  return s.reach;
}

typedef struct {
  double alpha, kappa, mu, muB, R, T_star_n, T_star_o, tau_n, tau_o;
  double tol2, c, c_lr;
  int level;
} rhs_par_t;

/* Physical source H for one cell: rhs_hybrid_high_llf.m4:159-237
 * (identical in all nine variants). */
static void cell_source(const view_t *w, const view_t *T, double H0, int i, int j,
                        const double *dxy, const rhs_par_t *P, double H[5]) {
  const double dx = dxy[0], dy = dxy[1];
#define U(a, b) (at(w, a, b, 1) / at(w, a, b, 0))
#define V(a, b) (at(w, a, b, 2) / at(w, a, b, 0))
#define TT(a, b) at(T, a, b, 0)
  H[0] = H0;

  double u_dprime_x = (U(i - 1, j) - 2 * U(i, j) + U(i + 1, j)) / sq(dx);
  double u_dprime_y = (U(i, j - 1) - 2 * U(i, j) + U(i, j + 1)) / sq(dy);
  double v_dprime_xy = (V(i + 1, j + 1) - V(i + 1, j - 1) - V(i - 1, j + 1) + V(i - 1, j - 1)) / (4 * dx * dy);
  double phi_xx_prime_x = LIT[L_4_3] * u_dprime_x - LIT[L_2_3] * v_dprime_xy;
  double phi_xy_prime_y = u_dprime_y + v_dprime_xy;
  H[1] = P->muB * (u_dprime_x + v_dprime_xy) + P->mu * (phi_xx_prime_x + phi_xy_prime_y);

  double v_dprime_y = (V(i, j - 1) - 2 * V(i, j) + V(i, j + 1)) / sq(dy);
  double v_dprime_x = (V(i - 1, j) - 2 * V(i, j) + V(i + 1, j)) / sq(dx);
  double u_dprime_yx = (U(i + 1, j + 1) - U(i + 1, j - 1) - U(i - 1, j + 1) + U(i - 1, j - 1)) / (4 * dx * dy);
  double phi_yx_prime_x = u_dprime_yx + v_dprime_x;
  double phi_yy_prime_y = LIT[L_4_3] * v_dprime_y - LIT[L_2_3] * u_dprime_yx;
  H[2] = P->muB * (v_dprime_y + u_dprime_yx) + P->mu * (phi_yx_prime_x + phi_yy_prime_y);

  double u_prime_x = (U(i + 1, j) - U(i - 1, j)) / (2 * dx);
  double u_prime_y = (U(i, j + 1) - U(i, j - 1)) / (2 * dy);
  double v_prime_y = (V(i, j + 1) - V(i, j - 1)) / (2 * dy);
  double v_prime_x = (V(i + 1, j) - V(i - 1, j)) / (2 * dx);
  double t_prime_x = (TT(i + 1, j) - TT(i - 1, j)) / (2 * dx);
  double t_prime_y = (TT(i, j + 1) - TT(i, j - 1)) / (2 * dy);
  double t_dprime_x = (TT(i - 1, j) - 2 * TT(i, j) + TT(i + 1, j)) / sq(dx);
  double t_dprime_y = (TT(i, j - 1) - 2 * TT(i, j) + TT(i, j + 1)) / sq(dy);
  double w0 = at(w, i, j, 0);
  double temp = TT(i, j);
  double T_n = at(w, i, j, 4) / w0;
  double Mn = (temp - T_n) / P->tau_n;
  double tmp1 = P->T_star_n / (T_n);
  double c_vn = LIT[L_0_78] * P->R * sq(tmp1) * exp(-tmp1);
  double A_n = ((temp / T_n) - 1) * c_vn;
  double phi_xx = LIT[L_4_3] * u_prime_x - LIT[L_2_3] * v_prime_y;
  double phi_xy_yx = u_prime_y + v_prime_x;
  double phi_yy = LIT[L_4_3] * v_prime_y - LIT[L_2_3] * u_prime_x;
  double sigma = (P->muB / temp) * sq(u_prime_x + v_prime_y) +
                 (P->mu / (2 * temp)) * (sq(phi_xx) + 2 * sq(phi_xy_yx) + sq(phi_yy)) +
                 (P->kappa / sq(temp)) * (sq(t_prime_x) + sq(t_prime_y)) +
                 (w0 / temp) * (A_n * Mn);
  H[3] = sigma - (w0 / T_n) * c_vn * Mn +
         (-((P->kappa * sq(t_prime_x)) / sq(temp)) + ((P->kappa * t_dprime_x) / temp) -
          ((P->kappa * sq(t_prime_y)) / sq(temp)) + ((P->kappa * t_dprime_y) / temp));
  H[4] = (w0 / P->tau_n) * (temp - T_n);
#undef U
#undef V
#undef TT
}

/* Inviscid flux F (x) or G (y) of component k at cell (a,b):
 * rhs_hybrid_high_llf.m4:125-156. */
static double flux_x(const view_t *w, const view_t *p, int a, int b, int k) {
  double w0 = at(w, a, b, 0);
  double u = at(w, a, b, 1) / w0;
  switch (k) {
    case 0: return w0 * u;
    case 1: return w0 * sq(u) + at(p, a, b, 0);
    case 2: return w0 * u * (at(w, a, b, 2) / w0);
    case 3: return w0 * u * (at(w, a, b, 3) / w0);
    default: return w0 * u * (at(w, a, b, 4) / w0);
  }
}
static double flux_y(const view_t *w, const view_t *p, int a, int b, int k) {
  double w0 = at(w, a, b, 0);
  double v = at(w, a, b, 2) / w0;
  switch (k) {
    case 0: return w0 * v;
    case 1: return w0 * v * (at(w, a, b, 1) / w0);
    case 2: return w0 * sq(v) + at(p, a, b, 0);
    case 3: return w0 * v * (at(w, a, b, 3) / w0);
    default: return w0 * v * (at(w, a, b, 4) / w0);
  }
}

/* Hybrid switch for one cell: rhs_hybrid_high_llf.m4:239-256.
 * Returns 1 (WENO) or 0 (DRP). */
static int cell_sensor(const view_t *w, int i, int j, double ep, int level) {
  int tag = 1;
  if (level == 0) {
    tag = 0;
    for (int q = -1; q <= 1; q++) {
      double tmp1 = at(w, i + q, j, 0) - at(w, i - 1 + q, j, 0);
      double tmp2 = at(w, i + 1 + q, j, 0) - at(w, i + q, j, 0);
      double rx = smooth_ratio(tmp1, tmp2, ep);
      if (rx < LIT[L_0_9999]) tag = 1;
      tmp1 = at(w, i, j + q, 0) - at(w, i, j - 1 + q, 0);
      tmp2 = at(w, i, j + 1 + q, 0) - at(w, i, j + q, 0);
      double ry = smooth_ratio(tmp1, tmp2, ep);
      if (ry < LIT[L_0_9999]) tag = 1;
    }
  }
  return tag;
}

/* One interior cell of any rhs variant.  Writes KK(i,j,0..4); returns the
 * DRP(0)/WENO(1) decision actually taken. */
static int rhs_cell(const scheme_t *s, const view_t *w, const view_t *p, const view_t *T,
                    const view_t *H0, view_t *KK, int i, int j, int ifirst1,
                    const double *xylo, const double *dxy, const rhs_par_t *P, double ep) {
  const double dx = dxy[0], dy = dxy[1];
  double H[5];
  cell_source(w, T, at(H0, i, j, 0), i, j, dxy, P, H);

  int tag;
  if (s->hybrid) tag = cell_sensor(w, i, j, ep, P->level);
  else tag = s->weno ? 1 : 0;

  if (tag == 0) {
    /* DRP branch: rhs_hybrid_high_llf.m4:258-268, rhs_drp.m4:227-236 */
    const double a_1 = LIT[L_A1], a_2 = LIT[L_A2], a_3 = LIT[L_A3];
    for (int k = 0; k < 5; k++) {
      *ref(KK, i, j, k) =
          (-a_3 * flux_x(w, p, i - 3, j, k) - a_2 * flux_x(w, p, i - 2, j, k) - a_1 * flux_x(w, p, i - 1, j, k) +
           a_1 * flux_x(w, p, i + 1, j, k) + a_2 * flux_x(w, p, i + 2, j, k) + a_3 * flux_x(w, p, i + 3, j, k)) /
              (-dx) +
          (-a_3 * flux_y(w, p, i, j - 3, k) - a_2 * flux_y(w, p, i, j - 2, k) - a_1 * flux_y(w, p, i, j - 1, k) +
           a_1 * flux_y(w, p, i, j + 1, k) + a_2 * flux_y(w, p, i, j + 2, k) + a_3 * flux_y(w, p, i, j + 3, k)) /
              (-dy) +
          H[k];
    }
    return 0;
  }

  /* WENO branch. */
  double af, ag;
  if (s->llf) {
    /* rhs_hybrid_high_llf.m4:270-284 */
    af = 0;
    ag = 0;
    for (int q = -s->llf_reach; q <= s->llf_reach; q++) {
      double height = xylo[1] + dy * ((double)(j - ifirst1) + 0.5);
      double vf = fabs(at(w, i + q, j, 1) / at(w, i + q, j, 0)) + P->c + P->c_lr * height;
      if (vf > af) af = vf;
      height = xylo[1] + dy * ((double)(j - ifirst1 + q) + 0.5);
      double vg = fabs(at(w, i, j + q, 2) / at(w, i, j + q, 0)) + P->c + P->c_lr * height;
      if (vg > ag) ag = vg;
    }
  } else {
    af = P->alpha;
    ag = P->alpha;
  }

  for (int k = 0; k < 5; k++) {
    /* Split fluxes at offsets -3..3; index m+3.  rhs_hybrid_high_llf.m4:286-356 */
    double fpx[7], fmx[7], fpy[7], fmy[7];
    for (int m = -s->reach; m <= s->reach; m++) {
      double Fx = flux_x(w, p, i + m, j, k), Gy = flux_y(w, p, i, j + m, k);
      fpx[m + 3] = 0.5 * (Fx + af * at(w, i + m, j, k));
      fmx[m + 3] = 0.5 * (Fx - af * at(w, i + m, j, k));
      fpy[m + 3] = 0.5 * (Gy + ag * at(w, i, j + m, k));
      fmy[m + 3] = 0.5 * (Gy - ag * at(w, i, j + m, k));
    }
    double pf1, mf1, pf2, mf2, pg1, mg1, pg2, mg2;
    if (s->order == 5) {
      pf1 = orc_wenop5(fpx[0], fpx[1], fpx[2], fpx[3], fpx[4]);
      mf1 = orc_wenom5(fmx[1], fmx[2], fmx[3], fmx[4], fmx[5]);
      pf2 = orc_wenop5(fpx[1], fpx[2], fpx[3], fpx[4], fpx[5]);
      mf2 = orc_wenom5(fmx[2], fmx[3], fmx[4], fmx[5], fmx[6]);
      pg1 = orc_wenop5(fpy[0], fpy[1], fpy[2], fpy[3], fpy[4]);
      mg1 = orc_wenom5(fmy[1], fmy[2], fmy[3], fmy[4], fmy[5]);
      pg2 = orc_wenop5(fpy[1], fpy[2], fpy[3], fpy[4], fpy[5]);
      mg2 = orc_wenom5(fmy[2], fmy[3], fmy[4], fmy[5], fmy[6]);
    } else {
      /* rhs_hybrid_low_llf.m4:286-338 */
      pf1 = orc_wenop3(fpx[1], fpx[2], fpx[3]);
      mf1 = orc_wenom3(fmx[2], fmx[3], fmx[4]);
      pf2 = orc_wenop3(fpx[2], fpx[3], fpx[4]);
      mf2 = orc_wenom3(fmx[3], fmx[4], fmx[5]);
      pg1 = orc_wenop3(fpy[1], fpy[2], fpy[3]);
      mg1 = orc_wenom3(fmy[2], fmy[3], fmy[4]);
      pg2 = orc_wenop3(fpy[2], fpy[3], fpy[4]);
      mg2 = orc_wenom3(fmy[3], fmy[4], fmy[5]);
    }
    *ref(KK, i, j, k) = (pf1 + mf1 - pf2 - mf2) / dx + (pg1 + mg1 - pg2 - mg2) / dy + H[k];
  }
  return 1;
}

/* rhs over the box [b0..e0]x[b1..e1] of views.  ifirst1 / xylo are those of
 * the patch (or tile) the views describe. */
static void rhs_box(const scheme_t *s, const view_t *w, const view_t *p, const view_t *T, const view_t *H0,
                    view_t *KK, int b0, int e0, int b1, int e1, int ifirst1, const double *xylo,
                    const double *dxy, const rhs_par_t *P, int *switch_map, long sm_stride) {
  double ep = orc_sensor_eps(P->tol2);
  for (int j = b1; j <= e1; j++)
    for (int i = b0; i <= e0; i++) {
      int t = rhs_cell(s, w, p, T, H0, KK, i, j, ifirst1, xylo, dxy, P, ep);
      if (switch_map) switch_map[(long)(i - b0) + (long)(j - b1) * sm_stride] = t;
    }
}

int orc_rhs_variant(int variant, const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                    const int *gcw0, const int *gcw1, const double *xylo, const double *dxy,
                    const double *w_n, const double *H0, double *KK, const double *pres, const double *temp,
                    const double *alpha, const double *kappa, const double *mu, const double *muB,
                    const double *R, const double *T_star_n, const double *T_star_o, const double *tau_n,
                    const double *tau_o, const double *tolerance, const int *level, const double *c,
                    const double *c_lr, const int *nequ, int *switch_map) {
  scheme_t s = scheme_of(variant);
  if (s.order < 0 || *nequ != 5) return -1;
  if (*gcw0 < s.reach || *gcw1 < s.reach) return -2;
  int if0 = *ifirst0, il0 = *ilast0, if1 = *ifirst1, il1 = *ilast1;
  view_t w = patch_view(w_n, if0, il0, if1, il1, *gcw0, *gcw1);
  view_t p = patch_view(pres, if0, il0, if1, il1, *gcw0, *gcw1);
  view_t T = patch_view(temp, if0, il0, if1, il1, *gcw0, *gcw1);
  view_t h = patch_view(H0, if0, il0, if1, il1, 0, 0);
  view_t K = patch_view(KK, if0, il0, if1, il1, 0, 0);
  rhs_par_t P = {*alpha, *kappa, *mu, *muB, *R, *T_star_n, *T_star_o, *tau_n, *tau_o,
                 tolerance[2], *c, *c_lr, *level};
  rhs_box(&s, &w, &p, &T, &h, &K, if0, il0, if1, il1, if1, xylo, dxy, &P, switch_map, il0 - if0 + 1);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* runge_kutta, update_state, tag_cells.                                     */
/* ------------------------------------------------------------------------ */

/* runge_kutta.m4:57-66 on a box of views. */
static void rk_box(view_t *w, const view_t *ww, const view_t *KK, int b0, int e0, int b1, int e1,
                   double dt, double alpha_1, double alpha_2, double beta) {
  for (int k = 0; k < 5; k++)
    for (int j = b1; j <= e1; j++)
      for (int i = b0; i <= e0; i++)
        *ref(w, i, j, k) = alpha_1 * at(ww, i, j, k) + alpha_2 * at(w, i, j, k) + beta * dt * at(KK, i, j, k);
}

void orc_runge_kutta(const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                     const int *gcw0, const int *gcw1, const double *dt, const double *alpha_1,
                     const double *alpha_2, const double *beta, double *w_n, const double *ww,
                     const double *KK, const int *nequ) {
  int if0 = *ifirst0, il0 = *ilast0, if1 = *ifirst1, il1 = *ilast1;
  (void)nequ;
  view_t w = patch_view(w_n, if0, il0, if1, il1, *gcw0, *gcw1);
  view_t a = patch_view(ww, if0, il0, if1, il1, 0, 0);
  view_t K = patch_view(KK, if0, il0, if1, il1, 0, 0);
  rk_box(&w, &a, &K, if0, il0, if1, il1, *dt, *alpha_1, *alpha_2, *beta);
}

typedef struct {
  double rho0, c, gamma1, c_p, p0, T0, c_lr, c_v, R;
} state_par_t;

/* update_state.m4:87-107 on a box of views; ifirst1/xylo of the enclosing
 * patch or tile. */
static void state_box(const view_t *w, view_t *p, view_t *T, int b0, int e0, int b1, int e1, int ifirst1,
                      const double *dxy, const double *xylo, const state_par_t *S) {
  for (int j = b1; j <= e1; j++) {
    double xc1 = xylo[1] + dxy[1] * ((double)(j - ifirst1) + 0.5);
    for (int i = b0; i <= e0; i++) {
      double w0 = at(w, i, j, 0), w3 = at(w, i, j, 3);
      *ref(p, i, j, 0) = sq(S->c + S->c_lr * xc1) *
                             ((w0 - S->rho0) + ((S->gamma1 - 1) / (2 * S->rho0)) * sq(w0 - S->rho0) +
                              (w0 / S->c_p) * (w3 / w0)) +
                         S->p0;
      *ref(T, i, j, 0) = S->T0 * exp((w3 / w0 - S->R * log(S->rho0 / w0)) / S->c_v);
    }
  }
}

void orc_update_state(const double *dxy, const double *xylo, const int *ifirst0, const int *ilast0,
                      const int *ifirst1, const int *ilast1, const int *gcw0, const int *gcw1,
                      const double *w_n, double *pres, double *temp, const double *rho0, const double *c,
                      const double *gamma1, const double *c_p, const double *p0, const double *T0,
                      const double *c_lr, const double *c_v, const double *R_tilde, const double *M,
                      const double *R, const int *nequ) {
  (void)R_tilde; (void)M; (void)nequ;
  int if0 = *ifirst0, il0 = *ilast0, if1 = *ifirst1, il1 = *ilast1;
  view_t w = patch_view(w_n, if0, il0, if1, il1, *gcw0, *gcw1);
  view_t p = patch_view(pres, if0, il0, if1, il1, *gcw0, *gcw1);
  view_t T = patch_view(temp, if0, il0, if1, il1, *gcw0, *gcw1);
  state_par_t S = {*rho0, *c, *gamma1, *c_p, *p0, *T0, *c_lr, *c_v, *R};
  state_box(&w, &p, &T, if0, il0, if1, il1, if1, dxy, xylo, &S);
}

/* tag_cells.m4:58-108. */
void orc_tag_cells(const int *ifirst0, const int *ilast0, const int *ifirst1, const int *ilast1,
                   const int *gcw0, const int *gcw1, int *tags, const double *w_n, const double *H0,
                   const int *refine_tag_val, const double *tolerance, const int *nequ) {
  (void)nequ;
  int if0 = *ifirst0, il0 = *ilast0, if1 = *ifirst1, il1 = *ilast1;
  long nx = il0 - if0 + 1;
  view_t w = patch_view(w_n, if0, il0, if1, il1, *gcw0, *gcw1);
  view_t h = patch_view(H0, if0, il0, if1, il1, *gcw0, *gcw1);
#define TAG(i, j) tags[(long)((i) - if0) + (long)((j) - if1) * nx]
  for (int j = if1; j <= il1; j++)
    for (int i = if0; i <= il0; i++) TAG(i, j) = 0;
  double ep = orc_sensor_eps(tolerance[0]);
  for (int j = if1; j <= il1; j++)
    for (int i = if0; i <= il0; i++) {
      double rx = smooth_ratio(at(&h, i, j, 0) - at(&h, i - 1, j, 0), at(&h, i + 1, j, 0) - at(&h, i, j, 0), ep);
      if (rx < LIT[L_0_9999]) TAG(i, j) = *refine_tag_val;
    }
  for (int j = if1; j <= il1; j++)
    for (int i = if0; i <= il0; i++) {
      double ry = smooth_ratio(at(&h, i, j, 0) - at(&h, i, j - 1, 0), at(&h, i, j + 1, 0) - at(&h, i, j, 0), ep);
      if (ry < LIT[L_0_9999]) TAG(i, j) = *refine_tag_val;
    }
  ep = orc_sensor_eps(tolerance[1]);
  for (int j = if1; j <= il1; j++)
    for (int i = if0; i <= il0; i++) {
      double rx = smooth_ratio(at(&w, i, j, 0) - at(&w, i - 1, j, 0), at(&w, i + 1, j, 0) - at(&w, i, j, 0), ep);
      if (rx < LIT[L_0_9999]) TAG(i, j) = *refine_tag_val;
    }
  for (int j = if1; j <= il1; j++)
    for (int i = if0; i <= il0; i++) {
      double ry = smooth_ratio(at(&w, i, j, 0) - at(&w, i, j - 1, 0), at(&w, i, j + 1, 0) - at(&w, i, j, 0), ep);
      if (ry < LIT[L_0_9999]) TAG(i, j) = *refine_tag_val;
    }
#undef TAG
}

/* ------------------------------------------------------------------------ */
/* Physical boundary fill: enlightning.C:790-1013.                           */
/* ------------------------------------------------------------------------ */

/* Fills the ghost strips of w_n (5 comps), pres, temp, source of a patch.
 * touch[s] says whether side s (0=left,1=right,2=bottom,3=top) lies on the
 * physical boundary; reflect[s] = 1 for REFLECT, 0 for ABSORB.  Edge strips
 * span the patch's interior range along the side (SAMRAI codim-1 boundary
 * boxes); a corner ghost block is filled when either adjacent side touches
 * (codim-2 boxes), after all edges, copying along x as enlightning.C:948-1008
 * does. */
void orc_fill_physical_bc(int if0, int il0, int if1, int il1, int g0, int g1, const int *touch,
                          const int *reflect, double *w_n, double *pres, double *temp, double *source) {
  long nxg = il0 - if0 + 1 + 2 * g0, nyg = il1 - if1 + 1 + 2 * g1;
  long area = nxg * nyg;
  double *scal[3] = {pres, temp, source};
#define IDX(i, j) ((long)((j) - if1 + g1) * nxg + (long)((i) - if0 + g0))
#define COPY(dst, src, negk, neg)                                  \
  do {                                                             \
    for (int s_ = 0; s_ < 3; s_++)                                 \
      if (scal[s_]) scal[s_][dst] = scal[s_][src];                 \
    for (int k = 0; k < 5; k++) {                                  \
      w_n[k * area + (dst)] = w_n[k * area + (src)];               \
      if (k == (negk) && (neg)) w_n[k * area + (dst)] = -w_n[k * area + (dst)]; \
    }                                                              \
  } while (0)
  /* edges: left, right, bottom, top (enlightning.C:836-944) */
  if (touch[0])
    for (int j = if1; j <= il1; j++)
      for (int i = if0 - 1; i >= if0 - g0; i--) COPY(IDX(i, j), IDX(i + 1, j), 1, i == if0 - 1 && reflect[0]);
  if (touch[1])
    for (int j = if1; j <= il1; j++)
      for (int i = il0 + 1; i <= il0 + g0; i++) COPY(IDX(i, j), IDX(i - 1, j), 1, i == il0 + 1 && reflect[1]);
  if (touch[2])
    for (int j = if1 - 1; j >= if1 - g1; j--)
      for (int i = il0; i >= if0; i--) COPY(IDX(i, j), IDX(i, j + 1), 2, j == if1 - 1 && reflect[2]);
  if (touch[3])
    for (int j = il1 + 1; j <= il1 + g1; j++)
      for (int i = if0; i <= il0; i++) COPY(IDX(i, j), IDX(i, j - 1), 2, j == il1 + 1 && reflect[3]);
  /* corners (node boundaries 0:lower-left 1:lower-right 2:upper-left
   * 3:upper-right), enlightning.C:948-1008 */
  for (int node = 0; node < 4; node++) {
    int right = node & 1, upper = node >> 1;
    if (!(touch[right ? 1 : 0] || touch[upper ? 3 : 2])) continue;
    int jb = upper ? il1 + 1 : if1 - g1, je = upper ? il1 + g1 : if1 - 1;
    for (int j = jb; j <= je; j++) {
      if (!right) {
        for (int i = if0 - 1; i >= if0 - g0; i--) COPY(IDX(i, j), IDX(i + 1, j), 1, i == if0 - 1 && reflect[0]);
      } else {
        for (int i = il0 + 1; i <= il0 + g0; i++) COPY(IDX(i, j), IDX(i - 1, j), 1, i == il0 + 1 && reflect[1]);
      }
    }
  }
#undef COPY
#undef IDX
}

/* Absorbing envelopes, enlightning::setAbsorbingBoundaryConditions
 * (enlightning.C:630-786).  Sides with absorb[s] set (0 left, 1 right,
 * 2 bottom, 3 top, in that order) damp the interior cells near them toward
 * the ambient by tmp = -exp(-log(2.0) / W^2 * d^2) + 1, d the distance of
 * the cell centre from the domain edge.  Each side is walked from the
 * patch edge inward: a cell (row) is processed while the previous tmp is
 * more than 1e-8 from 1; the test before the first cell uses the edge
 * cell's own tmp, and along the left / right sides every row after the
 * first restarts with tmp = 0. */
static double orc_env(double d, double width) { return -exp(-log(2.0) / (width * width) * (d * d)) + 1; }

static void orc_damp(double *w_n, double *pres, double *temp, long area, long c, double tmp,
                     const enl_gas_t *gas) {
  pres[c] = (pres[c] - gas->p0) * tmp + gas->p0;
  temp[c] = (temp[c] - gas->T0) * tmp + gas->T0;
  w_n[c] = (w_n[c] - gas->rho0) * tmp + gas->rho0;
  w_n[area + c] = w_n[area + c] * tmp;
  w_n[2 * area + c] = w_n[2 * area + c] * tmp;
  w_n[3 * area + c] = w_n[3 * area + c] * tmp;
  w_n[4 * area + c] = (w_n[4 * area + c] - gas->T0 * gas->rho0) * tmp + gas->T0 * gas->rho0;
}

/* The patch box [if0, il0] x [if1, il1] inside arrays whose ghost box is
 * [af0 - g0, al0 + g0] x [af1 - g1, al1 + g1]. */
static void absorb_box(int if0, int il0, int if1, int il1, int af0, int al0, int af1, int al1, int g0, int g1,
                       const double *pxlo, const double *dx, const double *gxlo, const double *gxhi,
                       const int *absorb, double width, const enl_gas_t *gas, double *w_n, double *pres,
                       double *temp) {
  const long nxg = al0 - af0 + 1 + 2L * g0, area = nxg * (al1 - af1 + 1 + 2L * g1);
  const double ep = 1.0e-8;
#define C_(i, j) ((long)((j) - af1 + g1) * nxg + (long)((i) - af0 + g0))
#define XC0(i) (pxlo[0] + dx[0] * ((double)((i) - if0) + 0.5))
#define XC1(j) (pxlo[1] + dx[1] * ((double)((j) - if1) + 0.5))
  double tmp;
  if (absorb[0]) {
    tmp = orc_env(XC0(if0) - gxlo[0], width);
    for (int j = if1; j <= il1; j++) {
      for (int i = if0; fabs(1.0 - tmp) > ep && i <= il0; i++) {
        tmp = orc_env(XC0(i) - gxlo[0], width);
        orc_damp(w_n, pres, temp, area, C_(i, j), tmp, gas);
      }
      tmp = 0;
    }
  }
  if (absorb[1]) {
    tmp = orc_env(gxhi[0] - XC0(il0), width);
    for (int j = if1; j <= il1; j++) {
      for (int i = il0; fabs(1.0 - tmp) > ep && i >= if0; i--) {
        tmp = orc_env(gxhi[0] - XC0(i), width);
        orc_damp(w_n, pres, temp, area, C_(i, j), tmp, gas);
      }
      tmp = 0;
    }
  }
  if (absorb[2]) {
    tmp = orc_env(XC1(if1) - gxlo[1], width);
    for (int j = if1; fabs(1.0 - tmp) > ep && j <= il1; j++) {
      tmp = orc_env(XC1(j) - gxlo[1], width);
      for (int i = if0; i <= il0; i++) orc_damp(w_n, pres, temp, area, C_(i, j), tmp, gas);
    }
  }
  if (absorb[3]) {
    tmp = orc_env(gxhi[1] - XC1(il1), width);
    for (int j = il1; fabs(1.0 - tmp) > ep && j >= if1; j--) {
      tmp = orc_env(gxhi[1] - XC1(j), width);
      for (int i = if0; i <= il0; i++) orc_damp(w_n, pres, temp, area, C_(i, j), tmp, gas);
    }
  }
#undef C_
#undef XC0
#undef XC1
}

void orc_absorb(int if0, int il0, int if1, int il1, int g0, int g1, const double *pxlo, const double *dx,
                const double *gxlo, const double *gxhi, const int *absorb, double width, const enl_gas_t *gas,
                double *w_n, double *pres, double *temp) {
  absorb_box(if0, il0, if1, il1, if0, il0, if1, il1, g0, g1, pxlo, dx, gxlo, gxhi, absorb, width, gas, w_n, pres,
             temp);
}

/* Mass source of a patch, enlightning::addSource (enlightning.C:2140-2307),
 * cell by cell in the reference's loop order: POINT and SINE pulses and the
 * DRIVEN point overwrite the interior source; a LIGHTNING channel merges
 * each segment k in turn (end caps on their side of the normal lines
 * through the end points, the perpendicular Gaussian between them) as
 * v + t - (v / amp) * t.  source: interior cells, row stride sy. */
#define ORC_PI 3.14159265358979323846264338
int orc_add_source(int if0, int il0, int if1, int il1, const double *pxlo, const double *dx,
                   const orc_source_t *src, int iteration, int pulse_time0, double dt, double *source, long sy) {
  const double amp = src->amp, sp2 = src->spread * src->spread;
  if (src->mode == 20 && src->type != 10) return -2;
  for (int j = if1; j <= il1; j++) {
    const double xc1 = pxlo[1] + dx[1] * ((double)(j - if1) + 0.5);
    for (int i = if0; i <= il0; i++) {
      const double xc0 = pxlo[0] + dx[0] * ((double)(i - if0) + 0.5);
      double *o = source + (long)(j - if1) * sy + (i - if0);
      const double x0 = xc0 - src->pos[0], x1 = xc1 - src->pos[1];
      if (src->type == 10 && src->mode == 10) {
        *o = amp * exp(-log(2.0) / sp2 * (x0 * x0 + x1 * x1));
      } else if (src->type == 10) {
        *o = amp * exp(-log(2.0) / sp2 * (x0 * x0 + x1 * x1)) *
             sin(src->omega * (iteration + 1 - pulse_time0) * dt);
      } else if (src->type == 20) {
        double d = x0 - src->sine_amp * sin(src->sine_freq * x1);
        *o = amp * exp(-log(2.0) / sp2 * (d * d));
      }
    }
  }
  if (src->type != 30) return 0;
  /* the channel: segment by segment over the whole patch */
  for (int k = 0; k < src->nseg; k++) {
    const double *c = src->coords + 4 * k;
    const double theta = -atan((c[3] - c[1]) / (c[2] - c[0]));
    int orient = 0;
    if ((c[0] > c[2] && c[1] > c[3]) || (c[0] <= c[2] && c[1] >= c[3])) orient = 1;
    else if ((c[0] >= c[2] && c[1] <= c[3]) || (c[0] < c[2] && c[1] < c[3])) orient = 2;
    for (int j = if1; j <= il1; j++) {
      const double xc1 = pxlo[1] + dx[1] * ((double)(j - if1) + 0.5);
      for (int i = if0; i <= il0; i++) {
        const double xc0 = pxlo[0] + dx[0] * ((double)(i - if0) + 0.5);
        const double l0 = tan(ORC_PI / 2.0 - theta) * (xc0 - c[0]) + c[1];
        const double l1 = tan(ORC_PI / 2.0 - theta) * (xc0 - c[2]) + c[3];
        double *o = source + (long)(j - if1) * sy + (i - if0);
        if ((orient == 1 && xc1 >= l0) || (orient == 2 && xc1 <= l0)) {
          double t = amp * exp(-log(2.0) / sp2 * ((xc0 - c[0]) * (xc0 - c[0]) + (xc1 - c[1]) * (xc1 - c[1])));
          *o = *o + t - (*o / amp) * t;
        }
        if ((orient == 1 && xc1 <= l1) || (orient == 2 && xc1 >= l1)) {
          double t = amp * exp(-log(2.0) / sp2 * ((xc0 - c[2]) * (xc0 - c[2]) + (xc1 - c[3]) * (xc1 - c[3])));
          *o = *o + t - (*o / amp) * t;
        }
        if ((orient == 1 && xc1 < l0 && xc1 > l1) || (orient == 2 && xc1 > l0 && xc1 < l1)) {
          double d = sin(theta) * (xc0 - c[0]) + cos(theta) * (xc1 - c[1]);
          double t = amp * exp(-log(2.0) / sp2 * (d * d));
          *o = *o + t - (*o / amp) * t;
        }
      }
    }
  }
  return 0;
}

/* Microphones, enlightning::recordAudio (enlightning.C:2414-2436): the
 * pressure of cell ((int)((x-xlo0)/dx0), (int)((y-xlo1)/dx1)) of the
 * interior array (row stride sy) for each microphone inside the patch
 * (bounds inclusive); hit[k] says whether it was. */
void orc_sample_mics(int if0, int il0, int if1, int il1, const double *pxlo, const double *pxhi, const double *dx,
                     int nmics, const double *mx, const double *my, const double *pres, long sy, double *out,
                     int *hit) {
  (void)if0; (void)il0; (void)if1; (void)il1;
  for (int k = 0; k < nmics; k++) {
    hit[k] = 0;
    if (mx[k] >= pxlo[0] && mx[k] <= pxhi[0] && my[k] >= pxlo[1] && my[k] <= pxhi[1]) {
      int i = (int)((mx[k] - pxlo[0]) / dx[0]);
      int j = (int)((my[k] - pxlo[1]) / dx[1]);
      out[k] = pres[(long)j * sy + i];
      hit[k] = 1;
    }
  }
}

/* enlightning::readLightningInput (enlightning.C:2333-2389): lines read with
 * fgets(line, 80) and sscanf("%f %f %f %f"); an unparsed line repeats the
 * previous values.  Returns the segment count (coords may be NULL). */
int orc_read_lightning(const char *path, double *coords, int max_segments) {
  FILE *fr = fopen(path, "rt");
  if (!fr) return -1;
  char line[80];
  float x1 = 0, y1 = 0, x2 = 0, y2 = 0;
  int n = 0;
  while (fgets(line, 80, fr) != NULL) {
    if (sscanf(line, "%f %f %f %f", &x1, &y1, &x2, &y2) < 0) { /* values kept */ }
    if (coords && n < max_segments) {
      coords[4 * n] = x1;
      coords[4 * n + 1] = y1;
      coords[4 * n + 2] = x2;
      coords[4 * n + 3] = y2;
    }
    n++;
  }
  fclose(fr);
  return n;
}

/* ------------------------------------------------------------------------ */
/* Single-patch MOL driver (the shape of SAMRAI's advanceHierarchy around    */
/* enlightning::singleStep, for one patch covering the whole domain).        */
/* ------------------------------------------------------------------------ */

static void rhs_params_from(const enl_gas_t *g, int level, rhs_par_t *P) {
  P->alpha = g->weno_alpha; P->kappa = g->kappa; P->mu = g->mu; P->muB = g->muB; P->R = g->R;
  P->T_star_n = g->T_star_n; P->T_star_o = g->T_star_o; P->tau_n = g->tau_n; P->tau_o = g->tau_o;
  P->tol2 = g->tag_tolerance[2]; P->c = g->c; P->c_lr = g->c_lr; P->level = level;
}
static void state_params_from(const enl_gas_t *g, state_par_t *S) {
  S->rho0 = g->rho0; S->c = g->c; S->gamma1 = g->gamma1; S->c_p = g->c_p; S->p0 = g->p0;
  S->T0 = g->T0; S->c_lr = g->c_lr; S->c_v = g->c_v; S->R = g->R;
}

/* Advances one patch (all four sides on the physical boundary) by nsteps
 * steps of an `order`-stage MOL Runge-Kutta scheme.  tile > 0 splits every
 * kernel call into tile x tile sub-boxes processed by OpenMP threads (one
 * sub-box per thread at a time, the patch decomposition a SAMRAI run with
 * that largest patch size would give); tile <= 0 runs whole-patch calls.
 * ww and KK are caller-provided interior scratch of nx*ny*5 doubles. */
int orc_advance_patch(int variant, int if0, int il0, int if1, int il1, int g, const double *xylo,
                      const double *dxy, const enl_gas_t *gas, const int *reflect, int order,
                      const double *alpha_1, const double *alpha_2, const double *beta, double dt,
                      int nsteps, int level, double *w_n, double *pres, double *temp, double *source,
                      double *ww, double *KK, int tile, int nthreads, double absorb_width) {
  scheme_t s = scheme_of(variant);
  if (s.order < 0) return -1;
  if (g < s.reach) return -2;
  int touch[4] = {1, 1, 1, 1};
  long nx = il0 - if0 + 1, ny = il1 - if1 + 1;
  view_t w = patch_view(w_n, if0, il0, if1, il1, g, g);
  view_t p = patch_view(pres, if0, il0, if1, il1, g, g);
  view_t T = patch_view(temp, if0, il0, if1, il1, g, g);
  view_t h = patch_view(source, if0, il0, if1, il1, g, g);
  view_t a = patch_view(ww, if0, il0, if1, il1, 0, 0);
  view_t K = patch_view(KK, if0, il0, if1, il1, 0, 0);
  rhs_par_t P;
  state_par_t S;
  rhs_params_from(gas, level, &P);
  state_params_from(gas, &S);
  int tsz = tile > 0 ? tile : (int)(nx > ny ? nx : ny);
  int ntx = (int)((nx + tsz - 1) / tsz), nty = (int)((ny + tsz - 1) / tsz);
  int ntiles = ntx * nty;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int step = 0; step < nsteps; step++) {
    /* current -> scratch: ww holds the state at the start of the step */
    for (int k = 0; k < 5; k++)
      for (int j = if1; j <= il1; j++)
        for (int i = if0; i <= il0; i++) *ref(&a, i, j, k) = at(&w, i, j, k);
    for (int st = 0; st < order; st++) {
      orc_fill_physical_bc(if0, il0, if1, il1, g, g, touch, reflect, w_n, pres, temp, source);
      /* rhs reads the source at interior points only (interior context) */
#pragma omp parallel for schedule(dynamic, 1) if (tile > 0)
      for (int t = 0; t < ntiles; t++) {
        int b0 = if0 + (t % ntx) * tsz, b1 = if1 + (t / ntx) * tsz;
        int e0 = b0 + tsz - 1 < il0 ? b0 + tsz - 1 : il0;
        int e1 = b1 + tsz - 1 < il1 ? b1 + tsz - 1 : il1;
        double txlo[2] = {xylo[0] + dxy[0] * (b0 - if0), xylo[1] + dxy[1] * (b1 - if1)};
        rhs_box(&s, &w, &p, &T, &h, &K, b0, e0, b1, e1, b1, txlo, dxy, &P, NULL, 0);
      }
#pragma omp parallel for schedule(dynamic, 1) if (tile > 0)
      for (int t = 0; t < ntiles; t++) {
        int b0 = if0 + (t % ntx) * tsz, b1 = if1 + (t / ntx) * tsz;
        int e0 = b0 + tsz - 1 < il0 ? b0 + tsz - 1 : il0;
        int e1 = b1 + tsz - 1 < il1 ? b1 + tsz - 1 : il1;
        double txlo[2] = {xylo[0] + dxy[0] * (b0 - if0), xylo[1] + dxy[1] * (b1 - if1)};
        rk_box(&w, &a, &K, b0, e0, b1, e1, dt, alpha_1[st], alpha_2[st], beta[st]);
        state_box(&w, &p, &T, b0, e0, b1, e1, b1, dxy, txlo, &S);
      }
      if (absorb_width > 0 && (!reflect[0] || !reflect[1] || !reflect[2] || !reflect[3])) {
        /* envelopes on the ABSORB sides, per patch (tile) as singleStep
         * applies them (enlightning.C:581-582); the domain is this patch */
        int ab[4] = {!reflect[0], !reflect[1], !reflect[2], !reflect[3]};
        double gxhi[2] = {xylo[0] + dxy[0] * (double)nx, xylo[1] + dxy[1] * (double)ny};
        for (int t = 0; t < ntiles; t++) {
          int b0 = if0 + (t % ntx) * tsz, b1 = if1 + (t / ntx) * tsz;
          int e0 = b0 + tsz - 1 < il0 ? b0 + tsz - 1 : il0;
          int e1 = b1 + tsz - 1 < il1 ? b1 + tsz - 1 : il1;
          double txlo[2] = {xylo[0] + dxy[0] * (b0 - if0), xylo[1] + dxy[1] * (b1 - if1)};
          absorb_box(b0, e0, b1, e1, if0, il0, if1, il1, g, g, txlo, dxy, xylo, gxhi, ab, absorb_width, gas, w_n,
                     pres, temp);
        }
      }
    }
  }
  return 0;
}

/* The literal table, for the pin tests. */
void orc_literals(double *out /* L_COUNT = 16 */) {
  for (int k = 0; k < L_COUNT; k++) out[k] = LIT[k];
}
