// flops.cc -- counting-double restatement of the oracle (TEST
// INFRASTRUCTURE: run by tests/test_flop_counts.py, never by the product).
//
// enl_oracle.c is compiled here with every `double` replaced by cnt_t, a
// double that counts the floating-point operations applied to it, and its
// kernels are run on single cells.  The counts are therefore those of the
// reference's arithmetic as the oracle restates it (the same expressions,
// the same order).  Counting rule: +, -, *, /, fabs and comparisons count
// 1 each; exp and log are counted apart; negation, copies and conversions
// count 0.
//
// Output: one line per configuration,
//   <name> add sub mul div abs cmp exp log flops
// with flops = add + sub + mul + div + abs + cmp.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace {
struct Counts {
  long add, sub, mul, div, abs, cmp, exp, log;
};
Counts g_n;
}  // namespace

struct cnt_t {
  double v;
  cnt_t() : v(0) {}
  cnt_t(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
  explicit operator int() const { return (int)v; }
  cnt_t operator-() const { return cnt_t(-v); }
  cnt_t& operator+=(const cnt_t& o) { g_n.add++; v += o.v; return *this; }
  cnt_t& operator-=(const cnt_t& o) { g_n.sub++; v -= o.v; return *this; }
  cnt_t& operator*=(const cnt_t& o) { g_n.mul++; v *= o.v; return *this; }
  cnt_t& operator/=(const cnt_t& o) { g_n.div++; v /= o.v; return *this; }
};
static inline cnt_t operator+(cnt_t a, cnt_t b) { g_n.add++; return cnt_t(a.v + b.v); }
static inline cnt_t operator-(cnt_t a, cnt_t b) { g_n.sub++; return cnt_t(a.v - b.v); }
static inline cnt_t operator*(cnt_t a, cnt_t b) { g_n.mul++; return cnt_t(a.v * b.v); }
static inline cnt_t operator/(cnt_t a, cnt_t b) { g_n.div++; return cnt_t(a.v / b.v); }
static inline bool operator<(cnt_t a, cnt_t b) { g_n.cmp++; return a.v < b.v; }
static inline bool operator>(cnt_t a, cnt_t b) { g_n.cmp++; return a.v > b.v; }
static inline bool operator<=(cnt_t a, cnt_t b) { g_n.cmp++; return a.v <= b.v; }
static inline bool operator>=(cnt_t a, cnt_t b) { g_n.cmp++; return a.v >= b.v; }
static inline bool operator==(cnt_t a, cnt_t b) { g_n.cmp++; return a.v == b.v; }
static inline bool operator!=(cnt_t a, cnt_t b) { g_n.cmp++; return a.v != b.v; }
static inline cnt_t fabs(cnt_t a) { g_n.abs++; return cnt_t(::fabs(a.v)); }
static inline cnt_t exp(cnt_t a) { g_n.exp++; return cnt_t(::exp(a.v)); }
static inline cnt_t log(cnt_t a) { g_n.log++; return cnt_t(::log(a.v)); }
// used only by the source / absorb code of the oracle, never counted below
static inline cnt_t sin(cnt_t a) { return cnt_t(::sin(a.v)); }
static inline cnt_t cos(cnt_t a) { return cnt_t(::cos(a.v)); }
static inline cnt_t tan(cnt_t a) { return cnt_t(::tan(a.v)); }
static inline cnt_t atan(cnt_t a) { return cnt_t(::atan(a.v)); }

#define double cnt_t
#include "enl_oracle.c"
#undef double

namespace {

Counts take() {
  Counts c = g_n;
  memset(&g_n, 0, sizeof g_n);
  return c;
}

void print(const char* name, const Counts& c) {
  long fl = c.add + c.sub + c.mul + c.div + c.abs + c.cmp;
  printf("%s %ld %ld %ld %ld %ld %ld %ld %ld %ld\n", name, c.add, c.sub, c.mul, c.div, c.abs, c.cmp, c.exp, c.log,
         fl);
}

const int G = 3;           // ghost width of the test patches
const int N = 2 + 2 * G;   // the largest ghosted side

// A patch of nx x 1 cells with a density step after its first cell (WENO at
// level 0) or with density uniform along x (DRP at level 0).
void state(cnt_t* w, cnt_t* p, cnt_t* T, bool step, int nx) {
  const int Nx = nx + 2 * G, Ny = 1 + 2 * G, sk = Nx * Ny;
  for (int j = 0; j < Ny; j++)
    for (int i = 0; i < Nx; i++) {
      int c = i + Nx * j;
      double rho = 1.2 + (step && i >= G + 1 ? 0.3 : 0.0) + 1e-3 * j;
      w[c] = rho;
      w[c + sk] = 3.0 + 0.1 * i;
      w[c + 2 * sk] = -2.0 + 0.2 * j;
      w[c + 3 * sk] = rho * 5.0;
      w[c + 4 * sk] = rho * 290.0;
      p[c] = 101325.0 + 10.0 * i;
      T[c] = 294.0 + 0.5 * j;
    }
}

// The rhs over the first nx cells of row 0 of that patch; sm gets the
// scheme of each.
Counts rhs_run(int variant, int level, bool step, int nx, int* sm) {
  cnt_t w[5 * N * N], p[N * N], T[N * N], H0[2] = {cnt_t(2.0), cnt_t(2.0)}, KK[10];
  state(w, p, T, step, nx);
  const int z = 0, il0 = nx - 1, g = G, nequ = 5;
  cnt_t xlo[2] = {cnt_t(0.0), cnt_t(0.0)}, dx[2] = {cnt_t(1.0 / 512), cnt_t(1.0 / 256)};
  cnt_t alpha = 360.0, kappa = 0.02624, mu = 1.846e-5, muB = 1.1076e-5, R = 287.06, Tsn = 3352.0, Tso = 2239.0,
        tau_n = 1.1e-3, tau_o = 5e-5, c = 343.8, c_lr = 0.0;
  cnt_t tol[3] = {cnt_t(1.5), cnt_t(1e-2), cnt_t(5e-3)};
  take();
  orc_rhs_variant(variant, &z, &il0, &z, &z, &g, &g, xlo, dx, w, H0, KK, p, T, &alpha, &kappa, &mu, &muB, &R, &Tsn,
                  &Tso, &tau_n, &tau_o, tol, &level, &c, &c_lr, &nequ, sm);
  return take();
}

Counts diff(const Counts& two, const Counts& one) {
  Counts d;
  d.add = two.add - one.add;
  d.sub = two.sub - one.sub;
  d.mul = two.mul - one.mul;
  d.div = two.div - one.div;
  d.abs = two.abs - one.abs;
  d.cmp = two.cmp - one.cmp;
  d.exp = two.exp - one.exp;
  d.log = two.log - one.log;
  return d;
}

// update_state (update_state.m4:87-107) over nx x 1 cells
Counts state_run(int nx) {
  cnt_t w[5 * N * N], p[N * N], T[N * N];
  state(w, p, T, false, nx);
  const int z = 0, il0 = nx - 1, g = G, nequ = 5;
  cnt_t dx[2] = {cnt_t(1.0 / 512), cnt_t(1.0 / 512)}, xlo[2] = {cnt_t(0.0), cnt_t(0.0)};
  cnt_t rho0 = 1.2, c = 343.8, gamma1 = 1.4, c_p = 1010.0, p0 = 101325.0, T0 = 294.1, c_lr = 0.0, c_v = 720.4,
        Rt = 8314.32, M = 28.9644, R = 287.06;
  take();
  orc_update_state(dx, xlo, &z, &il0, &z, &z, &g, &g, w, p, T, &rho0, &c, &gamma1, &c_p, &p0, &T0, &c_lr, &c_v, &Rt,
                   &M, &R, &nequ);
  return take();
}

// tag_cells (tag_cells.m4:58-108) over nx x 1 cells
Counts tag_run(int nx) {
  cnt_t w[5 * N * N], p[N * N], T[N * N], H0[N * N];
  state(w, p, T, true, nx);
  for (int k = 0; k < N * N; k++) H0[k] = 0.5 * (k % 3);
  const int z = 0, il0 = nx - 1, g = G, nequ = 5, val = 1;
  cnt_t tol[3] = {cnt_t(1.5), cnt_t(1e-2), cnt_t(5e-3)};
  int tags[2] = {0, 0};
  take();
  orc_tag_cells(&z, &il0, &z, &z, &g, &g, tags, w, H0, &val, tol, &nequ);
  return take();
}

// Per-cell counts: a two-cell run less a one-cell run (the sensor's
// epsilon is formed once per call, rhs_hybrid_high_llf.m4:114-115).
Counts rhs_cell(int variant, int level, bool step, int* weno) {
  int sm1[2] = {-1, -1}, sm2[2] = {-1, -1};
  Counts one = rhs_run(variant, level, step, 1, sm1);
  Counts two = rhs_run(variant, level, step, 2, sm2);
  if (sm2[0] != sm2[1] || sm1[0] != sm2[0]) {
    fprintf(stderr, "cells took different schemes\n");
    exit(1);
  }
  *weno = sm1[0];
  return diff(two, one);
}

}  // namespace

int main() {
  static const char* names[9] = {"hybrid_high_llf", "hybrid_high", "hybrid_low_llf", "hybrid_low", "weno_high_llf",
                                 "weno_high",       "weno_low_llf", "weno_low",       "drp"};
  // per variant: the cells a level-0 patch can take (DRP and WENO, after the
  // sensor where there is one) and a cell above level 0 (no sensor)
  char buf[96], last[96] = "";
  for (int v = 0; v < 9; v++) {
    for (int level = 0; level < 2; level++) {
      for (int step = 0; step < 2; step++) {
        int weno = -1;
        Counts n = rhs_cell(v, level, step != 0, &weno);
        snprintf(buf, sizeof buf, "rhs.%s.level%d.%s", names[v], level ? 1 : 0, weno ? "weno" : "drp");
        if (strcmp(buf, last)) print(buf, n);
        strcpy(last, buf);
      }
    }
  }
  {
    // one interior cell of runge_kutta (runge_kutta.m4:57-66)
    cnt_t w[5 * N * N], ww[5], KK[5];
    for (int k = 0; k < 5 * N * N; k++) w[k] = 1.0 + k;
    for (int k = 0; k < 5; k++) ww[k] = KK[k] = 2.0 + k;
    const int z = 0, g = G, nequ = 5;
    cnt_t dt = 1e-5, a1 = 0.75, a2 = 0.25, b = 0.25;
    take();
    orc_runge_kutta(&z, &z, &z, &z, &g, &g, &dt, &a1, &a2, &b, w, ww, KK, &nequ);
    print("runge_kutta", take());
  }
  print("update_state", diff(state_run(2), state_run(1)));
  print("tag_cells", diff(tag_run(2), tag_run(1)));
  return 0;
}
