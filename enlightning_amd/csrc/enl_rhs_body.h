// enl_rhs_body.h -- device bodies of the rhs kernels (fortran/rhs_*.m4),
// shared by the single-patch kernels (enl_rhs.hip) and the batched
// hierarchy kernels (enl_batch.hip).  See enl_rhs.hip for the design.
#pragma once
#include <cmath>

#include "enl_device.h"
#include "enl_kernels.h"

#pragma clang fp contract(off)

namespace enl {

// A BX x BY tile of interior cells plus a 3-cell halo, staged in LDS.
template <int BX, int BY>
struct TileShape {
  static constexpr int halo = 3;
  static constexpr int TW = BX + 2 * halo;
  static constexpr int TH = BY + 2 * halo;
  static constexpr int TN = TW * TH;
  static constexpr int threads = BX * BY;
};

// Workgroups of a 1-D grid are dispatched round-robin over the 8 XCDs
// (workgroup b runs on XCD b % 8), each XCD with its own L2.  xcd_tile maps
// the dispatch index b of n workgroups to a tile index such that each XCD
// works through one contiguous range of tiles: neighbouring tiles, whose
// halos overlap, then share an L2.  A bijection on [0, n).
__device__ inline int xcd_tile(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

constexpr int kStrip = 32;
// Tile (tx, ty) of workgroup b on an nbx x nby tile grid.  The linear tile
// order runs through vertical strips kStrip tiles wide, row-major inside
// a strip, and xcd_tile gives each XCD a contiguous range of it.  A tile's
// upper neighbour then follows it kStrip tiles later on the same XCD,
// while the halo rows they share are still in that XCD's L2; in row-major
// order a whole row of tiles (128 at 4096^2) passes between them, more data
// than one XCD's 4 MB L2 keeps.
__device__ inline int2 strip_tile(int b, int nbx, int nby) {
  const int t = xcd_tile(b, nbx * nby);
  const int full = kStrip * nby;
  const int s = t / full, r = t - s * full;
  const int w = min(kStrip, nbx - s * kStrip);
  return make_int2(s * kStrip + r % w, r / w);
}

// LDS of one staged tile.  w[1..4] (the raw momenta etc.) are only staged by
// the dense kernel; w[0] is rho.
template <int BX, int BY>
struct TileLds {
  static constexpr int TN = TileShape<BX, BY>::TN;
  double* w[5];
  double *p, *T, *u, *v, *s, *tn;
};

__host__ __device__ inline void fill_rhs_constants(RhsArgs& a) {
  a.inv_dx2 = 1.0 / (a.dx * a.dx);
  a.inv_dy2 = 1.0 / (a.dy * a.dy);
  a.inv_4dxdy = 1.0 / (4 * a.dx * a.dy);
  a.inv_2dx = 1.0 / (2 * a.dx);
  a.inv_2dy = 1.0 / (2 * a.dy);
  a.inv_mdx = 1.0 / (-a.dx);
  a.inv_mdy = 1.0 / (-a.dy);
  a.inv_dx = 1.0 / a.dx;
  a.inv_dy = 1.0 / a.dy;
  // operands of the source and update_state quotients that are the same
  // for every cell (div_nr's range conditions)
  a.fast_div = den_ok(a.tau_n) && num_ok(a.T_star_n) && num_ok(a.mu) && num_ok(a.muB) && num_ok(a.kappa);
  a.eos_k1 = (a.gamma1 - 1) / (2 * a.rho0);
  a.eos_ok = num_ok(a.rho0) && den_ok(a.c_p) && den_ok(a.c_v);
  a.ep_pre = a.ep > 0 ? 5e-5 * a.ep : -1.0;
}

// a / c for a grid constant c; a * (1/c) when 1/c is exact (POW2).
template <bool POW2>
__device__ inline double divc(double a, double c, double inv_c) {
  if constexpr (POW2)
    return a * inv_c;
  else
    return a / c;
}

// Pointers into the LDS tile; offsets are relative to a cell centre, row
// stride S (the staged tile width).  w[k] is only staged by the dense kernel.
struct Tile {
  const double* w0;
  const double* w[5];
  const double* p;
  const double* T;
  const double* u;
  const double* v;
  const double* s;
  const double* tn;
};

// Inviscid fluxes F (x) and G (y), rhs_hybrid_high_llf.m4:125-156.
__device__ inline double flux_f(double w0, double u, double v, double s, double tn, double p, int k) {
#pragma clang fp contract(fast)
  switch (k) {
    case 0: return w0 * u;
    case 1: return w0 * sq(u) + p;
    case 2: return w0 * u * v;
    case 3: return w0 * u * s;
    default: return w0 * u * tn;
  }
}
__device__ inline double flux_g(double w0, double u, double v, double s, double tn, double p, int k) {
#pragma clang fp contract(fast)
  switch (k) {
    case 0: return w0 * v;
    case 1: return w0 * v * u;
    case 2: return w0 * sq(v) + p;
    case 3: return w0 * v * s;
    default: return w0 * v * tn;
  }
}
__device__ inline double tflux_f(const Tile& t, int e, int k) {
  return flux_f(t.w0[e], t.u[e], t.v[e], t.s[e], t.tn[e], t.p[e], k);
}
__device__ inline double tflux_g(const Tile& t, int e, int k) {
  return flux_g(t.w0[e], t.u[e], t.v[e], t.s[e], t.tn[e], t.p[e], k);
}

// What a cell's rhs becomes (the FUSED parameter of the kernel bodies):
// kRhs stores it; kStage applies the runge_kutta update and update_state and
// stores the new state; kStageBc does that and, when the arguments' bc_fold
// asks for it, also writes the physical boundary ghosts (the single-patch
// stage: the batched kernels' patches fill ghosts by schedules).
constexpr int kRhs = 0, kStage = 1, kStageBc = 2;

// Result of one cell: store rhs (plain), or apply the runge_kutta update
// and update_state and store the new state (fused stage).
// The fused stage's operands of the runge_kutta update at cell (i, j):
// ww (wa) and the stage input w (wb).
__device__ inline void rk_operands(const RhsArgs& a, int i, int j, double wa[5], double wb[5]) {
  const Geom& G = a.geom;
  const long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    wa[k] = a.ww[a.ww_off + i + j * a.ww_sy + k * a.ww_sk];
    wb[k] = a.w[g + k * G.sk];
  }
}

// The physical boundary fold of the fused stage (a.bc_fold nonzero): the
// ghost cells whose boundary copy comes from interior cell (i, j), as
// bc_edge_kernel and bc_corner_kernel fill them (enlightning.C:836-1008):
// every layer of an edge strip is the nearest interior cell, with u negated
// on a reflecting left/right side and v on a reflecting bottom/top side;
// each corner block is the corner cell with both applicable negations.
// put(d, neg_u, neg_v) receives the ghost's offset in the ghosted arrays.
template <class Put>
__device__ inline void bc_fold_ghosts(const RhsArgs& a, int i, int j, Put put) {
  const Geom& G = a.geom;
  const int f = a.bc_fold;
  const bool lo0 = i == 0, hi0 = i == G.nx - 1, lo1 = j == 0, hi1 = j == G.ny - 1;
  if (!(lo0 || hi0 || lo1 || hi1)) return;
  const bool ru_lo = f & 2, ru_hi = f & 4, rv_lo = f & 8, rv_hi = f & 16;
  const long row = (long)(j + G.g1) * G.sy, col = i + G.g0;
  if (lo0)
    for (int l = 0; l < G.g0; l++) put(row + G.g0 - 1 - l, ru_lo, false);
  if (hi0)
    for (int l = 0; l < G.g0; l++) put(row + G.g0 + G.nx + l, ru_hi, false);
  if (lo1)
    for (int l = 0; l < G.g1; l++) put((long)(G.g1 - 1 - l) * G.sy + col, false, rv_lo);
  if (hi1)
    for (int l = 0; l < G.g1; l++) put((long)(G.g1 + G.ny + l) * G.sy + col, false, rv_hi);
  for (int cx = 0; cx < 2; cx++)
    for (int cy = 0; cy < 2; cy++) {
      if (!(cx ? hi0 : lo0) || !(cy ? hi1 : lo1)) continue;
      for (int lj = 0; lj < G.g1; lj++)
        for (int li = 0; li < G.g0; li++) {
          const int gi = cx ? G.g0 + G.nx + li : G.g0 - 1 - li;
          const int gj = cy ? G.g1 + G.ny + lj : G.g1 - 1 - lj;
          put((long)gi + (long)gj * G.sy, cx ? ru_hi : ru_lo, cy ? rv_hi : rv_lo);
        }
    }
}

template <int FUSED>
__device__ inline void cell_output(const RhsArgs& a, int i, int j, const double K[5]) {
  const Geom& G = a.geom;
  const long c = (long)i + (long)j * G.nx;
  const long ik = (long)G.nx * G.ny;
  if constexpr (!FUSED) {
#pragma unroll
    for (int k = 0; k < 5; k++) a.KK[c + k * ik] = K[k];
  } else {
    const long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
    // the ten loads before the first store, so that they are in flight
    // together (w_out aliases neither, which the compiler cannot know)
    double wa[5], wb[5], wn[5];
    rk_operands(a, i, j, wa, wb);
#pragma unroll
    for (int k = 0; k < 5; k++) {
      wn[k] = a.alpha_1 * wa[k] + a.alpha_2 * wb[k] + a.bdt * K[k];
      a.w_out[g + k * G.sk] = wn[k];
    }
    const double xc1 = a.ylo + a.dy * ((double)(j) + 0.5);
    double p, T;
    eos_p_T_nr(a.c, a.c_lr, a.rho0, a.gamma1, a.c_p, a.p0, a.T0, a.c_v, a.R, a.eos_k1, a.eos_ok, xc1, wn[0], wn[3], p,
               T);
    a.p_out[g] = p;
    a.T_out[g] = T;
    if (FUSED == kStageBc && a.bc_fold)
      bc_fold_ghosts(a, i, j, [&](long d, bool neg_u, bool neg_v) {
        a.w_out[d] = wn[0];
        a.w_out[d + G.sk] = neg_u ? -wn[1] : wn[1];
        a.w_out[d + 2 * G.sk] = neg_v ? -wn[2] : wn[2];
        a.w_out[d + 3 * G.sk] = wn[3];
        a.w_out[d + 4 * G.sk] = wn[4];
        a.p_out[d] = p;
        a.T_out[d] = T;
      });
  }
}

// Physical source H, rhs_hybrid_high_llf.m4:159-237.  The thirteen
// quotients by tau_n, T_n, temp, 2*temp and temp^2 are formed as n / d
// (NR false) or as n times one reciprocal per denominator (NR true: n *
// rcp_nr(d), within 1.5 ulp of n / d where d passes den_ok); the NR form
// returns false, its H to be discarded, where T_n or temp^2 does not.
template <int S, bool P, bool NR>
__device__ inline bool cell_source_q(const Tile& t, int e, double H0, const RhsArgs& a, double H[5]) {
#pragma clang fp contract(fast)
  const double dx = a.dx, dy = a.dy;
  const double* U = t.u;
  const double* V = t.v;
  const double* TT = t.T;
  const double dx2 = sq(dx), dy2 = sq(dy), dxy4 = 4 * dx * dy, dx_2 = 2 * dx, dy_2 = 2 * dy;
  H[0] = H0;

  double u_dprime_x = divc<P>(U[e - 1] - 2 * U[e] + U[e + 1], dx2, a.inv_dx2);
  double u_dprime_y = divc<P>(U[e - S] - 2 * U[e] + U[e + S], dy2, a.inv_dy2);
  double v_dprime_xy = divc<P>(V[e + 1 + S] - V[e + 1 - S] - V[e - 1 + S] + V[e - 1 - S], dxy4, a.inv_4dxdy);
  double phi_xx_prime_x = kC4_3 * u_dprime_x - kC2_3 * v_dprime_xy;
  double phi_xy_prime_y = u_dprime_y + v_dprime_xy;
  H[1] = a.muB * (u_dprime_x + v_dprime_xy) + a.mu * (phi_xx_prime_x + phi_xy_prime_y);

  double v_dprime_y = divc<P>(V[e - S] - 2 * V[e] + V[e + S], dy2, a.inv_dy2);
  double v_dprime_x = divc<P>(V[e - 1] - 2 * V[e] + V[e + 1], dx2, a.inv_dx2);
  double u_dprime_yx = divc<P>(U[e + 1 + S] - U[e + 1 - S] - U[e - 1 + S] + U[e - 1 - S], dxy4, a.inv_4dxdy);
  double phi_yx_prime_x = u_dprime_yx + v_dprime_x;
  double phi_yy_prime_y = kC4_3 * v_dprime_y - kC2_3 * u_dprime_yx;
  H[2] = a.muB * (v_dprime_y + u_dprime_yx) + a.mu * (phi_yx_prime_x + phi_yy_prime_y);

  double u_prime_x = divc<P>(U[e + 1] - U[e - 1], dx_2, a.inv_2dx);
  double u_prime_y = divc<P>(U[e + S] - U[e - S], dy_2, a.inv_2dy);
  double v_prime_y = divc<P>(V[e + S] - V[e - S], dy_2, a.inv_2dy);
  double v_prime_x = divc<P>(V[e + 1] - V[e - 1], dx_2, a.inv_2dx);
  double t_prime_x = divc<P>(TT[e + 1] - TT[e - 1], dx_2, a.inv_2dx);
  double t_prime_y = divc<P>(TT[e + S] - TT[e - S], dy_2, a.inv_2dy);
  double t_dprime_x = divc<P>(TT[e - 1] - 2 * TT[e] + TT[e + 1], dx2, a.inv_dx2);
  double t_dprime_y = divc<P>(TT[e - S] - 2 * TT[e] + TT[e + S], dy2, a.inv_dy2);
  double w0 = t.w0[e];
  double temp = TT[e];
  double T_n = t.tn[e];
  const double temp2 = sq(temp);
  const double nx1 = a.kappa * sq(t_prime_x), nx2 = a.kappa * t_dprime_x;
  const double ny1 = a.kappa * sq(t_prime_y), ny2 = a.kappa * t_dprime_y;
  bool ok = true;
  double rtau = 0, rTn = 0, rT = 0, r2T = 0, rT2 = 0;
  if constexpr (NR) {
    // |temp^2| in [2^-50, 2^50] puts temp and 2*temp in range too; tau_n
    // passed den_ok on the host (fast_div)
    ok = den_ok(temp2) && den_ok(T_n);
    rtau = rcp_nr(a.tau_n);
    rTn = rcp_nr(T_n);
    rT = rcp_nr(temp);
    r2T = rcp_nr(2 * temp);
    rT2 = rcp_nr(temp2);
  }
  auto dv = [](double n, double d, double r) { return NR ? n * r : n / d; };
  double Mn = dv(temp - T_n, a.tau_n, rtau);
  double tmp1 = dv(a.T_star_n, T_n, rTn);
  double c_vn = kC078 * a.R * sq(tmp1) * exp(-tmp1);
  double A_n = (dv(temp, T_n, rTn) - 1) * c_vn;
  double phi_xx = kC4_3 * u_prime_x - kC2_3 * v_prime_y;
  double phi_xy_yx = u_prime_y + v_prime_x;
  double phi_yy = kC4_3 * v_prime_y - kC2_3 * u_prime_x;
  double sigma = dv(a.muB, temp, rT) * sq(u_prime_x + v_prime_y) +
                 dv(a.mu, 2 * temp, r2T) * (sq(phi_xx) + 2 * sq(phi_xy_yx) + sq(phi_yy)) +
                 dv(a.kappa, temp2, rT2) * (sq(t_prime_x) + sq(t_prime_y)) + dv(w0, temp, rT) * (A_n * Mn);
  H[3] = sigma - dv(w0, T_n, rTn) * c_vn * Mn +
         (-dv(nx1, temp2, rT2) + dv(nx2, temp, rT) - dv(ny1, temp2, rT2) + dv(ny2, temp, rT));
  H[4] = dv(w0, a.tau_n, rtau) * (temp - T_n);
  return ok;
}

template <int S, bool P>
__device__ inline void cell_source(const Tile& t, int e, double H0, const RhsArgs& a, double H[5]) {
  if (a.fast_div && cell_source_q<S, P, true>(t, e, H0, a, H)) return;
  cell_source_q<S, P, false>(t, e, H0, a, H);
}

// r = (2|t1 t2| + ep) / (t1^2 + t2^2 + ep) < float(0.9999)?  The numerator
// and denominator are formed exactly as in rhs_hybrid_high_llf.m4:246; the
// comparison of their IEEE quotient with the threshold is decided from a
// single-precision estimate (relative error < 3e-7) whenever that estimate
// is more than 1e-6 away from the threshold, else from the exact quotient.
__device__ inline bool ratio_below(double tmp1, double tmp2, double ep) {
  double num = 2 * fabs(tmp1 * tmp2) + ep;
  double den = sq(tmp1) + sq(tmp2) + ep;
  // den >= num up to rounding (t1^2 + t2^2 >= 2|t1 t2|), so these two bounds
  // keep both within a few ulps of [1e-30, 1e30], normal in single precision
  if (num > 1e-30 && den < 1e30) {
    float q = (float)num * __builtin_amdgcn_rcpf((float)den);
    if (q < (float)(kThresh * (1 - 1e-6))) return true;
    if (q > (float)(kThresh * (1 + 1e-6))) return false;
  }
  return num / den < kThresh;
}

// Hybrid density sensor, rhs_hybrid_high_llf.m4:239-256.  1 = WENO.  The
// triples of q = -1, 0, 1 share four density differences per direction:
// d[m] = rho(i-1+m) - rho(i-2+m), the triple of q being (d[q+1], d[q+2]),
// the reference's tmp1, tmp2 (same subtraction, same bits).  When the
// squares of all eight sum below ep_pre = 5e-5 * ep (ep > 0), every ratio
// (2|t1 t2| + ep) / (t1^2 + t2^2 + ep) is at least 1 / (1 + 5e-5), rounding
// included, far above float(0.9999): the cell is DRP and the six ratios
// need not be formed.  Otherwise they decide as in the reference.
template <int S>
__device__ inline int cell_sensor(const double* r, int e, double ep, double ep_pre, int level) {
  if (level != 0) return 1;
  double dx[4], dy[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    dx[m] = r[e - 1 + m] - r[e - 2 + m];
    dy[m] = r[e + (m - 1) * S] - r[e + (m - 2) * S];
  }
  {
#pragma clang fp contract(fast)
    double s = 0;
#pragma unroll
    for (int m = 0; m < 4; m++) s += dx[m] * dx[m] + dy[m] * dy[m];
    if (s < ep_pre) return 0;
  }
  int tag = 0;
#pragma unroll
  for (int q = -1; q <= 1; q++) {
    if (ratio_below(dx[q + 1], dx[q + 2], ep)) tag = 1;
    if (ratio_below(dy[q + 1], dy[q + 2], ep)) tag = 1;
  }
  return tag;
}

// DRP branch, rhs_hybrid_high_llf.m4:258-268, without the source:
// out[k] = (F_x)/(-dx) + (G_y)/(-dy); the caller adds H[k] (the reference
// sums the three in that order).  The x stencil is finished before the y
// stencil's loads are issued (sched_barrier), which keeps the twelve
// points' values from being live at once.
template <int S, bool P>
__device__ inline void cell_drp(const Tile& t, int e, const RhsArgs& a, double out[5]) {
#pragma clang fp contract(fast)
  const double mdx = -a.dx, mdy = -a.dy;
  double fx[5];
#pragma unroll
  for (int k = 0; k < 5; k++)
    fx[k] = -kA3 * tflux_f(t, e - 3, k) - kA2 * tflux_f(t, e - 2, k) - kA1 * tflux_f(t, e - 1, k) +
            kA1 * tflux_f(t, e + 1, k) + kA2 * tflux_f(t, e + 2, k) + kA3 * tflux_f(t, e + 3, k);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double gy = -kA3 * tflux_g(t, e - 3 * S, k) - kA2 * tflux_g(t, e - 2 * S, k) - kA1 * tflux_g(t, e - S, k) +
                kA1 * tflux_g(t, e + S, k) + kA2 * tflux_g(t, e + 2 * S, k) + kA3 * tflux_g(t, e + 3 * S, k);
    out[k] = divc<P>(fx[k], mdx, a.inv_mdx) + divc<P>(gy, mdy, a.inv_mdy);
  }
}

// One direction of the WENO branch for component k: split fluxes at the
// 2R+1 points and the four reconstructions pf1, mf1, pf2, mf2
// (rhs_hybrid_high_llf.m4:286-354); returns pf1 + mf1 - pf2 - mf2.
template <int ORDER>
__device__ inline double weno_dir(const double* F, const double* W, double alpha) {
#pragma clang fp contract(fast)
  double fp[7], fm[7];
  constexpr int R = ORDER == 5 ? 3 : 2;
#pragma unroll
  for (int m = -R; m <= R; m++) {
    fp[m + 3] = 0.5 * (F[m + 3] + alpha * W[m + 3]);
    fm[m + 3] = 0.5 * (F[m + 3] - alpha * W[m + 3]);
  }
  double p1, m1, p2, m2;
  if constexpr (ORDER == 5) {
    p1 = wenop5(fp[0], fp[1], fp[2], fp[3], fp[4]);
    m1 = wenom5(fm[1], fm[2], fm[3], fm[4], fm[5]);
    p2 = wenop5(fp[1], fp[2], fp[3], fp[4], fp[5]);
    m2 = wenom5(fm[2], fm[3], fm[4], fm[5], fm[6]);
  } else {
    p1 = wenop3(fp[1], fp[2], fp[3]);
    m1 = wenom3(fm[2], fm[3], fm[4]);
    p2 = wenop3(fp[2], fp[3], fp[4]);
    m2 = wenom3(fm[3], fm[4], fm[5]);
  }
  return p1 + m1 - p2 - m2;
}

// Per-cell LLF wave speeds, rhs_hybrid_high_llf.m4:270-284: alpha_f over
// |u| in the x window at this row's height, alpha_g over |v| in the y
// window at each point's height.  `u` and `v` hold the window (index m+3).
template <int REACH>
__device__ inline void llf_alpha(const double* u, const double* v, int j, const RhsArgs& a, double& af,
                                 double& ag) {
#pragma clang fp contract(fast)
  af = 0;
  ag = 0;
#pragma unroll
  for (int q = -REACH; q <= REACH; q++) {
    double height = a.ylo + a.dy * ((double)(j) + 0.5);
    double vf = fabs(u[q + 3]) + a.c + a.c_lr * height;
    if (vf > af) af = vf;
    height = a.ylo + a.dy * ((double)(j + q) + 0.5);
    double vg = fabs(v[q + 3]) + a.c + a.c_lr * height;
    if (vg > ag) ag = vg;
  }
}

// WENO branch of one cell out of an LDS tile that holds the raw w_k.
template <int V, int S, bool P>
__device__ inline void cell_weno_tile(const Tile& t, int e, int j, const RhsArgs& a, const double H[5],
                                      double out[5]) {
#pragma clang fp contract(fast)
  using Sc = Scheme<V>;
  constexpr int R = Sc::order == 5 ? 3 : 2;
  double af, ag;
  if constexpr (Sc::llf) {
    double uu[7], vv[7];
#pragma unroll
    for (int q = -R; q <= R; q++) {
      uu[q + 3] = t.u[e + q];
      vv[q + 3] = t.v[e + q * S];
    }
    llf_alpha<R>(uu, vv, j, a, af, ag);
  } else {
    af = a.alpha;
    ag = a.alpha;
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double F[7], Wx[7], G[7], Wy[7];
#pragma unroll
    for (int m = -R; m <= R; m++) {
      F[m + 3] = tflux_f(t, e + m, k);
      Wx[m + 3] = t.w[k][e + m];
      G[m + 3] = tflux_g(t, e + m * S, k);
      Wy[m + 3] = t.w[k][e + m * S];
    }
    double x = weno_dir<Sc::order>(F, Wx, af);
    double y = weno_dir<Sc::order>(G, Wy, ag);
    out[k] = divc<P>(x, a.dx, a.inv_dx) + divc<P>(y, a.dy, a.inv_dy) + H[k];
  }
}

// One thread's share of the staging loads of a tile + halo (stage_load),
// held in registers until stage_store writes them to LDS.
template <int BX, int BY>
struct StageRegs {
  static constexpr int kIt = (TileShape<BX, BY>::TN + TileShape<BX, BY>::threads - 1) / TileShape<BX, BY>::threads;
  double v[kIt][7];  // w_n(5), p, T
  bool in[kIt];
};

// The global loads of the tile + halo at (i0, j0), all issued before any is
// used.  A cell outside the tile or the ghost box reads element 0 (always in
// the arrays) and is marked not to be stored; no interior cell of the patch
// reads it.
// Loads slots [IT0, IT1) of the thread's share (all of them by default).
template <int BX, int BY, int IT0 = 0, int IT1 = StageRegs<BX, BY>::kIt>
__device__ inline void stage_load(const RhsArgs& a, int i0, int j0, StageRegs<BX, BY>& R) {
  using Sh = TileShape<BX, BY>;
  const Geom& G = a.geom;
  const int tid = threadIdx.x + threadIdx.y * BX;
  // the whole tile and halo inside the ghost box (every tile of a grid
  // whose sides are multiples of the tile's, with g >= halo): no per-cell
  // bounds test
  const bool inside = i0 - Sh::halo >= -G.g0 && i0 + Sh::TW - Sh::halo <= G.nx + G.g0 &&
                      j0 - Sh::halo >= -G.g1 && j0 + Sh::TH - Sh::halo <= G.ny + G.g1;
#pragma unroll
  for (int it = IT0; it < IT1; it++) {
    const int e = tid + it * Sh::threads;
    const int gi = i0 + (e % Sh::TW) - Sh::halo, gj = j0 + (e / Sh::TW) - Sh::halo;
    R.in[it] = e < Sh::TN && (inside || (gi >= -G.g0 && gi < G.nx + G.g0 && gj >= -G.g1 && gj < G.ny + G.g1));
    const long gidx = R.in[it] ? (long)(gi + G.g0) + (long)(gj + G.g1) * G.sy : 0;
#pragma unroll
    for (int k = 0; k < 5; k++) R.v[it][k] = a.w[gidx + k * G.sk];
    R.v[it][5] = a.p[gidx];
    R.v[it][6] = a.T[gidx];
  }
}

// rho, p, T and the quotients (and the raw w_k if RAW) of the loaded cells
// into the LDS tile.
template <int BX, int BY, bool RAW>
__device__ inline void stage_store(const StageRegs<BX, BY>& R, TileLds<BX, BY>& L) {
  using Sh = TileShape<BX, BY>;
  const int tid = threadIdx.x + threadIdx.y * BX;
  double *s0 = L.w[0], *sp = L.p, *sT = L.T, *su = L.u, *sv = L.v, *ss = L.s, *stn = L.tn;
  const auto& v = R.v;
#pragma unroll
  for (int it = 0; it < R.kIt; it++) {
    if (!R.in[it]) continue;
    const int e = tid + it * Sh::threads;
    const double w0 = v[it][0], w1 = v[it][1], w2 = v[it][2], w3 = v[it][3], w4 = v[it][4];
    s0[e] = w0;
    sp[e] = v[it][5];
    sT[e] = v[it][6];
    // the four quotients through one reciprocal of w0 (two Newton steps,
    // within an ulp of 1/w0; the IEEE reciprocal outside den_ok's range)
    const double iw0 = den_ok(w0) ? rcp_nr(w0) : 1.0 / w0;
    su[e] = w1 * iw0;
    sv[e] = w2 * iw0;
    ss[e] = w3 * iw0;
    stn[e] = w4 * iw0;
    if constexpr (RAW) {
      L.w[1][e] = w1;
      L.w[2][e] = w2;
      L.w[3][e] = w3;
      L.w[4][e] = w4;
    }
  }
}

// Stage rho, p, T and the quotients (and the raw w_k if RAW) of the tile +
// halo.  Cells outside the ghosted array are skipped.
template <int BX, int BY, bool RAW>
__device__ inline void stage_tile(const RhsArgs& a, int i0, int j0, TileLds<BX, BY>& L) {
  StageRegs<BX, BY> R;
  stage_load(a, i0, j0, R);
  stage_store<BX, BY, RAW>(R, L);
}

// The per-cell inputs of a split tile that come from global memory besides
// the staged tile: the source H0 and, where a switch was decided elsewhere,
// its decision (2: the sensor decides).  Loaded apart so that a kernel can
// issue them before prefetching its next tile.
struct SplitIn {
  double h0;
  int d;
};
template <int BX, int BY>
__device__ inline SplitIn split_inputs(const RhsArgs& a, int i0, int j0) {
  const Geom& G = a.geom;
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  SplitIn s = {0.0, 2};
  if (i < G.nx && j < G.ny) {
    s.h0 = a.H0[(long)i + (long)j * a.h0_sy];
    if (a.decide) s.d = (int)a.decide[(long)i + (long)j * G.nx];
  }
  return s;
}

// Split tile: source H and sensor for every cell, DRP cells finished here,
// WENO cells (sensor on level 0; every cell on finer levels, where the
// reference switches the sensor off, and for rhs_weno_*) keep H in their rhs
// slots and are appended to the list as (patch << 32 | cell): one atomic per
// wave.  split_compute runs on a staged tile.
template <int V, bool P, int FUSED, int BX, int BY>
__device__ inline void split_compute(const RhsArgs& a, int i0, int j0, unsigned pid, TileLds<BX, BY>& L,
                                     const SplitIn& in);
template <int V, bool P, int FUSED, int BX, int BY>
__device__ inline void split_body(const RhsArgs& a, int i0, int j0, unsigned pid, TileLds<BX, BY>& L) {
  stage_tile<BX, BY, false>(a, i0, j0, L);
  const SplitIn in = split_inputs<BX, BY>(a, i0, j0);
  __syncthreads();
  split_compute<V, P, FUSED, BX, BY>(a, i0, j0, pid, L, in);
}
template <int V, bool P, int FUSED, int BX, int BY>
__device__ inline void split_compute(const RhsArgs& a, int i0, int j0, unsigned pid, TileLds<BX, BY>& L,
                                     const SplitIn& in) {
  using Sc = Scheme<V>;
  using Sh = TileShape<BX, BY>;
  const Geom& G = a.geom;
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const bool valid = i < G.nx && j < G.ny;
  int tag = 0;
  if (valid) {
    Tile t;
    t.w0 = L.w[0];
    t.p = L.p;
    t.T = L.T;
    t.u = L.u;
    t.v = L.v;
    t.s = L.s;
    t.tn = L.tn;
    const int e = (threadIdx.y + Sh::halo) * Sh::TW + threadIdx.x + Sh::halo;
    const long icell = (long)i + (long)j * G.nx;
    const long ik = (long)G.nx * G.ny;
    if constexpr (Sc::hybrid) {
      tag = in.d == 2 ? cell_sensor<Sh::TW>(t.w0, e, a.ep, a.ep_pre, a.level) : in.d;
    } else
      tag = Sc::drp ? 0 : 1;
    // Hybrid DRP cells: the stencil part first, so that H is not live
    // across it (the DRP-only variant keeps source then stencil)
    double out[5] = {0, 0, 0, 0, 0};
    double H[5];
    if constexpr (Sc::hybrid) {
      if (tag == 0) cell_drp<Sh::TW, P>(t, e, a, out);
      __builtin_amdgcn_sched_barrier(0);
      cell_source<Sh::TW, P>(t, e, in.h0, a, H);
    } else {
      cell_source<Sh::TW, P>(t, e, in.h0, a, H);
      if (tag == 0) cell_drp<Sh::TW, P>(t, e, a, out);
    }
    if (tag == 0) {
#pragma unroll
      for (int k = 0; k < 5; k++) out[k] = out[k] + H[k];
      cell_output<FUSED>(a, i, j, out);
    } else {
#pragma unroll
      for (int k = 0; k < 5; k++) a.KK[icell + k * ik] = H[k];
    }
    if (a.switch_map) a.switch_map[icell] = tag;
  }
  if constexpr (Sc::weno) {
    unsigned long long mask = __ballot(tag);
    if (mask) {
      int lane = __lane_id();
      int leader = __ffsll((unsigned long long)mask) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(a.count, (unsigned)__popcll(mask));
      base = __shfl(base, leader);
      if (tag)
        a.list[base + __popcll(mask & __lanemask_lt())] =
            ((unsigned long long)pid << 32) | (unsigned)(i + j * G.nx);
    }
  }
}

// One listed WENO cell: stencil from global memory, direction by direction,
// then rhs = X + Y + H (H waits in the rhs slots).  Per direction only rho
// and the normal velocity of the 2R+1 points stay live across the five
// components; the other quotient a component's flux needs (v, s or T_n for
// F; u, s or T_n for G) is used by that component alone and formed there.
template <int V, bool P, int FUSED>
__device__ inline double list_dir(const RhsArgs& a, long g, int j, int dir, int k, const double* w0,
                                  const double* un, double alpha) {
#pragma clang fp contract(fast)
  using Sc = Scheme<V>;
  constexpr int R = Sc::order == 5 ? 3 : 2;
  const Geom& G = a.geom;
  const long step = dir == 0 ? 1 : G.sy;
  double F[7], W[7];
#pragma unroll
  for (int m = -R; m <= R; m++) {
    const long gm = g + m * step;
    const double r = w0[m + 3], q = un[m + 3];
    W[m + 3] = a.w[gm + k * G.sk];
    if (k == 0) {
      F[m + 3] = r * q;
    } else if ((k == 1 && dir == 0) || (k == 2 && dir == 1)) {
      F[m + 3] = r * sq(q) + a.p[gm];  // normal momentum flux
    } else {
      // transverse momentum, entropy or N2 energy: (rho*q) * (w_k / rho)
      F[m + 3] = r * q * (W[m + 3] / r);
    }
  }
  double d = weno_dir<Sc::order>(F, W, alpha);
  return dir == 0 ? divc<P>(d, a.dx, a.inv_dx) : divc<P>(d, a.dy, a.inv_dy);
}

// One listed WENO cell: stencil from global memory, direction by direction,
// then rhs = X + Y + H (H waits in the rhs slots).  Per direction only rho
// and the normal velocity of the 2R+1 points stay live across the five
// components; the other quotient a component's flux needs (v, s or T_n for
// F; u, s or T_n for G) is used by that component alone and formed there.
// The component loop is kept rolled to bound register use.
template <int V, bool P, int FUSED>
__device__ inline void list_cell(const RhsArgs& a, int cell) {
#pragma clang fp contract(fast)
  using Sc = Scheme<V>;
  constexpr int R = Sc::order == 5 ? 3 : 2;
  const Geom& G = a.geom;
  const long ik = (long)G.nx * G.ny;
  const int i = cell % G.nx, j = cell / G.nx;
  const long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
  double X[5], Y[5];
#pragma unroll
  for (int dir = 0; dir < 2; dir++) {
    const long step = dir == 0 ? 1 : G.sy;
    const long kn = dir == 0 ? G.sk : 2 * G.sk;  // component of the normal momentum
    double w0[7], un[7];
#pragma unroll
    for (int m = -R; m <= R; m++) {
      long gm = g + m * step;
      w0[m + 3] = a.w[gm];
      un[m + 3] = a.w[gm + kn] / w0[m + 3];
    }
    double alpha = a.alpha;
    if constexpr (Sc::llf) {
      // the x window gives alpha_f (height of row j), the y window alpha_g
      alpha = 0;
#pragma unroll
      for (int m = -R; m <= R; m++) {
        double height = a.ylo + a.dy * ((double)(dir == 0 ? j : j + m) + 0.5);
        double vv = fabs(un[m + 3]) + a.c + a.c_lr * height;
        if (vv > alpha) alpha = vv;
      }
    }
    double r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
#pragma unroll 1
    for (int k = 0; k < 5; k++) {
      double d = list_dir<V, P, FUSED>(a, g, j, dir, k, w0, un, alpha);
      r0 = k == 0 ? d : r0;
      r1 = k == 1 ? d : r1;
      r2 = k == 2 ? d : r2;
      r3 = k == 3 ? d : r3;
      r4 = k == 4 ? d : r4;
    }
    double* D = dir == 0 ? X : Y;
    D[0] = r0;
    D[1] = r1;
    D[2] = r2;
    D[3] = r3;
    D[4] = r4;
  }
  double K[5];
#pragma unroll
  for (int k = 0; k < 5; k++) K[k] = X[k] + Y[k] + a.KK[(long)cell + k * ik];  // the slot holds H(k)
  cell_output<FUSED>(a, i, j, K);
}

// One (direction, component) of a listed WENO cell: the quantity list_cell
// forms as X[k] (dir 0) or Y[k] (dir 1), by the same operations.  The five
// lanes of one (cell, direction) share the window: lane k loads points k and
// k + 5 (rho and the normal velocity, one quotient each) and the others
// reach them by wavefront shuffles; the LLF speed is each lane's maximum
// over its own points, then the maximum over the five lanes by shuffles
// (rhs_hybrid_high_llf.m4:270-284: the max over the window, order-free).
// gbase is the wave lane of the group's component 0; the group's lanes are
// active together.
template <int V, bool P>
__device__ inline double list_item(const RhsArgs& a, int cell, int dir, int k, int gbase) {
#pragma clang fp contract(fast)
  using Sc = Scheme<V>;
  constexpr int R = Sc::order == 5 ? 3 : 2;
  const Geom& G = a.geom;
  const int i = cell % G.nx, j = cell / G.nx;
  const long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
  const long step = dir == 0 ? 1 : G.sy;
  const long kn = dir == 0 ? G.sk : 2 * G.sk;
  // this lane's points of the window: pi = k and pi = k + 5 (point m = pi - R)
  const bool two = k + 5 <= 2 * R;
  double wa, ua, wb = 0.0, ub = 0.0;
  {
    const long ga = g + (long)(k - R) * step;
    wa = a.w[ga];
    ua = a.w[ga + kn] / wa;
    if (two) {
      const long gb = g + (long)(k + 5 - R) * step;
      wb = a.w[gb];
      ub = a.w[gb + kn] / wb;
    }
  }
  double w0[7], un[7];
#pragma unroll
  for (int pi = 0; pi <= 2 * R; pi++) {
    const int src = gbase + pi % 5;
    w0[pi - R + 3] = __shfl(pi < 5 ? wa : wb, src);
    un[pi - R + 3] = __shfl(pi < 5 ? ua : ub, src);
  }
  double alpha = a.alpha;
  if constexpr (Sc::llf) {
    // the x window at row j's height, the y window at each point's height
    auto speed = [&](int pi, double u) {
      const double height = a.ylo + a.dy * ((double)(dir == 0 ? j : j + pi - R) + 0.5);
      return fabs(u) + a.c + a.c_lr * height;
    };
    double mine = 0;
    const double sa = speed(k, ua);
    if (sa > mine) mine = sa;
    if (two) {
      const double sb = speed(k + 5, ub);
      if (sb > mine) mine = sb;
    }
    alpha = 0;
#pragma unroll
    for (int d = 0; d < 5; d++) {
      const double o = __shfl(mine, gbase + d);
      if (o > alpha) alpha = o;
    }
  }
  return list_dir<V, P, false>(a, g, j, dir, k, w0, un, alpha);
}

// The WENO list, ten lanes per cell (two directions x five components) so a
// short list still spreads over the machine.  A wave holds six cells (lanes
// 0..59; a cell's ten lanes never straddle two waves, so they can shuffle):
// lane (dir, k) forms its term, lane (0, k) takes the y term of lane (1, k)
// by a shuffle, forms rhs_k = X[k] + Y[k] + H[k] (list_cell's order) and
// writes component k the way cell_output does; in the fused form lane (0,
// 0) takes the new rho*s from lane (0, 3) for p and T.  The waves of a
// block work apart (no LDS, no barrier).  Entries are (patch << 32) | (i +
// j*nx); args_of(patch) gives the patch's arguments.  A wave serves its
// lanes patch by patch (consecutive entries come from one tile: usually a
// single pass) so the arguments stay uniform.
constexpr int kListCellsPerWave = 6;
constexpr int kListCellsPerBlock = 4 * kListCellsPerWave;

template <int V, bool P, int FUSED, class ArgsOf>
__device__ inline void list_block(ArgsOf args_of, const unsigned long long* list, unsigned n, bool fold = false) {
  const int t = threadIdx.x;
  const int lane = t & 63, slot = lane / 10, item = lane % 10, k = item % 5;
  const int lc = (t >> 6) * kListCellsPerWave + slot;
  const int gbase = lane - k;  // lane (dir, 0) of this lane's group
  for (unsigned c0 = blockIdx.x * kListCellsPerBlock; c0 < n; c0 += gridDim.x * kListCellsPerBlock) {
    const unsigned q = c0 + lc;
    bool todo = slot < kListCellsPerWave && q < n;
    const unsigned long long e = todo ? list[q] : 0ull;
    const int pid = (int)(e >> 32);
    for (;;) {
      const unsigned long long m = __ballot(todo);
      if (!m) break;
      const int cur = __builtin_amdgcn_readfirstlane(__shfl(pid, __ffsll((long long)m) - 1));
      if (todo && pid == cur) {
        const RhsArgs& a = args_of(cur);
        const int cell = (int)(e & 0xffffffffu);
        const double term = list_item<V, P>(a, cell, item / 5, k, gbase);
        const double y = __shfl(term, item < 5 ? lane + 5 : lane);
        if (item < 5) {
          const Geom& G = a.geom;
          const long ik = (long)G.nx * G.ny;
          const double K = term + y + a.KK[(long)cell + k * ik];
          const int i = cell % G.nx, j = cell / G.nx;
          if constexpr (!FUSED) {
            a.KK[(long)cell + k * ik] = K;
          } else {
            const long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
            const double wa = a.ww[a.ww_off + i + j * a.ww_sy + k * a.ww_sk];
            const double wb = a.w[g + k * G.sk];
            const double wn = a.alpha_1 * wa + a.alpha_2 * wb + a.bdt * K;
            a.w_out[g + k * G.sk] = wn;
            const double wn3 = __shfl(wn, gbase + 3);
            if (k == 0) {
              const double xc1 = a.ylo + a.dy * ((double)(j) + 0.5);
              double p, T;
              eos_p_T_nr(a.c, a.c_lr, a.rho0, a.gamma1, a.c_p, a.p0, a.T0, a.c_v, a.R, a.eos_k1, a.eos_ok, xc1, wn,
                         wn3, p, T);
              a.p_out[g] = p;
              a.T_out[g] = T;
            }
          }
        }
        todo = false;
      }
    }
  }
  if constexpr (FUSED == kStageBc) {
    // The boundary fold (fold: the arguments' bc_fold is set), as a second
    // pass so that it adds no live registers to the one above: each lane
    // (0, k) revisits its cells and copies the value it stored there
    // (component k; p and T on lane k = 0) into the ghosts.
    if (fold && item < 5 && slot < kListCellsPerWave)
      for (unsigned q = blockIdx.x * kListCellsPerBlock + lc; q < n; q += gridDim.x * kListCellsPerBlock) {
        const unsigned long long e = list[q];
        const RhsArgs& a = args_of((int)(e >> 32));
        const Geom& G = a.geom;
        const int cell = (int)(e & 0xffffffffu);
        const int i = cell % G.nx, j = cell / G.nx;
        if (i != 0 && i != G.nx - 1 && j != 0 && j != G.ny - 1) continue;
        const long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
        double* w = a.w_out + (long)k * G.sk;
        const double v = w[g];
        bc_fold_ghosts(a, i, j, [&](long d, bool neg_u, bool neg_v) {
          w[d] = (k == 1 && neg_u) || (k == 2 && neg_v) ? -v : v;
        });
        if (k == 0) {
          const double p = a.p_out[g], T = a.T_out[g];
          bc_fold_ghosts(a, i, j, [&](long d, bool, bool) {
            a.p_out[d] = p;
            a.T_out[d] = T;
          });
        }
      }
  }
}

// Dense tile: every cell on the WENO path.
template <int V, bool P, int FUSED, int BX, int BY>
__device__ inline void dense_body(const RhsArgs& a, int i0, int j0, TileLds<BX, BY>& L) {
  using Sh = TileShape<BX, BY>;
  const Geom& G = a.geom;
  stage_tile<BX, BY, true>(a, i0, j0, L);
  __syncthreads();
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  if (i >= G.nx || j >= G.ny) return;
  Tile t;
  t.w0 = L.w[0];
  for (int k = 0; k < 5; k++) t.w[k] = L.w[k];
  t.p = L.p;
  t.T = L.T;
  t.u = L.u;
  t.v = L.v;
  t.s = L.s;
  t.tn = L.tn;
  const int e = (threadIdx.y + Sh::halo) * Sh::TW + threadIdx.x + Sh::halo;
  double H[5], out[5];
  cell_source<Sh::TW, P>(t, e, a.H0[(long)i + (long)j * a.h0_sy], a, H);
  cell_weno_tile<V, Sh::TW, P>(t, e, j, a, H, out);
  cell_output<FUSED>(a, i, j, out);
  if (a.switch_map) a.switch_map[(long)i + (long)j * G.nx] = 1;
}

// Dense tile on two thread planes: threadIdx.z = 0 forms the source H and
// the x direction of its cell's five components, threadIdx.z = 1 the y
// direction.  The x plane's results meet the y plane's in LDS (xs: X then H,
// 10 x BX*BY doubles) and the y plane forms each component as (X + Y) + H,
// the reference's order, and writes the cell.  One direction per thread
// needs about half the registers of both, so more waves fit a SIMD.
template <int V, bool P, int FUSED, int BX, int BY>
__device__ inline void dense2_body(const RhsArgs& a, int i0, int j0, TileLds<BX, BY>& L, double* xs) {
#pragma clang fp contract(fast)
  using Sh = TileShape<BX, BY>;
  using Sc = Scheme<V>;
  const Geom& G = a.geom;
  constexpr int NT = 2 * Sh::threads;
  constexpr int S = Sh::TW;
  constexpr int R = Sc::order == 5 ? 3 : 2;
  const int tid = threadIdx.x + BX * (threadIdx.y + BY * threadIdx.z);
  {
    // staging by both planes; every load of a thread before its first store
    constexpr int kIt = (Sh::TN + NT - 1) / NT;
    const bool inside = i0 - Sh::halo >= -G.g0 && i0 + Sh::TW - Sh::halo <= G.nx + G.g0 &&
                        j0 - Sh::halo >= -G.g1 && j0 + Sh::TH - Sh::halo <= G.ny + G.g1;
    double v[kIt][7];
    bool in[kIt];
#pragma unroll
    for (int it = 0; it < kIt; it++) {
      const int e = tid + it * NT;
      const int gi = i0 + (e % Sh::TW) - Sh::halo, gj = j0 + (e / Sh::TW) - Sh::halo;
      in[it] = e < Sh::TN && (inside || (gi >= -G.g0 && gi < G.nx + G.g0 && gj >= -G.g1 && gj < G.ny + G.g1));
      const long gidx = in[it] ? (long)(gi + G.g0) + (long)(gj + G.g1) * G.sy : 0;
#pragma unroll
      for (int k = 0; k < 5; k++) v[it][k] = a.w[gidx + k * G.sk];
      v[it][5] = a.p[gidx];
      v[it][6] = a.T[gidx];
    }
#pragma unroll
    for (int it = 0; it < kIt; it++) {
      if (!in[it]) continue;
      const int e = tid + it * NT;
      const double w0 = v[it][0];
      L.w[0][e] = w0;
      L.p[e] = v[it][5];
      L.T[e] = v[it][6];
      const double iw0 = den_ok(w0) ? rcp_nr(w0) : 1.0 / w0;
      L.u[e] = v[it][1] * iw0;
      L.v[e] = v[it][2] * iw0;
      L.s[e] = v[it][3] * iw0;
      L.tn[e] = v[it][4] * iw0;
#pragma unroll
      for (int k = 1; k < 5; k++) L.w[k][e] = v[it][k];
    }
  }
  __syncthreads();
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const bool live = i < G.nx && j < G.ny;
  const int c = threadIdx.x + BX * threadIdx.y;
  const int e = (threadIdx.y + Sh::halo) * S + threadIdx.x + Sh::halo;
  Tile t;
  t.w0 = L.w[0];
  for (int k = 0; k < 5; k++) t.w[k] = L.w[k];
  t.p = L.p;
  t.T = L.T;
  t.u = L.u;
  t.v = L.v;
  t.s = L.s;
  t.tn = L.tn;
  const bool yplane = threadIdx.z != 0;
  double Y[5];
  if (live && !yplane) {
    {
      double H[5];
      cell_source<S, P>(t, e, a.H0[(long)i + (long)j * a.h0_sy], a, H);
#pragma unroll
      for (int k = 0; k < 5; k++) xs[(5 + k) * Sh::threads + c] = H[k];
    }
    // x: LLF speed over the x window at this row's height
    double af;
    if constexpr (Sc::llf) {
      af = 0;
      const double height = a.ylo + a.dy * ((double)(j) + 0.5);
#pragma unroll
      for (int q = -R; q <= R; q++) {
        const double vf = fabs(t.u[e + q]) + a.c + a.c_lr * height;
        if (vf > af) af = vf;
      }
    } else {
      af = a.alpha;
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
      double F[7], W[7];
#pragma unroll
      for (int m = -R; m <= R; m++) {
        F[m + 3] = tflux_f(t, e + m, k);
        W[m + 3] = t.w[k][e + m];
      }
      xs[k * Sh::threads + c] = divc<P>(weno_dir<Sc::order>(F, W, af), a.dx, a.inv_dx);
    }
  } else if (live) {
    // y: LLF speed over the y window at each point's height
    double ag;
    if constexpr (Sc::llf) {
      ag = 0;
#pragma unroll
      for (int q = -R; q <= R; q++) {
        const double height = a.ylo + a.dy * ((double)(j + q) + 0.5);
        const double vg = fabs(t.v[e + q * S]) + a.c + a.c_lr * height;
        if (vg > ag) ag = vg;
      }
    } else {
      ag = a.alpha;
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
      double Gy[7], W[7];
#pragma unroll
      for (int m = -R; m <= R; m++) {
        Gy[m + 3] = tflux_g(t, e + m * S, k);
        W[m + 3] = t.w[k][e + m * S];
      }
      Y[k] = divc<P>(weno_dir<Sc::order>(Gy, W, ag), a.dy, a.inv_dy);
    }
  }
  __syncthreads();
  if (!live || !yplane) return;
  double K[5];
#pragma unroll
  for (int k = 0; k < 5; k++) K[k] = (xs[k * Sh::threads + c] + Y[k]) + xs[(5 + k) * Sh::threads + c];
  cell_output<FUSED>(a, i, j, K);
  if (a.switch_map) a.switch_map[(long)i + (long)j * G.nx] = 1;
}

}  // namespace enl
