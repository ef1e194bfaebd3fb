// enl_batch32.hip -- the FP32-perturbation form of the batched hierarchy.
//
// The state is split into an FP64 ambient that depends on height only and
// a perturbation.  Per level the ambient is one record per row (rho, rho*s,
// rho*T_n given by the caller; p and T from update_state's equation of state,
// update_state.m4:87-107; s = rho*s/rho, T_n = rho*T_n/rho and (c+c_lr*y)^2):
//   * the carried state (the start of each RK step) is the FP64 perturbation
//     W' = w_n - w_amb, interior only;
//   * the RK stage states that the stencils read are FP32 perturbations of
//     w_n (5 components), p and T, ghosted, in three blocks as in the FP64
//     form;
//   * each stage evaluates the rhs of rhs_hybrid_high_llf.m4 in FP32 from the
//     perturbations.  Quantities whose ambient is constant along a stencil
//     use the perturbation only (the constant cancels in every difference and
//     in WENO's shift-invariant reconstruction); along y the ambient
//     differences are formed in FP64 per row and rounded once;
//   * the RK update of the state is accumulated in FP64,
//       W'_new = alpha_1 W' + alpha_2 w'_stage + beta dt (K32 + dK),
//     stored as FP32 for the next stage and, after the last stage, as the
//     FP64 carried state;
//   * dK = K64(ambient) - K32(ambient), computed once by the FP64 kernels and
//     this file's kernel on the ambient alone, keeps the ambient's own
//     tendency (walls, coarse-to-fine refinement of the stratified profile)
//     at its FP64 value and takes the FP32 rounding of the ambient out of the
//     perturbation's tendency.
// Ghost fills and the coarsen act on full values in FP64 (ambient of the
// source row + perturbation, the FP64 form's operator, minus the ambient of
// the destination row), so they are the FP64 form's operators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "enl_batch.h"
#include "enl_device.h"
#include "enl_rhs_body.h"
#include "enl_switch.h"

#pragma clang fp contract(off)

namespace enl {

namespace {

constexpr int kRec = 8;  // ambient record: rho, rho*s, rho*T_n, p, T, s, T_n, (c + c_lr*y)^2

struct Patch32 {
  long blk[3];  // FP32 stage blocks (floats from pool32): w'(5), p', T', 7*sk each
  long res[3];  // their rounding residuals (bytes from res8), interior w'(5), 5*nx*ny each
  long dk;      // FP32 dK (floats from pool32), 5*nx*ny
  long wc;      // FP64 carried perturbation (doubles from wc), 5*nx*ny
  long amb;     // ambient record of ghosted row 0
  int nx, ny, g, level;
  long sy, sk;
};

// Per level row, what the stage kernel needs of the ambient: the FP64
// values of the components whose ambient varies along y, and FP32 values
// and FP64 differences rounded once.
struct RowF {
  double aW[3];  // rho, rho*s, rho*T_n
  double aS64, aTn64;  // s and T_n in FP64: their row differences are formed
                       // in FP64 and rounded once
  float aRho, aS, aTn, aT, aD;  // rho, s, T_n, T, rho - rho0
  float cy2, TmTn, pA;          // (c + c_lr*y)^2, T - T_n, p - p0
  float d1Rho, d1T, d2T, pad;   // rho(r+1)-rho(r), T(r+1)-T(r-1), T(r-1)-2T(r)+T(r+1)
};
static_assert(sizeof(RowF) % 8 == 0, "RowF is copied word by word and holds doubles");

// Ambient value of stored component k (w_n 0..4, p, T) from a row record.
__host__ __device__ inline double amb_of(const double* rec, int k) {
  switch (k) {
    case 0: return rec[0];
    case 3: return rec[1];
    case 4: return rec[2];
    case 5: return rec[3];
    case 6: return rec[4];
    default: return 0.0;  // momenta: the ambient is at rest
  }
}

struct Args32 {
  const float* in;   // stage input block
  float* out;        // stage output block
  const signed char* rin;  // rounding residuals of the input block's interior w'
  signed char* rout;       // and of the output block's
  double* wc;        // carried FP64 perturbation, interior
  const float* dk;   // ambient correction, interior (nullptr in eval mode)
  float* kout;       // eval mode: K32 stored here
  const double* h0;  // source, ghosted
  const double* amb; // record of ghosted row 0
  const RowF* rows;  // row table entry of ghosted row 0
  const unsigned char* decide;  // level 0 with the learned switch: 0 DRP, 1 WENO, 2 sensor
  unsigned char* amb_weno;      // eval mode, level 0: the sensor's choice on the ambient alone
  int nx, ny, g, last;
  long sy, sk;
  float ylo, dy;
  float inv_dx, inv_dy, inv_mdx, inv_mdy, inv_dx2, inv_dy2, inv_4dxdy, inv_2dx, inv_2dy;
  float c, c_lr, mu, muB, kappa, R, T_star_n, inv_tau_n, inv_cp, inv_cv, a_eos, ep;
  double alpha_1, alpha_2, bdt;
  double rho0;
};

__device__ inline float sqf(float x) { return x * x; }
__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// Stage states are stored in FP32, but the RK update combines a stage's
// input (alpha_2 w') with FP64 terms: rounding w' to FP32 there would add a
// fresh error of up to 2^-24 of the perturbation at every stage.  Each
// interior value of w' therefore carries its rounding residual, the FP64
// value less the stored float, as a signed byte in units of 1/254 of the
// float's ulp; float + residual stands for the FP64 value to within
// 2^-24 / 508 of it.
__device__ inline double ulp32(float f) {
  const unsigned e = __float_as_uint(f) & 0x7f800000u;  // 2^E, E the float's exponent
  return e ? (double)__uint_as_float(e) * 0x1p-23 : 0x1p-149;
}
__device__ inline signed char res_enc(double v, float f) {
  if (!__builtin_isfinite(f)) return 0;
  double r = rint((v - (double)f) * (254.0 / ulp32(f)));
  r = r > 127.0 ? 127.0 : (r < -127.0 ? -127.0 : r);
  return (signed char)r;
}
__device__ inline double res_dec(float f, signed char r) { return (double)f + (double)r * (ulp32(f) * (1.0 / 254.0)); }

// WENO5 reconstructions in FP32, the formulas of rhs_hybrid_high_llf.m4:365-433
// (float literals as there), weights through reciprocals.
__device__ inline float wenom5f(float fgm2, float fgm1, float fg0, float fgp1, float fgp2) {
  const float b0 = (13.0f / 12.0f) * sqf(fg0 - 2 * fgp1 + fgp2) + 0.25f * sqf(3 * fg0 - 4 * fgp1 + fgp2);
  const float b1 = (13.0f / 12.0f) * sqf(fgm1 - 2 * fg0 + fgp1) + 0.25f * sqf(fgm1 - fgp1);
  const float b2 = (13.0f / 12.0f) * sqf(fgm2 - 2 * fgm1 + fg0) + 0.25f * sqf(fgm2 - 4 * fgm1 + 3 * fg0);
  const float z0 = 0.1f * rcpf(sqf(10.0e-6f + b0));
  const float z1 = 0.6f * rcpf(sqf(10.0e-6f + b1));
  const float z2 = 0.3f * rcpf(sqf(10.0e-6f + b2));
  const float is = rcpf(z0 + z1 + z2);
  const float f0 = (1.0f / 6.0f) * (11 * fg0 - 7 * fgp1 + 2 * fgp2);
  const float f1 = (1.0f / 6.0f) * (2 * fgm1 + 5 * fg0 - fgp1);
  const float f2 = (1.0f / 6.0f) * (-fgm2 + 5 * fgm1 + 2 * fg0);
  return (z0 * is) * f0 + (z1 * is) * f1 + (z2 * is) * f2;
}
__device__ inline float wenop5f(float fgm2, float fgm1, float fg0, float fgp1, float fgp2) {
  const float b0 = (13.0f / 12.0f) * sqf(fg0 - 2 * fgp1 + fgp2) + 0.25f * sqf(3 * fg0 - 4 * fgp1 + fgp2);
  const float b1 = (13.0f / 12.0f) * sqf(fgm1 - 2 * fg0 + fgp1) + 0.25f * sqf(fgm1 - fgp1);
  const float b2 = (13.0f / 12.0f) * sqf(fgm2 - 2 * fgm1 + fg0) + 0.25f * sqf(fgm2 - 4 * fgm1 + 3 * fg0);
  const float z0 = 0.3f * rcpf(sqf(10.0e-6f + b0));
  const float z1 = 0.6f * rcpf(sqf(10.0e-6f + b1));
  const float z2 = 0.1f * rcpf(sqf(10.0e-6f + b2));
  const float is = rcpf(z0 + z1 + z2);
  const float f0 = (1.0f / 6.0f) * (2 * fg0 + 5 * fgp1 - fgp2);
  const float f1 = (1.0f / 6.0f) * (-fgm1 + 5 * fg0 + 2 * fgp1);
  const float f2 = (1.0f / 6.0f) * (2 * fgm2 - 7 * fgm1 + 11 * fg0);
  return (z0 * is) * f0 + (z1 * is) * f1 + (z2 * is) * f2;
}

// pf1 + mf1 - pf2 - mf2 of one direction (rhs_hybrid_high_llf.m4:286-354),
// from the stencil's differences to its centre point,
// dF[m] = F[m] - F[3] and dW[m] = W[m] - W[3].  The weights see only
// differences and every candidate's coefficients sum to 6 x (1.0/6.0), so
// shifting the seven values by one constant shifts pf1 and pf2 (mf1 and
// mf2) alike and leaves pf1 - pf2 + mf1 - mf2 as it was; with the small
// shifted values FP32 rounds at the size of the differences, not of the
// fluxes.
__device__ inline float weno_dir32s(const float* dF, const float* dW, float alpha) {
  float fp[7], fm[7];
#pragma unroll
  for (int m = 0; m < 7; m++) {
    fp[m] = 0.5f * (dF[m] + alpha * dW[m]);
    fm[m] = 0.5f * (dF[m] - alpha * dW[m]);
  }
  const float p1 = wenop5f(fp[0], fp[1], fp[2], fp[3], fp[4]);
  const float m1 = wenom5f(fm[1], fm[2], fm[3], fm[4], fm[5]);
  const float p2 = wenop5f(fp[1], fp[2], fp[3], fp[4], fp[5]);
  const float m2 = wenom5f(fm[2], fm[3], fm[4], fm[5], fm[6]);
  return (p1 - p2) + (m1 - m2);
}

// FP64 WENO5 with reciprocals refined by one Newton step, for the
// components whose ambient varies along the stencil (rho, rho*s, rho*T_n
// along y): there the LLF term alpha*w carries the stratified ambient, much
// larger than the perturbation, so its reconstruction is kept in FP64.
__device__ inline double rcpd(double x) {
  double r = __builtin_amdgcn_rcp(x);
  return r * (2.0 - x * r);
}
__device__ inline double wenom5d(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
  const double b0 = kC13_12 * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  const double b1 = kC13_12 * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  const double b2 = kC13_12 * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  const double z0 = kW01 * rcpd(sq(kEpsWeno + b0));
  const double z1 = kW06 * rcpd(sq(kEpsWeno + b1));
  const double z2 = kW03 * rcpd(sq(kEpsWeno + b2));
  const double is = rcpd(z0 + z1 + z2);
  const double f0 = kC1_6 * (11 * fg0 - 7 * fgp1 + 2 * fgp2);
  const double f1 = kC1_6 * (2 * fgm1 + 5 * fg0 - fgp1);
  const double f2 = kC1_6 * (-fgm2 + 5 * fgm1 + 2 * fg0);
  return (z0 * is) * f0 + (z1 * is) * f1 + (z2 * is) * f2;
}
__device__ inline double wenop5d(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
  const double b0 = kC13_12 * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  const double b1 = kC13_12 * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  const double b2 = kC13_12 * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  const double z0 = kW03 * rcpd(sq(kEpsWeno + b0));
  const double z1 = kW06 * rcpd(sq(kEpsWeno + b1));
  const double z2 = kW01 * rcpd(sq(kEpsWeno + b2));
  const double is = rcpd(z0 + z1 + z2);
  const double f0 = kC1_6 * (2 * fg0 + 5 * fgp1 - fgp2);
  const double f1 = kC1_6 * (-fgm1 + 5 * fg0 + 2 * fgp1);
  const double f2 = kC1_6 * (2 * fgm2 - 7 * fgm1 + 11 * fg0);
  return (z0 * is) * f0 + (z1 * is) * f1 + (z2 * is) * f2;
}
__device__ inline double weno_dir64(const double* F, const double* W, double alpha) {
  double fp[7], fm[7];
#pragma unroll
  for (int m = 0; m < 7; m++) {
    fp[m] = 0.5 * (F[m] + alpha * W[m]);
    fm[m] = 0.5 * (F[m] - alpha * W[m]);
  }
  return wenop5d(fp[0], fp[1], fp[2], fp[3], fp[4]) + wenom5d(fm[1], fm[2], fm[3], fm[4], fm[5]) -
         wenop5d(fp[1], fp[2], fp[3], fp[4], fp[5]) - wenom5d(fm[2], fm[3], fm[4], fm[5], fm[6]);
}

// LDS of one tile: raw perturbations, derived quantities, per-row ambient.
template <int BX, int BY>
struct Lds32 {
  static constexpr int TW = BX + 6, TH = BY + 6, TN = TW * TH;
  float w[5][TN];  // rho', rho*u, rho*v, (rho*s)', (rho*T_n)'
  float p[TN], T[TN];
  float rho[TN], u[TN], v[TN], s[TN], tn[TN];  // full rho, u, v; s', T_n'
  RowF row[TH];  // per staged row
};

// The y fluxes of the components whose ambient depends on height (rho,
// rho*s, rho*T_n), in FP64 for the FP64 reconstruction: the stored
// momentum times the row's FP64 ambient s or T_n plus the perturbation
// (rho v s = (rho v) s), so that the ambient's row-to-row change is not
// rounded to FP32 (a rounding that stays with the row and biases every
// step alike).
template <int BX, int BY>
__device__ inline double flux_y64(const Lds32<BX, BY>& L, int e, int r, int k) {
  const double m = (double)L.w[2][e];
  switch (k) {
    case 0: return m;
    case 3: return m * (L.row[r].aS64 + (double)L.s[e]);
    default: return m * (L.row[r].aTn64 + (double)L.tn[e]);
  }
}

// F(e1) - F(e2) of two cells of one row (x) and G(e1) - G(e2) of two cells
// of one column (y, rows r1 and r2), formed difference first: the parts
// that are stored values (the momenta for the mass flux, p' in the normal
// momentum flux, the ambient entropy and T_n of a row times the momentum)
// are differenced before the products of perturbations are added, so a
// smooth field's nearly cancelling fluxes lose no more than the rounding of
// their difference.  The fluxes are those of rhs_hybrid_high_llf.m4:125-156
// on full values (ambient + perturbation); the mass flux is the stored
// momentum rather than rho times its quotient u.  Along y the ambient s and
// T_n change from row to row: that change is formed in FP64 and rounded
// once (each row's FP32 value carries its own fixed rounding, which the
// difference of two rows would keep at full size).
template <int BX, int BY>
__device__ inline float dflux_x(const Lds32<BX, BY>& L, int e1, int e2, int r, int k) {
  const float m1 = L.w[1][e1], m2 = L.w[1][e2];
  switch (k) {
    case 0: return m1 - m2;
    case 1: return (L.p[e1] - L.p[e2]) + (m1 * L.u[e1] - m2 * L.u[e2]);
    case 2: return m1 * L.v[e1] - m2 * L.v[e2];
    case 3: return L.row[r].aS * (m1 - m2) + (m1 * L.s[e1] - m2 * L.s[e2]);
    default: return L.row[r].aTn * (m1 - m2) + (m1 * L.tn[e1] - m2 * L.tn[e2]);
  }
}
template <int BX, int BY>
__device__ inline float dflux_y(const Lds32<BX, BY>& L, int e1, int e2, int r1, int r2, int k) {
  const float m1 = L.w[2][e1], m2 = L.w[2][e2];
  switch (k) {
    case 0: return m1 - m2;
    case 1: return m1 * L.u[e1] - m2 * L.u[e2];
    case 2: return ((L.p[e1] - L.p[e2]) + (L.row[r1].pA - L.row[r2].pA)) + (m1 * L.v[e1] - m2 * L.v[e2]);
    case 3:
      return (L.row[r1].aS * (m1 - m2) + (float)(L.row[r1].aS64 - L.row[r2].aS64) * m2) +
             (m1 * L.s[e1] - m2 * L.s[e2]);
    default:
      return (L.row[r1].aTn * (m1 - m2) + (float)(L.row[r1].aTn64 - L.row[r2].aTn64) * m2) +
             (m1 * L.tn[e1] - m2 * L.tn[e2]);
  }
}

// One tile of one patch: rhs in FP32, FP64 RK accumulation, FP32 EOS.
// SENSOR: level 0 of a hybrid variant (DRP where the density sensor allows,
// or as the learned switch's decision map says where one is given).
// EVAL: store K32 into kout instead of stepping.
template <int BX, int BY, bool SENSOR, bool EVAL>
__device__ __attribute__((always_inline)) inline void stage32_body(const Args32& a, int i0, int j0, Lds32<BX, BY>& L) {
  using L_t = Lds32<BX, BY>;
  constexpr int TW = L_t::TW, TH = L_t::TH, TN = L_t::TN, S = TW;
  const int tid = threadIdx.x + threadIdx.y * BX;
  const int jr = j0 - 3;  // local row of staged row 0 (>= -g since g >= 3)
  // per-row ambient of the staged rows inside the ghosted array
  {
    constexpr int W = sizeof(RowF) / 4;
    const int r_hi = min(TH, a.ny + a.g - jr);
    const unsigned* src = (const unsigned*)(a.rows + (jr + a.g));
    unsigned* dst = (unsigned*)L.row;
    for (int q = tid; q < r_hi * W; q += BX * BY) dst[q] = src[q];
  }
  __syncthreads();
  // every load of the thread's staged cells is issued before the first is
  // used; a cell outside the tile or the ghost box reads element 0 and is
  // not stored
  constexpr int kIt = (TN + BX * BY - 1) / (BX * BY);
  float v[kIt][7];
  bool in[kIt];
#pragma unroll
  for (int it = 0; it < kIt; it++) {
    const int e = tid + it * BX * BY;
    const int gi = i0 + (e % TW) - 3, gj = jr + e / TW;
    in[it] = e < TN && gi >= -a.g && gi < a.nx + a.g && gj >= -a.g && gj < a.ny + a.g;
    const long gx = in[it] ? (long)(gi + a.g) + (long)(gj + a.g) * a.sy : 0;
#pragma unroll
    for (int k = 0; k < 7; k++) v[it][k] = a.in[gx + k * a.sk];
  }
#pragma unroll
  for (int it = 0; it < kIt; it++) {
    const int e = tid + it * BX * BY;
    if (in[it]) {
      const int r = e / TW;
      float w[5];
#pragma unroll
      for (int k = 0; k < 5; k++) L.w[k][e] = w[k] = v[it][k];
      L.p[e] = v[it][5];
      L.T[e] = v[it][6];
      const float rho = L.row[r].aRho + w[0];
      const float ir = rcpf(rho);
      L.rho[e] = rho;
      L.u[e] = w[1] * ir;
      L.v[e] = w[2] * ir;
      L.s[e] = (w[3] - L.row[r].aS * w[0]) * ir;
      L.tn[e] = (w[4] - L.row[r].aTn * w[0]) * ir;
    }
  }
  __syncthreads();
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  const int r = threadIdx.y + 3;
  const int e = r * TW + threadIdx.x + 3;
  const float* U = L.u;
  const float* V = L.v;
  const float* Tp = L.T;
  // ---- physical source H (rhs_hybrid_high_llf.m4:159-237), perturbation form
  float H[5];
  const long gx = (long)(i + a.g) + (long)(j + a.g) * a.sy;
  H[0] = (float)a.h0[gx];
  {
    const float u_dprime_x = (U[e - 1] - 2 * U[e] + U[e + 1]) * a.inv_dx2;
    const float u_dprime_y = (U[e - S] - 2 * U[e] + U[e + S]) * a.inv_dy2;
    const float v_dprime_xy = (V[e + 1 + S] - V[e + 1 - S] - V[e - 1 + S] + V[e - 1 - S]) * a.inv_4dxdy;
    const float phi_xx_prime_x = (4.0f / 3.0f) * u_dprime_x - (2.0f / 3.0f) * v_dprime_xy;
    const float phi_xy_prime_y = u_dprime_y + v_dprime_xy;
    H[1] = a.muB * (u_dprime_x + v_dprime_xy) + a.mu * (phi_xx_prime_x + phi_xy_prime_y);
    const float v_dprime_y = (V[e - S] - 2 * V[e] + V[e + S]) * a.inv_dy2;
    const float v_dprime_x = (V[e - 1] - 2 * V[e] + V[e + 1]) * a.inv_dx2;
    const float u_dprime_yx = (U[e + 1 + S] - U[e + 1 - S] - U[e - 1 + S] + U[e - 1 - S]) * a.inv_4dxdy;
    const float phi_yx_prime_x = u_dprime_yx + v_dprime_x;
    const float phi_yy_prime_y = (4.0f / 3.0f) * v_dprime_y - (2.0f / 3.0f) * u_dprime_yx;
    H[2] = a.muB * (v_dprime_y + u_dprime_yx) + a.mu * (phi_yx_prime_x + phi_yy_prime_y);
    const float u_prime_x = (U[e + 1] - U[e - 1]) * a.inv_2dx;
    const float u_prime_y = (U[e + S] - U[e - S]) * a.inv_2dy;
    const float v_prime_y = (V[e + S] - V[e - S]) * a.inv_2dy;
    const float v_prime_x = (V[e + 1] - V[e - 1]) * a.inv_2dx;
    const float t_prime_x = (Tp[e + 1] - Tp[e - 1]) * a.inv_2dx;
    const float t_prime_y = ((Tp[e + S] - Tp[e - S]) + L.row[r].d1T) * a.inv_2dy;
    const float t_dprime_x = (Tp[e - 1] - 2 * Tp[e] + Tp[e + 1]) * a.inv_dx2;
    const float t_dprime_y = ((Tp[e - S] - 2 * Tp[e] + Tp[e + S]) + L.row[r].d2T) * a.inv_dy2;
    const float w0 = L.rho[e];
    const float temp = L.row[r].aT + Tp[e];
    const float T_n = L.row[r].aTn + L.tn[e];
    const float dTn = (Tp[e] - L.tn[e]) + L.row[r].TmTn;  // T - T_n
    const float itemp = rcpf(temp), iTn = rcpf(T_n);
    const float Mn = dTn * a.inv_tau_n;
    const float tmp1 = a.T_star_n * iTn;
    const float c_vn = 0.78f * a.R * sqf(tmp1) * __expf(-tmp1);
    const float A_n = (dTn * iTn) * c_vn;
    const float phi_xx = (4.0f / 3.0f) * u_prime_x - (2.0f / 3.0f) * v_prime_y;
    const float phi_xy_yx = u_prime_y + v_prime_x;
    const float phi_yy = (4.0f / 3.0f) * v_prime_y - (2.0f / 3.0f) * u_prime_x;
    const float it2 = sqf(itemp);
    const float sigma = (a.muB * itemp) * sqf(u_prime_x + v_prime_y) +
                        (a.mu * 0.5f * itemp) * (sqf(phi_xx) + 2 * sqf(phi_xy_yx) + sqf(phi_yy)) +
                        (a.kappa * it2) * (sqf(t_prime_x) + sqf(t_prime_y)) + (w0 * itemp) * (A_n * Mn);
    H[3] = sigma - (w0 * iTn) * c_vn * Mn +
           (-(a.kappa * sqf(t_prime_x)) * it2 + (a.kappa * t_dprime_x) * itemp -
            (a.kappa * sqf(t_prime_y)) * it2 + (a.kappa * t_dprime_y) * itemp);
    H[4] = (w0 * a.inv_tau_n) * dTn;
  }
  // ---- switch (rhs_hybrid_high_llf.m4:239-256): density differences
  bool weno = true;
  if constexpr (SENSOR) {
    const unsigned char d = a.decide ? a.decide[(long)i + (long)j * a.nx] : 2;
    weno = d == 1;
    const float* R0 = L.w[0];
#pragma unroll
    for (int q = -1; q <= 1 && d == 2; q++) {
      float t1 = R0[e + q] - R0[e - 1 + q], t2 = R0[e + 1 + q] - R0[e + q];
      float num = 2 * fabsf(t1 * t2) + a.ep, den = sqf(t1) + sqf(t2) + a.ep;
      if (num / den < 0.9999f) weno = true;
      t1 = (R0[e + q * S] - R0[e + (q - 1) * S]) + L.row[r + q - 1].d1Rho;
      t2 = (R0[e + (q + 1) * S] - R0[e + q * S]) + L.row[r + q].d1Rho;
      num = 2 * fabsf(t1 * t2) + a.ep;
      den = sqf(t1) + sqf(t2) + a.ep;
      if (num / den < 0.9999f) weno = true;
    }
  }
  float K[5];
  if (!weno) {
    // DRP (rhs_hybrid_high_llf.m4:258-268) as a1 (F1 - F-1) + a2 (F2 - F-2)
    // + a3 (F3 - F-3), each pair differenced first (dflux_x / dflux_y)
    const float a1 = 0.79926643f, a2 = -0.18941314f, a3 = 0.02651995f;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const float fx = a1 * dflux_x(L, e + 1, e - 1, r, k) + a2 * dflux_x(L, e + 2, e - 2, r, k) +
                       a3 * dflux_x(L, e + 3, e - 3, r, k);
      const float gy = a1 * dflux_y(L, e + S, e - S, r + 1, r - 1, k) +
                       a2 * dflux_y(L, e + 2 * S, e - 2 * S, r + 2, r - 2, k) +
                       a3 * dflux_y(L, e + 3 * S, e - 3 * S, r + 3, r - 3, k);
      K[k] = fx * a.inv_mdx + gy * a.inv_mdy + H[k];
    }
  } else {
    // WENO5 with per-cell LLF speeds (rhs_hybrid_high_llf.m4:270-356)
    float af = 0, ag = 0;
    const float hj = a.ylo + a.dy * ((float)j + 0.5f);
#pragma unroll
    for (int q = -3; q <= 3; q++) {
      const float vf = fabsf(U[e + q]) + a.c + a.c_lr * hj;
      af = vf > af ? vf : af;
      const float vg = fabsf(V[e + q * S]) + a.c + a.c_lr * (a.ylo + a.dy * ((float)(j + q) + 0.5f));
      ag = vg > ag ? vg : ag;
    }
#pragma unroll 1
    for (int k = 0; k < 5; k++) {
      float F[7], W[7];
#pragma unroll
      for (int m = 0; m < 7; m++) {
        F[m] = dflux_x(L, e + m - 3, e, r, k);
        W[m] = L.w[k][e + m - 3] - L.w[k][e];
      }
      const float x = weno_dir32s(F, W, af);
      float y;
      if (k == 1 || k == 2) {
        // momenta: no ambient, FP32
#pragma unroll
        for (int m = 0; m < 7; m++) {
          F[m] = dflux_y(L, e + (m - 3) * S, e, r + m - 3, r, k);
          W[m] = L.w[k][e + (m - 3) * S] - L.w[k][e];
        }
        y = weno_dir32s(F, W, ag);
      } else {
        // full values: FP64 ambient row + FP32 perturbation
        const int c = k == 0 ? 0 : k - 2;
        double Fd[7], Wd[7];
#pragma unroll
        for (int m = 0; m < 7; m++) {
          Fd[m] = flux_y64(L, e + (m - 3) * S, r + m - 3, k);
          Wd[m] = L.row[r + m - 3].aW[c] + (double)L.w[k][e + (m - 3) * S];
        }
        y = (float)weno_dir64(Fd, Wd, (double)ag);
      }
      const float kk = x * a.inv_dx + y * a.inv_dy + H[k];
      K[0] = k == 0 ? kk : K[0];
      K[1] = k == 1 ? kk : K[1];
      K[2] = k == 2 ? kk : K[2];
      K[3] = k == 3 ? kk : K[3];
      K[4] = k == 4 ? kk : K[4];
    }
  }
  const long ic = (long)i + (long)j * a.nx, nn = (long)a.nx * a.ny;
  if constexpr (EVAL) {
#pragma unroll
    for (int k = 0; k < 5; k++) a.kout[ic + k * nn] = K[k];
    if (SENSOR && a.amb_weno) a.amb_weno[ic] = weno ? 1 : 0;
    return;
  } else {
    // ---- FP64 RK accumulation (runge_kutta.m4:60-63) of the perturbation
    // the loads before the first store, so that they are in flight together
    double wn[5], wcv[5], win[5];
    float dkv[5];
    signed char rv[5] = {0, 0, 0, 0, 0};
    const bool use_in = a.alpha_2 != 0.0;  // uniform over the launch
#pragma unroll
    for (int k = 0; k < 5; k++) {
      dkv[k] = a.dk[ic + k * nn];
      wcv[k] = a.wc[ic + k * nn];
      if (use_in) rv[k] = a.rin[ic + k * nn];
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
      // the stage input at the FP64 value its RK update formed
      win[k] = res_dec(L.w[k][e], rv[k]);
      const double kd = (double)K[k] + (double)dkv[k];
      wn[k] = a.alpha_1 * wcv[k] + a.alpha_2 * win[k] + a.bdt * kd;
      const float wf = (float)wn[k];
      a.out[gx + k * a.sk] = wf;
      a.rout[ic + k * nn] = res_enc(wn[k], wf);
      if (a.last) a.wc[ic + k * nn] = wn[k];
    }
    // ---- update_state (update_state.m4:87-107) in perturbation form
    const float rp = (float)wn[0], rsp = (float)wn[3];
    const float pp = L.row[r].cy2 * (rp * (1.0f + a.a_eos * (2.0f * L.row[r].aD + rp)) + rsp * a.inv_cp);
    const float rho = L.row[r].aRho + rp;
    const float sp = (rsp - L.row[r].aS * rp) * rcpf(rho);
    const float Tpn = L.row[r].aT * expm1f((sp + a.R * log1pf(rp * rcpf(L.row[r].aRho))) * a.inv_cv);
    a.out[gx + 5 * a.sk] = pp;
    a.out[gx + 6 * a.sk] = Tpn;
  }
}

// PACKED = false: one tile per workgroup, tiles[blockIdx.x] (make_tiles
// stores them in dispatch order).  PACKED = true: workgroup blockIdx.x works
// through its pack, tiles packs[w] .. packs[w + 1] - 1, with a barrier
// between tiles for the LDS; the LDS is addressed from a base the compiler
// cannot see through in each iteration, so that it does not hoist a whole
// tile's stencil addresses out of the loop into registers.
template <int BX, int BY, bool SENSOR, bool EVAL, bool PACKED>
__global__ void __launch_bounds__(BX* BY) __attribute__((amdgpu_waves_per_eu(4))) stage32_kernel(const int4* __restrict__ tiles, const int* __restrict__ packs,
                                                         const Args32* __restrict__ args) {
  __shared__ Lds32<BX, BY> L;
  if constexpr (!PACKED) {
    const int4 tl = tiles[blockIdx.x];
    stage32_body<BX, BY, SENSOR, EVAL>(args[tl.x], tl.y, tl.z, L);
  } else {
    const int q1 = packs[blockIdx.x + 1];
#pragma clang loop unroll(disable)
    for (int q = packs[blockIdx.x]; q < q1; q++) {
      const int4 tl = tiles[q];
      int z;
      asm volatile("s_mov_b32 %0, 0" : "=s"(z));
      Lds32<BX, BY>& Lq = *(Lds32<BX, BY>*)((char*)&L + z);
      stage32_body<BX, BY, SENSOR, EVAL>(args[tl.x], tl.y, tl.z, Lq);
      __syncthreads();
    }
  }
}

// ---- ghost fill and coarsen on full values ----------------------------------

struct Fill32 {
  float* pool;
  const Patch32* patches;
  const double* amb;
  int buf, cbuf;
  int reflect[4];
};

__device__ inline const double* row_rec(const Fill32& f, const Patch32& P, long local) {
  return f.amb + (P.amb + local / P.sy) * kRec;
}

__device__ inline void fill32_copy(const Fill32& f, int q, const int* dp, const int* dl, const int* sp,
                                   const int* sl) {
  const Patch32& D = f.patches[dp[q]];
  const Patch32& Sd = f.patches[sp[q]];
  float* d = f.pool + D.blk[f.buf] + dl[q];
  const float* s = f.pool + Sd.blk[f.buf] + sl[q];
  // the same global cell: the same row, the same ambient
  for (int k = 0; k < 7; k++) d[k * D.sk] = s[k * Sd.sk];
}

__device__ inline void fill32_interp(const Fill32& f, int q, const int* dp, const int* dl, const int* sp,
                                     const int* sl, const double* fx, const double* fy) {
  const Patch32& D = f.patches[dp[q]];
  const Patch32& C = f.patches[sp[q]];
  float* d = f.pool + D.blk[f.buf] + dl[q];
  const float* c = f.pool + C.blk[f.cbuf] + sl[q];
  const double* rf = row_rec(f, D, dl[q]);
  const double* rc = row_rec(f, C, sl[q]);
  const double ax = fx[q], ay = fy[q];
  for (int k = 0; k < 7; k++) {
    const float* ck = c + k * C.sk;
    const double A0 = amb_of(rc, k), Am = amb_of(rc - kRec, k), Ap = amb_of(rc + kRec, k);
    const double c0 = A0 + (double)ck[0];
    const double v = c0 + ax * (0.5 * ((A0 + (double)ck[1]) - (A0 + (double)ck[-1]))) +
                     ay * (0.5 * ((Ap + (double)ck[C.sy]) - (Am + (double)ck[-C.sy])));
    d[k * D.sk] = (float)(v - amb_of(rf, k));
  }
}

__device__ inline void fill32_bc(const Fill32& f, int q, const int* pp, const int* dl, const int* sl,
                                 const int* side) {
  const Patch32& P = f.patches[pp[q]];
  float* base = f.pool + P.blk[f.buf];
  const int sd = side[q];
  const int negk = f.reflect[sd - 1] ? (sd <= 2 ? 1 : 2) : -1;
  const double* rd = row_rec(f, P, dl[q]);
  const double* rs = row_rec(f, P, sl[q]);
  for (int k = 0; k < 7; k++) {
    double v = amb_of(rs, k) + (double)base[sl[q] + k * P.sk];
    v = k == negk ? -v : v;
    base[dl[q] + k * P.sk] = (float)(v - amb_of(rd, k));
  }
}

// copies, refinement and edge strips of a level in one launch (disjoint
// destinations, interior sources), corners after them
__global__ void __launch_bounds__(256) fill32_edges_kernel(Fill32 f, LevelFill L) {
  const int n = L.ncopy + L.ninterp + L.nedge;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    if (q < L.ncopy)
      fill32_copy(f, q, L.c_dp, L.c_dl, L.c_sp, L.c_sl);
    else if (q < L.ncopy + L.ninterp)
      fill32_interp(f, q - L.ncopy, L.i_dp, L.i_dl, L.i_sp, L.i_sl, L.i_fx, L.i_fy);
    else
      fill32_bc(f, q - L.ncopy - L.ninterp, L.e_p, L.e_d, L.e_s, L.e_n);
  }
}

__global__ void __launch_bounds__(256) fill32_corner_kernel(Fill32 f, int n, const int* pp, const int* dl,
                                                             const int* sl, const int* side) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    fill32_bc(f, q, pp, dl, sl, side);
}

// Fine-to-coarse average of block 0 and the carried state, on full values.
__global__ void __launch_bounds__(256) coarsen32_kernel(Fill32 f, double* wc, signed char* res8, int r, int n,
                                                         const int* cp, const int* cl, const int* fp,
                                                         const int* fl) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const Patch32& C = f.patches[cp[q]];
    const Patch32& F = f.patches[fp[q]];
    float* cb = f.pool + C.blk[0] + cl[q];
    const float* fb = f.pool + F.blk[0] + fl[q];
    const double* rc = row_rec(f, C, cl[q]);
    const double* rf0 = row_rec(f, F, fl[q]);
    const double inv = 1.0 / (double)(r * r);
    const long ci = (long)(cl[q] / C.sy - C.g) * C.nx + (cl[q] % C.sy - C.g);
    const long fi0 = (long)(fl[q] / F.sy - F.g) * F.nx + (fl[q] % F.sy - F.g);
    const long nc = (long)C.nx * C.ny, nf = (long)F.nx * F.ny;
    for (int k = 0; k < 7; k++) {
      double s = 0;
      for (int fj = 0; fj < r; fj++)
        for (int fi = 0; fi < r; fi++) {
          const double A = amb_of(rf0 + fj * kRec, k);
          const double pv = k < 5 ? wc[F.wc + fi0 + fi + (long)fj * F.nx + k * nf]
                                  : (double)fb[fi + (long)fj * F.sy + k * F.sk];
          s += A + pv;
        }
      const double v = s * inv - amb_of(rc, k);
      const float vf = (float)v;
      if (k < 5) {
        wc[C.wc + ci + k * nc] = v;
        res8[C.res[0] + ci + k * nc] = res_enc(v, vf);
      }
      cb[k * C.sk] = vf;
    }
  }
}

// FP64 block -> FP32 block 0 (+ carried state, at the patch's offset p.wc
// of the carried-state pool, and the interior's rounding residuals), and
// back.
__global__ void __launch_bounds__(256) to32_kernel(const double* w64, float* w32, double* wc, signed char* res8,
                                                    const double* amb, Patch32 p) {
  const long nall = p.sk;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nall; q += (long)gridDim.x * blockDim.x) {
    const double* rec = amb + (p.amb + q / p.sy) * kRec;
    const int jj = (int)(q / p.sy) - p.g, ii = (int)(q % p.sy) - p.g;
    const bool inner = ii >= 0 && ii < p.nx && jj >= 0 && jj < p.ny;
    for (int k = 0; k < 7; k++) {
      const double v = w64[q + k * p.sk] - amb_of(rec, k);
      const float vf = (float)v;
      w32[q + k * p.sk] = vf;
      if (inner && k < 5) {
        const long c = ii + (long)jj * p.nx + k * (long)p.nx * p.ny;
        wc[p.wc + c] = v;
        res8[p.res[0] + c] = res_enc(v, vf);
      }
    }
  }
}

__global__ void __launch_bounds__(256) to64_kernel(double* w64, const float* w32, const double* wc,
                                                    const double* amb, Patch32 p) {
  const long nall = p.sk;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nall; q += (long)gridDim.x * blockDim.x) {
    const double* rec = amb + (p.amb + q / p.sy) * kRec;
    const int jj = (int)(q / p.sy) - p.g, ii = (int)(q % p.sy) - p.g;
    const bool inner = ii >= 0 && ii < p.nx && jj >= 0 && jj < p.ny;
    for (int k = 0; k < 7; k++) {
      const double pv =
          (inner && k < 5) ? wc[p.wc + ii + (long)jj * p.nx + k * (long)p.nx * p.ny] : (double)w32[q + k * p.sk];
      w64[q + k * p.sk] = amb_of(rec, k) + pv;
    }
  }
}

// FP64 state block := the ambient on every cell (p, T from the records).
__global__ void __launch_bounds__(256) ambient64_kernel(double* w64, const double* amb, Patch32 p) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < p.sk; q += (long)gridDim.x * blockDim.x) {
    const double* rec = amb + (p.amb + q / p.sy) * kRec;
    for (int k = 0; k < 7; k++) w64[q + k * p.sk] = amb_of(rec, k);
  }
}

__global__ void __launch_bounds__(256) zero32_kernel(float* w32, long n) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) w32[q] = 0.0f;
}

// dK = K64 - K32 on the interior (K64 in the w_n components of an FP64 block).
__global__ void __launch_bounds__(256) dk_kernel(float* dk, const double* k64, Patch32 p) {
  const long nn = (long)p.nx * p.ny;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < 5 * nn; q += (long)gridDim.x * blockDim.x) {
    const int k = (int)(q / nn);
    const long c = q % nn;
    const long i = c % p.nx, j = c / p.nx;
    const double kk = k64[(i + p.g) + (j + p.g) * p.sy + k * p.sk];
    dk[q] = (float)(kk - (double)dk[q]);
  }
}

}  // namespace

struct Fp32Form {
  float* pool32 = nullptr;
  signed char* res8 = nullptr;
  double* wc = nullptr;
  double* amb = nullptr;
  RowF* rows = nullptr;
  double rho0 = 0, p0 = 0;  // of the gas the ambient was set with
  Patch32* dp = nullptr;
  Args32* d_args = nullptr;  // kMaxStages x npatch
  Args32* d_eval = nullptr;  // npatch
  std::vector<Patch32> hp;
  std::vector<Args32> h_args;
  std::vector<double> amb_h[kMaxLevels];  // records per level
  int amb_lo[kMaxLevels] = {0};
  bool ready = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> key;
  std::vector<void*> allocs;
  // the learned switch on level 0 (enl_batch_fp32_set_switch)
  Switch* sw = nullptr;
  float lo = 0, hi = 1;
  int every_stage = 1;
  unsigned long long* totals = nullptr;
  unsigned char* dmap = nullptr;  // decision maps of the level-0 patches
  unsigned char* ambmap = nullptr;  // per level-0 cell: the sensor's choice on the ambient alone
  std::vector<long> dmap_off;     // per patch, -1 above level 0
  SwPatch* swtab = nullptr;       // 3 x n0 entries: the level-0 patches per stage block
  int n0 = 0, max_side0 = 0;
  long cells0 = 0;
  long fcells0 = 0;  // ghosted cells of the level-0 patches (the switch's feature map)
  double sw_thr = 0;
  float sw_inv_r = 0, sw_inv_p = 0;
};

namespace {

Fp32Form* form(Batch* b) {
  if (!b->f32) b->f32 = new Fp32Form();
  return b->f32;
}

void fill32_level(Batch* b, int level, int buf, int cbuf, const int* reflect, hipStream_t s) {
  Fp32Form* F = b->f32;
  Fill32 f;
  f.pool = F->pool32;
  f.patches = F->dp;
  f.amb = F->amb;
  f.buf = buf;
  f.cbuf = cbuf;
  for (int k = 0; k < 4; k++) f.reflect[k] = reflect[k];
  const LevelFill& L = b->fill[level];
  const long n = (long)L.ncopy + L.ninterp + L.nedge;
  if (n) hipLaunchKernelGGL(fill32_edges_kernel, dim3(grid_for(n)), dim3(256), 0, s, f, L);
  if (L.ncorner)
    hipLaunchKernelGGL(fill32_corner_kernel, dim3(grid_for(L.ncorner)), dim3(256), 0, s, f, L.ncorner, L.k_p, L.k_d,
                       L.k_s, L.k_n);
}

template <int BX, int BY, bool EVAL>
void stage32_launch(Batch* b, const Args32* args, int lev, hipStream_t s) {
  const int np = b->level_npack[lev];
  if (np <= 0) return;
  const int4* tiles = b->tiles + b->level_tile_begin[lev];
  const int* packs = b->packs + b->level_pack_begin[lev];
  const bool packed = b->pack_mode[lev] != ENL_PACK_SINGLE;
  if (lev == 0 && packed)
    hipLaunchKernelGGL((stage32_kernel<BX, BY, true, EVAL, true>), dim3(np), dim3(BX, BY), 0, s, tiles, packs, args);
  else if (lev == 0)
    hipLaunchKernelGGL((stage32_kernel<BX, BY, true, EVAL, false>), dim3(np), dim3(BX, BY), 0, s, tiles, packs, args);
  else if (packed)
    hipLaunchKernelGGL((stage32_kernel<BX, BY, false, EVAL, true>), dim3(np), dim3(BX, BY), 0, s, tiles, packs, args);
  else
    hipLaunchKernelGGL((stage32_kernel<BX, BY, false, EVAL, false>), dim3(np), dim3(BX, BY), 0, s, tiles, packs,
                       args);
}

template <bool EVAL>
int stage32(Batch* b, const Args32* args, int lev, hipStream_t s) {
  const int bx = b->tile_bx[lev], by = b->tile_by[lev];
  if (bx == 32 && by == 8) stage32_launch<32, 8, EVAL>(b, args, lev, s);
  else if (bx == 16 && by == 16) stage32_launch<16, 16, EVAL>(b, args, lev, s);
  else if (bx == 16 && by == 8) stage32_launch<16, 8, EVAL>(b, args, lev, s);
  else if (bx == 8 && by == 8) stage32_launch<8, 8, EVAL>(b, args, lev, s);
  else return set_error(-2, "unsupported tile shape");
  return 0;
}

Args32 base32(Batch* b, int p, const enl_gas_t* gas) {
  const Fp32Form* F = b->f32;
  const PatchDev& pd = b->hp[p];
  const Patch32& q = F->hp[p];
  Args32 a;
  memset(&a, 0, sizeof a);
  a.h0 = b->pool + pd.src;
  a.amb = F->amb + q.amb * kRec;
  a.rows = F->rows + q.amb;
  a.wc = F->wc + q.wc;
  a.dk = F->pool32 + q.dk;
  a.nx = pd.nx;
  a.ny = pd.ny;
  a.g = pd.g;
  a.sy = pd.sy;
  a.sk = pd.sk;
  a.ylo = (float)pd.ylo;
  a.dy = (float)pd.dy;
  a.inv_dx = (float)(1.0 / pd.dx);
  a.inv_dy = (float)(1.0 / pd.dy);
  a.inv_mdx = (float)(1.0 / -pd.dx);
  a.inv_mdy = (float)(1.0 / -pd.dy);
  a.inv_dx2 = (float)(1.0 / (pd.dx * pd.dx));
  a.inv_dy2 = (float)(1.0 / (pd.dy * pd.dy));
  a.inv_4dxdy = (float)(1.0 / (4 * pd.dx * pd.dy));
  a.inv_2dx = (float)(1.0 / (2 * pd.dx));
  a.inv_2dy = (float)(1.0 / (2 * pd.dy));
  a.c = (float)gas->c;
  a.c_lr = (float)gas->c_lr;
  a.mu = (float)gas->mu;
  a.muB = (float)gas->muB;
  a.kappa = (float)gas->kappa;
  a.R = (float)gas->R;
  a.T_star_n = (float)gas->T_star_n;
  a.inv_tau_n = (float)(1.0 / gas->tau_n);
  a.inv_cp = (float)(1.0 / gas->c_p);
  a.inv_cv = (float)(1.0 / gas->c_v);
  a.a_eos = (float)((gas->gamma1 - 1) / (2 * gas->rho0));
  a.ep = (float)sensor_eps(gas->tag_tolerance[2]);
  a.rho0 = gas->rho0;
  return a;
}

template <typename T>
int alloc(Fp32Form* F, T** p, size_t n, const char* what) {
  hipError_t e = batch_malloc((void**)p, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess) return check_hip(e, what);
  F->allocs.push_back(*p);
  return 0;
}

void drop_graph32(Fp32Form* F) {
  if (F->exec) (void)hipGraphExecDestroy(F->exec);
  if (F->graph) (void)hipGraphDestroy(F->graph);
  F->exec = nullptr;
  F->graph = nullptr;
}

// Buffers, ambient records and patch table of the FP32 form.
int setup32(Batch* b) {
  Fp32Form* F = b->f32;
  for (void* a : F->allocs) batch_free(a);
  F->allocs.clear();
  drop_graph32(F);
  F->ready = false;
  F->hp.assign(b->npatch, Patch32());
  long off32 = 0, off64 = 0, off8 = 0, amb_total = 0;
  long amb_base[kMaxLevels];
  for (int l = 0; l < kMaxLevels; l++) {
    amb_base[l] = amb_total;
    amb_total += (long)F->amb_h[l].size() / kRec;
  }
  for (int p = 0; p < b->npatch; p++) {
    const PatchDev& pd = b->hp[p];
    const enl_patch_desc_t& d = b->desc[p];
    Patch32& q = F->hp[p];
    q.nx = pd.nx;
    q.ny = pd.ny;
    q.g = pd.g;
    q.level = pd.level;
    q.sy = pd.sy;
    q.sk = pd.sk;
    for (int k = 0; k < 3; k++) {
      q.blk[k] = off32;
      off32 += 7 * pd.sk;
      off32 = (off32 + 63) & ~63L;
      q.res[k] = off8;
      off8 += 5L * pd.nx * pd.ny;
      off8 = (off8 + 255) & ~255L;
    }
    q.dk = off32;
    off32 += 5L * pd.nx * pd.ny;
    off32 = (off32 + 63) & ~63L;
    q.wc = off64;
    off64 += 5L * pd.nx * pd.ny;
    off64 = (off64 + 31) & ~31L;
    const long nrec = (long)F->amb_h[pd.level].size() / kRec;
    const long r0 = (long)(d.box[2] - pd.g) - F->amb_lo[pd.level];
    if (nrec == 0 || r0 < 0 || r0 + pd.ny + 2 * pd.g > nrec)
      return set_error(-2, "ambient rows of level %d do not cover patch %d", pd.level, p);
    q.amb = amb_base[pd.level] + r0;
  }
  int rc = alloc(F, &F->pool32, (size_t)off32, "hipMalloc(FP32 pool)");
  if (!rc) rc = alloc(F, &F->res8, (size_t)off8, "hipMalloc(FP32 rounding residuals)");
  if (!rc) rc = check_hip(hipMemset(F->res8, 0, (size_t)off8), "FP32 rounding residuals clear");
  if (!rc) rc = alloc(F, &F->wc, (size_t)off64, "hipMalloc(FP32 carried state)");
  if (!rc) rc = alloc(F, &F->amb, (size_t)amb_total * kRec, "hipMalloc(ambient)");
  if (!rc) rc = alloc(F, &F->rows, (size_t)amb_total, "hipMalloc(ambient rows)");
  if (!rc) rc = alloc(F, &F->dp, (size_t)b->npatch, "hipMalloc(FP32 patch table)");
  if (!rc) rc = alloc(F, &F->d_args, (size_t)kMaxStages * b->npatch, "hipMalloc(FP32 arguments)");
  if (!rc) rc = alloc(F, &F->d_eval, (size_t)b->npatch, "hipMalloc(FP32 eval arguments)");
  if (rc) return rc;
  for (int l = 0; l < kMaxLevels; l++)
    if (!F->amb_h[l].empty())
      rc = rc ? rc
              : check_hip(hipMemcpy(F->amb + amb_base[l] * kRec, F->amb_h[l].data(),
                                    sizeof(double) * F->amb_h[l].size(), hipMemcpyHostToDevice),
                          "ambient upload");
  if (!rc) {
    std::vector<RowF> rows((size_t)amb_total);
    for (int l = 0; l < kMaxLevels; l++) {
      const long nrec = (long)F->amb_h[l].size() / kRec;
      for (long q = 0; q < nrec; q++) {
        const double* rec = &F->amb_h[l][(size_t)q * kRec];
        const bool up = q + 1 < nrec, dn = q > 0;
        RowF& R = rows[(size_t)(amb_base[l] + q)];
        memset(&R, 0, sizeof R);
        R.aW[0] = rec[0];
        R.aW[1] = rec[1];
        R.aW[2] = rec[2];
        R.aS64 = rec[5];
        R.aTn64 = rec[6];
        R.aRho = (float)rec[0];
        R.aS = (float)rec[5];
        R.aTn = (float)rec[6];
        R.aT = (float)rec[4];
        R.aD = (float)(rec[0] - F->rho0);
        R.cy2 = (float)rec[7];
        R.TmTn = (float)(rec[4] - rec[6]);
        R.pA = (float)(rec[3] - F->p0);
        R.d1Rho = up ? (float)(rec[kRec + 0] - rec[0]) : 0.0f;
        R.d1T = (up && dn) ? (float)(rec[kRec + 4] - rec[-kRec + 4]) : 0.0f;
        R.d2T = (up && dn) ? (float)((rec[-kRec + 4] - 2 * rec[4]) + rec[kRec + 4]) : 0.0f;
      }
    }
    rc = check_hip(hipMemcpy(F->rows, rows.data(), sizeof(RowF) * rows.size(), hipMemcpyHostToDevice),
                   "ambient row upload");
  }
  if (!rc)
    rc = check_hip(hipMemcpy(F->dp, F->hp.data(), sizeof(Patch32) * F->hp.size(), hipMemcpyHostToDevice),
                   "FP32 patch table upload");
  // decision maps of the level-0 patches, and per stage block the table
  // the learned switch reads them from
  F->dmap_off.assign((size_t)b->npatch, -1L);
  F->n0 = 0;
  F->max_side0 = 0;
  F->cells0 = 0;
  F->fcells0 = 0;
  long offd = 0;
  for (int p = 0; p < b->npatch; p++) {
    const Patch32& q = F->hp[p];
    if (q.level != 0) continue;
    F->dmap_off[p] = offd;
    offd += (long)q.nx * q.ny;
    offd = (offd + 255) & ~255L;
    F->n0++;
    F->cells0 += (long)q.nx * q.ny;
    F->max_side0 = std::max(F->max_side0, std::max(q.nx, q.ny));
  }
  if (!rc) rc = alloc(F, &F->dmap, (size_t)offd, "hipMalloc(decision maps)");
  if (!rc) rc = alloc(F, &F->ambmap, (size_t)offd, "hipMalloc(ambient sensor maps)");
  if (!rc) rc = alloc(F, &F->swtab, (size_t)3 * F->n0, "hipMalloc(switch tables)");
  if (!rc) {
    std::vector<SwPatch> t((size_t)3 * F->n0);
    int q0 = 0;
    for (int p = 0; p < b->npatch; p++) {
      if (F->dmap_off[p] < 0) continue;
      const Patch32& q = F->hp[p];
      const char* row0 = (const char*)(F->rows + q.amb);
      for (int k = 0; k < 3; k++) {
        SwPatch& S = t[(size_t)k * F->n0 + q0];
        S.rho = F->pool32 + q.blk[k];
        S.p = F->pool32 + q.blk[k] + 5 * q.sk;
        S.drho = (const float*)(row0 + offsetof(RowF, d1Rho));
        S.decide = F->dmap + F->dmap_off[p];
        S.amb_weno = F->ambmap + F->dmap_off[p];
        S.nx = q.nx;
        S.ny = q.ny;
        S.g = q.g;
        S.rs = (int)(sizeof(RowF) / sizeof(float));
        S.sy = q.sy;
        S.foff = F->fcells0;
      }
      F->fcells0 += q.sk;
      q0++;
    }
    rc = check_hip(hipMemcpy(F->swtab, t.data(), sizeof(SwPatch) * t.size(), hipMemcpyHostToDevice),
                   "switch table upload");
  }
  if (!rc && F->sw) rc = switch_reserve(F->sw, F->cells0, F->fcells0);
  F->h_args.clear();
  return rc;
}

int load32(Batch* b, hipStream_t s) {
  Fp32Form* F = b->f32;
  for (int p = 0; p < b->npatch; p++) {
    const Patch32& q = F->hp[p];
    hipLaunchKernelGGL(to32_kernel, dim3(grid_for(q.sk)), dim3(256), 0, s, b->pool + b->hp[p].state[0],
                       F->pool32 + q.blk[0], F->wc, F->res8, F->amb, q);
  }
  return check_hip(hipGetLastError(), "FP32 load");
}

int prepare_args32(Batch* b, const enl_gas_t* gas, int order, const double* a1, const double* a2,
                   const double* be, double dt, hipStream_t s) {
  Fp32Form* F = b->f32;
  std::vector<Args32> t((size_t)order * b->npatch);
  for (int st = 0; st < order; st++)
    for (int p = 0; p < b->npatch; p++) {
      Args32 a = base32(b, p, gas);
      const Patch32& q = F->hp[p];
      a.in = F->pool32 + q.blk[stage_in(st, order)];
      a.out = F->pool32 + q.blk[stage_out(st, order)];
      a.rin = F->res8 + q.res[stage_in(st, order)];
      a.rout = F->res8 + q.res[stage_out(st, order)];
      a.alpha_1 = a1[st];
      a.alpha_2 = a2[st];
      a.bdt = be[st] * dt;
      a.last = st == order - 1;
      if (F->sw && F->dmap_off[p] >= 0) a.decide = F->dmap + F->dmap_off[p];
      t[(size_t)st * b->npatch + p] = a;
    }
  if (F->sw) {
    const double ep = sensor_eps(gas->tag_tolerance[2]);
    F->sw_thr = 5e-5 * (double)(float)ep;  // of the FP32 sensor's epsilon
    F->sw_inv_r = (float)(1.0 / std::sqrt(ep));
    F->sw_inv_p = (float)(1.0 / (gas->c * gas->c * std::sqrt(ep)));
  }
  const size_t bytes = sizeof(Args32) * t.size();
  if (F->h_args.size() == t.size() && memcmp(F->h_args.data(), t.data(), bytes) == 0) return 0;
  F->h_args.swap(t);
  hipError_t e = hipMemcpyAsync(F->d_args, F->h_args.data(), bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return check_hip(e, "FP32 argument upload");
  return check_hip(hipStreamSynchronize(s), "FP32 argument upload sync");
}

int enqueue32(Batch* b, const int* reflect, int order, int nsteps, hipStream_t s) {
  Fp32Form* F = b->f32;
  for (int step = 0; step < nsteps; step++) {
    for (int st = 0; st < order; st++) {
      const int in = stage_in(st, order), out = stage_out(st, order);
      const Args32* args = F->d_args + (size_t)st * b->npatch;
      for (int lev = 0; lev < b->nlevels; lev++) {
        const bool refreshed = st > 0 && lev == 0 && b->nlevels > 1;
        if (!refreshed) fill32_level(b, lev, in, out, reflect, s);
        if (lev == 0 && F->sw && (F->every_stage || st == 0)) {
          // the learned switch decides level 0's cells from this stage's input
          int rc = switch_decide_batch(F->sw, F->swtab + (size_t)in * F->n0, F->n0, F->max_side0, F->cells0,
                                       F->fcells0, F->sw_thr, F->sw_inv_r, F->sw_inv_p, F->lo, F->hi, s);
          if (!rc && F->totals) rc = switch_accumulate(F->sw, F->totals, s);
          if (rc) return rc;
        }
        int rc = stage32<false>(b, args, lev, s);
        if (rc) return rc;
        if (lev + 1 < b->nlevels) fill32_level(b, lev, out, out, reflect, s);
      }
    }
    if (b->ncoarsen) {
      Fill32 f;
      f.pool = F->pool32;
      f.patches = F->dp;
      f.amb = F->amb;
      f.buf = 0;
      f.cbuf = 0;
      for (int k = 0; k < 4; k++) f.reflect[k] = reflect[k];
      hipLaunchKernelGGL(coarsen32_kernel, dim3(grid_for(b->ncoarsen)), dim3(256), 0, s, f, F->wc, F->res8, b->ratio,
                         b->ncoarsen, b->co_cp, b->co_cl, b->co_fp, b->co_fl);
    }
  }
  return check_hip(hipGetLastError(), "FP32 stage launch");
}

}  // namespace

void fp32_release(Batch* b) {
  if (!b->f32) return;
  drop_graph32(b->f32);
  for (void* a : b->f32->allocs) batch_free(a);
  delete b->f32;
  b->f32 = nullptr;
}

int fp32_advance(Batch* b, int mode, const enl_gas_t* gas, const int* reflect, int order, const double* a1,
                 const double* a2, const double* be, double dt, int nsteps, hipStream_t s) {
  Fp32Form* F = b->f32;
  if (!F || !F->ready) return set_error(-2, "FP32 form not prepared (enl_batch_fp32_prepare)");
  if (order < 2) return set_error(-2, "the FP32 form needs an RK scheme of at least 2 stages");
  int rc = prepare_args32(b, gas, order, a1, a2, be, dt, s);
  if (rc) return rc;
  if (mode == ENL_BATCH_FP32) return enqueue32(b, reflect, order, nsteps, s);
  std::vector<unsigned char> key;
  auto put = [&key](const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    key.insert(key.end(), c, c + n);
  };
  put(&order, sizeof order);
  put(reflect, 4 * sizeof(int));
  put(b->tile_bx, sizeof b->tile_bx);
  put(b->tile_by, sizeof b->tile_by);
  put(b->pack_mode, sizeof b->pack_mode);
  put(b->pack_wpc, sizeof b->pack_wpc);
  put(&b->tiles, sizeof b->tiles);
  put(&b->packs, sizeof b->packs);
  put(&F->sw, sizeof F->sw);
  put(&F->lo, sizeof F->lo);
  put(&F->hi, sizeof F->hi);
  put(&F->every_stage, sizeof F->every_stage);
  put(&F->totals, sizeof F->totals);
  put(&F->sw_thr, sizeof F->sw_thr);
  if (!(F->exec && key == F->key)) {
    drop_graph32(F);
    hipStream_t cs;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return check_hip(e, "capture stream");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return check_hip(e, "hipStreamBeginCapture");
    }
    rc = enqueue32(b, reflect, order, 1, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&F->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      return check_hip(e, "hipGraphInstantiate");
    }
    F->graph = g;
    F->key = key;
  }
  for (int k = 0; k < nsteps; k++) {
    hipError_t e = hipGraphLaunch(F->exec, s);
    if (e != hipSuccess) return check_hip(e, "hipGraphLaunch");
  }
  return 0;
}

}  // namespace enl

using namespace enl;

extern "C" {

int enl_batch_set_ambient(void* handle, int level, int row_lo, int nrows, const double* rho, const double* rhos,
                          const double* rhoTn, const enl_gas_t* gas) {
  Batch* b = (Batch*)handle;
  if (!b || level < 0 || level >= kMaxLevels || nrows <= 0 || !rho || !rhos || !rhoTn || !gas)
    return set_error(-2, "enl_batch_set_ambient: bad arguments");
  int ref = -1;
  for (int p = 0; p < b->npatch && ref < 0; p++)
    if (b->desc[p].level == level) ref = p;
  if (ref < 0) return set_error(-2, "enl_batch_set_ambient: no patch on level %d", level);
  const enl_patch_desc_t& d = b->desc[ref];
  Fp32Form* F = form(b);
  F->ready = false;
  F->rho0 = gas->rho0;
  F->p0 = gas->p0;
  F->amb_lo[level] = row_lo;
  std::vector<double>& A = F->amb_h[level];
  A.assign((size_t)nrows * kRec, 0.0);
  for (int r = 0; r < nrows; r++) {
    const double y = d.xlo[1] + d.dx[1] * ((double)(row_lo + r - d.box[2]) + 0.5);
    const double w0 = rho[r], w3 = rhos[r], w4 = rhoTn[r];
    const double cy = gas->c + gas->c_lr * y;
    // update_state.m4:90-105, as eos_p_T evaluates it
    const double pr = (cy * cy) * ((w0 - gas->rho0) + ((gas->gamma1 - 1) / (2 * gas->rho0)) * ((w0 - gas->rho0) * (w0 - gas->rho0)) +
                                   (w0 / gas->c_p) * (w3 / w0)) + gas->p0;
    const double T = gas->T0 * std::exp((w3 / w0 - gas->R * std::log(gas->rho0 / w0)) / gas->c_v);
    double* rec = &A[(size_t)r * kRec];
    rec[0] = w0;
    rec[1] = w3;
    rec[2] = w4;
    rec[3] = pr;
    rec[4] = T;
    rec[5] = w3 / w0;
    rec[6] = w4 / w0;
    rec[7] = cy * cy;
  }
  return 0;
}

int enl_batch_fp32_prepare(void* handle, int variant, const enl_gas_t* gas, const int* reflect) {
  Batch* b = (Batch*)handle;
  if (!b || !gas || !reflect) return set_error(-2, "enl_batch_fp32_prepare: bad arguments");
  if (int rc = ensure_tiles(b)) return rc;
  if (variant != kHybridHighLlf) return set_error(-1, "the FP32 form implements rhs_hybrid_high_llf only");
  if (!b->f32) return set_error(-2, "enl_batch_fp32_prepare: no ambient set");
  for (int l = 0; l < b->nlevels; l++)
    if (b->level_cells[l] > 0 && b->f32->amb_h[l].empty()) return set_error(-2, "no ambient for level %d", l);
  int rc = setup32(b);
  if (rc) return rc;
  Fp32Form* F = b->f32;
  hipStream_t s = lib_stream();
  // K64 of the ambient: FP64 block 2 := ambient, its rhs into block 1
  for (int p = 0; p < b->npatch; p++)
    hipLaunchKernelGGL(ambient64_kernel, dim3(grid_for(F->hp[p].sk)), dim3(256), 0, s,
                       b->pool + b->hp[p].state[2], F->amb, F->hp[p]);
  rc = batch_eval_rhs64(b, variant, gas, reflect, 2, 1, s);
  if (rc) return rc;
  // K32 of the ambient: FP32 block 2 := 0 perturbation, ghosts filled, K32 into dk
  for (int p = 0; p < b->npatch; p++)
    hipLaunchKernelGGL(zero32_kernel, dim3(grid_for(7 * F->hp[p].sk)), dim3(256), 0, s, F->pool32 + F->hp[p].blk[2],
                       7 * F->hp[p].sk);
  std::vector<Args32> ev(b->npatch);
  for (int p = 0; p < b->npatch; p++) {
    Args32 a = base32(b, p, gas);
    a.in = F->pool32 + F->hp[p].blk[2];
    a.out = nullptr;
    a.kout = F->pool32 + F->hp[p].dk;
    if (F->dmap_off[p] >= 0) a.amb_weno = F->ambmap + F->dmap_off[p];
    ev[p] = a;
  }
  rc = check_hip(hipMemcpy(F->d_eval, ev.data(), sizeof(Args32) * ev.size(), hipMemcpyHostToDevice),
                 "FP32 eval arguments");
  for (int lev = 0; lev < b->nlevels && !rc; lev++) {
    fill32_level(b, lev, 2, 2, reflect, s);
    rc = stage32<true>(b, F->d_eval, lev, s);
  }
  if (rc) return rc;
  for (int p = 0; p < b->npatch; p++)
    hipLaunchKernelGGL(dk_kernel, dim3(grid_for(5L * F->hp[p].nx * F->hp[p].ny)), dim3(256), 0, s,
                       F->pool32 + F->hp[p].dk, b->pool + b->hp[p].state[1], F->hp[p]);
  rc = load32(b, s);
  if (!rc) rc = check_hip(hipStreamSynchronize(s), "enl_batch_fp32_prepare");
  F->ready = rc == 0;
  return rc;
}

int enl_batch_fp32_set_switch(void* handle, void* sw, float lo, float hi, int every_stage,
                              unsigned long long* totals) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "enl_batch_fp32_set_switch: no batch");
  if (sw && !(lo <= hi)) return set_error(-2, "enl_batch_fp32_set_switch: lo > hi");
  // the features of a 5 x 5 window read one cell past it
  if (sw && b->g < 3) return set_error(-2, "enl_batch_fp32_set_switch: the switch's feature window needs 3 ghost cells");
  if (totals && !is_device_ptr(totals)) return set_error(-4, "enl_batch_fp32_set_switch: totals must be device memory");
  Fp32Form* F = form(b);
  F->sw = (Switch*)sw;
  F->lo = sw ? lo : 0.0f;
  F->hi = sw ? hi : 1.0f;
  F->every_stage = every_stage ? 1 : 0;
  F->totals = sw ? totals : nullptr;
  if (F->sw && F->ready) return switch_reserve(F->sw, F->cells0, F->fcells0);
  return 0;
}

int enl_batch_fp32_decisions(void* handle, int pid, unsigned char* out) {
  Batch* b = (Batch*)handle;
  if (!b || !b->f32 || !b->f32->ready || !out) return set_error(-2, "enl_batch_fp32_decisions: bad arguments");
  Fp32Form* F = b->f32;
  if (pid < 0 || pid >= b->npatch || F->dmap_off[pid] < 0)
    return set_error(-2, "enl_batch_fp32_decisions: patch %d is not on level 0", pid);
  const Patch32& q = F->hp[pid];
  hipStream_t s = lib_stream();
  hipError_t e = hipMemcpyAsync(out, F->dmap + F->dmap_off[pid], (size_t)q.nx * q.ny, hipMemcpyDefault, s);
  if (e != hipSuccess) return check_hip(e, "enl_batch_fp32_decisions copy");
  return check_hip(hipStreamSynchronize(s), "enl_batch_fp32_decisions sync");
}

int enl_batch_fp32_load(void* handle) {
  Batch* b = (Batch*)handle;
  if (!b || !b->f32 || !b->f32->ready) return set_error(-2, "FP32 form not prepared");
  return load32(b, lib_stream());
}

int enl_batch_fp32_store(void* handle) {
  Batch* b = (Batch*)handle;
  if (!b || !b->f32 || !b->f32->ready) return set_error(-2, "FP32 form not prepared");
  Fp32Form* F = b->f32;
  hipStream_t s = lib_stream();
  for (int p = 0; p < b->npatch; p++) {
    const Patch32& q = F->hp[p];
    hipLaunchKernelGGL(to64_kernel, dim3(grid_for(q.sk)), dim3(256), 0, s, b->pool + b->hp[p].state[0],
                       F->pool32 + q.blk[0], F->wc, F->amb, q);
  }
  return check_hip(hipGetLastError(), "enl_batch_fp32_store");
}

}  // extern "C"
