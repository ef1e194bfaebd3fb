// enl_device.h -- device-side building blocks of the per-patch update path
// for gfx950 (CDNA4).  Every expression keeps the operation order of the
// reference Fortran (fortran/rhs_*.m4, update_state.m4, tag_cells.m4) so
// that, compiled with -ffp-contract=off and IEEE division, results agree
// with the reference arithmetic except for exp/log rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace enl {

// Fortran default-REAL literals are folded in single precision and widened.
constexpr double kC13_12 = (double)(13.0f / 12.0f);
constexpr double kW01 = (double)0.1f;
constexpr double kW06 = (double)0.6f;
constexpr double kW03 = (double)0.3f;
constexpr double kEpsWeno = (double)10.0e-6f;
constexpr double kC1_6 = (double)(1.0f / 6.0f);
constexpr double kC1_3 = (double)(1.0f / 3.0f);
constexpr double kC2_3 = (double)(2.0f / 3.0f);
constexpr double kC4_3 = (double)(4.0f / 3.0f);
constexpr double kA1 = (double)0.79926643f;
constexpr double kA2 = (double)-0.18941314f;
constexpr double kA3 = (double)0.02651995f;
constexpr double kC078 = (double)0.78f;
constexpr double kC09 = (double)0.9f;
constexpr double kThresh = (double)0.9999f;

// rhs variants, one per fortran/rhs_*.m4 file.
enum Variant : int {
  kHybridHighLlf = 0,
  kHybridHigh = 1,
  kHybridLowLlf = 2,
  kHybridLow = 3,
  kWenoHighLlf = 4,
  kWenoHigh = 5,
  kWenoLowLlf = 6,
  kWenoLow = 7,
  kDrp = 8,
  kNumVariants = 9
};

// Compile-time description of a variant (what differs between the nine
// Fortran files; the flux and source blocks are common to all of them).
template <int V>
struct Scheme {
  static constexpr bool hybrid = V <= kHybridLow;
  static constexpr bool drp = V <= kHybridLow || V == kDrp;
  static constexpr bool weno = V != kDrp;
  static constexpr int order = (V == kHybridLowLlf || V == kHybridLow || V == kWenoLowLlf || V == kWenoLow) ? 3 : 5;
  static constexpr bool llf = (V == kHybridHighLlf || V == kHybridLowLlf || V == kWenoHighLlf || V == kWenoLowLlf);
  // flux stencil half width: rhs_weno_low*.m4 build F(0:4) (q=-2..2)
  static constexpr int reach = (V == kWenoLowLlf || V == kWenoLow) ? 2 : 3;
  // LLF window half width: q=-3..3 (*_high_llf) or -2..2 (*_low_llf)
  static constexpr int llf_reach = order == 5 ? 3 : 2;
};

__host__ __device__ inline double sq(double x) { return x * x; }

// The WENO combination sum_k (zeta_k / sigma) f_k with zeta_k = c_k / D_k
// and sigma = sum_k zeta_k, evaluated with one division: with
// n_k = c_k * prod_{j != k} D_j the weights are n_k / sum_j n_j, and the
// combination is f_1 + (sum_k n_k (f_k - f_1)) / (sum_k n_k).  The large
// common part f_1 enters once, through the last addition, so the rounding
// error is that of the small differences plus one rounding of the sum:
// against an extended-precision evaluation it is 1.5-2.4x smaller in RMS
// than the reference's expressions on smooth, stratified and shocked data,
// with no larger bias.  Where the products leave the finite range (D beyond
// ~1e100, or non-finite data) the reference's expressions are evaluated
// instead.
__device__ inline double weno_combine(double c0, double c1, double c2, double D0, double D1, double D2, double f0,
                                      double f1, double f2) {
#pragma clang fp contract(fast)
  const double n0 = c0 * (D1 * D2), n1 = c1 * (D0 * D2), n2 = c2 * (D0 * D1);
  const double r = f1 + (n0 * (f0 - f1) + n2 * (f2 - f1)) / (n0 + n1 + n2);
  if (__builtin_isfinite(r)) return r;
  const double z0 = c0 / D0, z1 = c1 / D1, z2 = c2 / D2;
  const double sigma = z0 + z1 + z2;
  return (z0 / sigma) * f0 + (z1 / sigma) * f1 + (z2 / sigma) * f2;
}
__device__ inline double weno_combine(double c0, double c1, double D0, double D1, double f0, double f1) {
#pragma clang fp contract(fast)
  const double n0 = c0 * D1, n1 = c1 * D0;
  const double r = f1 + n0 * (f0 - f1) / (n0 + n1);
  if (__builtin_isfinite(r)) return r;
  const double z0 = c0 / D0, z1 = c1 / D1;
  const double sigma = z0 + z1;
  return (z0 / sigma) * f0 + (z1 / sigma) * f1;
}

// fortran/rhs_hybrid_high_llf.m4:365-398
__device__ inline double wenom5(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
#pragma clang fp contract(fast)
  double beta0 = kC13_12 * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  double beta1 = kC13_12 * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  double beta2 = kC13_12 * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  double f0 = kC1_6 * (11 * fg0 - 7 * fgp1 + 2 * fgp2);
  double f1 = kC1_6 * (2 * fgm1 + 5 * fg0 - fgp1);
  double f2 = kC1_6 * (-fgm2 + 5 * fgm1 + 2 * fg0);
  return weno_combine(kW01, kW06, kW03, sq(kEpsWeno + beta0), sq(kEpsWeno + beta1), sq(kEpsWeno + beta2), f0, f1,
                      f2);
}

// fortran/rhs_hybrid_high_llf.m4:400-433
__device__ inline double wenop5(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
#pragma clang fp contract(fast)
  double beta0 = kC13_12 * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  double beta1 = kC13_12 * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  double beta2 = kC13_12 * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  double f0 = kC1_6 * (2 * fg0 + 5 * fgp1 - fgp2);
  double f1 = kC1_6 * (-fgm1 + 5 * fg0 + 2 * fgp1);
  double f2 = kC1_6 * (2 * fgm2 - 7 * fgm1 + 11 * fg0);
  return weno_combine(kW03, kW06, kW01, sq(kEpsWeno + beta0), sq(kEpsWeno + beta1), sq(kEpsWeno + beta2), f0, f1,
                      f2);
}

// fortran/rhs_hybrid_low_llf.m4:349-373
__device__ inline double wenom3(double fgm1, double fg0, double fgp1) {
#pragma clang fp contract(fast)
  double beta0 = sq(fgp1 - fg0);
  double beta1 = sq(fg0 - fgm1);
  double f0 = 0.5 * (3 * fg0 - fgp1);
  double f1 = 0.5 * (fgm1 + fg0);
  return weno_combine(kC1_3, kC2_3, sq(kEpsWeno + beta0), sq(kEpsWeno + beta1), f0, f1);
}

// fortran/rhs_hybrid_low_llf.m4:375-399
__device__ inline double wenop3(double fgm1, double fg0, double fgp1) {
#pragma clang fp contract(fast)
  double beta0 = sq(fgp1 - fg0);
  double beta1 = sq(fg0 - fgm1);
  double f0 = 0.5 * (fg0 + fgp1);
  double f1 = 0.5 * (-fgm1 + 3 * fg0);
  return weno_combine(kC2_3, kC1_3, sq(kEpsWeno + beta0), sq(kEpsWeno + beta1), f0, f1);
}

// Smoothness ratio of the hybrid sensor and of tag_cells
// (rhs_hybrid_high_llf.m4:246, tag_cells.m4:71).
__device__ inline double smooth_ratio(double tmp1, double tmp2, double ep) {
  return (2 * fabs(tmp1 * tmp2) + ep) / (sq(tmp1) + sq(tmp2) + ep);
}

// ep = 0.9*tol*(xi**2)/(1-0.9*(tol**2)), xi = 1 (rhs_hybrid_high_llf.m4:114-115)
__host__ __device__ inline double sensor_eps(double tol) {
  double xi = 1;
  return kC09 * tol * sq(xi) / (1 - kC09 * sq(tol));
}

// FP64 division through a shared reciprocal.  hipcc expands n / d into
// v_div_scale of both operands (range scaling), v_rcp_f64 and two Newton
// steps on the reciprocal, the quotient n*r, its residual, the corrected
// quotient (v_div_fmas) and v_div_fixup (special operands).  When neither
// operand needs scaling, both v_div_scale return their input, v_div_fmas is
// a plain fma and v_div_fixup returns the quotient: rcp_nr and div_nr below
// are then that same sequence, so div_nr(n, d, rcp_nr(d)) == n / d bit for
// bit, and one reciprocal serves every quotient by d.  No scaling happens
// when |d| is in [2^-50, 2^50] (den_ok) and n is zero or |n| is in
// [2^-960, 2^700] (num_ok): the exponents differ by less than 768, the
// exponent of n is above the sequence's tiny-numerator bound (2^-969) and
// 1/d, n/d are normal.  A zero numerator gives +0 (n / d gives -0 for
// n = -0, d > 0; the two compare equal).
__device__ inline double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
__device__ inline double div_nr(double n, double d, double r) {
  const double q = n * r;
  const double e = fma(-d, q, n);
  return fma(e, r, q);
}
__host__ __device__ inline bool den_ok(double d) {
  const double m = fabs(d);
  return m >= 0x1p-50 && m <= 0x1p50;
}
__host__ __device__ inline bool num_ok(double n) {
  const double m = fabs(n);
  return (m >= 0x1p-960 && m <= 0x1p700) || n == 0;
}

// Pressure and temperature of one cell, update_state.m4:88-105 (xc1 is the
// cell-centre height xylo(1)+dxy(1)*(j-ifirst1+0.5)).
__device__ inline void eos_p_T(double c, double c_lr, double rho0, double gamma1, double c_p, double p0, double T0,
                               double c_v, double R, double xc1, double w0, double w3, double& p, double& T) {
  p = sq(c + c_lr * xc1) * ((w0 - rho0) + ((gamma1 - 1) / (2 * rho0)) * sq(w0 - rho0) + (w0 / c_p) * (w3 / w0)) + p0;
  T = T0 * exp((w3 / w0 - R * log(rho0 / w0)) / c_v);
}

// eos_p_T with its quotients through rcp_nr / div_nr where the operands
// allow it, else through eos_p_T; the same p and T.  k1 is
// (gamma1 - 1) / (2 * rho0) formed on the host (IEEE, as the device forms
// it); consts_ok says rho0 passes num_ok and c_p, c_v pass den_ok.
__device__ inline void eos_p_T_nr(double c, double c_lr, double rho0, double gamma1, double c_p, double p0,
                                  double T0, double c_v, double R, double k1, bool consts_ok, double xc1, double w0,
                                  double w3, double& p, double& T) {
  if (consts_ok && den_ok(w0) && num_ok(w3)) {
    const double rw = rcp_nr(w0);
    const double s = div_nr(w3, w0, rw);
    const double x = s - R * log(div_nr(rho0, w0, rw));
    if (num_ok(x)) {
      p = sq(c + c_lr * xc1) * ((w0 - rho0) + k1 * sq(w0 - rho0) + div_nr(w0, c_p, rcp_nr(c_p)) * s) + p0;
      T = T0 * exp(div_nr(x, c_v, rcp_nr(c_v)));
      return;
    }
  }
  eos_p_T(c, c_lr, rho0, gamma1, c_p, p0, T0, c_v, R, xc1, w0, w3, p, T);
}

}  // namespace enl
