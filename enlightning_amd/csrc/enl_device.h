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

// fortran/rhs_hybrid_high_llf.m4:365-398
__device__ inline double wenom5(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
#pragma clang fp contract(fast)
  double beta0 = kC13_12 * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  double beta1 = kC13_12 * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  double beta2 = kC13_12 * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  double zeta0 = kW01 / sq(kEpsWeno + beta0);
  double zeta1 = kW06 / sq(kEpsWeno + beta1);
  double zeta2 = kW03 / sq(kEpsWeno + beta2);
  double sigma = zeta0 + zeta1 + zeta2;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma, omega2 = zeta2 / sigma;
  double f0 = kC1_6 * (11 * fg0 - 7 * fgp1 + 2 * fgp2);
  double f1 = kC1_6 * (2 * fgm1 + 5 * fg0 - fgp1);
  double f2 = kC1_6 * (-fgm2 + 5 * fgm1 + 2 * fg0);
  return omega0 * f0 + omega1 * f1 + omega2 * f2;
}

// fortran/rhs_hybrid_high_llf.m4:400-433
__device__ inline double wenop5(double fgm2, double fgm1, double fg0, double fgp1, double fgp2) {
#pragma clang fp contract(fast)
  double beta0 = kC13_12 * sq(fg0 - 2 * fgp1 + fgp2) + 0.25 * sq(3 * fg0 - 4 * fgp1 + fgp2);
  double beta1 = kC13_12 * sq(fgm1 - 2 * fg0 + fgp1) + 0.25 * sq(fgm1 - fgp1);
  double beta2 = kC13_12 * sq(fgm2 - 2 * fgm1 + fg0) + 0.25 * sq(fgm2 - 4 * fgm1 + 3 * fg0);
  double zeta0 = kW03 / sq(kEpsWeno + beta0);
  double zeta1 = kW06 / sq(kEpsWeno + beta1);
  double zeta2 = kW01 / sq(kEpsWeno + beta2);
  double sigma = zeta0 + zeta1 + zeta2;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma, omega2 = zeta2 / sigma;
  double f0 = kC1_6 * (2 * fg0 + 5 * fgp1 - fgp2);
  double f1 = kC1_6 * (-fgm1 + 5 * fg0 + 2 * fgp1);
  double f2 = kC1_6 * (2 * fgm2 - 7 * fgm1 + 11 * fg0);
  return omega0 * f0 + omega1 * f1 + omega2 * f2;
}

// fortran/rhs_hybrid_low_llf.m4:349-373
__device__ inline double wenom3(double fgm1, double fg0, double fgp1) {
#pragma clang fp contract(fast)
  double beta0 = sq(fgp1 - fg0);
  double beta1 = sq(fg0 - fgm1);
  double zeta0 = kC1_3 / sq(kEpsWeno + beta0);
  double zeta1 = kC2_3 / sq(kEpsWeno + beta1);
  double sigma = zeta0 + zeta1;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma;
  double f0 = 0.5 * (3 * fg0 - fgp1);
  double f1 = 0.5 * (fgm1 + fg0);
  return omega0 * f0 + omega1 * f1;
}

// fortran/rhs_hybrid_low_llf.m4:375-399
__device__ inline double wenop3(double fgm1, double fg0, double fgp1) {
#pragma clang fp contract(fast)
  double beta0 = sq(fgp1 - fg0);
  double beta1 = sq(fg0 - fgm1);
  double zeta0 = kC2_3 / sq(kEpsWeno + beta0);
  double zeta1 = kC1_3 / sq(kEpsWeno + beta1);
  double sigma = zeta0 + zeta1;
  double omega0 = zeta0 / sigma, omega1 = zeta1 / sigma;
  double f0 = 0.5 * (fg0 + fgp1);
  double f1 = 0.5 * (-fgm1 + 3 * fg0);
  return omega0 * f0 + omega1 * f1;
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

// Pressure and temperature of one cell, update_state.m4:88-105 (xc1 is the
// cell-centre height xylo(1)+dxy(1)*(j-ifirst1+0.5)).
__device__ inline void eos_p_T(double c, double c_lr, double rho0, double gamma1, double c_p, double p0, double T0,
                               double c_v, double R, double xc1, double w0, double w3, double& p, double& T) {
  p = sq(c + c_lr * xc1) * ((w0 - rho0) + ((gamma1 - 1) / (2 * rho0)) * sq(w0 - rho0) + (w0 / c_p) * (w3 / w0)) + p0;
  T = T0 * exp((w3 / w0 - R * log(rho0 / w0)) / c_v);
}

}  // namespace enl
