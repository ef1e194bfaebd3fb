// enl_kernels.hip -- gfx950 kernels of Enlightning's per-patch update path.
//
//   rhs            fortran/rhs_*.m4 (nine variants, one template instance each)
//   runge_kutta    fortran/runge_kutta.m4:57-66
//   update_state   fortran/update_state.m4:87-107
//   tag_cells      fortran/tag_cells.m4:58-108
//   physical BC    enlightning.C:790-1013
//
// rhs design: one workgroup of 256 lanes (4 wave64) owns a 32 x 8 tile of
// interior cells.  The tile plus a 3-cell halo (38 x 14 cells) is staged once
// into LDS with coalesced row loads: the seven conserved/state fields and the
// four per-cell velocity-like quotients w_k/w_0.  Those quotients are the
// only divisions of the flux stencil, so they are computed once per staged
// cell instead of once per use (the reference recomputes them per stencil
// point; the value is the same IEEE quotient either way).  Each lane then
// evaluates its cell entirely out of LDS.
#include "enl_device.h"
#include "enl_kernels.h"

#pragma clang fp contract(off)

namespace enl {

namespace {

constexpr int kBX = 32;
constexpr int kBY = 8;
constexpr int kHalo = 3;
constexpr int kTW = kBX + 2 * kHalo;
constexpr int kTH = kBY + 2 * kHalo;
constexpr int kTN = kTW * kTH;

// Pointers into the LDS tile; element offsets are relative to a cell centre.
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

// Inviscid fluxes, rhs_hybrid_high_llf.m4:125-156.
__device__ inline double flux_f(const Tile& t, int e, int k) {
  double w0 = t.w0[e], u = t.u[e];
  switch (k) {
    case 0: return w0 * u;
    case 1: return w0 * sq(u) + t.p[e];
    case 2: return w0 * u * t.v[e];
    case 3: return w0 * u * t.s[e];
    default: return w0 * u * t.tn[e];
  }
}
__device__ inline double flux_g(const Tile& t, int e, int k) {
  double w0 = t.w0[e], v = t.v[e];
  switch (k) {
    case 0: return w0 * v;
    case 1: return w0 * v * t.u[e];
    case 2: return w0 * sq(v) + t.p[e];
    case 3: return w0 * v * t.s[e];
    default: return w0 * v * t.tn[e];
  }
}

// Physical source H, rhs_hybrid_high_llf.m4:159-237.  e is the centre cell,
// S the tile row stride.
__device__ inline void cell_source(const Tile& t, int e, int S, double H0, const RhsArgs& a, double H[5]) {
  const double dx = a.dx, dy = a.dy;
  const double* U = t.u;
  const double* V = t.v;
  const double* TT = t.T;
  H[0] = H0;

  double u_dprime_x = (U[e - 1] - 2 * U[e] + U[e + 1]) / sq(dx);
  double u_dprime_y = (U[e - S] - 2 * U[e] + U[e + S]) / sq(dy);
  double v_dprime_xy = (V[e + 1 + S] - V[e + 1 - S] - V[e - 1 + S] + V[e - 1 - S]) / (4 * dx * dy);
  double phi_xx_prime_x = kC4_3 * u_dprime_x - kC2_3 * v_dprime_xy;
  double phi_xy_prime_y = u_dprime_y + v_dprime_xy;
  H[1] = a.muB * (u_dprime_x + v_dprime_xy) + a.mu * (phi_xx_prime_x + phi_xy_prime_y);

  double v_dprime_y = (V[e - S] - 2 * V[e] + V[e + S]) / sq(dy);
  double v_dprime_x = (V[e - 1] - 2 * V[e] + V[e + 1]) / sq(dx);
  double u_dprime_yx = (U[e + 1 + S] - U[e + 1 - S] - U[e - 1 + S] + U[e - 1 - S]) / (4 * dx * dy);
  double phi_yx_prime_x = u_dprime_yx + v_dprime_x;
  double phi_yy_prime_y = kC4_3 * v_dprime_y - kC2_3 * u_dprime_yx;
  H[2] = a.muB * (v_dprime_y + u_dprime_yx) + a.mu * (phi_yx_prime_x + phi_yy_prime_y);

  double u_prime_x = (U[e + 1] - U[e - 1]) / (2 * dx);
  double u_prime_y = (U[e + S] - U[e - S]) / (2 * dy);
  double v_prime_y = (V[e + S] - V[e - S]) / (2 * dy);
  double v_prime_x = (V[e + 1] - V[e - 1]) / (2 * dx);
  double t_prime_x = (TT[e + 1] - TT[e - 1]) / (2 * dx);
  double t_prime_y = (TT[e + S] - TT[e - S]) / (2 * dy);
  double t_dprime_x = (TT[e - 1] - 2 * TT[e] + TT[e + 1]) / sq(dx);
  double t_dprime_y = (TT[e - S] - 2 * TT[e] + TT[e + S]) / sq(dy);
  double w0 = t.w0[e];
  double temp = TT[e];
  double T_n = t.tn[e];
  double Mn = (temp - T_n) / a.tau_n;
  double tmp1 = a.T_star_n / (T_n);
  double c_vn = kC078 * a.R * sq(tmp1) * exp(-tmp1);
  double A_n = ((temp / T_n) - 1) * c_vn;
  double phi_xx = kC4_3 * u_prime_x - kC2_3 * v_prime_y;
  double phi_xy_yx = u_prime_y + v_prime_x;
  double phi_yy = kC4_3 * v_prime_y - kC2_3 * u_prime_x;
  double sigma = (a.muB / temp) * sq(u_prime_x + v_prime_y) +
                 (a.mu / (2 * temp)) * (sq(phi_xx) + 2 * sq(phi_xy_yx) + sq(phi_yy)) +
                 (a.kappa / sq(temp)) * (sq(t_prime_x) + sq(t_prime_y)) + (w0 / temp) * (A_n * Mn);
  H[3] = sigma - (w0 / T_n) * c_vn * Mn +
         (-((a.kappa * sq(t_prime_x)) / sq(temp)) + ((a.kappa * t_dprime_x) / temp) -
          ((a.kappa * sq(t_prime_y)) / sq(temp)) + ((a.kappa * t_dprime_y) / temp));
  H[4] = (w0 / a.tau_n) * (temp - T_n);
}

// Hybrid density sensor, rhs_hybrid_high_llf.m4:239-256.
__device__ inline int cell_sensor(const Tile& t, int e, int S, double ep, int level) {
  int tag = 1;
  if (level == 0) {
    tag = 0;
    const double* r = t.w0;
#pragma unroll
    for (int q = -1; q <= 1; q++) {
      double rx = smooth_ratio(r[e + q] - r[e - 1 + q], r[e + 1 + q] - r[e + q], ep);
      if (rx < kThresh) tag = 1;
      double ry = smooth_ratio(r[e + q * S] - r[e + (q - 1) * S], r[e + (q + 1) * S] - r[e + q * S], ep);
      if (ry < kThresh) tag = 1;
    }
  }
  return tag;
}

template <int V>
__device__ inline void cell_drp(const Tile& t, int e, int S, const RhsArgs& a, const double H[5], double out[5]) {
  // rhs_hybrid_high_llf.m4:258-268
#pragma unroll
  for (int k = 0; k < 5; k++) {
    out[k] = (-kA3 * flux_f(t, e - 3, k) - kA2 * flux_f(t, e - 2, k) - kA1 * flux_f(t, e - 1, k) +
              kA1 * flux_f(t, e + 1, k) + kA2 * flux_f(t, e + 2, k) + kA3 * flux_f(t, e + 3, k)) /
                 (-a.dx) +
             (-kA3 * flux_g(t, e - 3 * S, k) - kA2 * flux_g(t, e - 2 * S, k) - kA1 * flux_g(t, e - S, k) +
              kA1 * flux_g(t, e + S, k) + kA2 * flux_g(t, e + 2 * S, k) + kA3 * flux_g(t, e + 3 * S, k)) /
                 (-a.dy) +
             H[k];
  }
}

template <int V>
__device__ inline void cell_weno(const Tile& t, int e, int S, int j, const RhsArgs& a, const double H[5],
                                 double out[5]) {
  using Sc = Scheme<V>;
  double af, ag;
  if (Sc::llf) {
    // rhs_hybrid_high_llf.m4:270-284: per-cell max wave speed over the window
    af = 0;
    ag = 0;
#pragma unroll
    for (int q = -Sc::llf_reach; q <= Sc::llf_reach; q++) {
      double height = a.ylo + a.dy * ((double)(j) + 0.5);
      double vf = fabs(t.u[e + q]) + a.c + a.c_lr * height;
      if (vf > af) af = vf;
      height = a.ylo + a.dy * ((double)(j + q) + 0.5);
      double vg = fabs(t.v[e + q * S]) + a.c + a.c_lr * height;
      if (vg > ag) ag = vg;
    }
  } else {
    af = a.alpha;
    ag = a.alpha;
  }
  constexpr int R = Sc::order == 5 ? 3 : 2;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double* wk = t.w[k];
    double fpx[7], fmx[7], fpy[7], fmy[7];
#pragma unroll
    for (int m = -R; m <= R; m++) {
      double Fx = flux_f(t, e + m, k), Gy = flux_g(t, e + m * S, k);
      fpx[m + 3] = 0.5 * (Fx + af * wk[e + m]);
      fmx[m + 3] = 0.5 * (Fx - af * wk[e + m]);
      fpy[m + 3] = 0.5 * (Gy + ag * wk[e + m * S]);
      fmy[m + 3] = 0.5 * (Gy - ag * wk[e + m * S]);
    }
    double pf1, mf1, pf2, mf2, pg1, mg1, pg2, mg2;
    if (Sc::order == 5) {
      pf1 = wenop5(fpx[0], fpx[1], fpx[2], fpx[3], fpx[4]);
      mf1 = wenom5(fmx[1], fmx[2], fmx[3], fmx[4], fmx[5]);
      pf2 = wenop5(fpx[1], fpx[2], fpx[3], fpx[4], fpx[5]);
      mf2 = wenom5(fmx[2], fmx[3], fmx[4], fmx[5], fmx[6]);
      pg1 = wenop5(fpy[0], fpy[1], fpy[2], fpy[3], fpy[4]);
      mg1 = wenom5(fmy[1], fmy[2], fmy[3], fmy[4], fmy[5]);
      pg2 = wenop5(fpy[1], fpy[2], fpy[3], fpy[4], fpy[5]);
      mg2 = wenom5(fmy[2], fmy[3], fmy[4], fmy[5], fmy[6]);
    } else {
      pf1 = wenop3(fpx[1], fpx[2], fpx[3]);
      mf1 = wenom3(fmx[2], fmx[3], fmx[4]);
      pf2 = wenop3(fpx[2], fpx[3], fpx[4]);
      mf2 = wenom3(fmx[3], fmx[4], fmx[5]);
      pg1 = wenop3(fpy[1], fpy[2], fpy[3]);
      mg1 = wenom3(fmy[2], fmy[3], fmy[4]);
      pg2 = wenop3(fpy[2], fpy[3], fpy[4]);
      mg2 = wenom3(fmy[3], fmy[4], fmy[5]);
    }
    out[k] = (pf1 + mf1 - pf2 - mf2) / a.dx + (pg1 + mg1 - pg2 - mg2) / a.dy + H[k];
  }
}

template <int V>
__global__ void __launch_bounds__(kBX* kBY) rhs_tile_kernel(RhsArgs a) {
  __shared__ double sw[5][kTN];
  __shared__ double sp[kTN], sT[kTN], su[kTN], sv[kTN], ss[kTN], stn[kTN];
  const Geom& G = a.geom;
  const int tid = threadIdx.x + threadIdx.y * kBX;
  const int i0 = blockIdx.x * kBX, j0 = blockIdx.y * kBY;

  // Stage the tile and its halo (cells outside the ghosted array are skipped;
  // no interior cell of the patch reads them).
  for (int e = tid; e < kTN; e += kBX * kBY) {
    int gi = i0 + (e % kTW) - kHalo, gj = j0 + (e / kTW) - kHalo;
    if (gi >= -G.g0 && gi < G.nx + G.g0 && gj >= -G.g1 && gj < G.ny + G.g1) {
      long gidx = (long)(gi + G.g0) + (long)(gj + G.g1) * G.sy;
      double w0 = a.w[gidx], w1 = a.w[gidx + G.sk], w2 = a.w[gidx + 2 * G.sk];
      double w3 = a.w[gidx + 3 * G.sk], w4 = a.w[gidx + 4 * G.sk];
      sw[0][e] = w0;
      sw[1][e] = w1;
      sw[2][e] = w2;
      sw[3][e] = w3;
      sw[4][e] = w4;
      sp[e] = a.p[gidx];
      sT[e] = a.T[gidx];
      su[e] = w1 / w0;
      sv[e] = w2 / w0;
      ss[e] = w3 / w0;
      stn[e] = w4 / w0;
    }
  }
  __syncthreads();

  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  if (i >= G.nx || j >= G.ny) return;
  Tile t;
  t.w0 = sw[0];
  for (int k = 0; k < 5; k++) t.w[k] = sw[k];
  t.p = sp;
  t.T = sT;
  t.u = su;
  t.v = sv;
  t.s = ss;
  t.tn = stn;
  const int S = kTW;
  const int e = (threadIdx.y + kHalo) * kTW + threadIdx.x + kHalo;
  const long icell = (long)i + (long)j * G.nx;
  const long ik = (long)G.nx * G.ny;

  double H[5];
  cell_source(t, e, S, a.H0[(long)i + (long)j * a.h0_sy], a, H);
  using Sc = Scheme<V>;
  int tag = Sc::hybrid ? cell_sensor(t, e, S, a.ep, a.level) : (Sc::weno ? 1 : 0);
  double out[5];
  if (Sc::drp && tag == 0) {
    cell_drp<V>(t, e, S, a, H, out);
  } else {
    cell_weno<V>(t, e, S, j, a, H, out);
  }
#pragma unroll
  for (int k = 0; k < 5; k++) a.KK[icell + k * ik] = out[k];
  if (a.switch_map) a.switch_map[icell] = tag;
}

// runge_kutta.m4:57-66.  blockIdx.y = row j + ny * component k.
__global__ void __launch_bounds__(256) rk_kernel(RkArgs a) {
  const Geom& G = a.geom;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;
  for (int row = blockIdx.y; row < G.ny * 5; row += gridDim.y) {
    int j = row % G.ny, k = row / G.ny;
    long gidx = (long)(i + G.g0) + (long)(j + G.g1) * G.sy + (long)k * G.sk;
    long iidx = (long)i + (long)j * G.nx + (long)k * G.nx * G.ny;
    a.w[gidx] = a.alpha_1 * a.ww[iidx] + a.alpha_2 * a.w[gidx] + a.beta * a.dt * a.KK[iidx];
  }
}

// update_state.m4:87-107.
__global__ void __launch_bounds__(256) state_kernel(StateArgs a) {
  const Geom& G = a.geom;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;
  for (int j = blockIdx.y; j < G.ny; j += gridDim.y) {
    double xc1 = a.ylo + a.dy * ((double)(j) + 0.5);
    long gidx = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
    double w0 = a.w[gidx], w3 = a.w[gidx + 3 * G.sk];
    a.p[gidx] = sq(a.c + a.c_lr * xc1) *
                    ((w0 - a.rho0) + ((a.gamma1 - 1) / (2 * a.rho0)) * sq(w0 - a.rho0) + (w0 / a.c_p) * (w3 / w0)) +
                a.p0;
    a.T[gidx] = a.T0 * exp((w3 / w0 - a.R * log(a.rho0 / w0)) / a.c_v);
  }
}

// tag_cells.m4:58-108: the four passes only ever set tags, so a cell's tag
// is refine_tag_val if any of the four ratios is below the threshold.
__global__ void __launch_bounds__(256) tag_kernel(TagArgs a) {
  const Geom& G = a.geom;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;
  const double* h = a.H0;
  const double* r = a.w;
  const long S = G.sy;
  for (int j = blockIdx.y; j < G.ny; j += gridDim.y) {
    long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
    bool t = false;
    t |= smooth_ratio(h[g] - h[g - 1], h[g + 1] - h[g], a.ep0) < kThresh;
    t |= smooth_ratio(h[g] - h[g - S], h[g + S] - h[g], a.ep0) < kThresh;
    t |= smooth_ratio(r[g] - r[g - 1], r[g + 1] - r[g], a.ep1) < kThresh;
    t |= smooth_ratio(r[g] - r[g - S], r[g + S] - r[g], a.ep1) < kThresh;
    a.tags[(long)i + (long)j * G.nx] = t ? a.refine_tag_val : 0;
  }
}

__device__ inline void bc_copy(const BcArgs& a, long dst, long src, int negk) {
  const Geom& G = a.geom;
#pragma unroll
  for (int s = 0; s < 3; s++)
    if (a.f[s]) a.f[s][dst] = a.f[s][src];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double v = a.w[src + k * G.sk];
    a.w[dst + k * G.sk] = (k == negk) ? -v : v;
  }
}

// Edge strips, enlightning.C:836-944.  Every ghost layer ends up equal to
// the first interior cell (the reference copies layer by layer outward),
// with the normal momentum negated on REFLECT sides.  One lane per ghost
// cell of a strip: blockIdx.y picks the side.
__global__ void __launch_bounds__(256) bc_edge_kernel(BcArgs a) {
  const Geom& G = a.geom;
  int side = blockIdx.y;
  if (!a.touch[side]) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int along = side < 2 ? G.ny : G.nx;
  int width = side < 2 ? G.g0 : G.g1;
  if (n >= along * width) return;
  int t = n % along, l = n / along;  // tangential index, layer (0 = nearest)
  int i, j, si, sj, negk;
  switch (side) {
    case 0: i = -1 - l; j = t; si = 0; sj = t; negk = 1; break;
    case 1: i = G.nx + l; j = t; si = G.nx - 1; sj = t; negk = 1; break;
    case 2: i = t; j = -1 - l; si = t; sj = 0; negk = 2; break;
    default: i = t; j = G.ny + l; si = t; sj = G.ny - 1; negk = 2; break;
  }
  if (!a.reflect[side]) negk = -1;
  long dst = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
  long src = (long)(si + G.g0) + (long)(sj + G.g1) * G.sy;
  bc_copy(a, dst, src, negk);
}

// Corner blocks (node boundaries), enlightning.C:948-1008: after the edges,
// each corner ghost copies along x from the column next to the interior
// (i = 0 for the left corners, nx-1 for the right ones) and negates x
// momentum on a REFLECT left/right side.  blockIdx.y = node 0..3
// (lower-left, lower-right, upper-left, upper-right).
__global__ void __launch_bounds__(256) bc_corner_kernel(BcArgs a) {
  const Geom& G = a.geom;
  int node = blockIdx.y;
  int right = node & 1, upper = node >> 1;
  if (!(a.touch[right ? 1 : 0] || a.touch[upper ? 3 : 2])) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.g0 * G.g1) return;
  int li = n % G.g0, lj = n / G.g0;
  int i = right ? G.nx + li : -1 - li;
  int j = upper ? G.ny + lj : -1 - lj;
  int si = right ? G.nx - 1 : 0;
  int negk = a.reflect[right ? 1 : 0] ? 1 : -1;
  long dst = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
  long src = (long)(si + G.g0) + (long)(j + G.g1) * G.sy;
  bc_copy(a, dst, src, negk);
}

__global__ void __launch_bounds__(256) copy_interior_kernel(double* ww, const double* w, Geom G) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;
  for (int row = blockIdx.y; row < G.ny * 5; row += gridDim.y) {
    int j = row % G.ny, k = row / G.ny;
    ww[(long)i + (long)j * G.nx + (long)k * G.nx * G.ny] =
        w[(long)(i + G.g0) + (long)(j + G.g1) * G.sy + (long)k * G.sk];
  }
}

// Rows per launch of the row-looping elementwise kernels (grid.y is capped;
// each block strides over the remaining rows).
inline unsigned rows_grid(int rows) { return (unsigned)(rows < 16384 ? rows : 16384); }

template <int V>
hipError_t launch_rhs_t(const RhsArgs& a, hipStream_t s) {
  dim3 block(kBX, kBY);
  dim3 grid((a.geom.nx + kBX - 1) / kBX, (a.geom.ny + kBY - 1) / kBY);
  hipLaunchKernelGGL(rhs_tile_kernel<V>, grid, block, 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_rhs(int variant, const RhsArgs& a, hipStream_t s) {
  if (a.geom.nx <= 0 || a.geom.ny <= 0) return hipSuccess;
  switch (variant) {
    case kHybridHighLlf: return launch_rhs_t<kHybridHighLlf>(a, s);
    case kHybridHigh: return launch_rhs_t<kHybridHigh>(a, s);
    case kHybridLowLlf: return launch_rhs_t<kHybridLowLlf>(a, s);
    case kHybridLow: return launch_rhs_t<kHybridLow>(a, s);
    case kWenoHighLlf: return launch_rhs_t<kWenoHighLlf>(a, s);
    case kWenoHigh: return launch_rhs_t<kWenoHigh>(a, s);
    case kWenoLowLlf: return launch_rhs_t<kWenoLowLlf>(a, s);
    case kWenoLow: return launch_rhs_t<kWenoLow>(a, s);
    case kDrp: return launch_rhs_t<kDrp>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_runge_kutta(const RkArgs& a, hipStream_t s) {
  if (a.geom.nx <= 0 || a.geom.ny <= 0) return hipSuccess;
  dim3 grid((a.geom.nx + 255) / 256, rows_grid(a.geom.ny * 5));
  hipLaunchKernelGGL(rk_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_update_state(const StateArgs& a, hipStream_t s) {
  if (a.geom.nx <= 0 || a.geom.ny <= 0) return hipSuccess;
  dim3 grid((a.geom.nx + 255) / 256, rows_grid(a.geom.ny));
  hipLaunchKernelGGL(state_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_tag_cells(const TagArgs& a, hipStream_t s) {
  if (a.geom.nx <= 0 || a.geom.ny <= 0) return hipSuccess;
  dim3 grid((a.geom.nx + 255) / 256, rows_grid(a.geom.ny));
  hipLaunchKernelGGL(tag_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_physical_bc(const BcArgs& a, hipStream_t s) {
  const Geom& G = a.geom;
  if (G.nx <= 0 || G.ny <= 0) return hipSuccess;
  int longest = (G.nx > G.ny ? G.nx : G.ny) * (G.g0 > G.g1 ? G.g0 : G.g1);
  if (longest > 0) {
    hipLaunchKernelGGL(bc_edge_kernel, dim3((longest + 255) / 256, 4), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (G.g0 * G.g1 > 0) {
    hipLaunchKernelGGL(bc_corner_kernel, dim3((G.g0 * G.g1 + 255) / 256, 4), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_copy_interior(double* ww, const double* w, const Geom& g, hipStream_t s) {
  if (g.nx <= 0 || g.ny <= 0) return hipSuccess;
  dim3 grid((g.nx + 255) / 256, rows_grid(g.ny * 5));
  hipLaunchKernelGGL(copy_interior_kernel, grid, dim3(256), 0, s, ww, w, g);
  return hipGetLastError();
}

}  // namespace enl
