// enl_rhs.hip -- gfx950 kernels of the reference's rhs subroutine
// (fortran/rhs_*.m4, nine variants; one template instance per variant).
//
// Three kernels, chosen per call by variant and level:
//
//   rhs_split_kernel    hybrid variants on level 0, and rhs_drp.  A 256-lane
//                       workgroup owns a 32 x 8 tile; tile + 3-cell halo of
//                       rho, p, T and the four quotients u=w1/w0, v=w2/w0,
//                       s=w3/w0, Tn=w4/w0 is staged in LDS (29.8 KB).  Each
//                       lane computes the physical source H and the sensor;
//                       DRP cells finish here.  WENO cells store H in their
//                       rhs slots and are appended (wave ballot + one atomic
//                       per wave) to a compact list, so no wave runs the ~10x
//                       more expensive WENO path for a few of its lanes.
//   rhs_weno_list_kernel  one lane per listed WENO cell, stencil read from
//                       global memory (the flagged cells cluster along shock
//                       fronts, so the reads hit L2), direction by direction.
//   rhs_dense_kernel    every cell WENO (rhs_weno_* variants, and hybrid
//                       variants on levels > 0 where the sensor is off):
//                       tile staging as above plus the raw w_k, all lanes on
//                       the WENO path.
//
// All three keep the reference operation order (rhs_hybrid_high_llf.m4),
// compile with -ffp-contract=off and use IEEE division, so each result is
// the reference's value.  Two exact rewrites speed them up:
//   * when dx and dy are powers of two, x/c for the grid constants c (dx^2,
//     4 dx dy, 2 dx, -dx, ...) equals x*(1/c) exactly, since 1/c is exact;
//   * the sensor only needs the sign of r - 0.9999f: a single-precision
//     estimate decides it unless it lies within 1e-6 of the threshold, where
//     the exact double quotient is computed (same decision either way).
#include "enl_rhs_body.h"

#pragma clang fp contract(off)

namespace enl {
namespace {

constexpr int kBX = 32;
constexpr int kBY = 8;
using Sh = TileShape<kBX, kBY>;

// LDS for one tile: `raw` adds the four raw momenta etc. (dense kernel).
#define ENL_TILE_LDS(L, raw)                                             \
  __shared__ double lds_w0[Sh::TN], lds_p[Sh::TN], lds_T[Sh::TN];        \
  __shared__ double lds_u[Sh::TN], lds_v[Sh::TN], lds_s[Sh::TN], lds_tn[Sh::TN]; \
  TileLds<kBX, kBY> L;                                                   \
  L.w[0] = lds_w0;                                                       \
  L.p = lds_p;                                                           \
  L.T = lds_T;                                                           \
  L.u = lds_u;                                                           \
  L.v = lds_v;                                                           \
  L.s = lds_s;                                                           \
  L.tn = lds_tn;                                                         \
  L.w[1] = L.w[2] = L.w[3] = L.w[4] = nullptr;                           \
  if (raw) {                                                             \
    __shared__ double lds_raw[4][Sh::TN];                                \
    for (int k_ = 1; k_ < 5; k_++) L.w[k_] = lds_raw[k_ - 1];            \
  }

// Tile origin of this workgroup: a 1-D grid of nbx * nby tiles, in XCD
// chunks of vertical strips (strip_tile).
#define ENL_TILE_ORIGIN(a, i0, j0)                                      \
  const int nbx_ = (a.geom.nx + kBX - 1) / kBX;                         \
  const int nby_ = (a.geom.ny + kBY - 1) / kBY;                         \
  const int2 t_ = strip_tile((int)blockIdx.x, nbx_, nby_);              \
  const int i0 = t_.x * kBX, j0 = t_.y * kBY;

template <int V, bool P, int FUSED>
__global__ void __launch_bounds__(Sh::threads) __attribute__((amdgpu_waves_per_eu(5))) rhs_split_kernel(RhsArgs a) {
  ENL_TILE_LDS(L, false)
  ENL_TILE_ORIGIN(a, i0, j0)
  split_body<V, P, FUSED, kBX, kBY>(a, i0, j0, 0, L);
}
// The fused form carries the runge_kutta and update_state epilogue.  Both
// split kernels are capped at 5 waves per SIMD (<= 102 VGPRs), what the LDS
// of a tile allows anyway (5 tiles of 29.8 KB per CU); the cap also gives
// the better schedule (0.74 vs 0.82 ms per fused 4096^2 stage).
template <int V, bool P, int FUSED>
__global__ void __launch_bounds__(Sh::threads) __attribute__((amdgpu_waves_per_eu(5)))
rhs_split_fused_kernel(RhsArgs a) {
  ENL_TILE_LDS(L, false)
  ENL_TILE_ORIGIN(a, i0, j0)
  split_body<V, P, FUSED, kBX, kBY>(a, i0, j0, 0, L);
}

// WENO cells of the list, ten lanes each (list_block).
template <int V, bool P, int FUSED>
__global__ void __launch_bounds__(256) rhs_weno_list_kernel(RhsArgs a) {
  list_block<V, P, FUSED>([&](int) -> const RhsArgs& { return a; }, a.list, *a.count, a.bc_fold != 0);
}

// Every cell WENO: the two directions on two thread planes (dense2_body),
// block (kBX, kBY, 2).
template <int V, bool P, int FUSED>
__global__ void __launch_bounds__(2 * Sh::threads) __attribute__((amdgpu_waves_per_eu(4))) rhs_dense_kernel(RhsArgs a) {
  ENL_TILE_LDS(L, true)
  __shared__ double lds_x[10 * Sh::threads];
  ENL_TILE_ORIGIN(a, i0, j0)
  dense2_body<V, P, FUSED, kBX, kBY>(a, i0, j0, L, lds_x);
}

__global__ void reset_count_kernel(unsigned* count) { *count = 0; }

bool is_pow2(double x) {
  if (!(x > 0) || !std::isfinite(x)) return false;
  int e;
  return std::frexp(x, &e) == 0.5;
}

template <int V, bool P, int F>
hipError_t launch_t(const RhsArgs& a, hipStream_t s) {
  using Sc = Scheme<V>;
  const Geom& G = a.geom;
  dim3 block(kBX, kBY);
  dim3 grid((unsigned)(((G.nx + kBX - 1) / kBX) * ((G.ny + kBY - 1) / kBY)));
  if constexpr (!Sc::drp) {
    // rhs_weno_*: every cell on the WENO path
    hipLaunchKernelGGL((rhs_dense_kernel<V, P, F>), grid, dim3(kBX, kBY, 2), 0, s, a);
  } else if constexpr (!Sc::hybrid) {
    // rhs_drp: every cell on the DRP path
    if constexpr (F)
      hipLaunchKernelGGL((rhs_split_fused_kernel<V, P, F>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((rhs_split_kernel<V, P, F>), grid, block, 0, s, a);
  } else if (a.level != 0) {
    // hybrid above level 0: the sensor is off, every cell WENO
    hipLaunchKernelGGL((rhs_dense_kernel<V, P, F>), grid, dim3(kBX, kBY, 2), 0, s, a);
  } else {
    if (!a.list || !a.count) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_count_kernel, dim3(1), dim3(1), 0, s, a.count);
    if constexpr (F)
      hipLaunchKernelGGL((rhs_split_fused_kernel<V, P, F>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((rhs_split_kernel<V, P, F>), grid, block, 0, s, a);
    // a grid that covers the machine; blocks stride over the list
    long cells = (long)G.nx * G.ny;
    long blocks = (cells + kListCellsPerBlock - 1) / kListCellsPerBlock;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL((rhs_weno_list_kernel<V, P, F>), dim3((unsigned)blocks), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

template <int V, int F>
hipError_t launch_v(const RhsArgs& a, hipStream_t s) {
  if (is_pow2(a.dx) && is_pow2(a.dy)) return launch_t<V, true, F>(a, s);
  return launch_t<V, false, F>(a, s);
}

template <int F>
hipError_t launch_any(int variant, const RhsArgs& a0, hipStream_t s) {
  if (a0.geom.nx <= 0 || a0.geom.ny <= 0) return hipSuccess;
  RhsArgs a = a0;
  fill_rhs_constants(a);
  switch (variant) {
    case kHybridHighLlf: return launch_v<kHybridHighLlf, F>(a, s);
    case kHybridHigh: return launch_v<kHybridHigh, F>(a, s);
    case kHybridLowLlf: return launch_v<kHybridLowLlf, F>(a, s);
    case kHybridLow: return launch_v<kHybridLow, F>(a, s);
    case kWenoHighLlf: return launch_v<kWenoHighLlf, F>(a, s);
    case kWenoHigh: return launch_v<kWenoHigh, F>(a, s);
    case kWenoLowLlf: return launch_v<kWenoLowLlf, F>(a, s);
    case kWenoLow: return launch_v<kWenoLow, F>(a, s);
    case kDrp: return launch_v<kDrp, F>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

bool rhs_needs_list(int variant, int level) {
  return variant <= kHybridLow && level == 0;
}

hipError_t launch_rhs(int variant, const RhsArgs& a, hipStream_t s) { return launch_any<kRhs>(variant, a, s); }

hipError_t launch_stage(int variant, const RhsArgs& a, hipStream_t s) {
  if (!a.ww || !a.w_out || !a.p_out || !a.T_out || a.w_out == a.w) return hipErrorInvalidValue;
  return launch_any<kStageBc>(variant, a, s);
}

}  // namespace enl
