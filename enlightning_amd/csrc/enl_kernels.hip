// enl_kernels.hip -- gfx950 kernels of Enlightning's per-patch update path.
//
//   runge_kutta    fortran/runge_kutta.m4:57-66
//   update_state   fortran/update_state.m4:87-107
//   tag_cells      fortran/tag_cells.m4:58-108
//   physical BC    enlightning.C:790-1013
//
// The rhs kernels live in enl_rhs.hip.
#include <stdint.h>

#include "enl_device.h"
#include "enl_kernels.h"

#pragma clang fp contract(off)

namespace enl {

namespace {

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

// 16-byte vector forms of the streaming kernels: each lane owns two
// consecutive cells of a row (one double2 per array and component).  Valid
// when nx and g0 are even and every base pointer is 16-byte aligned, so the
// interior of each ghosted row starts on a 16-byte boundary.
__global__ void __launch_bounds__(256) rk_kernel_v(RkArgs a) {
  const Geom& G = a.geom;
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= G.nx / 2) return;
  const double bdt = a.beta * a.dt;  // beta*dt*KK groups as (beta*dt)*KK
  const long ik = (long)G.nx * G.ny;
  for (int row = blockIdx.y; row < G.ny * 5; row += gridDim.y) {
    const int j = row % G.ny, k = row / G.ny;
    const long g = (long)(2 * ip + G.g0) + (long)(j + G.g1) * G.sy + k * G.sk;
    const long c = (long)(2 * ip) + (long)j * G.nx + k * ik;
    double2 w = *(const double2*)(a.w + g);
    // ww and KK are read once per step: non-temporal loads keep them from
    // displacing w, which the next kernels read again
    double2 v, K;
    v.x = __builtin_nontemporal_load(a.ww + c);
    v.y = __builtin_nontemporal_load(a.ww + c + 1);
    K.x = __builtin_nontemporal_load(a.KK + c);
    K.y = __builtin_nontemporal_load(a.KK + c + 1);
    double2 o;
    o.x = a.alpha_1 * v.x + a.alpha_2 * w.x + bdt * K.x;
    o.y = a.alpha_1 * v.y + a.alpha_2 * w.y + bdt * K.y;
    *(double2*)(a.w + g) = o;
  }
}

__device__ inline void state_cell(const StateArgs& a, double xc1, double w0, double w3, double& p, double& T) {
  eos_p_T(a.c, a.c_lr, a.rho0, a.gamma1, a.c_p, a.p0, a.T0, a.c_v, a.R, xc1, w0, w3, p, T);
}

__global__ void __launch_bounds__(256) state_kernel_v(StateArgs a) {
  const Geom& G = a.geom;
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= G.nx / 2) return;
  for (int j = blockIdx.y; j < G.ny; j += gridDim.y) {
    const double xc1 = a.ylo + a.dy * ((double)(j) + 0.5);
    const long g = (long)(2 * ip + G.g0) + (long)(j + G.g1) * G.sy;
    double2 w0 = *(const double2*)(a.w + g);
    double2 w3 = *(const double2*)(a.w + g + 3 * G.sk);
    double2 p, T;
    state_cell(a, xc1, w0.x, w3.x, p.x, T.x);
    state_cell(a, xc1, w0.y, w3.y, p.y, T.y);
    *(double2*)(a.p + g) = p;
    *(double2*)(a.T + g) = T;
  }
}

// blockIdx.y = row j + ny * component k (one row of one component per
// block and iteration: more, shorter blocks stream faster than five
// components per block).
__global__ void __launch_bounds__(256) copy_interior_kernel_v(double* ww, const double* w, Geom G) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= G.nx / 2) return;
  const long ik = (long)G.nx * G.ny;
  for (int row = blockIdx.y; row < G.ny * 5; row += gridDim.y) {
    const int j = row % G.ny, k = row / G.ny;
    const long g = (long)(2 * ip + G.g0) + (long)(j + G.g1) * G.sy + k * G.sk;
    const long c = (long)(2 * ip) + (long)j * G.nx + k * ik;
    *(double2*)(ww + c) = *(const double2*)(w + g);
  }
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }
bool vec_ok(const Geom& G) { return (G.nx % 2) == 0 && (G.g0 % 2) == 0; }

// update_state.m4:87-107.
__global__ void __launch_bounds__(256) state_kernel(StateArgs a) {
  const Geom& G = a.geom;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;
  for (int j = blockIdx.y; j < G.ny; j += gridDim.y) {
    double xc1 = a.ylo + a.dy * ((double)(j) + 0.5);
    long gidx = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
    state_cell(a, xc1, a.w[gidx], a.w[gidx + 3 * G.sk], a.p[gidx], a.T[gidx]);
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

}  // namespace

hipError_t launch_runge_kutta(const RkArgs& a, hipStream_t s) {
  if (a.geom.nx <= 0 || a.geom.ny <= 0) return hipSuccess;
  if (vec_ok(a.geom) && aligned16(a.w) && aligned16(a.ww) && aligned16(a.KK)) {
    dim3 grid((a.geom.nx / 2 + 255) / 256, rows_grid(a.geom.ny * 5));
    hipLaunchKernelGGL(rk_kernel_v, grid, dim3(256), 0, s, a);
    return hipGetLastError();
  }
  dim3 grid((a.geom.nx + 255) / 256, rows_grid(a.geom.ny * 5));
  hipLaunchKernelGGL(rk_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_update_state(const StateArgs& a, hipStream_t s) {
  if (a.geom.nx <= 0 || a.geom.ny <= 0) return hipSuccess;
  if (vec_ok(a.geom) && aligned16(a.w) && aligned16(a.p) && aligned16(a.T)) {
    dim3 grid((a.geom.nx / 2 + 255) / 256, rows_grid(a.geom.ny));
    hipLaunchKernelGGL(state_kernel_v, grid, dim3(256), 0, s, a);
    return hipGetLastError();
  }
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
  if (vec_ok(g) && aligned16(ww) && aligned16(w)) {
    dim3 grid((g.nx / 2 + 255) / 256, rows_grid(g.ny * 5));
    hipLaunchKernelGGL(copy_interior_kernel_v, grid, dim3(256), 0, s, ww, w, g);
    return hipGetLastError();
  }
  dim3 grid((g.nx + 255) / 256, rows_grid(g.ny * 5));
  hipLaunchKernelGGL(copy_interior_kernel, grid, dim3(256), 0, s, ww, w, g);
  return hipGetLastError();
}

}  // namespace enl
