// enl_kernels.h -- launch interface of the HIP kernels of the per-patch
// update path (internal C++ API; the exported C ABI sits in enl_abi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "enlightning_hip.h"
#include <stdint.h>

namespace enl {

// Geometry of one patch in the SAMRAI CellData layout: interior nx x ny,
// ghost widths g0, g1; ghosted arrays have row stride sy = nx + 2*g0 and
// component stride sk = sy * (ny + 2*g1).  Interior-only arrays use
// row stride nx and component stride nx*ny.
struct Geom {
  int nx, ny, g0, g1;
  long sy, sk;
};

inline Geom make_geom(int ifirst0, int ilast0, int ifirst1, int ilast1, int g0, int g1) {
  Geom g;
  g.nx = ilast0 - ifirst0 + 1;
  g.ny = ilast1 - ifirst1 + 1;
  g.g0 = g0;
  g.g1 = g1;
  g.sy = g.nx + 2L * g0;
  g.sk = g.sy * (g.ny + 2L * g1);
  return g;
}

struct RhsArgs {
  const double* w;   // ghosted, 5 components
  const double* p;   // ghosted
  const double* T;   // ghosted
  const double* H0;  // source at interior cell (0,0); row stride h0_sy
  long h0_sy;        // nx for an interior array, sy for a ghosted one
  double* KK;        // interior, 5 components
  int* switch_map;   // optional: interior int map of the DRP(0)/WENO(1) choice
  // optional, level 0 of the hybrid variants: interior map of a switch
  // decided elsewhere (0 DRP, 1 WENO, 2 undecided: the sensor decides)
  const unsigned char* decide = nullptr;
  Geom geom;
  double dx, dy, ylo;
  double alpha, kappa, mu, muB, R, T_star_n, tau_n, c, c_lr;
  double ep;  // sensor epsilon from tag_tolerance(2)
  // 5e-5 * ep when ep > 0, else -1: a cell whose eight sensor differences
  // have squares summing below it is DRP (filled by fill_rhs_constants)
  double ep_pre;
  int level;
  // WENO-cell list of the split hybrid path (nx*ny ints) and its length;
  // needed when rhs_needs_list(variant, level)
  unsigned long long* list;  // entries (patch << 32) | (i + j*nx)
  unsigned* count;
  // reciprocals of the grid constants (filled by launch_rhs / launch_stage)
  double inv_dx2, inv_dy2, inv_4dxdy, inv_2dx, inv_2dy, inv_mdx, inv_mdy, inv_dx, inv_dy;
  // launch_stage only: instead of storing rhs, each cell applies the
  // runge_kutta update (runge_kutta.m4:60-63) and the state update
  // (update_state.m4:87-107) and stores the new state into w_out, p_out,
  // T_out (ghosted arrays, interior written; must not alias w, p, T).
  const double* ww;  // state at the start of the step: element (i,j,k) at
                     // ww[ww_off + i + j*ww_sy + k*ww_sk]
  long ww_off, ww_sy, ww_sk;
  double* w_out;
  double* p_out;
  double* T_out;
  double alpha_1, alpha_2, bdt;  // bdt = beta*dt
  double rho0, gamma1, c_p, p0, T0, c_v;
  // filled with the reciprocals: (gamma1 - 1) / (2 * rho0), and whether the
  // per-run operands of the source / update_state quotients admit div_nr
  double eos_k1;
  bool fast_div, eos_ok;
  // fused stage on a patch whose four sides are physical: nonzero makes the
  // cells next to a side also write their new state into that side's ghost
  // layers, as launch_physical_bc would before the next stage; bit 1 + k set
  // when side k (left, right, bottom, top) reflects
  int bc_fold = 0;
};

struct RkArgs {
  double* w;          // ghosted, updated on the interior
  const double* ww;   // interior
  const double* KK;   // interior
  Geom geom;
  double alpha_1, alpha_2, beta, dt;
};

struct StateArgs {
  const double* w;  // ghosted
  double* p;        // ghosted, interior written
  double* T;        // ghosted, interior written
  Geom geom;
  double dy, ylo;
  double rho0, c, gamma1, c_p, p0, T0, c_lr, c_v, R;
};

struct TagArgs {
  int* tags;          // interior
  const double* w;    // ghosted
  const double* H0;   // ghosted
  Geom geom;
  int refine_tag_val;
  double ep0, ep1;    // from tag_tolerance(0), tag_tolerance(1)
};

struct BcArgs {
  double* w;       // ghosted, 5 components
  double* f[3];    // ghosted scalars (pres, temp, source); entries may be null
  Geom geom;
  int touch[4];    // side on the physical boundary: left, right, bottom, top
  int reflect[4];  // REFLECT (1) or ABSORB (0) per side
};

hipError_t launch_rhs(int variant, const RhsArgs& a, hipStream_t s);
// One fused method-of-lines stage: rhs + runge_kutta + update_state.
hipError_t launch_stage(int variant, const RhsArgs& a, hipStream_t s);
bool rhs_needs_list(int variant, int level);
hipError_t launch_runge_kutta(const RkArgs& a, hipStream_t s);
hipError_t launch_update_state(const StateArgs& a, hipStream_t s);
hipError_t launch_tag_cells(const TagArgs& a, hipStream_t s);
hipError_t launch_physical_bc(const BcArgs& a, hipStream_t s);
// ww(interior) <- w(interior), 5 components
hipError_t launch_copy_interior(double* ww, const double* w, const Geom& g, hipStream_t s);

// Library runtime state shared by the exported entry points (enl_abi.hip).
hipStream_t lib_stream();

// The learned switch (enl_switch.hip): decisions of one patch state (0 DRP,
// 1 WENO, 2 undecided) and the running counts of its decisions.
struct Switch;
int switch_decide(Switch* sw, const Geom& G, const double* w_n, const double* pres, const enl_gas_t* gas, float lo,
                  float hi, unsigned char* decide, hipStream_t s);
int switch_accumulate(Switch* sw, unsigned long long* totals, hipStream_t s);
int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);
bool is_device_ptr(const void* p);
// Absorbing envelopes (enl_absorb.hip): the cells each ABSORB side touches
// (normal index range, interior-relative; left / right may skip the first
// row) and their tmp table, built on the host; uploaded to a per-device
// table and applied by one launch per side.
struct AbsorbSide {
  int lo = 0, hi = -1;
  bool skip_first = false;
  int off = 0;
};
struct AbsorbPlan {
  AbsorbSide side[4];
  std::vector<double> tab;
};
int absorb_plan(const Geom& G, const double* patch_xlo, const double* dx, const double* grid_xlo,
                const double* grid_xhi, const int* absorb, double width, AbsorbPlan* out);
int absorb_upload(const AbsorbPlan& plan, const double** dtab, hipStream_t s);
hipError_t launch_absorb(const AbsorbPlan& plan, const double* dtab, double* w, double* p, double* T, const Geom& G,
                         const enl_gas_t* gas, hipStream_t s);
// Device workspace of the current device (grow-only, owned by the library):
// the WENO-cell list for `cells` cells and its counter.
int rhs_workspace(long cells, unsigned long long** list, unsigned** count);

}  // namespace enl
