// enl_driver.hip -- host-side patch loop around the kernels for a patch that
// covers the whole domain: the per-patch sequence of enlightning::singleStep
// (enlightning.C:459-583) inside the SAMRAI MethodOfLinesIntegrator stage loop
// (copy current -> scratch, then per stage: ghost fill incl. the physical
// boundary of enlightning.C:790-1013, rhs, runge_kutta, update_state).
// Everything stays on the device; the host only enqueues launches.
#include <hip/hip_runtime.h>

#include <vector>

#include "enl_device.h"
#include "enl_kernels.h"
#include "enlightning_hip.h"

#pragma clang fp contract(off)

using namespace enl;

namespace {

// Per-kernel-class HIP event timing of the patch loop (enl_profile_start /
// enl_profile_stop).  Events are recorded on the launch stream around each
// launch; elapsed times are summed when the profile is stopped.
struct KernelTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<int> kind;  // kind of the launch bracketed by events 2i, 2i+1
  bool overflow = false;

  void start(int capacity) {
    while ((int)pool.size() < 2 * capacity) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) break;
      pool.push_back(e);
    }
    used = 0;
    kind.clear();
    overflow = false;
    on = true;
  }
  // Returns the event pair index or -1.
  long begin(int k, hipStream_t s) {
    if (!on) return -1;
    if (used + 2 > pool.size()) {
      overflow = true;
      return -1;
    }
    long idx = (long)used / 2;
    (void)hipEventRecord(pool[used], s);
    used += 2;
    kind.push_back(k);
    return idx;
  }
  void end(long idx, hipStream_t s) {
    if (idx >= 0) (void)hipEventRecord(pool[2 * idx + 1], s);
  }
};
KernelTimer g_timer;

}  // namespace

extern "C" int enl_profile_start(int capacity) {
  if (capacity < 1) return set_error(-2, "profile capacity must be positive");
  g_timer.start(capacity);
  return 0;
}

extern "C" int enl_profile_stop(double* ms, int* count, int nkinds) {
  g_timer.on = false;
  for (int k = 0; k < nkinds; k++) {
    ms[k] = 0;
    count[k] = 0;
  }
  for (size_t i = 0; i < g_timer.kind.size(); i++) {
    hipEvent_t a = g_timer.pool[2 * i], b = g_timer.pool[2 * i + 1];
    if (hipEventSynchronize(b) != hipSuccess) return set_error(-5, "hipEventSynchronize");
    float t = 0;
    if (hipEventElapsedTime(&t, a, b) != hipSuccess) return set_error(-5, "hipEventElapsedTime");
    int k = g_timer.kind[i];
    if (k < nkinds) {
      ms[k] += t;
      count[k] += 1;
    }
  }
  return g_timer.overflow ? set_error(-7, "profile capacity exceeded; later launches were not timed") : 0;
}

// Absorbing envelopes of a patch that covers the whole domain: the sides
// whose reflect flag is 0 (ABSORB) when absorb_width > 0; the domain is
// [xlo, xlo + n*dx] in each direction.
static int whole_domain_absorb(const Geom& G, const double* xlo, const double* dx, const int* reflect, double absorb_width,
                               AbsorbPlan* plan, const double** dtab, bool* on, hipStream_t s) {
  int ab[4];
  bool any = false;
  for (int k = 0; k < 4; k++) {
    ab[k] = reflect[k] == 0;
    any = any || ab[k];
  }
  *on = any && absorb_width > 0;
  if (!*on) return 0;
  const double xhi[2] = {xlo[0] + dx[0] * (double)G.nx, xlo[1] + dx[1] * (double)G.ny};
  int rc = absorb_plan(G, xlo, dx, xlo, xhi, ab, absorb_width, plan);
  return rc ? rc : absorb_upload(*plan, dtab, s);
}

extern "C" int enl_advance_patch(int variant, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw,
                                 const double* xlo, const double* dx, const enl_gas_t* gas, const int* reflect,
                                 int order, const double* alpha_1, const double* alpha_2, const double* beta,
                                 double dt, int nsteps, int level, double* w_n, double* pres, double* temp,
                                 double* source, double* ww, double* KK, double absorb_width) {
  if (variant < 0 || variant >= kNumVariants) return set_error(-1, "unknown rhs variant");
  int need = (variant == kWenoLowLlf || variant == kWenoLow) ? 2 : 3;
  if (gcw < need) return set_error(-2, "ghost width below the stencil reach");
  if (ilast0 < ifirst0 || ilast1 < ifirst1) return set_error(-2, "empty or invalid patch box");
  if (order < 1 || order > 8) return set_error(-2, "RK order out of range");
  const double* arrays[6] = {w_n, pres, temp, source, ww, KK};
  for (int i = 0; i < 6; i++)
    if (!is_device_ptr(arrays[i])) return set_error(-4, "enl_advance_patch needs device arrays");
  hipStream_t s = lib_stream();
  Geom G = make_geom(ifirst0, ilast0, ifirst1, ilast1, gcw, gcw);

  RhsArgs ra = {};
  ra.w = w_n;
  ra.p = pres;
  ra.T = temp;
  // rhs reads the source through the interior context (enlightning.C:475-477):
  // point it at interior cell (0,0) of the ghosted source array.
  ra.H0 = source + G.g0 + (long)G.g1 * G.sy;
  ra.h0_sy = G.sy;
  ra.KK = KK;
  ra.switch_map = nullptr;
  ra.geom = G;
  ra.dx = dx[0];
  ra.dy = dx[1];
  ra.ylo = xlo[1];
  ra.alpha = gas->weno_alpha;
  ra.kappa = gas->kappa;
  ra.mu = gas->mu;
  ra.muB = gas->muB;
  ra.R = gas->R;
  ra.T_star_n = gas->T_star_n;
  ra.tau_n = gas->tau_n;
  ra.c = gas->c;
  ra.c_lr = gas->c_lr;
  ra.ep = sensor_eps(gas->tag_tolerance[2]);
  ra.level = level;
  ra.list = nullptr;
  ra.count = nullptr;
  if (rhs_needs_list(variant, level)) {
    int rc = rhs_workspace((long)G.nx * G.ny, &ra.list, &ra.count);
    if (rc) return rc;
  }


  RkArgs rk;
  rk.w = w_n;
  rk.ww = ww;
  rk.KK = KK;
  rk.geom = G;
  rk.dt = dt;

  StateArgs st;
  st.w = w_n;
  st.p = pres;
  st.T = temp;
  st.geom = G;
  st.dy = dx[1];
  st.ylo = xlo[1];
  st.rho0 = gas->rho0;
  st.c = gas->c;
  st.gamma1 = gas->gamma1;
  st.c_p = gas->c_p;
  st.p0 = gas->p0;
  st.T0 = gas->T0;
  st.c_lr = gas->c_lr;
  st.c_v = gas->c_v;
  st.R = gas->R;

  BcArgs bc;
  bc.w = w_n;
  bc.f[0] = pres;
  bc.f[1] = temp;
  bc.f[2] = source;
  bc.geom = G;
  for (int k = 0; k < 4; k++) {
    bc.touch[k] = 1;
    bc.reflect[k] = reflect[k];
  }

  AbsorbPlan plan;
  const double* dtab = nullptr;
  bool absorb_on = false;
  int rc = whole_domain_absorb(G, xlo, dx, reflect, absorb_width, &plan, &dtab, &absorb_on, s);
  if (rc) return rc;
  long t;
  hipError_t e = hipSuccess;
  for (int step = 0; step < nsteps && e == hipSuccess; step++) {
    t = g_timer.begin(ENL_KERNEL_COPY, s);
    e = launch_copy_interior(ww, w_n, G, s);
    g_timer.end(t, s);
    for (int stage = 0; stage < order && e == hipSuccess; stage++) {
      rk.alpha_1 = alpha_1[stage];
      rk.alpha_2 = alpha_2[stage];
      rk.beta = beta[stage];
      t = g_timer.begin(ENL_KERNEL_BC, s);
      e = launch_physical_bc(bc, s);
      g_timer.end(t, s);
      if (e != hipSuccess) break;
      t = g_timer.begin(ENL_KERNEL_RHS, s);
      e = launch_rhs(variant, ra, s);
      g_timer.end(t, s);
      if (e != hipSuccess) break;
      t = g_timer.begin(ENL_KERNEL_RK, s);
      e = launch_runge_kutta(rk, s);
      g_timer.end(t, s);
      if (e != hipSuccess) break;
      t = g_timer.begin(ENL_KERNEL_STATE, s);
      e = launch_update_state(st, s);
      g_timer.end(t, s);
      if (absorb_on && e == hipSuccess) e = launch_absorb(plan, dtab, w_n, pres, temp, G, gas, s);
    }
  }
  if (e != hipSuccess) rc = check_hip(e, "patch loop launch");
  return rc;
}

extern "C" size_t enl_fused_scratch_bytes(int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw) {
  Geom G = make_geom(ifirst0, ilast0, ifirst1, ilast1, gcw, gcw);
  return sizeof(double) * (size_t)(2 * 7 * G.sk + 5L * G.nx * G.ny);
}

namespace {
// The learned switch of a fused advance, when one is given.
struct LearnedSwitch {
  Switch* sw = nullptr;
  float lo = 0, hi = 1;
  bool every_stage = true;
  unsigned char* decide = nullptr;
  unsigned long long* totals = nullptr;
};

int advance_fused(const LearnedSwitch& L, int variant, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw,
                  const double* xlo, const double* dx, const enl_gas_t* gas, const int* reflect, int order,
                  const double* alpha_1, const double* alpha_2, const double* beta, double dt, int nsteps, int level,
                  double* w_n, double* pres, double* temp, double* source, void* scratch, size_t scratch_bytes,
                  double absorb_width);
}  // namespace

extern "C" int enl_advance_patch_fused(int variant, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw,
                                       const double* xlo, const double* dx, const enl_gas_t* gas, const int* reflect,
                                       int order, const double* alpha_1, const double* alpha_2, const double* beta,
                                       double dt, int nsteps, int level, double* w_n, double* pres, double* temp,
                                       double* source, void* scratch, size_t scratch_bytes, double absorb_width) {
  return advance_fused(LearnedSwitch(), variant, ifirst0, ilast0, ifirst1, ilast1, gcw, xlo, dx, gas, reflect, order,
                       alpha_1, alpha_2, beta, dt, nsteps, level, w_n, pres, temp, source, scratch, scratch_bytes,
                       absorb_width);
}

extern "C" int enl_advance_patch_learned(void* sw, float lo, float hi, int every_stage, unsigned char* decide,
                                         unsigned long long* totals, int variant, int ifirst0, int ilast0,
                                         int ifirst1, int ilast1, int gcw, const double* xlo, const double* dx,
                                         const enl_gas_t* gas, const int* reflect, int order, const double* alpha_1,
                                         const double* alpha_2, const double* beta, double dt, int nsteps,
                                         double* w_n, double* pres, double* temp, double* source, void* scratch,
                                         size_t scratch_bytes, double absorb_width) {
  if (!sw || !decide) return set_error(-2, "enl_advance_patch_learned: no switch or decision map");
  if (!(variant == kHybridHighLlf || variant == kHybridHigh || variant == kHybridLowLlf || variant == kHybridLow))
    return set_error(-1, "the learned switch replaces the sensor of the hybrid variants only");
  if (!(lo <= hi)) return set_error(-2, "enl_advance_patch_learned: lo > hi");
  // the features of a 5 x 5 window read one cell past it: 3 ghost cells
  if (gcw < 3) return set_error(-2, "enl_advance_patch_learned: the switch's feature window needs 3 ghost cells");
  if (!is_device_ptr(decide) || (totals && !is_device_ptr(totals)))
    return set_error(-4, "enl_advance_patch_learned needs device decision and count arrays");
  LearnedSwitch L;
  L.sw = (Switch*)sw;
  L.lo = lo;
  L.hi = hi;
  L.every_stage = every_stage != 0;
  L.decide = decide;
  L.totals = totals;
  return advance_fused(L, variant, ifirst0, ilast0, ifirst1, ilast1, gcw, xlo, dx, gas, reflect, order, alpha_1,
                       alpha_2, beta, dt, nsteps, 0, w_n, pres, temp, source, scratch, scratch_bytes, absorb_width);
}

namespace {
int advance_fused(const LearnedSwitch& L, int variant, int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw,
                  const double* xlo, const double* dx, const enl_gas_t* gas, const int* reflect, int order,
                  const double* alpha_1, const double* alpha_2, const double* beta, double dt, int nsteps, int level,
                  double* w_n, double* pres, double* temp, double* source, void* scratch, size_t scratch_bytes,
                  double absorb_width) {
  if (variant < 0 || variant >= kNumVariants) return set_error(-1, "unknown rhs variant");
  int need = (variant == kWenoLowLlf || variant == kWenoLow) ? 2 : 3;
  if (gcw < need) return set_error(-2, "ghost width below the stencil reach");
  if (ilast0 < ifirst0 || ilast1 < ifirst1) return set_error(-2, "empty or invalid patch box");
  if (order < 1 || order > 8) return set_error(-2, "RK order out of range");
  if (scratch_bytes < enl_fused_scratch_bytes(ifirst0, ilast0, ifirst1, ilast1, gcw))
    return set_error(-2, "fused scratch too small");
  const void* arrays[5] = {w_n, pres, temp, source, scratch};
  for (int i = 0; i < 5; i++)
    if (!is_device_ptr(arrays[i])) return set_error(-4, "enl_advance_patch_fused needs device arrays");
  hipStream_t s = lib_stream();
  Geom G = make_geom(ifirst0, ilast0, ifirst1, ilast1, gcw, gcw);

  // states: A = caller's arrays, B and C in the scratch, then the rhs slots
  struct State {
    double *w, *p, *T;
  };
  double* sc = (double*)scratch;
  State A{w_n, pres, temp};
  State B{sc, sc + 5 * G.sk, sc + 6 * G.sk};
  State C{sc + 7 * G.sk, sc + 12 * G.sk, sc + 13 * G.sk};
  double* KK = sc + 14 * G.sk;

  RhsArgs ra = {};
  ra.H0 = source + G.g0 + (long)G.g1 * G.sy;
  ra.h0_sy = G.sy;
  ra.KK = KK;
  ra.switch_map = nullptr;
  ra.geom = G;
  ra.dx = dx[0];
  ra.dy = dx[1];
  ra.ylo = xlo[1];
  ra.alpha = gas->weno_alpha;
  ra.kappa = gas->kappa;
  ra.mu = gas->mu;
  ra.muB = gas->muB;
  ra.R = gas->R;
  ra.T_star_n = gas->T_star_n;
  ra.tau_n = gas->tau_n;
  ra.c = gas->c;
  ra.c_lr = gas->c_lr;
  ra.ep = sensor_eps(gas->tag_tolerance[2]);
  ra.level = level;
  ra.list = nullptr;
  ra.count = nullptr;
  if (rhs_needs_list(variant, level)) {
    int rc = rhs_workspace((long)G.nx * G.ny, &ra.list, &ra.count);
    if (rc) return rc;
  }
  ra.ww = A.w;  // the step's start state is A's interior
  ra.ww_off = G.g0 + (long)G.g1 * G.sy;
  ra.ww_sy = G.sy;
  ra.ww_sk = G.sk;
  ra.rho0 = gas->rho0;
  ra.gamma1 = gas->gamma1;
  ra.c_p = gas->c_p;
  ra.p0 = gas->p0;
  ra.T0 = gas->T0;
  ra.c_v = gas->c_v;

  BcArgs bc;
  bc.f[2] = source;
  bc.geom = G;
  for (int k = 0; k < 4; k++) {
    bc.touch[k] = 1;
    bc.reflect[k] = reflect[k];
  }

  AbsorbPlan plan;
  const double* dtab = nullptr;
  bool absorb_on = false;
  if (int rc = whole_domain_absorb(G, xlo, dx, reflect, absorb_width, &plan, &dtab, &absorb_on, s)) return rc;
  // Without the sponge every stage's output is final where the boundary
  // reads it, so the stage kernels write the next stage's physical ghosts
  // themselves and only the first stage's input (and the source) needs the
  // boundary launch; the sponge changes the output after the stage, so
  // with it the boundary runs before every stage.
  if (!absorb_on) {
    ra.bc_fold = 1;
    for (int k = 0; k < 4; k++) ra.bc_fold |= (reflect[k] ? 1 : 0) << (1 + k);
  }
  hipError_t e = hipSuccess;
  long t;
  for (int step = 0; step < nsteps && e == hipSuccess; step++) {
    State in = A;
    for (int stage = 0; stage < order && e == hipSuccess; stage++) {
      // the last stage writes A (in is B or C then), the others alternate;
      // a one-stage scheme goes through B and is copied back
      State out = (stage == order - 1 && order > 1) ? A : ((stage % 2 == 0) ? B : C);
      if (!ra.bc_fold || (step == 0 && stage == 0)) {
        bc.w = in.w;
        bc.f[0] = in.p;
        bc.f[1] = in.T;
        t = g_timer.begin(ENL_KERNEL_BC, s);
        e = launch_physical_bc(bc, s);
        g_timer.end(t, s);
        bc.f[2] = nullptr;  // the source is constant: its ghosts are filled once
        if (e != hipSuccess) break;
      }
      ra.w = in.w;
      ra.p = in.p;
      ra.T = in.T;
      ra.w_out = out.w;
      ra.p_out = out.p;
      ra.T_out = out.T;
      ra.alpha_1 = alpha_1[stage];
      ra.alpha_2 = alpha_2[stage];
      ra.bdt = beta[stage] * dt;
      if (L.sw) {
        // the learned switch decides level 0's cells from the stage input
        // (every stage, or the step's first and its map kept)
        if (L.every_stage || stage == 0) {
          int rc = switch_decide(L.sw, G, in.w, in.p, gas, L.lo, L.hi, L.decide, s);
          if (!rc && L.totals) rc = switch_accumulate(L.sw, L.totals, s);
          if (rc) return rc;
        }
        ra.decide = L.decide;
      }
      t = g_timer.begin(ENL_KERNEL_STAGE, s);
      e = launch_stage(variant, ra, s);
      g_timer.end(t, s);
      if (absorb_on && e == hipSuccess) e = launch_absorb(plan, dtab, out.w, out.p, out.T, G, gas, s);
      in = out;
    }
    if (order == 1 && e == hipSuccess) {
      // copy B back into A (w_n, pres, temp are separate arrays)
      e = hipMemcpyAsync(A.w, B.w, sizeof(double) * 5 * G.sk, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(A.p, B.p, sizeof(double) * G.sk, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(A.T, B.T, sizeof(double) * G.sk, hipMemcpyDeviceToDevice, s);
    }
  }
  return e == hipSuccess ? 0 : check_hip(e, "fused patch loop launch");
}
}  // namespace
