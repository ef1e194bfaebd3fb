// enl_abi.hip -- the exported C ABI (include/enlightning_hip.h): the
// Fortran-mangled drop-in kernels, pointer-kind dispatch with host staging,
// the error side channel and the library stream.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "enl_device.h"
#include "enl_kernels.h"
#include "enlightning_hip.h"

#pragma clang fp contract(off)

// The WENO list of enl_rhs_variant_noalloc, which allocates nothing:
// kNoAllocMaxCells list entries, then the list counter in the next slot.
// Every GPU holds its own copy until the process ends.  In host code the
// name `scratch` is a host-side stand-in, not a device address: host code
// uses only the address hipGetSymbolAddress returns for the current device.
static constexpr long kNoAllocMaxCells = 1L << 22;  // 32 MiB: a 2048 x 2048 patch
static __device__ __align__(256) unsigned long long scratch[kNoAllocMaxCells + 32];

namespace enl {

namespace {
hipStream_t g_stream = nullptr;
thread_local int t_err_code = 0;
thread_local char t_err_msg[512] = {0};
}  // namespace

hipStream_t lib_stream() { return g_stream; }

int set_error(int code, const char* fmt, ...) {
  t_err_code = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err_msg, sizeof t_err_msg, fmt, ap);
  va_end(ap);
  static const bool abort_on_error = [] {
    const char* e = getenv("ENL_ABORT_ON_ERROR");
    return e && e[0] == '1';
  }();
  if (abort_on_error) {
    fprintf(stderr, "enlightning_hip: %s\n", t_err_msg);
    abort();
  }
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return set_error(-100 - (int)e, "%s: %s", what, hipGetErrorString(e));
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
         attr.type == hipMemoryTypeUnified;
}

namespace {
// Per-device WENO-list workspace of the rhs kernels.
struct RhsWorkspace {
  unsigned long long* list = nullptr;
  unsigned* count = nullptr;
  long cap = 0;
};
std::mutex g_ws_mu;
RhsWorkspace g_ws[64];
}  // namespace

int rhs_workspace(long cells, unsigned long long** list, unsigned** count) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(-5, "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_ws_mu);
  RhsWorkspace& w = g_ws[dev];
  if (!w.count) {
    hipError_t e = hipMalloc((void**)&w.count, sizeof(unsigned));
    if (e != hipSuccess) return check_hip(e, "hipMalloc(rhs counter)");
  }
  if (w.cap < cells) {
    if (w.list) (void)hipFree(w.list);
    w.list = nullptr;
    w.cap = 0;
    hipError_t e = hipMalloc((void**)&w.list, sizeof(unsigned long long) * (size_t)cells);
    if (e != hipSuccess) return check_hip(e, "hipMalloc(rhs list)");
    w.cap = cells;
  }
  *list = w.list;
  *count = w.count;
  return 0;
}

namespace {
// Per-device address of `scratch`, looked up on the first call on a device.
std::mutex g_na_mu;
void* g_na_ptr[64] = {nullptr};

int noalloc_workspace(long cells, unsigned long long** list, unsigned** count) {
  if (cells > kNoAllocMaxCells)
    return set_error(-2, "enl_rhs_variant_noalloc: %ld cells exceed its fixed workspace of %ld cells", cells,
                     kNoAllocMaxCells);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(-5, "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return check_hip(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_na_ptr[dev] = p;
  }
  *list = (unsigned long long*)g_na_ptr[dev];
  *count = (unsigned*)(*list + kNoAllocMaxCells);
  return 0;
}

// Device mirrors for host-pointer calls.  Grow-only buffers, one per argument
// slot, guarded by one mutex (host-pointer calls are serialized; device-pointer
// calls take no lock).
struct Staging {
  std::mutex mu;
  void* buf[8] = {nullptr};
  size_t cap[8] = {0};
  int get(int slot, size_t bytes, void** out) {
    if (cap[slot] < bytes) {
      if (buf[slot]) (void)hipFree(buf[slot]);
      buf[slot] = nullptr;
      cap[slot] = 0;
      hipError_t e = hipMalloc(&buf[slot], bytes);
      if (e != hipSuccess) return check_hip(e, "hipMalloc(staging)");
      cap[slot] = bytes;
    }
    *out = buf[slot];
    return 0;
  }
};
Staging& staging() {
  static Staging* s = new Staging();  // never destroyed: outlives static teardown
  return *s;
}

// One array argument of a drop-in call: device pointers pass through, host
// pointers are mirrored (copied in if `in`, copied back if `out`).
struct Arg {
  void* host = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  bool staged = false, out = false;
};

int bind(Arg& a, int slot, const void* p, size_t bytes, bool in, bool out) {
  a.host = const_cast<void*>(p);
  a.bytes = bytes;
  a.out = out;
  if (bytes == 0) {
    a.dev = a.host;
    return 0;
  }
  if (!p) return set_error(-3, "null array argument (slot %d)", slot);
  if (is_device_ptr(p)) {
    a.dev = a.host;
    return 0;
  }
  a.staged = true;
  int rc = staging().get(slot, bytes, &a.dev);
  if (rc) return rc;
  if (in) return check_hip(hipMemcpyAsync(a.dev, p, bytes, hipMemcpyHostToDevice, g_stream), "stage in");
  return 0;
}

int finish(Arg* args, int n) {
  bool any = false;
  for (int i = 0; i < n; i++) {
    if (args[i].staged && args[i].out) {
      int rc = check_hip(hipMemcpyAsync(args[i].host, args[i].dev, args[i].bytes, hipMemcpyDeviceToHost, g_stream),
                         "stage out");
      if (rc) return rc;
    }
    any |= args[i].staged;
  }
  if (any) return check_hip(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
  return 0;
}

bool any_host(const void* const* ps, int n) {
  for (int i = 0; i < n; i++)
    if (ps[i] && !is_device_ptr(ps[i])) return true;
  return false;
}

int check_box(int if0, int il0, int if1, int il1, int g0, int g1, int need, int nequ) {
  if (nequ != 5) return set_error(-1, "nequ = %d; the kernels are built for NEQU = 5 (enlightning.h:48)", nequ);
  if (il0 < if0 - 1 || il1 < if1 - 1) return set_error(-2, "invalid box [%d,%d]x[%d,%d]", if0, il0, if1, il1);
  if (g0 < need || g1 < need)
    return set_error(-2, "ghost width (%d,%d) below the stencil reach %d", g0, g1, need);
  return 0;
}

bool il_invalid(int first, int last) { return last < first - 1; }

int ghost_need(int variant) {
  switch (variant) {
    case kWenoLowLlf:
    case kWenoLow: return 2;
    default: return 3;
  }
}

int do_rhs(int variant, const int* ifirst0, const int* ilast0, const int* ifirst1, const int* ilast1,
           const int* gcw0, const int* gcw1, const double* xlo, const double* dx, const double* w_n,
           const double* source, double* rhs, const double* pres, const double* temp, const double* weno_alpha,
           const double* kappa, const double* mu, const double* muB, const double* R, const double* T_star_n,
           const double* T_star_o, const double* tau_n, const double* tau_o, const double* tag_tolerance,
           const int* level, const double* c, const double* c_lr, const int* nequ, int* switch_map,
           bool noalloc = false) {
  (void)T_star_o;
  (void)tau_o;
  if (variant < 0 || variant >= kNumVariants) return set_error(-1, "unknown rhs variant %d", variant);
  int rc = check_box(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1, ghost_need(variant), *nequ);
  if (rc) return rc;
  Geom G = make_geom(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1);
  if (noalloc && (long)G.nx * G.ny > kNoAllocMaxCells)
    return set_error(-2, "enl_rhs_variant_noalloc: %ld cells exceed its fixed workspace of %ld cells",
                     (long)G.nx * G.ny, kNoAllocMaxCells);
  if (G.nx == 0 || G.ny == 0) return 0;
  size_t ng = (size_t)G.sk * sizeof(double), ni = (size_t)G.nx * G.ny * sizeof(double);
  const void* ptrs[6] = {w_n, source, rhs, pres, temp, switch_map};
  if (noalloc && any_host(ptrs, 6))
    return set_error(-3, "enl_rhs_variant_noalloc: every array must be device memory (host arrays need staging)");
  std::unique_lock<std::mutex> lock(staging().mu, std::defer_lock);
  if (any_host(ptrs, 6)) lock.lock();
  Arg args[6];
  if ((rc = bind(args[0], 0, w_n, 5 * ng, true, false))) return rc;
  if ((rc = bind(args[1], 1, source, ni, true, false))) return rc;
  if ((rc = bind(args[2], 2, rhs, 5 * ni, false, true))) return rc;
  if ((rc = bind(args[3], 3, pres, ng, true, false))) return rc;
  if ((rc = bind(args[4], 4, temp, ng, true, false))) return rc;
  if (switch_map && (rc = bind(args[5], 5, switch_map, (size_t)G.nx * G.ny * sizeof(int), false, true))) return rc;
  RhsArgs a = {};
  a.w = (const double*)args[0].dev;
  a.H0 = (const double*)args[1].dev;
  a.h0_sy = G.nx;
  a.KK = (double*)args[2].dev;
  a.p = (const double*)args[3].dev;
  a.T = (const double*)args[4].dev;
  a.switch_map = switch_map ? (int*)args[5].dev : nullptr;
  a.geom = G;
  a.dx = dx[0];
  a.dy = dx[1];
  a.ylo = xlo[1];
  a.alpha = *weno_alpha;
  a.kappa = *kappa;
  a.mu = *mu;
  a.muB = *muB;
  a.R = *R;
  a.T_star_n = *T_star_n;
  a.tau_n = *tau_n;
  a.c = *c;
  a.c_lr = *c_lr;
  a.ep = sensor_eps(tag_tolerance[2]);
  a.level = *level;
  a.list = nullptr;
  a.count = nullptr;
  if (rhs_needs_list(variant, a.level) &&
      (rc = noalloc ? noalloc_workspace((long)G.nx * G.ny, &a.list, &a.count)
                    : rhs_workspace((long)G.nx * G.ny, &a.list, &a.count)))
    return rc;
  if ((rc = check_hip(launch_rhs(variant, a, g_stream), "rhs launch"))) return rc;
  return finish(args, switch_map ? 6 : 5);
}

int default_variant() {
  static const int v = [] {
    const char* e = getenv("ENL_RHS_VARIANT");
    if (!e || !*e) return (int)kHybridHighLlf;
    static const char* names[kNumVariants] = {"hybrid_high_llf", "hybrid_high", "hybrid_low_llf",
                                              "hybrid_low",      "weno_high_llf", "weno_high",
                                              "weno_low_llf",    "weno_low",      "drp"};
    for (int i = 0; i < kNumVariants; i++)
      if (!strcmp(e, names[i])) return i;
    char* end = nullptr;
    long n = strtol(e, &end, 10);
    if (end && *end == 0 && n >= 0 && n < kNumVariants) return (int)n;
    fprintf(stderr, "enlightning_hip: ENL_RHS_VARIANT=%s not recognised; using hybrid_high_llf\n", e);
    return (int)kHybridHighLlf;
  }();
  return v;
}

}  // namespace
}  // namespace enl

using namespace enl;

extern "C" {

#define ENL_RHS_FWD                                                                                        \
  ifirst0, ilast0, ifirst1, ilast1, gcw0, gcw1, xlo, dx, w_n, source, rhs, pres, temp, weno_alpha, kappa, mu, \
      muB, R, T_star_n, T_star_o, tau_n, tau_o, tag_tolerance, level, c, c_lr, nequ

int enl_rhs_variant(int variant, ENL_RHS_ARGS, int* switch_map) {
  return do_rhs(variant, ENL_RHS_FWD, switch_map);
}

int enl_rhs_variant_noalloc(int variant, ENL_RHS_ARGS, int* switch_map) {
  return do_rhs(variant, ENL_RHS_FWD, switch_map, true);
}

void rhs_(ENL_RHS_ARGS) { (void)do_rhs(default_variant(), ENL_RHS_FWD, nullptr); }
void rhs_hybrid_high_llf_(ENL_RHS_ARGS) { (void)do_rhs(kHybridHighLlf, ENL_RHS_FWD, nullptr); }
void rhs_hybrid_high_(ENL_RHS_ARGS) { (void)do_rhs(kHybridHigh, ENL_RHS_FWD, nullptr); }
void rhs_hybrid_low_llf_(ENL_RHS_ARGS) { (void)do_rhs(kHybridLowLlf, ENL_RHS_FWD, nullptr); }
void rhs_hybrid_low_(ENL_RHS_ARGS) { (void)do_rhs(kHybridLow, ENL_RHS_FWD, nullptr); }
void rhs_weno_high_llf_(ENL_RHS_ARGS) { (void)do_rhs(kWenoHighLlf, ENL_RHS_FWD, nullptr); }
void rhs_weno_high_(ENL_RHS_ARGS) { (void)do_rhs(kWenoHigh, ENL_RHS_FWD, nullptr); }
void rhs_weno_low_llf_(ENL_RHS_ARGS) { (void)do_rhs(kWenoLowLlf, ENL_RHS_FWD, nullptr); }
void rhs_weno_low_(ENL_RHS_ARGS) { (void)do_rhs(kWenoLow, ENL_RHS_FWD, nullptr); }
void rhs_drp_(ENL_RHS_ARGS) { (void)do_rhs(kDrp, ENL_RHS_FWD, nullptr); }

void runge_kutta_(const int* ifirst0, const int* ilast0, const int* ifirst1, const int* ilast1, const int* gcw0,
                  const int* gcw1, const double* dt, const double* alpha_1, const double* alpha_2, const double* beta,
                  double* w_n, const double* w, const double* rhs, const int* nequ) {
  if (check_box(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1, 0, *nequ)) return;
  Geom G = make_geom(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1);
  if (G.nx == 0 || G.ny == 0) return;
  size_t ng = (size_t)G.sk * sizeof(double), ni = (size_t)G.nx * G.ny * sizeof(double);
  const void* ptrs[3] = {w_n, w, rhs};
  std::unique_lock<std::mutex> lock(staging().mu, std::defer_lock);
  if (any_host(ptrs, 3)) lock.lock();
  Arg args[3];
  if (bind(args[0], 0, w_n, 5 * ng, true, true) || bind(args[1], 1, w, 5 * ni, true, false) ||
      bind(args[2], 2, rhs, 5 * ni, true, false))
    return;
  RkArgs a;
  a.w = (double*)args[0].dev;
  a.ww = (const double*)args[1].dev;
  a.KK = (const double*)args[2].dev;
  a.geom = G;
  a.alpha_1 = *alpha_1;
  a.alpha_2 = *alpha_2;
  a.beta = *beta;
  a.dt = *dt;
  if (check_hip(launch_runge_kutta(a, g_stream), "runge_kutta launch")) return;
  (void)finish(args, 3);
}

void update_state_(const double* dx, const double* xlo, const int* ifirst0, const int* ilast0, const int* ifirst1,
                   const int* ilast1, const int* gcw0, const int* gcw1, const double* w_n, double* pres, double* temp,
                   const double* rho0, const double* c, const double* gamma1, const double* c_p, const double* p0,
                   const double* T0, const double* c_lr, const double* c_v, const double* R_tilde, const double* M,
                   const double* R, const int* nequ) {
  (void)R_tilde;
  (void)M;
  if (check_box(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1, 0, *nequ)) return;
  Geom G = make_geom(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1);
  if (G.nx == 0 || G.ny == 0) return;
  size_t ng = (size_t)G.sk * sizeof(double);
  const void* ptrs[3] = {w_n, pres, temp};
  std::unique_lock<std::mutex> lock(staging().mu, std::defer_lock);
  if (any_host(ptrs, 3)) lock.lock();
  Arg args[3];
  // pres/temp are copied in as well: only their interior is written, the
  // ghosts must come back unchanged.
  if (bind(args[0], 0, w_n, 5 * ng, true, false) || bind(args[1], 1, pres, ng, true, true) ||
      bind(args[2], 2, temp, ng, true, true))
    return;
  StateArgs a;
  a.w = (const double*)args[0].dev;
  a.p = (double*)args[1].dev;
  a.T = (double*)args[2].dev;
  a.geom = G;
  a.dy = dx[1];
  a.ylo = xlo[1];
  a.rho0 = *rho0;
  a.c = *c;
  a.gamma1 = *gamma1;
  a.c_p = *c_p;
  a.p0 = *p0;
  a.T0 = *T0;
  a.c_lr = *c_lr;
  a.c_v = *c_v;
  a.R = *R;
  if (check_hip(launch_update_state(a, g_stream), "update_state launch")) return;
  (void)finish(args, 3);
}

void tag_cells_(const int* ifirst0, const int* ilast0, const int* ifirst1, const int* ilast1, const int* gcw0,
                const int* gcw1, int* tags, const double* w_n, const double* source, const int* refine_tag_val,
                const double* tag_tolerance, const int* nequ) {
  if (check_box(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1, 1, *nequ)) return;
  Geom G = make_geom(*ifirst0, *ilast0, *ifirst1, *ilast1, *gcw0, *gcw1);
  if (G.nx == 0 || G.ny == 0) return;
  size_t ng = (size_t)G.sk * sizeof(double);
  const void* ptrs[3] = {tags, w_n, source};
  std::unique_lock<std::mutex> lock(staging().mu, std::defer_lock);
  if (any_host(ptrs, 3)) lock.lock();
  Arg args[3];
  if (bind(args[0], 0, tags, (size_t)G.nx * G.ny * sizeof(int), false, true) ||
      bind(args[1], 1, w_n, 5 * ng, true, false) || bind(args[2], 2, source, ng, true, false))
    return;
  TagArgs a;
  a.tags = (int*)args[0].dev;
  a.w = (const double*)args[1].dev;
  a.H0 = (const double*)args[2].dev;
  a.geom = G;
  a.refine_tag_val = *refine_tag_val;
  a.ep0 = sensor_eps(tag_tolerance[0]);
  a.ep1 = sensor_eps(tag_tolerance[1]);
  if (check_hip(launch_tag_cells(a, g_stream), "tag_cells launch")) return;
  (void)finish(args, 3);
}

int enl_set_stream(void* hip_stream) {
  g_stream = (hipStream_t)hip_stream;
  return 0;
}

int enl_sync(void) { return check_hip(hipStreamSynchronize(g_stream), "hipStreamSynchronize"); }

int enl_last_error(char* buf, size_t n) {
  int code = t_err_code;
  if (buf && n) {
    strncpy(buf, t_err_msg, n - 1);
    buf[n - 1] = 0;
  }
  t_err_code = 0;
  t_err_msg[0] = 0;
  return code;
}

const char* enl_version(void) { return "enlightning_hip 0.1 gfx950"; }

int enl_fill_physical_bc(int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw0, int gcw1, const int* touch,
                         const int* reflect, double* w_n, double* pres, double* temp, double* source) {
  if (il_invalid(ifirst0, ilast0) || il_invalid(ifirst1, ilast1))
    return set_error(-2, "invalid box [%d,%d]x[%d,%d]", ifirst0, ilast0, ifirst1, ilast1);
  double* ptrs[4] = {w_n, pres, temp, source};
  for (int i = 0; i < 4; i++)
    if (ptrs[i] && !is_device_ptr(ptrs[i])) return set_error(-4, "enl_fill_physical_bc needs device arrays");
  if (!w_n) return set_error(-3, "w_n is null");
  BcArgs a;
  a.w = w_n;
  a.f[0] = pres;
  a.f[1] = temp;
  a.f[2] = source;
  a.geom = make_geom(ifirst0, ilast0, ifirst1, ilast1, gcw0, gcw1);
  for (int s = 0; s < 4; s++) {
    a.touch[s] = touch[s];
    a.reflect[s] = reflect[s];
  }
  return check_hip(launch_physical_bc(a, g_stream), "physical bc launch");
}

}  // extern "C"
