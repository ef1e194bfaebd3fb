// Absorbing envelopes: enlightning::setAbsorbingBoundaryConditions
// (enlightning.C:630-786) on the device.
//
// Each side s with bc_type ABSORB damps the perturbation of the interior
// cells near it: v <- (v - v0) * tmp + v0 for pres, temp, rho and rho*T_n
// (v0 = p0, T0, rho0, T0*rho0), v <- v * tmp for the momenta and rho*s, with
// tmp = -exp(-log(2.0) / W^2 * d^2) + 1 (d: distance of the cell centre from
// the domain edge, W: bc_absorb_width), sides in the order left, right,
// bottom, top.  The reference walks each side from the patch edge inward
// and stops after the first cell (row) whose |1 - tmp| <= 1e-8; along the
// left and right sides every row after the first starts with tmp reset to
// 0, so its first cell is always processed.  tmp depends only on the index
// normal to the side, so the processed cells of a side form a rectangle
// (less the first row on the left / right sides when the first cell already
// passes the test): the host finds the rectangles and the tmp table with
// the reference's own expressions (same libm as the reference's host code),
// and one kernel launch per side applies them.
#include <cmath>
#include <mutex>
#include <vector>

#include "enl_kernels.h"

namespace enl {

namespace {

constexpr double kAbsorbEp = 1.0e-8;  // enlightning.C:667

// tmp of the reference at distance d from the domain edge (enlightning.C:676)
double envelope(double d, double width) { return -std::exp(-std::log(2.0) / (width * width) * (d * d)) + 1; }

struct AbsorbArgs {
  double* w;
  double* p;
  double* T;
  Geom geom;
  double rho0, p0, T0, T0rho0;
  const double* tab;  // tmp by normal index
  int side;           // 0 left, 1 right, 2 bottom, 3 top
  int lo, hi;         // normal index range
  int t0;             // first tangential index processed
  int nt;             // tangential count
};

__global__ void __launch_bounds__(256) absorb_kernel(AbsorbArgs a) {
  const int nn = a.hi - a.lo + 1;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nn * a.nt) return;
  // left / right: normal = i (fastest, contiguous), tangential = j
  int i, j;
  const int u = (int)(q % nn) + a.lo, v = (int)(q / nn) + a.t0;
  if (a.side < 2) {
    i = u;
    j = v;
  } else {
    // bottom / top: tangential = i fastest
    const int uu = (int)(q % a.nt) + a.t0, vv = (int)(q / a.nt) + a.lo;
    i = uu;
    j = vv;
  }
  const double tmp = a.tab[a.side < 2 ? i : j];
  const Geom& G = a.geom;
  const long g = (long)(i + G.g0) + (long)(j + G.g1) * G.sy;
  a.p[g] = (a.p[g] - a.p0) * tmp + a.p0;
  a.T[g] = (a.T[g] - a.T0) * tmp + a.T0;
  a.w[g] = (a.w[g] - a.rho0) * tmp + a.rho0;
  a.w[g + G.sk] = a.w[g + G.sk] * tmp;
  a.w[g + 2 * G.sk] = a.w[g + 2 * G.sk] * tmp;
  a.w[g + 3 * G.sk] = a.w[g + 3 * G.sk] * tmp;
  a.w[g + 4 * G.sk] = (a.w[g + 4 * G.sk] - a.T0rho0) * tmp + a.T0rho0;
}

// Per device: the last uploaded tables and the key they were built for.
struct AbsorbCache {
  double* dev = nullptr;
  size_t cap = 0;
  std::vector<double> host;  // last uploaded table
};
std::mutex g_ab_mu;
AbsorbCache g_ab[64];

}  // namespace

int absorb_plan(const Geom& G, const double* patch_xlo, const double* dx, const double* grid_xlo,
                const double* grid_xhi, const int* absorb, double width, AbsorbPlan* out) {
  const int nx = G.nx, ny = G.ny;
  AbsorbSide* sp = out->side;
  std::vector<double>& tab = out->tab;
  tab.clear();
  for (int s = 0; s < 4; s++) sp[s] = AbsorbSide();
  auto xc0 = [&](int i) { return patch_xlo[0] + dx[0] * ((double)i + 0.5); };
  auto xc1 = [&](int j) { return patch_xlo[1] + dx[1] * ((double)j + 0.5); };
  for (int s = 0; s < 4; s++) {
    if (!absorb[s]) continue;
    const int n = s < 2 ? nx : ny;
    AbsorbSide& P = sp[s];
    P.off = (int)tab.size();
    tab.resize(tab.size() + n, 0.0);
    double* t = tab.data() + P.off;
    // walk inward from the side's edge: the first index whose |1 - tmp|
    // passes the test is the last one processed
    const bool up = (s == 0 || s == 2);
    int last = up ? n - 1 : 0;
    for (int q = 0; q < n; q++) {
      const int k = up ? q : n - 1 - q;
      double d;
      if (s == 0) d = xc0(k) - grid_xlo[0];
      else if (s == 1) d = grid_xhi[0] - xc0(k);
      else if (s == 2) d = xc1(k) - grid_xlo[1];
      else d = grid_xhi[1] - xc1(k);
      t[k] = envelope(d, width);
      if (!(std::fabs(1.0 - t[k]) > kAbsorbEp)) {
        last = k;
        break;
      }
    }
    P.lo = up ? 0 : last;
    P.hi = up ? last : n - 1;
    // the pre-loop test on the edge cell (enlightning.C:675-676, 704-705,
    // 734-735, 763-764): fails -> the bottom / top side does nothing, the
    // left / right side skips its first row
    const int edge = up ? 0 : n - 1;
    const bool edge_passes = !(std::fabs(1.0 - t[edge]) > kAbsorbEp);
    if (s >= 2 && edge_passes) P.hi = P.lo - 1;
    if (s < 2) P.skip_first = edge_passes;
  }
  return 0;
}

int absorb_upload(const AbsorbPlan& plan, const double** dtab, hipStream_t s) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(-5, "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_ab_mu);
  AbsorbCache& c = g_ab[dev];
  const size_t n = plan.tab.size();
  if (n == 0) {
    *dtab = c.dev;
    return 0;
  }
  if (c.cap < n) {
    if (c.dev) (void)hipFree(c.dev);
    c.dev = nullptr;
    c.cap = 0;
    c.host.clear();
    hipError_t e = hipMalloc((void**)&c.dev, n * sizeof(double));
    if (e != hipSuccess) return check_hip(e, "hipMalloc(absorb table)");
    c.cap = n;
  }
  if (c.host != plan.tab) {
    // the previous table may still be read by queued launches
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(c.dev, plan.tab.data(), n * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return check_hip(e, "absorb table upload");
    c.host = plan.tab;
  }
  *dtab = c.dev;
  return 0;
}

hipError_t launch_absorb(const AbsorbPlan& plan, const double* dtab, double* w, double* p, double* T, const Geom& G,
                         const enl_gas_t* gas, hipStream_t s) {
  for (int side = 0; side < 4; side++) {
    const AbsorbSide& P = plan.side[side];
    if (P.hi < P.lo) continue;
    AbsorbArgs a;
    a.w = w;
    a.p = p;
    a.T = T;
    a.geom = G;
    a.rho0 = gas->rho0;
    a.p0 = gas->p0;
    a.T0 = gas->T0;
    a.T0rho0 = gas->T0 * gas->rho0;
    a.tab = dtab + P.off;
    a.side = side;
    a.lo = P.lo;
    a.hi = P.hi;
    const int ntan = side < 2 ? G.ny : G.nx;
    a.t0 = (side < 2 && P.skip_first) ? 1 : 0;
    a.nt = ntan - a.t0;
    const long cells = (long)(P.hi - P.lo + 1) * a.nt;
    if (cells <= 0) continue;
    hipLaunchKernelGGL(absorb_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace enl

using namespace enl;

extern "C" {

int enl_absorb(int ifirst0, int ilast0, int ifirst1, int ilast1, int gcw0, int gcw1, const double* patch_xlo,
               const double* dx, const double* grid_xlo, const double* grid_xhi, const int* absorb, double width,
               const enl_gas_t* gas, double* w_n, double* pres, double* temp) {
  if (ilast0 < ifirst0 || ilast1 < ifirst1)
    return set_error(-2, "invalid box [%d,%d]x[%d,%d]", ifirst0, ilast0, ifirst1, ilast1);
  if (!patch_xlo || !dx || !grid_xlo || !grid_xhi || !absorb || !gas) return set_error(-3, "null argument");
  const double* ptrs[3] = {w_n, pres, temp};
  for (int i = 0; i < 3; i++)
    if (!ptrs[i] || !is_device_ptr(ptrs[i])) return set_error(-4, "enl_absorb needs device arrays");
  if (!(absorb[0] || absorb[1] || absorb[2] || absorb[3])) return 0;
  const Geom G = make_geom(ifirst0, ilast0, ifirst1, ilast1, gcw0, gcw1);
  AbsorbPlan plan;
  int rc = absorb_plan(G, patch_xlo, dx, grid_xlo, grid_xhi, absorb, width, &plan);
  if (rc) return rc;
  const double* dtab = nullptr;
  hipStream_t s = lib_stream();
  if ((rc = absorb_upload(plan, &dtab, s))) return rc;
  return check_hip(launch_absorb(plan, dtab, w_n, pres, temp, G, gas, s), "absorb launch");
}

}  // extern "C"
