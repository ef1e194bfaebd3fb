// Source injection: enlightning::addSource (enlightning.C:2140-2307), the
// lightning-channel file reader (enlightning::readLightningInput,
// enlightning.C:2333-2389) and the microphones' pressure read
// (enlightning::recordAudio, enlightning.C:2392-2506) for the device.
//
// addSource writes the mass source of every interior cell of a patch:
//   PULSE POINT      amp * exp(-log 2 / s^2 * (x0^2 + x1^2))
//   PULSE SINE       amp * exp(-log 2 / s^2 * (x0 - A sin(f x1))^2)
//   PULSE LIGHTNING  per channel segment, the Gaussian of the distance to the
//                    segment (a cap around each end point on its side of the
//                    segment's normal through it, the perpendicular distance
//                    between them), merged into the cell's value v as
//                    v + t - (v / amp) * t, segment by segment
//   DRIVEN POINT     the POINT pulse times sin(omega * (it + 1 - t0) * dt)
// (x0, x1 the cell centre relative to src_pos, s = src_spread).  The
// per-segment quantities (angle, the slope of the normal lines, sin and cos
// of the angle, the orientation) are formed on the host with the reference's
// expressions; the device evaluates the per-cell expressions in the
// reference's order without contraction.
#include <cmath>
#include <cstdio>
#include <vector>

#include "enl_kernels.h"

namespace enl {

namespace {

constexpr int kPoint = 10, kSine = 20, kLightning = 30;  // enlightning.C:55-57
constexpr int kPulse = 10, kDriven = 20;                 // enlightning.C:60-61
constexpr double kPi = 3.14159265358979323846264338;     // enlightning.C:70

// One channel segment with its host-formed quantities.
struct Segment {
  double x0, y0, x1, y1;  // end points (p_lightning_coords[4k .. 4k+3])
  double slope;           // tan(PI/2 - theta)
  double sn, cs;          // sin(theta), cos(theta)
  int orientation;        // 1 or 2 (0: no cell qualifies)
};

struct SourceArgs {
  double* out;  // interior cell (0, 0)
  long sy;      // row stride of out
  int nx, ny;
  double xlo0, xlo1, dx0, dx1;
  int type;
  double amp, k;  // k = -log(2.0) / SQUARE(spread)
  double pos0, pos1, sine_amp, sine_freq;
  double drive;  // DRIVEN: the sin(...) factor; PULSE: unused
  bool driven;
  const Segment* seg;
  int nseg;
};

__global__ void __launch_bounds__(256) source_kernel(SourceArgs a) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)a.nx * a.ny) return;
  const int i = (int)(q % a.nx), j = (int)(q / a.nx);
  const double xc0 = a.xlo0 + a.dx0 * ((double)i + 0.5);
  const double xc1 = a.xlo1 + a.dx1 * ((double)j + 0.5);
  double* o = a.out + (long)i + (long)j * a.sy;
  if (a.type == kPoint) {
    const double x0 = xc0 - a.pos0, x1 = xc1 - a.pos1;
    double v = a.amp * exp(a.k * (x0 * x0 + x1 * x1));
    if (a.driven) v = v * a.drive;
    *o = v;
  } else if (a.type == kSine) {
    const double x0 = xc0 - a.pos0, x1 = xc1 - a.pos1;
    const double d = x0 - a.sine_amp * sin(a.sine_freq * x1);
    *o = a.amp * exp(a.k * (d * d));
  } else {
    double v = *o;
    for (int k = 0; k < a.nseg; k++) {
      const Segment s = a.seg[k];
      const double line0 = s.slope * (xc0 - s.x0) + s.y0;
      const double line1 = s.slope * (xc0 - s.x1) + s.y1;
      if ((s.orientation == 1 && xc1 >= line0) || (s.orientation == 2 && xc1 <= line0)) {
        const double dx = xc0 - s.x0, dy = xc1 - s.y0;
        const double t = a.amp * exp(a.k * (dx * dx + dy * dy));
        v = v + t - (v / a.amp) * t;
      }
      if ((s.orientation == 1 && xc1 <= line1) || (s.orientation == 2 && xc1 >= line1)) {
        const double dx = xc0 - s.x1, dy = xc1 - s.y1;
        const double t = a.amp * exp(a.k * (dx * dx + dy * dy));
        v = v + t - (v / a.amp) * t;
      }
      if ((s.orientation == 1 && xc1 < line0 && xc1 > line1) || (s.orientation == 2 && xc1 > line0 && xc1 < line1)) {
        const double d = s.sn * (xc0 - s.x0) + s.cs * (xc1 - s.y0);
        const double t = a.amp * exp(a.k * (d * d));
        v = v + t - (v / a.amp) * t;
      }
    }
    *o = v;
  }
}

// Segments of the channel with the quantities of enlightning.C:2215-2232.
std::vector<Segment> segments(const double* c, int n) {
  std::vector<Segment> v((size_t)n);
  for (int k = 0; k < n; k++) {
    const double x0 = c[4 * k], y0 = c[4 * k + 1], x1 = c[4 * k + 2], y1 = c[4 * k + 3];
    const double theta = -std::atan((y1 - y0) / (x1 - x0));
    int orientation = 0;
    if ((x0 > x1 && y0 > y1) || (x0 <= x1 && y0 >= y1))
      orientation = 1;
    else if ((x0 >= x1 && y0 <= y1) || (x0 < x1 && y0 < y1))
      orientation = 2;
    Segment& s = v[(size_t)k];
    s.x0 = x0;
    s.y0 = y0;
    s.x1 = x1;
    s.y1 = y1;
    s.slope = std::tan(kPi / 2.0 - theta);
    s.sn = std::sin(theta);
    s.cs = std::cos(theta);
    s.orientation = orientation;
  }
  return v;
}

// Pressure at each microphone inside the patch (enlightning::recordAudio,
// enlightning.C:2414-2436): cell i = (int)((x - xlo) / dx) of the interior
// array, the bounds inclusive as in the reference.
__global__ void __launch_bounds__(64) mic_kernel(const double* pres, long sy, int nx, int ny, double xlo0,
                                                 double xlo1, double xhi0, double xhi1, double dx0, double dx1,
                                                 int nmics, const double* mx, const double* my, double* out,
                                                 int* hit) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nmics) return;
  const double x = mx[k], y = my[k];
  if (x >= xlo0 && x <= xhi0 && y >= xlo1 && y <= xhi1) {
    const int i = (int)((x - xlo0) / dx0), j = (int)((y - xlo1) / dx1);
    // the reference reads cell (i, j) even on the upper bound (i == nx);
    // the device clamps it into the patch instead of reading past it
    const int ic = i < nx ? i : nx - 1, jc = j < ny ? j : ny - 1;
    out[k] = pres[(long)ic + (long)jc * sy];
    hit[k] = 1;
  } else {
    hit[k] = 0;
  }
}

}  // namespace

}  // namespace enl

using namespace enl;

extern "C" {

int enl_add_source(int ifirst0, int ilast0, int ifirst1, int ilast1, const double* patch_xlo, const double* dx,
                   const enl_source_t* src, int iteration, int pulse_time0, double dt, double* source, long sy) {
  if (ilast0 < ifirst0 || ilast1 < ifirst1)
    return set_error(-2, "invalid box [%d,%d]x[%d,%d]", ifirst0, ilast0, ifirst1, ilast1);
  if (!patch_xlo || !dx || !src) return set_error(-3, "null argument");
  if (!source || !is_device_ptr(source)) return set_error(-4, "enl_add_source needs a device source array");
  const int nx = ilast0 - ifirst0 + 1, ny = ilast1 - ifirst1 + 1;
  if (sy < nx) return set_error(-2, "source row stride below the patch width");
  SourceArgs a;
  a.out = source;
  a.sy = sy;
  a.nx = nx;
  a.ny = ny;
  a.xlo0 = patch_xlo[0];
  a.xlo1 = patch_xlo[1];
  a.dx0 = dx[0];
  a.dx1 = dx[1];
  a.type = src->type;
  a.amp = src->amp;
  a.k = -std::log(2.0) / (src->spread * src->spread);
  a.pos0 = src->pos[0];
  a.pos1 = src->pos[1];
  a.sine_amp = src->sine_amp;
  a.sine_freq = src->sine_freq;
  a.driven = src->mode == kDriven;
  a.drive = a.driven ? std::sin(src->omega * (double)(iteration + 1 - pulse_time0) * dt) : 0.0;
  a.seg = nullptr;
  a.nseg = 0;
  if (src->mode != kPulse && src->mode != kDriven) return set_error(-2, "unknown source mode %d", src->mode);
  if (src->type != kPoint && src->type != kSine && src->type != kLightning)
    return set_error(-2, "unknown source type %d", src->type);
  // enlightning.C:2301-2302
  if (a.driven && src->type != kPoint) return set_error(-2, "DRIVEN SINE or LIGHTNING source mode not implemented");
  hipStream_t s = lib_stream();
  Segment* dseg = nullptr;
  if (src->type == kLightning) {
    if (src->nseg < 0 || (src->nseg > 0 && !src->coords)) return set_error(-3, "lightning segments missing");
    if (src->nseg == 0) return 0;
    const std::vector<Segment> h = segments(src->coords, src->nseg);
    hipError_t e = hipMallocAsync((void**)&dseg, sizeof(Segment) * h.size(), s);
    if (e == hipSuccess) e = hipMemcpyAsync(dseg, h.data(), sizeof(Segment) * h.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // h goes out of scope
    if (e != hipSuccess) return check_hip(e, "lightning segments upload");
    a.seg = dseg;
    a.nseg = src->nseg;
  }
  const long cells = (long)nx * ny;
  hipLaunchKernelGGL(source_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (dseg) {
    hipError_t f = hipFreeAsync(dseg, s);
    if (e == hipSuccess) e = f;
  }
  return check_hip(e, "source launch");
}

int enl_sample_mics(int ifirst0, int ilast0, int ifirst1, int ilast1, const double* patch_xlo, const double* patch_xhi,
                    const double* dx, int nmics, const double* mic_x, const double* mic_y, const double* pres, long sy,
                    double* out, int* hit) {
  if (ilast0 < ifirst0 || ilast1 < ifirst1)
    return set_error(-2, "invalid box [%d,%d]x[%d,%d]", ifirst0, ilast0, ifirst1, ilast1);
  if (nmics <= 0) return 0;
  if (!patch_xlo || !patch_xhi || !dx) return set_error(-3, "null argument");
  const void* dev[5] = {mic_x, mic_y, pres, out, hit};
  for (int i = 0; i < 5; i++)
    if (!dev[i] || !is_device_ptr(dev[i])) return set_error(-4, "enl_sample_mics needs device arrays");
  const int nx = ilast0 - ifirst0 + 1, ny = ilast1 - ifirst1 + 1;
  if (sy < nx) return set_error(-2, "pressure row stride below the patch width");
  hipLaunchKernelGGL(mic_kernel, dim3((unsigned)((nmics + 63) / 64)), dim3(64), 0, lib_stream(), pres, sy, nx, ny,
                     patch_xlo[0], patch_xlo[1], patch_xhi[0], patch_xhi[1], dx[0], dx[1], nmics, mic_x, mic_y, out,
                     hit);
  return check_hip(hipGetLastError(), "mic gather launch");
}

// enlightning::readLightningInput: one segment per line read with
// fgets(line, 80) and sscanf("%f %f %f %f"); a line that does not parse
// repeats the previous segment's values (zeros before the first).
int enl_read_lightning(const char* path, double* coords, int max_segments) {
  if (!path) return set_error(-3, "null path");
  FILE* fr = fopen(path, "rt");
  if (!fr) return set_error(-6, "can't open '%s' for reading", path);
  char line[80];
  float x1 = 0, y1 = 0, x2 = 0, y2 = 0;
  int n = 0;
  while (fgets(line, 80, fr) != NULL) {
    (void)sscanf(line, "%f %f %f %f", &x1, &y1, &x2, &y2);
    if (coords && n < max_segments) {
      coords[4 * n] = x1;
      coords[4 * n + 1] = y1;
      coords[4 * n + 2] = x2;
      coords[4 * n + 3] = y2;
    }
    n++;
  }
  fclose(fr);
  if (n < 1) return set_error(-6, "file '%s' is empty", path);
  return n;
}

}  // extern "C"
