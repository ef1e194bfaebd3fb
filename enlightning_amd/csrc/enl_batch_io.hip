// enl_batch_io.hip -- the SOLN patch data of a batched hierarchy in one
// buffer: the VisIt plot quantities (registerModelVariables,
// enlightning.C:295-378: "w_n #0".."w_n #4", "pressure", "temperature",
// "source", each an interior per patch) and a checkpoint's patch data (the
// same variables over the ghosted boxes, SAMRAI's restart of the SOLN
// context, enlightning.C:259-281).  One launch moves every selected plane of
// every patch between the pool and a buffer laid out patch after patch in
// pid order, plane after plane (w_n 0..4, pressure, temperature, source),
// row by row; a host buffer is reached through one copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "enl_batch.h"
#include "enlightning_hip.h"

namespace enl {
namespace {

constexpr int kPlanes = 8;  // w_n(5), pressure, temperature, source

// Offset of plane k in the pool (state block 0 holds w_n, p, T).
__device__ __host__ inline long plane_base(const PatchDev& p, int k) {
  return k < 7 ? p.state[0] + (long)k * p.sk : p.src;
}

// blockIdx.y = patch; its planes (bits of `planes`) at buf + off[patch].
template <bool GATHER>
__global__ void __launch_bounds__(256) move_kernel(double* pool, const PatchDev* __restrict__ dp, int planes,
                                                   int ghosted, const long* __restrict__ off, double* buf) {
  const PatchDev pd = dp[blockIdx.y];
  const long cells = ghosted ? pd.sk : (long)pd.nx * pd.ny;
  const int np = __popc(planes);
  double* b = buf + off[blockIdx.y];
  for (long q = blockIdx.x * 256L + threadIdx.x; q < np * cells; q += (long)gridDim.x * 256L) {
    const int n = (int)(q / cells);
    const long c = q - n * cells;
    // the n-th set bit of planes
    int k = 0, seen = -1;
    for (; k < kPlanes; k++)
      if ((planes >> k) & 1)
        if (++seen == n) break;
    const long l = ghosted ? c : (c % pd.nx + pd.g) + (c / pd.nx + pd.g) * pd.sy;
    double* v = pool + plane_base(pd, k) + l;
    if (GATHER)
      b[q] = *v;
    else
      *v = b[q];
  }
}

int move(Batch* b, int planes, int ghosted, double* data, long count, bool gather) {
  if (planes <= 0 || planes >= (1 << kPlanes) || (ghosted != 0 && ghosted != 1))
    return set_error(-2, "enl_batch_%s: bad planes 0x%x or ghosted %d", gather ? "gather" : "scatter", planes,
                     ghosted);
  if (b->npatch > 65535) return set_error(-2, "enl_batch_%s: more than 65535 patches", gather ? "gather" : "scatter");
  const int np = __builtin_popcount((unsigned)planes);
  std::vector<long> off((size_t)b->npatch);
  long total = 0, most = 0;
  for (int p = 0; p < b->npatch; p++) {
    const PatchDev& pd = b->hp[(size_t)p];
    const long cells = ghosted ? pd.sk : (long)pd.nx * pd.ny;
    off[(size_t)p] = total;
    total += np * cells;
    most = std::max(most, np * cells);
  }
  if (count != total)
    return set_error(-2, "enl_batch_%s: the buffer holds %ld values, the patches %ld", gather ? "gather" : "scatter",
                     count, total);
  hipStream_t s = lib_stream();
  const bool dev = is_device_ptr(data);
  const size_t obytes = sizeof(long) * off.size();
  // the offsets, then (for a host buffer) the staged values, in one buffer
  int rc = reserve(b, &b->stage, &b->stage_cap, obytes + (dev ? 0 : sizeof(double) * (size_t)total));
  if (rc) return rc;
  long* doff = (long*)b->stage;
  double* dbuf = dev ? data : (double*)((char*)b->stage + obytes);
  hipError_t e = hipMemcpyAsync(doff, off.data(), obytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !gather && !dev)
    e = hipMemcpyAsync(dbuf, data, sizeof(double) * (size_t)total, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return check_hip(e, gather ? "enl_batch_gather upload" : "enl_batch_scatter upload");
  const unsigned bx = (unsigned)std::min<long>((most + 255) / 256, 256);
  if (gather)
    move_kernel<true><<<dim3(bx, b->npatch), 256, 0, s>>>(b->pool, b->dp, planes, ghosted, doff, dbuf);
  else
    move_kernel<false><<<dim3(bx, b->npatch), 256, 0, s>>>(b->pool, b->dp, planes, ghosted, doff, dbuf);
  e = hipGetLastError();
  if (e == hipSuccess && gather && !dev)
    e = hipMemcpyAsync(data, dbuf, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost, s);
  // the offsets' (and a host buffer's) copies read host memory that the
  // caller may release on return
  const hipError_t e2 = hipStreamSynchronize(s);
  if (e != hipSuccess) return check_hip(e, gather ? "enl_batch_gather" : "enl_batch_scatter");
  return check_hip(e2, gather ? "enl_batch_gather sync" : "enl_batch_scatter sync");
}

}  // namespace
}  // namespace enl

using namespace enl;

extern "C" {

int enl_batch_gather(void* handle, int planes, int ghosted, double* out, long count) {
  Batch* b = (Batch*)handle;
  if (!b || !out) return set_error(-3, "enl_batch_gather: null argument");
  return move(b, planes, ghosted, out, count, true);
}

int enl_batch_scatter(void* handle, int planes, const double* in, long count) {
  Batch* b = (Batch*)handle;
  if (!b || !in) return set_error(-3, "enl_batch_scatter: null argument");
  return move(b, planes, 1, const_cast<double*>(in), count, false);
}

}  // extern "C"
