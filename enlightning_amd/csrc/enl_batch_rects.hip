// enl_batch_rects.hip -- a batched hierarchy's schedules from rectangles.
//
// A regrid (regridAllFinerLevels, main.C:489-502) gives the hierarchy new
// fine boxes, and with them new ghost-fill, coarsen and transfer schedules.
// Listed cell by cell they run to about a million items at the AMR
// workload (1024^2 base, ~560 fine patches); as rectangles of cells whose
// addresses are affine in the cell's offset (enl_rect_t) they are a few
// thousand.  The host plans and uploads the rectangles; the device expands
// a level's fill rectangles and the coarsen rectangles into the per-cell
// lists the stage fills read (fill_edges_kernel / fill32_edges_kernel,
// coarsen_kernel / coarsen32_kernel: the same items, in rectangle order),
// and moves the regrid transfer straight from its rectangles.  The level-0
// tags of a regrid go into a map of the level's index space in one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "enl_batch.h"
#include "enl_device.h"
#include "enlightning_hip.h"

#pragma clang fp contract(off)

namespace enl {
namespace {

// ---- device memory cache ------------------------------------------------------
struct MemCache {
  std::mutex mu;
  std::map<int, std::multimap<size_t, void*>> idle;         // device -> block size -> block
  std::unordered_map<void*, std::pair<int, size_t>> live;  // block -> (device, block size)
};

MemCache& cache() {
  static MemCache* c = new MemCache();  // never destroyed: outlives static teardown
  return *c;
}

// Small requests in 256-byte granules; large ones with 1/16 headroom in 2 MiB
// granules, so that the next regrid's slightly larger pool still fits.
size_t block_size(size_t bytes) {
  if (bytes <= ((size_t)1 << 20)) return std::max<size_t>(256, (bytes + 255) & ~(size_t)255);
  const size_t g = (size_t)2 << 20, s = bytes + bytes / 16;
  return (s + g - 1) / g * g;
}

}  // namespace

hipError_t batch_malloc(void** p, size_t bytes) {
  *p = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  MemCache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  std::multimap<size_t, void*>& idle = c.idle[dev];
  // the smallest idle block that holds the request, unless a quarter (or
  // 1 MiB) larger than it
  auto it = idle.lower_bound(bytes);
  if (it != idle.end() && it->first <= std::max(bytes + bytes / 4, bytes + ((size_t)1 << 20))) {
    *p = it->second;
    c.live[*p] = {dev, it->first};
    idle.erase(it);
    return hipSuccess;
  }
  const size_t bs = block_size(bytes);
  e = hipMalloc(p, bs);
  if (e != hipSuccess) {
    // out of memory: hand the idle blocks of this device back and retry
    (void)hipGetLastError();
    for (auto& kv : idle) (void)hipFree(kv.second);
    idle.clear();
    e = hipMalloc(p, bs);
    if (e != hipSuccess) {
      *p = nullptr;
      return e;
    }
  }
  c.live[*p] = {dev, bs};
  return hipSuccess;
}

int reserve(Batch* b, void** buf, size_t* cap, size_t bytes) {
  if (*buf && *cap >= bytes) return 0;
  release(b, *buf);
  *buf = nullptr;
  *cap = 0;
  hipError_t e = batch_malloc(buf, bytes);
  if (e != hipSuccess) {
    *buf = nullptr;
    return check_hip(e, "hipMalloc(schedule lists)");
  }
  b->allocations.push_back(*buf);
  *cap = bytes;
  return 0;
}

void batch_free(void* p) {
  if (!p) return;
  // as hipFree: queued work that may still read the block ends first
  (void)hipDeviceSynchronize();
  MemCache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.live.find(p);
  if (it == c.live.end()) {
    (void)hipFree(p);
    return;
  }
  c.idle[it->second.first].emplace(it->second.second, p);
  c.live.erase(it);
}

namespace {

constexpr int kCopy = 0, kRefine = 1, kSide1 = 2;  // kSide1 + s - 1: physical side s = 1..4

long rows_of(const PatchDev& P) { return P.sk / P.sy; }

// the ni x nj block whose first cell is local l lies in P's ghosted box
bool block_in(const PatchDev& P, long l, long ni, long nj) {
  if (l < 0 || l >= P.sk || ni < 1 || nj < 1) return false;
  const long col = l % P.sy, row = l / P.sy;
  return col + ni <= P.sy && row + nj <= rows_of(P);
}

// the coarse cells a refine rectangle reads, with their +-1 neighbours in
// both directions, lie in C's ghosted box
bool refine_in(const PatchDev& C, const enl_rect_t& R, int r) {
  const int pa = R.phase & 255, pb = R.phase >> 8;
  if (pa >= r || pb >= r || R.ni < 1 || R.nj < 1 || R.sl < 0 || R.sl >= C.sk) return false;
  const long span_i = (pa + R.ni - 1) / r + 1, span_j = (pb + R.nj - 1) / r + 1;
  const long col = R.sl % C.sy, row = R.sl / C.sy;
  return col >= 1 && col + span_i <= C.sy - 1 && row >= 1 && row + span_j <= rows_of(C) - 1;
}

// source of a physical-side rectangle: the patch's edge column (sides 1, 2)
// or edge row (sides 3, 4) alongside it
bool side_in(const PatchDev& P, const enl_rect_t& R) {
  const int side = R.kind - kSide1 + 1;
  return side <= 2 ? block_in(P, R.sl, 1, R.nj) : block_in(P, R.sl, R.ni, 1);
}

// Per rectangle, blockIdx.x strided: its cells into the level's per-cell
// lists at the rectangle's offset in its list (off[q]).  Rectangles before
// n1 are the first pass (copy, refine, edge lists), the rest the corners.
__global__ void __launch_bounds__(256) expand_fill_kernel(const enl_rect_t* __restrict__ rects,
                                                          const int* __restrict__ off, int nrect, int n1,
                                                          const PatchDev* __restrict__ pat, int r, LevelFill L) {
  for (int q = blockIdx.x; q < nrect; q += gridDim.x) {
    const enl_rect_t R = rects[q];
    const long dsy = pat[R.dp].sy, ssy = pat[R.sp].sy;
    const int n = R.ni * R.nj, o = off[q];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      const int a = c % R.ni, b = c / R.ni, x = o + c;
      const int dst = R.dl + a + b * (int)dsy;
      if (R.kind == kCopy) {
        L.c_dp[x] = R.dp;
        L.c_dl[x] = dst;
        L.c_sp[x] = R.sp;
        L.c_sl[x] = R.sl + a + b * (int)ssy;
      } else if (R.kind == kRefine) {
        const int pa = (R.phase & 255) + a, pb = (R.phase >> 8) + b;
        const int ia = pa / r, ja = pb / r;
        L.i_dp[x] = R.dp;
        L.i_dl[x] = dst;
        L.i_sp[x] = R.sp;
        L.i_sl[x] = R.sl + ia + ja * (int)ssy;
        L.i_fx[x] = ((double)(pa - ia * r) + 0.5) / r - 0.5;
        L.i_fy[x] = ((double)(pb - ja * r) + 0.5) / r - 0.5;
      } else {
        const int side = R.kind - kSide1 + 1;
        const int src = side <= 2 ? R.sl + b * (int)dsy : R.sl + a;
        if (q < n1) {
          L.e_p[x] = R.dp;
          L.e_d[x] = dst;
          L.e_s[x] = src;
          L.e_n[x] = side;
        } else {
          L.k_p[x] = R.dp;
          L.k_d[x] = dst;
          L.k_s[x] = src;
          L.k_n[x] = side;
        }
      }
    }
  }
}

// Coarsen rectangles (coarse cells) into the per-cell coarsen list.
__global__ void __launch_bounds__(256) expand_coarsen_kernel(const enl_rect_t* __restrict__ rects,
                                                             const int* __restrict__ off, int nrect,
                                                             const PatchDev* __restrict__ pat, int r, int* cp, int* cl,
                                                             int* fp, int* fl) {
  for (int q = blockIdx.x; q < nrect; q += gridDim.x) {
    const enl_rect_t R = rects[q];
    const long csy = pat[R.dp].sy, fsy = pat[R.sp].sy;
    const int n = R.ni * R.nj, o = off[q];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      const int a = c % R.ni, b = c / R.ni, x = o + c;
      cp[x] = R.dp;
      cl[x] = R.dl + a + b * (int)csy;
      fp[x] = R.sp;
      fl[x] = R.sl + a * r + b * r * (int)fsy;
    }
  }
}

// Regrid transfer from rectangles: w_n, pres, temp (state block 0) and the
// source of each cell, copied from the old patch of the same level or
// refined from the old coarser level (the arithmetic of transfer_kernel).
__global__ void __launch_bounds__(256) transfer_rect_kernel(double* dpool, const PatchDev* __restrict__ dpat,
                                                            const double* spool, const PatchDev* __restrict__ spat,
                                                            const enl_rect_t* __restrict__ rects, int nrect, int r) {
  for (int q = blockIdx.x; q < nrect; q += gridDim.x) {
    const enl_rect_t R = rects[q];
    const PatchDev& D = dpat[R.dp];
    const PatchDev& S = spat[R.sp];
    const int n = R.ni * R.nj;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      const int a = c % R.ni, b = c / R.ni;
      const long dl = R.dl + a + b * D.sy;
      double* d = dpool + D.state[0] + dl;
      double* dsrc = dpool + D.src + dl;
      if (R.kind == kCopy) {
        const long sl = R.sl + a + b * S.sy;
        const double* s = spool + S.state[0] + sl;
        for (int k = 0; k < 7; k++) d[k * D.sk] = s[k * S.sk];
        *dsrc = spool[S.src + sl];
      } else {
        const int pa = (R.phase & 255) + a, pb = (R.phase >> 8) + b;
        const int ia = pa / r, ja = pb / r;
        const long sl = R.sl + ia + ja * S.sy;
        const double ax = ((double)(pa - ia * r) + 0.5) / r - 0.5;
        const double ay = ((double)(pb - ja * r) + 0.5) / r - 0.5;
        const double* s = spool + S.state[0] + sl;
        for (int k = 0; k < 7; k++) d[k * D.sk] = lin_refine(s + k * S.sk, S.sy, ax, ay);
        *dsrc = lin_refine(spool + S.src + sl, S.sy, ax, ay);
      }
    }
  }
}

__global__ void __launch_bounds__(256) zero_map_kernel(int* map, long n) {
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) map[q] = 0;
}

// tag_cells (tag_cells.m4:58-108) on every patch of `level` (blockIdx.y =
// patch), each tag written at its cell of the level's index space: the
// tag_kernel test, cell by cell.
__global__ void __launch_bounds__(256) tag_map_kernel(const double* pool, const PatchDev* __restrict__ dp, int level,
                                                      int* map, long nmap, int val, double ep0, double ep1) {
  const PatchDev pd = dp[blockIdx.y];
  if (pd.level != level) return;
  const double* h = pool + pd.src;
  const double* w = pool + pd.state[0];
  const long S = pd.sy, n = (long)pd.nx * pd.ny;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
    const long i = q % pd.nx, j = q / pd.nx;
    const long g = (i + pd.g) + (j + pd.g) * S;
    bool t = false;
    t |= smooth_ratio(h[g] - h[g - 1], h[g + 1] - h[g], ep0) < kThresh;
    t |= smooth_ratio(h[g] - h[g - S], h[g + S] - h[g], ep0) < kThresh;
    t |= smooth_ratio(w[g] - w[g - 1], w[g + 1] - w[g], ep1) < kThresh;
    t |= smooth_ratio(w[g] - w[g - S], w[g + S] - w[g], ep1) < kThresh;
    map[(pd.bj0 + j) * nmap + pd.bi0 + i] = t ? val : 0;
  }
}


// Uploads n rects and their offsets (off: n entries, or none) into the
// batch's staging buffer; the caller synchronizes the stream before the
// buffer is used again.
int upload_rects(Batch* b, const enl_rect_t* rects, const std::vector<int>& off, int n, char** dev, hipStream_t s) {
  const size_t rb = sizeof(enl_rect_t) * (size_t)n, ob = sizeof(int) * off.size();
  *dev = nullptr;
  int rc = reserve(b, &b->stage, &b->stage_cap, rb + ob);
  if (rc) return rc;
  *dev = (char*)b->stage;
  hipError_t e = hipMemcpyAsync(*dev, rects, rb, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && ob) e = hipMemcpyAsync(*dev + rb, off.data(), ob, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return check_hip(e, "upload rectangles");
  return 0;
}

void drop_graph_of(Batch* b) {
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
  }
}

}  // namespace
}  // namespace enl

using namespace enl;

extern "C" {

int enl_batch_trim(void) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return check_hip(e, "enl_batch_trim");
  MemCache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  for (auto& kv : c.idle[dev]) (void)hipFree(kv.second);
  c.idle[dev].clear();
  return 0;
}

int enl_batch_set_fill_rects(void* handle, int level, int n1, int n2, const enl_rect_t* rects) {
  Batch* b = (Batch*)handle;
  if (!b || level < 0 || level >= kMaxLevels) return set_error(-2, "enl_batch_set_fill_rects: bad batch or level");
  if (n1 < 0 || n2 < 0 || (n1 + n2 > 0 && !rects)) return set_error(-3, "enl_batch_set_fill_rects: bad rectangles");
  const int n = n1 + n2;
  const std::vector<PatchDev>& hp = b->hp;
  const int np = (int)hp.size();
  std::vector<int> off((size_t)n);
  long nc = 0, ni = 0, ne = 0, nk = 0;
  for (int q = 0; q < n; q++) {
    const enl_rect_t& R = rects[q];
    if (R.dp < 0 || R.dp >= np || R.sp < 0 || R.sp >= np || hp[R.dp].level != level ||
        !block_in(hp[R.dp], R.dl, R.ni, R.nj))
      return set_error(-2, "enl_batch_set_fill_rects: level %d rectangle %d out of range", level, q);
    const PatchDev& S = hp[R.sp];
    bool ok;
    if (R.kind == kCopy)
      ok = q < n1 && S.level == level && block_in(S, R.sl, R.ni, R.nj);
    else if (R.kind == kRefine)
      ok = q < n1 && S.level == level - 1 && refine_in(S, R, b->ratio);
    else
      ok = R.kind >= kSide1 && R.kind < kSide1 + 4 && R.sp == R.dp && side_in(S, R) && (q < n1 || R.kind < kSide1 + 2);
    if (!ok) return set_error(-2, "enl_batch_set_fill_rects: level %d rectangle %d out of range", level, q);
    const long cells = (long)R.ni * R.nj;
    long& cnt = q >= n1 ? nk : (R.kind == kCopy ? nc : (R.kind == kRefine ? ni : ne));
    if (cnt + cells > (1L << 30)) return set_error(-2, "enl_batch_set_fill_rects: too many cells");
    off[(size_t)q] = (int)cnt;
    cnt += cells;
  }
  drop_graph_of(b);
  const size_t ints = 4 * (size_t)(nc + ni + ne + nk);
  const size_t ibytes = (sizeof(int) * ints + 255) & ~(size_t)255;
  const size_t bytes = ibytes + sizeof(double) * 2 * (size_t)ni;
  int rc = reserve(b, &b->fill_buf[level], &b->fill_cap[level], std::max<size_t>(bytes, 256));
  if (rc) return rc;
  LevelFill& L = b->fill[level];
  L = LevelFill();
  L.ncopy = (int)nc;
  L.ninterp = (int)ni;
  L.nedge = (int)ne;
  L.ncorner = (int)nk;
  int* p = (int*)b->fill_buf[level];
  int** lists[16] = {&L.c_dp, &L.c_dl, &L.c_sp, &L.c_sl, &L.i_dp, &L.i_dl, &L.i_sp, &L.i_sl,
                     &L.e_p,  &L.e_d,  &L.e_s,  &L.e_n,  &L.k_p,  &L.k_d,  &L.k_s,  &L.k_n};
  const long len[4] = {nc, ni, ne, nk};
  for (int a = 0; a < 16; a++) {
    *lists[a] = p;
    p += len[a / 4];
  }
  L.i_fx = (double*)((char*)b->fill_buf[level] + ibytes);
  L.i_fy = L.i_fx + ni;
  if (n == 0) return 0;
  hipStream_t s = lib_stream();
  char* dev = nullptr;
  rc = upload_rects(b, rects, off, n, &dev, s);
  if (rc) return rc;
  hipLaunchKernelGGL(expand_fill_kernel, dim3(std::min(n, 4096)), dim3(256), 0, s, (const enl_rect_t*)dev,
                     (const int*)(dev + sizeof(enl_rect_t) * (size_t)n), n, n1, b->dp, b->ratio, L);
  hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(s);
  if (e != hipSuccess) return check_hip(e, "expand fill rectangles");
  return check_hip(e2, "expand fill rectangles sync");
}

int enl_batch_set_coarsen_rects(void* handle, int n, const enl_rect_t* rects) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (n < 0 || (n > 0 && !rects)) return set_error(-3, "enl_batch_set_coarsen_rects: bad rectangles");
  const std::vector<PatchDev>& hp = b->hp;
  const int np = (int)hp.size(), r = b->ratio;
  std::vector<int> off((size_t)n);
  long total = 0;
  for (int q = 0; q < n; q++) {
    const enl_rect_t& R = rects[q];
    if (R.dp < 0 || R.dp >= np || R.sp < 0 || R.sp >= np || R.kind != kCopy ||
        hp[R.sp].level != hp[R.dp].level + 1 || !block_in(hp[R.dp], R.dl, R.ni, R.nj) ||
        !block_in(hp[R.sp], R.sl, (long)R.ni * r, (long)R.nj * r))
      return set_error(-2, "enl_batch_set_coarsen_rects: rectangle %d out of range", q);
    off[(size_t)q] = (int)total;
    total += (long)R.ni * R.nj;
    if (total > (1L << 30)) return set_error(-2, "enl_batch_set_coarsen_rects: too many cells");
  }
  drop_graph_of(b);
  int rc = reserve(b, &b->co_buf, &b->co_cap, std::max<size_t>(sizeof(int) * 4 * (size_t)total, 256));
  if (rc) return rc;
  int* p = (int*)b->co_buf;
  b->ncoarsen = (int)total;
  b->co_cp = p;
  b->co_cl = p + total;
  b->co_fp = p + 2 * total;
  b->co_fl = p + 3 * total;
  if (n == 0) return 0;
  hipStream_t s = lib_stream();
  char* dev = nullptr;
  rc = upload_rects(b, rects, off, n, &dev, s);
  if (rc) return rc;
  hipLaunchKernelGGL(expand_coarsen_kernel, dim3(std::min(n, 4096)), dim3(256), 0, s, (const enl_rect_t*)dev,
                     (const int*)(dev + sizeof(enl_rect_t) * (size_t)n), n, b->dp, r, b->co_cp, b->co_cl, b->co_fp,
                     b->co_fl);
  hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(s);
  if (e != hipSuccess) return check_hip(e, "expand coarsen rectangles");
  return check_hip(e2, "expand coarsen rectangles sync");
}

int enl_batch_transfer_rects(void* new_handle, void* old_handle, const int* reflect, int nsame, const int* s_new,
                             const int* s_old, int nrect, const enl_rect_t* rects) {
  Batch* nb = (Batch*)new_handle;
  Batch* ob = (Batch*)old_handle;
  if (!nb || !ob || nb == ob) return set_error(-2, "enl_batch_transfer_rects: needs two distinct batches");
  if (!reflect) return set_error(-3, "enl_batch_transfer_rects: null reflect");
  if (nb->g != ob->g || nb->ratio != ob->ratio)
    return set_error(-2, "enl_batch_transfer_rects: ghost width or ratio differ");
  if (nsame < 0 || nrect < 0 || (nsame && !(s_new && s_old)) || (nrect && !rects))
    return set_error(-3, "enl_batch_transfer_rects: bad lists");
  const int r = nb->ratio;
  std::vector<enl_rect_t> all;
  all.reserve((size_t)nrect + 4 * (size_t)nsame);
  // unchanged patches: their interiors, in bands of at most 1024 cells
  for (int q = 0; q < nsame; q++) {
    if (s_new[q] < 0 || s_new[q] >= nb->npatch || s_old[q] < 0 || s_old[q] >= ob->npatch)
      return set_error(-2, "enl_batch_transfer_rects: patch pair %d names a bad patch", q);
    const enl_patch_desc_t &D = nb->desc[s_new[q]], &S = ob->desc[s_old[q]];
    if (D.level != S.level || memcmp(D.box, S.box, sizeof D.box) != 0)
      return set_error(-2, "enl_batch_transfer_rects: patch pair %d differs in level or box", q);
    const PatchDev &pd = nb->hp[s_new[q]], &ps = ob->hp[s_old[q]];
    const int band = std::max(1, 1024 / pd.nx);
    for (int j = 0; j < pd.ny; j += band)
      all.push_back(enl_rect_t{s_new[q], s_old[q], (int)(pd.g + (j + pd.g) * pd.sy), (int)(ps.g + (j + ps.g) * ps.sy),
                               pd.nx, std::min(band, pd.ny - j), kCopy, 0});
  }
  for (int q = 0; q < nrect; q++) {
    const enl_rect_t& R = rects[q];
    if (R.dp < 0 || R.dp >= nb->npatch || R.sp < 0 || R.sp >= ob->npatch)
      return set_error(-2, "enl_batch_transfer_rects: rectangle %d names a bad patch", q);
    const PatchDev &D = nb->hp[R.dp], &S = ob->hp[R.sp];
    const bool ok = block_in(D, R.dl, R.ni, R.nj) &&
                    (R.kind == kCopy ? S.level == D.level && block_in(S, R.sl, R.ni, R.nj)
                                     : R.kind == kRefine && S.level == D.level - 1 && refine_in(S, R, r));
    if (!ok) return set_error(-2, "enl_batch_transfer_rects: rectangle %d out of range", q);
    all.push_back(R);
  }
  hipStream_t s = lib_stream();
  // the old hierarchy's ghosts (source, then state block 0) feed the
  // refinement stencils at its patch edges
  for (int lev = 0; lev < ob->nlevels; lev++) fill_level(ob, lev, -1, -1, 1, reflect, s);
  fill_all(ob, 0, reflect, s);
  const int n = (int)all.size();
  if (n > 0) {
    char* dev = nullptr;
    int rc = upload_rects(nb, all.data(), std::vector<int>(), n, &dev, s);
    if (rc) return rc;
    hipLaunchKernelGGL(transfer_rect_kernel, dim3(std::min(n, 8192)), dim3(256), 0, s, nb->pool, nb->dp, ob->pool,
                       ob->dp, (const enl_rect_t*)dev, n, r);
    hipError_t e = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e != hipSuccess) return check_hip(e, "transfer rectangles launch");
    if (e2 != hipSuccess) return check_hip(e2, "transfer rectangles sync");
  }
  // the new hierarchy's ghosts: source (and its interior copy), then state
  int rc = enl_batch_prepare(nb, reflect);
  if (rc) return rc;
  fill_all(nb, 0, reflect, s);
  return check_hip(hipGetLastError(), "enl_batch_transfer_rects");
}

int enl_batch_tag_map(void* handle, int level, const int* reflect, int refine_tag_val, const double* tag_tolerance,
                      int* map, int nmap) {
  Batch* b = (Batch*)handle;
  if (!b || !reflect || !tag_tolerance || !map) return set_error(-3, "enl_batch_tag_map: null argument");
  if (level < 0 || level >= b->nlevels || nmap <= 0) return set_error(-2, "enl_batch_tag_map: bad level or size");
  if (b->npatch > 65535) return set_error(-2, "enl_batch_tag_map: more than 65535 patches");
  long most = 0;
  for (const PatchDev& p : b->hp) {
    if (p.level != level) continue;
    if (p.bi0 < 0 || p.bj0 < 0 || p.bi0 + p.nx > nmap || p.bj0 + p.ny > nmap)
      return set_error(-2, "enl_batch_tag_map: a level-%d patch lies outside the %d x %d map", level, nmap, nmap);
    most = std::max(most, (long)p.nx * p.ny);
  }
  const long cells = (long)nmap * nmap;
  hipStream_t s = lib_stream();
  const bool dev = is_device_ptr(map);
  int* out = map;
  if (!dev) {
    int rc = reserve(b, &b->tagmap, &b->tagmap_cap, sizeof(int) * (size_t)cells);
    if (rc) return rc;
    out = (int*)b->tagmap;
  }
  // the stencil reads the state's ghosts; the source's are filled by
  // enl_batch_prepare
  fill_all(b, 0, reflect, s);
  hipLaunchKernelGGL(zero_map_kernel, dim3(grid_for(cells)), dim3(256), 0, s, out, cells);
  if (most > 0) {
    const unsigned bx = (unsigned)std::min<long>((most + 255) / 256, 64);
    hipLaunchKernelGGL(tag_map_kernel, dim3(bx, b->npatch), dim3(256), 0, s, b->pool, b->dp, level, out, (long)nmap,
                       refine_tag_val, sensor_eps(tag_tolerance[0]), sensor_eps(tag_tolerance[1]));
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !dev) e = hipMemcpyAsync(map, out, sizeof(int) * (size_t)cells, hipMemcpyDeviceToHost, s);
  const hipError_t e2 = dev ? hipSuccess : hipStreamSynchronize(s);
  if (e != hipSuccess) return check_hip(e, "enl_batch_tag_map");
  return check_hip(e2, "enl_batch_tag_map sync");
}

}  // extern "C"
