// enl_batch.hip -- all patches of an AMR hierarchy advanced together.
//
// The reference advances a hierarchy patch by patch: for every RK stage
// SAMRAI fills the ghosts of each level (same-level copies, coarse-to-fine
// LINEAR_REFINE, setPhysicalBoundaryConditions, enlightning.C:790-1013) and
// calls enlightning::singleStep on each patch (enlightning.C:459-583); after
// the step it averages fine data onto the coarse level (CONSERVATIVE_COARSEN).
// Small patches (8..64 cells a side) make that launch-bound on a GPU.
//
// Here every patch lives in one device pool, described by a device patch
// table.  Each stage runs level by level, as SAMRAI's method-of-lines loop
// does (fill a level's ghosts, then step its patches, then the next level,
// whose coarse-to-fine ghosts therefore see the coarse level already
// advanced by this stage):
//   * the ghost fill of the level from a precomputed schedule (copy /
//     linear-refine / physical-BC passes: one launch each; the refine reads
//     coarse cells through the coarse patches' ghosts, refreshed after the
//     coarse level's stage, i.e. coarse interior data as SAMRAI's
//     temporary coarse level provides);
//   * one persistent launch of the fused stage kernel over the tile list of
//     all patches of the level (rhs + runge_kutta + update_state per cell;
//     WENO cells -- the sensor's picks on level 0, every cell above it -- go
//     to a compacted list);
//   * one persistent launch over that WENO list;
// and the step ends with one coarsen launch.  The stage sequence is captured
// once into a hipGraph and replayed.  A plain mode advances the same
// hierarchy through the drop-in entry points, patch by patch, for comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "enl_batch.h"
#include "enl_rhs_body.h"
#include "enlightning_hip.h"

#pragma clang fp contract(off)

namespace enl {
namespace {


// Per stage and patch the kernels read a prebuilt argument block (pointers
// into the pool, grid reciprocals, gas constants, RK coefficients) through a
// restrict pointer: uniform scalar loads, as for the single-patch kernels'
// kernel-argument block, instead of per-tile arithmetic held in registers.
#define ENL_BATCH_LDS(BX, BY, L, raw)                                        \
  using ShB = TileShape<BX, BY>;                                             \
  __shared__ double lds_w0[ShB::TN], lds_p[ShB::TN], lds_T[ShB::TN];         \
  __shared__ double lds_u[ShB::TN], lds_v[ShB::TN], lds_s[ShB::TN], lds_tn[ShB::TN]; \
  TileLds<BX, BY> L;                                                         \
  L.w[0] = lds_w0;                                                           \
  L.p = lds_p;                                                               \
  L.T = lds_T;                                                               \
  L.u = lds_u;                                                               \
  L.v = lds_v;                                                               \
  L.s = lds_s;                                                               \
  L.tn = lds_tn;                                                             \
  L.w[1] = L.w[2] = L.w[3] = L.w[4] = nullptr;                               \
  if (raw) {                                                                 \
    __shared__ double lds_raw[4][ShB::TN];                                   \
    for (int k_ = 1; k_ < 5; k_++) L.w[k_] = lds_raw[k_ - 1];                \
  }

// One tile per workgroup: sensor + DRP cells finish here, WENO cells go to
// the list (split_body).  The fused epilogue is capped at 128 VGPRs.
// The tile's LDS pointers re-based by an offset the compiler cannot see
// through (a scalar zero written by inline asm in each iteration of a tile
// loop): LDS addresses then depend on the iteration, so the loop-invariant
// code motion does not hoist the stencil's LDS addresses of a whole tile
// out of the loop and hold them in registers across it.
template <int BX, int BY>
__device__ inline TileLds<BX, BY> rebased(const TileLds<BX, BY>& L) {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  TileLds<BX, BY> R;
  for (int k = 0; k < 5; k++) R.w[k] = L.w[k] ? L.w[k] + z : nullptr;
  R.p = L.p + z;
  R.T = L.T + z;
  R.u = L.u + z;
  R.v = L.v + z;
  R.s = L.s + z;
  R.tn = L.tn + z;
  return R;
}

// PACKED = false: one tile per workgroup, tiles[blockIdx.x] (make_tiles
// stores them in dispatch order).  PACKED = true: workgroup blockIdx.x works
// through its pack, tiles packs[w] .. packs[w + 1] - 1, with a barrier
// between tiles for the LDS.
template <int V, bool P, int BX, int BY, bool PACKED>
__global__ void __launch_bounds__(BX* BY) __attribute__((amdgpu_waves_per_eu(5)))
batch_split_kernel(const int4* __restrict__ tiles, const int* __restrict__ packs, const RhsArgs* __restrict__ args) {
  ENL_BATCH_LDS(BX, BY, L, false)
  if constexpr (!PACKED) {
    const int4 tl = tiles[blockIdx.x];
    split_body<V, P, kStage, BX, BY>(args[tl.x], tl.y, tl.z, (unsigned)tl.x, L);
  } else {
    const int q1 = packs[blockIdx.x + 1];
#pragma clang loop unroll(disable)
    for (int q = packs[blockIdx.x]; q < q1; q++) {
      const int4 tl = tiles[q];
      TileLds<BX, BY> Lq = rebased(L);
      split_body<V, P, kStage, BX, BY>(args[tl.x], tl.y, tl.z, (unsigned)tl.x, Lq);
      __syncthreads();
    }
  }
}

// Every cell of the tile WENO (levels above 0, rhs_weno_* variants): the
// two directions on two thread planes (dense2_body), block (BX, BY, 2).
template <int V, bool P, int BX, int BY, bool PACKED>
__global__ void __launch_bounds__(2 * BX * BY) __attribute__((amdgpu_waves_per_eu(4)))
batch_dense_kernel(const int4* __restrict__ tiles, const int* __restrict__ packs, const RhsArgs* __restrict__ args) {
  ENL_BATCH_LDS(BX, BY, L, true)
  __shared__ double lds_x[10 * BX * BY];
  if constexpr (!PACKED) {
    const int4 tl = tiles[blockIdx.x];
    dense2_body<V, P, kStage, BX, BY>(args[tl.x], tl.y, tl.z, L, lds_x);
  } else {
    const int q1 = packs[blockIdx.x + 1];
#pragma clang loop unroll(disable)
    for (int q = packs[blockIdx.x]; q < q1; q++) {
      const int4 tl = tiles[q];
      TileLds<BX, BY> Lq = rebased(L);
      int z;
      asm volatile("s_mov_b32 %0, 0" : "=s"(z));
      dense2_body<V, P, kStage, BX, BY>(args[tl.x], tl.y, tl.z, Lq, lds_x + z);
      __syncthreads();
    }
  }
}

// The listed WENO cells of the level, ten lanes each (list_block).
template <int V, bool P>
__global__ void __launch_bounds__(256) batch_list_kernel(const RhsArgs* __restrict__ args,
                                                         const unsigned long long* __restrict__ list,
                                                         const unsigned* __restrict__ count) {
  list_block<V, P, kStage>([=](int pid) -> const RhsArgs& { return args[pid]; }, list, *count);
}

// Zeroes the WENO-list counters of every (stage, level) of a step.
__global__ void batch_reset_kernel(unsigned* count, int n) {
  if ((int)threadIdx.x < n) count[threadIdx.x] = 0;
}


// ---- ghost fill and coarsen -------------------------------------------------

struct FillArgs {
  double* pool;
  const PatchDev* patches;
  int buf;    // state block filled, or -1 for the source
  int cbuf;   // state block the coarse-to-fine refine reads (-1: source)
  int ncomp;  // 7 for a state block, 1 for the source
  int reflect[4];
};

__device__ inline double* fill_base(const FillArgs& f, const PatchDev& p) {
  return f.pool + (f.buf >= 0 ? p.state[f.buf] : p.src);
}
__device__ inline const double* coarse_base(const FillArgs& f, const PatchDev& p) {
  return f.pool + (f.cbuf >= 0 ? p.state[f.cbuf] : p.src);
}

// same-level copy: every component of the cell
__device__ inline void fill_copy(const FillArgs& f, int q, const int* dp, const int* dl, const int* sp, const int* sl) {
  const PatchDev& D = f.patches[dp[q]];
  const PatchDev& S = f.patches[sp[q]];
  double* d = fill_base(f, D) + dl[q];
  const double* s = fill_base(f, S) + sl[q];
  for (int k = 0; k < f.ncomp; k++) d[k * D.sk] = s[k * S.sk];
}

// linear refinement from the coarse cell holding the fine ghost cell:
// c + fx*(c[i+1]-c[i-1])/2 + fy*(c[j+1]-c[j-1])/2, (fx, fy) the fine cell
// centre's offset from the coarse centre in coarse cell widths
__device__ inline void fill_interp(const FillArgs& f, int q, const int* dp, const int* dl, const int* sp,
                                   const int* sl, const double* fx, const double* fy) {
  const PatchDev& D = f.patches[dp[q]];
  const PatchDev& C = f.patches[sp[q]];
  double* d = fill_base(f, D) + dl[q];
  const double* c = coarse_base(f, C) + sl[q];
  const double ax = fx[q], ay = fy[q];
  for (int k = 0; k < f.ncomp; k++) d[k * D.sk] = lin_refine(c + k * C.sk, C.sy, ax, ay);
}

// physical boundary: copy within the patch; side 1..4 = left, right, bottom,
// top selects the REFLECT flag and the momentum component it negates
__device__ inline void fill_bc(const FillArgs& f, int q, const int* pp, const int* dl, const int* sl,
                               const int* side) {
  const PatchDev& P = f.patches[pp[q]];
  double* base = fill_base(f, P);
  const int sd = side[q];
  const int negk = (f.ncomp > 2 && f.reflect[sd - 1]) ? (sd <= 2 ? 1 : 2) : -1;
  for (int k = 0; k < f.ncomp; k++) {
    double v = base[sl[q] + k * P.sk];
    base[dl[q] + k * P.sk] = k == negk ? -v : v;
  }
}

// Same-level copies, coarse-to-fine refinement and the physical edge strips
// of one level in one launch: their destinations are distinct ghost cells
// and every source is an interior cell (of a neighbour, of the coarse level,
// of the patch itself), so the three lists do not depend on each other.
__global__ void __launch_bounds__(256) fill_edges_kernel(FillArgs f, LevelFill L) {
  const int n = L.ncopy + L.ninterp + L.nedge;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    if (q < L.ncopy)
      fill_copy(f, q, L.c_dp, L.c_dl, L.c_sp, L.c_sl);
    else if (q < L.ncopy + L.ninterp)
      fill_interp(f, q - L.ncopy, L.i_dp, L.i_dl, L.i_sp, L.i_sl, L.i_fx, L.i_fy);
    else
      fill_bc(f, q - L.ncopy - L.ninterp, L.e_p, L.e_d, L.e_s, L.e_n);
  }
}

// corner blocks: they copy ghost cells the edge pass wrote, so they follow it
__global__ void __launch_bounds__(256) fill_corner_kernel(FillArgs f, int n, const int* pp, const int* dl,
                                                           const int* sl, const int* side) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    fill_bc(f, q, pp, dl, sl, side);
}

// Regrid transfer: the SOLN variables (w_n, pres, temp in state block 0,
// and the source) of each listed cell of a new hierarchy, copied from the
// old hierarchy's patch of the same level that held the cell, or refined
// from the old coarser level where none did.
struct TransferArgs {
  double* dpool;
  const PatchDev* dpat;
  const double* spool;
  const PatchDev* spat;
  int ncopy, ninterp;
  const int *c_dp, *c_dl, *c_sp, *c_sl;
  const int *i_dp, *i_dl, *i_sp, *i_sl;
  const double *i_fx, *i_fy;
};

__global__ void __launch_bounds__(256) transfer_kernel(TransferArgs t) {
  const int n = t.ncopy + t.ninterp;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const bool copy = q < t.ncopy;
    const int r = copy ? q : q - t.ncopy;
    const PatchDev& D = t.dpat[copy ? t.c_dp[r] : t.i_dp[r]];
    const PatchDev& S = t.spat[copy ? t.c_sp[r] : t.i_sp[r]];
    const long dl = copy ? t.c_dl[r] : t.i_dl[r];
    const long sl = copy ? t.c_sl[r] : t.i_sl[r];
    double* d = t.dpool + D.state[0] + dl;
    const double* c = t.spool + S.state[0] + sl;
    double* dsrc = t.dpool + D.src + dl;
    const double* csrc = t.spool + S.src + sl;
    if (copy) {
      for (int k = 0; k < 7; k++) d[k * D.sk] = c[k * S.sk];
      *dsrc = *csrc;
    } else {
      const double ax = t.i_fx[r], ay = t.i_fy[r];
      for (int k = 0; k < 7; k++) d[k * D.sk] = lin_refine(c + k * S.sk, S.sy, ax, ay);
      *dsrc = lin_refine(csrc, S.sy, ax, ay);
    }
  }
}

// Regrid transfer of whole patches whose box did not change: the interior
// of old patch s_old[y] into new patch s_new[y] (same shape).
__global__ void __launch_bounds__(256) transfer_patch_kernel(double* dpool, const PatchDev* dpat, const double* spool,
                                                              const PatchDev* spat, const int* s_new,
                                                              const int* s_old, int npairs) {
  for (int y = blockIdx.y; y < npairs; y += gridDim.y) {
    const PatchDev& D = dpat[s_new[y]];
    const PatchDev& S = spat[s_old[y]];
    const int n = D.nx * D.ny;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
      const int i = c % D.nx, j = c / D.nx;
      const long dl = (i + D.g) + (long)(j + D.g) * D.sy;
      const long sl = (i + S.g) + (long)(j + S.g) * S.sy;
      const double* src = spool + S.state[0] + sl;
      double* dst = dpool + D.state[0] + dl;
      for (int k = 0; k < 7; k++) dst[k * D.sk] = src[k * S.sk];
      dpool[D.src + dl] = spool[S.src + sl];
    }
  }
}

// fine-to-coarse average of the r x r fine cells over a coarse cell
__global__ void __launch_bounds__(256) coarsen_kernel(double* pool, const PatchDev* patches, int buf, int r, int n,
                                                       const int* cp, const int* cl, const int* fp, const int* fl) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const PatchDev& C = patches[cp[q]];
    const PatchDev& F = patches[fp[q]];
    double* c = pool + C.state[buf] + cl[q];
    const double* fb = pool + F.state[buf] + fl[q];
    const double inv = 1.0 / (double)(r * r);
    for (int k = 0; k < 7; k++) {
      double s = 0;
      for (int fj = 0; fj < r; fj++)
        for (int fi = 0; fi < r; fi++) s += fb[fi + fj * F.sy + k * F.sk];
      c[k * C.sk] = s * inv;
    }
  }
}

__global__ void copy_interior_block_kernel(double* dst, const double* src, PatchDev p) {
  // state A interior (w_n) -> the patch's interior ww slots (rhs layout)
  const int n = p.nx * p.ny;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 5 * n; q += gridDim.x * blockDim.x) {
    int k = q / n, c = q % n, i = c % p.nx, j = c / p.nx;
    dst[q] = src[(i + p.g) + (long)(j + p.g) * p.sy + (long)k * p.sk];
  }
}

bool pow2(double x) {
  if (!(x > 0) || !std::isfinite(x)) return false;
  int e;
  return std::frexp(x, &e) == 0.5;
}

}  // namespace

void fill_level(Batch* b, int level, int buf, int cbuf, int ncomp, const int* reflect, hipStream_t s) {
  FillArgs f;
  f.pool = b->pool;
  f.patches = b->dp;
  f.buf = buf;
  f.cbuf = cbuf;
  f.ncomp = ncomp;
  for (int k = 0; k < 4; k++) f.reflect[k] = reflect[k];
  const LevelFill& L = b->fill[level];
  const long n = (long)L.ncopy + L.ninterp + L.nedge;
  if (n) hipLaunchKernelGGL(fill_edges_kernel, dim3(grid_for(n)), dim3(256), 0, s, f, L);
  if (L.ncorner)
    hipLaunchKernelGGL(fill_corner_kernel, dim3(grid_for(L.ncorner)), dim3(256), 0, s, f, L.ncorner, L.k_p, L.k_d,
                       L.k_s, L.k_n);
}

void fill_all(Batch* b, int buf, const int* reflect, hipStream_t s) {
  for (int lev = 0; lev < b->nlevels; lev++) fill_level(b, lev, buf, buf, 7, reflect, s);
}

void release(Batch* b, void* p) {
  if (!p) return;
  batch_free(p);
  for (auto& a : b->allocations)
    if (a == p) a = nullptr;
}

namespace {
int make_tiles(Batch* b);
}  // namespace

int ensure_tiles(Batch* b) { return b->tiles_dirty ? make_tiles(b) : 0; }


namespace {


// Tiles of every level (patch by patch, row-major inside a patch) and their
// packing into workgroups.  Workgroup w of a launch runs on XCD w % 8 (the
// dispatcher deals workgroups round-robin over the 8 XCDs), so each XCD is
// given one contiguous range of the level's tiles, and its workgroups split
// that range:
//   ENL_PACK_SINGLE   one tile per workgroup, in the range's order (the
//                     dispatch order of xcd_tile);
//   ENL_PACK_STRIDED  a persistent grid of wpc waves per SIMD, the j-th
//                     workgroup of an XCD taking tiles j, j + P, j + 2P, ...
//                     of its range (P workgroups on the XCD): neighbouring
//                     tiles run side by side, as dispatched singly;
//   ENL_PACK_CHUNKED  the same grid, each workgroup taking one contiguous
//                     run of the range, the runs of equal estimated cost (a
//                     tile's valid cells plus half its area for the staging
//                     and the halo): a workgroup walks along its patches.
// The reordered tiles go to b->tiles, the pack offsets to b->packs.
int make_tiles(Batch* b) {
  if (!b->ncu) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || n <= 0)
      n = 256;
    b->ncu = n;
  }
  std::vector<int4> t, T;
  std::vector<int> packs;
  std::vector<double> cost;
  std::vector<int> cut;
  t.reserve(b->ntiles > 0 ? (size_t)b->ntiles : 1024);
  for (int lev = 0; lev < kMaxLevels; lev++) {
    b->level_tile_begin[lev] = (int)t.size();
    b->level_pack_begin[lev] = (int)packs.size();
    const int bx = b->tile_bx[lev], by = b->tile_by[lev];
    const int mode = b->pack_mode[lev];
    T.clear();
    cost.clear();
    for (int p = 0; p < b->npatch; p++)
      if (b->hp[p].level == lev)
        for (int j0 = 0; j0 < b->hp[p].ny; j0 += by)
          for (int i0 = 0; i0 < b->hp[p].nx; i0 += bx) {
            T.push_back(make_int4(p, i0, j0, 0));
            if (mode == ENL_PACK_CHUNKED)
              cost.push_back(std::min(bx, b->hp[p].nx - i0) * std::min(by, b->hp[p].ny - j0) + 0.5 * bx * by);
          }
    const int nt = (int)T.size();
    if (nt == 0) {
      b->level_npack[lev] = 0;
      continue;
    }
    const int q8 = nt / 8, r8 = nt % 8;
    auto range0 = [&](int x) { return x * q8 + std::min(x, r8); };  // XCD x: tiles [range0(x), range0(x + 1))
    // persistent grid: pack_wpc[lev] waves of 64 lanes per SIMD, 4 SIMDs per
    // CU, in workgroups of one tile's bx*by lanes
    const long waves = 4L * b->ncu * b->pack_wpc[lev];
    const int G = mode == ENL_PACK_SINGLE ? nt
                                          : (int)std::min<long>(nt, std::max<long>(8, waves * 64 / (bx * by)));
    auto npk = [&](int x) { return x < G ? (G - 1 - x) / 8 + 1 : 0; };  // packs w = x, x + 8, ... < G on XCD x
    // chunked: per XCD the tile index where each of its packs starts
    cut.assign((size_t)G + 8, 0);
    std::vector<int> cut0(9, 0);
    if (mode == ENL_PACK_CHUNKED) {
      int o = 0;
      for (int x = 0; x < 8; x++) {
        cut0[(size_t)x] = o;
        const int a0 = range0(x), a1 = range0(x + 1), P = npk(x);
        double total = 0, acc = 0;
        for (int k = a0; k < a1; k++) total += cost[(size_t)k];
        int jj = 0;
        cut[(size_t)o] = a0;
        for (int k = a0; k < a1; k++) {
          // the next pack starts when this one has its share of the cost
          while (jj + 1 < P && acc >= total * (jj + 1) / P) cut[(size_t)(o + ++jj)] = k;
          acc += cost[(size_t)k];
        }
        while (jj + 1 < P) cut[(size_t)(o + ++jj)] = a1;
        cut[(size_t)(o + P)] = a1;
        o += P + 1;
      }
    }
    const int base = (int)t.size();
    int npack = 0;
    for (int w = 0; w < G; w++) {
      const int x = w & 7, jw = w >> 3, P = npk(x), a0 = range0(x), a1 = range0(x + 1);
      const int start = (int)t.size();
      if (mode == ENL_PACK_CHUNKED) {
        const int c0 = cut[(size_t)(cut0[(size_t)x] + jw)], c1 = cut[(size_t)(cut0[(size_t)x] + jw + 1)];
        for (int k = c0; k < c1; k++) t.push_back(T[(size_t)k]);
      } else {
        // single (P = the XCD's tile count: one tile per pack) or strided
        for (int k = a0 + jw; k < a1; k += P) t.push_back(T[(size_t)k]);
      }
      if ((int)t.size() == start) continue;
      packs.push_back(start - base);
      npack++;
    }
    packs.push_back((int)t.size() - base);
    b->level_npack[lev] = npack;
  }
  b->level_tile_begin[kMaxLevels] = (int)t.size();
  b->level_pack_begin[kMaxLevels] = (int)packs.size();
  // tiles and pack offsets in one allocation and one upload, after queued
  // launches that read the previous ones
  (void)hipDeviceSynchronize();
  const size_t tb = sizeof(int4) * t.size(), pb = sizeof(int) * packs.size();
  int rc = reserve(b, &b->tile_buf, &b->tile_cap, tb + pb);
  if (rc) return rc;
  b->tiles = (int4*)b->tile_buf;
  b->packs = (int*)((char*)b->tile_buf + tb);
  b->ntiles = (int)t.size();
  std::vector<char> h(tb + pb);
  memcpy(h.data(), t.data(), tb);
  memcpy(h.data() + tb, packs.data(), pb);
  rc = check_hip(hipMemcpy(b->tile_buf, h.data(), tb + pb, hipMemcpyHostToDevice), "upload tiles");
  if (!rc) b->tiles_dirty = false;
  return rc;
}


// Arguments common to every patch: gas constants and the WENO list.
RhsArgs base_args(Batch* b, const enl_gas_t* gas) {
  RhsArgs base = {};
  memset(&base, 0, sizeof base);
  base.alpha = gas->weno_alpha;
  base.kappa = gas->kappa;
  base.mu = gas->mu;
  base.muB = gas->muB;
  base.R = gas->R;
  base.T_star_n = gas->T_star_n;
  base.tau_n = gas->tau_n;
  base.c = gas->c;
  base.c_lr = gas->c_lr;
  base.ep = sensor_eps(gas->tag_tolerance[2]);
  base.rho0 = gas->rho0;
  base.gamma1 = gas->gamma1;
  base.c_p = gas->c_p;
  base.p0 = gas->p0;
  base.T0 = gas->T0;
  base.c_v = gas->c_v;
  base.list = b->list;
  base.switch_map = nullptr;
  return base;
}

// One patch's arguments: read state block `in`, write block `out`, ww in
// block `ww`, WENO-list counter slot `slot`.
void patch_args(Batch* b, int p, int in, int out, int ww, int slot, RhsArgs& a) {
  const PatchDev& pd = b->hp[p];
  a.geom.nx = pd.nx;
  a.geom.ny = pd.ny;
  a.geom.g0 = pd.g;
  a.geom.g1 = pd.g;
  a.geom.sy = pd.sy;
  a.geom.sk = pd.sk;
  const double* w = b->pool + pd.state[in];
  a.w = w;
  a.p = w + 5 * pd.sk;
  a.T = w + 6 * pd.sk;
  a.H0 = b->pool + pd.src + pd.g + pd.g * pd.sy;
  a.h0_sy = pd.sy;
  a.KK = b->pool + pd.rhs;
  a.dx = pd.dx;
  a.dy = pd.dy;
  a.ylo = pd.ylo;
  a.level = pd.level;
  a.count = b->count + slot * kMaxLevels + pd.level;
  fill_rhs_constants(a);
  a.ww = b->pool + pd.state[ww];
  a.ww_off = pd.g + pd.g * pd.sy;
  a.ww_sy = pd.sy;
  a.ww_sk = pd.sk;
  double* o = b->pool + pd.state[out];
  a.w_out = o;
  a.p_out = o + 5 * pd.sk;
  a.T_out = o + 6 * pd.sk;
}

// Argument blocks of every (stage, patch), built on the host (the same
// fill_rhs_constants as the single-patch launch) and uploaded when they
// change; the kernels of stage st read b->d_args + st*npatch.
int prepare_args(Batch* b, const enl_gas_t* gas, int order, const double* a1, const double* a2, const double* be,
                 double dt, hipStream_t s) {
  // the table depends on the batch's fixed layout and on these inputs only
  std::vector<unsigned char> key(sizeof *gas + sizeof order + sizeof dt + 3 * sizeof(double) * (size_t)order);
  {
    unsigned char* k = key.data();
    memcpy(k, gas, sizeof *gas);
    k += sizeof *gas;
    memcpy(k, &order, sizeof order);
    k += sizeof order;
    memcpy(k, &dt, sizeof dt);
    k += sizeof dt;
    for (const double* v : {a1, a2, be}) {
      memcpy(k, v, sizeof(double) * (size_t)order);
      k += sizeof(double) * (size_t)order;
    }
  }
  if (!b->h_args.empty() && key == b->args_key) return 0;
  b->args_key = key;
  std::vector<RhsArgs> t((size_t)order * b->npatch, base_args(b, gas));
  for (int st = 0; st < order; st++) {
    for (int p = 0; p < b->npatch; p++) {
      RhsArgs& a = t[(size_t)st * b->npatch + p];
      a.alpha_1 = a1[st];
      a.alpha_2 = a2[st];
      a.bdt = be[st] * dt;
      patch_args(b, p, stage_in(st, order), stage_out(st, order), 0, st, a);
    }
  }
  const size_t bytes = sizeof(RhsArgs) * t.size();
  if (b->h_args.size() == t.size() && memcmp(b->h_args.data(), t.data(), bytes) == 0) return 0;
  b->h_args.swap(t);
  hipError_t e = hipMemcpyAsync(b->d_args, b->h_args.data(), bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return check_hip(e, "batch argument upload");
  return check_hip(hipStreamSynchronize(s), "batch argument upload sync");
}

template <int V, bool P, int BX, int BY>
void stage_launch(Batch* b, const RhsArgs* args, int st, int lev, hipStream_t s) {
  using Sc = Scheme<V>;
  const int np = b->level_npack[lev];
  if (np <= 0) return;
  const int4* tiles = b->tiles + b->level_tile_begin[lev];
  const int* packs = b->packs + b->level_pack_begin[lev];
  const bool packed = b->pack_mode[lev] != ENL_PACK_SINGLE;
  if (!Sc::drp || (Sc::hybrid && lev != 0)) {
    // rhs_weno_*, and hybrid variants above level 0 (the sensor is off)
    if (packed)
      hipLaunchKernelGGL((batch_dense_kernel<V, P, BX, BY, true>), dim3(np), dim3(BX, BY, 2), 0, s, tiles, packs, args);
    else
      hipLaunchKernelGGL((batch_dense_kernel<V, P, BX, BY, false>), dim3(np), dim3(BX, BY, 2), 0, s, tiles, packs,
                         args);
    return;
  }
  if (packed)
    hipLaunchKernelGGL((batch_split_kernel<V, P, BX, BY, true>), dim3(np), dim3(BX, BY), 0, s, tiles, packs, args);
  else
    hipLaunchKernelGGL((batch_split_kernel<V, P, BX, BY, false>), dim3(np), dim3(BX, BY), 0, s, tiles, packs, args);
  if (Sc::hybrid)
    hipLaunchKernelGGL((batch_list_kernel<V, P>), dim3(grid_for(b->level_cells[lev], kListCellsPerBlock, 4096)), dim3(256), 0, s,
                       args, (const unsigned long long*)b->list,
                       (const unsigned*)(b->count + st * kMaxLevels + lev));
}

template <int V, bool P>
int stage_shape(Batch* b, const RhsArgs* args, int st, int lev, hipStream_t s) {
  const int bx = b->tile_bx[lev], by = b->tile_by[lev];
  if (bx == 32 && by == 8) stage_launch<V, P, 32, 8>(b, args, st, lev, s);
  else if (bx == 16 && by == 16) stage_launch<V, P, 16, 16>(b, args, st, lev, s);
  else if (bx == 16 && by == 8) stage_launch<V, P, 16, 8>(b, args, st, lev, s);
  else if (bx == 8 && by == 8) stage_launch<V, P, 8, 8>(b, args, st, lev, s);
  else return set_error(-2, "unsupported tile shape");
  return 0;
}

template <bool P>
int stage_variant(Batch* b, int variant, const RhsArgs* args, int st, int lev, hipStream_t s) {
  switch (variant) {
    case kHybridHighLlf: return stage_shape<kHybridHighLlf, P>(b, args, st, lev, s);
    case kHybridHigh: return stage_shape<kHybridHigh, P>(b, args, st, lev, s);
    case kHybridLowLlf: return stage_shape<kHybridLowLlf, P>(b, args, st, lev, s);
    case kHybridLow: return stage_shape<kHybridLow, P>(b, args, st, lev, s);
    case kWenoHighLlf: return stage_shape<kWenoHighLlf, P>(b, args, st, lev, s);
    case kWenoHigh: return stage_shape<kWenoHigh, P>(b, args, st, lev, s);
    case kWenoLowLlf: return stage_shape<kWenoLowLlf, P>(b, args, st, lev, s);
    case kWenoLow: return stage_shape<kWenoLow, P>(b, args, st, lev, s);
    case kDrp: return stage_shape<kDrp, P>(b, args, st, lev, s);
    default: return set_error(-1, "unknown rhs variant");
  }
}

// Enqueues nsteps batched steps on stream s (argument blocks already
// prepared).  Per stage and level: fill the level's ghosts on the stage
// input (its coarse-to-fine refine reading the coarser level's stage
// output), step the level's tiles, and refresh the level's ghosts on its
// output for the next finer level's refine.  Level 0 has no coarse-to-fine
// part, so after stage 0 its refresh is already the next stage's fill.
// Absorbing envelopes of one level: one launch per absorbing side (left,
// right, bottom, top), one workgroup per patch rectangle.
__global__ void __launch_bounds__(256) batch_absorb_kernel(double* pool, const PatchDev* __restrict__ dp,
                                                           const AbsRect* __restrict__ rects,
                                                           const double* __restrict__ tab, int side, int block,
                                                           double rho0, double p0, double T0, double T0rho0) {
  const AbsRect r = rects[blockIdx.x];
  const PatchDev pd = dp[r.patch];
  double* w = pool + pd.state[block];
  double* p = w + 5 * pd.sk;
  double* T = w + 6 * pd.sk;
  const int nn = r.hi - r.lo + 1;
  const long n = (long)nn * r.nt;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    int i, j;
    if (side < 2) {
      i = (int)(q % nn) + r.lo;
      j = (int)(q / nn) + r.t0;
    } else {
      i = (int)(q % r.nt) + r.t0;
      j = (int)(q / r.nt) + r.lo;
    }
    const double tmp = tab[r.off + (side < 2 ? i : j)];
    const long g = (long)(i + pd.g) + (long)(j + pd.g) * pd.sy;
    p[g] = (p[g] - p0) * tmp + p0;
    T[g] = (T[g] - T0) * tmp + T0;
    w[g] = (w[g] - rho0) * tmp + rho0;
    w[g + pd.sk] = w[g + pd.sk] * tmp;
    w[g + 2 * pd.sk] = w[g + 2 * pd.sk] * tmp;
    w[g + 3 * pd.sk] = w[g + 3 * pd.sk] * tmp;
    w[g + 4 * pd.sk] = (w[g + 4 * pd.sk] - T0rho0) * tmp + T0rho0;
  }
}

}  // namespace

int absorb_level(Batch* b, int lev, int block, const int* reflect, const enl_gas_t* gas, hipStream_t s) {
  if (!b->absorb) return 0;
  for (int side = 0; side < 4; side++) {
    if (reflect[side]) continue;
    const int r0 = b->ab_begin[side][lev], r1 = b->ab_begin[side][lev + 1];
    if (r1 <= r0) continue;
    hipLaunchKernelGGL(batch_absorb_kernel, dim3((unsigned)(r1 - r0)), dim3(256), 0, s, b->pool, b->dp,
                       b->ab_rects + r0, b->ab_tab, side, block, gas->rho0, gas->p0, gas->T0, gas->T0 * gas->rho0);
  }
  return check_hip(hipGetLastError(), "batched absorb launch");
}

namespace {

int enqueue_batched(Batch* b, int variant, const enl_gas_t* gas, const int* reflect, int order, int nsteps,
                    hipStream_t s) {
  for (int step = 0; step < nsteps; step++) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, b->count, kMaxStages * kMaxLevels);
    for (int st = 0; st < order; st++) {
      const int in = stage_in(st, order), out = stage_out(st, order);
      const RhsArgs* args = b->d_args + (size_t)st * b->npatch;
      for (int lev = 0; lev < b->nlevels; lev++) {
        const bool refreshed = st > 0 && lev == 0 && b->nlevels > 1;
        if (!refreshed) fill_level(b, lev, in, out, 7, reflect, s);
        int rc = b->all_pow2 ? stage_variant<true>(b, variant, args, st, lev, s)
                             : stage_variant<false>(b, variant, args, st, lev, s);
        if (!rc) rc = absorb_level(b, lev, out, reflect, gas, s);
        if (rc) return rc;
        if (lev + 1 < b->nlevels) fill_level(b, lev, out, out, 7, reflect, s);
      }
    }
    if (order == 1) {
      for (int p = 0; p < b->npatch; p++) {
        const PatchDev& pd = b->hp[p];
        hipError_t e = hipMemcpyAsync(b->pool + pd.state[0], b->pool + pd.state[1], sizeof(double) * 7 * pd.sk,
                                      hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return check_hip(e, "RK1 state copy");
      }
    }
    if (b->ncoarsen)
      hipLaunchKernelGGL(coarsen_kernel, dim3(grid_for(b->ncoarsen)), dim3(256), 0, s, b->pool, b->dp, 0, b->ratio,
                         b->ncoarsen, b->co_cp, b->co_cl, b->co_fp, b->co_fl);
  }
  return check_hip(hipGetLastError(), "batched stage launch");
}

}  // namespace

// The rhs of every patch of state block `in`, stored into the w_n
// components of block `out`: the batched stage kernels with alpha_1 =
// alpha_2 = 0 and beta*dt = 1 (0*ww + 0*w + 1*K).  Ghosts of `in` are
// filled first, level by level.  Uses WENO-list counter slot kMaxStages-1
// and a temporary argument table.
int batch_eval_rhs64(Batch* b, int variant, const enl_gas_t* gas, const int* reflect, int in, int out,
                     hipStream_t s) {
  if (int rc = ensure_tiles(b)) return rc;
  std::vector<RhsArgs> t((size_t)b->npatch, base_args(b, gas));
  for (int p = 0; p < b->npatch; p++) {
    t[p].alpha_1 = 0;
    t[p].alpha_2 = 0;
    t[p].bdt = 1;
    patch_args(b, p, in, out, 0, kMaxStages - 1, t[p]);
  }
  RhsArgs* d = nullptr;
  hipError_t e = batch_malloc((void**)&d, sizeof(RhsArgs) * t.size());
  if (e != hipSuccess) return check_hip(e, "hipMalloc(eval arguments)");
  int rc = check_hip(hipMemcpy(d, t.data(), sizeof(RhsArgs) * t.size(), hipMemcpyHostToDevice), "eval arguments");
  if (!rc) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, b->count, kMaxStages * kMaxLevels);
    for (int lev = 0; lev < b->nlevels && !rc; lev++) {
      fill_level(b, lev, in, in, 7, reflect, s);
      rc = b->all_pow2 ? stage_variant<true>(b, variant, d, kMaxStages - 1, lev, s)
                       : stage_variant<false>(b, variant, d, kMaxStages - 1, lev, s);
    }
  }
  if (!rc) rc = check_hip(hipStreamSynchronize(s), "batch_eval_rhs64");
  batch_free(d);
  return rc;
}

namespace {

// The reference's way: the drop-in entry points called patch by patch
// (SAMRAI's copy current -> scratch, then per stage the ghost fill and
// rhs_ / runge_kutta_ / update_state_ on every patch).  ww goes into state
// block 1 (interior layout), the rhs into the rhs slots.
int enqueue_plain(Batch* b, int variant, const enl_gas_t* gas, const int* reflect, int order, const double* a1,
                  const double* a2, const double* be, double dt, int nsteps, hipStream_t s) {
  const int nequ = 5;
  const int nv = variant;
  for (int step = 0; step < nsteps; step++) {
    for (int p = 0; p < b->npatch; p++) {
      const PatchDev& pd = b->hp[p];
      hipLaunchKernelGGL(copy_interior_block_kernel, dim3(grid_for(5L * pd.nx * pd.ny)), dim3(256), 0, s,
                         b->pool + pd.state[1], b->pool + pd.state[0], pd);
    }
    for (int st = 0; st < order; st++) {
      for (int lev = 0; lev < b->nlevels; lev++) {
      fill_level(b, lev, 0, 0, 7, reflect, s);
      for (int p = 0; p < b->npatch; p++) {
        const PatchDev& pd = b->hp[p];
        if (pd.level != lev) continue;
        const enl_patch_desc_t& d = b->desc[p];
        double* w = b->pool + pd.state[0];
        double* pres = w + 5 * pd.sk;
        double* temp = w + 6 * pd.sk;
        double* ww = b->pool + pd.state[1];
        double* KK = b->pool + pd.rhs;
        double* src = b->pool + pd.src_int;  // the drop-in takes the source interior-only
        int rc = enl_rhs_variant(nv, &d.box[0], &d.box[1], &d.box[2], &d.box[3], &b->g, &b->g, d.xlo, d.dx, w, src,
                                 KK, pres, temp, &gas->weno_alpha, &gas->kappa, &gas->mu, &gas->muB, &gas->R,
                                 &gas->T_star_n, &gas->T_star_o, &gas->tau_n, &gas->tau_o, gas->tag_tolerance,
                                 &d.level, &gas->c, &gas->c_lr, &nequ, nullptr);
        if (rc) return rc;
        runge_kutta_(&d.box[0], &d.box[1], &d.box[2], &d.box[3], &b->g, &b->g, &dt, &a1[st], &a2[st], &be[st], w,
                     ww, KK, &nequ);
        update_state_(d.dx, d.xlo, &d.box[0], &d.box[1], &d.box[2], &d.box[3], &b->g, &b->g, w, pres, temp,
                      &gas->rho0, &gas->c, &gas->gamma1, &gas->c_p, &gas->p0, &gas->T0, &gas->c_lr, &gas->c_v,
                      &gas->R_tilde, &gas->M, &gas->R, &nequ);
        char msg[8];
        int code = enl_last_error(msg, sizeof msg);
        if (code) return set_error(code, "plain per-patch call failed");
        if (b->absorb) {
          // singleStep's last call (enlightning.C:581-582), patch by patch
          const int sides[4] = {!reflect[0], !reflect[1], !reflect[2], !reflect[3]};
          rc = enl_absorb(d.box[0], d.box[1], d.box[2], d.box[3], b->g, b->g, d.xlo, d.dx, b->ab_glo, b->ab_ghi,
                          sides, b->ab_width, gas, w, pres, temp);
          if (rc) return rc;
        }
      }
      if (lev + 1 < b->nlevels) fill_level(b, lev, 0, 0, 7, reflect, s);
      }
    }
    if (b->ncoarsen)
      hipLaunchKernelGGL(coarsen_kernel, dim3(grid_for(b->ncoarsen)), dim3(256), 0, s, b->pool, b->dp, 0, b->ratio,
                         b->ncoarsen, b->co_cp, b->co_cl, b->co_fp, b->co_fl);
  }
  return check_hip(hipGetLastError(), "plain hierarchy launch");
}

}  // namespace
}  // namespace enl

using namespace enl;

extern "C" {

int enl_batch_create(int npatch, const enl_patch_desc_t* patches, int gcw, int ratio, void** handle) {
  *handle = nullptr;
  if (npatch <= 0 || gcw < 3) return set_error(-2, "enl_batch_create: need patches and gcw >= 3");
  Batch* b = new Batch();
  b->npatch = npatch;
  b->g = gcw;
  b->ratio = ratio;
  b->desc.assign(patches, patches + npatch);
  long off = 0;
  for (int p = 0; p < npatch; p++) {
    const enl_patch_desc_t& d = patches[p];
    PatchDev pd;
    pd.nx = d.box[1] - d.box[0] + 1;
    pd.ny = d.box[3] - d.box[2] + 1;
    if (pd.nx <= 0 || pd.ny <= 0 || d.level < 0 || d.level > 7) {
      delete b;
      return set_error(-2, "enl_batch_create: bad patch %d", p);
    }
    pd.g = gcw;
    pd.level = d.level;
    pd.bi0 = d.box[0];
    pd.bj0 = d.box[2];
    pd.sy = pd.nx + 2L * gcw;
    pd.sk = pd.sy * (pd.ny + 2L * gcw);
    for (int s = 0; s < 3; s++) {
      pd.state[s] = off;
      off += 7 * pd.sk;
      off = (off + 31) & ~31L;  // 256-byte aligned blocks
    }
    pd.src = off;
    off += pd.sk;
    off = (off + 31) & ~31L;
    pd.src_int = off;
    off += (long)pd.nx * pd.ny;
    off = (off + 31) & ~31L;
    pd.rhs = off;
    off += 5L * pd.nx * pd.ny;
    off = (off + 31) & ~31L;
    pd.dx = d.dx[0];
    pd.dy = d.dx[1];
    pd.ylo = d.xlo[1];
    b->all_pow2 = b->all_pow2 && pow2(pd.dx) && pow2(pd.dy);
    b->hp.push_back(pd);
    b->cells += (long)pd.nx * pd.ny;
    b->level_cells[d.level] += (long)pd.nx * pd.ny;
    if (d.level + 1 > b->nlevels) b->nlevels = d.level + 1;
  }
  b->pool_doubles = (size_t)off;
  hipError_t e = batch_malloc((void**)&b->pool, sizeof(double) * b->pool_doubles);
  if (e != hipSuccess) {
    delete b;
    return check_hip(e, "hipMalloc(batch pool)");
  }
  b->allocations.push_back(b->pool);
  (void)hipMemset(b->pool, 0, sizeof(double) * b->pool_doubles);
  int rc = upload(b, &b->dp, b->hp.data(), npatch);
  if (!rc) {
    e = batch_malloc((void**)&b->list, sizeof(unsigned long long) * (size_t)b->cells);
    if (e == hipSuccess) b->allocations.push_back(b->list);
    if (e == hipSuccess) e = batch_malloc((void**)&b->count, sizeof(unsigned) * kMaxStages * kMaxLevels);
    if (e == hipSuccess) b->allocations.push_back(b->count);
    if (e == hipSuccess) e = hipMemset(b->count, 0, sizeof(unsigned) * kMaxStages * kMaxLevels);
    if (e == hipSuccess) e = batch_malloc((void**)&b->d_args, sizeof(RhsArgs) * (size_t)kMaxStages * npatch);
    if (e == hipSuccess) b->allocations.push_back(b->d_args);
    if (e != hipSuccess) rc = check_hip(e, "hipMalloc(batch list)");
  }
  for (int lev = 0; lev < kMaxLevels; lev++) {
    b->tile_bx[lev] = 32;
    b->tile_by[lev] = 8;
  }
  b->tiles_dirty = true;  // built by the first launch (ensure_tiles)
  if (rc) {
    for (void* a : b->allocations)
      if (a) batch_free(a);
    delete b;
    return rc;
  }
  *handle = b;
  return 0;
}

namespace {
// Host checks of schedule items before they reach a kernel.  A (patch,
// local) pair must name a cell of the patch's ghosted box; a refinement
// source also needs its +-1 cell and +-1 row neighbours in that box (a
// local in the first or last column passes a flat-index bound but its i+-1
// reads would wrap into the neighbouring row).
bool cell_ok(const std::vector<PatchDev>& hp, int p, long l) {
  return p >= 0 && p < (int)hp.size() && l >= 0 && l < hp[p].sk;
}
bool stencil_ok(const PatchDev& S, long l) {
  if (l < 0 || l >= S.sk) return false;
  const long col = l % S.sy, row = l / S.sy, rows = S.sk / S.sy;
  return col >= 1 && col <= S.sy - 2 && row >= 1 && row <= rows - 2;
}
// the r x r fine block whose lower-left cell is l lies in the fine box
bool block_ok(const PatchDev& F, long l, int r) {
  if (l < 0 || l >= F.sk) return false;
  const long col = l % F.sy, row = l / F.sy, rows = F.sk / F.sy;
  return col + r <= F.sy && row + r <= rows;
}
bool null_list(long n, const void* a, const void* b, const void* c, const void* d) {
  return n > 0 && !(a && b && c && d);
}
}  // namespace

int enl_batch_set_fill(void* handle, int level, int ncopy, const int* c_dp, const int* c_dl, const int* c_sp,
                       const int* c_sl, int ninterp, const int* i_dp, const int* i_dl, const int* i_sp,
                       const int* i_sl, const double* i_fx, const double* i_fy, int nedge, const int* e_p,
                       const int* e_d, const int* e_s, const int* e_n, int ncorner, const int* k_p, const int* k_d,
                       const int* k_s, const int* k_n) {
  Batch* b = (Batch*)handle;
  if (!b || level < 0 || level >= 8) return set_error(-2, "enl_batch_set_fill: bad level");
  if (ncopy < 0 || ninterp < 0 || nedge < 0 || ncorner < 0 || null_list(ncopy, c_dp, c_dl, c_sp, c_sl) ||
      null_list(ninterp, i_dp, i_dl, i_sp, i_sl) || (ninterp > 0 && !(i_fx && i_fy)) ||
      null_list(nedge, e_p, e_d, e_s, e_n) || null_list(ncorner, k_p, k_d, k_s, k_n))
    return set_error(-3, "enl_batch_set_fill: bad item lists");
  const std::vector<PatchDev>& hp = b->hp;
  for (int q = 0; q < ncopy; q++)
    if (!cell_ok(hp, c_dp[q], c_dl[q]) || !cell_ok(hp, c_sp[q], c_sl[q]) || hp[c_dp[q]].level != level ||
        hp[c_sp[q]].level != level)
      return set_error(-2, "enl_batch_set_fill: level %d copy item %d out of range", level, q);
  for (int q = 0; q < ninterp; q++)
    if (!cell_ok(hp, i_dp[q], i_dl[q]) || !cell_ok(hp, i_sp[q], i_sl[q]) || hp[i_dp[q]].level != level ||
        hp[i_sp[q]].level != level - 1 || !stencil_ok(hp[i_sp[q]], i_sl[q]))
      return set_error(-2, "enl_batch_set_fill: level %d refine item %d out of range", level, q);
  for (int q = 0; q < nedge + ncorner; q++) {
    const bool edge = q < nedge;
    const int r = edge ? q : q - nedge;
    const int pp = edge ? e_p[r] : k_p[r], side = edge ? e_n[r] : k_n[r];
    if (!cell_ok(hp, pp, edge ? e_d[r] : k_d[r]) || !cell_ok(hp, pp, edge ? e_s[r] : k_s[r]) || side < 1 ||
        side > 4 || hp[pp].level != level)
      return set_error(-2, "enl_batch_set_fill: level %d %s item %d out of range", level, edge ? "edge" : "corner", r);
  }
  LevelFill& L = b->fill[level];
  L.ncopy = ncopy;
  L.ninterp = ninterp;
  L.nedge = nedge;
  L.ncorner = ncorner;
  int rc = 0;
  rc = rc ? rc : upload(b, &L.c_dp, c_dp, ncopy);
  rc = rc ? rc : upload(b, &L.c_dl, c_dl, ncopy);
  rc = rc ? rc : upload(b, &L.c_sp, c_sp, ncopy);
  rc = rc ? rc : upload(b, &L.c_sl, c_sl, ncopy);
  rc = rc ? rc : upload(b, &L.i_dp, i_dp, ninterp);
  rc = rc ? rc : upload(b, &L.i_dl, i_dl, ninterp);
  rc = rc ? rc : upload(b, &L.i_sp, i_sp, ninterp);
  rc = rc ? rc : upload(b, &L.i_sl, i_sl, ninterp);
  rc = rc ? rc : upload(b, &L.i_fx, i_fx, ninterp);
  rc = rc ? rc : upload(b, &L.i_fy, i_fy, ninterp);
  rc = rc ? rc : upload(b, &L.e_p, e_p, nedge);
  rc = rc ? rc : upload(b, &L.e_d, e_d, nedge);
  rc = rc ? rc : upload(b, &L.e_s, e_s, nedge);
  rc = rc ? rc : upload(b, &L.e_n, e_n, nedge);
  rc = rc ? rc : upload(b, &L.k_p, k_p, ncorner);
  rc = rc ? rc : upload(b, &L.k_d, k_d, ncorner);
  rc = rc ? rc : upload(b, &L.k_s, k_s, ncorner);
  rc = rc ? rc : upload(b, &L.k_n, k_n, ncorner);
  return rc;
}

int enl_batch_set_coarsen(void* handle, int n, const int* cp, const int* cl, const int* fp, const int* fl) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (n < 0 || null_list(n, cp, cl, fp, fl)) return set_error(-3, "enl_batch_set_coarsen: bad item lists");
  for (int q = 0; q < n; q++)
    if (!cell_ok(b->hp, cp[q], cl[q]) || !cell_ok(b->hp, fp[q], fl[q]) ||
        b->hp[fp[q]].level != b->hp[cp[q]].level + 1 || !block_ok(b->hp[fp[q]], fl[q], b->ratio))
      return set_error(-2, "enl_batch_set_coarsen: item %d out of range", q);
  b->ncoarsen = n;
  int rc = 0;
  rc = rc ? rc : upload(b, &b->co_cp, cp, n);
  rc = rc ? rc : upload(b, &b->co_cl, cl, n);
  rc = rc ? rc : upload(b, &b->co_fp, fp, n);
  rc = rc ? rc : upload(b, &b->co_fl, fl, n);
  return rc;
}

namespace {
bool tile_ok(int bx, int by) {
  return (bx == 32 && by == 8) || (bx == 16 && by == 16) || (bx == 16 && by == 8) || (bx == 8 && by == 8);
}
void drop_graph(Batch* b) {
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
  }
}
}  // namespace

int enl_batch_set_tiles(void* handle, int bx, int by) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (!tile_ok(bx, by)) return set_error(-2, "unsupported tile shape %dx%d", bx, by);
  drop_graph(b);
  for (int lev = 0; lev < kMaxLevels; lev++) {
    b->tile_bx[lev] = bx;
    b->tile_by[lev] = by;
  }
  b->tiles_dirty = true;
  return 0;
}

int enl_batch_set_level_tiles(void* handle, int level, int bx, int by) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (level < 0 || level >= kMaxLevels) return set_error(-2, "bad level %d", level);
  if (!tile_ok(bx, by)) return set_error(-2, "unsupported tile shape %dx%d", bx, by);
  drop_graph(b);
  b->tile_bx[level] = bx;
  b->tile_by[level] = by;
  b->tiles_dirty = true;
  return 0;
}

int enl_batch_set_level_packing(void* handle, int level, int mode, int wpc) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (level < 0 || level >= kMaxLevels) return set_error(-2, "bad level %d", level);
  if (mode < ENL_PACK_SINGLE || mode > ENL_PACK_CHUNKED || wpc < 1 || wpc > 32)
    return set_error(-2, "enl_batch_set_level_packing: bad mode %d or waves per SIMD %d", mode, wpc);
  drop_graph(b);
  b->pack_mode[level] = mode;
  b->pack_wpc[level] = wpc;
  b->tiles_dirty = true;
  return 0;
}

int enl_batch_patch_state(void* handle, int pid, double** w_n, double** pres, double** temp, double** source) {
  Batch* b = (Batch*)handle;
  if (!b || pid < 0 || pid >= b->npatch) return set_error(-2, "bad patch id");
  const PatchDev& pd = b->hp[pid];
  double* w = b->pool + pd.state[0];
  if (w_n) *w_n = w;
  if (pres) *pres = w + 5 * pd.sk;
  if (temp) *temp = w + 6 * pd.sk;
  if (source) *source = b->pool + pd.src;
  return 0;
}

int enl_batch_copy_state(void* handle, int pid, int to_device, double* w_n, double* pres, double* temp,
                         double* source) {
  Batch* b = (Batch*)handle;
  if (!b || pid < 0 || pid >= b->npatch) return set_error(-2, "bad patch id");
  const PatchDev& pd = b->hp[pid];
  double* w = b->pool + pd.state[0];
  double* dev[4] = {w, w + 5 * pd.sk, w + 6 * pd.sk, b->pool + pd.src};
  double* host[4] = {w_n, pres, temp, source};
  size_t n[4] = {5 * (size_t)pd.sk, (size_t)pd.sk, (size_t)pd.sk, (size_t)pd.sk};
  hipStream_t s = lib_stream();
  for (int a = 0; a < 4; a++) {
    if (!host[a]) continue;
    hipError_t e = to_device ? hipMemcpyAsync(dev[a], host[a], sizeof(double) * n[a], hipMemcpyHostToDevice, s)
                             : hipMemcpyAsync(host[a], dev[a], sizeof(double) * n[a], hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return check_hip(e, "enl_batch_copy_state");
  }
  return check_hip(hipStreamSynchronize(s), "enl_batch_copy_state sync");
}

namespace {
// The interior of patch blockIdx.y's ghosted source into its interior copy.
__global__ void __launch_bounds__(256) source_interior_kernel(double* pool, const PatchDev* __restrict__ dp) {
  const PatchDev pd = dp[blockIdx.y];
  const long n = (long)pd.nx * pd.ny;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)gridDim.x * 256L) {
    const long i = q % pd.nx, j = q / pd.nx;
    pool[pd.src_int + q] = pool[pd.src + (i + pd.g) + (j + pd.g) * pd.sy];
  }
}
}  // namespace

int enl_batch_prepare(void* handle, const int* reflect) {
  // ghosts of the source (constant over the run) and its interior copy for
  // the plain mode's drop-in calls
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (b->npatch > 65535) return set_error(-2, "enl_batch_prepare: more than 65535 patches");
  hipStream_t s = lib_stream();
  for (int lev = 0; lev < b->nlevels; lev++) fill_level(b, lev, -1, -1, 1, reflect, s);
  long most = 0;
  for (const PatchDev& pd : b->hp) most = std::max(most, (long)pd.nx * pd.ny);
  const unsigned bx = (unsigned)std::min<long>((most + 255) / 256, 64);
  source_interior_kernel<<<dim3(bx, b->npatch), 256, 0, s>>>(b->pool, b->dp);
  return check_hip(hipGetLastError(), "enl_batch_prepare");
}

namespace {
// Zeroes, for patch blockIdx.y, its ghosted source and the interior copy
// that follows it in the pool (and the alignment padding between them).
__global__ void __launch_bounds__(256) zero_source_kernel(double* pool, const PatchDev* __restrict__ dp) {
  const PatchDev pd = dp[blockIdx.y];
  const long b = pd.src, e = pd.src_int + (long)pd.nx * pd.ny;
  for (long q = b + blockIdx.x * 256L + threadIdx.x; q < e; q += (long)gridDim.x * 256L) pool[q] = 0.0;
}
}  // namespace

int enl_batch_zero_source(void* handle) {
  // enlightning::setSourceToZero (enlightning.C:2310-2329), called by the
  // main loop after a step that added a source (main.C:476-480).  A kernel,
  // not a memset, so the call is also safe inside a captured graph.
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (b->npatch > 65535) return set_error(-2, "enl_batch_zero_source: more than 65535 patches");
  long most = 0;
  for (const PatchDev& pd : b->hp) most = std::max(most, pd.src_int + (long)pd.nx * pd.ny - pd.src);
  const unsigned bx = (unsigned)std::min<long>((most + 255) / 256, 64);
  zero_source_kernel<<<dim3(bx, b->npatch), 256, 0, lib_stream()>>>(b->pool, b->dp);
  return check_hip(hipGetLastError(), "enl_batch_zero_source");
}

int enl_batch_fill(void* handle, const int* reflect) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  fill_all(b, 0, reflect, lib_stream());
  return check_hip(hipGetLastError(), "enl_batch_fill");
}

int enl_batch_update_state(void* handle, const enl_gas_t* gas) {
  // p, T of every patch interior from its w_n (update_state.m4)
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  hipStream_t s = lib_stream();
  for (int p = 0; p < b->npatch; p++) {
    const PatchDev& pd = b->hp[p];
    StateArgs a;
    double* w = b->pool + pd.state[0];
    a.w = w;
    a.p = w + 5 * pd.sk;
    a.T = w + 6 * pd.sk;
    a.geom.nx = pd.nx;
    a.geom.ny = pd.ny;
    a.geom.g0 = a.geom.g1 = pd.g;
    a.geom.sy = pd.sy;
    a.geom.sk = pd.sk;
    a.dy = pd.dy;
    a.ylo = pd.ylo;
    a.rho0 = gas->rho0;
    a.c = gas->c;
    a.gamma1 = gas->gamma1;
    a.c_p = gas->c_p;
    a.p0 = gas->p0;
    a.T0 = gas->T0;
    a.c_lr = gas->c_lr;
    a.c_v = gas->c_v;
    a.R = gas->R;
    hipError_t e = launch_update_state(a, s);
    if (e != hipSuccess) return check_hip(e, "batch update_state");
  }
  return 0;
}

int enl_batch_advance(void* handle, int mode, int variant, const enl_gas_t* gas, const int* reflect, int order,
                      const double* alpha_1, const double* alpha_2, const double* beta, double dt, int nsteps) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  if (int rc = ensure_tiles(b)) return rc;
  if (variant < 0 || variant >= kNumVariants) return set_error(-1, "unknown rhs variant");
  if (order < 1 || order > kMaxStages) return set_error(-2, "RK order out of range");
  hipStream_t s = lib_stream();
  if (mode == ENL_BATCH_PLAIN) return enqueue_plain(b, variant, gas, reflect, order, alpha_1, alpha_2, beta, dt, nsteps, s);
  if (mode == ENL_BATCH_FP32 || mode == ENL_BATCH_FP32_GRAPH) {
    if (b->absorb && !(reflect[0] && reflect[1] && reflect[2] && reflect[3]))
      return set_error(-2, "absorbing envelopes are not available in the FP32-perturbation form");
    return fp32_advance(b, mode, gas, reflect, order, alpha_1, alpha_2, beta, dt, nsteps, s);
  }
  if (mode != ENL_BATCH_FUSED && mode != ENL_BATCH_GRAPH) return set_error(-2, "unknown batch mode");
  int prc = prepare_args(b, gas, order, alpha_1, alpha_2, beta, dt, s);
  if (prc) return prc;
  if (mode == ENL_BATCH_FUSED) return enqueue_batched(b, variant, gas, reflect, order, nsteps, s);
  // capture one step into a hipGraph (re-captured when the step changes) and
  // replay it nsteps times
  std::vector<unsigned char> key;
  auto put = [&key](const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    key.insert(key.end(), c, c + n);
  };
  put(&variant, sizeof variant);
  put(&order, sizeof order);
  put(&dt, sizeof dt);
  put(gas, sizeof *gas);
  put(reflect, 4 * sizeof(int));
  put(alpha_1, order * sizeof(double));
  put(alpha_2, order * sizeof(double));
  put(beta, order * sizeof(double));
  put(b->tile_bx, sizeof b->tile_bx);
  put(b->tile_by, sizeof b->tile_by);
  put(b->pack_mode, sizeof b->pack_mode);
  put(b->pack_wpc, sizeof b->pack_wpc);
  put(&b->absorb, sizeof b->absorb);
  put(&b->ab_rects, sizeof b->ab_rects);
  bool same = b->graph_exec && key == b->graph_key;
  if (!same) {
    if (b->graph_exec) {
      (void)hipGraphExecDestroy(b->graph_exec);
      (void)hipGraphDestroy(b->graph);
      b->graph_exec = nullptr;
      b->graph = nullptr;
    }
    hipStream_t cs;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return check_hip(e, "capture stream");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return check_hip(e, "hipStreamBeginCapture");
    }
    int rc = enqueue_batched(b, variant, gas, reflect, order, 1, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      return check_hip(e, "hipGraphInstantiate");
    }
    b->graph = g;
    b->graph_key = key;
  }
  for (int step = 0; step < nsteps; step++) {
    hipError_t e = hipGraphLaunch(b->graph_exec, s);
    if (e != hipSuccess) return check_hip(e, "hipGraphLaunch");
  }
  return 0;
}

int enl_batch_set_absorb(void* handle, const double* grid_xlo, const double* grid_xhi, double width) {
  Batch* b = (Batch*)handle;
  if (!b) return set_error(-2, "null batch");
  // queued steps may still read the old tables
  if (b->ab_rects || b->ab_tab) (void)hipStreamSynchronize(lib_stream());
  // a captured step holds the old table pointers and per-side grid sizes:
  // it is re-captured on the next graph advance
  drop_graph(b);
  if (b->ab_rects) {
    batch_free(b->ab_rects);
    b->ab_rects = nullptr;
  }
  if (b->ab_tab) {
    batch_free(b->ab_tab);
    b->ab_tab = nullptr;
  }
  for (int s = 0; s < 4; s++)
    for (int l = 0; l < 9; l++) b->ab_begin[s][l] = 0;
  b->absorb = width > 0;
  if (!b->absorb) return 0;
  if (!grid_xlo || !grid_xhi) return set_error(-3, "enl_batch_set_absorb: null domain bounds");
  for (int d = 0; d < 2; d++) {
    b->ab_glo[d] = grid_xlo[d];
    b->ab_ghi[d] = grid_xhi[d];
  }
  b->ab_width = width;
  // every patch's rectangles for all four sides (which sides apply is
  // decided by the reflect flags of each advance), grouped by side and level
  std::vector<double> tab;
  std::vector<AbsRect> rects[4][8];
  const int all[4] = {1, 1, 1, 1};
  for (int p = 0; p < b->npatch; p++) {
    const enl_patch_desc_t& d = b->desc[p];
    const Geom G = make_geom(d.box[0], d.box[1], d.box[2], d.box[3], b->g, b->g);
    AbsorbPlan plan;
    int rc = absorb_plan(G, d.xlo, d.dx, grid_xlo, grid_xhi, all, width, &plan);
    if (rc) return rc;
    const int base = (int)tab.size();
    tab.insert(tab.end(), plan.tab.begin(), plan.tab.end());
    for (int s = 0; s < 4; s++) {
      const AbsorbSide& P = plan.side[s];
      if (P.hi < P.lo) continue;
      AbsRect r;
      r.patch = p;
      r.lo = P.lo;
      r.hi = P.hi;
      r.t0 = (s < 2 && P.skip_first) ? 1 : 0;
      r.nt = (s < 2 ? G.ny : G.nx) - r.t0;
      r.off = base + P.off;
      if (r.nt > 0) rects[s][d.level].push_back(r);
    }
  }
  std::vector<AbsRect> flat;
  for (int s = 0; s < 4; s++) {
    for (int l = 0; l < 8; l++) {
      b->ab_begin[s][l] = (int)flat.size();
      flat.insert(flat.end(), rects[s][l].begin(), rects[s][l].end());
    }
    b->ab_begin[s][8] = (int)flat.size();
  }
  // levels are contiguous inside a side's block, so ab_begin[s][l + 1] ends
  // level l
  hipError_t e = hipSuccess;
  if (!flat.empty()) e = batch_malloc((void**)&b->ab_rects, sizeof(AbsRect) * flat.size());
  if (e == hipSuccess && !tab.empty()) e = batch_malloc((void**)&b->ab_tab, sizeof(double) * tab.size());
  if (e == hipSuccess && !flat.empty())
    e = hipMemcpy(b->ab_rects, flat.data(), sizeof(AbsRect) * flat.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !tab.empty())
    e = hipMemcpy(b->ab_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
  return check_hip(e, "enl_batch_set_absorb upload");
}

int enl_batch_transfer(void* new_handle, void* old_handle, const int* reflect, int nsame, const int* s_new,
                       const int* s_old, int ncopy, const int* c_dp, const int* c_dl, const int* c_sp,
                       const int* c_sl, int ninterp, const int* i_dp, const int* i_dl, const int* i_sp,
                       const int* i_sl, const double* i_fx, const double* i_fy) {
  Batch* nb = (Batch*)new_handle;
  Batch* ob = (Batch*)old_handle;
  if (!nb || !ob || nb == ob) return set_error(-2, "enl_batch_transfer: needs two distinct batches");
  if (!reflect) return set_error(-3, "enl_batch_transfer: null reflect");
  if (nb->g != ob->g || nb->ratio != ob->ratio) return set_error(-2, "enl_batch_transfer: ghost width or ratio differ");
  if (nsame < 0 || ncopy < 0 || ninterp < 0 || (nsame && !(s_new && s_old)) ||
      (ncopy && !(c_dp && c_dl && c_sp && c_sl)) || (ninterp && !(i_dp && i_dl && i_sp && i_sl && i_fx && i_fy)))
    return set_error(-3, "enl_batch_transfer: bad item lists");
  for (int q = 0; q < nsame; q++) {
    if (s_new[q] < 0 || s_new[q] >= nb->npatch || s_old[q] < 0 || s_old[q] >= ob->npatch)
      return set_error(-2, "enl_batch_transfer: patch pair %d names a bad patch", q);
    const enl_patch_desc_t &D = nb->desc[s_new[q]], &S = ob->desc[s_old[q]];
    if (D.level != S.level || memcmp(D.box, S.box, sizeof D.box) != 0)
      return set_error(-2, "enl_batch_transfer: patch pair %d differs in level or box", q);
  }
  // every item must address cells of its patches' ghosted boxes, with the
  // refinement stencil (+-1 cell, +-1 row) inside the coarse patch's box
  for (int q = 0; q < ncopy; q++) {
    if (c_dp[q] < 0 || c_dp[q] >= nb->npatch || c_sp[q] < 0 || c_sp[q] >= ob->npatch)
      return set_error(-2, "enl_batch_transfer: copy item %d names a bad patch", q);
    if (nb->hp[c_dp[q]].level != ob->hp[c_sp[q]].level || !cell_ok(nb->hp, c_dp[q], c_dl[q]) ||
        !cell_ok(ob->hp, c_sp[q], c_sl[q]))
      return set_error(-2, "enl_batch_transfer: copy item %d out of range", q);
  }
  for (int q = 0; q < ninterp; q++) {
    if (i_dp[q] < 0 || i_dp[q] >= nb->npatch || i_sp[q] < 0 || i_sp[q] >= ob->npatch)
      return set_error(-2, "enl_batch_transfer: refine item %d names a bad patch", q);
    const PatchDev &D = nb->hp[i_dp[q]], &S = ob->hp[i_sp[q]];
    if (S.level != D.level - 1 || !cell_ok(nb->hp, i_dp[q], i_dl[q]) || !stencil_ok(S, i_sl[q]))
      return set_error(-2, "enl_batch_transfer: refine item %d out of range", q);
  }
  hipStream_t s = lib_stream();
  // the old hierarchy's ghosts (source, then state block 0) feed the
  // refinement stencils at its patch edges
  for (int lev = 0; lev < ob->nlevels; lev++) fill_level(ob, lev, -1, -1, 1, reflect, s);
  fill_all(ob, 0, reflect, s);
  const long n = (long)ncopy + ninterp;
  if (n + nsame > 0) {
    const size_t ibytes = sizeof(int) * (size_t)(2 * nsame + 4 * ncopy + 4 * ninterp);
    const size_t bytes = ibytes + sizeof(double) * (size_t)(2 * ninterp);
    char* buf = nullptr;
    hipError_t e = batch_malloc((void**)&buf, bytes);
    if (e != hipSuccess) return check_hip(e, "hipMalloc(transfer items)");
    std::vector<char> h(bytes);
    int* hi = (int*)h.data();
    const int* lists[10] = {c_dp, c_dl, c_sp, c_sl, i_dp, i_dl, i_sp, i_sl, s_new, s_old};
    for (int a = 0; a < 10; a++) {
      const int cnt = a < 4 ? ncopy : (a < 8 ? ninterp : nsame);
      if (cnt) memcpy(hi, lists[a], sizeof(int) * (size_t)cnt);
      hi += cnt;
    }
    double* hd = (double*)(h.data() + ibytes);
    if (ninterp) {
      memcpy(hd, i_fx, sizeof(double) * (size_t)ninterp);
      memcpy(hd + ninterp, i_fy, sizeof(double) * (size_t)ninterp);
    }
    e = hipMemcpyAsync(buf, h.data(), bytes, hipMemcpyHostToDevice, s);
    TransferArgs t;
    t.dpool = nb->pool;
    t.dpat = nb->dp;
    t.spool = ob->pool;
    t.spat = ob->dp;
    t.ncopy = ncopy;
    t.ninterp = ninterp;
    const int* di = (const int*)buf;
    t.c_dp = di;
    t.c_dl = di + ncopy;
    t.c_sp = di + 2 * ncopy;
    t.c_sl = di + 3 * ncopy;
    t.i_dp = di + 4 * ncopy;
    t.i_dl = t.i_dp + ninterp;
    t.i_sp = t.i_dp + 2 * ninterp;
    t.i_sl = t.i_dp + 3 * ninterp;
    t.i_fx = (const double*)(buf + ibytes);
    t.i_fy = t.i_fx + ninterp;
    const int* sn = t.i_dp + 4 * ninterp;
    if (e == hipSuccess && n > 0) {
      hipLaunchKernelGGL(transfer_kernel, dim3(grid_for(n)), dim3(256), 0, s, t);
      e = hipGetLastError();
    }
    if (e == hipSuccess && nsame > 0) {
      const unsigned gy = (unsigned)(nsame < 65535 ? nsame : 65535);
      hipLaunchKernelGGL(transfer_patch_kernel, dim3(16, gy), dim3(256), 0, s, nb->pool, nb->dp, ob->pool, ob->dp,
                         sn, sn + nsame, nsame);
      e = hipGetLastError();
    }
    // the staging copy reads h, and the kernel reads buf: wait for both
    const hipError_t e2 = hipStreamSynchronize(s);
    batch_free(buf);
    if (e != hipSuccess) return check_hip(e, "transfer launch");
    if (e2 != hipSuccess) return check_hip(e2, "transfer sync");
  }
  // the new hierarchy's ghosts: source (and its interior copy), then state
  int rc = enl_batch_prepare(nb, reflect);
  if (rc) return rc;
  fill_all(nb, 0, reflect, s);
  return check_hip(hipGetLastError(), "enl_batch_transfer");
}

int enl_batch_tag_cells(void* handle, int level, const int* reflect, int refine_tag_val,
                        const double* tag_tolerance, int* tags) {
  Batch* b = (Batch*)handle;
  if (!b || !reflect || !tag_tolerance || !tags) return set_error(-3, "enl_batch_tag_cells: null argument");
  if (level < 0 || level >= b->nlevels) return set_error(-2, "enl_batch_tag_cells: bad level %d", level);
  long total = 0;
  for (const PatchDev& p : b->hp)
    if (p.level == level) total += (long)p.nx * p.ny;
  if (total == 0) return 0;
  hipStream_t s = lib_stream();
  const bool dev = is_device_ptr(tags);
  int* out = tags;
  if (!dev) {
    hipError_t e = batch_malloc((void**)&out, sizeof(int) * (size_t)total);
    if (e != hipSuccess) return check_hip(e, "hipMalloc(tags)");
  }
  // the stencil reads the state's ghosts; the source's are filled by
  // enl_batch_prepare
  fill_all(b, 0, reflect, s);
  hipError_t e = hipSuccess;
  long off = 0;
  for (const PatchDev& p : b->hp) {
    if (p.level != level) continue;
    TagArgs a;
    a.tags = out + off;
    a.w = b->pool + p.state[0];
    a.H0 = b->pool + p.src;
    a.geom.nx = p.nx;
    a.geom.ny = p.ny;
    a.geom.g0 = a.geom.g1 = p.g;
    a.geom.sy = p.sy;
    a.geom.sk = p.sk;
    a.refine_tag_val = refine_tag_val;
    a.ep0 = sensor_eps(tag_tolerance[0]);
    a.ep1 = sensor_eps(tag_tolerance[1]);
    if ((e = launch_tag_cells(a, s)) != hipSuccess) break;
    off += (long)p.nx * p.ny;
  }
  if (e == hipSuccess && !dev) e = hipMemcpyAsync(tags, out, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, s);
  const hipError_t e2 = dev ? hipSuccess : hipStreamSynchronize(s);
  if (!dev) batch_free(out);
  if (e != hipSuccess) return check_hip(e, "enl_batch_tag_cells");
  return check_hip(e2, "enl_batch_tag_cells sync");
}

int enl_batch_destroy(void* handle) {
  Batch* b = (Batch*)handle;
  if (!b) return 0;
  (void)hipDeviceSynchronize();
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  fp32_release(b);
  if (b->ab_rects) batch_free(b->ab_rects);
  if (b->ab_tab) batch_free(b->ab_tab);
  for (void* a : b->allocations)
    if (a) batch_free(a);
  delete b;
  return 0;
}

}  // extern "C"
