// enl_batch.h -- the batched hierarchy's host-side state, shared by the FP64
// form (enl_batch.hip) and the FP32-perturbation form (enl_batch32.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "enl_kernels.h"
#include "enlightning_hip.h"

namespace enl {

constexpr int kMaxStages = 8;
constexpr int kMaxLevels = 8;

// One patch in device memory.  Offsets are in doubles from the pool base.
struct PatchDev {
  int nx, ny, g, level;
  int bi0, bj0;     // the box's lower corner in the level's index space
  long sy, sk;      // ghosted row / component strides
  long state[3];    // three ghosted state blocks [w_n(5), pres, temp], 7*sk each
  long src;         // ghosted source
  long src_int;     // interior copy of the source (the drop-in's layout)
  long rhs;         // interior rhs slots, 5*nx*ny
  double dx, dy, ylo;
};

// Device copies of one level's fill schedule.
struct LevelFill {
  int ncopy = 0, ninterp = 0, nedge = 0, ncorner = 0;
  int *c_dp = nullptr, *c_dl = nullptr, *c_sp = nullptr, *c_sl = nullptr;
  int *i_dp = nullptr, *i_dl = nullptr, *i_sp = nullptr, *i_sl = nullptr;
  double *i_fx = nullptr, *i_fy = nullptr;
  int *e_p = nullptr, *e_d = nullptr, *e_s = nullptr, *e_n = nullptr;
  int *k_p = nullptr, *k_d = nullptr, *k_s = nullptr, *k_n = nullptr;
};

struct Fp32Form;  // enl_batch32.hip

// One patch's rectangle of cells touched by one absorbing side (see
// enl_absorb.hip): normal index range [lo, hi], tangential [t0, t0 + nt),
// tmp at tab[off + normal index].
struct AbsRect {
  int patch, lo, hi, t0, nt, off;
};

struct Batch {
  int npatch = 0, g = 0, ratio = 0, nlevels = 0;
  std::vector<PatchDev> hp;
  std::vector<enl_patch_desc_t> desc;
  double* pool = nullptr;
  size_t pool_doubles = 0;
  PatchDev* dp = nullptr;
  int4* tiles = nullptr;
  int ntiles = 0;
  int tile_bx[8], tile_by[8];     // tile shape per level
  int level_tile_begin[9] = {0};  // tiles sorted by level
  // How a level's tiles are packed into the workgroups of its stage launch
  // (ENL_PACK_*): one tile per workgroup, or a persistent grid of
  // pack_wpc[level] waves per SIMD, each workgroup working through a pack
  // of tiles.
  // packs + level_pack_begin[level] holds the level's npack + 1 pack
  // offsets into its (reordered) tiles.
  int pack_mode[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int pack_wpc[8] = {4, 4, 4, 4, 4, 4, 4, 4};
  int ncu = 0;
  int* packs = nullptr;
  void* tile_buf = nullptr;  // tiles, then packs, in one allocation
  size_t tile_cap = 0;
  int level_pack_begin[9] = {0};
  int level_npack[8] = {0};
  bool tiles_dirty = true;  // tiles and packs are (re)built by ensure_tiles before the next launch
  // reusable device buffers for uploads of rectangles and for tag maps
  void* stage = nullptr;
  size_t stage_cap = 0;
  void* tagmap = nullptr;
  size_t tagmap_cap = 0;
  bool all_pow2 = true;
  LevelFill fill[8];
  void* fill_buf[8] = {nullptr};  // one allocation per level holding its lists (rectangle path)
  size_t fill_cap[8] = {0};
  void* co_buf = nullptr;         // the coarsen lists (rectangle path)
  size_t co_cap = 0;
  int ncoarsen = 0;
  int *co_cp = nullptr, *co_cl = nullptr, *co_fp = nullptr, *co_fl = nullptr;
  unsigned long long* list = nullptr;
  unsigned* count = nullptr;
  long cells = 0;
  long level_cells[8] = {0};
  RhsArgs* d_args = nullptr;  // kMaxStages x npatch argument blocks
  std::vector<RhsArgs> h_args;
  std::vector<unsigned char> args_key;  // the inputs h_args was built from
  std::vector<void*> allocations;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<unsigned char> graph_key;
  Fp32Form* f32 = nullptr;  // the FP32-perturbation form's buffers, once set up
  // absorbing envelopes (enl_batch_set_absorb): per side the rectangles of
  // every patch, grouped by level, and their tmp tables
  bool absorb = false;
  double ab_glo[2] = {0, 0}, ab_ghi[2] = {0, 0}, ab_width = 0;
  AbsRect* ab_rects = nullptr;
  double* ab_tab = nullptr;
  int ab_begin[4][9] = {{0}};
};

// Device memory of the batched hierarchies: blocks a destroyed batch frees
// stay in a per-device cache and serve the next batch of a similar size (a
// regrid replaces a ~600 MB pool every few steps).  batch_free waits for the
// device first, as hipFree does.
hipError_t batch_malloc(void** p, size_t bytes);
void batch_free(void* p);

inline unsigned grid_for(long n, int block = 256, long cap = 4096) {
  long b = (n + block - 1) / block;
  if (b < 1) b = 1;
  return (unsigned)(b > cap ? cap : b);
}

template <typename T>
inline int upload(Batch* b, T** dst, const T* src, long n) {
  *dst = nullptr;
  if (n <= 0) return 0;
  hipError_t e = batch_malloc((void**)dst, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return check_hip(e, "hipMalloc(batch table)");
  b->allocations.push_back(*dst);
  return check_hip(hipMemcpy(*dst, src, sizeof(T) * (size_t)n, hipMemcpyHostToDevice), "upload batch table");
}

// State blocks of RK stage st of an order-stage step: stage 0 reads block 0
// (the step's start state), stages alternate blocks 1 and 2, the last stage
// writes block 0.
inline int stage_out(int st, int order) { return (st == order - 1 && order > 1) ? 0 : (st % 2 == 0 ? 1 : 2); }
inline int stage_in(int st, int order) { return st == 0 ? 0 : stage_out(st - 1, order); }

// Absorbing envelopes of one level's patches in state block `block`, on the
// sides whose reflect flag is 0, in the reference's side order.
int absorb_level(Batch* b, int lev, int block, const int* reflect, const enl_gas_t* gas, hipStream_t s);

// Linear refinement of the coarse cell at c (row stride sy) to the fine
// centre (ax, ay) coarse widths from its centre, unlimited centred slopes.
__device__ inline double lin_refine(const double* c, long sy, double ax, double ay) {
  return c[0] + ax * (0.5 * (c[1] - c[-1])) + ay * (0.5 * (c[sy] - c[-sy]));
}

// Builds the tiles and packs of every level when a tile shape or packing
// changed since the last build (or none was built): called by the entry
// points that launch stage kernels, outside any graph capture.
int ensure_tiles(Batch* b);

// Device memory for `bytes` in *buf, reusing it when it holds enough (the
// allocation is the batch's, freed with it).
int reserve(Batch* b, void** buf, size_t* cap, size_t bytes);

// Fills the ghosts of every level of state block `buf`, level by level.
void fill_all(Batch* b, int buf, const int* reflect, hipStream_t s);

// Frees device memory the batch owns (batch_free) and drops it from its
// allocation list.
void release(Batch* b, void* p);

// Ghost fill of one level of FP64 state block `buf` (-1: the source) from the
// level's schedule; the coarse-to-fine refine reads block `cbuf` of the
// coarser level.
void fill_level(Batch* b, int level, int buf, int cbuf, int ncomp, const int* reflect, hipStream_t s);

// The FP64 rhs of every patch of state block `in` into the w_n components of
// block `out` (ghosts of `in` filled first).
int batch_eval_rhs64(Batch* b, int variant, const enl_gas_t* gas, const int* reflect, int in, int out,
                     hipStream_t s);

// Steps of the FP32-perturbation form (modes ENL_BATCH_FP32 / _FP32_GRAPH).
int fp32_advance(Batch* b, int mode, const enl_gas_t* gas, const int* reflect, int order, const double* a1,
                 const double* a2, const double* be, double dt, int nsteps, hipStream_t s);

// Frees the FP32 form's buffers (no-op when it was never set up).
void fp32_release(Batch* b);

}  // namespace enl
