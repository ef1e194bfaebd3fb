// enl_regrid.hip -- the host side of a regrid in native code: clustering of
// the base level's tags into fine boxes, the ghost-fill schedule of each
// level, the fine-to-coarse (coarsen) schedule and the regrid transfer
// schedule (gridding_algorithm->regridAllFinerLevels, main.C:489-502, and
// the refine/coarsen schedules enlightning.C:259-288 registers).  The lists
// are those enlightning_amd/amr.py builds with numpy, item for item and in
// the same order; a plan keeps them until enl_plan_free.
//
// Beside the per-cell lists a plan holds the same schedules as rectangles
// (Rect: a block of cells whose addresses are affine in the cell's offset
// inside the block).  A regrid builds and uploads only the rectangles, a few
// thousand where the lists run to a million items; the device expands them
// into its per-cell lists (enl_batch_set_fill_rects, enl_batch_transfer_rects).
// The per-cell lists are built on first request (enl_plan_get), for the
// CPU twin and the tests.
//
// Every cell address is a (patch id, local index) pair, local index
// (i - ifirst0 + g) + Nx * (j - ifirst1 + g) inside the patch's ghosted box.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "enl_kernels.h"
#include "enlightning_hip.h"

namespace enl {
namespace {

struct Box {
  int level, i0, i1, j0, j1;
  int nx() const { return i1 - i0 + 1; }
  int ny() const { return j1 - j0 + 1; }
};

struct Owner;

struct Plan {
  int n0 = 0, ratio = 0, g = 0;
  std::vector<Box> boxes;  // pid order
  // owner maps per level, built once (shared with a later plan whose base
  // level is the same)
  std::map<int, std::shared_ptr<const Owner>> owners;
  std::map<std::string, std::vector<int>> ints;
  std::map<std::string, std::vector<double>> dbls;
  std::set<std::string> built;  // per-cell lists built so far ("fill@L", "coarsen")
  int nlevels() const {
    int m = 0;
    for (const Box& b : boxes) m = std::max(m, b.level);
    return boxes.empty() ? 0 : m + 1;
  }
  long ncells(int level) const {
    long n = n0;
    for (int l = 0; l < level; l++) n *= ratio;
    return n;
  }
};

inline int local(const Box& b, int g, int i, int j) { return (i - b.i0 + g) + (b.nx() + 2 * g) * (j - b.j0 + g); }

// A pool of host threads kept for the process (created on first use, never
// joined: a regrid runs several parallel loops of a millisecond or less, and
// starting threads for each cost as much as the loop).  A worker that has
// finished a loop polls for the next one for 200 us before it sleeps, since
// a regrid's loops follow each other closely and a wake-up from sleep can
// take as long as the loop.
class Pool {
 public:
  static Pool& get() {
    static Pool* p = new Pool();  // deliberately never destroyed
    return *p;
  }
  int size() const { return (int)workers_.size() + 1; }
  // job(k) for k in [0, n), items handed out one at a time; the caller works too
  void run(int n, int want, const std::function<void(int)>& job) {
    {
      std::lock_guard<std::mutex> lk(m_);
      job_ = &job;
      n_ = n;
      next_.store(0);
      active_ = std::min(want - 1, (int)workers_.size());
      pending_ = active_;
      gen_.fetch_add(1, std::memory_order_release);
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [&] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  Pool() {
    const int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    for (int t = 0; t + 1 < nt; t++) workers_.emplace_back([this, t] { loop(t); });
    for (std::thread& w : workers_) w.detach();
  }
  void drain() {
    for (int k; (k = next_.fetch_add(1)) < n_;) (*job_)(k);
  }
  void loop(int id) {
    in_pool() = true;
    unsigned long seen = 0;
    for (;;) {
      const auto t0 = std::chrono::steady_clock::now();
      while (gen_.load(std::memory_order_acquire) == seen &&
             std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(200))
        __builtin_ia32_pause();
      std::unique_lock<std::mutex> lk(m_);
      cv_.wait(lk, [&] { return gen_.load(std::memory_order_acquire) != seen; });
      seen = gen_.load(std::memory_order_acquire);
      if (id >= active_) continue;
      lk.unlock();
      drain();
      lk.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }

 public:
  static bool& in_pool() {
    static thread_local bool flag = false;
    return flag;
  }

 private:
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int n_ = 0, active_ = 0, pending_ = 0;
  std::atomic<unsigned long> gen_{0};
  std::atomic<int> next_{0};
};

// f(k) for k in [0, n) on up to 16 host threads (the pool), items handed
// out one at a time; every list below is written at offsets fixed
// beforehand, so the result does not depend on the threads.  Serial inside
// a pool thread and for short loops.
template <class F>
void parallel_for(int n, F f) {
  if (n <= 0) return;
  const int nt = std::min(16, std::max(1, n / 8));
  if (nt <= 1 || Pool::in_pool()) {
    for (int k = 0; k < n; k++) f(k);
    return;
  }
  static std::mutex one_at_a_time;  // the pool runs one loop at a time
  std::lock_guard<std::mutex> lk(one_at_a_time);
  const std::function<void(int)> job = [&](int k) { f(k); };
  Pool::in_pool() = true;  // a loop nested in f runs serially on this thread too
  Pool::get().run(n, nt, job);
  Pool::in_pool() = false;
}

// Exclusive prefix sums of counts[1..n] in place: counts[k] becomes the
// offset of item k, counts[n] the total.
long prefix(std::vector<long>& counts) {
  for (size_t k = 1; k < counts.size(); k++) counts[k] += counts[k - 1];
  return counts.back();
}

// Tile starts along one axis; a last tile narrower than g joins the one
// before it.
std::vector<int> cuts(int n0, int tile0, int g) {
  std::vector<int> s;
  for (int v = 0; v < n0; v += tile0) s.push_back(v);
  if (s.size() > 1 && n0 - s.back() < g) s.pop_back();
  s.push_back(n0);
  return s;
}

// The patch of a level holding each cell, at the resolution of the level
// below when every box of the level is aligned to the ratio (s = ratio).
// Owner maps of a few MB are built at every regrid; their buffers are kept
// for the next plan instead of being returned to the OS and faulted in again.
struct MapPool {
  std::mutex mu;
  std::vector<std::vector<int>> idle;
};
MapPool& map_pool() {
  static MapPool* p = new MapPool();  // never destroyed: outlives static teardown
  return *p;
}
std::vector<int> take_map(size_t n) {
  MapPool& p = map_pool();
  {
    std::lock_guard<std::mutex> lk(p.mu);
    for (size_t k = 0; k < p.idle.size(); k++)
      if (p.idle[k].size() == n) {
        std::vector<int> v = std::move(p.idle[k]);
        p.idle.erase(p.idle.begin() + (long)k);
        return v;
      }
  }
  return std::vector<int>(n);
}
void give_map(std::vector<int>&& v) {
  if (v.size() < ((size_t)1 << 16)) return;
  MapPool& p = map_pool();
  std::lock_guard<std::mutex> lk(p.mu);
  if (p.idle.size() < 4) p.idle.push_back(std::move(v));
}

struct Owner {
  std::vector<int> map;
  long n = 0, m = 0;  // level cells a side, map cells a side (n / s)
  int s = 1, sh = -1;  // s = 2^sh when s is a power of two
  int at(long i, long j) const {
    if (sh >= 0) return map[(size_t)((j >> sh) * m + (i >> sh))];
    return map[(size_t)((j / s) * m + i / s)];
  }
  Owner() = default;
  Owner(Owner&&) = default;
  Owner& operator=(Owner&&) = default;
  Owner(const Owner&) = delete;
  ~Owner() {
    if (!map.empty()) give_map(std::move(map));
  }
};

int log2_exact(long v) {
  int k = 0;
  while ((1L << k) < v) k++;
  return (1L << k) == v ? k : -1;
}

Owner build_owner(const Plan& P, int level) {
  Owner o;
  o.n = P.ncells(level);
  const int r = P.ratio;
  bool aligned = level > 0 && o.n % r == 0;
  for (const Box& b : P.boxes)
    if (b.level == level && aligned)
      aligned = b.i0 % r == 0 && (b.i1 + 1) % r == 0 && b.j0 % r == 0 && (b.j1 + 1) % r == 0;
  o.s = aligned ? r : 1;
  o.sh = log2_exact(o.s);
  const long m = o.n / o.s;
  o.m = m;
  o.map = take_map((size_t)(m * m));
  parallel_for((int)m, [&](int j) { std::fill(o.map.begin() + (long)j * m, o.map.begin() + (long)(j + 1) * m, -1); });
  parallel_for((int)P.boxes.size(), [&](int pid) {
    const Box& b = P.boxes[(size_t)pid];
    if (b.level != level) return;
    for (long j = b.j0 / o.s; j < (b.j1 + 1) / o.s; j++)
      for (long i = b.i0 / o.s; i < (b.i1 + 1) / o.s; i++) o.map[(size_t)(j * m + i)] = pid;
  });
  return o;
}

const Owner& owner_map(Plan& P, int level) {
  auto it = P.owners.find(level);
  if (it == P.owners.end()) it = P.owners.emplace(level, std::make_shared<const Owner>(build_owner(P, level))).first;
  return *it->second;
}

// The same base boxes (pid by pid) in two plans.
bool same_base(const Plan& a, const Plan& b) {
  if (a.n0 != b.n0 || a.ratio != b.ratio || a.g != b.g) return false;
  size_t x = 0, y = 0;
  for (;;) {
    while (x < a.boxes.size() && a.boxes[x].level != 0) x++;
    while (y < b.boxes.size() && b.boxes[y].level != 0) y++;
    const bool ea = x == a.boxes.size(), eb = y == b.boxes.size();
    if (ea || eb) return ea && eb;
    const Box &p = a.boxes[x], &q = b.boxes[y];
    if (x != y || p.i0 != q.i0 || p.i1 != q.i1 || p.j0 != q.j0 || p.j1 != q.j1) return false;
    x++;
    y++;
  }
}

// Fine boxes from the coarse tags: per block of fine_max/ratio coarse cells
// the bounding box of the tags dilated by nbuf cells (a square, clipped to
// the domain) inside the block, grown to fine_min/ratio coarse cells inside
// the block, extended to the domain edge when within one coarse cell of it
// (amr.buffer_tags, amr.cluster_fine_boxes).  The bounding box of the
// dilated tags inside a block is that of the dilated squares of the tags
// within nbuf cells of it, each clipped to the block, so the dilated map is
// never formed.  Blocks in parallel, boxes in block order.
void cluster(const int* tags, int n0, int nbuf, int r, int fine_max, int fine_min, std::vector<Box>& out) {
  const int blk = fine_max / r;
  const int cmin = std::max(1, fine_min / r);
  const int nb = (n0 + blk - 1) / blk;
  std::vector<Box> found((size_t)nb * nb, Box{-1, 0, 0, 0, 0});
  parallel_for(nb * nb, [&](int kb) {
    const int kj = kb / nb, ki = kb % nb;
    const int bi = ki * blk, bj = kj * blk;
    const int bi1 = std::min(bi + blk, n0) - 1, bj1 = std::min(bj + blk, n0) - 1;
    int i0 = n0, i1 = -1, j0 = n0, j1 = -1;
    const int ja = std::max(0, bj - nbuf), jb = std::min(n0 - 1, bj1 + nbuf);
    const int ia = std::max(0, bi - nbuf), ib = std::min(n0 - 1, bi1 + nbuf);
    for (int j = ja; j <= jb; j++) {
      const int* t = tags + (size_t)j * n0;
      for (int i = ia; i <= ib; i++)
        if (t[i]) {
          i0 = std::min(i0, std::max(i - nbuf, bi));
          i1 = std::max(i1, std::min(i + nbuf, bi1));
          j0 = std::min(j0, std::max(j - nbuf, bj));
          j1 = std::max(j1, std::min(j + nbuf, bj1));
        }
    }
    if (i1 < 0) return;
    while (i1 - i0 + 1 < cmin && (i1 < bi1 || i0 > bi)) {
      if (i1 < bi1)
        i1++;
      else
        i0--;
    }
    while (j1 - j0 + 1 < cmin && (j1 < bj1 || j0 > bj)) {
      if (j1 < bj1)
        j1++;
      else
        j0--;
    }
    if (i0 == 1 && bi == 0) i0 = 0;
    if (j0 == 1 && bj == 0) j0 = 0;
    if (i1 == n0 - 2 && bi1 == n0 - 1) i1 = n0 - 1;
    if (j1 == n0 - 2 && bj1 == n0 - 1) j1 = n0 - 1;
    found[(size_t)kb] = Box{1, i0 * r, (i1 + 1) * r - 1, j0 * r, (j1 + 1) * r - 1};
  });
  for (const Box& b : found)
    if (b.level == 1) out.push_back(b);
}

std::string lv(const char* name, int level) { return std::string(name) + "@" + std::to_string(level); }

// Ghost-fill schedule of one level (amr.make_fill).  Returns nonzero when a
// ghost cell inside the domain has no source.
int make_fill(Plan& P, int level) {
  const int g = P.g, r = P.ratio;
  const long n = P.ncells(level);
  const Owner& own = owner_map(P, level);
  const Owner& coarse = owner_map(P, level > 0 ? level - 1 : level);
  std::vector<int>&cp = P.ints[lv("fill.copy.p", level)], &cd = P.ints[lv("fill.copy.dst", level)],
  &cs = P.ints[lv("fill.copy.sp", level)], &cl = P.ints[lv("fill.copy.sl", level)];
  std::vector<int>&ip = P.ints[lv("fill.interp.p", level)], &id = P.ints[lv("fill.interp.dst", level)],
  &ic_ = P.ints[lv("fill.interp.cp", level)], &il = P.ints[lv("fill.interp.cl", level)];
  std::vector<double>&fx = P.dbls[lv("fill.interp.fx", level)], &fy = P.dbls[lv("fill.interp.fy", level)];
  std::vector<int>&ep = P.ints[lv("fill.edge.p", level)], &ed = P.ints[lv("fill.edge.dst", level)],
  &es = P.ints[lv("fill.edge.src", level)], &en = P.ints[lv("fill.edge.neg", level)];
  std::vector<int>&kp = P.ints[lv("fill.corner.p", level)], &kd = P.ints[lv("fill.corner.dst", level)],
  &ks = P.ints[lv("fill.corner.src", level)], &kn = P.ints[lv("fill.corner.neg", level)];
  const int rsh = log2_exact(r);
  std::vector<int> pats;
  for (size_t q = 0; q < P.boxes.size(); q++)
    if (P.boxes[q].level == level) pats.push_back((int)q);
  const int np = (int)pats.size();
  // the ghost cells of a patch inside the domain, in list order: the rows
  // below and above the box over the ghosted width, then the columns left
  // and right of it over the box rows
  auto visit = [&](const Box& b, auto&& f) {
    auto in = [&](int i, int j) { return i >= 0 && i < n && j >= 0 && j < n; };
    for (int h = 0; h < 2; h++)
      for (int l = 0; l < g; l++) {
        const int j = h == 0 ? b.j0 - g + l : b.j1 + 1 + l;
        for (int i = b.i0 - g; i <= b.i1 + g; i++)
          if (in(i, j)) f(i, j);
      }
    for (int j = b.j0; j <= b.j1; j++)
      for (int h = 0; h < 2; h++)
        for (int l = 0; l < g; l++) {
          const int i = h == 0 ? b.i0 - g + l : b.i1 + 1 + l;
          if (in(i, j)) f(i, j);
        }
  };
  // pass 1: copies and refinements per patch; pass 2: the items, written
  // at each patch's offsets
  std::vector<long> nc((size_t)np + 1, 0), ni((size_t)np + 1, 0);
  parallel_for(np, [&](int k) {
    long c = 0, m = 0;
    visit(P.boxes[(size_t)pats[(size_t)k]], [&](int i, int j) { (own.at(i, j) >= 0 ? c : m)++; });
    nc[(size_t)k + 1] = c;
    ni[(size_t)k + 1] = m;
  });
  const long tc = prefix(nc), ti = prefix(ni);
  if (level == 0 && ti > 0) return set_error(-2, "a level-0 ghost cell inside the domain has no owner");
  for (auto* v : {&cp, &cd, &cs, &cl}) v->resize((size_t)tc);
  for (auto* v : {&ip, &id, &ic_, &il}) v->resize((size_t)ti);
  fx.resize((size_t)ti);
  fy.resize((size_t)ti);
  std::atomic<int> outside(0);
  parallel_for(np, [&](int k) {
    const int pid = pats[(size_t)k];
    const Box& b = P.boxes[(size_t)pid];
    size_t oc = (size_t)nc[(size_t)k], oi = (size_t)ni[(size_t)k];
    visit(b, [&](int i, int j) {
      const int dst = local(b, g, i, j);
      const int o = own.at(i, j);
      if (o >= 0) {
        cp[oc] = pid;
        cd[oc] = dst;
        cs[oc] = o;
        cl[oc] = local(P.boxes[(size_t)o], g, i, j);
        oc++;
        return;
      }
      const int icc = rsh >= 0 ? i >> rsh : i / r, jcc = rsh >= 0 ? j >> rsh : j / r;
      const int c = coarse.at(icc, jcc);
      if (c < 0) {
        outside.store(1);
        return;
      }
      ip[oi] = pid;
      id[oi] = dst;
      ic_[oi] = c;
      il[oi] = local(P.boxes[(size_t)c], g, icc, jcc);
      fx[oi] = ((double)(i - icc * r) + 0.5) / r - 0.5;
      fy[oi] = ((double)(j - jcc * r) + 0.5) / r - 0.5;
      oi++;
    });
  });
  if (outside.load()) return set_error(-2, "a level-%d ghost cell lies outside level %d", level, level - 1);
  for (size_t q = 0; q < P.boxes.size(); q++) {
    const Box& b = P.boxes[q];
    if (b.level != level) continue;
    const int pid = (int)q;
    const bool touch[4] = {b.i0 == 0, b.i1 == n - 1, b.j0 == 0, b.j1 == n - 1};
    if (!(touch[0] || touch[1] || touch[2] || touch[3])) continue;
    auto edge = [&](int dst, int src, int neg) {
      ep.push_back(pid);
      ed.push_back(dst);
      es.push_back(src);
      en.push_back(neg);
    };
    if (touch[0])
      for (int l = 1; l <= g; l++)
        for (int j = b.j0; j <= b.j1; j++) edge(local(b, g, b.i0 - l, j), local(b, g, b.i0, j), 1);
    if (touch[1])
      for (int l = 1; l <= g; l++)
        for (int j = b.j0; j <= b.j1; j++) edge(local(b, g, b.i1 + l, j), local(b, g, b.i1, j), 2);
    if (touch[2])
      for (int l = 1; l <= g; l++)
        for (int i = b.i0; i <= b.i1; i++) edge(local(b, g, i, b.j0 - l), local(b, g, i, b.j0), 3);
    if (touch[3])
      for (int l = 1; l <= g; l++)
        for (int i = b.i0; i <= b.i1; i++) edge(local(b, g, i, b.j1 + l), local(b, g, i, b.j1), 4);
    for (int right = 0; right < 2; right++)
      for (int upper = 0; upper < 2; upper++) {
        const int ci0 = right ? b.i1 + 1 : b.i0 - g, cj0 = upper ? b.j1 + 1 : b.j0 - g;
        int nout = 0;
        for (int jj = cj0; jj < cj0 + g; jj++)
          for (int ii = ci0; ii < ci0 + g; ii++) nout += (ii < 0 || ii >= n || jj < 0 || jj >= n);
        if (nout == 0) continue;
        if (nout != g * g) return set_error(-2, "corner block straddles the domain edge");
        const int col = right ? b.i1 : b.i0;
        for (int jj = cj0; jj < cj0 + g; jj++)
          for (int ii = ci0; ii < ci0 + g; ii++) {
            kp.push_back(pid);
            kd.push_back(local(b, g, ii, jj));
            ks.push_back(local(b, g, col, jj));
            kn.push_back(right ? 2 : 1);
          }
      }
  }
  return 0;
}

// Coarse cells under each fine patch (amr.make_coarsen).
int make_coarsen(Plan& P) {
  const int r = P.ratio, g = P.g;
  const Owner& own0 = owner_map(P, 0);
  std::vector<int>&cp = P.ints["coarsen.cp"], &cl = P.ints["coarsen.cl"], &fp = P.ints["coarsen.fp"],
  &fl = P.ints["coarsen.fl"];
  std::vector<int> fine;
  for (size_t q = 0; q < P.boxes.size(); q++)
    if (P.boxes[q].level == 1) fine.push_back((int)q);
  std::vector<long> off(fine.size() + 1, 0);
  for (size_t k = 0; k < fine.size(); k++) {
    const Box& f = P.boxes[(size_t)fine[k]];
    off[k + 1] = (long)((f.i1 + 1) / r - f.i0 / r) * ((f.j1 + 1) / r - f.j0 / r);
  }
  const long total = prefix(off);
  for (auto* v : {&cp, &cl, &fp, &fl}) v->resize((size_t)total);
  std::atomic<int> outside(0);
  parallel_for((int)fine.size(), [&](int k) {
    const int q = fine[(size_t)k];
    const Box& f = P.boxes[(size_t)q];
    size_t o = (size_t)off[(size_t)k];
    for (int jc = f.j0 / r; jc <= (f.j1 + 1) / r - 1; jc++)
      for (int ic = f.i0 / r; ic <= (f.i1 + 1) / r - 1; ic++, o++) {
        const int c = own0.at(ic, jc);
        if (c < 0) {
          outside.store(1);
          return;
        }
        cp[o] = c;
        cl[o] = local(P.boxes[(size_t)c], g, ic, jc);
        fp[o] = q;
        fl[o] = local(f, g, ic * r, jc * r);
      }
  });
  if (outside.load()) return set_error(-2, "a fine patch covers a cell outside level 0");
  return 0;
}

// ---- rectangles ---------------------------------------------------------------
// One rectangle of a schedule, 8 ints (enl_rect_t in the public header): ni x nj
// cells; cell (a, b) has destination local dl + a + b*Nx(dp).  Its source:
//   kind 0 (copy):   sl + a + b*Nx(sp), the same cell in patch sp;
//   kind 1 (refine): the coarse cell sl + (pa/r) + (pb/r)*Nx(sp) of patch sp,
//                    pa = (phase & 255) + a, pb = (phase >> 8) + b, at offsets
//                    ((pa % r) + 0.5)/r - 0.5 and ((pb % r) + 0.5)/r - 0.5;
//   kind 1+s (physical side s = 1..4, sp = dp): sl + b*Nx for the left and
//                    right sides, sl + a for bottom and top.
// A coarsen rectangle lists coarse cells: coarse dl + a + b*Nx(dp) averages
// the r x r fine block whose lower-left cell is sl + a*r + b*r*Nx(sp).
struct Rect {
  int dp, sp, dl, sl, ni, nj, kind, phase;
};
static_assert(sizeof(Rect) == 8 * sizeof(int), "Rect is eight ints");

constexpr long long kSkip = -1;
inline long long key_of(int kind, int src) { return ((long long)kind << 32) | (unsigned)src; }

// The cells [ia, ib] x [ja, jb] in rectangles of one key: the runs of one
// key along each row, merged with the run of the row below when both ends
// and the key agree.  Cells whose key is kSkip are left out.
template <class Key, class Emit>
void decompose(int ia, int ib, int ja, int jb, Key key, Emit emit) {
  struct Run {
    int a, b, j0;
    long long k;
    bool open;
  };
  std::vector<Run> cur, nxt;
  for (int j = ja; j <= jb; j++) {
    nxt.clear();
    for (int i = ia; i <= ib;) {
      const long long k = key(i, j);
      int e = i;
      while (e < ib && key(e + 1, j) == k) e++;
      if (k != kSkip) {
        Run r{i, e, j, k, true};
        for (Run& o : cur)
          if (o.open && o.a == i && o.b == e && o.k == k) {
            r.j0 = o.j0;
            o.open = false;
            break;
          }
        nxt.push_back(r);
      }
      i = e + 1;
    }
    for (const Run& o : cur)
      if (o.open) emit(o.a, o.b, o.j0, j - 1, o.k);
    cur.swap(nxt);
  }
  for (const Run& o : cur) emit(o.a, o.b, o.j0, jb, o.k);
}

void put_rects(std::vector<int>& out, const std::vector<Rect>& r) {
  const size_t at = out.size();
  out.resize(at + 8 * r.size());
  if (!r.empty()) memcpy(out.data() + at, r.data(), sizeof(Rect) * r.size());
}

// Ghost-fill schedule of one level as rectangles: "fill.rects@L" holds the
// copies, refinements and physical edge strips (one pass: disjoint
// destinations, interior sources), then the corner blocks (a second pass:
// they copy edge ghosts); "fill.rects_n@L" = {first pass, corners}.  The
// same cells, sources and coefficients as make_fill's lists.
int make_fill_rects(Plan& P, int level) {
  const int g = P.g, r = P.ratio;
  const long n = P.ncells(level);
  const Owner& own = owner_map(P, level);
  const Owner& coarse = owner_map(P, level > 0 ? level - 1 : level);
  std::vector<int> pats;
  for (size_t q = 0; q < P.boxes.size(); q++)
    if (P.boxes[q].level == level) pats.push_back((int)q);
  const int np = (int)pats.size();
  std::vector<std::vector<Rect>> per((size_t)np);
  std::atomic<int> bad(0);
  parallel_for(np, [&](int k) {
    const int pid = pats[(size_t)k];
    const Box& b = P.boxes[(size_t)pid];
    auto key = [&](int i, int j) -> long long {
      if (i < 0 || i >= n || j < 0 || j >= n) return kSkip;  // the physical passes
      const int o = own.at(i, j);
      if (o >= 0) return key_of(0, o);
      const int c = level > 0 ? coarse.at(i / r, j / r) : -1;
      if (c < 0) {
        bad.store(level > 0 ? 2 : 1);
        return kSkip;
      }
      return key_of(1, c);
    };
    std::vector<Rect>& out = per[(size_t)k];
    auto emit = [&](int a, int bb, int j0, int j1, long long kk) {
      const int kind = (int)(kk >> 32), src = (int)(kk & 0xffffffffLL);
      const Box& S = P.boxes[(size_t)src];
      Rect R{pid, src, local(b, g, a, j0), 0, bb - a + 1, j1 - j0 + 1, kind, 0};
      if (kind == 0) {
        R.sl = local(S, g, a, j0);
      } else {
        const int ic = a / r, jc = j0 / r;
        R.sl = local(S, g, ic, jc);
        R.phase = (a - ic * r) | ((j0 - jc * r) << 8);
      }
      out.push_back(R);
    };
    // the list order of make_fill: rows below and above, then the columns
    // left and right of the box rows
    decompose(b.i0 - g, b.i1 + g, b.j0 - g, b.j0 - 1, key, emit);
    decompose(b.i0 - g, b.i1 + g, b.j1 + 1, b.j1 + g, key, emit);
    decompose(b.i0 - g, b.i0 - 1, b.j0, b.j1, key, emit);
    decompose(b.i1 + 1, b.i1 + g, b.j0, b.j1, key, emit);
  });
  if (bad.load() == 1) return set_error(-2, "a level-0 ghost cell inside the domain has no owner");
  if (bad.load() == 2) return set_error(-2, "a level-%d ghost cell lies outside level %d", level, level - 1);
  std::vector<Rect> first, corners;
  for (const auto& v : per) first.insert(first.end(), v.begin(), v.end());
  for (int pid : pats) {
    const Box& b = P.boxes[(size_t)pid];
    const bool touch[4] = {b.i0 == 0, b.i1 == n - 1, b.j0 == 0, b.j1 == n - 1};
    if (touch[0]) first.push_back(Rect{pid, pid, local(b, g, b.i0 - g, b.j0), local(b, g, b.i0, b.j0), g, b.ny(), 2, 0});
    if (touch[1]) first.push_back(Rect{pid, pid, local(b, g, b.i1 + 1, b.j0), local(b, g, b.i1, b.j0), g, b.ny(), 3, 0});
    if (touch[2]) first.push_back(Rect{pid, pid, local(b, g, b.i0, b.j0 - g), local(b, g, b.i0, b.j0), b.nx(), g, 4, 0});
    if (touch[3]) first.push_back(Rect{pid, pid, local(b, g, b.i0, b.j1 + 1), local(b, g, b.i0, b.j1), b.nx(), g, 5, 0});
    if (!(touch[0] || touch[1] || touch[2] || touch[3])) continue;
    for (int right = 0; right < 2; right++)
      for (int upper = 0; upper < 2; upper++) {
        const int ci0 = right ? b.i1 + 1 : b.i0 - g, cj0 = upper ? b.j1 + 1 : b.j0 - g;
        int nout = 0;
        for (int jj = cj0; jj < cj0 + g; jj++)
          for (int ii = ci0; ii < ci0 + g; ii++) nout += (ii < 0 || ii >= n || jj < 0 || jj >= n);
        if (nout == 0) continue;
        if (nout != g * g) return set_error(-2, "corner block straddles the domain edge");
        // the patch's edge column at each of the block's rows, the left or
        // right side's negation
        const int col = right ? b.i1 : b.i0;
        corners.push_back(Rect{pid, pid, local(b, g, ci0, cj0), local(b, g, col, cj0), g, g, right ? 3 : 2, 0});
      }
  }
  std::vector<int>& out = P.ints[lv("fill.rects", level)];
  out.clear();
  put_rects(out, first);
  put_rects(out, corners);
  P.ints[lv("fill.rects_n", level)] = {(int)first.size(), (int)corners.size()};
  return 0;
}

// Fine-to-coarse schedule as rectangles of coarse cells ("coarsen.rects"):
// per fine patch its coarse footprint, split by the level-0 patch holding
// each coarse cell.
int make_coarsen_rects(Plan& P) {
  const int r = P.ratio, g = P.g;
  const Owner& own0 = owner_map(P, 0);
  std::vector<int> fine;
  for (size_t q = 0; q < P.boxes.size(); q++)
    if (P.boxes[q].level == 1) fine.push_back((int)q);
  std::vector<std::vector<Rect>> per(fine.size());
  std::atomic<int> outside(0);
  parallel_for((int)fine.size(), [&](int k) {
    const int q = fine[(size_t)k];
    const Box& f = P.boxes[(size_t)q];
    auto key = [&](int ic, int jc) -> long long {
      const int c = own0.at(ic, jc);
      if (c < 0) {
        outside.store(1);
        return kSkip;
      }
      return key_of(0, c);
    };
    decompose(f.i0 / r, (f.i1 + 1) / r - 1, f.j0 / r, (f.j1 + 1) / r - 1, key,
              [&](int a, int bb, int j0, int j1, long long kk) {
                const int c = (int)(kk & 0xffffffffLL);
                per[(size_t)k].push_back(Rect{c, q, local(P.boxes[(size_t)c], g, a, j0), local(f, g, a * r, j0 * r),
                                              bb - a + 1, j1 - j0 + 1, 0, 0});
              });
  });
  if (outside.load()) return set_error(-2, "a fine patch covers a cell outside level 0");
  std::vector<Rect> all;
  for (const auto& v : per) all.insert(all.end(), v.begin(), v.end());
  std::vector<int>& out = P.ints["coarsen.rects"];
  out.clear();
  put_rects(out, all);
  return 0;
}

// The per-cell lists of a level's fill / of the coarsen schedule, built on
// first request.
int ensure_cells(Plan& P, const char* what, int level) {
  const std::string tag = strcmp(what, "fill") == 0 ? lv("fill", level) : std::string(what);
  if (P.built.count(tag)) return 0;
  int rc = 0;
  if (strcmp(what, "fill") == 0) {
    if (level < 0 || level >= P.nlevels()) return 0;
    rc = make_fill(P, level);
  } else if (P.nlevels() > 1) {
    rc = make_coarsen(P);
  }
  if (!rc) P.built.insert(tag);
  return rc;
}

int schedules(Plan& P, const Plan* previous) {
  const int L = P.nlevels();
  const bool reuse = previous && same_base(P, *previous);
  if (reuse) {
    auto it = previous->owners.find(0);
    if (it != previous->owners.end()) P.owners[0] = it->second;
  }
  for (int lev = 0; lev < L; lev++) {
    if (lev == 0 && reuse && previous->ints.count(lv("fill.rects", 0))) {
      // the same base boxes: the level-0 schedule is the previous plan's
      for (const char* name : {"fill.rects", "fill.rects_n"}) P.ints[lv(name, 0)] = previous->ints.at(lv(name, 0));
    } else {
      int rc = make_fill_rects(P, lev);
      if (rc) return rc;
    }
  }
  return L > 1 ? make_coarsen_rects(P) : 0;
}

}  // namespace
}  // namespace enl

using namespace enl;

extern "C" {

int enl_plan_build(int n0, int ratio, int g, int tile0, const int* tags, int tag_buffer, int fine_max, int fine_min,
                   const void* previous, void** plan) {
  if (!plan || n0 <= 0 || ratio < 2 || g < 1 || tile0 <= 0 || tag_buffer < 0 || fine_max < ratio || fine_min < 1)
    return set_error(-2, "enl_plan_build: bad arguments");
  *plan = nullptr;
  Plan* P = new (std::nothrow) Plan();
  if (!P) return set_error(-5, "enl_plan_build: out of host memory");
  P->n0 = n0;
  P->ratio = ratio;
  P->g = g;
  const std::vector<int> c = cuts(n0, tile0, g);
  for (size_t a = 0; a + 1 < c.size(); a++)
    for (size_t b = 0; b + 1 < c.size(); b++) P->boxes.push_back(Box{0, c[b], c[b + 1] - 1, c[a], c[a + 1] - 1});
  if (tags) cluster(tags, n0, tag_buffer, ratio, fine_max, fine_min, P->boxes);
  int rc = schedules(*P, (const Plan*)previous);
  if (rc) {
    delete P;
    return rc;
  }
  *plan = P;
  return 0;
}

int enl_plan_from_boxes(int n0, int ratio, int g, int npatch, const int* boxes, void** plan) {
  if (!plan || n0 <= 0 || ratio < 2 || g < 1 || npatch <= 0 || !boxes)
    return set_error(-2, "enl_plan_from_boxes: bad arguments");
  *plan = nullptr;
  Plan* P = new (std::nothrow) Plan();
  if (!P) return set_error(-5, "enl_plan_from_boxes: out of host memory");
  P->n0 = n0;
  P->ratio = ratio;
  P->g = g;
  for (int q = 0; q < npatch; q++) {
    const int* b = boxes + 5 * q;
    P->boxes.push_back(Box{b[0], b[1], b[2], b[3], b[4]});
  }
  int rc = schedules(*P, nullptr);
  if (rc) {
    delete P;
    return rc;
  }
  *plan = P;
  return 0;
}

int enl_plan_transfer(void* plan, void* old_plan, int n_old, const int* old_boxes) {
  Plan* P = (Plan*)plan;
  if (!P || (!old_plan && (n_old <= 0 || !old_boxes))) return set_error(-2, "enl_plan_transfer: bad arguments");
  Plan given;
  Plan& O = old_plan ? *(Plan*)old_plan : given;
  if (!old_plan) {
    given.n0 = P->n0;
    given.ratio = P->ratio;
    given.g = P->g;
    for (int q = 0; q < n_old; q++) {
      const int* b = old_boxes + 5 * q;
      given.boxes.push_back(Box{b[0], b[1], b[2], b[3], b[4]});
    }
  } else if (O.n0 != P->n0 || O.ratio != P->ratio || O.g != P->g) {
    return set_error(-2, "enl_plan_transfer: the plans differ in base grid, ratio or ghost width");
  }
  const int r = P->ratio, g = P->g;
  std::map<std::vector<int>, int> old_box;
  for (size_t q = 0; q < O.boxes.size(); q++) {
    const Box& b = O.boxes[q];
    old_box.emplace(std::vector<int>{b.level, b.i0, b.i1, b.j0, b.j1}, (int)q);
  }
  std::vector<int>&sn = P->ints["transfer.same.new"], &so = P->ints["transfer.same.old"];
  std::vector<int>&cdp = P->ints["transfer.copy.dp"], &cdl = P->ints["transfer.copy.dl"],
  &csp = P->ints["transfer.copy.sp"], &csl = P->ints["transfer.copy.sl"];
  std::vector<int>&idp = P->ints["transfer.interp.dp"], &idl = P->ints["transfer.interp.dl"],
  &isp = P->ints["transfer.interp.sp"], &isl = P->ints["transfer.interp.sl"];
  std::vector<double>&ifx = P->dbls["transfer.interp.fx"], &ify = P->dbls["transfer.interp.fy"];
  for (auto* v : {&sn, &so, &cdp, &cdl, &csp, &csl, &idp, &idl, &isp, &isl}) v->clear();
  ifx.clear();
  ify.clear();
  std::vector<int> rest;
  int maxlev = 0;
  for (size_t q = 0; q < P->boxes.size(); q++) {
    const Box& b = P->boxes[q];
    auto it = old_box.find(std::vector<int>{b.level, b.i0, b.i1, b.j0, b.j1});
    if (it == old_box.end()) {
      rest.push_back((int)q);
      maxlev = std::max(maxlev, b.level);
    } else {
      sn.push_back((int)q);
      so.push_back(it->second);
    }
  }
  const int oL = O.nlevels();
  std::vector<const Owner*> own((size_t)oL);
  for (int l = 0; l < oL; l++) own[(size_t)l] = &owner_map(O, l);
  const int rsh = log2_exact(r);
  for (int lev = 0; lev <= maxlev; lev++) {
    std::vector<int> pats;
    for (int q : rest)
      if (P->boxes[(size_t)q].level == lev) pats.push_back(q);
    if (pats.empty()) continue;
    const Owner* ow = lev < oL ? own[(size_t)lev] : nullptr;
    // copies of the whole level first, then its refinements; pass 1 counts
    // per patch, pass 2 writes at the patch's offsets
    const int np = (int)pats.size();
    std::vector<long> nc((size_t)np + 1, 0), ni((size_t)np + 1, 0);
    parallel_for(np, [&](int k) {
      const Box& b = P->boxes[(size_t)pats[(size_t)k]];
      long c = 0;
      if (ow)
        for (int j = b.j0; j <= b.j1; j++)
          for (int i = b.i0; i <= b.i1; i++) c += ow->at(i, j) >= 0;
      nc[(size_t)k + 1] = c;
      ni[(size_t)k + 1] = (long)b.nx() * b.ny() - c;
    });
    const long tc = prefix(nc), ti = prefix(ni);
    if (ti > 0 && (lev == 0 || lev - 1 >= oL))
      return set_error(-2, "enl_plan_transfer: new level-%d cells have no old source", lev);
    const size_t bc = cdp.size(), bi = idp.size();
    for (auto* v : {&cdp, &cdl, &csp, &csl}) v->resize(bc + (size_t)tc);
    for (auto* v : {&idp, &idl, &isp, &isl}) v->resize(bi + (size_t)ti);
    ifx.resize(bi + (size_t)ti);
    ify.resize(bi + (size_t)ti);
    const Owner* oc = ti > 0 ? own[(size_t)(lev - 1)] : nullptr;
    std::atomic<int> outside(0);
    parallel_for(np, [&](int k) {
      const int q = pats[(size_t)k];
      const Box& b = P->boxes[(size_t)q];
      size_t xc = bc + (size_t)nc[(size_t)k], xi = bi + (size_t)ni[(size_t)k];
      for (int j = b.j0; j <= b.j1; j++)
        for (int i = b.i0; i <= b.i1; i++) {
          const int dl = local(b, g, i, j);
          const int o = ow ? ow->at(i, j) : -1;
          if (o >= 0) {
            cdp[xc] = q;
            cdl[xc] = dl;
            csp[xc] = o;
            csl[xc] = local(O.boxes[(size_t)o], g, i, j);
            xc++;
            continue;
          }
          const int icc = rsh >= 0 ? i >> rsh : i / r, jcc = rsh >= 0 ? j >> rsh : j / r;
          const int c = oc->at(icc, jcc);
          if (c < 0) {
            outside.store(1);
            return;
          }
          idp[xi] = q;
          idl[xi] = dl;
          isp[xi] = c;
          isl[xi] = local(O.boxes[(size_t)c], g, icc, jcc);
          ifx[xi] = ((double)(i - icc * r) + 0.5) / r - 0.5;
          ify[xi] = ((double)(j - jcc * r) + 0.5) / r - 0.5;
          xi++;
        }
    });
    if (outside.load()) return set_error(-2, "enl_plan_transfer: a new fine cell lies outside the old coarser level");
  }
  return 0;
}

// The regrid transfer as rectangles ("transfer.rects"): the "same" pairs as
// enl_plan_transfer lists them, and every interior cell of the other new
// patches in rectangles of one source -- copies from the old patch of the
// same level holding the cells, refinements from the old coarser level
// elsewhere -- cut into blocks of at most 1024 cells.
int enl_plan_transfer_rects(void* plan, void* old_plan, int n_old, const int* old_boxes) {
  Plan* P = (Plan*)plan;
  if (!P || (!old_plan && (n_old <= 0 || !old_boxes))) return set_error(-2, "enl_plan_transfer_rects: bad arguments");
  Plan given;
  Plan& O = old_plan ? *(Plan*)old_plan : given;
  if (!old_plan) {
    given.n0 = P->n0;
    given.ratio = P->ratio;
    given.g = P->g;
    for (int q = 0; q < n_old; q++) {
      const int* b = old_boxes + 5 * q;
      given.boxes.push_back(Box{b[0], b[1], b[2], b[3], b[4]});
    }
  } else if (O.n0 != P->n0 || O.ratio != P->ratio || O.g != P->g) {
    return set_error(-2, "enl_plan_transfer_rects: the plans differ in base grid, ratio or ghost width");
  }
  const int r = P->ratio, g = P->g;
  std::map<std::vector<int>, int> old_box;
  for (size_t q = 0; q < O.boxes.size(); q++) {
    const Box& b = O.boxes[q];
    old_box.emplace(std::vector<int>{b.level, b.i0, b.i1, b.j0, b.j1}, (int)q);
  }
  std::vector<int>&sn = P->ints["transfer.same.new"], &so = P->ints["transfer.same.old"];
  sn.clear();
  so.clear();
  std::vector<int> rest;
  for (size_t q = 0; q < P->boxes.size(); q++) {
    const Box& b = P->boxes[q];
    auto it = old_box.find(std::vector<int>{b.level, b.i0, b.i1, b.j0, b.j1});
    if (it == old_box.end()) {
      rest.push_back((int)q);
    } else {
      sn.push_back((int)q);
      so.push_back(it->second);
    }
  }
  const int oL = O.nlevels();
  std::vector<const Owner*> own((size_t)oL);
  for (int l = 0; l < oL; l++) own[(size_t)l] = &owner_map(O, l);
  std::vector<std::vector<Rect>> per(rest.size());
  std::atomic<int> bad(0);
  parallel_for((int)rest.size(), [&](int k) {
    const int q = rest[(size_t)k];
    const Box& b = P->boxes[(size_t)q];
    const int lev = b.level;
    const Owner* ow = lev < oL ? own[(size_t)lev] : nullptr;
    const Owner* oc = lev > 0 && lev - 1 < oL ? own[(size_t)(lev - 1)] : nullptr;
    auto key = [&](int i, int j) -> long long {
      const int o = ow ? ow->at(i, j) : -1;
      if (o >= 0) return key_of(0, o);
      const int c = oc ? oc->at(i / r, j / r) : -1;
      if (c < 0) {
        bad.store(oc ? 2 : 1);
        return kSkip;
      }
      return key_of(1, c);
    };
    decompose(b.i0, b.i1, b.j0, b.j1, key, [&](int a, int bb, int j0, int j1, long long kk) {
      const int kind = (int)(kk >> 32), src = (int)(kk & 0xffffffffLL);
      const Box& S = O.boxes[(size_t)src];
      const int ni = bb - a + 1;
      const int band = std::max(1, 1024 / ni);  // rows per block
      for (int jb = j0; jb <= j1; jb += band) {
        const int je = std::min(j1, jb + band - 1);
        Rect R{q, src, local(b, g, a, jb), 0, ni, je - jb + 1, kind, 0};
        if (kind == 0) {
          R.sl = local(S, g, a, jb);
        } else {
          const int ic = a / r, jc = jb / r;
          R.sl = local(S, g, ic, jc);
          R.phase = (a - ic * r) | ((jb - jc * r) << 8);
        }
        per[(size_t)k].push_back(R);
      }
    });
  });
  if (bad.load() == 1) return set_error(-2, "enl_plan_transfer_rects: new cells have no old source");
  if (bad.load() == 2)
    return set_error(-2, "enl_plan_transfer_rects: a new fine cell lies outside the old coarser level");
  std::vector<int>& out = P->ints["transfer.rects"];
  out.clear();
  for (const auto& v : per) put_rects(out, v);
  return 0;
}

int enl_plan_get(const void* plan, const char* name, int level, const void** data, long* count) {
  Plan* P = (Plan*)plan;
  if (!P || !name || !data || !count) return set_error(-2, "enl_plan_get: bad arguments");
  *data = nullptr;
  *count = 0;
  // per-cell fill and coarsen lists are built on first request
  if (!strncmp(name, "fill.", 5) && strncmp(name, "fill.rects", 10)) {
    int rc = ensure_cells(*P, "fill", level);
    if (rc) return rc;
  } else if (!strncmp(name, "coarsen.", 8) && strcmp(name, "coarsen.rects")) {
    int rc = ensure_cells(*P, "coarsen", 0);
    if (rc) return rc;
  }
  if (!strcmp(name, "boxes")) {
    static_assert(sizeof(Box) == 5 * sizeof(int), "Box is five ints");
    *data = P->boxes.data();
    *count = (long)P->boxes.size() * 5;
    return 0;
  }
  const std::string key = strncmp(name, "fill.", 5) == 0 ? lv(name, level) : std::string(name);
  auto it = P->ints.find(key);
  if (it != P->ints.end()) {
    *data = it->second.data();
    *count = (long)it->second.size();
    return 0;
  }
  auto dt = P->dbls.find(key);
  if (dt != P->dbls.end()) {
    *data = dt->second.data();
    *count = (long)dt->second.size();
    return 1;  // doubles
  }
  return 0;  // an empty list
}

int enl_plan_free(void* plan) {
  delete (Plan*)plan;
  return 0;
}

}  // extern "C"
