"""Two-level AMR hierarchy and its ghost-fill schedule, built on the host.

This stands in for the parts of SAMRAI the reference relies on around the
per-patch kernels (not vendored, so none of it is parity-pinned):

* level 0: the base grid cut into patches of at most `tile0` cells a side
  (input.txt:114-116, largest_patch_size level_0 = 64); a remainder
  narrower than the ghost width joins the tile before it;
* level 1: refine ratio r; tags from tag_cells on level 0, buffered by
  `tag_buffer` cells (input.txt:8), clustered into boxes whose fine sides lie
  in [8, 64]: the domain is cut into blocks of 64/r coarse cells and each
  block's tagged cells become its bounding box (grown to 8/r coarse cells
  inside the block; a last block narrower than that keeps its width);
  boxes within one coarse cell of the domain edge are extended to touch it
  when their block reaches the edge, as SAMRAI's gridding does so that
  ghost strips never straddle the edge (boxes never leave their block, so
  they never overlap; with the ghost width at most r, a one-coarse-cell gap
  to the edge holds a whole corner block);
* ghost fill for every patch, per level, in three passes:
    1. ghost cells inside the domain: copy from the patch of the same level
       that holds the cell, else linear refinement of level-0 data (cell-
       centred slopes from the neighbouring coarse cells, unlimited: the
       LINEAR_REFINE operator registered at enlightning.C:259-281);
    2. edge strips outside the domain: the reference's physical boundary
       (enlightning.C:836-944), a copy of the nearest interior cell with the
       normal momentum negated at REFLECT walls;
    3. corner blocks outside the domain: copy along x from the patch's first
       or last interior column, negating x momentum at a REFLECT left/right
       wall (enlightning.C:948-1008), after passes 1 and 2;
* the fine-to-coarse average of w_n, p, T after each step
  (CONSERVATIVE_COARSEN, enlightning.C:282-288);
* at a regrid (gridding_algorithm->regridAllFinerLevels, main.C:489-502) the
  transfer of the SOLN variables (w_n, p, T, source) into the new
  hierarchy's patch interiors: from the old patch of the same level that
  held the cell, else by the same linear refinement from the old coarser
  level (make_transfer).

All cell addresses are (patch id, local index) pairs with local index
(i - ifirst0 + g) + Nx * (j - ifirst1 + g) inside the patch's ghosted box.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Patch:
    pid: int
    level: int
    box: tuple  # ifirst0, ilast0, ifirst1, ilast1 in the level's index space
    g: int

    @property
    def nx(self):
        return self.box[1] - self.box[0] + 1

    @property
    def ny(self):
        return self.box[3] - self.box[2] + 1

    @property
    def Nx(self):
        return self.nx + 2 * self.g

    @property
    def Ny(self):
        return self.ny + 2 * self.g

    def local(self, i, j):
        return (i - self.box[0] + self.g) + self.Nx * (j - self.box[2] + self.g)


@dataclass
class Hierarchy:
    n0: int
    ratio: int
    g: int
    L: float
    patches: list = field(default_factory=list)
    # ghost-fill schedule per level: dicts of int32/float64 numpy arrays
    fill: list = field(default_factory=list)
    coarsen: dict = None
    # owner map per level (level index space -> pid or -1), built on demand
    owners: dict = field(default_factory=dict, repr=False, compare=False)
    # the native plan (enl_plan_*) whose lists fill and coarsen view, if any
    plan: object = field(default=None, repr=False, compare=False)

    @property
    def nlevels(self):
        return 1 + max(p.level for p in self.patches)

    def dx(self, level):
        return self.L / (self.n0 * self.ratio ** level)

    def ncells(self, level):
        return self.n0 * self.ratio ** level

    def level_patches(self, level):
        return [p for p in self.patches if p.level == level]

    def interior_cells(self, level=None):
        return int(sum(p.nx * p.ny for p in self.patches if level is None or p.level == level))


def _cuts(n0, tile0, g):
    """Tile starts along one axis; a last tile narrower than the ghost width
    g joins the one before it (a patch within g cells of the domain edge
    but not on it would have corner ghost blocks partly outside the domain)."""
    starts = list(range(0, n0, tile0))
    if len(starts) > 1 and n0 - starts[-1] < g:
        starts.pop()
    return starts + [n0]


def base_boxes(n0, tile0, g=0):
    cuts = _cuts(n0, tile0, g)
    boxes = []
    for j0, j1 in zip(cuts[:-1], cuts[1:]):
        for i0, i1 in zip(cuts[:-1], cuts[1:]):
            boxes.append((i0, i1 - 1, j0, j1 - 1))
    return boxes


def buffer_tags(tags, nbuf):
    """Grow the tagged set by nbuf cells (a square neighbourhood: a dilation
    along rows, then along columns)."""
    out = tags.astype(bool)
    for axis in (1, 0):
        src = out.copy()
        for d in range(1, nbuf + 1):
            if axis == 1:
                out[:, d:] |= src[:, :-d]
                out[:, :-d] |= src[:, d:]
            else:
                out[d:, :] |= src[:-d, :]
                out[:-d, :] |= src[d:, :]
    return out


def cluster_fine_boxes(tags, ratio, fine_max=64, fine_min=8):
    """Coarse tag map (n0 x n0, [j, i]) -> fine boxes (fine index space)."""
    n0 = tags.shape[0]
    blk = fine_max // ratio
    cmin = max(1, fine_min // ratio)
    # per block of blk x blk coarse cells the bounding box of its tags,
    # all blocks at once: t[bj, jj, bi, ii]
    nb = -(-n0 // blk)
    t = np.zeros((nb * blk, nb * blk), bool)
    t[:n0, :n0] = tags != 0
    t = t.reshape(nb, blk, nb, blk)
    cols = t.any(axis=1)  # [bj, bi, ii]
    rows = t.any(axis=3)  # [bj, jj, bi]
    ilo, ihi = cols.argmax(axis=2), blk - 1 - cols[:, :, ::-1].argmax(axis=2)
    jlo, jhi = rows.argmax(axis=1), blk - 1 - rows[:, ::-1, :].argmax(axis=1)
    boxes = []
    for kj, ki in zip(*np.nonzero(cols.any(axis=2))):
        bj, bi = int(kj) * blk, int(ki) * blk
        i0, i1 = bi + int(ilo[kj, ki]), bi + int(ihi[kj, ki])
        j0, j1 = bj + int(jlo[kj, ki]), bj + int(jhi[kj, ki])
        # at least cmin coarse cells a side where the block allows it; a box
        # never leaves its block, so boxes of different blocks cannot overlap
        # (the last block of a row is narrower than cmin when n0 % blk is
        # smaller: its box is then the block's width)
        bi1, bj1 = min(bi + blk, n0) - 1, min(bj + blk, n0) - 1
        while i1 - i0 + 1 < cmin and (i1 < bi1 or i0 > bi):
            if i1 < bi1:
                i1 += 1
            else:
                i0 -= 1
        while j1 - j0 + 1 < cmin and (j1 < bj1 or j0 > bj):
            if j1 < bj1:
                j1 += 1
            else:
                j0 -= 1
        # touch the domain edge when within one coarse cell of it, where the
        # block edge is the domain edge
        if i0 == 1 and bi == 0:
            i0 = 0
        if j0 == 1 and bj == 0:
            j0 = 0
        if i1 == n0 - 2 and bi1 == n0 - 1:
            i1 = n0 - 1
        if j1 == n0 - 2 and bj1 == n0 - 1:
            j1 = n0 - 1
        boxes.append((i0 * ratio, (i1 + 1) * ratio - 1, j0 * ratio, (j1 + 1) * ratio - 1))
    return boxes


def _owner(h, level):
    """The level's owner map, kept on the hierarchy: (map, s), the patch
    holding cell (i, j) (or -1) at map[j // s, i // s].  s is the refinement
    ratio when every box of a level above 0 is aligned to it (the clustered
    boxes are), which keeps the map at the coarser level's size, else 1."""
    if level not in h.owners:
        pats = h.level_patches(level)
        n, r = h.ncells(level), h.ratio
        aligned = level > 0 and n % r == 0 and all(
            p.box[0] % r == 0 and (p.box[1] + 1) % r == 0 and p.box[2] % r == 0 and (p.box[3] + 1) % r == 0
            for p in pats)
        s = r if aligned else 1
        own = np.full((n // s, n // s), -1, np.int32)
        for p in pats:
            own[p.box[2] // s:(p.box[3] + 1) // s, p.box[0] // s:(p.box[1] + 1) // s] = p.pid
        h.owners[level] = (own, s)
    return h.owners[level]


def _owner_at(h, level, i, j):
    """Owners of the cells (i, j) of a level (index arrays inside it)."""
    own, s = _owner(h, level)
    return own[j, i] if s == 1 else own[j // s, i // s]


class NativePlan:
    """A plan of the library's regrid planner (enl_plan_*, enl_regrid.hip):
    the same boxes and lists as the numpy builders here, built in C++.
    Arrays taken from it view its memory and keep it alive."""

    def __init__(self, handle):
        from . import _lib
        self.lib = _lib.load()
        self.handle = handle

    def get(self, name, level=0, copy=False):
        from . import _lib
        data, count = ctypes.c_void_p(), ctypes.c_long()
        kind = self.lib.enl_plan_get(self.handle, name.encode(), int(level), ctypes.byref(data), ctypes.byref(count))
        _lib.check(0 if kind in (0, 1) else kind)
        ct = ctypes.c_double if kind == 1 else ctypes.c_int32
        if count.value == 0:
            return np.zeros(0, np.float64 if kind == 1 else np.int32)
        buf = (ct * count.value).from_address(data.value)
        buf._owner = self  # the plan outlives the view
        a = np.ctypeslib.as_array(buf)
        return a.copy() if copy else a

    def fill(self, level):
        f = lambda n: self.get("fill." + n, level)  # noqa: E731
        return {"copy": tuple(f("copy." + k) for k in ("p", "dst", "sp", "sl")),
                "interp": tuple(f("interp." + k) for k in ("p", "dst", "cp", "cl", "fx", "fy")),
                "edge": tuple(f("edge." + k) for k in ("p", "dst", "src", "neg")),
                "corner": tuple(f("corner." + k) for k in ("p", "dst", "src", "neg"))}

    def coarsen(self):
        return tuple(self.get("coarsen." + k) for k in ("cp", "cl", "fp", "fl"))

    def fill_rects(self, level):
        """(rects (n, 8) int32 enl_rect_t rows, n1, n2): the level's ghost-fill
        schedule as rectangles, n1 first-pass rows then n2 corner rows."""
        n1, n2 = (int(v) for v in self.get("fill.rects_n", level))
        return self.get("fill.rects", level).reshape(-1, 8), n1, n2

    def coarsen_rects(self):
        return self.get("coarsen.rects").reshape(-1, 8)

    def __del__(self):
        try:
            if self.handle:
                self.lib.enl_plan_free(self.handle)
                self.handle = None
        except Exception:
            pass


def _boxes_array(h):
    return np.ascontiguousarray([[p.level] + list(p.box) for p in h.patches], dtype=np.int32).reshape(-1, 5)


class _LazyFill:
    """The per-level fill lists of a native plan, built (by the plan) on
    first access; a level may be lent by another hierarchy's lists."""

    def __init__(self, plan, nlevels):
        self._plan, self._n, self._have, self._lent = plan, nlevels, {}, {}

    def lend(self, level, other):
        self._lent[level] = other

    def __len__(self):
        return self._n

    def __getitem__(self, level):
        if level < 0:
            level += self._n
        if not 0 <= level < self._n:
            raise IndexError(level)
        if level in self._lent:
            return self._lent[level][level]
        if level not in self._have:
            self._have[level] = self._plan.fill(level)
        return self._have[level]

    def __iter__(self):
        return (self[lev] for lev in range(self._n))


class _LazyCoarsen:
    """The per-cell coarsen lists (cp, cl, fp, fl) of a native plan, built on
    first access."""

    def __init__(self, plan):
        self._plan, self._lists = plan, None

    def _get(self):
        if self._lists is None:
            self._lists = self._plan.coarsen()
        return self._lists

    def __len__(self):
        return 4

    def __getitem__(self, k):
        return self._get()[k]

    def __iter__(self):
        return iter(self._get())


def _from_plan(plan, n0, ratio, g, L):
    h = Hierarchy(n0=n0, ratio=ratio, g=g, L=L)
    rows = plan.get("boxes").reshape(-1, 5).tolist()
    h.patches = [Patch(k, row[0], tuple(row[1:]), g) for k, row in enumerate(rows)]
    h.fill = _LazyFill(plan, h.nlevels)
    h.coarsen = _LazyCoarsen(plan) if h.nlevels > 1 else None
    h.plan = plan
    return h


def build(n0, ratio, g, L, coarse_tags, tile0=64, tag_buffer=2, fine_max=64, fine_min=8, previous=None, native=True):
    """Hierarchy with level 0 = base patches and, if any cell is tagged, a
    level-1 of clustered boxes; with its ghost-fill and coarsen schedules.
    At a regrid, `previous` (the hierarchy being replaced) lends its level-0
    fill schedule when its base patches are the same.  native: the library's
    planner (enl_plan_build) builds it; native=False: the numpy builders
    below, which state the same lists (the tests hold the two equal)."""
    if native:
        from . import _lib
        lib = _lib.load()
        tags = None if coarse_tags is None else np.ascontiguousarray(coarse_tags, dtype=np.int32)
        prev = previous.plan.handle if previous is not None and previous.plan is not None else None
        handle = ctypes.c_void_p()
        _lib.check(lib.enl_plan_build(n0, ratio, g, tile0, None if tags is None else ctypes.c_void_p(tags.ctypes.data),
                                      tag_buffer, fine_max, fine_min, prev, ctypes.byref(handle)))
        h = _from_plan(NativePlan(handle), n0, ratio, g, L)
        if (previous is not None and (previous.n0, previous.ratio, previous.g) == (n0, ratio, g)
                and [p.box for p in previous.level_patches(0)] == [p.box for p in h.level_patches(0)]
                and [p.pid for p in previous.level_patches(0)] == [p.pid for p in h.level_patches(0)]):
            if isinstance(h.fill, _LazyFill):
                h.fill.lend(0, previous.fill)  # the same lists: lent, as the numpy builder lends them
            else:
                h.fill[0] = previous.fill[0]
            if 0 in previous.owners:
                h.owners[0] = previous.owners[0]
        return h
    h = Hierarchy(n0=n0, ratio=ratio, g=g, L=L)
    for b in base_boxes(n0, tile0, g):
        h.patches.append(Patch(len(h.patches), 0, b, g))
    if coarse_tags is not None and np.any(coarse_tags):
        tb = buffer_tags(coarse_tags, tag_buffer)
        for b in cluster_fine_boxes(tb, ratio, fine_max, fine_min):
            h.patches.append(Patch(len(h.patches), 1, b, g))
    return _schedules(h, previous)


def from_boxes(n0, ratio, g, L, boxes, native=True):
    """Hierarchy from (level, ifirst0, ilast0, ifirst1, ilast1) rows in pid
    order (a restart's boxes), with its schedules."""
    if native:
        from . import _lib
        lib = _lib.load()
        b = np.ascontiguousarray(boxes, dtype=np.int32).reshape(-1, 5)
        handle = ctypes.c_void_p()
        _lib.check(lib.enl_plan_from_boxes(n0, ratio, g, len(b), ctypes.c_void_p(b.ctypes.data),
                                           ctypes.byref(handle)))
        return _from_plan(NativePlan(handle), n0, ratio, g, L)
    h = Hierarchy(n0=n0, ratio=ratio, g=g, L=L)
    for row in boxes:
        lev, b = int(row[0]), tuple(int(v) for v in row[1:5])
        h.patches.append(Patch(len(h.patches), lev, b, g))
    return _schedules(h, None)


def _schedules(h, previous):
    same_base = (previous is not None and (previous.n0, previous.ratio, previous.g) == (h.n0, h.ratio, h.g)
                 and [p.box for p in previous.level_patches(0)] == [p.box for p in h.level_patches(0)]
                 and [p.pid for p in previous.level_patches(0)] == [p.pid for p in h.level_patches(0)])
    if same_base and 0 in previous.owners:
        h.owners[0] = previous.owners[0]
    h.fill = [previous.fill[0] if (lev == 0 and same_base) else make_fill(h, lev) for lev in range(h.nlevels)]
    h.coarsen = make_coarsen(h) if h.nlevels > 1 else None
    return h


def _ghost_cells(pats, g):
    """Ghost cells of every patch of a level as flat arrays (k, i, j), k the
    index into pats: per patch the rows below and above the box (full
    ghosted width), then the columns left and right of it (box rows)."""
    ks, I, J = [], [], []
    for k, p in enumerate(pats):
        i0, i1, j0, j1 = p.box
        ci = np.arange(i0 - g, i1 + g + 1)
        rows = np.concatenate([np.arange(j0 - g, j0), np.arange(j1 + 1, j1 + g + 1)])
        cols = np.concatenate([np.arange(i0 - g, i0), np.arange(i1 + 1, i1 + g + 1)])
        cj = np.arange(j0, j1 + 1)
        I.append(np.tile(ci, rows.size))
        J.append(np.repeat(rows, ci.size))
        I.append(np.tile(cols, cj.size))
        J.append(np.repeat(cj, cols.size))
        ks.append(2 * g * (ci.size + cj.size))
    if not ks:
        z = np.zeros(0, np.int64)
        return z, z, z
    return np.repeat(np.arange(len(ks)), ks), np.concatenate(I), np.concatenate(J)


def _locals(box0, box2, Nx, g, i, j):
    return ((i - box0 + g) + Nx * (j - box2 + g)).astype(np.int32)


def make_fill(h, level):
    """Ghost-fill schedule of one level (see the module docstring).  Items
    of one kind are independent (distinct destinations, interior sources),
    so their order is free: here patch by patch, strip by strip."""
    n = h.ncells(level)
    g = h.g
    pats = h.level_patches(level)
    # per global pid: box origin and ghosted row length
    B0 = np.array([p.box[0] for p in h.patches], np.int64)
    B2 = np.array([p.box[2] for p in h.patches], np.int64)
    NX = np.array([p.Nx for p in h.patches], np.int64)
    pid = np.array([p.pid for p in pats], np.int64)
    r = h.ratio
    K, I, J = _ghost_cells(pats, g)
    inside = (I >= 0) & (I < n) & (J >= 0) & (J < n)
    K, I, J = K[inside], I[inside], J[inside]
    dpid = pid[K]
    dst = _locals(B0[dpid], B2[dpid], NX[dpid], g, I, J)
    owner = _owner_at(h, level, I, J).astype(np.int64)
    have = owner >= 0
    o = owner[have]
    copy = (dpid[have].astype(np.int32), dst[have], o.astype(np.int32),
            _locals(B0[o], B2[o], NX[o], g, I[have], J[have]))
    ui, uj, up, ud = I[~have], J[~have], dpid[~have], dst[~have]
    if ui.size:
        assert level > 0, "a level-0 ghost cell inside the domain has no owner"
        ic, jc = ui // r, uj // r
        c = _owner_at(h, level - 1, ic, jc).astype(np.int64)
        assert (c >= 0).all()
        interp = (up.astype(np.int32), ud, c.astype(np.int32), _locals(B0[c], B2[c], NX[c], g, ic, jc),
                  ((ui - ic * r) + 0.5) / r - 0.5, ((uj - jc * r) + 0.5) / r - 0.5)
    else:
        interp = tuple(np.zeros(0, np.int32) for _ in range(4)) + (np.zeros(0), np.zeros(0))
    edge_p, edge_dst, edge_src, edge_neg = [], [], [], []
    corn_p, corn_dst, corn_src, corn_neg = [], [], [], []
    for p in pats:
        i0, i1, j0, j1 = p.box
        touch = (i0 == 0, i1 == n - 1, j0 == 0, j1 == n - 1)
        if not any(touch):
            continue
        # physical boundary: edge strips over the patch's own tangential range
        if touch[0]:
            jj, ll = np.meshgrid(np.arange(j0, j1 + 1), np.arange(1, g + 1))
            edge_p.append(np.full(jj.size, p.pid, np.int32))
            edge_dst.append(p.local(i0 - ll, jj).ravel().astype(np.int32))
            edge_src.append(p.local(np.full_like(jj, i0), jj).ravel().astype(np.int32))
            edge_neg.append(np.full(jj.size, 1, np.int32))
        if touch[1]:
            jj, ll = np.meshgrid(np.arange(j0, j1 + 1), np.arange(1, g + 1))
            edge_p.append(np.full(jj.size, p.pid, np.int32))
            edge_dst.append(p.local(i1 + ll, jj).ravel().astype(np.int32))
            edge_src.append(p.local(np.full_like(jj, i1), jj).ravel().astype(np.int32))
            edge_neg.append(np.full(jj.size, 2, np.int32))
        if touch[2]:
            ii, ll = np.meshgrid(np.arange(i0, i1 + 1), np.arange(1, g + 1))
            edge_p.append(np.full(ii.size, p.pid, np.int32))
            edge_dst.append(p.local(ii, j0 - ll).ravel().astype(np.int32))
            edge_src.append(p.local(ii, np.full_like(ii, j0)).ravel().astype(np.int32))
            edge_neg.append(np.full(ii.size, 3, np.int32))
        if touch[3]:
            ii, ll = np.meshgrid(np.arange(i0, i1 + 1), np.arange(1, g + 1))
            edge_p.append(np.full(ii.size, p.pid, np.int32))
            edge_dst.append(p.local(ii, j1 + ll).ravel().astype(np.int32))
            edge_src.append(p.local(ii, np.full_like(ii, j1)).ravel().astype(np.int32))
            edge_neg.append(np.full(ii.size, 4, np.int32))
        # corner blocks that lie outside the domain
        for right in (0, 1):
            for upper in (0, 1):
                ci = np.arange(i1 + 1, i1 + g + 1) if right else np.arange(i0 - g, i0)
                cj = np.arange(j1 + 1, j1 + g + 1) if upper else np.arange(j0 - g, j0)
                CI, CJ = np.meshgrid(ci, cj)
                outside = (CI < 0) | (CI >= n) | (CJ < 0) | (CJ >= n)
                if not outside.any():
                    continue
                assert outside.all(), "corner block straddles the domain edge"
                col = i1 if right else i0
                corn_p.append(np.full(CI.size, p.pid, np.int32))
                corn_dst.append(p.local(CI, CJ).ravel().astype(np.int32))
                corn_src.append(p.local(np.full_like(CI, col), CJ).ravel().astype(np.int32))
                corn_neg.append(np.full(CI.size, 2 if right else 1, np.int32))

    def cat(lst, dtype):
        return np.concatenate(lst).astype(dtype) if lst else np.zeros(0, dtype)

    return {
        "copy": copy,
        "interp": interp,
        # neg: side code 1..4 = left, right, bottom, top (which wall's REFLECT
        # flag applies and which momentum it negates)
        "edge": (cat(edge_p, np.int32), cat(edge_dst, np.int32), cat(edge_src, np.int32), cat(edge_neg, np.int32)),
        "corner": (cat(corn_p, np.int32), cat(corn_dst, np.int32), cat(corn_src, np.int32),
                   cat(corn_neg, np.int32)),
    }


def _interior_cells(pats):
    """Interior cells of every patch as flat arrays (k, i, j), k the index
    into pats, row by row."""
    ks, I, J = [], [], []
    for k, p in enumerate(pats):
        ci, cj = np.arange(p.box[0], p.box[1] + 1), np.arange(p.box[2], p.box[3] + 1)
        I.append(np.tile(ci, cj.size))
        J.append(np.repeat(cj, ci.size))
        ks.append(ci.size * cj.size)
    if not ks:
        z = np.zeros(0, np.int64)
        return z, z, z
    return np.repeat(np.arange(len(ks)), ks), np.concatenate(I), np.concatenate(J)


def _patch_arrays(h):
    """Per global pid: box origin and ghosted row length."""
    return (np.array([p.box[0] for p in h.patches], np.int64), np.array([p.box[2] for p in h.patches], np.int64),
            np.array([p.Nx for p in h.patches], np.int64))


def make_coarsen(h):
    """(coarse patch, coarse local, fine patch, fine local of the lower-left
    fine cell) for every coarse cell under a fine patch."""
    r = h.ratio
    B0, B2, NX = _patch_arrays(h)
    coarse_boxes = []
    fine = h.level_patches(1)
    for f in fine:
        i0, i1, j0, j1 = f.box
        coarse_boxes.append(Patch(f.pid, 0, (i0 // r, (i1 + 1) // r - 1, j0 // r, (j1 + 1) // r - 1), f.g))
    K, IC, JC = _interior_cells(coarse_boxes)
    fp = np.array([f.pid for f in fine], np.int64)[K]
    cp = _owner_at(h, 0, IC, JC).astype(np.int64)
    g = h.g
    return (cp.astype(np.int32), _locals(B0[cp], B2[cp], NX[cp], g, IC, JC), fp.astype(np.int32),
            _locals(B0[fp], B2[fp], NX[fp], g, IC * r, JC * r))


def transfer_rects(new, old):
    """The regrid transfer from hierarchy `old` to `new` as the library's
    planner states it in rectangles (enl_plan_transfer_rects): the "same"
    pairs (new pid, old pid) and an (n, 8) int32 array of enl_rect_t rows
    (copies from old patches of the same level, refinements from the old
    coarser level) for enl_batch_transfer_rects.  Both need native plans."""
    if (new.n0, new.ratio, new.g) != (old.n0, old.ratio, old.g):
        raise ValueError("transfer_rects: the hierarchies differ in base grid, ratio or ghost width")
    if new.plan is None:
        raise ValueError("transfer_rects: the new hierarchy has no native plan")
    from . import _lib
    ob = _boxes_array(old)
    rc = new.plan.lib.enl_plan_transfer_rects(new.plan.handle, old.plan.handle if old.plan is not None else None,
                                              len(ob), ctypes.c_void_p(ob.ctypes.data))
    if rc:
        raise ValueError("transfer_rects: %s" % _lib.last_error()[1])
    t = lambda n: new.plan.get("transfer." + n, copy=True)  # noqa: E731
    return {"same": (t("same.new"), t("same.old")), "rects": t("rects").reshape(-1, 8)}


def make_transfer(new, old, native=True):
    """Regrid transfer schedule from hierarchy `old` to `new` (same base
    grid, ratio and ghost width).  A new patch with the level and box of an
    old one is a "same" pair (new pid, old pid), copied whole; for every
    interior cell of the other new patches, either a copy item (new patch,
    new local) <- (old patch of the same level holding the cell, its local),
    or, where no old patch of that level held it, an interp item from the
    old coarser level's cell with the fine centre's offset (fx, fy) in
    coarse widths.  Ghosts are not listed: the new hierarchy's own ghost
    fill takes them afterwards."""
    if (new.n0, new.ratio, new.g) != (old.n0, old.ratio, old.g):
        raise ValueError("make_transfer: the hierarchies differ in base grid, ratio or ghost width")
    if native and new.plan is not None:
        from . import _lib
        ob = _boxes_array(old)
        rc = new.plan.lib.enl_plan_transfer(new.plan.handle, old.plan.handle if old.plan is not None else None,
                                            len(ob), ctypes.c_void_p(ob.ctypes.data))
        if rc:
            raise ValueError("make_transfer: %s" % _lib.last_error()[1])
        t = lambda n: new.plan.get("transfer." + n, copy=True)  # noqa: E731
        return {"same": (t("same.new"), t("same.old")),
                "copy": tuple(t("copy." + k) for k in ("dp", "dl", "sp", "sl")),
                "interp": tuple(t("interp." + k) for k in ("dp", "dl", "sp", "sl", "fx", "fy"))}
    r, g = new.ratio, new.g
    old_box = {(p.level, p.box): p.pid for p in old.patches}
    s_new, s_old, rest = [], [], []
    for p in new.patches:
        q = old_box.get((p.level, p.box))
        if q is None:
            rest.append(p)
        else:
            s_new.append(p.pid)
            s_old.append(q)
    nB0, nB2, nNX = _patch_arrays(new)
    oB0, oB2, oNX = _patch_arrays(old)
    copy = [[], [], [], []]
    interp = [[], [], [], [], [], []]
    for lev in sorted({p.level for p in rest}):
        pats = [p for p in rest if p.level == lev]
        K, I, J = _interior_cells(pats)
        dp = np.array([p.pid for p in pats], np.int64)[K]
        dl = _locals(nB0[dp], nB2[dp], nNX[dp], g, I, J)
        if lev < old.nlevels:
            owner = _owner_at(old, lev, I, J).astype(np.int64)
        else:
            owner = np.full(I.size, -1, np.int64)
        have = owner >= 0
        o = owner[have]
        for lst, a in zip(copy, (dp[have], dl[have], o, _locals(oB0[o], oB2[o], oNX[o], g, I[have], J[have]))):
            lst.append(a)
        if (~have).any():
            if lev == 0 or lev - 1 >= old.nlevels:
                raise ValueError("make_transfer: new level-%d cells have no old source" % lev)
            ui, uj = I[~have], J[~have]
            ic, jc = ui // r, uj // r
            c = _owner_at(old, lev - 1, ic, jc).astype(np.int64)
            if (c < 0).any():
                raise ValueError("make_transfer: a new fine cell lies outside the old coarser level")
            for lst, a in zip(interp, (dp[~have], dl[~have], c, _locals(oB0[c], oB2[c], oNX[c], g, ic, jc),
                                       ((ui - ic * r) + 0.5) / r - 0.5, ((uj - jc * r) + 0.5) / r - 0.5)):
                lst.append(a)

    def cat(lst, dtype):
        return np.concatenate(lst).astype(dtype) if lst else np.zeros(0, dtype)

    return {
        "same": (np.array(s_new, np.int32), np.array(s_old, np.int32)),
        "copy": tuple(cat(x, np.int32) for x in copy),
        "interp": tuple(cat(x, np.int32) for x in interp[:4]) + tuple(cat(x, np.float64) for x in interp[4:]),
    }
