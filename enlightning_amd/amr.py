"""Two-level AMR hierarchy and its ghost-fill schedule, built on the host.

This stands in for the parts of SAMRAI the reference relies on around the
per-patch kernels (not vendored, so none of it is parity-pinned):

* level 0: the base grid cut into patches of at most `tile0` cells a side
  (input.txt:114-116, largest_patch_size level_0 = 64);
* level 1: refine ratio r; tags from tag_cells on level 0, buffered by
  `tag_buffer` cells (input.txt:8), clustered into boxes whose fine sides lie
  in [8, 64]: the domain is cut into blocks of 64/r coarse cells and each
  block's tagged cells become its bounding box (grown to 8/r coarse cells);
  boxes within one coarse cell of the domain edge are extended to touch it,
  as SAMRAI's gridding does so that ghost strips never straddle the edge;
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


def base_boxes(n0, tile0):
    boxes = []
    for j0 in range(0, n0, tile0):
        for i0 in range(0, n0, tile0):
            boxes.append((i0, min(i0 + tile0, n0) - 1, j0, min(j0 + tile0, n0) - 1))
    return boxes


def buffer_tags(tags, nbuf):
    """Grow the tagged set by nbuf cells (a square neighbourhood)."""
    t = tags.astype(bool)
    out = t.copy()
    for dj in range(-nbuf, nbuf + 1):
        for di in range(-nbuf, nbuf + 1):
            if di == 0 and dj == 0:
                continue
            sh = np.zeros_like(t)
            ys = slice(max(dj, 0), t.shape[0] + min(dj, 0))
            yd = slice(max(-dj, 0), t.shape[0] + min(-dj, 0))
            xs = slice(max(di, 0), t.shape[1] + min(di, 0))
            xd = slice(max(-di, 0), t.shape[1] + min(-di, 0))
            sh[yd, xd] = t[ys, xs]
            out |= sh
    return out


def cluster_fine_boxes(tags, ratio, fine_max=64, fine_min=8):
    """Coarse tag map (n0 x n0, [j, i]) -> fine boxes (fine index space)."""
    n0 = tags.shape[0]
    blk = fine_max // ratio
    cmin = max(1, fine_min // ratio)
    boxes = []
    for bj in range(0, n0, blk):
        for bi in range(0, n0, blk):
            sub = tags[bj:bj + blk, bi:bi + blk]
            if not sub.any():
                continue
            jj, ii = np.nonzero(sub)
            i0, i1 = bi + ii.min(), bi + ii.max()
            j0, j1 = bj + jj.min(), bj + jj.max()
            # at least cmin coarse cells a side, kept inside the block
            bi1, bj1 = min(bi + blk, n0) - 1, min(bj + blk, n0) - 1
            while i1 - i0 + 1 < cmin:
                if i1 < bi1:
                    i1 += 1
                else:
                    i0 -= 1
            while j1 - j0 + 1 < cmin:
                if j1 < bj1:
                    j1 += 1
                else:
                    j0 -= 1
            # touch the domain edge when within one coarse cell of it (the
            # block edge coincides with the domain edge there)
            if i0 == 1:
                i0 = 0
            if j0 == 1:
                j0 = 0
            if i1 == n0 - 2:
                i1 = n0 - 1
            if j1 == n0 - 2:
                j1 = n0 - 1
            boxes.append((i0 * ratio, (i1 + 1) * ratio - 1, j0 * ratio, (j1 + 1) * ratio - 1))
    return boxes


def _owner_map(patches, n):
    own = np.full((n, n), -1, np.int32)
    for p in patches:
        own[p.box[2]:p.box[3] + 1, p.box[0]:p.box[1] + 1] = p.pid
    return own


def build(n0, ratio, g, L, coarse_tags, tile0=64, tag_buffer=2, fine_max=64, fine_min=8):
    """Hierarchy with level 0 = base patches and, if any cell is tagged, a
    level-1 of clustered boxes; with its ghost-fill and coarsen schedules."""
    h = Hierarchy(n0=n0, ratio=ratio, g=g, L=L)
    for b in base_boxes(n0, tile0):
        h.patches.append(Patch(len(h.patches), 0, b, g))
    if coarse_tags is not None and np.any(coarse_tags):
        tb = buffer_tags(coarse_tags, tag_buffer)
        for b in cluster_fine_boxes(tb, ratio, fine_max, fine_min):
            h.patches.append(Patch(len(h.patches), 1, b, g))
    h.fill = [make_fill(h, lev) for lev in range(h.nlevels)]
    h.coarsen = make_coarsen(h) if h.nlevels > 1 else None
    return h


def make_fill(h, level):
    """Ghost-fill schedule of one level (see the module docstring)."""
    n = h.ncells(level)
    pats = h.level_patches(level)
    own = _owner_map(pats, n)
    byid = {p.pid: p for p in h.patches}
    if level > 0:
        coarse = h.level_patches(level - 1)
        cown = _owner_map(coarse, h.ncells(level - 1))
    copy_dst_p, copy_dst_l, copy_src_p, copy_src_l = [], [], [], []
    int_dst_p, int_dst_l, int_src_p, int_src_l, int_fx, int_fy = [], [], [], [], [], []
    edge_p, edge_dst, edge_src, edge_neg = [], [], [], []
    corn_p, corn_dst, corn_src, corn_neg = [], [], [], []
    r = h.ratio
    for p in pats:
        g = p.g
        i0, i1, j0, j1 = p.box
        I, J = np.meshgrid(np.arange(i0 - g, i1 + g + 1), np.arange(j0 - g, j1 + g + 1))
        ghost = ~((I >= i0) & (I <= i1) & (J >= j0) & (J <= j1))
        inside = (I >= 0) & (I < n) & (J >= 0) & (J < n)
        gi, gj = I[ghost & inside], J[ghost & inside]
        owner = own[gj, gi]
        dst = p.local(gi, gj)
        have = owner >= 0
        for q in np.unique(owner[have]):
            sel = have & (owner == q)
            Q = byid[int(q)]
            copy_dst_p.append(np.full(sel.sum(), p.pid, np.int32))
            copy_dst_l.append(dst[sel].astype(np.int32))
            copy_src_p.append(np.full(sel.sum(), Q.pid, np.int32))
            copy_src_l.append(Q.local(gi[sel], gj[sel]).astype(np.int32))
        if (~have).any():
            assert level > 0, "a level-0 ghost cell inside the domain has no owner"
            ui, uj = gi[~have], gj[~have]
            ic, jc = ui // r, uj // r
            cowner = cown[jc, ic]
            assert (cowner >= 0).all()
            for q in np.unique(cowner):
                sel = cowner == q
                C = byid[int(q)]
                int_dst_p.append(np.full(sel.sum(), p.pid, np.int32))
                int_dst_l.append(p.local(ui[sel], uj[sel]).astype(np.int32))
                int_src_p.append(np.full(sel.sum(), C.pid, np.int32))
                int_src_l.append(C.local(ic[sel], jc[sel]).astype(np.int32))
                int_fx.append(((ui[sel] - ic[sel] * r) + 0.5) / r - 0.5)
                int_fy.append(((uj[sel] - jc[sel] * r) + 0.5) / r - 0.5)
        # physical boundary: edge strips over the patch's own tangential range
        touch = (i0 == 0, i1 == n - 1, j0 == 0, j1 == n - 1)
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
        "copy": (cat(copy_dst_p, np.int32), cat(copy_dst_l, np.int32), cat(copy_src_p, np.int32),
                 cat(copy_src_l, np.int32)),
        "interp": (cat(int_dst_p, np.int32), cat(int_dst_l, np.int32), cat(int_src_p, np.int32),
                   cat(int_src_l, np.int32), cat(int_fx, np.float64), cat(int_fy, np.float64)),
        # neg: side code 1..4 = left, right, bottom, top (which wall's REFLECT
        # flag applies and which momentum it negates)
        "edge": (cat(edge_p, np.int32), cat(edge_dst, np.int32), cat(edge_src, np.int32), cat(edge_neg, np.int32)),
        "corner": (cat(corn_p, np.int32), cat(corn_dst, np.int32), cat(corn_src, np.int32),
                   cat(corn_neg, np.int32)),
    }


def make_coarsen(h):
    """(coarse patch, coarse local, fine patch, fine local of the lower-left
    fine cell) for every coarse cell under a fine patch."""
    r = h.ratio
    coarse = h.level_patches(0)
    cown = _owner_map(coarse, h.ncells(0))
    byid = {p.pid: p for p in h.patches}
    cp, cl, fp, fl = [], [], [], []
    for f in h.level_patches(1):
        i0, i1, j0, j1 = f.box
        IC, JC = np.meshgrid(np.arange(i0 // r, (i1 + 1) // r), np.arange(j0 // r, (j1 + 1) // r))
        IC, JC = IC.ravel(), JC.ravel()
        owner = cown[JC, IC]
        for q in np.unique(owner):
            sel = owner == q
            C = byid[int(q)]
            cp.append(np.full(sel.sum(), C.pid, np.int32))
            cl.append(C.local(IC[sel], JC[sel]).astype(np.int32))
            fp.append(np.full(sel.sum(), f.pid, np.int32))
            fl.append(f.local(IC[sel] * r, JC[sel] * r).astype(np.int32))
    return tuple(np.concatenate(x).astype(np.int32) for x in (cp, cl, fp, fl))


def make_transfer(new, old):
    """Regrid transfer schedule from hierarchy `old` to `new` (same base
    grid, ratio and ghost width): for every interior cell of every new
    patch, either a copy item (new patch, new local) <- (old patch of the
    same level holding the cell, its local), or, where no old patch of that
    level held it, an interp item from the old coarser level's cell with the
    fine centre's offset (fx, fy) in coarse widths.  Ghosts are not listed:
    the new hierarchy's own ghost fill takes them afterwards."""
    if (new.n0, new.ratio, new.g) != (old.n0, old.ratio, old.g):
        raise ValueError("make_transfer: the hierarchies differ in base grid, ratio or ghost width")
    r = new.ratio
    c_dp, c_dl, c_sp, c_sl = [], [], [], []
    i_dp, i_dl, i_sp, i_sl, i_fx, i_fy = [], [], [], [], [], []
    byid = {p.pid: p for p in old.patches}
    owners = {}
    for lev in range(new.nlevels):
        if lev < old.nlevels:
            owners[lev] = _owner_map(old.level_patches(lev), old.ncells(lev))
    for p in new.patches:
        lev = p.level
        I, J = np.meshgrid(np.arange(p.box[0], p.box[1] + 1), np.arange(p.box[2], p.box[3] + 1))
        I, J = I.ravel(), J.ravel()
        owner = owners[lev][J, I] if lev in owners else np.full(I.size, -1, np.int32)
        have = owner >= 0
        for q in np.unique(owner[have]):
            sel = have & (owner == q)
            Q = byid[int(q)]
            c_dp.append(np.full(sel.sum(), p.pid, np.int32))
            c_dl.append(p.local(I[sel], J[sel]).astype(np.int32))
            c_sp.append(np.full(sel.sum(), Q.pid, np.int32))
            c_sl.append(Q.local(I[sel], J[sel]).astype(np.int32))
        if (~have).any():
            if lev == 0 or lev - 1 not in owners:
                raise ValueError("make_transfer: new level-%d cells have no old source" % lev)
            ui, uj = I[~have], J[~have]
            ic, jc = ui // r, uj // r
            cown = owners[lev - 1][jc, ic]
            if (cown < 0).any():
                raise ValueError("make_transfer: a new fine cell lies outside the old coarser level")
            for q in np.unique(cown):
                sel = cown == q
                C = byid[int(q)]
                i_dp.append(np.full(sel.sum(), p.pid, np.int32))
                i_dl.append(p.local(ui[sel], uj[sel]).astype(np.int32))
                i_sp.append(np.full(sel.sum(), C.pid, np.int32))
                i_sl.append(C.local(ic[sel], jc[sel]).astype(np.int32))
                i_fx.append(((ui[sel] - ic[sel] * r) + 0.5) / r - 0.5)
                i_fy.append(((uj[sel] - jc[sel] * r) + 0.5) / r - 0.5)

    def cat(lst, dtype):
        return np.concatenate(lst).astype(dtype) if lst else np.zeros(0, dtype)

    return {
        "copy": (cat(c_dp, np.int32), cat(c_dl, np.int32), cat(c_sp, np.int32), cat(c_sl, np.int32)),
        "interp": (cat(i_dp, np.int32), cat(i_dl, np.int32), cat(i_sp, np.int32), cat(i_sl, np.int32),
                   cat(i_fx, np.float64), cat(i_fy, np.float64)),
    }
