"""The two-level hierarchy builder and its ghost-fill schedule (CPU).

SAMRAI's refine / coarsen operators are not vendored, so the schedule is
checked against what the operators must do on fields whose answer is known:
copies are exact, linear refinement reproduces linear fields, the
fine-to-coarse average of a linear field is its value at the coarse centre,
and the physical boundary follows the reference's rules (nearest interior
cell, normal momentum negated at REFLECT walls, corners along x).
"""
import numpy as np
import pytest

import oracle
from enlightning_amd import amr
from enlightning_amd.gas import Gas, RK3
from enlightning_amd.hierarchy import level_tags, patch_initial_state, refine_from_base
from enlightning_amd.workloads import Workload
from oracle.amr import HostHierarchy

WL = Workload("amr_test_64", n=64, seed=7, npulses=4, amp_range=(0.3, 0.8), sigma_range=(0.03, 0.06))


@pytest.fixture(scope="module")
def hier():
    gas = Gas()
    tags = level_tags(gas, WL, oracle.tag_cells)
    assert tags.any()
    return amr.build(WL.n, 4, 4, 1.0, tags, tile0=16)


def test_fine_boxes(hier):
    h = hier
    fine = h.level_patches(1)
    assert len(fine) >= 4
    n1 = h.ncells(1)
    cover = np.zeros((n1, n1), np.int32)
    for p in fine:
        i0, i1, j0, j1 = p.box
        assert 8 <= p.nx <= 64 and 8 <= p.ny <= 64
        assert i0 % 4 == 0 and j0 % 4 == 0 and (i1 + 1) % 4 == 0 and (j1 + 1) % 4 == 0
        assert 0 <= i0 and i1 < n1 and 0 <= j0 and j1 < n1
        # never closer than one coarse cell (g fine cells) to the edge unless touching
        assert i0 == 0 or i0 >= 4
        assert j0 == 0 or j0 >= 4
        cover[j0:j1 + 1, i0:i1 + 1] += 1
    assert cover.max() == 1  # no overlap
    # every buffered tag is covered
    tags = amr.buffer_tags(level_tags(Gas(), WL, oracle.tag_cells), 2)
    jj, ii = np.nonzero(tags)
    assert (cover[4 * jj, 4 * ii] == 1).all()


@pytest.mark.parametrize("n0,tile0", [(65, 64), (65, 16), (66, 16), (67, 16), (97, 32), (70, 64)])
def test_odd_base_sizes_build_without_overlap(n0, tile0):
    """Base sizes that leave a remainder block: n0 = 65 with tags at coarse
    columns 62 and 64 (blocks of 16 coarse cells end at 63, the last block
    is column 64 alone), plus random tag maps.  The fine boxes stay inside
    their blocks (no overlap), lie in the domain, and build() succeeds (its
    schedule asserts that no corner block straddles the domain edge)."""
    g, r = 4, 4
    rng = np.random.default_rng(n0 * 100 + tile0)
    maps = []
    t = np.zeros((n0, n0), np.int32)
    t[40:42, 62] = 1
    t[40:42, 64] = 1
    maps.append((t, 0))
    for _ in range(3):
        maps.append(((rng.random((n0, n0)) < 0.01).astype(np.int32), int(rng.integers(0, 3))))
    t = np.zeros((n0, n0), np.int32)
    t[n0 - 2:, :] = 1  # along the top edge and the last row of blocks
    t[:, n0 - 1] = 1
    maps.append((t, 1))
    for tags, buf in maps:
        h = amr.build(n0, r, g, 1.0, tags, tile0=tile0, tag_buffer=buf)
        n1 = h.ncells(1) if h.nlevels > 1 else 0
        cover = np.zeros((n1, n1), np.int32)
        for p in h.level_patches(1):
            i0, i1, j0, j1 = p.box
            assert 0 <= i0 <= i1 < n1 and 0 <= j0 <= j1 < n1
            cover[j0:j1 + 1, i0:i1 + 1] += 1
        assert cover.max(initial=0) <= 1
        if buf == 0 and h.nlevels > 1:
            jj, ii = np.nonzero(tags)
            assert (cover[r * jj, r * ii] == 1).all()
        base = np.zeros((n0, n0), np.int32)
        for p in h.level_patches(0):
            assert p.nx >= g and p.ny >= g
            base[p.box[2]:p.box[3] + 1, p.box[0]:p.box[1] + 1] += 1
        assert (base == 1).all()


def _linear_states(h, a, bx, by):
    """Per patch: w_k = a_k + bx_k*x + by_k*y on all ghosted cell centres."""
    states = []
    for p in h.patches:
        dx = h.dx(p.level)
        xc = (np.arange(p.box[0] - p.g, p.box[1] + p.g + 1) + 0.5) * dx
        yc = (np.arange(p.box[2] - p.g, p.box[3] + p.g + 1) + 0.5) * dx
        X, Y = np.meshgrid(xc, yc)
        states.append(np.stack([a[k] + bx[k] * X + by[k] * Y for k in range(7)]))
    return states


def test_fill_reproduces_linear_fields_and_reflects(hier):
    h = hier
    hh = HostHierarchy(h, Gas(), reflect=(1, 1, 1, 1))
    a = np.array([1.2, 0.3, -0.2, 0.05, 350.0, 101325.0, 294.0])
    bx = np.array([0.1, 2.0, 0.5, 0.01, 3.0, 100.0, 1.0])
    by = np.array([-0.05, 1.0, -1.5, 0.02, -2.0, 50.0, -0.5])
    exact = _linear_states(h, a, bx, by)
    for k, p in enumerate(h.patches):
        st = exact[k].copy()
        g = p.g
        ghost = np.ones((p.Ny, p.Nx), bool)
        ghost[g:-g, g:-g] = False
        st[:, ghost] = np.nan  # ghosts must all be written by the fill
        o, sk = hh.off[k], hh.sk[k]
        hh.pool[o:o + 7 * sk] = st.ravel()
    hh.fill_all()
    for k, p in enumerate(h.patches):
        o, sk = hh.off[k], hh.sk[k]
        got = hh.pool[o:o + 7 * sk].reshape(7, p.Ny, p.Nx)
        assert np.isfinite(got).all(), p
        n = h.ncells(p.level)
        I = np.arange(p.box[0] - p.g, p.box[1] + p.g + 1)[None, :]
        J = np.arange(p.box[2] - p.g, p.box[3] + p.g + 1)[:, None]
        inside = (I >= 0) & (I < n) & (J >= 0) & (J < n)
        if p.level > 0:
            # linear refinement is exact while its coarse stencil (+-1 coarse
            # cell) stays inside the domain; next to the edge it sees the
            # coarse physical-boundary ghosts (zero-order copies)
            r = h.ratio
            inside &= (I >= r) & (I < n - r) & (J >= r) & (J < n - r)
        for c in range(7):
            e = exact[k][c]
            scale = np.abs(e).max()
            assert np.abs(got[c][inside] - e[inside]).max() <= 1e-12 * scale, (p.pid, c)
        # outside the domain (reference rules): edge strips and corner blocks
        # beyond an x wall hold the nearest in-domain cell (clamped index)
        # with x momentum negated beyond x walls and y momentum beyond y
        # walls; a corner block beyond a y wall only (its columns outside the
        # patch's range) is filled along x from the patch's first / last
        # column, i.e. the bottom / top strip value there, with x momentum
        # negated as well (enlightning.C:948-1008)
        dx = h.dx(p.level)
        Ib = I + 0 * J
        Jb = J + 0 * I
        out_x = (Ib < 0) | (Ib >= n)
        out_y = (Jb < 0) | (Jb >= n)
        node = out_y & ~out_x & ((Ib < p.box[0]) | (Ib > p.box[1]))
        col = np.where(Ib < p.box[0], p.box[0], p.box[1])
        Ic = np.where(node, col, np.clip(Ib, 0, n - 1))
        Jc = np.clip(Jb, 0, n - 1)
        out = out_x | out_y
        # the source cell lies in the patch's ghosted box (interior, or an
        # in-domain ghost already checked above): the rule is an exact copy
        li = Ic - (p.box[0] - p.g)
        lj = Jc - (p.box[2] - p.g)
        for c in range(7):
            sign = np.ones(Ic.shape)
            if c == 1:
                sign = np.where(out_x | node, -1.0, 1.0)
            if c == 2:
                sign = np.where(out_y, -1.0, 1.0)
            want = got[c][lj, li] * sign
            if out.any():
                assert np.array_equal(got[c][out], want[out]), (p.pid, c)
            # and beyond the walls of the domain the values are those of the
            # exact field where the source is an interior cell
            src_interior = out & (li >= p.g) & (li < p.g + p.nx) & (lj >= p.g) & (lj < p.g + p.ny)
            if src_interior.any():
                ref = (a[c] + bx[c] * (Ic + 0.5) * dx + by[c] * (Jc + 0.5) * dx) * sign
                assert np.abs(got[c][src_interior] - ref[src_interior]).max() <= 1e-12 * np.abs(ref).max()


def test_coarsen_of_linear_field(hier):
    h = hier
    hh = HostHierarchy(h, Gas())
    a = np.array([1.2, 0.3, -0.2, 0.05, 350.0, 101325.0, 294.0])
    bx = np.array([0.1, 2.0, 0.5, 0.01, 3.0, 100.0, 1.0])
    by = np.array([-0.05, 1.0, -1.5, 0.02, -2.0, 50.0, -0.5])
    exact = _linear_states(h, a, bx, by)
    for k, p in enumerate(h.patches):
        o, sk = hh.off[k], hh.sk[k]
        st = exact[k].copy()
        if p.level == 0:
            st[...] = 0.0  # coarse cells under fine patches must be overwritten
        hh.pool[o:o + 7 * sk] = st.ravel()
    hh.coarsen()
    cp, cl, fp, fl = h.coarsen
    for q in range(0, len(cp), 7):
        k = cp[q]
        got = hh.pool[hh.off[k] + cl[q] + hh.sk[k] * np.arange(7)]
        want = exact[k].reshape(7, -1)[:, cl[q]]
        assert np.allclose(got, want, rtol=1e-13, atol=1e-12)


def test_twin_steps_stay_finite_and_fine_level_is_weno(hier):
    gas = Gas()
    h = hier
    hh = HostHierarchy(h, gas)
    hh.load_initial([patch_initial_state(h, p, WL, gas) for p in h.patches])
    dt = 0.5 * h.dx(1) / gas.c  # one dt for all levels, from the finest dx
    hh.advance(2, dt, RK3)
    for w, p, T in hh.states():
        assert np.isfinite(w).all() and np.isfinite(p).all()


def test_refine_from_base_linear_and_windows(hier):
    """Initial state of an evolved hierarchy: level-0 patches take their
    window of the base grid; level-1 interiors are the linear refinement,
    exact for a linear base field."""
    h = hier
    n0, g = h.n0, 4
    x = (np.arange(-g, n0 + g) + 0.5) * h.dx(0)
    X, Y = np.meshgrid(x, x)
    a = np.array([1.2, 0.3, -0.2, 0.05, 350.0])
    bx = np.array([0.1, 2.0, 0.5, 0.01, 3.0])
    by = np.array([-0.05, 1.0, -1.5, 0.02, -2.0])
    wb = np.stack([a[k] + bx[k] * X + by[k] * Y for k in range(5)])
    for p in h.patches:
        w = refine_from_base(h, p, wb)
        assert w.shape == (5, p.Ny, p.Nx)
        if p.level == 0:
            i0, i1, j0, j1 = p.box
            assert np.array_equal(w, wb[:, j0:j1 + 2 * g + 1, i0:i1 + 2 * g + 1])
        else:
            dx = h.dx(1)
            xc = (np.arange(p.box[0], p.box[1] + 1) + 0.5) * dx
            yc = (np.arange(p.box[2], p.box[3] + 1) + 0.5) * dx
            XX, YY = np.meshgrid(xc, yc)
            for k in range(5):
                exact = a[k] + bx[k] * XX + by[k] * YY
                got = w[k, g:-g, g:-g]
                assert np.abs(got - exact).max() <= 1e-12 * np.abs(exact).max(), (p.pid, k)
