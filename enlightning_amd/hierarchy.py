"""A two-level hierarchy resident in HBM, driven through the batch C ABI
(enl_batch_*): build from an amr.Hierarchy, load the initial state, advance
in plain (drop-in per patch), fused or graph mode, read states back."""
import ctypes

import numpy as np

from . import _lib
from .gas import RK3
from .workloads import pulses_at, state_from_overpressure

MODES = {"plain": 0, "fused": 1, "graph": 2}


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def level_tags(gas, wl, tag_fn):
    """Tags of the workload's initial state on the base grid (one patch
    covering the domain), from tag_fn(box, g, tags, w, source, val, tol)."""
    n, g = wl.n, wl.g
    xc = (np.arange(-g, n + g) + 0.5) * wl.dx
    w = np.ascontiguousarray(state_from_overpressure(pulses_at(wl, xc, xc), gas))
    src = np.zeros((n + 2 * g, n + 2 * g))
    tags = np.zeros((n, n), np.int32)
    tag_fn((0, n - 1, 0, n - 1), g, tags, w, src, 1, gas.tag_tolerance)
    return tags


def patch_initial_state(h, patch, wl, gas):
    """w on the patch's ghosted cell centres (analytic initial condition)."""
    dx = h.dx(patch.level)
    i0, i1, j0, j1 = patch.box
    g = patch.g
    xc = (np.arange(i0 - g, i1 + g + 1) + 0.5) * dx
    yc = (np.arange(j0 - g, j1 + g + 1) + 0.5) * dx
    return np.ascontiguousarray(state_from_overpressure(pulses_at(wl, xc, yc), gas))


class DeviceHierarchy:
    def __init__(self, h, gas, reflect=(1, 1, 1, 1)):
        self.h, self.gas, self.reflect = h, gas, tuple(reflect)
        self.lib = _lib.load()
        n = len(h.patches)
        desc = (_lib.EnlPatchDesc * n)()
        for k, p in enumerate(h.patches):
            d = desc[k]
            d.level = p.level
            for q in range(4):
                d.box[q] = p.box[q]
            dx = h.dx(p.level)
            d.xlo[0] = p.box[0] * dx
            d.xlo[1] = p.box[2] * dx
            d.dx[0] = dx
            d.dx[1] = dx
        handle = ctypes.c_void_p()
        _lib.check(self.lib.enl_batch_create(n, desc, h.g, h.ratio, ctypes.byref(handle)))
        self.handle = handle
        self._keep = []
        for lev, f in enumerate(h.fill):
            c, it, e, k = f["copy"], f["interp"], f["edge"], f["corner"]
            arrs = [np.ascontiguousarray(a) for a in c + it + e + k]
            self._keep += arrs
            P = [_p(a) for a in arrs]
            _lib.check(self.lib.enl_batch_set_fill(
                self.handle, lev, len(c[0]), *P[0:4], len(it[0]), *P[4:10], len(e[0]), *P[10:14],
                len(k[0]), *P[14:18]))
        if h.coarsen is not None:
            arrs = [np.ascontiguousarray(a) for a in h.coarsen]
            self._keep += arrs
            _lib.check(self.lib.enl_batch_set_coarsen(self.handle, len(arrs[0]), *[_p(a) for a in arrs]))

    def _reflect(self):
        return (ctypes.c_int * 4)(*self.reflect)

    def set_tiles(self, bx, by):
        _lib.check(self.lib.enl_batch_set_tiles(self.handle, bx, by))

    def load_initial(self, wl, states=None):
        """Upload the analytic initial state (or given per-patch w arrays),
        fill ghosts, derive p, T and fill again."""
        gas = self.gas
        for k, p in enumerate(self.h.patches):
            w = patch_initial_state(self.h, p, wl, gas) if states is None else np.ascontiguousarray(states[k])
            src = np.zeros((p.Ny, p.Nx))
            _lib.check(self.lib.enl_batch_copy_state(self.handle, k, 1, _p(w), None, None, _p(src)))
        refl = self._reflect()
        _lib.check(self.lib.enl_batch_prepare(self.handle, refl))
        _lib.check(self.lib.enl_batch_fill(self.handle, refl))
        gs = gas.to_struct()
        _lib.check(self.lib.enl_batch_update_state(self.handle, ctypes.byref(gs)))
        _lib.check(self.lib.enl_batch_fill(self.handle, refl))

    def advance(self, nsteps, dt, mode="graph", variant="hybrid_high_llf", rk=RK3):
        vid = _lib.VARIANT_ID[variant] if isinstance(variant, str) else int(variant)
        a1, a2, b = rk
        D = ctypes.c_double * len(a1)
        gs = self.gas.to_struct()
        _lib.check(self.lib.enl_batch_advance(self.handle, MODES[mode], vid, ctypes.byref(gs), self._reflect(),
                                              len(a1), D(*a1), D(*a2), D(*b), float(dt), int(nsteps)))

    def states(self):
        """Per patch (w, p, T) host copies of the current state."""
        out = []
        for k, p in enumerate(self.h.patches):
            w = np.empty((5, p.Ny, p.Nx))
            pr = np.empty((p.Ny, p.Nx))
            T = np.empty((p.Ny, p.Nx))
            _lib.check(self.lib.enl_batch_copy_state(self.handle, k, 0, _p(w), _p(pr), _p(T), None))
            out.append((w, pr, T))
        return out

    def close(self):
        if self.handle:
            self.lib.enl_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
