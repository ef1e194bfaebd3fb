"""Device-resident patch and thin Python calls into the C ABI.

torch provides device memory and the stream; all arithmetic runs in the HIP
kernels of libenlightning_hip.so.  Arrays use the SAMRAI CellData layout:
w_n is (5, Ny, Nx) C-order, scalars (Ny, Nx), interior arrays (.., ny, nx).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .gas import RK3

_c_int = ctypes.c_int
_c_double = ctypes.c_double


def _ptr(x):
    """Raw address of a torch tensor or numpy array (None -> NULL)."""
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        assert x.is_contiguous()
        return ctypes.c_void_p(x.data_ptr())
    assert x.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(x.ctypes.data)


def _iref(v):
    return ctypes.byref(_c_int(int(v)))


def _dref(v):
    return ctypes.byref(_c_double(float(v)))


def _darr(vals):
    return (_c_double * len(vals))(*[float(v) for v in vals])


def use_torch_stream():
    """Route library launches onto torch's current stream."""
    lib = _lib.load()
    lib.enl_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def rhs(variant, box, g, xlo, dx, w, source, KK, p, T, gas, level, switch_map=None):
    """enl_rhs_variant on one patch.  box = (ifirst0, ilast0, ifirst1, ilast1)."""
    lib = _lib.load()
    vid = _lib.VARIANT_ID[variant] if isinstance(variant, str) else int(variant)
    tol = _darr(gas.tag_tolerance)
    rc = lib.enl_rhs_variant(
        vid, _iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]), _iref(g), _iref(g),
        _darr(xlo), _darr(dx), _ptr(w), _ptr(source), _ptr(KK), _ptr(p), _ptr(T),
        _dref(gas.weno_alpha), _dref(gas.kappa), _dref(gas.mu), _dref(gas.muB), _dref(gas.R),
        _dref(gas.T_star_n), _dref(gas.T_star_o), _dref(gas.tau_n), _dref(gas.tau_o), tol,
        _iref(level), _dref(gas.c), _dref(gas.c_lr), _iref(5), _ptr(switch_map))
    _lib.check(rc)


def runge_kutta(box, g, dt, a1, a2, b, w, ww, KK):
    lib = _lib.load()
    lib.runge_kutta_(_iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]), _iref(g), _iref(g),
                     _dref(dt), _dref(a1), _dref(a2), _dref(b), _ptr(w), _ptr(ww), _ptr(KK), _iref(5))
    _lib.check()


def update_state(box, g, xlo, dx, w, p, T, gas):
    lib = _lib.load()
    lib.update_state_(_darr(dx), _darr(xlo), _iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]),
                      _iref(g), _iref(g), _ptr(w), _ptr(p), _ptr(T), _dref(gas.rho0), _dref(gas.c),
                      _dref(gas.gamma1), _dref(gas.c_p), _dref(gas.p0), _dref(gas.T0), _dref(gas.c_lr),
                      _dref(gas.c_v), _dref(gas.R_tilde), _dref(gas.M), _dref(gas.R), _iref(5))
    _lib.check()


def tag_cells(box, g, tags, w, source, refine_tag_val, tolerance):
    lib = _lib.load()
    lib.tag_cells_(_iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]), _iref(g), _iref(g),
                   _ptr(tags), _ptr(w), _ptr(source), _iref(refine_tag_val), _darr(tolerance), _iref(5))
    _lib.check()


def fill_physical_bc(box, g, touch, reflect, w, p=None, T=None, source=None):
    lib = _lib.load()
    rc = lib.enl_fill_physical_bc(box[0], box[1], box[2], box[3], g, g, (_c_int * 4)(*touch),
                                  (_c_int * 4)(*reflect), _ptr(w), _ptr(p), _ptr(T), _ptr(source))
    _lib.check(rc)


def absorb(box, g, patch_xlo, dx, grid_xlo, grid_xhi, sides, width, gas, w, p, T):
    """Absorbing envelopes (enl_absorb, enlightning.C:630-786) on the sides
    with sides[s] set (left, right, bottom, top); device arrays."""
    lib = _lib.load()
    use_torch_stream()
    gs = gas.to_struct()
    rc = lib.enl_absorb(box[0], box[1], box[2], box[3], g, g, _darr(patch_xlo), _darr(dx), _darr(grid_xlo),
                        _darr(grid_xhi), (_c_int * 4)(*sides), float(width), ctypes.byref(gs), _ptr(w), _ptr(p),
                        _ptr(T))
    _lib.check(rc)


class DevicePatch:
    """One patch covering the whole domain, resident in HBM.

    Holds the ghosted w_n (5 comps), pressure, temperature, source and the
    interior scratch ww (state at the start of the step) and KK (rhs)."""

    def __init__(self, wl, gas, device="cuda"):
        self.wl, self.gas = wl, gas
        self.n, self.g = wl.n, wl.g
        self.box = (0, wl.n - 1, 0, wl.n - 1)
        self.xlo = (0.0, 0.0)
        self.dx = (wl.dx, wl.dx)
        N = wl.n + 2 * wl.g
        opts = dict(dtype=torch.float64, device=device)
        self.w = torch.empty((5, N, N), **opts)
        self.p = torch.empty((N, N), **opts)
        self.T = torch.empty((N, N), **opts)
        self.src = torch.empty((N, N), **opts)
        self.ww = torch.empty((5, wl.n, wl.n), **opts)
        self.KK = torch.empty((5, wl.n, wl.n), **opts)
        self.reflect = (1, 1, 1, 1)
        # absorbing envelopes on the sides whose reflect flag is 0, when > 0
        self.absorb_width = 0.0

    def load_host(self, w, p, T, src):
        self.w.copy_(torch.from_numpy(w))
        self.p.copy_(torch.from_numpy(p))
        self.T.copy_(torch.from_numpy(T))
        self.src.copy_(torch.from_numpy(src))

    def init_state(self, w, p, T, src):
        """Upload the initial w and derive p, T (update_state) and ghosts."""
        self.load_host(w, p, T, src)
        use_torch_stream()
        fill_physical_bc(self.box, self.g, (1, 1, 1, 1), self.reflect, self.w, self.p, self.T, self.src)
        update_state(self.box, self.g, self.xlo, self.dx, self.w, self.p, self.T, self.gas)
        fill_physical_bc(self.box, self.g, (1, 1, 1, 1), self.reflect, self.w, self.p, self.T, self.src)

    def advance(self, nsteps, dt, variant=None, level=0, rk=RK3):
        lib = _lib.load()
        use_torch_stream()
        variant = self.wl.variant if variant is None else variant
        vid = _lib.VARIANT_ID[variant] if isinstance(variant, str) else int(variant)
        a1, a2, b = rk
        gas_s = self.gas.to_struct()
        rc = lib.enl_advance_patch(
            vid, self.box[0], self.box[1], self.box[2], self.box[3], self.g, _darr(self.xlo),
            _darr(self.dx), ctypes.byref(gas_s), (_c_int * 4)(*self.reflect), len(a1), _darr(a1),
            _darr(a2), _darr(b), float(dt), int(nsteps), int(level), _ptr(self.w), _ptr(self.p),
            _ptr(self.T), _ptr(self.src), _ptr(self.ww), _ptr(self.KK), float(self.absorb_width))
        _lib.check(rc)

    def advance_fused(self, nsteps, dt, variant=None, level=0, rk=RK3):
        """Same steps as advance(), one fused kernel per stage (bit-identical
        interiors).  Without an absorbing envelope the returned ghosts of
        w, p, T hold the physical boundary of the returned state; with one
        they are not refreshed."""
        lib = _lib.load()
        use_torch_stream()
        variant = self.wl.variant if variant is None else variant
        vid = _lib.VARIANT_ID[variant] if isinstance(variant, str) else int(variant)
        nbytes = lib.enl_fused_scratch_bytes(self.box[0], self.box[1], self.box[2], self.box[3], self.g)
        if getattr(self, "_scratch", None) is None or self._scratch.numel() < nbytes:
            self._scratch = torch.empty(nbytes, dtype=torch.uint8, device=self.w.device)
        a1, a2, b = rk
        gas_s = self.gas.to_struct()
        rc = lib.enl_advance_patch_fused(
            vid, self.box[0], self.box[1], self.box[2], self.box[3], self.g, _darr(self.xlo),
            _darr(self.dx), ctypes.byref(gas_s), (_c_int * 4)(*self.reflect), len(a1), _darr(a1),
            _darr(a2), _darr(b), float(dt), int(nsteps), int(level), _ptr(self.w), _ptr(self.p),
            _ptr(self.T), _ptr(self.src), ctypes.c_void_p(self._scratch.data_ptr()), nbytes,
            float(self.absorb_width))
        _lib.check(rc)

    def to_host(self):
        torch.cuda.synchronize()
        return (self.w.cpu().numpy(), self.p.cpu().numpy(), self.T.cpu().numpy(), self.src.cpu().numpy())
