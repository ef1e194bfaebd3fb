"""ctypes wrapper of liborc.so, the CPU restatement of the reference kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Never the product path.

Signatures mirror enlightning_amd.patch so a parity test reads the same on
both sides.  All arrays are numpy float64 / int32 in the SAMRAI CellData
layout.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

VARIANTS = [
    "hybrid_high_llf", "hybrid_high", "hybrid_low_llf", "hybrid_low",
    "weno_high_llf", "weno_high", "weno_low_llf", "weno_low", "drp",
]

_lib = None
_I = ctypes.c_int
_D = ctypes.c_double
_P = ctypes.c_void_p


class OrcGas(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "weno_alpha", "c", "gamma1", "p0", "T0", "c_lr", "rho0", "R", "c_p", "T_star_n", "T_star_o",
        "mu", "muB", "kappa", "c_v", "R_tilde", "M", "tau_n", "tau_o")] + [
        ("tag_tolerance", ctypes.c_double * 3)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    for n in ("orc_wenom5", "orc_wenop5"):
        getattr(lib, n).argtypes = [_D] * 5
        getattr(lib, n).restype = _D
    for n in ("orc_wenom3", "orc_wenop3"):
        getattr(lib, n).argtypes = [_D] * 3
        getattr(lib, n).restype = _D
    lib.orc_sensor_eps.argtypes = [_D]
    lib.orc_sensor_eps.restype = _D
    lib.orc_rhs_ghost_need.argtypes = [_I]
    lib.orc_rhs_ghost_need.restype = _I
    lib.orc_literals.argtypes = [_P]
    lib.orc_literals.restype = None
    lib.orc_rhs_variant.argtypes = [_I] + [_P] * 28
    lib.orc_rhs_variant.restype = _I
    lib.orc_runge_kutta.argtypes = [_P] * 14
    lib.orc_runge_kutta.restype = None
    lib.orc_update_state.argtypes = [_P] * 23
    lib.orc_update_state.restype = None
    lib.orc_tag_cells.argtypes = [_P] * 12
    lib.orc_tag_cells.restype = None
    lib.orc_fill_physical_bc.argtypes = [_I] * 6 + [_P] * 6
    lib.orc_fill_physical_bc.restype = None
    lib.orc_advance_patch.argtypes = ([_I] * 6 + [_P, _P, ctypes.POINTER(OrcGas), _P, _I, _P, _P, _P, _D, _I, _I]
                                      + [_P] * 6 + [_I, _I, _D])
    lib.orc_advance_patch.restype = _I
    lib.orc_absorb.argtypes = [_I] * 6 + [_P] * 5 + [_D, ctypes.POINTER(OrcGas)] + [_P] * 3
    lib.orc_absorb.restype = None
    lib.orc_add_source.argtypes = [_I] * 4 + [_P, _P, ctypes.POINTER(OrcSource), _I, _I, _D, _P, ctypes.c_long]
    lib.orc_add_source.restype = _I
    lib.orc_read_lightning.argtypes = [ctypes.c_char_p, _P, _I]
    lib.orc_read_lightning.restype = _I
    lib.orc_sample_mics.argtypes = [_I] * 4 + [_P, _P, _P, _I, _P, _P, _P, ctypes.c_long, _P, _P]
    lib.orc_sample_mics.restype = None
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


def _iref(v):
    return ctypes.byref(_I(int(v)))


def _dref(v):
    return ctypes.byref(_D(float(v)))


def _darr(vals):
    return (_D * len(vals))(*[float(v) for v in vals])


def wenop5(*f):
    return load().orc_wenop5(*map(float, f))


def wenom5(*f):
    return load().orc_wenom5(*map(float, f))


def wenop3(*f):
    return load().orc_wenop3(*map(float, f))


def wenom3(*f):
    return load().orc_wenom3(*map(float, f))


def sensor_eps(tol):
    return load().orc_sensor_eps(float(tol))


def ghost_need(variant):
    vid = VARIANTS.index(variant) if isinstance(variant, str) else int(variant)
    return load().orc_rhs_ghost_need(vid)


def literals():
    out = np.zeros(16)
    load().orc_literals(_ptr(out))
    return out


def gas_struct(gas):
    s = OrcGas()
    for name, _ in OrcGas._fields_:
        if name == "tag_tolerance":
            for i in range(3):
                s.tag_tolerance[i] = gas.tag_tolerance[i]
        else:
            setattr(s, name, getattr(gas, name))
    return s


def rhs(variant, box, g, xlo, dx, w, source, KK, p, T, gas, level, switch_map=None):
    vid = VARIANTS.index(variant) if isinstance(variant, str) else int(variant)
    rc = load().orc_rhs_variant(
        vid, _iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]), _iref(g), _iref(g),
        _darr(xlo), _darr(dx), _ptr(w), _ptr(source), _ptr(KK), _ptr(p), _ptr(T),
        _dref(gas.weno_alpha), _dref(gas.kappa), _dref(gas.mu), _dref(gas.muB), _dref(gas.R),
        _dref(gas.T_star_n), _dref(gas.T_star_o), _dref(gas.tau_n), _dref(gas.tau_o),
        _darr(gas.tag_tolerance), _iref(level), _dref(gas.c), _dref(gas.c_lr), _iref(5), _ptr(switch_map))
    if rc != 0:
        raise ValueError("orc_rhs_variant failed (%d)" % rc)


def runge_kutta(box, g, dt, a1, a2, b, w, ww, KK):
    load().orc_runge_kutta(_iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]), _iref(g), _iref(g),
                           _dref(dt), _dref(a1), _dref(a2), _dref(b), _ptr(w), _ptr(ww), _ptr(KK), _iref(5))


def update_state(box, g, xlo, dx, w, p, T, gas):
    load().orc_update_state(_darr(dx), _darr(xlo), _iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]),
                            _iref(g), _iref(g), _ptr(w), _ptr(p), _ptr(T), _dref(gas.rho0), _dref(gas.c),
                            _dref(gas.gamma1), _dref(gas.c_p), _dref(gas.p0), _dref(gas.T0), _dref(gas.c_lr),
                            _dref(gas.c_v), _dref(gas.R_tilde), _dref(gas.M), _dref(gas.R), _iref(5))


def tag_cells(box, g, tags, w, source, refine_tag_val, tolerance):
    load().orc_tag_cells(_iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]), _iref(g), _iref(g),
                         _ptr(tags), _ptr(w), _ptr(source), _iref(refine_tag_val), _darr(tolerance), _iref(5))


def fill_physical_bc(box, g, touch, reflect, w, p=None, T=None, source=None):
    load().orc_fill_physical_bc(box[0], box[1], box[2], box[3], g, g, (_I * 4)(*touch), (_I * 4)(*reflect),
                                _ptr(w), _ptr(p), _ptr(T), _ptr(source))


class OrcSource(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("type", ctypes.c_int), ("amp", ctypes.c_double),
                ("spread", ctypes.c_double), ("pos", ctypes.c_double * 2), ("sine_amp", ctypes.c_double),
                ("sine_freq", ctypes.c_double), ("omega", ctypes.c_double), ("nseg", ctypes.c_int),
                ("coords", ctypes.c_void_p)]


def add_source(box, patch_xlo, dx, src, iteration, pulse_time0, dt, source, sy=None):
    """The patch's mass source (enlightning.C:2140-2307) into `source`
    (interior cells, row stride sy, default the patch width); src has the
    fields of enlightning_amd.source.Source."""
    s = OrcSource()
    s.mode, s.type, s.amp, s.spread = int(src.mode), int(src.type), float(src.amp), float(src.spread)
    s.pos[0], s.pos[1] = float(src.pos[0]), float(src.pos[1])
    s.sine_amp, s.sine_freq, s.omega = float(src.sine_amp), float(src.sine_freq), float(src.omega)
    keep = None
    if src.coords is not None:
        keep = np.ascontiguousarray(src.coords, dtype=np.float64).reshape(-1, 4)
        s.nseg, s.coords = keep.shape[0], keep.ctypes.data
    nx = box[1] - box[0] + 1
    rc = load().orc_add_source(box[0], box[1], box[2], box[3], _darr(patch_xlo), _darr(dx), ctypes.byref(s),
                               int(iteration), int(pulse_time0), float(dt), _ptr(source), int(sy if sy else nx))
    del keep
    if rc:
        raise ValueError("orc_add_source failed (%d)" % rc)


def sample_mics(box, patch_xlo, patch_xhi, dx, mic_x, mic_y, pres, sy=None):
    """Pressure at the microphones inside the patch (enlightning.C:2414-2436):
    (values, hit) for interior pressure `pres` (row stride sy)."""
    mx = np.ascontiguousarray(mic_x, dtype=np.float64)
    my = np.ascontiguousarray(mic_y, dtype=np.float64)
    out = np.zeros(len(mx))
    hit = np.zeros(len(mx), np.int32)
    nx = box[1] - box[0] + 1
    load().orc_sample_mics(box[0], box[1], box[2], box[3], _darr(patch_xlo), _darr(patch_xhi), _darr(dx), len(mx),
                           _ptr(mx), _ptr(my), _ptr(np.ascontiguousarray(pres).ravel()), int(sy if sy else nx),
                           _ptr(out), _ptr(hit))
    return out, hit


def read_lightning(path, max_segments=100000):
    buf = np.zeros((max_segments, 4))
    n = load().orc_read_lightning(str(path).encode(), buf.ctypes.data, max_segments)
    if n < 0:
        raise OSError("can't read %s" % path)
    return buf[:min(n, max_segments)].copy()


def absorb(box, g, patch_xlo, dx, grid_xlo, grid_xhi, sides, width, gas, w, p, T):
    """Absorbing envelopes on the sides with sides[s] set (left, right, bottom,
    top), enlightning.C:630-786."""
    gs = gas_struct(gas)
    load().orc_absorb(box[0], box[1], box[2], box[3], g, g, _darr(patch_xlo), _darr(dx), _darr(grid_xlo),
                      _darr(grid_xhi), (_I * 4)(*sides), float(width), ctypes.byref(gs), _ptr(w), _ptr(p), _ptr(T))


class HostPatch:
    """Host mirror of enlightning_amd.patch.DevicePatch driven by the oracle."""

    def __init__(self, wl, gas):
        self.wl, self.gas = wl, gas
        self.n, self.g = wl.n, wl.g
        self.box = (0, wl.n - 1, 0, wl.n - 1)
        self.xlo = (0.0, 0.0)
        self.dx = (wl.dx, wl.dx)
        self.reflect = (1, 1, 1, 1)
        self.absorb_width = 0.0  # envelopes on the sides whose reflect is 0 when > 0
        self.ww = np.zeros((5, wl.n, wl.n))
        self.KK = np.zeros((5, wl.n, wl.n))

    def init_state(self, w, p, T, src):
        self.w, self.p, self.T, self.src = w.copy(), p.copy(), T.copy(), src.copy()
        fill_physical_bc(self.box, self.g, (1, 1, 1, 1), self.reflect, self.w, self.p, self.T, self.src)
        update_state(self.box, self.g, self.xlo, self.dx, self.w, self.p, self.T, self.gas)
        fill_physical_bc(self.box, self.g, (1, 1, 1, 1), self.reflect, self.w, self.p, self.T, self.src)

    def advance(self, nsteps, dt, rk, variant=None, level=0, tile=0, nthreads=1):
        variant = self.wl.variant if variant is None else variant
        vid = VARIANTS.index(variant) if isinstance(variant, str) else int(variant)
        a1, a2, b = rk
        gs = gas_struct(self.gas)
        rc = load().orc_advance_patch(
            vid, self.box[0], self.box[1], self.box[2], self.box[3], self.g, _darr(self.xlo), _darr(self.dx),
            ctypes.byref(gs), (_I * 4)(*self.reflect), len(a1), _darr(a1), _darr(a2), _darr(b), float(dt),
            int(nsteps), int(level), _ptr(self.w), _ptr(self.p), _ptr(self.T), _ptr(self.src), _ptr(self.ww),
            _ptr(self.KK), int(tile), int(nthreads), float(self.absorb_width))
        if rc != 0:
            raise ValueError("orc_advance_patch failed (%d)" % rc)
