"""enl_rhs_variant_noalloc: the rhs entry point that allocates nothing.

Called through the C ABI with torch device memory, against enl_rhs_variant
(the same kernels with a workspace the library allocates): the rhs and the
switch map must be identical, bit for bit, on two calls in a row, the second
with a smaller patch (the fixed workspace keeps the first call's list, so the
second must clear its counter), and again after the allocating entry point
has run in between.  Host arrays and patches beyond the fixed workspace are
rejected with an error before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from enlightning_amd.gas import Gas
from tests.fields import random_state

pytestmark = pytest.mark.gpu


def _call(fn, variant, box, g, xlo, dx, w, src, KK, p, T, gas, level, sm):
    from enlightning_amd import _lib
    from enlightning_amd.patch import _dref, _iref, _ptr, _darr
    lib = _lib.load()
    vid = _lib.VARIANT_ID[variant]
    return getattr(lib, fn)(
        vid, _iref(box[0]), _iref(box[1]), _iref(box[2]), _iref(box[3]), _iref(g), _iref(g),
        _darr(xlo), _darr(dx), _ptr(w), _ptr(src), _ptr(KK), _ptr(p), _ptr(T),
        _dref(gas.weno_alpha), _dref(gas.kappa), _dref(gas.mu), _dref(gas.muB), _dref(gas.R),
        _dref(gas.T_star_n), _dref(gas.T_star_o), _dref(gas.tau_n), _dref(gas.tau_o), _darr(gas.tag_tolerance),
        _iref(level), _dref(gas.c), _dref(gas.c_lr), _iref(5), _ptr(sm))


def _both(variant, nx, ny, g, seed, level=0):
    import torch
    from enlightning_amd.patch import use_torch_stream
    use_torch_stream()
    gas = Gas(c_lr=-2.0e-3)
    w, p, T, src = random_state(nx, ny, g, seed=seed)
    box, xlo, dx = (3, 3 + nx - 1, -5, -5 + ny - 1), (0.25, 10.0), (1.0 / 512, 1.0 / 256)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    wd, pd, Td, sd = dev(w), dev(p), dev(T), dev(src)
    out = {}
    for fn in ("enl_rhs_variant_noalloc", "enl_rhs_variant"):
        KK = torch.full((5, ny, nx), np.nan, dtype=torch.float64, device="cuda")
        sm = torch.full((ny, nx), -1, dtype=torch.int32, device="cuda")
        rc = _call(fn, variant, box, g, xlo, dx, wd, sd, KK, pd, Td, gas, level, sm)
        assert rc == 0, (fn, rc)
        torch.cuda.synchronize()
        out[fn] = (KK.cpu().numpy(), sm.cpu().numpy())
    (Ka, sa), (Kb, sb) = out["enl_rhs_variant_noalloc"], out["enl_rhs_variant"]
    assert np.isfinite(Ka).all()
    assert np.array_equal(sa, sb)
    assert np.array_equal(Ka, Kb)
    return sa


def test_noalloc_equals_allocating_entry_point_twice_then_smaller():
    print("no-alloc rhs: 200x150 hybrid_high_llf level 0", flush=True)
    sm = _both("hybrid_high_llf", 200, 150, 4, seed=61)
    assert 0 < sm.sum() < sm.size  # the WENO list is used
    print("no-alloc rhs: 64x40 hybrid_low_llf level 0", flush=True)
    sm = _both("hybrid_low_llf", 64, 40, 3, seed=62)
    assert 0 < sm.sum() < sm.size
    print("no-alloc rhs: 96x72 weno_high_llf, 33x17 drp", flush=True)
    _both("weno_high_llf", 96, 72, 4, seed=63)
    _both("drp", 33, 17, 3, seed=64)
    print("no-alloc rhs: 40x30 hybrid_high_llf level 0 again", flush=True)
    _both("hybrid_high_llf", 40, 30, 4, seed=65)


def test_noalloc_rejects_host_arrays_and_oversized_patches():
    import torch
    from enlightning_amd import _lib
    gas = Gas()
    w, p, T, src = random_state(8, 8, 4, seed=66)
    KK = np.zeros((5, 8, 8))
    rc = _call("enl_rhs_variant_noalloc", "hybrid_high_llf", (0, 7, 0, 7), 4, (0, 0), (1, 1), w, src, KK, p, T,
               gas, 0, None)
    assert rc == -3
    code, msg = _lib.last_error()
    assert "device memory" in msg
    # 2049 x 2048 interior cells exceed the 2^22-cell workspace; the size is
    # checked before any array is touched
    tiny = torch.zeros(8, dtype=torch.float64, device="cuda")
    rc = _call("enl_rhs_variant_noalloc", "hybrid_high_llf", (0, 2048, 0, 2047), 4, (0, 0), (1, 1), tiny, tiny,
               tiny, tiny, tiny, gas, 0, None)
    assert rc == -2
    code, msg = _lib.last_error()
    assert "exceed" in msg
