"""Calls rhs_ (the reference's one entry point, whose scheme comes from
ENL_RHS_VARIANT, read once per process) and the rhs_<variant>_ aliases
through the C ABI, with device and with host arrays, and compares each
result bitwise with enl_rhs_variant's on the same inputs.

Run as a child process by tests/test_gpu_rhs_entry_points.py, so that the
environment variable is set before the library's first rhs_ call:
    python -m tests._rhs_entry_points <expected variant> <alias> ...
Prints one line per check and exits non-zero on a mismatch.
"""
import ctypes
import sys

import numpy as np
import torch

from enlightning_amd import _lib
from enlightning_amd.gas import Gas
from tests.fields import random_state

VARIANTS = ["hybrid_high_llf", "hybrid_high", "hybrid_low_llf", "hybrid_low", "weno_high_llf", "weno_high",
            "weno_low_llf", "weno_low", "drp"]


def args_for(box, g, xlo, dx, w, src, KK, p, T, gas, level, ptr):
    I = lambda v: ctypes.byref(ctypes.c_int(v))  # noqa: E731
    Dr = lambda v: ctypes.byref(ctypes.c_double(v))  # noqa: E731
    Da = lambda vals: (ctypes.c_double * len(vals))(*vals)  # noqa: E731
    keep = []  # the ctypes objects must outlive the call

    def k(o):
        keep.append(o)
        return o
    a = [k(I(box[0])), k(I(box[1])), k(I(box[2])), k(I(box[3])), k(I(g)), k(I(g)), k(Da(xlo)), k(Da(dx)),
         ptr(w), ptr(src), ptr(KK), ptr(p), ptr(T), k(Dr(gas.weno_alpha)), k(Dr(gas.kappa)), k(Dr(gas.mu)),
         k(Dr(gas.muB)), k(Dr(gas.R)), k(Dr(gas.T_star_n)), k(Dr(gas.T_star_o)), k(Dr(gas.tau_n)),
         k(Dr(gas.tau_o)), k(Da(gas.tag_tolerance)), k(I(level)), k(Dr(gas.c)), k(Dr(gas.c_lr)), k(I(5))]
    return a, keep


def run(fn_name, variant_id, arrays, on_device, geometry):
    lib = _lib.load()
    w, p, T, src = arrays
    box, g, xlo, dx, gas, level = geometry
    ny, nx = box[3] - box[2] + 1, box[1] - box[0] + 1
    if on_device:
        w_, p_, T_, s_ = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (w, p, T, src))
        KK = torch.zeros((5, ny, nx), dtype=torch.float64, device="cuda")
        ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    else:
        w_, p_, T_, s_ = (np.ascontiguousarray(a) for a in (w, p, T, src))
        KK = np.zeros((5, ny, nx))
        ptr = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    a, keep = args_for(box, g, xlo, dx, w_, s_, KK, p_, T_, gas, level, ptr)
    torch.cuda.synchronize()
    if fn_name == "enl_rhs_variant":
        rc = lib.enl_rhs_variant(variant_id, *a, None)
        assert rc == 0, rc
    else:
        getattr(lib, fn_name)(*a)
        _lib.check()
    torch.cuda.synchronize()
    return KK.cpu().numpy() if on_device else KK


def main(argv):
    expected = argv[0]
    aliases = argv[1:]
    gas = Gas()
    nx, ny, g = 48, 40, 4
    w, p, T, src = random_state(nx, ny, g, seed=7)
    geometry = ((0, nx - 1, 0, ny - 1), g, (0.0, 0.0), (1.0 / 512, 1.0 / 384), gas, 0)
    arrays = (w, p, T, src)
    bad = 0
    checks = [("rhs_", expected)] + [("rhs_%s_" % v, v) for v in aliases]
    for fn, v in checks:
        for on_device in (True, False):
            ref = run("enl_rhs_variant", VARIANTS.index(v), arrays, True, geometry)
            got = run(fn, None, arrays, on_device, geometry)
            same = np.array_equal(ref.view(np.uint64), got.view(np.uint64))
            print("%s %s arrays vs enl_rhs_variant(%s): %s" % (fn, "device" if on_device else "host", v,
                                                                "bitwise equal" if same else "DIFFERENT"),
                  flush=True)
            bad += 0 if same else 1
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
