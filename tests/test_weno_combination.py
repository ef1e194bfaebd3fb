"""The WENO combination of the device kernels (enl_device.h weno_combine)
against the reference's expression (rhs_hybrid_high_llf.m4:365-433), both
evaluated in float64 and compared with an extended-precision (x86 80-bit
long double) evaluation of the reference's expression.

The device forms sum_k (zeta_k / sigma) f_k, zeta_k = c_k / D_k, as
f_1 + (sum_k n_k (f_k - f_1)) / (sum_k n_k) with n_k = c_k * prod_{j != k} D_j:
one division instead of six.  On the reconstructed flux differences
p1 + m1 - p2 - m2 of a stratified exponential profile (the split fluxes'
large common part alpha*w over rounding-sized differences), a smooth wave and
a shock, its RMS error must not exceed the reference expression's.  This
is a model in numpy of the device arithmetic (no FMA contraction), not a
test of the device.
"""
import numpy as np
import pytest

L = np.longdouble
pytestmark = pytest.mark.skipif(np.finfo(L).eps >= np.finfo(np.float64).eps, reason="needs extended long double")

C13_12, EPS, C1_6 = (L(np.float32(v)) for v in (13.0 / 12.0, 10.0e-6, 1.0 / 6.0))
W01, W06, W03 = (L(np.float32(v)) for v in (0.1, 0.6, 0.3))


def parts(a, b, c, d, e, T, plus):
    b0 = T(C13_12) * (c - 2 * d + e) ** 2 + T(0.25) * (3 * c - 4 * d + e) ** 2
    b1 = T(C13_12) * (b - 2 * c + d) ** 2 + T(0.25) * (b - d) ** 2
    b2 = T(C13_12) * (a - 2 * b + c) ** 2 + T(0.25) * (a - 4 * b + 3 * c) ** 2
    if plus:  # wenop
        cs = (T(W03), T(W06), T(W01))
        f = (T(C1_6) * (2 * c + 5 * d - e), T(C1_6) * (-b + 5 * c + 2 * d), T(C1_6) * (2 * a - 7 * b + 11 * c))
    else:  # wenom
        cs = (T(W01), T(W06), T(W03))
        f = (T(C1_6) * (11 * c - 7 * d + 2 * e), T(C1_6) * (2 * b + 5 * c - d), T(C1_6) * (-a + 5 * b + 2 * c))
    D = [(T(EPS) + x) ** 2 for x in (b0, b1, b2)]
    return cs, D, f


def reference(cs, D, f):
    z = [c / d for c, d in zip(cs, D)]
    s = z[0] + z[1] + z[2]
    return (z[0] / s) * f[0] + (z[1] / s) * f[1] + (z[2] / s) * f[2]


def shifted(cs, D, f):
    n = (cs[0] * (D[1] * D[2]), cs[1] * (D[0] * D[2]), cs[2] * (D[0] * D[1]))
    return f[1] + (n[0] * (f[0] - f[1]) + n[2] * (f[2] - f[1])) / (n[0] + n[1] + n[2])


def divergence(form, T, fp, fm):
    """p1 + m1 - p2 - m2 at every cell (rhs_hybrid_high_llf.m4:286-354)."""
    out = []
    for i in range(3, len(fp) - 3):
        p1 = form(*parts(*fp[i - 3:i + 2], T, True))
        m1 = form(*parts(*fm[i - 2:i + 3], T, False))
        p2 = form(*parts(*fp[i - 2:i + 3], T, True))
        m2 = form(*parts(*fm[i - 1:i + 4], T, False))
        out.append(p1 + m1 - p2 - m2)
    return np.array(out)


def profiles():
    y = (np.arange(-3, 300) + 0.5) * (2000.0 / 4096)
    w = 353.0 * np.exp(-y / 8000.0)  # rho*T_n of the stratified ambient
    yield "stratified", 0.5 * 343.0 * w, -0.5 * 343.0 * w
    x = np.arange(-3, 250) * 1.0
    f = 6e4 + 7 * np.sin(0.07 * x)
    yield "wave", 0.5 * (f + 1e3 * np.cos(0.05 * x)), 0.5 * (f - 1e3 * np.cos(0.05 * x))
    f = np.where(x > 120, 6e4, 1e4) + 50 * np.sin(0.3 * x)
    yield "shock", 0.65 * f, -0.35 * f


@pytest.mark.parametrize("name", ["stratified", "wave", "shock"])
def test_shifted_combination_is_no_less_accurate(name):
    fp, fm = next((a, b) for n, a, b in profiles() if n == name)
    exact = divergence(reference, L, fp.astype(L), fm.astype(L))
    err_ref = divergence(reference, np.float64, fp, fm).astype(L) - exact
    err_new = divergence(shifted, np.float64, fp, fm).astype(L) - exact
    rms = lambda e: float(np.sqrt(np.mean(e ** 2)))
    assert rms(err_new) <= rms(err_ref), (name, rms(err_new), rms(err_ref))

