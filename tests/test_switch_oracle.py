"""The learned switch's CPU restatement (oracle/switch_mlp.py), pinned by
what does not come from its own formulas: central finite differences of the
loss for the gradients, closed forms for BF16 rounding and the features."""
import numpy as np

from oracle import switch_mlp as M


def test_bf16_rounding_closed_forms():
    # exactly representable values stay; ties go to even; 1 + 2^-8 rounds to 1
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0           # tie, even mantissa
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie, odd -> up
    assert M.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    assert M.bf16(-2.5) == -2.5
    x = np.random.default_rng(0).normal(size=1000)
    assert np.all(np.abs(M.bf16(x) - x) <= np.abs(x) * 2.0 ** -8)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    params = M.bf16(rng.normal(scale=0.2, size=M.N_PARAM))
    X = M.bf16(rng.normal(size=(40, M.N_IN)))
    y = (rng.random(40) < 0.4).astype(float)
    w = rng.uniform(0.5, 2.0, size=40)
    g = M.gradients(params, X, y, w, round_h1=False)
    idx = np.concatenate([rng.choice(M.N_PARAM, 300, replace=False), np.arange(M.N_PARAM - 97, M.N_PARAM)])
    h = 1e-6
    fd = np.empty(len(idx))
    for n, q in enumerate(idx):
        pp = params.copy(); pp[q] += h
        pm = params.copy(); pm[q] -= h
        fd[n] = (M.loss_exact_params(pp, X, y, w) - M.loss_exact_params(pm, X, y, w)) / (2 * h)
    err = np.abs(fd - g[idx])
    scale = np.abs(g[idx]).max()
    # a few parameters may straddle a ReLU kink; the bulk must agree closely
    assert np.median(err) <= 1e-8 * max(scale, 1.0)
    assert np.mean(err <= 1e-6 * scale + 1e-10) >= 0.97


def test_features_closed_forms():
    # linear rho and p ramps: every x difference equals the slope per cell,
    # every y difference the y slope
    n, g = 8, 4
    N = n + 2 * g
    jj, ii = np.mgrid[0:N, 0:N]
    eps, c = 4.5e-3, 343.0
    rho = 1.2 + 0.01 * ii - 0.02 * jj
    p = 101325.0 + 50.0 * ii + 25.0 * jj
    X = M.features(rho, p, [0, 9, 63], n, g, eps, c)
    want = M.bf16(np.array([np.float32(np.arcsinh(np.float64(np.float32(v)))) for v in (
        np.float32(0.01) * np.float32(1 / np.sqrt(eps)), np.float32(-0.02) * np.float32(1 / np.sqrt(eps)),
        np.float32(50.0) * np.float32(1 / (c * c * np.sqrt(eps))), np.float32(25.0) * np.float32(1 / (c * c * np.sqrt(eps))))]))
    for row in X:
        assert np.allclose(row.reshape(25, 4), want[None, :], rtol=2 ** -7, atol=0)
