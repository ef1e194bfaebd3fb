"""The learned WENO/DRP switch on the device.

* The network's forward pass, loss and gradients match the CPU restatement
  (oracle/switch_mlp.py, itself pinned by finite differences), and so do the
  features of real channel states.
* With every candidate left to the sensor (lo = 0, hi = 1) the learned path
  is bit-identical to the plain fused path: the prefilter only ever marks
  cells the sensor itself calls DRP, and the decisions reach the split
  kernel unchanged.
* Trained on plain-sensor maps of seeded channels, the network reproduces
  the sensor on held-out channels, and the hybrid solution with it stays
  close to the plain hybrid solution; the test reports what the trained
  network decides on its own compared with the untrained one.
"""
import numpy as np
import pytest
import torch

from tests import devops as D

pytestmark = pytest.mark.gpu

if D.DRY:
    pytest.skip("the learned switch has no CPU stand-in", allow_module_level=True)

from enlightning_amd import switch as S  # noqa: E402
from enlightning_amd.gas import Gas, sensor_eps  # noqa: E402
from oracle import switch_mlp as M  # noqa: E402

SMALL = S.ChannelWorkload("channel_test_128", n=128, L=32.0, seed=0, cfl=0.2, segments=8, seg_len=4.0,
                          dp_range=(2.0, 10.0))


def bits_to_float(X):
    u = X.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def test_forward_loss_gradients_match_oracle():
    rng = np.random.default_rng(3)
    params = S.init_params(5) + rng.normal(scale=0.05, size=S.N_PARAM).astype(np.float32)
    sw = S.Switch(params)
    n = 3000
    xf = M.bf16(rng.normal(size=(n, 100)))
    Xbits = (xf.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    X = torch.from_numpy(Xbits.copy()).cuda()
    y = (rng.random(n) < 0.3).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=n).astype(np.float32)
    p = sw.forward(X).cpu().numpy()
    z = M.forward(params, xf)[0]
    # the first hidden layer is rounded to BF16: where FP32 and FP64 sums
    # straddle a rounding boundary that unit moves by one BF16 ulp, so a few
    # samples differ by up to ~1e-3; the rest agree to FP32 accuracy
    e = np.abs(p - 1 / (1 + np.exp(-z)))
    assert np.percentile(e, 99) <= 2e-5 and e.max() <= 5e-3
    g, loss = sw.gradients(X, torch.from_numpy(y).cuda(), torch.from_numpy(w).cuda())
    g = g.cpu().numpy()
    g_ref = M.gradients(params, xf, y.astype(np.float64), w.astype(np.float64))
    assert abs(loss - M.loss(params, xf, y, w)) <= 1e-4 * max(1.0, abs(loss))
    ge = np.abs(g - g_ref) / np.abs(g_ref).max()
    assert np.percentile(ge, 99) <= 2e-4 and ge.max() <= 1e-2
    sw.close()


def test_features_match_oracle():
    gas = Gas()
    dp = S.channel_patch(SMALL, gas, seed=11)
    dp.advance_fused(5, SMALL.dt(gas))
    S.P.fill_physical_bc(dp.box, dp.g, (1, 1, 1, 1), dp.reflect, dp.w, dp.p, dp.T)
    cand = torch.nonzero(S.candidates(dp).flatten()).flatten()
    assert cand.numel() > 50
    cells = cand[torch.randperm(cand.numel(), device=cand.device)[:200]]
    sw = S.Switch()
    X = bits_to_float(sw.features(dp, cells))
    w = dp.w.cpu().numpy()
    pr = dp.p.cpu().numpy()
    ref = M.features(w[0], pr, cells.cpu().numpy(), SMALL.n, SMALL.g, sensor_eps(gas.tag_tolerance[2]), gas.c)
    # asinh on the device and in numpy may differ by one float32 ulp, which
    # can move a BF16 rounding by one BF16 ulp
    assert np.all(np.abs(X - ref) <= 2.0 ** -7 * np.abs(ref) + 1e-30)
    sw.close()


def test_undecided_switch_equals_plain():
    gas = Gas()
    a = S.channel_patch(SMALL, gas, seed=12)
    b = S.channel_patch(SMALL, gas, seed=12)
    dt = SMALL.dt(gas)
    sw = S.Switch()
    a.advance_fused(12, dt)
    counts = sw.advance(b, 12, dt, lo=0.0, hi=1.0)
    assert counts[0] > 0 and counts[1] == 0 and counts[2] == counts[0]
    for x, y in zip(a.to_host()[:3], b.to_host()[:3]):
        g = SMALL.g
        assert np.array_equal(x[..., g:-g, g:-g], y[..., g:-g, g:-g])
    sw.close()


def test_learned_path_needs_three_ghost_cells():
    """A cell's features read one cell past its 5 x 5 window: the learned
    advance refuses a patch with two ghost cells (enough for hybrid_low_llf
    on its own) before any launch."""
    import dataclasses
    gas = Gas()
    wl = dataclasses.replace(SMALL, g=2)
    b = S.channel_patch(wl, gas, seed=12)
    sw = S.Switch()
    with pytest.raises(RuntimeError, match="3 ghost cells"):
        sw.advance(b, 1, wl.dt(gas), lo=0.0, hi=1.0, variant="hybrid_low_llf")
    sw.close()


@pytest.fixture(scope="module")
def trained():
    gas = Gas()
    sw = S.Switch()
    X, y, w = S.build_samples(sw, SMALL, gas, seeds=range(20, 32), snapshots=(0, 10, 30), per_snapshot=4000)
    Xv, yv, wv = S.build_samples(sw, SMALL, gas, seeds=range(90, 93), snapshots=(0, 10, 30), per_snapshot=4000)
    untrained = S.evaluate(sw, Xv, yv)
    losses = S.train(sw, X, y, w, epochs=6, batch=2048, lr=3e-3)
    return sw, losses, untrained, S.evaluate(sw, Xv, yv)


def test_training_reduces_loss_and_learns_the_sensor(trained):
    sw, losses, untrained, ev = trained
    print("losses", losses, "untrained", untrained, "trained", ev)
    assert losses[-1] < 0.8 * losses[0]
    assert ev["accuracy"] >= 0.9
    assert ev["accuracy"] > untrained["accuracy"]
    # the network settles most rows itself, where the untrained one leaves
    # them to the sensor, and misses fewer shocks
    assert ev["to_sensor"] < 0.5 * untrained["to_sensor"]
    assert ev["missed_shocks"] <= untrained["missed_shocks"]


def test_learned_hybrid_close_to_plain(trained):
    sw, _, _, _ = trained
    gas = Gas()
    a = S.channel_patch(SMALL, gas, seed=95)
    b = S.channel_patch(SMALL, gas, seed=95)
    dt = SMALL.dt(gas)
    a.advance_fused(20, dt)
    cand, weno, undecided = sw.advance(b, 20, dt, lo=0.2, hi=0.8)
    print("candidates %d, network WENO %d, to the sensor %d" % (cand, weno, undecided))
    assert cand > 0 and undecided < cand
    wa, wb = a.to_host()[0], b.to_host()[0]
    g = SMALL.g
    for k in range(5):
        x, y = wa[k, g:-g, g:-g], wb[k, g:-g, g:-g]
        ref = x - x.mean() if k in (0, 3, 4) else x
        err = np.sqrt(np.sum((x - y) ** 2) / max(np.sum(ref ** 2), 1e-300))
        assert err <= 5e-2, (k, err)


def test_matrix_core_decisions_match_forward(trained):
    """The per-stage decisions (matrix-core pass over the compacted
    candidates) against the plain forward pass on the same features: every
    non-candidate is DRP, and candidates get the thresholded probability
    except within 1e-3 of a threshold (summation order differs)."""
    sw, _, _, _ = trained
    gas = Gas()
    dp = S.channel_patch(SMALL, gas, seed=96)
    dp.advance_fused(8, SMALL.dt(gas))
    S.P.fill_physical_bc(dp.box, dp.g, (1, 1, 1, 1), dp.reflect, dp.w, dp.p, dp.T)
    lo, hi = 0.2, 0.8
    d, counts = sw.decide(dp, lo, hi)
    cand = S.candidates(dp).flatten()
    d = d.flatten()
    assert counts[0] == int(cand.sum().item())
    assert int((d[~cand] != 0).sum().item()) == 0
    cells = torch.nonzero(cand).flatten()
    p = sw.forward(sw.features(dp, cells))
    ref = torch.where(p > hi, 1, torch.where(p < lo, 0, 2)).to(torch.uint8)
    got = d[cells]
    bad = got != ref
    near = ((p - lo).abs() < 1e-3) | ((p - hi).abs() < 1e-3)
    assert int((bad & ~near).sum().item()) == 0
    assert counts[1] == int((got == 1).sum().item()) and counts[2] == int((got == 2).sum().item())
