"""Source injection on the device (enl_add_source) against the oracle, and
the library's lightning-file reader against the oracle's.

The per-segment quantities (angle, normal slopes, sin, cos) and the DRIVEN
factor are formed on the host with the reference's expressions, so the
cell selection is the oracle's exactly; the per-cell exp and sin are the
device's, within a few ulps of the host's: 1e-14 of the amplitude.
"""
import numpy as np
import pytest

import oracle
from enlightning_amd.source import DRIVEN, LIGHTNING, POINT, PULSE, SINE, Source
from tests import devops as D

pytestmark = pytest.mark.gpu

if D.DRY:
    pytest.skip("source injection has no CPU stand-in", allow_module_level=True)

from enlightning_amd import source as S  # noqa: E402


def channel(seed, n=30, seg=0.02):
    """A tortuous channel of n segments from (0.5, 0.9) downwards."""
    rng = np.random.default_rng(seed)
    p = np.array([0.5, 0.9])
    out = []
    for _ in range(n):
        ang = np.deg2rad(-90 + rng.normal(0, 25))
        q = p + seg * np.array([np.cos(ang), np.sin(ang)])
        out.append([p[0], p[1], q[0], q[1]])
        p = q
    return np.array(out, dtype=np.float32).astype(np.float64)  # single-precision values, as read from file


CASES = {
    "point": Source(mode=PULSE, type=POINT, amp=5e5, spread=0.01, pos=(0.3, 0.4)),
    "sine": Source(mode=PULSE, type=SINE, amp=2e3, spread=0.02, pos=(0.5, 0.5), sine_amp=0.05, sine_freq=20.0),
    "lightning": Source(mode=PULSE, type=LIGHTNING, amp=4e4, spread=0.004, coords=channel(1)),
    "driven": Source(mode=DRIVEN, type=POINT, amp=5e5, spread=0.01, pos=(0.31, 0.47), omega=1000.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_add_source_matches_oracle(case):
    src = CASES[case]
    # a 200 x 150 patch at (40, 60) of a 256^2 grid on the unit square, in a
    # ghosted array (row stride 208, interior at (4, 4))
    box = (40, 239, 60, 209)
    dx = (1 / 256, 1 / 256)
    xlo = (40 / 256, 60 / 256)
    nx, ny, g = 200, 150, 4
    rng = np.random.default_rng(3)
    base = rng.uniform(0, 0.3 * src.amp, size=(ny + 2 * g, nx + 2 * g))  # LIGHTNING merges into it
    ref = base[g:-g, g:-g].copy()
    oracle.add_source(box, xlo, dx, src, 12, 3, 2e-5, ref)
    dev = D.dev(base)
    S.add_source(box, xlo, dx, src, 12, 3, 2e-5, dev[g:, g:], sy=nx + 2 * g)
    got = D.host(dev)
    assert np.abs(got[g:-g, g:-g] - ref).max() <= 1e-14 * src.amp, case
    # ghosts untouched
    mask = np.ones_like(base, bool)
    mask[g:-g, g:-g] = False
    assert np.array_equal(got[mask], base[mask])
    assert np.abs(ref).max() > 0.1 * src.amp * (0.3 if case == "driven" else 1)


def test_read_lightning_matches_oracle(tmp_path):
    f = tmp_path / "channel.txt"
    c = channel(5, n=12)
    f.write_text("".join("%.6f %.6f %.6f %.6f\n" % tuple(r) for r in c) + "\n")
    a = S.read_lightning(f)
    b = oracle.read_lightning(f)
    assert a.shape == (13, 4) and np.array_equal(a, b)


def test_mics_match_oracle():
    """Microphones inside, outside and on the lower bounds of a patch in a
    ghosted pressure array."""
    box = (16, 79, 8, 55)
    dx = (1 / 128, 1 / 128)
    xlo = (16 / 128, 8 / 128)
    xhi = (80 / 128, 56 / 128)
    nx, ny, g = 64, 48, 4
    rng = np.random.default_rng(8)
    p = 101325.0 + rng.standard_normal((ny + 2 * g, nx + 2 * g))
    mx = np.concatenate([rng.uniform(0.0, 0.7, 40), [xlo[0], 0.3]])
    my = np.concatenate([rng.uniform(0.0, 0.5, 40), [0.2, xlo[1]]])
    ref, rhit = oracle.sample_mics(box, xlo, xhi, dx, mx, my, p[g:-g, g:-g].copy())
    dp = D.dev(p)
    got, ghit = S.sample_mics(box, xlo, xhi, dx, mx, my, dp[g:, g:], sy=nx + 2 * g)
    got, ghit = D.host(got), D.host(ghit)
    assert np.array_equal(ghit, rhit) and rhit.sum() > 10 and rhit.sum() < len(mx)
    assert np.array_equal(got[ghit == 1], ref[rhit == 1])
