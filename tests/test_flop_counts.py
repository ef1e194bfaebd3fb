"""The algorithmic FLOP table the benchmark uses (enlightning_amd/flops.py)
against the counting-double restatement (oracle/flops.cc: the oracle's
source compiled with a double that counts its operations), and the counts
against independent tallies of the reference's expressions."""
import os
import subprocess

import pytest

from enlightning_amd import flops

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def counted():
    subprocess.run(["make", "-s", "-C", ORACLE, "flops"], check=True)
    out = subprocess.run([os.path.join(ORACLE, "flops")], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.splitlines():
        name, *vals = line.split()
        rows[name] = dict(zip(("add", "sub", "mul", "div", "abs", "cmp", "exp", "log", "flops"), map(int, vals)))
    return rows


def test_table_matches_counting_restatement(counted):
    for (variant, level, kind), n in flops.RHS_CELL_FLOPS.items():
        row = counted["rhs.%s.level%d.%s" % (variant, level, kind)]
        assert row["flops"] == n, (variant, level, kind)
        assert row["exp"] == flops.RHS_CELL_EXP and row["log"] == 0
    assert counted["runge_kutta"]["flops"] == flops.RUNGE_KUTTA_CELL_FLOPS
    assert counted["update_state"]["flops"] == flops.UPDATE_STATE_CELL_FLOPS
    assert (counted["update_state"]["exp"], counted["update_state"]["log"]) == (1, 1)
    assert counted["tag_cells"]["flops"] == flops.TAG_CELLS_CELL_FLOPS


def test_counts_against_hand_tallies(counted):
    # runge_kutta.m4:60-63 per component: alpha_1*ww + alpha_2*w_n + beta*dt*KK
    # = 4 multiplications and 2 additions, 5 components
    rk = counted["runge_kutta"]
    assert (rk["mul"], rk["add"], rk["flops"]) == (20, 10, 30)
    # one smoothness ratio (2|t1 t2| + ep)/(t1^2 + t2^2 + ep) and its test:
    # 2 differences, 4 multiplications (t1 t2, 2*, t1^2, t2^2), 1 fabs,
    # 3 additions, 1 division, 1 comparison = 12
    ratio = 12
    # tag_cells.m4:67-108: 4 ratios per cell
    tag = counted["tag_cells"]
    assert (tag["sub"], tag["mul"], tag["abs"], tag["add"], tag["div"], tag["cmp"]) == (8, 16, 4, 12, 4, 4)
    assert tag["flops"] == 4 * ratio
    # the hybrid sensor (rhs_hybrid_high_llf.m4:244-254): 6 ratios on top of
    # rhs_drp's cell, and on top of the sensor-free WENO cell above level 0
    assert counted["rhs.hybrid_high_llf.level0.drp"]["flops"] - counted["rhs.drp.level0.drp"]["flops"] == 6 * ratio
    assert (counted["rhs.hybrid_high_llf.level0.weno"]["flops"] -
            counted["rhs.hybrid_high_llf.level1.weno"]["flops"]) == 6 * ratio
    # the LLF window (rhs_hybrid_high_llf.m4:270-284): per q and direction a
    # division, fabs, |u| + c, c_lr*height and its addition, the height
    # (2 additions, 1 multiply) and a comparison = 9; 7 q for WENO5, 5 for WENO3
    assert (counted["rhs.weno_high_llf.level0.weno"]["flops"] -
            counted["rhs.weno_high.level0.weno"]["flops"]) == 7 * 2 * 9
    assert (counted["rhs.weno_low_llf.level0.weno"]["flops"] -
            counted["rhs.weno_low.level0.weno"]["flops"]) == 5 * 2 * 9
    # update_state.m4:90-105 per cell: (c + c_lr y)^2 (2 mul, 1 add),
    # (rho - rho0) twice, (gamma - 1), 2 rho0, the quotient, the square and
    # products (7 mul, 3 div, 3 add with + p0), T: 3 div, 1 sub, 2 mul
    us = counted["update_state"]
    assert (us["add"], us["sub"], us["mul"], us["div"], us["exp"], us["log"]) == (4, 4, 9, 6, 1, 1)
