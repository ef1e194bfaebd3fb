"""Algorithmic floating-point operations per cell of the reference kernels.

The reference's own arithmetic, counted operation by operation: +, -, *, /,
fabs and comparisons count 1 each, negation and copies 0; exp and log are
listed apart and not included.  Work done once per call or per row (the
sensor's epsilon, update_state's row height) is not in the per-cell
figures.  The numbers are what a counting-double run
of the CPU restatement of fortran/rhs_*.m4, runge_kutta.m4, update_state.m4
and tag_cells.m4 gives on single cells (tests/test_flop_counts.py checks
this table against it).  They count the reference's form of the arithmetic:
e.g. every flux evaluation recomputes its velocity quotients (168 of a DRP
cell's 590 operations are divisions), where the HIP kernels form each
quotient once per staged cell, so a kernel's achieved FLOP/s below means
reference-equivalent work per second, not issued instructions.

The rhs entries are keyed by (variant, level, scheme): level 0 carries the
hybrid sensor (rhs_hybrid_high_llf.m4:239-256) where the variant has one;
level 1 stands for every level above 0, where the hybrid variants take
WENO without the sensor.
"""

RHS_CELL_FLOPS = {
    ("hybrid_high_llf", 0, "drp"): 590,
    ("hybrid_high_llf", 0, "weno"): 3812,
    ("hybrid_high_llf", 1, "weno"): 3740,
    ("hybrid_high", 0, "drp"): 590,
    ("hybrid_high", 0, "weno"): 3686,
    ("hybrid_high", 1, "weno"): 3614,
    ("hybrid_low_llf", 0, "drp"): 590,
    ("hybrid_low_llf", 0, "weno"): 1896,
    ("hybrid_low_llf", 1, "weno"): 1824,
    ("hybrid_low", 0, "drp"): 590,
    ("hybrid_low", 0, "weno"): 1806,
    ("hybrid_low", 1, "weno"): 1734,
    ("weno_high_llf", 0, "weno"): 3740,
    ("weno_high_llf", 1, "weno"): 3740,
    ("weno_high", 0, "weno"): 3614,
    ("weno_high", 1, "weno"): 3614,
    ("weno_low_llf", 0, "weno"): 1632,
    ("weno_low_llf", 1, "weno"): 1632,
    ("weno_low", 0, "weno"): 1542,
    ("weno_low", 1, "weno"): 1542,
    ("drp", 0, "drp"): 518,
    ("drp", 1, "drp"): 518,
}
# exp count of every rhs cell (the N2 relaxation's c_vn, :217)
RHS_CELL_EXP = 1

RUNGE_KUTTA_CELL_FLOPS = 30    # runge_kutta.m4:57-66, 5 components
UPDATE_STATE_CELL_FLOPS = 23   # update_state.m4:87-107, plus 1 exp and 1 log
TAG_CELLS_CELL_FLOPS = 48      # tag_cells.m4:58-108

FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X, dense FP64 vector


def rhs_flops(variant, level, n_drp, n_weno):
    """Algorithmic FLOPs of one rhs launch over n_drp DRP and n_weno WENO cells."""
    lev = 0 if level == 0 else 1
    f = 0
    if n_drp:
        f += n_drp * RHS_CELL_FLOPS[(variant, lev, "drp")]
    if n_weno:
        f += n_weno * RHS_CELL_FLOPS[(variant, lev, "weno")]
    return f


def stage_flops(variant, level, n_drp, n_weno):
    """One fused stage (rhs + runge_kutta + update_state) over those cells."""
    return rhs_flops(variant, level, n_drp, n_weno) + (n_drp + n_weno) * (RUNGE_KUTTA_CELL_FLOPS +
                                                                           UPDATE_STATE_CELL_FLOPS)
