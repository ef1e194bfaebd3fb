"""Physical constants of the Enlightning{constants} block, mapped onto the
workloads' ideal-gas description.

Reference: input.txt:40-59 (constants), input.txt:4 (tag_tolerance),
enlightning.C:218-234 (tau_n, tau_o from T0, p0 and relative humidity).

Workload mapping (stated with every result): gamma = 1.4, rho0 = 1.2 kg/m^3,
p0 = 101325 Pa; c = sqrt(gamma*p0/rho0); T0 = p0/(rho0*R) with R = 287.06;
c_lr = 0; the remaining constants are those of input.txt.
"""
import math
from dataclasses import dataclass, field, asdict

import numpy as np

from ._lib import EnlGas


def relaxation_times(T0, p0, RH):
    """tau_n, tau_o of N2/O2 vibrational relaxation, enlightning.C:218-234
    (ISO 9613-1 relaxation frequencies with T_ref = T0, p_ref = p0)."""
    T_ref, p_ref = T0, p0
    zeta = (10.79586 * (1 - (273.16 / T0)) - 5.02808 * math.log10(T0 / 273.16)
            + (1.50474e-4) * (1 - 10 ** (-8.29692 * ((T0 / 273.16) - 1)))
            - (4.2873e-4) * (1 - 10 ** (-4.76955 * ((273.16 / T0) - 1)))
            - 2.2195983)
    p_vp = p_ref * 10 ** zeta
    h = (RH * p_vp) / p0
    eta = 4.17 * ((T_ref / T0) ** (1.0 / 3.0) - 1)
    f_n = (p0 / p_ref) * ((T_ref / T0) ** 0.5 * (9 + 280 * h * math.exp(-eta)))
    f_o = (p0 / p_ref) * (24 + (4.04e4) * h * ((0.02 + h) / (0.391 + h)))
    return 1 / (2 * math.pi * f_n), 1 / (2 * math.pi * f_o)


@dataclass
class Gas:
    gamma1: float = 1.4
    rho0: float = 1.2
    p0: float = 101325.0
    R: float = 287.06
    c_lr: float = 0.0
    weno_alpha: float = 360.0
    RH: float = 20.0
    c_p: float = 1010.0
    c_v: float = 720.4
    T_star_n: float = 3352.0
    T_star_o: float = 2239.0
    mu: float = 1.846e-5
    muB: float = 1.1076e-5
    kappa: float = 0.02624
    R_tilde: float = 8314.32
    M: float = 28.9644
    tag_tolerance: tuple = (1.5, 1e-2, 5e-3)
    c: float = field(default=None)
    T0: float = field(default=None)
    tau_n: float = field(default=None)
    tau_o: float = field(default=None)

    def __post_init__(self):
        if self.c is None:
            self.c = math.sqrt(self.gamma1 * self.p0 / self.rho0)
        if self.T0 is None:
            self.T0 = self.p0 / (self.rho0 * self.R)
        if self.tau_n is None or self.tau_o is None:
            self.tau_n, self.tau_o = relaxation_times(self.T0, self.p0, self.RH)

    def to_struct(self):
        s = EnlGas()
        for name, _ in EnlGas._fields_:
            if name == "tag_tolerance":
                for i in range(3):
                    s.tag_tolerance[i] = self.tag_tolerance[i]
            else:
                setattr(s, name, getattr(self, name))
        return s

    def as_dict(self):
        return asdict(self)


# SSP-RK3 (Shu-Osher) coefficients in the method-of-lines form
# w_n = alpha_1*w + alpha_2*w_n + beta*dt*K (input.txt:130-133 lists them to
# 10 digits; the exact thirds of matlab/hybrid.m:132-145 are used here).
RK3 = ((1.0, 0.75, 1.0 / 3.0), (0.0, 0.25, 2.0 / 3.0), (1.0, 0.25, 2.0 / 3.0))


def sensor_eps(tol):
    """The hybrid sensor's epsilon from tag_tolerance(2): 0.9f * tol * xi^2 /
    (1 - 0.9f * tol^2) with xi = 1 (rhs_hybrid_high_llf.m4:114-115; the
    literal 0.9 is single precision in the Fortran)."""
    c09 = float(np.float32(0.9))
    return c09 * tol / (1.0 - c09 * tol * tol)
