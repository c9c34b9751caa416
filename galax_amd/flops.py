"""Algorithmic work per unit, for roofline accounting.

Unit: one attempted Runge-Kutta step of one orbit (an orbit-step).
Counting rule: add, sub, mul = 1; fma = 2; div, sqrt, rsqrt, log1p, sincos
= 1 algorithmic operation each.

  F_step  = S * F_rhs + 2 * 6 * (nnz(A) + nnz(b) + nnz(b - b_hat)) + F_ctrl
  F_dense = 2 * (D - 1) * (n_q + n_p) + 2 * 3 * (n_q + n_p) + 24
            (one save inside a step from the method's dense output: Horner
            evaluation of the n_q position and n_p velocity weight
            polynomials of degree D - 1, their combination with the stage
            accelerations, the final three FMAs per component, and th)
  F_save  = S * F_rhs + 2 * 6 * (nnz(A) + nnz(b))      (a step without error)

nnz values are those of the implemented tableaux (checked against the
oracle's tableau in the tests).  F_rhs counts the closed-form field as
implemented (galax_amd/csrc/gx_potential.hpp):
  shared radius: r^2 (5), r (2), 1/r (1)                        =  8
  Miyamoto-Nagai: z^2+b^2 (2), rsqrt (1), S (1), a+S (1),
    R^2+A^2 (5), rsqrt (1), GM/D^3 (3), a_x, a_y (4), a_z (4)   = 22
  NFW: log1p(r/r_s) (2), 1/(r+r_s) (2), g (5), apply (6)         = 15
  Hernquist: 1/(r+r_s) (2), g (3), apply (6)                     = 11
  Plummer: rsqrt(r^2+b^2) (2), g (3), apply (6)                  = 11
  Kepler: g (3), apply (6)                                       =  9
so the Milky Way model (MN + NFW + 2 Hernquist) costs 8 + 22 + 15 + 22 = 67.
"""

from __future__ import annotations

NNZ = {
    # method: (stages, nnz(A), nnz(b), nnz(b - b_hat))
    "dopri8": (13, 59, 9, 9),
    "dopri5": (7, 20, 5, 6),
}
# dense output: (D, n_q, n_p) -- polynomial length, nonzero position and
# velocity weight rows over the step's stage slots (13 stages and the end
# stage for Dopri8, 7 for Dopri5; checked against the oracle in the tests)
NNZ_DENSE = {"dopri8": (6, 11, 10), "dopri5": (4, 5, 6)}
F_CTRL = 60
F_RHS = {"kepler": 9, "hernquist": 11, "plummer": 11, "nfw": 15, "miyamoto_nagai": 22, "shared_radius": 8}
F_RHS_MW = F_RHS["shared_radius"] + F_RHS["miyamoto_nagai"] + F_RHS["nfw"] + 2 * F_RHS["hernquist"]


def f_step(method: str = "dopri8", f_rhs: int = F_RHS_MW) -> int:
    S, a, b, e = NNZ[method]
    return S * f_rhs + 12 * (a + b + e) + F_CTRL


def f_dense(method: str = "dopri8") -> int:
    D, nq, npv = NNZ_DENSE[method]
    return 2 * (D - 1) * (nq + npv) + 2 * 3 * (nq + npv) + 24


def f_save(method: str = "dopri8", f_rhs: int = F_RHS_MW) -> int:
    S, a, b, _ = NNZ[method]
    return S * f_rhs + 12 * (a + b)


# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0

# HBM bytes of one k_integrate pass, measured with rocprofv3 (FETCH_SIZE and
# WRITE_SIZE, each scaled by the same counter's reading for a plain 1 GiB
# copy: x2.0 and x1.0), for bench.py's workloads at 1,048,576 orbits and 64
# saves.  The throughput workload's measurement and commands are kept in
# profiles/k_integrate_mw_uneven_traffic.txt.
MEASURED_TRAFFIC_BYTES = {"uneven": 3.8973e9}
