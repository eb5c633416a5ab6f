"""Algorithmic work model of the FP64 chemistry step (for roofline fractions).

Per integrated cell-step the work is
    F = F_SETUP + NFE*F_RHS + NJE*F_JAC + NCORR*F_CORR + NATT*F_STEP
with SLATEC's counters (NFE includes the N extra right-hand sides of every
finite-difference Jacobian, slatec/ddpst.f90:110; NJE per Jacobian; NCORR per
corrector iteration, slatec/ddcor.f90; NATT per step attempt, slatec/ddstp.f90:107).
The constants are the KEROMNES figures (9 species, 24 reactions):

* F_RHS   = 285  mass-action RHS: 95 concentration products + 2 x 95 accumulations
                 (chemical_kinetics_solver.f90:509-536)
* F_JAC   = 866  FD Jacobian bookkeeping + P = I - h*l0*J + 9x9 DGEFA, without its 9 RHS
* F_CORR  = 220  DDCOR: residual, DGESL, update, scaled norm
* F_STEP  = 375  predictor/retraction, error test, step and order selection (300-450)
* F_SETUP = 2081 rates, [M], thermo (1955) + 126 transcendentals counted as 1 FLOP each

Bytes per cell-step at the reference boundary: 89 in (T, rho, 9 Y as FP64 +
int8 marker) + 152 out (9 Y_prod, 9 Y_prod2, E_f_prod) = 241 B.
"""
F_RHS = 285
F_JAC = 866
F_CORR = 220
F_STEP = 375
F_SETUP = 2081

BYTES_BOUNDARY = 241          # chemistry step, per inner cell
BYTES_CLOSURE = 8 * (1 + 1 + 1 + 9 + 9 + 1 + 1 + 9 + 1 + 1) + 1   # rho,E_f,E_f_prod,Y,Y_prod,T in; E_f,Y,T,p out; bc


def step_flops(stats):
    """Algorithmic FLOPs of one nrg_chem_step call from its nrg_chem_stats."""
    return (stats.cells_integrated * F_SETUP + stats.nfe_sum * F_RHS + stats.nje_sum * F_JAC
            + stats.ncorr_sum * F_CORR + stats.natt_sum * F_STEP)


def step_bytes(stats):
    return stats.cells_inner * BYTES_BOUNDARY
