/*
 * nrg_surrogate.h -- learned chemistry step with a conservation/confidence
 * gate and FP64 fallback (gfx950).
 *
 * No counterpart in the reference, which integrates every cell with SLATEC
 * (computing_module/src/current_build/chemical_kinetics_solver.f90:175-358).
 * nrg_chem_surrogate_step_dev() is a drop-in for nrg_chem_step_dev(): same
 * fields, same outputs (Y_prod, Y_prod2, E_f_prod), so the split-step
 * closure is unchanged.  Per cell:
 *   1. gated cells (T < t_gate or bc != 0) are handled as the FP64 step does;
 *   2. the network (nrg_mlp.h) maps features
 *        f = (T, ln p, ln(Y_s + y_floor) for s = 1..S, ln dt)
 *      normalised as x_i = (f_i - in_shift_i) * in_scale_i, to the increments
 *      of one split step, (dT, dY_1..dY_S)_j = y_j * out_scale_j + out_shift_j;
 *   3. the gate accepts a cell only if
 *        - T, p, dt lie inside the trained domain,
 *        - cells the network sees as inert (|dT| <= dT_freeze and every
 *          |dY_s| <= dY_freeze) are accepted with a zero increment, so network
 *          noise below its resolution never enters the state (a perturbed
 *          radical pool would make the FP64 integrator resolve a stiff
 *          relaxation back to the slow manifold in later steps);
 *        - the network's |dT| is at most dT_max (confidence: fast chemistry,
 *          where a wrong reaction progress conserves elements and energy and
 *          so passes the checks below, goes to the FP64 integrator),
 *        - with rate_tol > 0, the answer (network or frozen) agrees with the
 *          explicit estimate of the step, C_s(t+dt) - C_s = omega_s(C, T) dt
 *          from the FP64 rate constants and mass action at the cell's own
 *          state, to rate_tol in every active species' mass fraction (a
 *          state outside the training data that the network calls nearly
 *          inert, such as radical-rich burnt gas recombining, fails it);
 *        - the raw dY conserves every element to elem_tol (relative to the
 *          atoms it moves) + elem_floor [mol/kg],
 *        - after projecting dY onto the element-conserving subspace, no
 *          mass fraction falls below -y_neg_tol,
 *        - the network's dT agrees within energy_tol [K] with the temperature
 *          that energy conservation gives for the projected dY (the same
 *          source application and e->T inversion as nrg_chem_closure_dev);
 *      accepted cells get the projected increments as Y_prod, Y_prod2,
 *      E_f_prod (so chemistry conserves elements and energy exactly);
 *   4. every rejected cell is integrated by the FP64 Gear-BDF kernels of
 *      nrg_chem.h, exactly as nrg_chem_step_dev would.
 */
#ifndef NRG_SURROGATE_H
#define NRG_SURROGATE_H

#include <stdint.h>

#include "nrg_chem.h"
#include "nrg_mlp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NRG_SUR_MAX_IN 16
#define NRG_SUR_MAX_OUT 16
#define NRG_SUR_MAX_ELEMENTS 4

/* per-cell status of an accepted surrogate answer (extends NRG_CELL_*) */
#define NRG_CELL_SURROGATE 6

/*
 * gate.screen: which cells the network sees.
 *   NONE   every live cell (the gate's checks decide, as listed above);
 *   RATES  an explicit-rate screen runs first: a live cell whose explicit
 *          estimate of the step, dY_s = omega_s(C, T) dt W_s / rho, has every
 *          |dY_s| <= dY_freeze and implies |dT| <= dT_freeze (at the cell's
 *          frozen cv) is accepted with the zero increment without the network;
 *          the network and the gate then run on the other live cells only;
 *   ONLY   the same screen, and every cell it does not freeze goes to the
 *          FP64 integrator: the method without the network (mlp and norm may
 *          be NULL, p is not read).
 * RATES and ONLY need dT_freeze > 0.  The screen's decision is the FP64
 * estimate's: a first pass evaluates the estimate with FP32 exponentials and a
 * bound on its distance from the FP64 estimate, and decides each cell whose
 * decision that bound settles; the FP64 estimate decides the others
 * (stats.cells_screen_fp64).  gate.screen_fp64 = 1 sends every live cell to
 * the FP64 estimate (same decisions, for comparison).
 */
#define NRG_SUR_SCREEN_NONE 0
#define NRG_SUR_SCREEN_RATES 1
#define NRG_SUR_SCREEN_ONLY 2

typedef struct nrg_surrogate_norm {
  float in_shift[NRG_SUR_MAX_IN], in_scale[NRG_SUR_MAX_IN];     /* S + 3 used */
  float out_shift[NRG_SUR_MAX_OUT], out_scale[NRG_SUR_MAX_OUT]; /* S + 1 used */
  float y_floor;
  int32_t reserved[7];
} nrg_surrogate_norm;

typedef struct nrg_surrogate_gate {
  double t_min, t_max, p_min, p_max, dt_min, dt_max;
  double elem_tol, elem_floor;
  double y_neg_tol;
  double energy_tol;
  double dT_max;                                     /* K                              */
  double dT_freeze, dY_freeze;                       /* K, mass fraction (dT_freeze 0: off) */
  int32_t n_elements;                                /* rows of `elements` (<= 4)      */
  int32_t screen;                                    /* NRG_SUR_SCREEN_*, below        */
  double rate_tol;                                   /* mass fraction (0: off), below  */
  int32_t screen_fp64;                               /* 1: FP64 estimate for every cell */
  int32_t reserved[3];
  double elements[NRG_SUR_MAX_ELEMENTS][32];         /* atoms of element e in species s */
} nrg_surrogate_gate;

typedef struct nrg_surrogate_stats {
  int64_t cells_inner, cells_gated, cells_accepted, cells_fallback;
  int64_t reject_domain, reject_elements, reject_negative, reject_energy, reject_confidence;
  int64_t cells_frozen;       /* accepted with a zero increment (inert)                 */
  double infer_ms;            /* featurise + network, HIP events on the handle's stream */
  double gate_ms;
  double fallback_ms;         /* FP64 integration of the rejected cells                */
  nrg_chem_stats fallback;    /* counters of that integration                          */
  int64_t reject_rate;        /* rejected by the rate-consistency test (rate_tol)      */
  int64_t cells_network;      /* cells the network ran on                              */
  double screen_ms;           /* the explicit-rate screen (SCREEN_RATES, SCREEN_ONLY)   */
  int64_t cells_screen_fp64;  /* cells the screen decided with the FP64 estimate        */
} nrg_surrogate_stats;

/*
 * One chemistry step with the surrogate.  Arguments as nrg_chem_step_dev, plus
 * the trained network, its normalisation and the gate; `p` is required (the
 * network reads ln p; not with NRG_SUR_SCREEN_ONLY).  The FP64 fallback runs
 * with the handle's integrator options: a handle created for the learned step
 * with a looser eps and NRG_JAC_ANALYTIC (eps 1e-6 keeps a 200-step run within
 * 1e-2 relative T and 1e-3 major-species mass fraction of the reference) gives
 * a faster fallback; the screen and the gate do not depend on those options.
 * Returns the FP64 fallback's fatal codes as nrg_chem_step_dev does.
 */
int nrg_chem_surrogate_step_dev(nrg_chem_t *h, nrg_mlp_t *mlp, const nrg_surrogate_norm *norm,
                                const nrg_surrogate_gate *gate, const nrg_grid *grid, double dt,
                                const double *T, const double *p, const double *rho,
                                const double *const *Y, const int8_t *bc,
                                double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                                int32_t *cell_status, nrg_surrogate_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* NRG_SURROGATE_H */
