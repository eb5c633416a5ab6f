/* Low-Mach (FDS) coupling of the chemistry step: the per-cell consumers of the
 * chemistry outputs in NRG's low-Mach solver (computing_module/src/
 * current_build/fds_low_mach_solver.f90) and its equation of state
 * (table_approximated_real_gas.f90:570-677), on device-resident fields.
 *
 * In that solver T comes from the static pressure and the density, not from
 * the energy; the chemistry step's Y_prod and E_f_prod enter the species
 * predictor/corrector and the velocity divergence.  The flux (gas-dynamic)
 * terms of those loops stay with the flow solver: the species update takes
 * them as an input field, and the divergence kernel returns the reactive
 * part only.
 *
 * Conventions as in nrg_chem.h: fields are FP64 device arrays with NRG's
 * ghosted column-major layout described by `grid`; Y-like arguments are host
 * arrays of S device pointers; cells with bc != 0 are skipped (left
 * untouched), as in the reference loops.  Calls run on the handle's stream
 * and return after it has drained.
 */
#ifndef NRG_LOW_MACH_H
#define NRG_LOW_MACH_H

#include "nrg_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* apply_state_equation_low_mach_fds (table_approximated_real_gas.f90:570-677).
 * Per inner cell with bc == 0:
 *   W = 1/sum_s Y_s/W_s -> mix_mol_mass;  X_s = Y_s W / W_s;
 *   predictor: p_stat_old = p_stat; p_stat = p_stat + dp_stat_dt*dt
 *   corrector: p_stat = 0.5*(p_stat + p_stat_old + dp_stat_dt*dt)
 *   T = p_stat/rho/R*W;  cp = sum X_s cp_s(T);  gamma = cp/(cp - R);
 *   h_s = (sum X_s h_s(T) - sum X_s h_s(298.15))/W;  p = p_stat;
 *   v_s = sqrt(gamma p_stat/rho);  e_i = p_stat/rho/(gamma - 1);
 *   E_f = e_i + v_d^2/2 with d the LAST dimension (the reference's loop over
 *   dimensions overwrites E_f; ndim = 0 leaves E_f untouched).
 * Required: rho, Y, dp_stat_dt, p_stat, p_stat_old, T, mix_mol_mass (and v
 * when ndim > 0).  p, mixture_cp, gamma, h_s, v_s, e_i, E_f may be NULL (not
 * written). */
int nrg_low_mach_state_equation_dev(nrg_chem_t *h, const nrg_grid *grid, double time_step,
                                    int32_t predictor, const double *rho,
                                    const double *const *Y, const double *const *v, int32_t ndim,
                                    const double *dp_stat_dt, const int8_t *bc, double *p_stat,
                                    double *p_stat_old, double *T, double *mix_mol_mass, double *p,
                                    double *mixture_cp, double *gamma, double *h_s, double *v_s,
                                    double *e_i, double *E_f);

/* Species update of calculate_interm_Y_predictor (fds_low_mach_solver.f90:
 * 906-925) and calculate_interm_Y_corrector (:3103-3125).  Per inner cell with
 * bc == 0 and species s, with f = rhs[s] (the caller's flux divergence, zero
 * if rhs is NULL):
 *   predictor: r = rho_a Y_a + f dt (+ Y_prod_diff dt) (+ Y_prod_chem dt)
 *   corrector: r = 0.5 (rho_a Y_a + rho_b Y_b) + f (dt/2)
 *                  (+ 0.5 Y_prod_diff dt) (+ 0.5 Y_prod_chem dt)
 * (a NULL source is a switched-off term); then r = max(r, 0), sum = sum_s r,
 * and if sum > DBL_MIN: rho_out = sum, Y_out = r/sum (else both untouched).
 * In the reference's predictor rho_a/Y_a are rho/Y and the outputs rho_int/
 * Y_int; in its corrector rho_a/Y_a are rho_old/Y_old, rho_b/Y_b are
 * rho_int/Y_int and the outputs rho/Y.  Outputs may alias inputs.
 * Lagrangian-particle sources are not part of this call. */
int nrg_low_mach_species_dev(nrg_chem_t *h, const nrg_grid *grid, double time_step,
                             int32_t corrector, const double *rho_a, const double *const *Y_a,
                             const double *rho_b, const double *const *Y_b,
                             const double *const *rhs, const double *const *Y_prod_diff,
                             const double *const *Y_prod_chem, const int8_t *bc, double *rho_out,
                             double *const *Y_out);

/* Reactive part of div_v in calculate_divergence_v (fds_low_mach_solver.f90:
 * 1057-1162): with W = mix_mol_mass, X_s = Y_s W/W_s, cp = sum X_s cp_s(T),
 * N = rho cp T/W and hs_s = (h_s(T) - h_s(298.15))/W_s,
 *   div_v = E_f_prod_chem/N + sum_s (W/W_s/rho - hs_s/N) Y_prod_chem_s
 * in the reference's order of operations.  This is the reference's div_v
 * when its other source terms are switched off and the face velocities are
 * zero; otherwise the flow solver adds its terms to it. */
int nrg_low_mach_chem_divergence_dev(nrg_chem_t *h, const nrg_grid *grid, const double *rho,
                                     const double *T, const double *const *Y,
                                     const double *mix_mol_mass, const double *E_f_prod_chem,
                                     const double *const *Y_prod_chem, const int8_t *bc,
                                     double *div_v);

#ifdef __cplusplus
}
#endif
#endif /* NRG_LOW_MACH_H */
