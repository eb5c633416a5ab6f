/*
 * nrg_oracle.h -- CPU restatement of NRG's chemistry step, for TESTS ONLY.
 *
 * This is the checker that the HIP path is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library (nrg_amd/libnrg_chem.so) never links or calls it.
 *
 * Every function restates the reference algorithm in plain C, keeping the
 * reference's floating-point evaluation order (built with -ffp-contract=off,
 * glibc libm, as the reference is built).  Citations are to
 * yakovenko-ivan/NRG; "sl/" abbreviates computing_module/src/third_party/slatec/.
 *
 * Parity pinning (what checks each part; tests/test_oracle_slatec.py,
 * tests/test_oracle_chem.py, tests/test_mechanism.py):
 *   or_dnrm2, or_dgefa, or_dgesl, or_ddcst, or_ddriv3
 *       bit-for-bit against SLATEC itself, compiled from the reference's own
 *       sources into oracle/_ref (oracle/Makefile): DNRM2 through all four
 *       of Lawson's phases, LU with exact zeros and singular matrices, the
 *       Gear tables, DDRIV3 on linear, Robertson (incl. MXSTEP exhaustion)
 *       and KEROMNES cells (identical y, NSTEP, NFE, NJE); plus closed forms
 *       (Euclidean norm, numpy solve, BDF l-vectors from prod(1 + x/i),
 *       expm of a stiff linear system, Robertson's known solution).
 *   or_h_molar / or_s_molar / or_cp_molar / or_e_molar
 *       NIST-JANAF values at 298.15, 1000 and 2000 K; cp = dh/dT,
 *       cp/T = ds/dT, e = h - RT; continuity at the 1000 K switch.
 *   or_mixture_concentration, or_rate_constants, or_rhs
 *       50-digit mpmath evaluation of the textbook forms (Arrhenius, Troe
 *       fall-off, K_c from NASA-7 Gibbs energies with dnu over species) for
 *       all 24 reactions at 700-3100 K; law of mass action; atom balance.
 *   or_chem_cell / or_chem_cells (composition of the above)
 *       scipy Radau at rtol 1e-12 on the same frozen-T system (1e-6 rel);
 *       gating (T < t_gate, bc != 0); atom and mass balance of the outputs.
 *   or_calculate_temperature, or_closure_cells, or_energy_from_temperature_cells
 *       bisection on e(T) (1e-6 K); round trip T -> E_f -> T; conservation of
 *       the mixture internal energy sum_s C_s e_s(T) over chemistry + closure
 *       (mass-consistent molar masses); p = rho R T / W.
 * Parity with the reference's Fortran chemistry module itself is unpinned by
 * execution: it does not compile with the available Fortran compiler without
 * source edits (namelist/derived-type name clash), see oracle/Makefile.
 */
#ifndef NRG_ORACLE_H
#define NRG_ORACLE_H

#include <stdint.h>
#include "../include/nrg_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NMAX 32

/* ---- LINPACK / BLAS-1 / SDRIVE pieces (0-based, column-major) ---------- */
double or_dnrm2(int n, const double *x);                          /* sl/dnrm2.f90:84-160  */
int    or_dgefa(double *a, int lda, int n, int *ipvt);            /* sl/dgefa.f90:70-116, returns INFO */
void   or_dgesl(const double *a, int lda, int n, const int *ipvt, double *b); /* sl/dgesl.f90:78-103, JOB=0 */
/* Gear tables EL(13,12), TQ(3,12) column-major, MINT = 2 (sl/ddcst.f90:39-86). */
void   or_ddcst(int maxord, double *el, double *tq);

typedef void (*or_rhs_fn)(void *ctx, int n, double t, const double *y, double *ydot);

typedef struct or_bdf_opts {
  double eps;     /* EPS (relative RMS tolerance)           */
  double ewt;     /* EWT(1), IERROR = 3 floor               */
  double hmax;    /* HMAX                                   */
  int mxstep;     /* MXSTEP                                 */
  int mxord;      /* MXORD (Gear: capped at 5)              */
} or_bdf_opts;

typedef struct or_bdf_out {
  int nstate;     /* 2 = reached TOUT, 3 = MXSTEP, 4 = EPS raised, 12 = fatal */
  int ierflg;     /* 0, 3, 4, 41, 42, 43                     */
  int nstep, nfe, nje, ncorr, natt, nqused;
  double hused, t;
} or_bdf_out;

/*
 * DDRIV3 with the chemistry solver's settings (NSTATE=1, NTASK=1, NROOT=0,
 * IERROR=3, MINT=2, MITER=2, IMPL=0): integrate y' = f(y) from 0 to tout in
 * place (chemical_kinetics_solver.f90:270-313; sl/ddriv3.f90:930-1449).
 */
void or_ddriv3(int n, double *y, double tout, const or_bdf_opts *o,
               or_rhs_fn f, void *ctx, or_bdf_out *out);

/* ---- thermochemistry (package_library/src/thermophysical_properties_class.f90) ---- */
double or_cp_molar(const nrg_chem_mech *m, double T, int s);      /* :225-242 */
double or_h_molar(const nrg_chem_mech *m, double T, int s);       /* :244-264 */
double or_s_molar(const nrg_chem_mech *m, double T, int s);       /* :266-286 */
double or_e_molar(const nrg_chem_mech *m, double T, int s);       /* :288-300 */
double or_mixture_cp_molar(const nrg_chem_mech *m, double T, const double *X);
double or_mixture_h_molar(const nrg_chem_mech *m, double T, const double *X);
double or_mixture_e_molar(const nrg_chem_mech *m, double T, const double *X);
double or_calculate_temperature(const nrg_chem_mech *m, double T, double e_molar, const double *X); /* :409-452 */

/* ---- kinetics (computing_module/src/current_build/chemical_kinetics_solver.f90) ---- */
void or_mixture_concentration(const nrg_chem_mech *m, const double *Y, double rho, double *Mconc); /* :235-248 */
void or_rate_constants(const nrg_chem_mech *m, double Tin, const double *Mconc,
                       double *kf, double *kr);                    /* :360-493 */
void or_rhs(const nrg_chem_mech *m, const double *kf, const double *kr, const double *Mconc,
            const double *Z, double *Zdot);                         /* :495-540 */

/* Per-cell chemistry right-hand side as a DDRIV3 callback (for driving the
 * reference's SLATEC build with the same f): ctx is or_chem_ctx_size() bytes. */
size_t or_chem_ctx_size(void);
void or_chem_ctx_init(void *ctx, const nrg_chem_mech *m, double T, double rho, const double *Y,
                      double *cin);
void or_chem_fun(void *ctx, int n, double t, const double *y, double *ydot);

typedef struct or_cell_out {
  int status;     /* NRG_CELL_* */
  or_bdf_out bdf;
} or_cell_out;

/* One cell of solve_chemical_kinetics (:228-346).  Returns the NRG_CELL_* status. */
int or_chem_cell(const nrg_chem_mech *m, const nrg_chem_opts *o, double dt,
                 double T, double rho, const double *Y, int8_t bc,
                 double *Y_prod, double *Y_prod2, double *E_f_prod, or_cell_out *out);

/* Whole grid, fields in NRG layout (see nrg_grid); nthreads > 1 uses OpenMP dynamic scheduling
 * like the reference loop (:216-222).  Returns the nrg_chem_step_dev status convention. */
int or_chem_step(const nrg_chem_mech *m, const nrg_chem_opts *o, const nrg_grid *g, double dt,
                 const double *T, const double *rho, const double *const *Y, const int8_t *bc,
                 double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                 int32_t *cell_status, nrg_chem_stats *stats, int nthreads);

/* Same step over a flat list of cells (SoA, stride 1): the workload of a sample. */
int or_chem_cells(const nrg_chem_mech *m, const nrg_chem_opts *o, int64_t ncells, double dt,
                  const double *T, const double *rho, const double *const *Y, const int8_t *bc,
                  double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                  int32_t *cell_status, int32_t *nstep, int32_t *nfe, int32_t *nje, int nthreads);

/* Legacy CABARET split closure + apply_state_equation for a flat list of cells
 * (cabaret_solver.f90:3316-3318,3386-3416; table_approximated_real_gas.f90:410-440). */
void or_closure_cells(const nrg_chem_mech *m, int64_t ncells, double dt,
                      const double *rho, const double *E_f_prod, const double *const *Y_prod,
                      const int8_t *bc, double *E_f, double *const *Y, double *T, double *p,
                      int nthreads);

/* E_f = (sum X_s (e_s(T) - h_s(T_ref)))/W, zero velocity (state setup). */
void or_energy_from_temperature_cells(const nrg_chem_mech *m, int64_t ncells, const double *T,
                                      const double *const *Y, double *E_f);

/* Table-approximated mode (cm/chemical_kinetics_solver.f90:250-346 with
 * ODE_solver = 'table_approximated'; interpolate_chemical_kinetics :650-685)
 * over a flat list of cells.  Table: T_t[rows], dC_t[s*rows + i] (Y_t(s+1, i+1)).
 * Gate as the SLATEC path (T >= t_gate and bc == 0); gated cells get
 * Y_prod = 0, E_f_prod = 0, Y_prod2 untouched. */
void or_table_cells(const nrg_chem_mech *m, int64_t ncells, double dt, double t_gate,
                    int32_t rows, const double *T_t, const double *dC_t,
                    const double *T, const double *const *Y, const int8_t *bc,
                    double *const *Y_prod, double *const *Y_prod2, double *E_f_prod,
                    int32_t *cell_status);

/* Low-Mach (FDS) coupling over a flat list of cells; cells with bc != 0 are
 * skipped (left untouched), as in the reference loops.
 *
 * or_low_mach_eos_cells: apply_state_equation_low_mach_fds
 *   (cm/table_approximated_real_gas.f90:570-677): W_mix, p_stat update
 *   (predictor: p_stat_old = p_stat, p_stat += dp_stat_dt dt; corrector:
 *   p_stat = (p_stat + p_stat_old + dp_stat_dt dt)/2), T = p_stat W_mix/(rho R),
 *   cp, gamma, sensible h_s, p, sound speed, e_i and E_f = e_i + v_d^2/2 of
 *   the LAST dimension d (the reference's loop overwrites E_f per dimension).
 *   Optional outputs (NULL: not written): p, mixture_cp, gamma, h_s, v_s, e_i, E_f.
 * or_low_mach_species_cells: the species update of calculate_interm_Y_predictor
 *   (cm/fds_low_mach_solver.f90:906-925) / _corrector (:3103-3125) with the
 *   flux divergence rhs supplied by the caller (NULL: zero) and the
 *   diffusion and chemistry sources (NULL: flag off).
 * or_low_mach_chem_divergence_cells: div_v of calculate_divergence_v
 *   (cm/fds_low_mach_solver.f90:1057-1162) with the reactive terms only (heat
 *   transfer, diffusion, radiation, particles off; zero face velocities). */
void or_low_mach_eos_cells(const nrg_chem_mech *m, int64_t ncells, double dt, int predictor,
                           const double *rho, const double *const *Y, const double *const *v,
                           int ndim, const double *dp_stat_dt, const int8_t *bc, double *p_stat,
                           double *p_stat_old, double *T, double *p, double *mix_mol_mass,
                           double *mixture_cp, double *gamma, double *h_s, double *v_s,
                           double *e_i, double *E_f);
void or_low_mach_species_cells(const nrg_chem_mech *m, int64_t ncells, double dt, int corrector,
                               const double *rho_a, const double *const *Y_a,
                               const double *rho_b, const double *const *Y_b,
                               const double *const *rhs, const double *const *Y_prod_diff,
                               const double *const *Y_prod_chem, const int8_t *bc,
                               double *rho_out, double *const *Y_out);
void or_low_mach_chem_divergence_cells(const nrg_chem_mech *m, int64_t ncells, const double *rho,
                                       const double *T, const double *const *Y,
                                       const double *mix_mol_mass, const double *E_f_prod_chem,
                                       const double *const *Y_prod_chem, const int8_t *bc,
                                       double *div_v);

void or_default_opts(nrg_chem_opts *o);

#ifdef __cplusplus
}
#endif
#endif
