// topo_api.h -- the C-ABI entry points that work on a handle, as one list.
//
// The kernels are compiled once per reaction topology (nrg_chem.hip built with
// a different topology_<name>.h each time).  Every such build exposes the
// entry points below through a table of function pointers (nrg_topo_ops); the
// C-ABI in nrg_api.cpp holds a handle's table and forwards each call to it.
// NRG_TOPO_ENTRY_POINTS(X) expands X(name, parameters, arguments) for every
// entry point that returns a status and takes a handle first; H(h) names the
// handle argument as the including file needs it.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/nrg_chem.h"
#include "../../include/nrg_low_mach.h"
#include "../../include/nrg_surrogate.h"

#define NRG_STEP_PARAMS                                                                       \
  ::nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T, const double *p,        \
      const double *rho, const double *const *Y, const int8_t *bc, double *const *Y_prod,   \
      double *const *Y_prod2, double *E_f_prod, int32_t *cell_status, nrg_chem_stats *stats
#define NRG_STEP_ARGS H(h), grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status, stats

#define NRG_TOPO_ENTRY_POINTS(X)                                                                  \
  X(nrg_chem_step_dev, (NRG_STEP_PARAMS), (NRG_STEP_ARGS))                                        \
  X(nrg_chem_step_dev_noalloc, (NRG_STEP_PARAMS), (NRG_STEP_ARGS))                                \
  X(nrg_chem_step, (NRG_STEP_PARAMS), (NRG_STEP_ARGS))                                            \
  X(nrg_chem_table_step, (NRG_STEP_PARAMS), (NRG_STEP_ARGS))                                      \
  X(nrg_chem_table_step_dev, (NRG_STEP_PARAMS), (NRG_STEP_ARGS))                                  \
  X(nrg_chem_step_dev_diag,                                                                       \
    (NRG_STEP_PARAMS, int32_t *nstep, int32_t *nfe, int32_t *nje), (NRG_STEP_ARGS, nstep, nfe, nje)) \
  X(nrg_chem_closure_dev,                                                                         \
    (::nrg_chem_t *h, const nrg_grid *grid, double dt, const double *rho, const double *const *v, \
     int32_t ndim, const double *E_f_prod, const double *const *Y_prod, const int8_t *bc,         \
     double *E_f, double *const *Y, double *T, double *p),                                        \
    (H(h), grid, dt, rho, v, ndim, E_f_prod, Y_prod, bc, E_f, Y, T, p))                           \
  X(nrg_state_equation_dev,                                                                       \
    (::nrg_chem_t *h, const nrg_grid *grid, const double *rho, const double *const *v,            \
     int32_t ndim, const double *E_f, const double *const *Y, const int8_t *bc, double *T,        \
     double *p),                                                                                  \
    (H(h), grid, rho, v, ndim, E_f, Y, bc, T, p))                                                 \
  X(nrg_energy_from_temperature_dev,                                                              \
    (::nrg_chem_t *h, const nrg_grid *grid, const double *T, const double *const *Y,              \
     const double *const *v, int32_t ndim, double *E_f),                                          \
    (H(h), grid, T, Y, v, ndim, E_f))                                                             \
  X(nrg_chem_autotune_dev,                                                                        \
    (::nrg_chem_t *h, const nrg_grid *grid, double dt, const double *T, const double *p,          \
     const double *rho, const double *const *Y, const int8_t *bc, double *const *Y_prod,          \
     double *const *Y_prod2, double *E_f_prod, nrg_chem_tuning *out),                             \
    (H(h), grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, out))                           \
  X(nrg_chem_set_table,                                                                           \
    (::nrg_chem_t *h, int32_t rows, const double *T_t, const double *dC_t),                       \
    (H(h), rows, T_t, dC_t))                                                                      \
  X(nrg_low_mach_state_equation_dev,                                                              \
    (::nrg_chem_t *h, const nrg_grid *grid, double time_step, int32_t predictor,                  \
     const double *rho, const double *const *Y, const double *const *v, int32_t ndim,             \
     const double *dp_stat_dt, const int8_t *bc, double *p_stat, double *p_stat_old, double *T,   \
     double *mix_mol_mass, double *p, double *mixture_cp, double *gamma, double *h_s,             \
     double *v_s, double *e_i, double *E_f),                                                      \
    (H(h), grid, time_step, predictor, rho, Y, v, ndim, dp_stat_dt, bc, p_stat, p_stat_old, T,    \
     mix_mol_mass, p, mixture_cp, gamma, h_s, v_s, e_i, E_f))                                     \
  X(nrg_low_mach_species_dev,                                                                     \
    (::nrg_chem_t *h, const nrg_grid *grid, double time_step, int32_t corrector,                  \
     const double *rho_a, const double *const *Y_a, const double *rho_b,                          \
     const double *const *Y_b, const double *const *rhs, const double *const *Y_prod_diff,        \
     const double *const *Y_prod_chem, const int8_t *bc, double *rho_out, double *const *Y_out),  \
    (H(h), grid, time_step, corrector, rho_a, Y_a, rho_b, Y_b, rhs, Y_prod_diff, Y_prod_chem, bc, \
     rho_out, Y_out))                                                                             \
  X(nrg_low_mach_chem_divergence_dev,                                                             \
    (::nrg_chem_t *h, const nrg_grid *grid, const double *rho, const double *T,                   \
     const double *const *Y, const double *mix_mol_mass, const double *E_f_prod_chem,             \
     const double *const *Y_prod_chem, const int8_t *bc, double *div_v),                          \
    (H(h), grid, rho, T, Y, mix_mol_mass, E_f_prod_chem, Y_prod_chem, bc, div_v))                 \
  X(nrg_chem_surrogate_step_dev,                                                                  \
    (::nrg_chem_t *h, nrg_mlp_t *mlp, const nrg_surrogate_norm *norm,                             \
     const nrg_surrogate_gate *gate, const nrg_grid *grid, double dt, const double *T,            \
     const double *p, const double *rho, const double *const *Y, const int8_t *bc,                \
     double *const *Y_prod, double *const *Y_prod2, double *E_f_prod, int32_t *cell_status,       \
     nrg_surrogate_stats *stats),                                                                 \
    (H(h), mlp, norm, gate, grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status,   \
     stats))

// One compiled topology.  `check` returns NRG_CHEM_OK when the mechanism's
// reaction list is the one its kernels were built for, NRG_CHEM_E_MECH when
// it is another, or an argument error; `create` builds a handle (returned as
// the opaque type, to be passed back to the same table only).
struct nrg_topo_ops {
  const char *name;
  int (*check)(const nrg_chem_mech *mech);
  int (*create)(const nrg_chem_mech *mech, const nrg_chem_opts *opts, ::nrg_chem_t **out);
  void (*destroy)(::nrg_chem_t *h);
  void *(*stream)(::nrg_chem_t *h);
#define NRG_TOPO_OP_FIELD(name, params, args) int(*name) params;
  NRG_TOPO_ENTRY_POINTS(NRG_TOPO_OP_FIELD)
#undef NRG_TOPO_OP_FIELD
};

// The tables of the topologies this library is built with (one per
// nrg_chem.hip build, see the Makefile).
const nrg_topo_ops *nrg_topo_ops_keromnes();
const nrg_topo_ops *nrg_topo_ops_zhang();
const nrg_topo_ops *nrg_topo_ops_tereza();
const nrg_topo_ops *nrg_topo_ops_konnov();
const nrg_topo_ops *nrg_topo_ops_acetylene();
