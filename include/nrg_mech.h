/* Mechanism and thermochemistry files read the way NRG reads them, for hosts
 * that are not NRG itself (NRG hands its own arrays to nrg_chem_create).
 *
 * Replaces chemical_properties%set_properties
 * (package_library/src/chemical_properties_class.f90:117-181, reaction parser
 * :289-451 and :453-635, species :685-768) and the NASA-7 / molar-mass part of
 * thermophysical_properties (package_library/src/thermophysical_properties_class.f90:
 * 137-179, get_a_coeffs :1226-1311, molar masses :1313-1341), including the
 * reference's reading rules:
 *   - a reaction with a species not in the SPECIES list is dropped;
 *   - list-directed reads: an unquoted '/' ends a record, so unquoted
 *     efficiency lines ("H2/2.5/ ...") are not read as efficiencies;
 *   - the reads after a reaction line use iostat: a TROE line with three
 *     values takes the next record's first item as the fourth value, which
 *     fails to convert, leaves Troe(4) = 0 and skips that reaction's
 *     efficiency record;
 *   - the enhanced-efficiency flag, once set by a "+M" or "(+M)" reaction,
 *     stays set for the rest of the file;
 *   - A is scaled by 1e-6 / 1e-12 for two / three left-hand terms, A_low by
 *     the count plus one, E_act and E_act_low by the energy unit
 *     ("J.mol", "kJ.mol", "cal.mol", "kcal.mol");
 *   - NASA-7 coefficients are multiplied by R = 8.31446261815324 J/(mol K).
 * The arrays have the reference's Fortran layout (see nrg_chem_mech).
 */
#ifndef NRG_MECH_H
#define NRG_MECH_H

#include "nrg_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nrg_mech_file nrg_mech_file;   /* owns the loaded arrays */

/* Loads the three files.  E_act_units NULL means "cal.mol" (the reference
 * tests' setting); default_enhanced_efficiencies is the reference's
 * chemical_data namelist default (1.0 in its tests).  Returns
 * NRG_CHEM_E_FORMAT (with nrg_chem_last_error) for a file that cannot be
 * read, a missing SPECIES/END or REACTIONS section, an unreadable NASA-7
 * block or a species without a molar mass. */
int nrg_mech_load(const char *mechanism_path, const char *thermo_path,
                  const char *molar_masses_path, double default_enhanced_efficiencies,
                  const char *E_act_units, nrg_mech_file **out);

/* Loads a task set-up folder the way NRG's main program does
 * (computing_module/src/current_build/main.f90:170-171): the namelist
 *   <dir>/chemical_data.inf         &chemical_properties chemical_mechanism_file_name,
 *                                   default_enhanced_efficiencies, E_act_units /
 * (package_library/src/chemical_properties_class.f90:77-97), the namelist
 *   <dir>/thermophysical_setup.inf  &thermo_properties thermo_data_file_name,
 *                                   transport_data_file_name, molar_masses_data_file_name /
 * (package_library/src/thermophysical_properties_class.f90:97-134), then the
 * named files <dir>/chemical_mechanisms/<mechanism> and
 * <dir>/thermophysical_data/<thermo>, <dir>/thermophysical_data/<molar masses>
 * through nrg_mech_load.  The transport file is not read (the chemistry step
 * does not use it).  Names and groups are case-insensitive; quoted values lose
 * their trailing blanks.  A missing file, group or variable is NRG_CHEM_E_FORMAT. */
int nrg_mech_load_task_setup(const char *task_setup_dir, nrg_mech_file **out);

/* Points *view at the loaded arrays (valid until nrg_mech_free), ready for
 * nrg_chem_create. */
int nrg_mech_view(const nrg_mech_file *m, nrg_chem_mech *view);

/* Name of species s (0-based), or NULL when s is out of range. */
const char *nrg_mech_species_name(const nrg_mech_file *m, int32_t s);

/* Reaction line r (0-based) as read from the file, or NULL. */
const char *nrg_mech_reaction_string(const nrg_mech_file *m, int32_t r);

void nrg_mech_free(nrg_mech_file *m);

#ifdef __cplusplus
}
#endif
#endif /* NRG_MECH_H */
