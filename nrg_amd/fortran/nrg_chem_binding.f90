! nrg_chem_binding -- ISO_C_BINDING interface to libnrg_chem (include/nrg_chem.h).
!
! This is the Fortran side of the boundary: the bind(C) types mirror the C
! structs field for field, and the interfaces bind every entry point by its C
! name.  A replacement of NRG's chemical_kinetics_solver module
! (computing_module/src/current_build/chemical_kinetics_solver.f90) uses it
! to hand its fields to the MI355X step by c_loc, without copies:
!   - chemical_kinetics_solver_c(manager)      -> nrg_chem_create   (:111-173)
!   - this%solve_chemical_kinetics(time_step)  -> nrg_chem_step / nrg_chem_step_dev (:175-358)
! Arrays keep the reference's layout: column-major, ghosted bounds, one
! allocation per species (Y%pr(s)%cells), 1-based index tables.
module nrg_chem_binding
  use iso_c_binding
  implicit none
  private

  ! NRG_CHEM_ABI_VERSION of the header (Fortran names are case-blind, so the
  ! constant cannot share the name of the function nrg_chem_abi_version)
  integer(c_int), parameter, public :: NRG_CHEM_ABI = 1
  integer(c_int), parameter, public :: NRG_CHEM_OK = 0, NRG_CHEM_E_ARG = 1, NRG_CHEM_E_MECH = 2, &
                                       NRG_CHEM_E_HIP = 3, NRG_CHEM_E_NO_DEVICE = 4, &
                                       NRG_CHEM_E_STEP_ZERO = 41, NRG_CHEM_E_REPEATED_FAIL = 42
  integer(c_int), parameter, public :: NRG_CELL_OK = 0, NRG_CELL_GATED = 1, NRG_CELL_MXSTEP = 2, &
                                       NRG_CELL_STEP_ZERO = 3, NRG_CELL_REPEATED_FAIL = 4, &
                                       NRG_CELL_SINGULAR = 5
  integer(c_int), parameter, public :: NRG_JAC_FD_SLATEC = 0, NRG_JAC_ANALYTIC = 1
  integer(c_int), parameter, public :: NRG_FORM_PLAIN = 0, NRG_FORM_PERSISTENT = 1
  integer(c_int), parameter, public :: NRG_ORDER_NATURAL = 0, NRG_ORDER_STIFFNESS = 1
  integer(c_int), parameter, public :: NRG_NOALLOC_MAX_CELLS = 48 * 131072

  ! struct nrg_chem_mech: the arrays of chemical_properties / thermophysical_properties
  type, bind(C), public :: nrg_chem_mech
    integer(c_int32_t) :: species_number, reactions_number
    type(c_ptr) :: reactions_type, chemical_coeffs
    type(c_ptr) :: A, beta, E_act, A_low, beta_low, E_act_low
    type(c_ptr) :: Troe_coeffs, enhanced_efficiencies, a_coeffs, molar_masses
  end type

  type, bind(C), public :: nrg_chem_opts
    real(c_double) :: eps, ewt, hmax, t_gate
    integer(c_int32_t) :: mxstep, mxord, jacobian, device, form, jac_batch_pct, waves_per_cu
    integer(c_int32_t) :: cell_order, use_graph, chunk_log2
    integer(c_int32_t) :: errtest_gather_pct, refill_min_idle
  end type

  type, bind(C), public :: nrg_grid
    integer(c_int32_t) :: lo(3), hi(3), inner_lo(3), inner_hi(3)
  end type

  type, bind(C), public :: nrg_chem_stats
    integer(c_int64_t) :: cells_inner, cells_integrated
    integer(c_int64_t) :: nstep_sum, nfe_sum, nje_sum, ncorr_sum, natt_sum
    integer(c_int32_t) :: nstep_max, nfe_max
    integer(c_int64_t) :: status_count(8)
    real(c_double) :: kernel_ms
  end type

  ! nrg_chem_autotune_dev: one trial and the table of trials
  integer(c_int), parameter, public :: NRG_TUNE_MAX_TRIALS = 24
  type, bind(C), public :: nrg_chem_tune_trial
    integer(c_int32_t) :: form, cell_order, jac_batch_pct, waves_per_cu
    integer(c_int64_t) :: chunk_cells
    real(c_double) :: kernel_ms
    integer(c_int32_t) :: errtest_gather_pct, refill_min_idle
  end type

  type, bind(C), public :: nrg_chem_tuning
    type(nrg_chem_tune_trial) :: best
    integer(c_int32_t) :: n_trials, reserved
    type(nrg_chem_tune_trial) :: trials(NRG_TUNE_MAX_TRIALS)
  end type

  public :: nrg_chem_default_opts, nrg_chem_strerror, nrg_chem_last_error, nrg_chem_abi_version
  public :: nrg_chem_autotune_dev
  public :: nrg_chem_create, nrg_chem_destroy, nrg_chem_step, nrg_chem_step_dev
  public :: nrg_chem_closure_dev, nrg_state_equation_dev, nrg_energy_from_temperature_dev
  public :: nrg_chem_stream, nrg_chem_message
  public :: nrg_chem_set_table, nrg_chem_table_step, nrg_chem_table_step_dev
  public :: nrg_low_mach_state_equation_dev, nrg_low_mach_species_dev
  public :: nrg_low_mach_chem_divergence_dev
  public :: nrg_chem_step_dev_noalloc, nrg_chem_topology, nrg_chem_topology_name

  interface
    subroutine nrg_chem_default_opts(opts) bind(C, name="nrg_chem_default_opts")
      import :: nrg_chem_opts
      type(nrg_chem_opts), intent(out) :: opts
    end subroutine

    function nrg_chem_strerror(status) bind(C, name="nrg_chem_strerror") result(msg)
      import :: c_int, c_ptr
      integer(c_int), value :: status
      type(c_ptr) :: msg
    end function

    function nrg_chem_last_error() bind(C, name="nrg_chem_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function

    function nrg_chem_abi_version() bind(C, name="nrg_chem_abi_version") result(v)
      import :: c_int
      integer(c_int) :: v
    end function

    function nrg_chem_create(mech, opts, handle) bind(C, name="nrg_chem_create") result(status)
      import :: c_int, c_ptr, nrg_chem_mech, nrg_chem_opts
      type(nrg_chem_mech), intent(in) :: mech
      type(nrg_chem_opts), intent(in) :: opts
      type(c_ptr), intent(out) :: handle
      integer(c_int) :: status
    end function

    subroutine nrg_chem_destroy(handle) bind(C, name="nrg_chem_destroy")
      import :: c_ptr
      type(c_ptr), value :: handle
    end subroutine

    ! host fields (staged through device buffers by the library)
    function nrg_chem_step(handle, grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, &
                           cell_status, stats) bind(C, name="nrg_chem_step") result(status)
      import :: c_int, c_ptr, c_double, nrg_grid, nrg_chem_stats
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      type(c_ptr), value :: T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status
      type(nrg_chem_stats), intent(out) :: stats
      integer(c_int) :: status
    end function

    ! device-resident fields
    function nrg_chem_step_dev(handle, grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, &
                               cell_status, stats) bind(C, name="nrg_chem_step_dev") result(status)
      import :: c_int, c_ptr, c_double, nrg_grid, nrg_chem_stats
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      type(c_ptr), value :: T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status
      type(nrg_chem_stats), intent(out) :: stats
      integer(c_int) :: status
    end function

    ! nrg_chem_step_dev without any device allocation (no reference counterpart:
    ! for callers that cannot allocate device memory; at most
    ! NRG_NOALLOC_MAX_CELLS inner cells, one call at a time per device)
    function nrg_chem_step_dev_noalloc(handle, grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, &
                                       E_f_prod, cell_status, stats) &
        bind(C, name="nrg_chem_step_dev_noalloc") result(status)
      import :: c_int, c_ptr, c_double, nrg_grid, nrg_chem_stats
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      type(c_ptr), value :: T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status
      type(nrg_chem_stats), intent(out) :: stats
      integer(c_int) :: status
    end function

    ! name of the reaction-topology build serving the handle (C string)
    function nrg_chem_topology(handle) bind(C, name="nrg_chem_topology") result(name)
      import :: c_ptr
      type(c_ptr), value :: handle
      type(c_ptr) :: name
    end function

    ! read_chemical_kinetics_table (chemical_kinetics_solver.f90:608-648):
    ! T_t(rows) and dC_t(rows, S) in Fortran order (species-major in C).
    function nrg_chem_set_table(handle, rows, T_t, dC_t) bind(C, name="nrg_chem_set_table") &
        result(status)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: rows
      real(c_double), intent(in) :: T_t(*), dC_t(*)
      integer(c_int) :: status
    end function

    ! ODE_solver = 'table_approximated' (chemical_kinetics_solver.f90:250-346), host fields
    function nrg_chem_table_step(handle, grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, &
                                 E_f_prod, cell_status, stats) &
        bind(C, name="nrg_chem_table_step") result(status)
      import :: c_int, c_ptr, c_double, nrg_grid, nrg_chem_stats
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      type(c_ptr), value :: T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status
      type(nrg_chem_stats), intent(out) :: stats
      integer(c_int) :: status
    end function

    ! include/nrg_low_mach.h: apply_state_equation_low_mach_fds
    ! (table_approximated_real_gas.f90:570-677) on device fields
    function nrg_low_mach_state_equation_dev(handle, grid, dt, predictor, rho, Y, v, ndim, &
                                             dp_stat_dt, bc, p_stat, p_stat_old, T, &
                                             mix_mol_mass, p, mixture_cp, gamma, h_s, v_s, e_i, &
                                             E_f) bind(C, name="nrg_low_mach_state_equation_dev") &
        result(status)
      import :: c_int, c_int32_t, c_ptr, c_double, nrg_grid
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      integer(c_int32_t), value :: predictor, ndim
      type(c_ptr), value :: rho, Y, v, dp_stat_dt, bc, p_stat, p_stat_old, T, mix_mol_mass
      type(c_ptr), value :: p, mixture_cp, gamma, h_s, v_s, e_i, E_f
      integer(c_int) :: status
    end function

    ! species predictor / corrector (fds_low_mach_solver.f90:906-925, 3103-3125)
    function nrg_low_mach_species_dev(handle, grid, dt, corrector, rho_a, Y_a, rho_b, Y_b, rhs, &
                                      Y_prod_diff, Y_prod_chem, bc, rho_out, Y_out) &
        bind(C, name="nrg_low_mach_species_dev") result(status)
      import :: c_int, c_int32_t, c_ptr, c_double, nrg_grid
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      integer(c_int32_t), value :: corrector
      type(c_ptr), value :: rho_a, Y_a, rho_b, Y_b, rhs, Y_prod_diff, Y_prod_chem, bc
      type(c_ptr), value :: rho_out, Y_out
      integer(c_int) :: status
    end function

    ! reactive part of div_v (fds_low_mach_solver.f90:1057-1162)
    function nrg_low_mach_chem_divergence_dev(handle, grid, rho, T, Y, mix_mol_mass, &
                                              E_f_prod_chem, Y_prod_chem, bc, div_v) &
        bind(C, name="nrg_low_mach_chem_divergence_dev") result(status)
      import :: c_int, c_ptr, nrg_grid
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      type(c_ptr), value :: rho, T, Y, mix_mol_mass, E_f_prod_chem, Y_prod_chem, bc, div_v
      integer(c_int) :: status
    end function

    ! the same on device-resident fields
    function nrg_chem_table_step_dev(handle, grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, &
                                     E_f_prod, cell_status, stats) &
        bind(C, name="nrg_chem_table_step_dev") result(status)
      import :: c_int, c_ptr, c_double, nrg_grid, nrg_chem_stats
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      type(c_ptr), value :: T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, cell_status
      type(nrg_chem_stats), intent(out) :: stats
      integer(c_int) :: status
    end function

    function nrg_chem_closure_dev(handle, grid, dt, rho, v, ndim, E_f_prod, Y_prod, bc, E_f, Y, &
                                  T, p) bind(C, name="nrg_chem_closure_dev") result(status)
      import :: c_int, c_int32_t, c_ptr, c_double, nrg_grid
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      type(c_ptr), value :: rho, v
      integer(c_int32_t), value :: ndim
      type(c_ptr), value :: E_f_prod, Y_prod, bc, E_f, Y, T, p
      integer(c_int) :: status
    end function

    function nrg_state_equation_dev(handle, grid, rho, v, ndim, E_f, Y, bc, T, p) &
        bind(C, name="nrg_state_equation_dev") result(status)
      import :: c_int, c_int32_t, c_ptr, nrg_grid
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      type(c_ptr), value :: rho, v
      integer(c_int32_t), value :: ndim
      type(c_ptr), value :: E_f, Y, bc, T, p
      integer(c_int) :: status
    end function

    function nrg_energy_from_temperature_dev(handle, grid, T, Y, v, ndim, E_f) &
        bind(C, name="nrg_energy_from_temperature_dev") result(status)
      import :: c_int, c_int32_t, c_ptr, nrg_grid
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      type(c_ptr), value :: T, Y, v
      integer(c_int32_t), value :: ndim
      type(c_ptr), value :: E_f
      integer(c_int) :: status
    end function

    function nrg_chem_autotune_dev(handle, grid, dt, T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod, &
                                   tuning) bind(C, name="nrg_chem_autotune_dev") result(status)
      import :: c_int, c_ptr, c_double, nrg_grid, nrg_chem_tuning
      type(c_ptr), value :: handle
      type(nrg_grid), intent(in) :: grid
      real(c_double), value :: dt
      type(c_ptr), value :: T, p, rho, Y, bc, Y_prod, Y_prod2, E_f_prod
      type(nrg_chem_tuning), intent(out) :: tuning
      integer(c_int) :: status
    end function

    function nrg_chem_stream(handle) bind(C, name="nrg_chem_stream") result(stream)
      import :: c_ptr
      type(c_ptr), value :: handle
      type(c_ptr) :: stream
    end function

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! Name of the kernel build that serves a handle ("KEROMNES", "ZHANG", ...).
  function nrg_chem_topology_name(handle) result(name)
    type(c_ptr), intent(in) :: handle
    character(len=:), allocatable :: name
    name = c_string(nrg_chem_topology(handle))
  end function

  ! Text of a status code followed by the library's last detailed message.
  function nrg_chem_message(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=:), allocatable :: msg
    msg = c_string(nrg_chem_strerror(status))
    if (status /= NRG_CHEM_OK) msg = msg // ': ' // c_string(nrg_chem_last_error())
  end function

  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    do i = 1, n
      s(i:i) = chars(i)
    end do
  end function

end module nrg_chem_binding
