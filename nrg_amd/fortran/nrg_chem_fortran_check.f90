! Fortran caller of libnrg_chem through nrg_chem_binding, holding its fields
! the way NRG's data manager does: one ghosted, column-major allocation per
! scalar field and per species (Y%pr(s)%cells, package_library/src/
! field_vector_class.f90), int8 boundary markers with the same bounds
! (package_library/src/boundary_conditions_class.f90:19), and the mechanism
! arrays of chemical_properties / thermophysical_properties in Fortran order.
!
! usage: nrg_chem_fortran_check <input.bin> <output.bin>
!   input  (stream, native endian): int32 S, R, nx, ny, nz; real64 dt;
!          reactions_type(R), chemical_coeffs(4,R,2) (int32);
!          A, beta, E_act, A_low, beta_low, E_act_low (R); Troe_coeffs(R,4);
!          enhanced_efficiencies(R,S); a_coeffs(7,2,S); molar_masses(S);
!          then on the ghosted grid: T, rho, Y(1..S) (real64), bc (int8)
!   output: int32 status; int64 cells_integrated, nstep_sum, nfe_sum, nje_sum;
!           E_f_prod, Y_prod(1..S), Y_prod2(1..S) on the ghosted grid
! Exit status: 0 step done, 3 no gfx950 device (after the host-side checks),
! 1 any other failure.
program nrg_chem_fortran_check
  use iso_c_binding
  use nrg_chem_binding
  implicit none

  type :: field
    real(c_double), allocatable :: cells(:,:,:)
  end type

  integer(c_int32_t) :: S, R, nx, ny, nz
  real(c_double) :: dt
  integer(c_int32_t), allocatable, target :: reactions_type(:), chemical_coeffs(:,:,:)
  real(c_double), allocatable, target :: A(:), beta(:), E_act(:), A_low(:), beta_low(:), E_act_low(:)
  real(c_double), allocatable, target :: Troe(:,:), eff(:,:), a_coeffs(:,:,:), W(:)
  type(field), allocatable, target :: Y(:), Y_prod(:), Y_prod2(:)
  real(c_double), allocatable, target :: T(:,:,:), rho(:,:,:), E_f_prod(:,:,:)
  integer(c_int8_t), allocatable, target :: bc(:,:,:)
  type(c_ptr), allocatable, target :: pY(:), pYp(:), pYp2(:)
  type(nrg_chem_mech) :: mech, bad
  type(nrg_chem_opts) :: opts
  type(nrg_grid) :: grid
  type(nrg_chem_stats) :: stats
  type(c_ptr) :: handle
  integer(c_int) :: status
  integer :: u, sp, klo, khi
  character(len=4096) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  ! ---- host-side contract checks (no device needed) ----------------------
  if (nrg_chem_abi_version() /= NRG_CHEM_ABI) then
    print '(a)', 'abi version mismatch'
    stop 1
  end if
  call nrg_chem_default_opts(opts)
  print '(a,es24.16,1x,es24.16,1x,es24.16,1x,i0,1x,i0)', 'defaults ', opts%eps, opts%ewt, &
        opts%hmax, opts%mxstep, opts%mxord
  if (opts%hmax /= real(1.0e-7, c_double)) then       ! HMAX = 1.0E-7, a default-real literal
    print '(a)', 'HMAX default differs from the reference literal'
    stop 1
  end if
  bad%species_number = 0
  bad%reactions_number = 0
  status = nrg_chem_create(bad, opts, handle)
  print '(a,i0,a,a)', 'empty-mechanism status ', status, ' ', nrg_chem_message(status)
  if (status == NRG_CHEM_OK) stop 1

  ! ---- read the case -----------------------------------------------------
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) S, R, nx, ny, nz, dt
  allocate(reactions_type(R), chemical_coeffs(4, R, 2), A(R), beta(R), E_act(R), A_low(R), &
           beta_low(R), E_act_low(R), Troe(R, 4), eff(R, S), a_coeffs(7, 2, S), W(S))
  read(u) reactions_type, chemical_coeffs, A, beta, E_act, A_low, beta_low, E_act_low, Troe, &
          eff, a_coeffs, W
  ! computational_domain: ghost layer only in active dimensions
  klo = 1
  khi = 1
  if (nz > 1) then
    klo = 0
    khi = nz + 1
  end if
  allocate(T(0:nx+1, 0:ny+1, klo:khi), rho(0:nx+1, 0:ny+1, klo:khi), &
           E_f_prod(0:nx+1, 0:ny+1, klo:khi), bc(0:nx+1, 0:ny+1, klo:khi))
  allocate(Y(S), Y_prod(S), Y_prod2(S), pY(S), pYp(S), pYp2(S))
  read(u) T, rho
  do sp = 1, S
    allocate(Y(sp)%cells(0:nx+1, 0:ny+1, klo:khi))
    allocate(Y_prod(sp)%cells(0:nx+1, 0:ny+1, klo:khi))
    allocate(Y_prod2(sp)%cells(0:nx+1, 0:ny+1, klo:khi))
    read(u) Y(sp)%cells
    Y_prod(sp)%cells = 0.0_c_double
    Y_prod2(sp)%cells = 7.0_c_double        ! gated cells must keep this
    pY(sp) = c_loc(Y(sp)%cells)
    pYp(sp) = c_loc(Y_prod(sp)%cells)
    pYp2(sp) = c_loc(Y_prod2(sp)%cells)
  end do
  read(u) bc
  close(u)
  E_f_prod = 0.0_c_double

  mech%species_number = S
  mech%reactions_number = R
  mech%reactions_type = c_loc(reactions_type)
  mech%chemical_coeffs = c_loc(chemical_coeffs)
  mech%A = c_loc(A)
  mech%beta = c_loc(beta)
  mech%E_act = c_loc(E_act)
  mech%A_low = c_loc(A_low)
  mech%beta_low = c_loc(beta_low)
  mech%E_act_low = c_loc(E_act_low)
  mech%Troe_coeffs = c_loc(Troe)
  mech%enhanced_efficiencies = c_loc(eff)
  mech%a_coeffs = c_loc(a_coeffs)
  mech%molar_masses = c_loc(W)

  grid%lo = [0, 0, klo]
  grid%hi = [nx + 1, ny + 1, khi]
  grid%inner_lo = [1, 1, 1]
  grid%inner_hi = [nx, ny, nz]

  ! ---- constructor and one step (chemical_kinetics_solver_c + solve_chemical_kinetics)
  status = nrg_chem_create(mech, opts, handle)
  print '(a,i0,a,a)', 'create status ', status, ' ', nrg_chem_message(status)
  if (status == NRG_CHEM_E_NO_DEVICE) stop 3
  if (status /= NRG_CHEM_OK) stop 1
  print '(a,a)', 'topology ', nrg_chem_topology_name(handle)
  status = nrg_chem_step(handle, grid, dt, c_loc(T), c_null_ptr, c_loc(rho), c_loc(pY), &
                         c_loc(bc), c_loc(pYp), c_loc(pYp2), c_loc(E_f_prod), c_null_ptr, stats)
  print '(a,i0,a,i0,a,i0)', 'step status ', status, ' integrated ', stats%cells_integrated, &
        ' nfe ', stats%nfe_sum
  call nrg_chem_destroy(handle)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) status, stats%cells_integrated, stats%nstep_sum, stats%nfe_sum, stats%nje_sum
  write(u) E_f_prod
  do sp = 1, S
    write(u) Y_prod(sp)%cells
  end do
  do sp = 1, S
    write(u) Y_prod2(sp)%cells
  end do
  close(u)
  if (status /= NRG_CHEM_OK) stop 1
end program nrg_chem_fortran_check
