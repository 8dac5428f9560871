! NetCDF-free OpenMP driver around the reference Fortran transport -- TEST INFRASTRUCTURE ONLY.
!
! Project-written. It reads a transport.bin v1 export (kernels/cpu/worker.f90:1102-1120
! layout: header, geometry, zunit, views, surface, albedo, elevation, BETA, species
! fractions, gas albedo, then per species LUT size, albedo, cosines, bound CDF, F11, F22,
! F33, F12, and with nmat 4 also F34, F44) and runs statistical rounds exactly as the reference worker does
! (kernels/cpu/worker.f90:1126-1204): per-thread seeds seed+thread+index
! (worker.f90:298-321), then one OpenMP REDUCTION(+) region around the unchanged
! monte_carlo subroutine (mc.f90:20) per round. Tallies of each round are appended to the
! output file as float64 in Fortran order (Stokes, then paths and weights); transport
! seconds per round go to stdout. With save_scats = 1 (optional sixth argument) the rounds
! run with the single/multiple-scattering decomposition (calc_radiances_and_scats, one
! thread) and each round also writes Stokes_sing, Stokes_mult and the three 3-D contribution
! arrays (nmat, nx, ny, 0:nz, nangles).
!
! Usage: monki_ref <transport.bin> <photons_per_column> <rounds> <seed> <output.bin> [save_scats]
program monki_ref
  use iso_fortran_env, only: real64, int64, output_unit
  use omp_lib
  use new_data_types
  implicit none

  character(len=4096) :: arg, table_path, output_path
  integer :: ppc, rounds, seed, nt, seed_size, myid, idx, iround, unit_in, unit_out, stat
  integer :: nx, ny, nz, nmat, n_angles, n_types, k_source, k_detector, k_register, is, n
  logical :: save_path, backward, save_scats
  real(real64) :: header(12), geometry(6), xunit, yunit, depol, vartheta0, phi0, phi
  real(real64), allocatable :: zunit(:), vartheta(:), surface_r(:,:), albedo(:,:), elev_r(:,:)
  real(real64), allocatable :: beta(:,:,:), ext(:,:,:,:), ssalb_gas(:), tau_gas(:,:)
  integer, allocatable :: surface(:,:), elev(:,:), blocked(:,:,:,:), seeds(:)
  real(real64), allocatable :: stokes(:,:,:,:), paths(:,:,:,:), weights(:,:,:)
  real(real64), allocatable :: contrib(:,:,:,:,:), sing(:,:,:,:,:), mult(:,:,:,:,:)
  real(real64), allocatable :: stokes_sing(:,:,:,:), stokes_mult(:,:,:,:)
  integer :: d1(5)
  real(real64) :: size_albedo(2), t0, t1
  integer(int64) :: n_photons
  type(new_scatterer_info_pol), allocatable :: scatterers(:)

  call get_command_argument(1, table_path)
  call get_command_argument(2, arg); read(arg, *) ppc
  call get_command_argument(3, arg); read(arg, *) rounds
  call get_command_argument(4, arg); read(arg, *) seed
  call get_command_argument(5, output_path)
  save_scats = .false.
  if (command_argument_count() >= 6) then
     call get_command_argument(6, arg)
     save_scats = trim(arg) == '1'
  end if

  open(newunit=unit_in, file=trim(table_path), access='stream', form='unformatted', &
       status='old', action='read', iostat=stat)
  if (stat /= 0) error stop 'cannot open transport tables'
  read(unit_in) header, geometry
  if (nint(header(1)) /= 1) error stop 'unsupported transport export version'
  nx = nint(header(2)); ny = nint(header(3)); nz = nint(header(4)); nmat = nint(header(5))
  n_angles = nint(header(6)); n_types = nint(header(7)); k_source = nint(header(8))
  k_detector = nint(header(9)); k_register = nint(header(10))
  save_path = nint(header(11)) == 1; backward = nint(header(12)) == 1
  xunit = geometry(1); yunit = geometry(2); depol = geometry(3)
  vartheta0 = geometry(4); phi0 = geometry(5); phi = geometry(6)

  allocate(zunit(nz), vartheta(n_angles), surface_r(nx, ny), albedo(nx, ny), elev_r(nx, ny))
  allocate(beta(nx, ny, nz), ext(n_types + 1, nx, ny, nz), ssalb_gas(nz))
  read(unit_in) zunit, vartheta, surface_r, albedo, elev_r, beta, ext, ssalb_gas
  allocate(scatterers(max(n_types, 1)))
  do is = 1, n_types
     read(unit_in) size_albedo
     n = nint(size_albedo(1))
     scatterers(is)%n_scat_angles = n
     allocate(scatterers(is)%scat_angles(n), scatterers(is)%CDF_bound(n))
     allocate(scatterers(is)%F11(n), scatterers(is)%F22(n), scatterers(is)%F33(n))
     allocate(scatterers(is)%F12(n), scatterers(is)%F34(n), scatterers(is)%F44(n))
     allocate(scatterers(is)%ssalb_aer)
     scatterers(is)%ssalb_aer = size_albedo(2)
     read(unit_in) scatterers(is)%scat_angles, scatterers(is)%CDF_bound
     read(unit_in) scatterers(is)%F11, scatterers(is)%F22, scatterers(is)%F33, scatterers(is)%F12
     if (nmat == 4) then
        read(unit_in) scatterers(is)%F34, scatterers(is)%F44
     else
        scatterers(is)%F34 = 0.d0
        scatterers(is)%F44 = scatterers(is)%F33
     end if
  end do
  close(unit_in)

  allocate(surface(nx, ny), elev(nx, ny), blocked(nx, ny, nz, n_angles), tau_gas(nz, n_angles))
  surface = nint(surface_r); elev = nint(elev_r); blocked = 0; tau_gas = 0.d0
  allocate(stokes(nmat, nx, ny, n_angles), paths(nx, ny, nz, n_angles), weights(nx, ny, n_angles))
  ! Dimensions of the decomposition arrays (nmat_1 .. n_angles_1 of monte_carlo): full size
  ! with save_scats, else placeholders.
  d1 = [1, 1, 1, 1, 1]
  if (save_scats) d1 = [nx, ny, nz, nmat, n_angles]
  allocate(contrib(d1(4), d1(1), d1(2), 0:d1(3), d1(5)), sing(d1(4), d1(1), d1(2), 0:d1(3), d1(5)))
  allocate(mult(d1(4), d1(1), d1(2), 0:d1(3), d1(5)))
  allocate(stokes_sing(d1(4), d1(1), d1(2), d1(5)), stokes_mult(d1(4), d1(1), d1(2), d1(5)))

  nt = 1
  !$OMP PARALLEL
  nt = omp_get_num_threads()
  !$OMP END PARALLEL
  if (save_scats .and. nt /= 1) error stop 'save_scats runs with one thread'
  if (mod(ppc, nt) /= 0) error stop 'photons per column must be divisible by threads'
  n_photons = int(ppc, int64) * int(nx, int64) * int(ny, int64) / nt

  call random_seed(size=seed_size)
  !$OMP PARALLEL PRIVATE(myid, seeds, idx)
  allocate(seeds(seed_size))
  myid = omp_get_thread_num()
  seeds = [(seed + myid + idx, idx = 1, seed_size)]
  call random_seed(put=seeds)
  deallocate(seeds)
  !$OMP END PARALLEL

  open(newunit=unit_out, file=trim(output_path), access='stream', form='unformatted', &
       status='replace', action='write')
  write(output_unit, '(A,6I8)') 'MONKI_REF_READY', nx, ny, nz, nmat, n_angles, nt
  flush(output_unit)

  do iround = 1, rounds
     stokes = 0.d0; paths = 0.d0; weights = 0.d0
     t0 = omp_get_wtime()
     if (save_scats) then
        contrib = 0.d0; sing = 0.d0; mult = 0.d0; stokes_sing = 0.d0; stokes_mult = 0.d0
        call monte_carlo(nx, ny, nz, nmat, n_angles, &
             d1(1), d1(2), d1(3), d1(4), d1(5), &
             nx, ny, nz, nmat, n_angles, &
             n_types, scatterers, &
             n_photons, &
             vartheta0, phi0, &
             vartheta, phi, &
             xunit, yunit, zunit, &
             surface, albedo, elev, &
             beta, ext, ssalb_gas, depol, &
             k_detector, &
             tau_gas, nz, &
             blocked, .false., &
             stokes, &
             .false., &
             paths, &
             weights, &
             .true., &
             contrib, sing, mult, &
             stokes_sing, stokes_mult, &
             k_register, k_source, backward)
     else
        !$OMP PARALLEL REDUCTION(+:stokes, paths, weights) &
        !$OMP& PRIVATE(contrib, sing, mult, stokes_sing, stokes_mult)
        call monte_carlo(nx, ny, nz, nmat, n_angles, &
             1, 1, 1, 1, 1, &
             nx, ny, nz, nmat, n_angles, &
             n_types, scatterers, &
             n_photons, &
             vartheta0, phi0, &
             vartheta, phi, &
             xunit, yunit, zunit, &
             surface, albedo, elev, &
             beta, ext, ssalb_gas, depol, &
             k_detector, &
             tau_gas, nz, &
             blocked, .false., &
             stokes, &
             save_path, &
             paths, &
             weights, &
             .false., &
             contrib, sing, mult, &
             stokes_sing, stokes_mult, &
             k_register, k_source, backward)
        !$OMP END PARALLEL
     end if
     t1 = omp_get_wtime()
     write(unit_out) stokes
     if (save_path .and. .not. save_scats) write(unit_out) paths, weights
     if (save_scats) write(unit_out) stokes_sing, stokes_mult, contrib, sing, mult
     write(output_unit, '(A,I8,F14.6)') 'MONKI_REF_ROUND', iround, t1 - t0
     flush(output_unit)
  end do
  close(unit_out)
end program monki_ref
