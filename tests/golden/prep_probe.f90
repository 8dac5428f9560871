! Prepared tables from the reference's unchanged preparation routines (project-written probe).
!
! Calls rdsize, rdinput (main_routines_cartesian_mod.f90:32-487), molecules_earth
! (main_routines_general_mod.f90:33-321, which calls get_O3_crabs.f90:15-157) and aerosols
! (main_routines_general_mod.f90:629-951) on one positional fixture, then forms the transport
! tables as kernels/cpu/worker.f90:933-1010 does and writes them, in float64, in the
! transport.bin v1 order of kernels/cpu/worker.f90:1102-1120. After that stream it appends
! bmsca(nz), bmabs(nz), depol, baer(nscatt_types,nx,ny,nz), alt, pres, temp (nz+1 levels) and
! vmrgas(ngases,nz+1).
!
! Arguments: monki_home input_file output_file albedo theta0 phi0 phi n_angles vza...
program prep_probe
  use iso_fortran_env, only: real64
  use physical_parameters
  use new_data_types
  use main_routines_general
  use main_routines_cartesian
  implicit none
  character(len=500) :: infile, outfile, arg
  character(len=300) :: monki_home
  character(len=500) :: bmsca_file, bmabs_file
  character(len=30), allocatable :: coefile(:), coelut(:)
  character(len=5) :: background
  character(len=4) :: cloud_input_type
  character(len=2) :: wave_type
  integer :: nx, ny, nz, nmat, nscatt_types, n_scat_angles, ngases, nclouds, n_angles
  integer :: gas_col_type, detector_location, source_location, register_location, n_rounds
  integer :: i, j, k, isc, unit
  integer(8) :: n_photons_per_column
  logical :: read_bmsca_from_file, fix_depol, read_bmabs_from_file, save_path, save_scats
  logical :: stop_converged
  real(real64) :: x_dim, y_dim, wave_min, wave_max, dwave, depol_fixed, tol_error
  real(real64) :: asurf, theta0, phi0, phi, wav, depol, x_unit, y_unit
  real(real64), allocatable :: vartheta(:), alt(:), pres(:), temp(:), vmrgas(:,:)
  real(real64), allocatable :: cloud_abundance(:,:,:,:), baer(:,:,:,:), bmsca(:), bmabs(:)
  real(real64), allocatable :: z_unit(:), beta(:,:,:), ext(:,:,:,:), norm(:,:,:,:)
  real(real64), allocatable :: ssalb_gas(:), albedo(:,:)
  type(new_scatterer_info_pol), allocatable :: tables(:)
  real(real64), parameter :: eps = 1.0D-16

  call get_command_argument(1, arg); monki_home = arg
  call get_command_argument(2, infile)
  call get_command_argument(3, outfile)
  call get_command_argument(4, arg); read(arg, *) asurf
  call get_command_argument(5, arg); read(arg, *) theta0
  call get_command_argument(6, arg); read(arg, *) phi0
  call get_command_argument(7, arg); read(arg, *) phi
  call get_command_argument(8, arg); read(arg, *) n_angles
  allocate(vartheta(n_angles))
  do i = 1, n_angles
     call get_command_argument(8 + i, arg); read(arg, *) vartheta(i)
  end do

  call rdsize(infile, nx, ny, nz, x_dim, y_dim, nmat, nscatt_types, n_scat_angles, wave_type, &
       wave_min, wave_max, dwave, ngases, nclouds)
  allocate(alt(nz + 1), pres(nz + 1), temp(nz + 1), vmrgas(ngases, nz + 1))
  allocate(cloud_abundance(nscatt_types, nx, ny, nz), baer(nscatt_types, nx, ny, nz))
  allocate(coefile(nscatt_types), coelut(nscatt_types), bmsca(nz), bmabs(nz), z_unit(nz))
  allocate(beta(nx, ny, nz), ext(nscatt_types + 1, nx, ny, nz), norm(nscatt_types + 1, nx, ny, nz))
  allocate(ssalb_gas(nz), albedo(nx, ny), tables(nscatt_types))

  do isc = 1, nscatt_types
     call allocate_and_initialize_new_scatterer_info_pol(tables(isc), n_scat_angles)
  end do

  call rdinput(infile, nx, ny, nz, nscatt_types, ngases, nclouds, vmrgas, pres, temp, alt, &
       read_bmsca_from_file, bmsca_file, depol_fixed, gas_col_type, fix_depol, &
       read_bmabs_from_file, bmabs_file, background, coefile, coelut, cloud_input_type, &
       cloud_abundance, n_photons_per_column, n_rounds, save_path, save_scats, stop_converged, &
       tol_error, detector_location, source_location, register_location)

  if (wave_type == 'wn') then
     wav = 1.D7 / wave_min
  else
     wav = wave_min
  end if

  x_unit = x_dim / real(nx, 8)
  y_unit = y_dim / real(ny, 8)
  do k = 1, nz
     z_unit(k) = (alt(k + 1) - alt(k)) / 1000.D0
  end do
  albedo = asurf

  call molecules_earth(nz, pres, temp, alt, wav, depol, gas_col_type, bmsca, monki_home, ngases, &
       vmrgas, bmabs)
  if (fix_depol) depol = depol_fixed
  call aerosols(nmat, nscatt_types, tables, monki_home, coelut, coefile, wav, cloud_input_type, &
       cloud_abundance, nx, ny, nz, baer)

  beta = 0.D0
  ext = 0.D0
  norm = 0.D0
  ssalb_gas = 1.D0
  do k = 1, nz
     do j = 1, ny
        do i = 1, nx
           ext(nscatt_types + 1, i, j, k) = (bmsca(k) + bmabs(k)) / z_unit(k)
           if ((bmsca(k) + bmabs(k)) > eps) ssalb_gas(k) = bmsca(k) / (bmsca(k) + bmabs(k))
        end do
     end do
  end do
  do isc = 1, nscatt_types
     ext(isc, :, :, :) = baer(isc, :, :, :) / spread(spread(z_unit, 1, ny), 1, nx)
  end do
  do isc = 1, nscatt_types + 1
     beta = beta + ext(isc, :, :, :)
  end do
  do isc = 1, nscatt_types + 1
     where (beta > eps) norm(isc, :, :, :) = ext(isc, :, :, :) / beta
  end do

  open(newunit=unit, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(unit) real([1, nx, ny, nz, nmat, n_angles, nscatt_types, merge(nz, 0, source_location == 1), &
       merge(nz, 0, detector_location == 1), merge(nz, 0, register_location == 1), &
       merge(1, 0, save_path), 0], real64)
  write(unit) x_unit, y_unit, depol, 180.D0 - theta0, phi0, phi
  write(unit) z_unit, vartheta
  write(unit) spread(spread(-1.D0, 1, nx), 2, ny), albedo, spread(spread(1.D0, 1, nx), 2, ny)
  write(unit) beta, norm, ssalb_gas
  do isc = 1, nscatt_types
     write(unit) real(tables(isc)%n_scat_angles, real64), tables(isc)%ssalb_aer
     write(unit) tables(isc)%scat_angles, tables(isc)%CDF_bound
     write(unit) tables(isc)%F11, tables(isc)%F22, tables(isc)%F33, tables(isc)%F12
  end do
  write(unit) bmsca, bmabs, depol, baer, alt, pres, temp, vmrgas
  close(unit)
end program prep_probe
