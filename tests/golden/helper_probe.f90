! Expected values from the reference's unchanged polarization helpers (project-written probe).
! Reads cases from stdin and prints results; see make_helper_vectors.py.
!
! Input:  ngeom, then ngeom lines "uo u1 phio phi1 depol I Q U"
!         nlut, then nlut lines "cos f11 f22 f33 f12 f34 f44" (ascending cosines), then ncos
!         cosines; then n4, then n4 lines "uo u1 phio phi1 depol I Q U V" (full Stokes vector)
! Output: per geometry line: cos_sca, C1 S1 C2 S2, Rayleigh F11 F22 F33 F12 (/4), P*S (3)
!         per cosine: mie_scat_matrix F11 F22 F33 F12 (/4)
!         per full-Stokes line (nmat 4): Rayleigh F44 (/4), Rayleigh P*S (4), tabulated F34 and
!         F44 (/4) at the line's scattering angle, tabulated P*S (4)
program helper_probe
  use iso_fortran_env, only: real64
  use mc_routines_general
  use new_data_types
  implicit none
  integer :: ngeom, nlut, ncos, n, n4
  real(real64) :: uo, u1, phio, phi1, depol, s(3), c, f(3,3), l1(3,3), l2(3,3), p(3,3)
  real(real64) :: s4(4), f4(4,4), g4(4,4), l14(4,4), l24(4,4), p4(4,4), q4(4,4)
  type(new_scatterer_info_pol) :: table

  read(*, *) ngeom
  do n = 1, ngeom
     read(*, *) uo, u1, phio, phi1, depol, s
     c = uo * u1 + sqrt(1 - uo**2) * sqrt(1 - u1**2) * cos(phio - phi1)
     call ray_scat_matrix(3, c, depol, f)
     call stokes_rot_mat(l1, l2, c, uo, u1, phio, phi1, 3)
     call matmul_stokes_rot(3, f, l1, l2, p)
     write(*, '(12ES25.16)') c, l1(2,2), l1(3,2), l2(2,2), l2(3,2), f(1,1), f(2,2), f(3,3), &
          f(1,2), matmul(p, s)
  end do
  read(*, *) nlut
  table%n_scat_angles = nlut
  allocate(table%scat_angles(nlut), table%F11(nlut), table%F22(nlut), table%F33(nlut))
  allocate(table%F12(nlut), table%F34(nlut), table%F44(nlut), table%CDF_bound(nlut))
  do n = 1, nlut
     read(*, *) table%scat_angles(n), table%F11(n), table%F22(n), table%F33(n), table%F12(n), &
          table%F34(n), table%F44(n)
  end do
  read(*, *) ncos
  do n = 1, ncos
     read(*, *) c
     call mie_scat_matrix(3, c, table, nlut, f)
     write(*, '(4ES25.16)') f(1,1), f(2,2), f(3,3), f(1,2)
  end do
  read(*, *) n4
  do n = 1, n4
     read(*, *) uo, u1, phio, phi1, depol, s4
     c = uo * u1 + sqrt(1 - uo**2) * sqrt(1 - u1**2) * cos(phio - phi1)
     call ray_scat_matrix(4, c, depol, f4)
     call stokes_rot_mat(l14, l24, c, uo, u1, phio, phi1, 4)
     call matmul_stokes_rot(4, f4, l14, l24, p4)
     call mie_scat_matrix(4, c, table, nlut, g4)
     call matmul_stokes_rot(4, g4, l14, l24, q4)
     write(*, '(11ES25.16)') f4(4,4), matmul(p4, s4), g4(3,4), g4(4,4), matmul(q4, s4)
  end do
end program helper_probe
