! ref_pes_h2ohe.f90 -- evaluates the reference's H2O-He potential
! (EXAMPLES_IO/SYS_TYPE_4/H2O_He_external_rot_wfn/PES_H2OHe/PES_H2O_He.f,
! compiled from its own source by oracle/Makefile) at given points, the way
! the reference calls it on the SYS_TYPE 4 grid.
!
! TEST INFRASTRUCTURE ONLY: tests/golden/make_h2ohe_vgrid.py runs it once
! to write a V-grid fixture.
!
! V_POT_TOP_ATOM passes (-beta, -gamma) to USER_DEFINED_PES
! (pes_sys_type.f:32-37), which hands (R, beta_in, gamma_in) to
! HeH2O_PES_routine; that converts the angles to degrees and calls
! arh2osapt(1, R, theta, phi, V) (R in bohr).  The routine's one-time
! POTENL set-up (MOLSCAT expansion tables) is not used by arh2osapt, which
! reads its own data files on its first call, so this driver calls
! arh2osapt with the same arguments.
!
! stdin: n, then n lines "R beta_in gamma_in" (bohr, rad: the arguments
! USER_DEFINED_PES receives); stdout: n values of V (cm^-1).
program ref_pes_h2ohe
  implicit none
  integer :: n, i
  real(8) :: r, b_in, g_in, theta, phi, v, pi
  pi = acos(-1.d0)
  read(*, *) n
  do i = 1, n
    read(*, *) r, b_in, g_in
    theta = b_in * 180.d0 / pi
    phi = g_in * 180.d0 / pi
    call arh2osapt(1, r, theta, phi, v)
    write(*, '(ES26.17E3)') v
  end do
end program ref_pes_h2ohe
