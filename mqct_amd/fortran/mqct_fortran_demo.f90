! mqct_fortran_demo.f90 -- a Fortran host driving libmqct_hip.so through
! mqct_hip_bridge, the way the MQCT driver would: diatom + atom basis and
! grid set up in Fortran (INI_ARRAYS / GRID_INTEGR order), matrix built on
! the GPU, MTRX_UF.DAT written with Fortran unformatted records in the
! PRINT_MIJ_USER layout (matrix_ini.f:6658-6769) beside the library's own
! writer, and a few CC-MQCT RK4 trajectories propagated on the GPU.
!
! Input: namelist &demo on the file named by the first argument.
! Output: MTRX_UF.DAT (Fortran), MTRX_UF_C.DAT (library), demo_out.bin
! (stream: the MIJ_SHIFTed Mat_el the trajectories used, Mat_el_der, the
! final sys_var of each trajectory, then the expansion-built Mat_el and
! Mat_el_der and the projected expansion_terms).
program mqct_fortran_demo
  use iso_c_binding
  use mqct_hip_bridge
  implicit none
  integer :: jmax = 2, n_beta = 16, n_r = 12, npar = 0, n_traj = 2
  real(8) :: rmin = 3d0, rmax = 20d0, be = 1.93d0, e_coll_cm = 500d0, mass_amu = 3.6d0, time_step = 20d0
  real(8) :: par(64) = 0d0, l_values(16) = 0d0
  namelist /demo/ jmax, n_beta, n_r, rmin, rmax, be, npar, par, e_coll_cm, mass_amu, time_step, n_traj, l_values
  real(8), parameter :: conv_unit_e = 1d0 / 8065.5448d0 / 27.21113845d0, amutoau = 1822.8885d0
  real(8), parameter :: pi = 4d0 * atan(1d0)
  character(len=256) :: fin
  integer :: states_size, number_of_channels, st, i, j, m, k, ir1, ir2, u
  integer(c_int64_t) :: total_size
  integer, allocatable :: j12(:), m12(:), indx_chann(:), j_ch(:), ind_mat(:, :), chan_qn(:)
  real(8), allocatable :: xb(:), wb(:), R_COM(:), Mat_el(:, :), Mat_el_der(:, :), Ej(:), sys_var(:, :)
  real(8), allocatable :: exp_terms(:, :), Mat_exp(:, :), Der_exp(:, :)
  integer, allocatable :: a_top(:, :)
  type(mqct_traj_t), allocatable :: traj(:)
  type(mqct_traj_stat_t), allocatable :: stats(:)
  type(c_ptr) :: h
  real(8) :: dt

  call get_command_argument(1, fin)
  open(newunit=u, file=trim(fin), status="old", action="read")
  read(u, nml=demo)
  close(u)

  ! channels j = 0..jmax, states (j, m = -j..j) channel-major
  number_of_channels = jmax + 1
  states_size = (jmax + 1)**2
  allocate(j_ch(number_of_channels), j12(states_size), m12(states_size), indx_chann(states_size))
  st = 0
  do i = 1, number_of_channels
    j_ch(i) = i - 1
    do m = -j_ch(i), j_ch(i)
      st = st + 1
      j12(st) = j_ch(i); m12(st) = m; indx_chann(st) = i
    end do
  end do
  ! coupled pairs st1 >= st2 with equal m, st1-major
  total_size = 0
  do i = 1, states_size
    do j = 1, i
      if (m12(i) == m12(j)) total_size = total_size + 1
    end do
  end do
  allocate(ind_mat(2, total_size))
  k = 0
  do i = 1, states_size
    do j = 1, i
      if (m12(i) == m12(j)) then
        k = k + 1
        ind_mat(1, k) = i; ind_mat(2, k) = j
      end if
    end do
  end do

  allocate(xb(n_beta), wb(n_beta), R_COM(n_r))
  call gauss_legendre(n_beta, xb, wb)
  do i = 1, n_r
    R_COM(i) = rmin + (rmax - rmin) / (n_r - 1d0) * (i - 1d0)
  end do
  ir1 = minloc(abs(R_COM - 6d0), 1)
  ir2 = minloc(abs(R_COM - 8d0), 1)

  allocate(Mat_el(n_r, total_size), Mat_el_der(n_r, total_size))
  call mqct_check(mqct_hip_create(0_c_int, h), "create")
  call mqct_gpu_matrix_make(h, 1, states_size, j12, m12, indx_chann, number_of_channels, j_ch, &
      n_beta, xb, wb, pi / 2d0, pi / 2d0, n_r, R_COM, total_size, ind_mat, conv_unit_e, &
      1d-12 * conv_unit_e, ir1, ir2, MQCT_PREC_FP64, Mat_el, Mat_el_der, &
      pes_kind=MQCT_PES_LEGENDRE, pes_par=par(1:npar))

  ! the same matrix through the PES expansion: the terms L = 0..6 of the
  ! Legendre potential projected from the handle's V grid
  ! (CALC_EXPANSION_TERMS), then EXPANSION_MATRIX_ELEMENT
  allocate(a_top(5, 7), exp_terms(n_r, 7), Mat_exp(n_r, total_size), Der_exp(n_r, total_size))
  a_top = 0
  do i = 1, 7
    a_top(1, i) = i - 1
  end do
  call mqct_gpu_expansion_matrix(h, 7, a_top, n_r, R_COM, exp_terms, .true., total_size, ind_mat, &
      conv_unit_e, 1d-12 * conv_unit_e, ir1, ir2, Mat_exp, Der_exp)

  ! MTRX_UF.DAT from Fortran
  open(newunit=u, file="MTRX_UF.DAT", form="unformatted", action="write", status="replace")
  write(u) 1
  write(u) number_of_channels
  write(u) states_size
  write(u) int(total_size)
  write(u) n_r
  do st = 1, states_size
    write(u) st, indx_chann(st), j12(st), m12(st), j_ch(indx_chann(st))
  end do
  do k = 1, int(total_size)
    write(u) k, ind_mat(1, k), ind_mat(2, k)
  end do
  do i = 1, n_r
    write(u) i, R_COM(i)
  end do
  do k = 1, int(total_size)
    write(u) k
    write(u) Mat_el(:, k)
  end do
  close(u)
  ! the same file from the library's writer
  allocate(chan_qn(number_of_channels))
  chan_qn = j_ch
  call mqct_check(mqct_hip_write_mtrx_uf("MTRX_UF_C.DAT" // c_null_char, 1_c_int, int(number_of_channels, c_int), &
      int(states_size, c_int), total_size, int(n_r, c_int), indx_chann, j12, m12, chan_qn, &
      reshape(ind_mat, [2 * int(total_size)]), R_COM, Mat_el), "write_mtrx_uf")

  ! MIJ_SHIFT (default YES): the driver's own compute-branch shift after the
  ! spline, before the propagation (matrix_ini.f:2226-2230)
  do i = 1, n_r
    Mat_el(i, :) = Mat_el(i, :) - Mat_el(n_r, :)
  end do

  ! CC-MQCT RK4 trajectories from j = 0 (state 1)
  allocate(Ej(states_size), traj(n_traj), stats(n_traj), sys_var(2 * states_size + 8, n_traj))
  do st = 1, states_size
    Ej(st) = be * j12(st) * (j12(st) + 1d0) * conv_unit_e
  end do
  dt = time_step / sqrt(e_coll_cm / 300d0)
  do i = 1, n_traj
    traj(i)%s_st = 1
    traj(i)%at_path = 0
    traj(i)%l_real = l_values(i)
    traj(i)%E_coll = e_coll_cm * conv_unit_e
    traj(i)%dt = dt
    traj(i)%eps = 1d-8
    traj(i)%hmin = 1d-4
    traj(i)%R_start = R_COM(n_r)
    traj(i)%time_lim = -1d0
    traj(i)%numb_oscl_prds = 1
    traj(i)%numb_orb_prds = 1
    traj(i)%max_steps = 200000
  end do
  call mqct_gpu_trajectories(h, MQCT_PROP_CC_RK4, states_size, Ej, j12, m12, total_size, ind_mat, n_r, &
      R_COM, Mat_el, Mat_el_der, mass_amu * amutoau, n_traj, traj, sys_var, stats)
  call mqct_check(mqct_hip_destroy(h), "destroy")

  open(newunit=u, file="demo_out.bin", form="unformatted", access="stream", action="write", status="replace")
  write(u) Mat_el, Mat_el_der, sys_var, Mat_exp, Der_exp, exp_terms
  close(u)
  write(*, '(A,I0,A,I0,A,I0,A)') "demo: ", states_size, " states, ", total_size, " pairs, ", n_traj, &
      " trajectories"
  do i = 1, n_traj
    write(*, '(A,I0,A,I0,A,ES12.4)') "  trajectory ", i, ": steps ", stats(i)%n_steps, ", norm error ", stats(i)%norm_err
  end do

contains

  ! Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n).
  subroutine gauss_legendre(n, x, w)
    integer, intent(in) :: n
    real(8), intent(out) :: x(n), w(n)
    integer :: i, k, it
    real(8) :: z, p0, p1, p2, dp, dz
    do i = 1, (n + 1) / 2
      z = cos(pi * (i - 0.25d0) / (n + 0.5d0))
      do it = 1, 100
        p0 = 1d0; p1 = z
        do k = 2, n
          p2 = ((2d0 * k - 1d0) * z * p1 - (k - 1d0) * p0) / k
          p0 = p1; p1 = p2
        end do
        if (n == 1) then
          p1 = z; p0 = 1d0
        end if
        dp = n * (z * p1 - p0) / (z * z - 1d0)
        dz = p1 / dp
        z = z - dz
        if (abs(dz) < 1d-15) exit
      end do
      x(i) = -z; x(n + 1 - i) = z
      w(i) = 2d0 / ((1d0 - z * z) * dp * dp); w(n + 1 - i) = w(i)
    end do
  end subroutine

end program
