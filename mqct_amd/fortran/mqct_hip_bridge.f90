! mqct_hip_bridge.f90 -- iso_c_binding interface of libmqct_hip.so for the
! MQCT Fortran driver.
!
! The interfaces follow include/mqct_hip.h one to one.  The drop-in
! wrappers at the end take the driver's own arrays:
!   mqct_gpu_matrix_make  replaces the k/R loop of MATRIX_MAKE
!                         (matrix_ini.f:1916-2016) and the V-grid fill of
!                         INI_POT_STORAGE (pot_comp.f:333-1788);
!   mqct_gpu_expansion_matrix replaces its expansion branch
!                         (matrix_ini.f:1950-1956, EXPANSION_MATRIX_ELEMENT)
!                         with CALC_EXPANSION_TERMS' projection if asked;
!   mqct_gpu_trajectories replaces the per-trajectory DOTRAJECT stepping
!                         (propagator.f:3205-4137) for a batch.
! A failing call ends the run with STOP and the library's message, as the
! reference stops on bad input.
module mqct_hip_bridge
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MQCT_OK = 0
  integer(c_int), parameter :: MQCT_PREC_FP64 = 0, MQCT_PREC_FP32 = 1, MQCT_PREC_BF16X3 = 2
  integer(c_int), parameter :: MQCT_PES_LEGENDRE = 0, MQCT_PES_REALSH = 1, MQCT_PES_SITES = 2
  integer(c_int), parameter :: MQCT_PROP_CC_RK4 = 0, MQCT_PROP_CC_ODEINT = 1, &
                               MQCT_PROP_AT_RK4 = 2, MQCT_PROP_AT_ODEINT = 3

  type, bind(C) :: mqct_grid_t
    integer(c_int) :: n_beta
    type(c_ptr) :: xb, wb
    real(c_double) :: xbm, xbl
    integer(c_int) :: n_gamma
    type(c_ptr) :: xg, wg
    real(c_double) :: xgm, xgl
    integer(c_int) :: n_alpha2
    type(c_ptr) :: xaa, waa
    real(c_double) :: xaam, xaal
    integer(c_int) :: n_beta2
    type(c_ptr) :: xbb, wbb
    real(c_double) :: xbbm, xbbl
    integer(c_int) :: n_gamma2
    type(c_ptr) :: xgg, wgg
    real(c_double) :: xggm, xggl
    integer(c_int) :: eq_gamma
    integer(c_int) :: ref_xz
  end type

  type, bind(C) :: mqct_traj_t
    integer(c_int) :: s_st, at_path
    real(c_double) :: l_real, E_coll, dt, eps, hmin, R_start, time_lim
    integer(c_int) :: numb_oscl_prds, numb_orb_prds, max_steps
  end type

  type, bind(C) :: mqct_traj_stat_t
    real(c_double) :: t_end
    integer(c_int) :: n_steps, n_rhs, fail
    real(c_double) :: norm_err
  end type

  interface
    integer(c_int) function mqct_hip_create(device, h) bind(C, name="mqct_hip_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: h
    end function
    integer(c_int) function mqct_hip_destroy(h) bind(C, name="mqct_hip_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mqct_hip_last_error(buf, n) bind(C, name="mqct_hip_last_error")
      import :: c_int, c_char
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function mqct_hip_version() bind(C, name="mqct_hip_version")
      import :: c_int
    end function
    integer(c_int) function mqct_hip_set_basis(h, coll_type, n_states, j12, m12, indx_chann, &
        n_channels, j1_ch, j2_ch, ldvec, eig1, eig2) bind(C, name="mqct_hip_set_basis")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: coll_type, n_states, n_channels, ldvec
      integer(c_int) :: j12(*), m12(*), indx_chann(*), j1_ch(*)
      type(c_ptr), value :: j2_ch, eig1, eig2
    end function
    integer(c_int) function mqct_hip_set_grid(h, g) bind(C, name="mqct_hip_set_grid")
      import :: c_int, c_ptr, mqct_grid_t
      type(c_ptr), value :: h
      type(mqct_grid_t) :: g
    end function
    integer(c_int) function mqct_hip_set_vgrid(h, n_r, R_com, V) bind(C, name="mqct_hip_set_vgrid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n_r
      real(c_double) :: R_com(*), V(*)
    end function
    integer(c_int) function mqct_hip_eval_pes(h, n_r, R_com, pes_kind, n_par, par) &
        bind(C, name="mqct_hip_eval_pes")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n_r, pes_kind, n_par
      real(c_double) :: R_com(*), par(*)
    end function
    integer(c_int) function mqct_hip_matrix_build(h, precision, total_size, ind_mat, conv_unit_e, &
        mij_zero_cut, ir_cut1, ir_cut2, mat_el, mat_el_der, skipped) bind(C, name="mqct_hip_matrix_build")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: precision, ir_cut1, ir_cut2
      integer(c_int64_t), value :: total_size
      integer(c_int) :: ind_mat(*)
      real(c_double), value :: conv_unit_e, mij_zero_cut
      real(c_double) :: mat_el(*), mat_el_der(*)
      type(c_ptr), value :: skipped
    end function
    integer(c_int) function mqct_hip_build_timing(h, ms5) bind(C, name="mqct_hip_build_timing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: ms5(5)
    end function
    integer(c_int) function mqct_hip_load_matrix(h, n_states, Ej, j12, m12, total_size, ind_mat, n_r, &
        R_com, mat_el, mat_el_der, mass, theta_on) bind(C, name="mqct_hip_load_matrix")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: n_states, n_r, theta_on
      integer(c_int64_t), value :: total_size
      real(c_double) :: Ej(*), R_com(*), mat_el(*), mat_el_der(*)
      integer(c_int) :: j12(*), m12(*), ind_mat(*)
      real(c_double), value :: mass
    end function
    integer(c_int) function mqct_hip_load_paths(h, n_paths, path_off, records) bind(C, name="mqct_hip_load_paths")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: n_paths
      integer(c_int64_t) :: path_off(*)
      real(c_double) :: records(*)
    end function
    integer(c_int) function mqct_hip_propagate(h, mode, n_traj, traj, batched, q_out, stat_out) &
        bind(C, name="mqct_hip_propagate")
      import :: c_int, c_ptr, c_double, mqct_traj_t, mqct_traj_stat_t
      type(c_ptr), value :: h
      integer(c_int), value :: mode, n_traj, batched
      type(mqct_traj_t) :: traj(*)
      real(c_double) :: q_out(*)
      type(mqct_traj_stat_t) :: stat_out(*)
    end function
    integer(c_int) function mqct_hip_write_mtrx_uf(path, coll_type, n_channels, n_states, total_size, &
        n_r, indx_chann, j12, m12, chan_qn, ind_mat, R_com, mat_el) bind(C, name="mqct_hip_write_mtrx_uf")
      import :: c_int, c_char, c_double, c_int64_t
      character(kind=c_char) :: path(*)
      integer(c_int), value :: coll_type, n_channels, n_states, n_r
      integer(c_int64_t), value :: total_size
      integer(c_int) :: indx_chann(*), j12(*), m12(*), chan_qn(*), ind_mat(*)
      real(c_double) :: R_com(*), mat_el(*)
    end function
    integer(c_int) function mqct_hip_set_mij_shift(h, on) bind(C, name="mqct_hip_set_mij_shift")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function mqct_hip_mij_spline(h, n_r, R_com, total_size, mat_el, mat_el_der) &
        bind(C, name="mqct_hip_mij_spline")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: n_r
      integer(c_int64_t), value :: total_size
      real(c_double) :: R_com(*), mat_el(*), mat_el_der(*)
    end function
    integer(c_int) function mqct_hip_use_built_matrix(h, Ej, mass, theta_on) bind(C, name="mqct_hip_use_built_matrix")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: Ej(*)
      real(c_double), value :: mass
      integer(c_int), value :: theta_on
    end function
    integer(c_int) function mqct_hip_set_identical(h, identical_pes, parity_state, parity_state_bk, same_level) &
        bind(C, name="mqct_hip_set_identical")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: identical_pes
      type(c_ptr), value :: parity_state, parity_state_bk, same_level
    end function
    integer(c_int) function mqct_hip_set_expansion_terms(h, n_terms, a_top) &
        bind(C, name="mqct_hip_set_expansion_terms")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: n_terms
      integer(c_int) :: a_top(*)
    end function
    integer(c_int) function mqct_hip_expansion_project(h, terms) bind(C, name="mqct_hip_expansion_project")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: terms(*)
    end function
    integer(c_int) function mqct_hip_expansion_build(h, n_r, R_com, terms, total_size, ind_mat, conv_unit_e, &
        mij_zero_cut, ir_cut1, ir_cut2, mat_el, mat_el_der, skipped) bind(C, name="mqct_hip_expansion_build")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: n_r, ir_cut1, ir_cut2
      integer(c_int64_t), value :: total_size
      type(c_ptr), value :: R_com, terms
      integer(c_int) :: ind_mat(*)
      real(c_double), value :: conv_unit_e, mij_zero_cut
      real(c_double) :: mat_el(*), mat_el_der(*)
      type(c_ptr), value :: skipped
    end function
    integer(c_int) function mqct_hip_cross_sections(h, n_traj, n_states, ldq, q, n_channels, indx_chann, &
        n_groups, weight, sigma, opacity) bind(C, name="mqct_hip_cross_sections")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n_traj, n_states, ldq, n_channels, n_groups
      real(c_double) :: q(*), weight(*), sigma(*)
      integer(c_int) :: indx_chann(*)
      type(c_ptr), value :: opacity
    end function
    integer(c_int) function mqct_hip_write_at_path(dir, j12, m12, l, parity, n_rec, records) &
        bind(C, name="mqct_hip_write_at_path")
      import :: c_int, c_char, c_double
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: j12, m12, l, parity, n_rec
      real(c_double) :: records(*)
    end function
    integer(c_int) function mqct_hip_read_at_path(dir, j12, m12, l, parity, cap, records, n_rec) &
        bind(C, name="mqct_hip_read_at_path")
      import :: c_int, c_char, c_double
      character(kind=c_char) :: dir(*)
      integer(c_int), value :: j12, m12, l, parity, cap
      real(c_double) :: records(*)
      integer(c_int) :: n_rec
    end function
  end interface

contains

  subroutine mqct_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, n
    if (rc == MQCT_OK) return
    n = mqct_hip_last_error(buf, 512)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    write(*, '(A,": ",A)') what, trim(msg)
    stop "MQCT GPU LIBRARY ERROR"
  end subroutine

  ! MATRIX_MAKE on the GPU for a diatom + atom basis (coll_type 1) with the
  ! potential given on the grid by the caller (V_2_1(n_beta, n_r), E_UNITS)
  ! or, when pes_par is present, by a synthetic potential.
  subroutine mqct_gpu_matrix_make(h, coll_type, states_size, j12, m12, indx_chann, number_of_channels, &
      j_ch, n_beta, xb, wb, xbm, xbl, n_r, R_COM, total_size, ind_mat, conv_unit_e, mij_zero_cut, &
      ir_cut1, ir_cut2, precision, Mat_el, Mat_el_der, V, pes_kind, pes_par)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: coll_type, states_size, number_of_channels, n_beta, n_r, ir_cut1, ir_cut2, precision
    integer, intent(in) :: j12(states_size), m12(states_size), indx_chann(states_size), j_ch(number_of_channels)
    integer(c_int64_t), intent(in) :: total_size
    integer, intent(in) :: ind_mat(2, total_size)
    real(c_double), intent(in), target :: xb(n_beta), wb(n_beta)
    real(c_double), intent(in) :: xbm, xbl, R_COM(n_r), conv_unit_e, mij_zero_cut
    real(c_double), intent(out) :: Mat_el(n_r, total_size), Mat_el_der(n_r, total_size)
    real(c_double), intent(in), optional :: V(n_beta, n_r)
    integer, intent(in), optional :: pes_kind
    real(c_double), intent(in), optional :: pes_par(:)
    type(mqct_grid_t) :: g
    real(c_double), target :: zero(1), one(1)
    integer(c_int), allocatable :: jj(:), mm(:), ic(:), jc(:), im(:)
    real(c_double), allocatable :: rr(:), vv(:), pp(:)
    if (coll_type /= 1) stop "mqct_gpu_matrix_make: this wrapper takes coll_type 1"
    allocate(jj(states_size), mm(states_size), ic(states_size), jc(number_of_channels), im(2 * total_size))
    jj = j12; mm = m12; ic = indx_chann; jc = j_ch
    im = reshape(ind_mat, [2 * total_size])
    call mqct_check(mqct_hip_set_basis(h, 1_c_int, int(states_size, c_int), jj, mm, ic, &
                    int(number_of_channels, c_int), jc, c_null_ptr, 0_c_int, c_null_ptr, c_null_ptr), "set_basis")
    zero = 0d0
    one = 1d0
    g%n_beta = n_beta; g%xb = c_loc(xb); g%wb = c_loc(wb); g%xbm = xbm; g%xbl = xbl
    g%n_gamma = 1; g%xg = c_loc(zero); g%wg = c_loc(one); g%xgm = 0d0; g%xgl = 1d0
    g%n_alpha2 = 1; g%xaa = c_loc(zero); g%waa = c_loc(one); g%xaam = 0d0; g%xaal = 1d0
    g%n_beta2 = 1; g%xbb = c_loc(zero); g%wbb = c_loc(one); g%xbbm = 0d0; g%xbbl = 1d0
    g%n_gamma2 = 1; g%xgg = c_loc(zero); g%wgg = c_loc(one); g%xggm = 0d0; g%xggl = 1d0
    g%eq_gamma = 1; g%ref_xz = 1
    call mqct_check(mqct_hip_set_grid(h, g), "set_grid")
    allocate(rr(n_r))
    rr = R_COM
    if (present(V)) then
      allocate(vv(n_beta * n_r))
      vv = reshape(V, [n_beta * n_r])
      call mqct_check(mqct_hip_set_vgrid(h, int(n_r, c_int), rr, vv), "set_vgrid")
    else
      allocate(pp(size(pes_par)))
      pp = pes_par
      call mqct_check(mqct_hip_eval_pes(h, int(n_r, c_int), rr, int(pes_kind, c_int), int(size(pp), c_int), pp), &
                      "eval_pes")
    end if
    call mqct_check(mqct_hip_matrix_build(h, int(precision, c_int), total_size, im, conv_unit_e, mij_zero_cut, &
                    int(ir_cut1, c_int), int(ir_cut2, c_int), Mat_el, Mat_el_der, c_null_ptr), "matrix_build")
  end subroutine

  ! MATRIX_MAKE's expansion branch (matrix_ini.f:1950-1956) on the GPU for the
  ! basis already on the handle: the terms A_TOP(5, nterms) and either the
  ! coefficients expansion_terms(n_r, nterms) as READ_EXPANSION_TERMS read
  ! them (project = .false.), or their projection from the handle's V grid,
  ! returned in expansion_terms (CALC_EXPANSION_TERMS, project = .true.).
  subroutine mqct_gpu_expansion_matrix(h, nterms, A_TOP, n_r, R_COM, expansion_terms, project, total_size, &
      ind_mat, conv_unit_e, mij_zero_cut, ir_cut1, ir_cut2, Mat_el, Mat_el_der)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: nterms, n_r, ir_cut1, ir_cut2
    integer, intent(in) :: A_TOP(5, nterms)
    real(c_double), intent(in), target :: R_COM(n_r)
    real(c_double), intent(inout), target :: expansion_terms(n_r, nterms)
    logical, intent(in) :: project
    integer(c_int64_t), intent(in) :: total_size
    integer, intent(in) :: ind_mat(2, total_size)
    real(c_double), intent(in) :: conv_unit_e, mij_zero_cut
    real(c_double), intent(out) :: Mat_el(n_r, total_size), Mat_el_der(n_r, total_size)
    integer(c_int), allocatable :: at(:), im(:)
    allocate(at(5 * nterms), im(2 * total_size))
    at = reshape(A_TOP, [5 * nterms])
    im = reshape(ind_mat, [2 * total_size])
    call mqct_check(mqct_hip_set_expansion_terms(h, int(nterms, c_int), at), "set_expansion_terms")
    if (project) call mqct_check(mqct_hip_expansion_project(h, expansion_terms), "expansion_project")
    call mqct_check(mqct_hip_expansion_build(h, int(n_r, c_int), c_loc(R_COM), c_loc(expansion_terms), &
                    total_size, im, conv_unit_e, mij_zero_cut, int(ir_cut1, c_int), int(ir_cut2, c_int), &
                    Mat_el, Mat_el_der, c_null_ptr), "expansion_build")
  end subroutine

  ! DOTRAJECT for a batch of CC-MQCT trajectories from the matrix in memory
  ! (READ_MTRX path): sys_var_out(2*states_size+8, n_traj) receives the
  ! final amplitudes and classical variables of each trajectory.
  subroutine mqct_gpu_trajectories(h, mode, states_size, Ej, j12, m12, total_size, ind_mat, n_r, R_COM, &
      Mat_el, Mat_el_der, massAB_M, n_traj, traj, sys_var_out, stats)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: mode, states_size, n_r, n_traj
    integer(c_int64_t), intent(in) :: total_size
    real(c_double), intent(in) :: Ej(states_size), R_COM(n_r), massAB_M
    real(c_double), intent(in) :: Mat_el(n_r, total_size), Mat_el_der(n_r, total_size)
    integer, intent(in) :: j12(states_size), m12(states_size), ind_mat(2, total_size)
    type(mqct_traj_t), intent(in) :: traj(n_traj)
    real(c_double), intent(out) :: sys_var_out(2 * states_size + 8, n_traj)
    type(mqct_traj_stat_t), intent(out) :: stats(n_traj)
    integer(c_int), allocatable :: jj(:), mm(:), im(:)
    real(c_double), allocatable :: e(:), rr(:), m1(:), m2(:)
    type(mqct_traj_t), allocatable :: tr(:)
    allocate(jj(states_size), mm(states_size), im(2 * total_size), e(states_size), rr(n_r))
    allocate(m1(n_r * total_size), m2(n_r * total_size), tr(n_traj))
    jj = j12; mm = m12; im = reshape(ind_mat, [2 * total_size]); e = Ej; rr = R_COM
    m1 = reshape(Mat_el, [n_r * total_size]); m2 = reshape(Mat_el_der, [n_r * total_size]); tr = traj
    call mqct_check(mqct_hip_load_matrix(h, int(states_size, c_int), e, jj, mm, total_size, im, int(n_r, c_int), &
                    rr, m1, m2, massAB_M, 0_c_int), "load_matrix")
    call mqct_check(mqct_hip_propagate(h, int(mode, c_int), int(n_traj, c_int), tr, 0_c_int, sys_var_out, stats), &
                    "propagate")
  end subroutine

end module mqct_hip_bridge
