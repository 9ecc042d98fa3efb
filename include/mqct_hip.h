/*
 * mqct_hip.h -- C ABI of libmqct_hip.so, the MI355X (gfx950) library that
 * replaces MQCT's two data-parallel hot paths on one GPU.
 *
 * Conventions (match the Fortran host so iso_c_binding can bind directly):
 *  - all pointers are host memory owned by the caller unless a name ends in
 *    _dev; arrays are in Fortran column-major order as the reference keeps
 *    them (Mat_el(n_r, total_size) is n_r contiguous R values per element);
 *  - state numbers, channel numbers and ind_mat entries are 1-based;
 *  - energies and matrix elements are Hartree, distances bohr, angles rad;
 *  - every call returns 0 on success and a negative code on failure; the
 *    message is available through mqct_hip_last_error (the Fortran bridge
 *    turns a failure into STOP "<message>", the reference's convention);
 *  - one host thread per handle; a handle owns its device buffers and its
 *    stream and is not re-entrant.
 */
#ifndef MQCT_HIP_H
#define MQCT_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mqct_hip mqct_hip_t;

enum {
  MQCT_OK = 0,
  MQCT_ERR_ARG = -1,       /* bad argument / unsupported configuration */
  MQCT_ERR_HIP = -2,       /* HIP runtime failure */
  MQCT_ERR_STATE = -3,     /* call order (e.g. build before set_basis) */
  MQCT_ERR_NOMEM = -4
};

/* precision forms of the matrix build */
enum {
  MQCT_PREC_FP64 = 0,      /* FP64 MFMA, FP64 accumulation */
  MQCT_PREC_FP32 = 1,      /* FP32 MFMA, FP32 chunk sums added in FP64 */
  MQCT_PREC_BF16X3 = 2,    /* split-BF16 hi*hi + hi*lo + lo*hi on MFMA */
  MQCT_PREC_SURROGATE = 3  /* reserved: fill V with mqct_hip_surrogate_vgrid */
};

/* synthetic potentials evaluated on the GPU (parameter layouts in
 * mqct_amd/workloads.py) */
enum { MQCT_PES_LEGENDRE = 0, MQCT_PES_REALSH = 1, MQCT_PES_SITES = 2 };

/* propagation modes (PROPAGATOR= keyword x AT_APPROX= keyword) */
enum {
  MQCT_PROP_CC_RK4 = 0,    /* DERIVS_BK + rk4, fixed step */
  MQCT_PROP_CC_ODEINT = 1, /* DERIVS_BK + odeint per outer step */
  MQCT_PROP_AT_RK4 = 2,    /* DERIVS_BK_adia + rk4 on a recorded path */
  MQCT_PROP_AT_ODEINT = 3  /* DERIVS_BK_adia + odeint over the path */
};

/* ---------------------------------------------------------------------- */
/* Library lifetime.  No reference counterpart (the reference has no device). */
int mqct_hip_create(int device, mqct_hip_t **h);
int mqct_hip_destroy(mqct_hip_t *h);
int mqct_hip_last_error(char *buf, int len);
int mqct_hip_version(void);

/* ---------------------------------------------------------------------- */
/* Matrix build: replaces the loop of MATRIX_MAKE over k and R that calls
 * INTEGRATOR_TEMP (MQCT_SOURCE_CODES/matrix_ini.f:1916-2016, element
 * integrand pot_comp.f:1790-3737) together with the V-grid fill of
 * INI_POT_STORAGE (pot_comp.f:333-1788).                                  */

/* Basis: the state tables INI_ARRAYS builds (matrix_ini.f:255-545) and the
 * channel quantum numbers and eigenvector table ASYM_TOP_VECTORS keeps
 * (pot_comp.f:1-254, M_VECTORS / M1_VECTORS / M2_VECTORS).
 * coll_type: 1 (diatom+atom), 4 (asym. top+atom), 0 (asym.+asym. top,
 * distinguishable).  eig1/eig2 are [n_channels][ldvec] row-major with c_K,
 * K=-J..J in the first 2J+1 slots (NULL for type 1; eig2 only for type 0). */
int mqct_hip_set_basis(mqct_hip_t *h, int coll_type, int n_states,
                       const int *j12, const int *m12, const int *indx_chann,
                       int n_channels, const int *j1_ch, const int *j2_ch,
                       int ldvec, const double *eig1, const double *eig2);

/* IDENTICAL= and IDENTICAL_PES= of a type-0 basis, after set_basis.
 * parity_state (+-1, the exchange parity P of each state) and
 * parity_state_bk ((-1)^j12 times the channel's inversion and kappa signs)
 * are INI_ARRAYS' arrays (matrix_ini.f:486-528); same_level[c] is 1 when
 * both partners of channel c are in the same level (j, Ka, Kc).  With
 * parity_state the pairs are those of equal m12 and equal P
 * (matrix_ini.f:556-630) and the basis is of identical partners, which the
 * reference requires to come with identical_pes; with parity_state NULL
 * the partners are distinguishable and identical_pes alone marks the term
 * list as that of an exchange-symmetric PES (EXPANSION_MATRIX_ELEMENT's
 * second symmetry term).  Identical partners are built through the
 * expansion (mqct_hip_expansion_build); mqct_hip_matrix_build refuses
 * them.  Clears the expansion terms. */
int mqct_hip_set_identical(mqct_hip_t *h, int identical_pes,
                           const int *parity_state, const int *parity_state_bk,
                           const int *same_level);

/* Angular quadrature: GRID_INTEGR module state (iotest.f:62-90) as
 * INI_POT_STORAGE fills it (pot_comp.f:374-382, 586-611, 1628-1709).  Nodes
 * x* on [-1,1] with weights w*, mapped as angle = x*m + x*l * x (or angle =
 * x for equidistant gamma/alpha grids, eq_gamma=1).  Unused dimensions have
 * n=1, x=0, w=1, m=0, l=1.  ref_xz=0 selects the REF_XZ=NO integrand of
 * pot_comp.f:2544-2553. */
typedef struct {
  int n_beta;   const double *xb;  const double *wb;  double xbm, xbl;
  int n_gamma;  const double *xg;  const double *wg;  double xgm, xgl;
  int n_alpha2; const double *xaa; const double *waa; double xaam, xaal;
  int n_beta2;  const double *xbb; const double *wbb; double xbbm, xbbl;
  int n_gamma2; const double *xgg; const double *wgg; double xggm, xggl;
  int eq_gamma;
  int ref_xz;
} mqct_grid_t;
int mqct_hip_set_grid(mqct_hip_t *h, const mqct_grid_t *g);

/* V on the grid from the host, in E_UNITS, in the POT_STORE layouts
 * (iotest.f:33-48): type 1 V_2_1(n_beta, n_r); type 4 V_3_1(n_gamma,
 * n_beta, n_r); type 0 V_3_3(n_gamma2, n_gamma1, n_beta2, n_beta1, n_alpha2,
 * n_r).  R_com in bohr (the PES was called with R_COM*conv_unit_r). */
int mqct_hip_set_vgrid(mqct_hip_t *h, int n_r, const double *R_com,
                       const double *V);

/* Same V grid computed on the GPU from a synthetic potential (stands in for
 * USER_DEFINED_PES called through pes_sys_type.f:2-97). */
int mqct_hip_eval_pes(mqct_hip_t *h, int n_r, const double *R_com,
                      int pes_kind, int n_par, const double *par);

/* The build.  Produces Mat_el(n_r, total_size) = quadrature * conv_unit_e,
 * zeroes element k when max(|M(ir_cut1)|, |M(ir_cut2)|) < mij_zero_cut
 * (Hartree; matrix_ini.f:1991-1997; ir_cut* 1-based), and Mat_el_der =
 * clamped-spline second derivatives (matrix_ini.f:2005-2016).  ind_mat is
 * 2 x total_size (matrix_ini.f:553-630) and must be INI_ARRAYS' order.
 * mat_el / mat_el_der may be NULL: the result then stays on the device
 * (mqct_hip_matrix_get copies it out; mqct_hip_use_built_matrix hands it to
 * the propagator).  skipped (optional, total_size bytes) receives the
 * K_SKIPPED_BY_ROUTINE mask. */
int mqct_hip_matrix_build(mqct_hip_t *h, int precision, int64_t total_size,
                          const int *ind_mat, double conv_unit_e,
                          double mij_zero_cut, int ir_cut1, int ir_cut2,
                          double *mat_el, double *mat_el_der,
                          uint8_t *skipped);
int mqct_hip_matrix_get(mqct_hip_t *h, double *mat_el, double *mat_el_der);

/* Per-stage device times of the last build, milliseconds:
 * [0] basis functions on the grid, [1] potential/weights, [2] contraction,
 * [3] cutoff+spline, [4] total.  The first build of a basis and grid in a
 * precision form also times the contraction variants on a sample of tiles
 * (mqct_hip_contract_variant); that time is not in [2] or [4]. */
int mqct_hip_build_timing(mqct_hip_t *h, double *ms5);

/* MIJ_SHIFT= keyword (iotest.f:3754 default YES, parsed at iotest.f:4867).
 * on=1 (the default of a new handle): the propagation sees M(R) - M(R_max).
 * mqct_hip_use_built_matrix then applies the compute branch's shift
 * (matrix_ini.f:2226-2230): the values are shifted after the spline, so
 * Mat_el_der stays that of the unshifted values.  The built matrix itself
 * (mqct_hip_matrix_get, MTRX_UF.DAT) is never shifted, as the reference
 * writes it before the shift. */
int mqct_hip_set_mij_shift(mqct_hip_t *h, int on);

/* READ_MIJ_USER's work after each column is read (matrix_ini.f:6603-6621),
 * on the GPU for the whole matrix: Mat_el(i,k) -= Mat_el(n_r,k) when the
 * handle's MIJ_SHIFT is on, then the clamped spline of the (shifted)
 * values into Mat_el_der with finite-difference end slopes.  mat_el is
 * updated in place; the result is what mqct_hip_load_matrix takes. */
int mqct_hip_mij_spline(mqct_hip_t *h, int n_r, const double *R_com,
                        int64_t total_size, double *mat_el, double *mat_el_der);

/* ---------------------------------------------------------------------- */
/* PES expansion.  The potential as a sum of analytic angular terms with
 * R-dependent coefficients: the term list A_TOP (CALC_EXPANSION,
 * pot_comp.f:3898-4693, or READ_EXPANSION_TERMS' TERMS= list,
 * pot_comp.f:6276-6341), the projection of the V grid on the terms
 * (CALC_EXPANSION_TERMS, pot_comp.f:4695-5920), and the matrix elements
 * from the coefficients (EXPANSION_MATRIX_ELEMENT, pot_comp.f:6465-8751, as
 * MATRIX_MAKE calls it with expansion_defined, matrix_ini.f:1950-1956).
 * Types 1 and 4, and type 0 for distinguishable partners, IDENTICAL_PES
 * and identical partners (mqct_hip_set_identical).                        */

/* a_top: n_terms rows of five ints, A_TOP(1:5, t): type 1 (L, 0, 0, 0, 0),
 * type 4 (L, NJU, 0, 0, 0), type 0 (L1, NJU1, L2, NJU2, L).  Rejected as
 * the reference rejects them: negative NJU (type 4) or NJU1, NJU1 = 0 with
 * NJU2 < 0, a broken triangle (L1, L2, L), a term given twice (also as its
 * (-NJU1, -NJU2) mirror).  Needs the basis; clears earlier coefficients. */
int mqct_hip_set_expansion_terms(mqct_hip_t *h, int n_terms, const int *a_top);

/* CALC_EXPANSION_TERMS on the V grid set by mqct_hip_set_vgrid or
 * mqct_hip_eval_pes: expansion_terms(n_r, n_terms), column-major as the
 * reference keeps it (terms[t * n_r + ir]); the coefficients stay on the
 * device for mqct_hip_expansion_build.  terms may be NULL. */
int mqct_hip_expansion_project(mqct_hip_t *h, double *terms);

/* The matrix from expansion coefficients: Mat_el = conv_unit_e *
 * EXPANSION_MATRIX_ELEMENT, then the cutoff and clamped spline of
 * mqct_hip_matrix_build (same arguments and outputs; the result also stays
 * on the device for mqct_hip_use_built_matrix).  terms != NULL: the caller's
 * coefficients on R_com (n_r points), as READ_EXPANSION_TERMS reads them
 * from PES_EXPAN_TERMS.DAT; terms == NULL: those of the last projection
 * (R_com may then be NULL).  mqct_hip_build_timing reports [0] angular
 * coefficients, [1] 0, [2] the coefficient GEMM, [3] cutoff + spline. */
int mqct_hip_expansion_build(mqct_hip_t *h, int n_r, const double *R_com,
                             const double *terms, int64_t total_size,
                             const int *ind_mat, double conv_unit_e,
                             double mij_zero_cut, int ir_cut1, int ir_cut2,
                             double *mat_el, double *mat_el_der,
                             uint8_t *skipped);

/* Device milliseconds of the expansion path: [0] term functions and
 * weights of the last projection, [1] its GEMM and reduction, [2] angular
 * coefficients of the last expansion build, [3] its coefficient GEMM. */
int mqct_hip_expansion_timing(mqct_hip_t *h, double *ms4);

/* ---------------------------------------------------------------------- */
/* Propagation: replaces the stepping of DOTRAJECT (propagator.f:3205-4137)
 * with DERIVS_BK / DERIVS_BK_adia (mqct_eqtns.f:112-499, 720-879), the
 * spline interpolation potential_data_bk / splint_bk (mqct_eqtns.f:501-611)
 * and rk4 / rkqs / rkck / odeint (suppl_routines.f:1076-1463) for a batch
 * of trajectories.                                                        */

/* Spline-coefficient residency: Mat_el / Mat_el_der as READ_MIJ_USER leaves
 * them (matrix_ini.f:5794-6640) plus the state energies Ej (Hartree) and
 * quantum numbers.  mass = massAB_M (a.u.).  theta_on = bikram_theta. */
int mqct_hip_load_matrix(mqct_hip_t *h, int n_states, const double *Ej,
                         const int *j12, const int *m12, int64_t total_size,
                         const int *ind_mat, int n_r, const double *R_com,
                         const double *mat_el, const double *mat_el_der,
                         double mass, int theta_on);
/* Same, taking the matrix of the last mqct_hip_matrix_build from HBM. */
int mqct_hip_use_built_matrix(mqct_hip_t *h, const double *Ej, double mass,
                              int theta_on);

/* Per-trajectory input (DOTRAJECT initial conditions, propagator.f:2431-
 * 2566, and the $SYSTEM integrator keywords). */
typedef struct {
  int s_st;            /* initial state, 1-based */
  int at_path;         /* AT modes: index of the recorded path */
  double l_real;       /* orbital angular momentum quantum number */
  double E_coll;       /* collision energy, Hartree */
  double dt;           /* TIME_STEP/sqrt(U/300) */
  double eps;          /* EPS_ODEINT */
  double hmin;         /* MIN_TMSTP */
  double R_start;      /* RMAX */
  double time_lim;     /* TIME_LIM (<=0: none) */
  int numb_oscl_prds;
  int numb_orb_prds;
  int max_steps;       /* safety cap on outer steps (<=0: none); the plain
                        * form and the batched CC form end the trajectory
                        * there with fail = 2 (the batched AT form runs
                        * every path to its end) */
} mqct_traj_t;

typedef struct {
  double t_end;
  int n_steps;
  int n_rhs;
  int fail;
  double norm_err;
} mqct_traj_stat_t;

/* Recorded classical paths for the AT modes (the AT_<j12>_<m12>_<L>_<p>
 * files of the first AT run, propagator.f:2994-3032): n_paths paths,
 * path p has records [path_off[p], path_off[p+1]) of 8 doubles each
 * (t, R, P_R, Theta, P_Theta, Phi, P_Phi, V). */
int mqct_hip_load_paths(mqct_hip_t *h, int n_paths, const int64_t *path_off,
                        const double *records);

/* Propagate n_traj trajectories.  q_out: n_traj x (2N+8) final sys_var
 * (AT modes fill the 2N amplitudes and leave the classical slots 0);
 * stat_out (optional): per-trajectory statistics.
 * batched=0: the one-trajectory-per-workgroup plain FP64 form;
 * batched=1: the interval-bucketed batched form with FP64 spline
 * coefficients -- AT modes, and CC modes, whose GEMM also forms dM/dR v
 * for the mean-field force and Coriolis sums (DERIVS_BK, mqct_eqtns.f:
 * 112-499) and whose step kernel runs DOTRAJECT's dt intervals and end
 * checks (propagator.f:3873-4050); batched=2: the same with the coefficients stored
 * in FP32 (amplitudes, products and time stepping stay FP64); batched=3:
 * FP32 coefficients and FP32 products -- the pipelined interval-GEMM forms
 * multiply them by the weighted amplitudes rounded to FP32 on f32 MFMA
 * with FP32 accumulation (the other GEMM forms keep FP64 products;
 * amplitudes and time stepping stay FP64). */
int mqct_hip_propagate(mqct_hip_t *h, int mode, int n_traj,
                       const mqct_traj_t *traj, int batched, double *q_out,
                       mqct_traj_stat_t *stat_out);
int mqct_hip_propagate_timing(mqct_hip_t *h, double *ms);

/* Progress of long batched propagations: when seconds > 0, the host loop
 * of the batched form prints the number of trajectories still running to
 * stderr at most every `seconds`, from the device counter it reads between
 * graph replays anyway (the reference prints per-trajectory progress to its
 * log instead, propagator.f).  0 (the default) prints nothing. */
int mqct_hip_set_progress(mqct_hip_t *h, double seconds);

/* Kernel timing of the batched form (a diagnostic with no counterpart in
 * the reference).  on=1: the following batched propagations record HIP
 * events on their stream around every launch of the interval GEMM (inside
 * the captured graph) and accumulate; on=0 stops.  Each call resets the
 * accumulators.  kernel_stats returns the GEMM launches, their summed
 * duration (ms), the algorithmic FLOPs they did (16*S2 per trajectory
 * right-hand side with R inside the grid, S2 = sum over m-blocks of nb^2)
 * and their algorithmic bytes (the knot matrices of every non-empty bucket,
 * 2*S2 values each, plus 64*N bytes of amplitudes in and products out per
 * right-hand side, 48*N where the products are stored in FP32). */
int mqct_hip_batched_profile(mqct_hip_t *h, int on);
/* The GEMM and step timing of the profiled runs since the last
 * batched_profile call, per group of event-timed iterations: rows of
 * (running trajectories at the start of the group, iterations, summed GEMM
 * ms, shape index, summed step-kernel ms); copies min(cap, n) rows of 5
 * doubles into rows and returns their number in n. */
int mqct_hip_batched_prof_trace(mqct_hip_t *h, int cap, double *rows, int *n);
/* The interval-GEMM candidates: shape index 0..22 (shape = -1: the shape
 * the last batched run started with, for the coefficient precision of that
 * run).  form: 0 LDS-staged per knot, 1 per-wave loads per knot, 2 the same
 * over the fragment-ordered knot tiles, 3 fragment-ordered work queue of
 * (knot, tile group, entry chunk) items over persistent workgroups, 4
 * LDS-staged work queue of (knot, 64-row group, entry chunk) items, 5
 * software-pipelined work queue of (knot, row group, entry chunk) items
 * (a row group: up to 128 rows of one m-block, contracted over the whole
 * block), 6 the same with the items handed out one at a time from a device
 * counter, largest row groups first, 7 form 2 over the list of knots with
 * bucket entries (a grid of two workgroup columns per running trajectory);
 * entries: bucket entries per super-tile (per wave for forms 1-3); k_rows:
 * K rows per LDS slice (0 without LDS); row_tiles: 16-row tiles per wave.
 * The runner chooses a shape per level floor(log2(running trajectories)) by
 * timing the candidates on the first iterations at that level. */
int mqct_hip_batched_tile(mqct_hip_t *h, int shape, int *form, int *entries,
                          int *k_rows, int *row_tiles);
/* The chosen shapes (indices into the candidate list, -1 none yet) of
 * levels 0..n-1 for the FP64 (fp32=0) or FP32 (fp32=1) knot matrices, or
 * the FP32 knot matrices with FP32 products (fp32=2), of the AT modes;
 * fp32 = 3, 4, 5: the same three forms of the CC modes. */
int mqct_hip_batched_tile_levels(mqct_hip_t *h, int fp32, int n, int *shapes);
/* Sets the shapes of levels 0..n-1 (n <= 24; -1: chosen by timing) for the
 * product form fp32 (0..5 as in batched_tile_levels), e.g. to repeat a
 * measured run's choices in a profiled one.  A batched run refuses a set
 * shape that does not fit its matrix (work-queue forms: at most 16383
 * knots). */
int mqct_hip_batched_set_tile_levels(mqct_hip_t *h, int fp32, int n, const int *shapes);
/* Fixes the interval-GEMM tile shape (index into the candidate list,
 * 0..22) for the following batched runs, or -1 to choose by timing again. */
int mqct_hip_batched_set_tile(mqct_hip_t *h, int shape);
/* Threads per trajectory of the batched step kernel: 256, 512 or 1024
 * (AT modes, at most 2048 states; the CC modes run 512), or 0 for the
 * automatic choice (512 when the states exceed 1024). */
int mqct_hip_batched_set_step_threads(mqct_hip_t *h, int threads);
int mqct_hip_batched_kernel_stats(mqct_hip_t *h, int64_t *launches,
                                  double *gemm_ms, double *gemm_flops,
                                  double *gemm_bytes);

/* Contraction kernel variants of the matrix build (no reference
 * counterpart: INTEGRATOR_TEMP is one scalar loop).  Per precision form the
 * library times its candidates -- grid points per LDS stage (kc) and waves
 * per workgroup -- on a sample of tiles at the first build of a basis and
 * grid and keeps the fastest.  set_contract_variant fixes candidate v
 * (0..n-1) of a precision for the following builds, or -1 to time again;
 * contract_variant returns the candidate in use (-1 none yet), its kc and
 * waves, and the number of candidates of that precision. */
int mqct_hip_set_contract_variant(mqct_hip_t *h, int precision, int v);
int mqct_hip_contract_variant(mqct_hip_t *h, int precision, int *v, int *kc,
                              int *waves, int *n_candidates);

/* CC_RK4 propagation that also writes the SAVE_TRAJECT records of the first
 * AT run (propagator.f:3457-3462, 4330-4333): per trajectory up to rec_cap
 * records of 8 doubles (t, R, P_R, Theta, P_Theta, Phi, P_Phi, V) into
 * records[n_traj][rec_cap][8], the count into rec_n[n_traj]. */
int mqct_hip_propagate_record(mqct_hip_t *h, int n_traj,
                              const mqct_traj_t *traj, int rec_cap,
                              double *records, int *rec_n, double *q_out,
                              mqct_traj_stat_t *stat_out);

/* Cross sections from final amplitudes: replaces the per-channel
 * probabilities of DOTRAJECT (propagator.f:4364-4383) and the sums of
 * INELAST_CALC (propagator.f:4752-4983) for a batch, on the GPU.
 *   P_n(t)      = sum over the states of channel n of |a_{t,s}|^2
 *   sigma[g][n] = sum_t weight[g][t] P_n(t)
 * q: n_traj rows of ldq doubles [Re a (n_states) | Im a (n_states) | ...]
 * (what mqct_hip_propagate returns, ldq = 2*n_states+8); indx_chann: the
 * channel of every state (1-based, INI_ARRAYS order); weight[g][t]: the
 * INELAST_CALC factors (2J+1) pi/(k^2 (2j+1)) of trajectory t summed over
 * its (J, l) pairs, one row per group (a collision energy for sigma(U), or
 * a J for the J-resolved table).  sigma: [n_groups][n_channels] in bohr^2
 * (the caller applies a0^2 and the billing factor, propagator.f:1198-1276);
 * opacity (optional): P[n_traj][n_channels].  Sums run in a fixed order,
 * so results repeat bit for bit. */
int mqct_hip_cross_sections(mqct_hip_t *h, int n_traj, int n_states, int ldq,
                            const double *q, int n_channels,
                            const int *indx_chann, int n_groups,
                            const double *weight, double *sigma,
                            double *opacity);

/* The same cross sections without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * All array arguments are device pointers the caller owns (q_dev,
 * indx_chann_dev, weight_dev in, sigma_dev and the optional opacity_dev
 * out); the working memory is a fixed 32 MiB device array inside the
 * library, one per GPU, held until the process ends.  A call needs
 * 8 * (min(64, n_traj) * n_groups * n_channels + n_traj * n_channels
 * [omitted when opacity_dev is given] + n_channels/2 + 3) bytes of it and is
 * refused with MQCT_ERR_ARG beyond 32 MiB.  Runs on the handle's stream and
 * returns when the results are in place.  Every call on one GPU uses the
 * same working memory: calls that overlap on one GPU (from several threads
 * or handles) race. */
int mqct_hip_cross_sections_noalloc(mqct_hip_t *h, int n_traj, int n_states,
                                    int ldq, const double *q_dev,
                                    int n_channels, const int *indx_chann_dev,
                                    int n_groups, const double *weight_dev,
                                    double *sigma_dev, double *opacity_dev);

/* ---------------------------------------------------------------------- */
/* MTRX_UF.DAT (Fortran sequential unformatted, 4-byte record markers):
 * PRINT_MIJ_USER (matrix_ini.f:6640-6780) writer and READ_MIJ_USER
 * (matrix_ini.f:5794-6640) reader.  chan_qn: per channel, the quantum
 * numbers the state records carry after (st, chann, j12, m12): type 1:
 * {j}; type 4: {j, ka, kc}; type 0: {j1, ka1, kc1, j2, ka2, kc2}. */
int mqct_hip_write_mtrx_uf(const char *path, int coll_type, int n_channels,
                           int n_states, int64_t total_size, int n_r,
                           const int *indx_chann, const int *j12,
                           const int *m12, const int *chan_qn,
                           const int *ind_mat, const double *R_com,
                           const double *mat_el);
/* Reads the header first (pass NULL arrays to get the sizes), then the
 * data.  hdr = {coll_type, n_channels, n_states, total_size, n_r}. */
int mqct_hip_read_mtrx_uf(const char *path, int64_t *hdr, int *indx_chann,
                          int *j12, int *m12, int *chan_qn, int *ind_mat,
                          double *R_com, double *mat_el);

/* The parallel-IO matrix directory of a split build (MATRIX_TRUNCATED,
 * UNFORMAT=NO or YES): MATRIX_INFO.DAT with the state tables
 * (bk_print_matrix_info, matrix_ini.f:6836-7144; types 1, 4, 0, with the
 * exchange parity column for identical partners), and the matrix as
 * n_files pairs MIJ_<k_st>_<k_fn>.DAT / Ind_<k_st>_<k_fn>.DAT over the
 * reference's rank split of 1..total_size, holding the elements whose |M|
 * at the two cutoff R points exceeds mij_zero_cut (Hartree), with
 * MATRIX_NONZERO_INDEX.DAT (the cut in cm^-1) and MATRIX_FILE_INDEX.DAT
 * (bk_print_matrix, matrix_ini.f:7147-7244; 1545-1575; 9053-9145).
 * mqct_hip_read_mij_dir is bk_read_matrix (matrix_ini.f:7246-7525) for one
 * reader: the non-zero elements in file order with their ind_mat pairs
 * (mat_el[n_nonzero][n_r], ind_mat[n_nonzero][2]); with mat_el and ind_mat
 * NULL only n_nonzero and the cut are returned.  MIJ_SHIFT and the spline
 * that follow the read are mqct_hip_mij_spline's. */
int mqct_hip_write_mij_info(const char *dir, int unformatted, int coll_type,
                            int n_channels, int n_states, int64_t total_size,
                            int n_r, const int *indx_chann, const int *j12,
                            const int *m12, const int *chan_qn,
                            const int *parity_state, const int *ind_mat,
                            const double *R_com);
int mqct_hip_write_mij_dir(const char *dir, int unformatted, int n_files,
                           int64_t total_size, int n_r, const int *ind_mat,
                           const double *mat_el, double mij_zero_cut,
                           int ir_cut1, int ir_cut2);
int mqct_hip_read_mij_dir(const char *dir, int unformatted, int n_r,
                          int64_t cap, int64_t *n_nonzero, double *cutoff_cm,
                          int *ind_mat, double *mat_el);

/* The classical paths of the first AT run, in the reference's files:
 * dir/AT_<j12>_<m12>_<l>_<parity> (one Fortran unformatted record of 8
 * REAL*8 -- t, R, P_R, Theta, P_Theta, Phi, P_Phi, V -- per step) and
 * dir/n_AT_<...> (one INTEGER*4 record: the number of records).  Replaces
 * SAVE_TRAJECT's writes (propagator.f:2838-2852 open, 3457-3462 and
 * 4330-4333 write) and the AT run's read-back (propagator.f:2994-3005).
 * records: [n_rec][8] as mqct_hip_propagate_record returns them.
 * read_at_path with records == NULL returns the count in n_rec only. */
int mqct_hip_write_at_path(const char *dir, int j12, int m12, int l,
                           int parity, int n_rec, const double *records);
int mqct_hip_read_at_path(const char *dir, int j12, int m12, int l,
                          int parity, int cap, double *records, int *n_rec);

/* ---------------------------------------------------------------------- */
/* Learned potential: an MLP (12 features -> 4 x 256 SiLU -> 1) trained and
 * evaluated by the library's own HIP kernels (BF16 weights, FP32
 * accumulation, Adam); it stands in for USER_DEFINED_PES on the SYS_TYPE 0
 * grid.  No reference counterpart: it plugs in where the reference calls
 * the user PES (pes_sys_type.f:88-97).  Geometries are (R, beta1, gamma1,
 * alpha2, beta2, gamma2); the network predicts y with V = scale*sinh(y) +
 * shift (cm^-1); below r_min the plain potential is kept. */
int mqct_hip_mlp_init(mqct_hip_t *h, uint64_t seed);
int mqct_hip_mlp_n_params(void);
int mqct_hip_mlp_get_params(mqct_hip_t *h, float *params);
int mqct_hip_mlp_set_params(mqct_hip_t *h, const float *params);
int mqct_hip_mlp_set_scaling(mqct_hip_t *h, double scale, double shift,
                             double r_min);
/* Training set kept on the device: geo[n][6], targets y[n]. */
int mqct_hip_mlp_set_data(mqct_hip_t *h, int n, const double *geo,
                          const float *y);
/* One Adam step on samples [start, start+n) of the training set; loss =
 * batch mean squared error before the step. */
int mqct_hip_mlp_train_step(mqct_hip_t *h, int start, int n, float lr,
                            float *loss);
/* Mean-squared-error gradient on samples [start, start+n), parameter order
 * W1,b1,W2,b2,W3,b3,W4,b4,W5,b5 with W as [out][in]. */
int mqct_hip_mlp_grad(mqct_hip_t *h, int start, int n, float *grad,
                      float *loss);
int mqct_hip_mlp_predict(mqct_hip_t *h, int n, const double *geo, float *y);
/* Fill the V grid of the current SYS_TYPE 0 grid from the network (then
 * build with MQCT_PREC_BF16X3 or any other precision). */
int mqct_hip_surrogate_vgrid(mqct_hip_t *h, int n_r, const double *R_com);
/* The synthetic potential at n arbitrary geometries (training targets). */
int mqct_hip_pes_points(mqct_hip_t *h, int pes_kind, int n_par,
                        const double *par, int n, const double *geo,
                        double *V);

/* ---------------------------------------------------------------------- */
/* Diagnostics: multiplies 16x16 tiles with one MFMA instruction to check the
 * lane layouts the kernels assume.  which: 0 f64 16x16x4, 1 f32 16x16x4,
 * 2 bf16 16x16x32 (inputs given as float, rounded to bf16).  A is 16xK,
 * B is Kx16 row-major, C out 16x16 row-major (as double). */
int mqct_hip_probe_mfma(int which, const double *A, const double *B,
                        double *C);

#ifdef __cplusplus
}
#endif
#endif
