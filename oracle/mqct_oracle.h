/*
 * mqct_oracle.h -- CPU FP64 restatement of the MQCT hot paths.
 *
 * TEST INFRASTRUCTURE ONLY.  The parity tests, the smoke check and the
 * CPU-baseline leg of the benchmark compare the HIP library against this
 * library.  Product code never links, loads or calls it.
 *
 * Every function names the reference lines it restates (paths relative to
 * the MQCT source tree, MQCT_SOURCE_CODES/).  Indices that the reference
 * keeps 1-based (state numbers, channel numbers, ind_mat) stay 1-based here.
 */
#ifndef MQCT_ORACLE_H
#define MQCT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- special functions ------------------------------------------------ */
void   or_logfact_init(int fact_max);
double or_logfact(int n);
void   or_jacobi_poly(int n, double alpha, double beta, double x, double *cx);
void   or_djkm(int j, int k, int m, double alpha, double beta, double gamma,
               double *re, double *im);
double or_plgndr(int l, int m, double x);
double or_w3js(int j1, int j2, int j3, int m1, int m2, int m3);
double or_cg(int j1, int j2, int j3, int m1, int m2, int m3);
void   or_gauleg(int n, double *x, double *w);

/* ---- wavefunctions ----------------------------------------------------- */
void or_wf_diat_top(int j, int m, double theta, double phi,
                    double *re, double *im);
void or_wf_asym_top(int J, const double *coeff, int M, double alpha,
                    double beta, double gamma, double *re, double *im);

/* ---- asymmetric top ----------------------------------------------------- */
double or_hamilt_kyro(int J1, int J2, int K1, int K2, double A, double B,
                      double C);
int    or_state(int J, int DK, int *KA, int *KC);
int    or_asym_top_eigen(double A, double B, double C, int J, double *w,
                         double *z);
int    or_eigen_vec(double A, double B, double C, int J, int KA, int KC,
                    double *E, double *zout);

/* ---- channel lists and state enumeration -------------------------------- */
int or_channels_type1(int jmin, int jmax, double Be, double De, int *j_ch,
                      double *E_ch);
int or_channels_type4(int jmin, int jmax, double A, double B, double C,
                      int *j_ch, int *ka_ch, int *kc_ch, double *E_ch);
int or_channels_type0(int jmin1, int jmax1, int jmin2, int jmax2,
                      double A1, double B1, double C1, double A2, double B2,
                      double C2, int *j1, int *ka1, int *kc1, int *j2,
                      int *ka2, int *kc2, double *E_ch);
int64_t or_ini_arrays(int coll_type, int n_ch, const int *j1_ch,
                      const int *j2_ch, int *j12, int *m12, int *indx_chann,
                      int *chann_indx, int64_t cap_pairs, int *ind_mat,
                      int *n_states_out);

/* ---- matrix build ------------------------------------------------------ */
typedef struct {
  int coll_type;              /* 1, 4 or 0 */
  int n_states;
  const int *j12, *m12, *indx_chann;  /* per state; indx_chann 1-based */
  const int *j1_ch, *j2_ch;   /* per channel (j2_ch only for type 0) */
  int ldvec;                  /* row length of eig1/eig2 */
  const double *eig1, *eig2;  /* [n_ch][ldvec]: c_K, K=-J..J */
  int ref_xz;                 /* 1 = XZ frame (default), 0 = XY/YZ quirk */
  /* angular grids: nodes as produced by gauleg / equidistant rule */
  int n_beta;  const double *xb, *wb;  double xbm, xbl;
  int n_gamma; const double *xg, *wg;  double xgm, xgl;
  int n_alpha2; const double *xaa, *waa; double xaam, xaal;
  int n_beta2;  const double *xbb, *wbb; double xbbm, xbbl;
  int n_gamma2; const double *xgg, *wgg; double xggm, xggl;
  int eq_gamma;               /* gamma/alpha grids equidistant (angle = x) */
  int n_r;
  const double *V;            /* type 1: [n_r][n_beta]
                                 type 4: [n_r][n_beta][n_gamma]
                                 type 0: [n_r][a2][b1][b2][g1][g2] */
  int abs_mode;               /* 1: sum |integrand| (condition scale) */
  /* type 0 identical partners (IDENTICAL=YES) and IDENTICAL_PES=YES */
  int identical, identical_pes;
  const int *parity_state;    /* per state, +-1 (matrix_ini.f:504-528) */
  const int *parity_state_bk; /* per state, +-1 */
  const int *ka1_ch, *kc1_ch, *ka2_ch, *kc2_ch;  /* per channel */
  /* planted mistakes for the tests that show a check fails (0 = none):
   * bit 0 runs the identical-partner m sum over -j1'..j1' only */
  int plant;
} or_matrix_ctx;

double or_integrator_temp(const or_matrix_ctx *c, const int *ind_mat,
                          int64_t k, int i_r);
void or_matrix_make(const or_matrix_ctx *c, int64_t total_size,
                    const int *ind_mat, const double *R, double conv_unit_e,
                    double mij_zero_cut_hartree, int ir_cut1, int ir_cut2,
                    double *mat_el, double *mat_el_der, uint8_t *skipped);
void or_mij_shift(int64_t total_size, int n, double *mat_el);
void or_read_mij_post(const double *R, int n, int64_t total_size, int shift,
                      double *mat_el, double *mat_el_der);
void or_spline(const double *x, const double *y, int n, double yp1,
               double ypn, double *y2);
void or_splint(const double *xa, const double *ya, const double *y2a, int n,
               double x, double *y, double *yprime);
int  or_nearest_r(const double *R, int n, double r0);

/* ---- PES expansion (mqct_oracle_exp.c) --------------------------------- */
/* Term rows are A_TOP(1:5, t); expansion coefficients terms[t * n_r + ir]. */
void or_expansion_term(int coll_type, const int *t, double b1, double g1,
                       double a2, double b2, double g2, double *re,
                       double *im);
void or_expansion_term_grid(const or_matrix_ctx *c, const int *t, double *out);
void or_calc_expansion_terms(const or_matrix_ctx *c, int n_terms,
                             const int *a_top, double *out);
void or_expansion_matrix_element(const or_matrix_ctx *c, int n_terms,
                                 const int *a_top, const double *terms,
                                 const int *ind_mat, int64_t k, double *out);
void or_expansion_matrix_make(const or_matrix_ctx *c, int n_terms,
                              const int *a_top, const double *terms,
                              int64_t total_size, const int *ind_mat,
                              const double *R, double conv_unit_e,
                              double mij_zero_cut_hartree, int ir_cut1,
                              int ir_cut2, double *mat_el, double *mat_el_der,
                              uint8_t *skipped);

/* ---- synthetic potentials --------------------------------------------- */
double or_pes_legendre(const double *par, double R, double theta);
double or_pes_realsh(const double *par, double R, double beta, double gamma);
double or_pes_sites(const double *par, double R, double b1, double g1,
                    double a2, double b2, double g2);
void or_vgrid(int coll_type, int pes_kind, const double *par,
              const or_matrix_ctx *c, const double *R, double *V);

/* ---- dynamics ---------------------------------------------------------- */
typedef struct {
  int n_states;
  int64_t total_size;
  const int *ind_mat;          /* [T][2], 1-based */
  const int *j12, *m12;
  const double *Ej;            /* Hartree */
  int n_r;
  const double *R;
  const double *mat;           /* [T][n_r]  (Mat_el column-major) */
  const double *mat_der;       /* [T][n_r]  y'' of the spline */
  double mass;                 /* massAB_M, a.u. */
  int theta_on;                /* bikram_theta */
  /* derived tables (filled by or_resize) */
  double *misc;                /* [T][4] */
  int *bk_indx;                /* [T] 0-based index into delta_E */
  double *delta_E;             /* [T] (first n_dE used) */
  int n_dE;
  /* adiabatic-trajectory path (AT) */
  int at_n;
  const double *at_t;          /* [at_n] */
  const double *at_var;        /* [7][at_n]: R,P_R,Theta,P_Theta,Phi,P_Phi,V */
  double *at_var_der;          /* [7][at_n]  y'' of the path splines */
  /* planted mistakes for the tests that show a check fails (0 = none):
   * bit 0 flips the mean-field force on P_R, bit 1 the first half of the
   * P_Phi Coriolis accumulator, bit 2 the quantum Coriolis coupling */
  int plant;
} or_dyn_ctx;

void or_resize(or_dyn_ctx *d);
void or_at_prepare(or_dyn_ctx *d);
void or_potential_data_bk(const or_dyn_ctx *d, double R, double *m,
                          double *dm);
void or_derivs_bk(const or_dyn_ctx *d, double t, const double *q,
                  double *dqdt);
void or_derivs_bk_adia(const or_dyn_ctx *d, double t, const double *q,
                       double *dqdt);
typedef void (*or_derivs_fn)(const or_dyn_ctx *, double, const double *,
                             double *);
void or_rk4(const or_dyn_ctx *d, const double *y, const double *dydx, int n,
            double x, double h, double *yout, or_derivs_fn f);
void or_rkck(const or_dyn_ctx *d, const double *y, const double *dydx, int n,
             double x, double h, double *yout, double *yerr, or_derivs_fn f);
void or_rkqs(const or_dyn_ctx *d, double *y, const double *dydx, int n,
             double *x, double htry, double eps, const double *yscal,
             double *hdid, double *hnext, double hmin, or_derivs_fn f);

typedef struct {
  int mode;           /* 0 CC-RK4, 1 CC-ODEINT(bk_int), 2 AT-RK4, 3 AT-ODEINT */
  int s_st;           /* initial state, 1-based */
  double l_real;      /* orbital quantum number */
  double E_coll;      /* collision energy, Hartree */
  double dt;          /* RK4 step (a.u.) */
  double eps;         /* odeint tolerance */
  double h1;          /* first odeint step */
  double hmin;        /* minimal odeint step */
  double R_start;     /* R_max_dist */
  double time_lim;    /* bk_int upper time */
  int numb_oscl_prds;
  int numb_orb_prds;
  int max_steps;      /* RK4 safety cap */
} or_traj_param;

typedef struct {
  double t_end;
  int n_steps;        /* accepted steps (RK4) or odeint steps */
  int n_rhs;          /* right-hand-side evaluations */
  int fail;
  double norm_err;    /* max |1 - sum|a|^2| over the run */
} or_traj_stat;

int or_dotraject(const or_dyn_ctx *d, const or_traj_param *p, double *q,
                 or_traj_stat *st, double *rec, int rec_cap, int *rec_n);

#ifdef __cplusplus
}
#endif
#endif
