/*
 * mqct_oracle_dyn.c -- CPU FP64 restatement of MQCT's propagation hot path:
 * resize tables, spline interpolation of M(R), the right-hand sides
 * DERIVS_BK / DERIVS_BK_adia, the Numerical-Recipes integrators and the
 * stepping loop of DOTRAJECT.
 *
 * TEST INFRASTRUCTURE ONLY (the reference results of the tests); the product
 * never calls it.
 *
 * Pins: or_derivs_bk on a two-state constant coupling reproduces the Rabi
 * formula; diagonal-only M gives a pure phase; the norm sum |a|^2 is
 * conserved by the quantum part; rk4/rkck reproduce exp(-i w t) for linear
 * problems to their order; spline interpolation is exact on cubics.
 * The single-state pair loop and the Coriolis form are restated line by
 * line, with the zero pad slots q(2N+7), q(2N+8) of the reference.
 */
#include "mqct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846264338327950288;

/* resize: mqct_eqtns.f:1-110.  Coriolis factors misc_bk(4,k), unique
 * Delta E table with the 1e-12 (REAL*4 literal) dedupe tolerance, the last
 * matching entry winning as in the reference's loop. */
void or_resize(or_dyn_ctx *d) {
  int64_t T = d->total_size;
  const float tolf = 1.E-12f;
  const double tol = (double)tolf;
  for (int64_t k = 0; k < T; k++) {
    int i = d->ind_mat[2 * k] - 1, j = d->ind_mat[2 * k + 1] - 1;
    int j_12 = d->j12[j], m_12 = d->m12[i];
    int m1 = m_12 + 1, m0 = m_12 - 1;
    d->misc[4 * k + 0] = sqrt(j_12 * (j_12 + 1.0) - m1 * (m1 - 1.0));
    d->misc[4 * k + 1] = sqrt(j_12 * (j_12 + 1.0) - m0 * (m0 + 1.0));
    j_12 = d->j12[i];
    m_12 = d->m12[j];
    m1 = m_12 + 1;
    m0 = m_12 - 1;
    d->misc[4 * k + 2] = sqrt(j_12 * (j_12 + 1.0) - m1 * (m1 - 1.0));
    d->misc[4 * k + 3] = sqrt(j_12 * (j_12 + 1.0) - m0 * (m0 + 1.0));
  }
  int cnt = 0;
  d->delta_E[0] = d->Ej[d->ind_mat[1] - 1] - d->Ej[d->ind_mat[0] - 1];
  cnt = 1;
  d->bk_indx[0] = 0;
  for (int64_t k = 1; k < T; k++) {
    int i = d->ind_mat[2 * k] - 1, j = d->ind_mat[2 * k + 1] - 1;
    double de = d->Ej[j] - d->Ej[i];
    int same = 0;
    for (int ii = 0; ii < cnt; ii++)
      if (fabs(de - d->delta_E[ii]) < tol) { same = 1; d->bk_indx[k] = ii; }
    if (!same) {
      d->delta_E[cnt] = de;
      d->bk_indx[k] = cnt;
      cnt++;
    }
  }
  d->n_dE = cnt;
}

/* Path splines of the recorded classical trajectory: propagator.f:3012-3031
 * (clamped with finite-difference end slopes, columns 1..7). */
void or_at_prepare(or_dyn_ctx *d) {
  int n = d->at_n;
  for (int c = 0; c < 7; c++) {
    const double *y = d->at_var + (size_t)c * n;
    double yp1 = (y[1] - y[0]) / (d->at_t[1] - d->at_t[0]);
    double ypn = (y[n - 1] - y[n - 2]) / (d->at_t[n - 1] - d->at_t[n - 2]);
    or_spline(d->at_t, y, n, yp1, ypn, d->at_var_der + (size_t)c * n);
  }
}

static void path_at(const or_dyn_ctx *d, int col, double t, double *v) {
  double yp;
  or_splint(d->at_t, d->at_var + (size_t)col * d->at_n,
            d->at_var_der + (size_t)col * d->at_n, d->at_n, t, v, &yp);
}

/* potential_data_bk + splint_bk: mqct_eqtns.f:501-611.  Below R_1 the
 * linear block is computed but then overwritten by the cubic extrapolation
 * of splint_bk (the RETURN is commented out); above R_n everything is 0. */
void or_potential_data_bk(const or_dyn_ctx *d, double x, double *m,
                          double *dm) {
  int n = d->n_r;
  int64_t T = d->total_size;
  if (x > d->R[n - 1]) {
    for (int64_t k = 0; k < T; k++) { m[k] = 0.0; if (dm) dm[k] = 0.0; }
    return;
  }
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    int k = (khi + klo) / 2;
    if (d->R[k] > x) khi = k; else klo = k;
  }
  double h = d->R[khi] - d->R[klo];
  double aa = (d->R[khi] - x) / h, bb = (x - d->R[klo]) / h;
  double cc1 = aa * aa * aa - aa, cc2 = bb * bb * bb - bb;
  double cc3 = 3 * aa * aa - 1.0, cc4 = 3 * bb * bb - 1.0;
  double cc5 = (h * h) / 6.0, cc6 = cc5 / h;
  for (int64_t k = 0; k < T; k++) {
    const double *y = d->mat + (size_t)k * n, *y2 = d->mat_der + (size_t)k * n;
    m[k] = aa * y[klo] + bb * y[khi] + (cc1 * y2[klo] + cc2 * y2[khi]) * cc5;
    if (dm)
      dm[k] = (y[khi] - y[klo]) / h + (-cc3 * y2[klo] + cc4 * y2[khi]) * cc6;
  }
}

/* Quantum pair loop shared by both right-hand sides: mqct_eqtns.f:190-259
 * (and 776-828).  Returns the per-pair phase factors through rp/ip. */
static void pair_loop(const or_dyn_ctx *d, double t, const double *q,
                      double *dqdt, const double *bk_mat, double *sinE,
                      double *cosE) {
  int N = d->n_states;
  for (int k = 0; k < d->n_dE; k++) {
    sinE[k] = sin(t * d->delta_E[k]);
    cosE[k] = cos(t * d->delta_E[k]);
  }
  for (int64_t k = 0; k < d->total_size; k++) {
    int i = d->ind_mat[2 * k] - 1, j = d->ind_mat[2 * k + 1] - 1;
    double M = bk_mat[k];
    double d_f = (i == j) ? 1.0 : 2.0;
    double iphase = sinE[d->bk_indx[k]], rphase = cosE[d->bk_indx[k]];
    double rf = -d_f * iphase, jf = d_f * rphase;
    dqdt[i] += (q[j] * rf + q[j + N] * jf) * M / 2.0;
    dqdt[i + N] += (q[j + N] * rf - q[j] * jf) * M / 2.0;
    dqdt[j] -= (q[i] * rf - q[i + N] * jf) * M / 2.0;
    dqdt[j + N] -= (q[i + N] * rf + q[i] * jf) * M / 2.0;
  }
}

/* State-diagonal Coriolis coupling m <-> m+-1 without the Theta terms:
 * mqct_eqtns.f:434-466 (and 845-877). */
static void coriolis_diag(const or_dyn_ctx *d, const double *q, double *dqdt,
                          double dPhidT) {
  int N = d->n_states;
  if (d->plant & 4) dPhidT = -dPhidT;
  for (int i = 0; i < N; i++) {
    double j_12 = d->j12[i], m_12 = d->m12[i];
    double m1 = m_12 + 1.0, m0 = m_12 - 1.0;
    double sqrt1 = sqrt(j_12 * (j_12 + 1.0) - m0 * (m0 + 1.0));
    double sqrt2 = sqrt(j_12 * (j_12 + 1.0) - m1 * (m1 - 1.0));
    if (i > 0) {
      dqdt[i] += dPhidT / 2.0 * sqrt1 * q[i - 1 + N];
      dqdt[i + N] -= dPhidT / 2.0 * sqrt1 * q[i - 1];
    }
    if (i < N - 1) {
      dqdt[i] += dPhidT / 2.0 * sqrt2 * q[i + 1 + N];
      dqdt[i + N] -= dPhidT / 2.0 * sqrt2 * q[i + 1];
    }
  }
}

/* DERIVS_BK: mqct_eqtns.f:112-499.  q = [Re a(N), Im a(N), R, P_R, Theta,
 * P_Theta, Phi, P_Phi, 0, 0]. */
void or_derivs_bk(const or_dyn_ctx *d, double t, const double *q,
                  double *dqdt) {
  int N = d->n_states;
  int64_t T = d->total_size;
  double mu = d->mass;
  memset(dqdt, 0, sizeof(double) * (2 * N + 8));
  double Rrr = q[2 * N], tet = q[2 * N + 2];
  double sin_tet = sin(tet), cos_tet = cos(tet);
  double dThetadT = 0.0;
  if (d->theta_on) dThetadT = q[2 * N + 3] / q[2 * N] / q[2 * N] / mu;
  double dPhidT = q[2 * N + 5] / q[2 * N] / q[2 * N] / mu / (sin_tet * sin_tet);
  double *bk_mat = (double *)malloc(sizeof(double) * T);
  double *bk_der = (double *)malloc(sizeof(double) * T);
  double *amp_phi = (double *)calloc((size_t)T, sizeof(double));
  double *amp_theta = (double *)calloc((size_t)T, sizeof(double));
  double *sinE = (double *)malloc(sizeof(double) * d->n_dE);
  double *cosE = (double *)malloc(sizeof(double) * d->n_dE);
  or_potential_data_bk(d, Rrr, bk_mat, bk_der);
  pair_loop(d, t, q, dqdt, bk_mat, sinE, cosE);
  const int PAD1 = 2 * N + 6, PAD2 = 2 * N + 7; /* q(2N+7), q(2N+8) */
  for (int64_t k = 0; k < T; k++) {
    int i = d->ind_mat[2 * k] - 1, j = d->ind_mat[2 * k + 1] - 1;
    double d_f = (i == j) ? 1.0 : 2.0;
    double iphase = sinE[d->bk_indx[k]], rphase = cosE[d->bk_indx[k]];
    double rf = -d_f * iphase, jf = d_f * rphase;
    const double *misc = d->misc + 4 * k;
    /* mean-field force on P_R: mqct_eqtns.f:263-266 */
    dqdt[2 * N + 1] -= ((q[i] * q[j] + q[i + N] * q[j + N]) * rphase -
                        (q[i + N] * q[j] - q[i] * q[j + N]) * iphase) *
                       d_f * bk_der[k] * ((d->plant & 1) ? -1.0 : 1.0);
    /* Coriolis accumulators: mqct_eqtns.f:270-359 */
    double j_12 = d->j12[j], m_12 = d->m12[i];
    double m1 = m_12 + 1.0, m0 = m_12 - 1.0;
    int s1, s11, s0, s00;
    if ((int)fabs(m1) <= (int)j_12) { s1 = j + 1; s11 = s1 + N; }
    else { s1 = PAD1; s11 = PAD2; }
    if ((int)fabs(m0) <= (int)j_12) { s0 = j - 1; s00 = s0 + N; }
    else { s0 = PAD1; s00 = PAD2; }
    amp_phi[k] += (-sin_tet * rf / 2.0 *
                       (misc[0] * (q[i] * q[s1] + q[i + N] * q[s11]) +
                        misc[1] * (q[i] * q[s0] + q[i + N] * q[s00])) -
                   sin_tet * (-jf) / 2.0 *
                       (misc[0] * (q[i + N] * q[s1] - q[i] * q[s11]) +
                        misc[1] * (q[i + N] * q[s0] - q[i] * q[s00]))) *
                  ((d->plant & 2) ? -1.0 : 1.0);
    if (d->theta_on)
      amp_theta[k] += d_f * rphase / 2.0 *
                          (misc[0] * (q[i] * q[s1] + q[i + N] * q[s11]) -
                           misc[1] * (q[i] * q[s0] + q[i + N] * q[s00])) +
                      d_f * (-iphase) / 2.0 *
                          (misc[0] * (q[i + N] * q[s1] - q[i] * q[s11]) -
                           misc[1] * (q[i + N] * q[s0] - q[i] * q[s00]));
    j_12 = d->j12[i];
    m_12 = d->m12[j];
    m1 = m_12 + 1;
    m0 = m_12 - 1;
    if ((int)fabs(m1) <= (int)j_12) { s1 = i + 1; s11 = s1 + N; }
    else { s1 = PAD1; s11 = PAD2; }
    if ((int)fabs(m0) <= (int)j_12) { s0 = i - 1; s00 = s0 + N; }
    else { s0 = PAD1; s00 = PAD2; }
    amp_phi[k] += -sin_tet * (-rf) / 2.0 *
                  (misc[2] * (q[s1] * q[j] + q[s11] * q[j + N]) +
                   misc[3] * (q[s0] * q[j] + q[s00] * q[j + N]));
    amp_phi[k] += -sin_tet * (-jf) / 2.0 *
                  (misc[2] * (q[s1] * q[j + N] - q[s11] * q[j]) +
                   misc[3] * (q[s0] * q[j + N] - q[s00] * q[j]));
    if (d->theta_on) {
      amp_theta[k] += d_f * rphase / 2.0 *
                      (misc[2] * (q[s1] * q[j] + q[s11] * q[j + N]) -
                       misc[3] * (q[s0] * q[j] + q[s00] * q[j + N]));
      amp_theta[k] += d_f * iphase / 2.0 *
                      (misc[2] * (q[s1] * q[j + N] - q[s11] * q[j]) -
                       misc[3] * (q[s0] * q[j + N] - q[s00] * q[j]));
    }
  }
  double sphi = 0.0, sthe = 0.0;
  for (int64_t k = 0; k < T; k++) {
    sphi += bk_mat[k] * amp_phi[k];
    sthe += bk_mat[k] * amp_theta[k];
  }
  dqdt[2 * N + 5] = sphi;
  if (d->theta_on) dqdt[2 * N + 3] = sthe;
  if (d->theta_on) {
    /* mqct_eqtns.f:389-432 */
    for (int i = 0; i < N; i++) {
      double j_12 = d->j12[i], m_12 = d->m12[i];
      double m1 = m_12 + 1.0, m0 = m_12 - 1.0;
      double sqrt1 = sqrt(j_12 * (j_12 + 1.0) - m0 * (m0 + 1.0));
      double sqrt2 = sqrt(j_12 * (j_12 + 1.0) - m1 * (m1 - 1.0));
      dqdt[i] += dPhidT * m_12 * cos_tet * q[i];
      dqdt[i + N] -= dPhidT * m_12 * cos_tet * q[i + N];
      if (i > 0) {
        dqdt[i] += (dPhidT * sin_tet * q[i - 1 + N] - dThetadT * q[i - 1]) * sqrt1 / 2.0;
        dqdt[i + N] -= (dPhidT * sin_tet * q[i - 1] - dThetadT * q[i - 1 + N]) * sqrt1 / 2.0;
      }
      if (i < N - 1) {
        dqdt[i] += (dPhidT * sin_tet * q[i + 1 + N] - dThetadT * q[i + 1]) * sqrt2 / 2.0;
        dqdt[i + N] -= (dPhidT * sin_tet * q[i + 1] - dThetadT * q[i + 1 + N]) * sqrt2 / 2.0;
      }
    }
  } else {
    coriolis_diag(d, q, dqdt, dPhidT);
  }
  /* classical equations: mqct_eqtns.f:473-495 */
  dqdt[2 * N] = q[2 * N + 1] / mu;
  dqdt[2 * N + 4] = q[2 * N + 5] / q[2 * N] / q[2 * N] / mu / (sin_tet * sin_tet);
  dqdt[2 * N + 1] += q[2 * N + 3] / q[2 * N] / q[2 * N] / mu * q[2 * N + 3] / q[2 * N] +
                     q[2 * N + 5] / q[2 * N] / q[2 * N] / mu * q[2 * N + 5] / q[2 * N] /
                         (sin_tet * sin_tet);
  if (d->theta_on) {
    dqdt[2 * N + 2] = q[2 * N + 3] / q[2 * N] / q[2 * N] / mu;
    dqdt[2 * N + 3] += q[2 * N + 5] / q[2 * N] / q[2 * N] / mu * q[2 * N + 5] /
                       (sin_tet * sin_tet * sin_tet) * cos_tet;
  }
  free(bk_mat); free(bk_der); free(amp_phi); free(amp_theta);
  free(sinE); free(cosE);
}

/* DERIVS_BK_adia: mqct_eqtns.f:720-879.  q holds the 2N amplitudes only;
 * R, Theta, P_Theta, P_Phi come from the recorded path. */
void or_derivs_bk_adia(const or_dyn_ctx *d, double t, const double *q,
                       double *dqdt) {
  int N = d->n_states;
  int64_t T = d->total_size;
  double Rrr, tet, ptet, pphi;
  path_at(d, 0, t, &Rrr);
  path_at(d, 2, t, &tet);
  path_at(d, 3, t, &ptet);
  path_at(d, 5, t, &pphi);
  (void)ptet;
  double sin_tet = sin(tet);
  double dPhidT = pphi / Rrr / Rrr / d->mass / (sin_tet * sin_tet);
  memset(dqdt, 0, sizeof(double) * 2 * N);
  double *bk_mat = (double *)malloc(sizeof(double) * T);
  double *sinE = (double *)malloc(sizeof(double) * d->n_dE);
  double *cosE = (double *)malloc(sizeof(double) * d->n_dE);
  or_potential_data_bk(d, Rrr, bk_mat, NULL);
  pair_loop(d, t, q, dqdt, bk_mat, sinE, cosE);
  coriolis_diag(d, q, dqdt, dPhidT);
  free(bk_mat); free(sinE); free(cosE);
}

/* rk4: suppl_routines.f:1076-1113. */
void or_rk4(const or_dyn_ctx *d, const double *y, const double *dydx, int n,
            double x, double h, double *yout, or_derivs_fn f) {
  double *dym = (double *)malloc(sizeof(double) * (n + 8));
  double *dyt = (double *)malloc(sizeof(double) * (n + 8));
  double *yt = (double *)calloc((size_t)n + 8, sizeof(double));
  double hh = h * 0.5, h6 = h / 6.0, xh = x + hh;
  for (int i = 0; i < n; i++) yt[i] = y[i] + hh * dydx[i];
  f(d, xh, yt, dyt);
  for (int i = 0; i < n; i++) yt[i] = y[i] + hh * dyt[i];
  f(d, xh, yt, dym);
  for (int i = 0; i < n; i++) {
    yt[i] = y[i] + h * dym[i];
    dym[i] = dyt[i] + dym[i];
  }
  f(d, x + h, yt, dyt);
  for (int i = 0; i < n; i++) yout[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
  free(dym); free(dyt); free(yt);
}

/* Cash-Karp step: suppl_routines.f:1289-1356.  A4 is the REAL*4 literal .6
 * of the reference. */
void or_rkck(const or_dyn_ctx *d, const double *y, const double *dydx, int n,
             double x, double h, double *yout, double *yerr, or_derivs_fn f) {
  const double A2 = .2, A3 = .3, A4 = (double)0.6f, A5 = 1.0, A6 = .875;
  const double B21 = .2, B31 = 3.0 / 40.0, B32 = 9.0 / 40.0, B41 = .3,
               B42 = -.9, B43 = 1.2, B51 = -11.0 / 54.0, B52 = 2.5,
               B53 = -70.0 / 27.0, B54 = 35.0 / 27.0,
               B61 = 1631.0 / 55296.0, B62 = 175.0 / 512.0,
               B63 = 575.0 / 13824.0, B64 = 44275.0 / 110592.0,
               B65 = 253.0 / 4096.0, C1 = 37.0 / 378.0, C3 = 250.0 / 621.0,
               C4 = 125.0 / 594.0, C6 = 512.0 / 1771.0;
  const double DC1 = C1 - 2825.0 / 27648.0, DC3 = C3 - 18575.0 / 48384.0,
               DC4 = C4 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0,
               DC6 = C6 - .25;
  size_t sz = sizeof(double) * (n + 8);
  double *ak2 = (double *)malloc(sz), *ak3 = (double *)malloc(sz),
         *ak4 = (double *)malloc(sz), *ak5 = (double *)malloc(sz),
         *ak6 = (double *)malloc(sz), *yt = (double *)calloc(n + 8, sizeof(double));
  for (int i = 0; i < n; i++) yt[i] = y[i] + B21 * h * dydx[i];
  f(d, x + A2 * h, yt, ak2);
  for (int i = 0; i < n; i++) yt[i] = y[i] + h * (B31 * dydx[i] + B32 * ak2[i]);
  f(d, x + A3 * h, yt, ak3);
  for (int i = 0; i < n; i++)
    yt[i] = y[i] + h * (B41 * dydx[i] + B42 * ak2[i] + B43 * ak3[i]);
  f(d, x + A4 * h, yt, ak4);
  for (int i = 0; i < n; i++)
    yt[i] = y[i] + h * (B51 * dydx[i] + B52 * ak2[i] + B53 * ak3[i] + B54 * ak4[i]);
  f(d, x + A5 * h, yt, ak5);
  for (int i = 0; i < n; i++)
    yt[i] = y[i] + h * (B61 * dydx[i] + B62 * ak2[i] + B63 * ak3[i] +
                        B64 * ak4[i] + B65 * ak5[i]);
  f(d, x + A6 * h, yt, ak6);
  for (int i = 0; i < n; i++)
    yout[i] = y[i] + h * (C1 * dydx[i] + C3 * ak3[i] + C4 * ak4[i] + C6 * ak6[i]);
  for (int i = 0; i < n; i++)
    yerr[i] = h * (DC1 * dydx[i] + DC3 * ak3[i] + DC4 * ak4[i] + DC5 * ak5[i] +
                   DC6 * ak6[i]);
  free(ak2); free(ak3); free(ak4); free(ak5); free(ak6); free(yt);
}

static int g_rhs_count = 0;

/* rkqs: suppl_routines.f:1229-1283. */
void or_rkqs(const or_dyn_ctx *d, double *y, const double *dydx, int n,
             double *x, double htry, double eps, const double *yscal,
             double *hdid, double *hnext, double hmin, or_derivs_fn f) {
  const double SAFETY = 0.9, PGROW = -.2, PSHRNK = -.25, ERRCON = 1.89e-4;
  double *yerr = (double *)malloc(sizeof(double) * (n + 8));
  double *ytemp = (double *)malloc(sizeof(double) * (n + 8));
  double h = htry;
  for (;;) {
    or_rkck(d, y, dydx, n, *x, h, ytemp, yerr, f);
    g_rhs_count += 5;
    if (h <= hmin) {
      *hdid = h;
      *x = *x + h;
      for (int i = 0; i < n; i++) y[i] = ytemp[i];
      break;
    }
    double errmax = 0.0;
    for (int i = 0; i < n; i++) {
      double e = fabs(yerr[i] / yscal[i]);
      if (e > errmax) errmax = e;
    }
    errmax = errmax / eps;
    if (errmax > 1.0) {
      double htemp = SAFETY * h * pow(errmax, PSHRNK);
      double a = fabs(htemp), b = 0.1 * fabs(h);
      h = copysign(a > b ? a : b, h);
      continue;
    }
    if (errmax > ERRCON) *hnext = SAFETY * h * pow(errmax, PGROW);
    else *hnext = 5.0 * h;
    *hdid = h;
    *x = *x + h;
    for (int i = 0; i < n; i++) y[i] = ytemp[i];
    break;
  }
  free(yerr); free(ytemp);
}

/* odeint (without the wall-clock interrupt): suppl_routines.f:1362-1463.
 * Returns 1 on MAXSTP exhaustion (fail_odeint). */
static int odeint(const or_dyn_ctx *d, double *ystart, int nvar, double x1,
                  double x2, double eps, double h1, double hmin,
                  double *hnext_out, or_derivs_fn f, int *nsteps) {
  const double TINY = 1.e-30;
  const int MAXSTP = 10000;
  double *y = (double *)calloc((size_t)nvar + 8, sizeof(double));
  double *dydx = (double *)calloc((size_t)nvar + 8, sizeof(double));
  double *yscal = (double *)malloc(sizeof(double) * (nvar + 8));
  double x = x1, h = copysign(h1, x2 - x1), hdid = 0.0, hnext = h;
  memcpy(y, ystart, sizeof(double) * nvar);
  int fail = 1, last_step = 0;
  for (int nstp = 1; nstp <= MAXSTP; nstp++) {
    f(d, x, y, dydx);
    g_rhs_count += 1;
    if (nsteps) (*nsteps)++;
    if (nstp == MAXSTP) last_step = 1;
    if (!last_step) {
      for (int i = 0; i < nvar; i++) yscal[i] = fabs(y[i]) + fabs(h * dydx[i]) + TINY;
      if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
      or_rkqs(d, y, dydx, nvar, &x, h, eps, yscal, &hdid, &hnext, hmin, f);
    } else {
      double hk4 = copysign(fabs(x2 - x), x2 - x1);
      or_rk4(d, y, dydx, nvar, x, hk4, y, f);
      g_rhs_count += 3;
      x = x2 + copysign(TINY, x2 - x1);
    }
    if ((x - x2) * (x2 - x1) >= 0.0) {
      memcpy(ystart, y, sizeof(double) * nvar);
      fail = 0;
      break;
    }
    if (fabs(hnext) < hmin) hnext = copysign(hmin, x2 - x1);
    h = hnext;
  }
  *hnext_out = hnext;
  free(y); free(dydx); free(yscal);
  return fail;
}

static double pot_energy(const or_dyn_ctx *d, double t, const double *q) {
  /* propagator.f:3383-3425 (REF_XZ frame): sum_k d_f M_k Re(v_i* v_j) */
  int N = d->n_states;
  int64_t T = d->total_size;
  double *m = (double *)malloc(sizeof(double) * T);
  or_potential_data_bk(d, q[2 * N], m, NULL);
  double pot = 0.0;
  for (int64_t k = 0; k < T; k++) {
    int i = d->ind_mat[2 * k] - 1, j = d->ind_mat[2 * k + 1] - 1;
    double de = d->delta_E[d->bk_indx[k]];
    double ip = sin(t * de), rp = cos(t * de);
    double df = (i == j) ? 1.0 : 2.0;
    pot += df * m[k] * ((q[i] * q[j] + q[i + N] * q[j + N]) * rp -
                        (q[i + N] * q[j] - q[i] * q[j + N]) * ip);
  }
  free(m);
  return pot;
}

static void record(double *rec, int cap, int *n, double t, const double *cl,
                   double pot) {
  if (!rec || *n >= cap) return;
  double *r = rec + 8 * (size_t)(*n);
  r[0] = t;
  for (int i = 0; i < 6; i++) r[1 + i] = cl[i];
  r[7] = pot;
  (*n)++;
}

/* One trajectory: DOTRAJECT (propagator.f:2283-4645).  Initial conditions
 * 2431-2566, step dispatch 3817-3943 (RK4 / plain odeint, bikram_int off),
 * termination 3995-4050, the SAVE_TRAJECT records 3457-3462 / 4330-4333.
 * q: in/out [2N+8] (CC) or [2N] (AT); for CC modes it is initialised here.
 * rec: optional [rec_cap][8] records of (t, R, P_R, Th, P_Th, Phi, P_Phi, V).*/
int or_dotraject(const or_dyn_ctx *d, const or_traj_param *p, double *q,
                 or_traj_stat *st, double *rec, int rec_cap, int *rec_n) {
  int N = d->n_states;
  int nv = 2 * N + 8;
  double mu = d->mass;
  memset(st, 0, sizeof(*st));
  if (rec_n) *rec_n = 0;
  g_rhs_count = 0;
  int at = (p->mode == 2 || p->mode == 3);
  double k_vec = sqrt(mu * 2.0 * p->E_coll);
  double l_orb = sqrt(p->l_real * (p->l_real + 1.0));
  double R_st = p->R_start;
  double dt = p->dt;
  double *dq = (double *)calloc((size_t)nv, sizeof(double));
  if (!at) {
    memset(q, 0, sizeof(double) * nv);
    q[p->s_st - 1] = 1.0;
    q[2 * N] = R_st;
    q[2 * N + 1] = -sqrt(k_vec * k_vec - l_orb * l_orb / (R_st * R_st));
    q[2 * N + 2] = acos(0.0);
    q[2 * N + 3] = l_orb * sin(0.0);
    q[2 * N + 4] = 0.0;
    q[2 * N + 5] = l_orb * cos(0.0);
  } else {
    memset(q, 0, sizeof(double) * 2 * N);
    q[p->s_st - 1] = 1.0;
  }
  double tcur = 0.0, norm_err = 0.0;
  double tmp_R1 = 0.0, tmp_R2 = 0.0;
  int vibration_cnt = 0, counter = 0, fail = 0, run = 1;
  double dt_step = dt;
  double t_end = at ? d->at_t[d->at_n - 1] : 0.0;
  while (run) {
    counter++;
    if (!at) record(rec, rec_cap, rec_n, tcur, q + 2 * N, pot_energy(d, tcur, q));
    double updated_R, updated_phi;
    if (p->mode == 0) {
      or_derivs_bk(d, tcur, q, dq);
      or_rk4(d, q, dq, nv, tcur, dt, q, or_derivs_bk);
      st->n_rhs += 4;
      st->n_steps++;
    } else if (p->mode == 1) {
      double hn;
      int ns = 0;
      g_rhs_count = 0;
      fail = odeint(d, q, nv, tcur, tcur + dt, p->eps, dt_step, p->hmin, &hn,
                    or_derivs_bk, &ns);
      st->n_rhs += g_rhs_count;
      st->n_steps += ns;
      dt_step = hn;
    } else if (p->mode == 2) {
      if (tcur + dt >= t_end) { dt = t_end - tcur; run = 0; }
      or_derivs_bk_adia(d, tcur, q, dq);
      or_rk4(d, q, dq, 2 * N, tcur, dt, q, or_derivs_bk_adia);
      st->n_rhs += 4;
      st->n_steps++;
    } else {
      double hn;
      int ns = 0;
      g_rhs_count = 0;
      fail = odeint(d, q, 2 * N, tcur, t_end, p->eps, dt_step, p->hmin, &hn,
                    or_derivs_bk_adia, &ns);
      st->n_rhs += g_rhs_count;
      st->n_steps += ns;
      run = 0;
      tcur = t_end;
    }
    double nrm = 0.0;
    for (int i = 0; i < N; i++) nrm += q[i] * q[i] + q[i + N] * q[i + N];
    if (fabs(1.0 - nrm) > norm_err) norm_err = fabs(1.0 - nrm);
    if (at) {
      if (p->mode == 2) tcur += dt;
      if (fail) break;
      continue;
    }
    updated_R = q[2 * N];
    updated_phi = q[2 * N + 4];
    if (fail) break;
    if (tcur > dt && counter > 2)
      if ((tmp_R2 - tmp_R1) < 0.0 && (tmp_R1 - updated_R) > 0.0) vibration_cnt++;
    tmp_R2 = tmp_R1;
    tmp_R1 = updated_R;
    tcur += dt;
    int period_cnt = (int)(fabs(updated_phi) / 2.0 / PI);
    if (p->numb_orb_prds <= period_cnt) run = 0;
    if (vibration_cnt >= p->numb_oscl_prds) run = 0;
    if (p->time_lim > 0.0 && tcur >= p->time_lim) run = 0;
    if (updated_R > R_st) run = 0;
    if (p->max_steps > 0 && st->n_steps >= p->max_steps) { run = 0; fail = 2; }
  }
  if (!at) record(rec, rec_cap, rec_n, tcur + dt, q + 2 * N, pot_energy(d, tcur, q));
  st->t_end = tcur;
  st->fail = fail;
  st->norm_err = norm_err;
  free(dq);
  return fail;
}
