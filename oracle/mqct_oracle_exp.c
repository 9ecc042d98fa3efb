/*
 * mqct_oracle_exp.c -- CPU FP64 restatement of the PES-expansion path:
 * the angular functions of the expansion terms (EXPANSION_TERM,
 * pot_comp.f:3739-3896), the projection of the V grid on them
 * (CALC_EXPANSION_TERMS, pot_comp.f:4695-5920), and the matrix elements
 * built from the expansion coefficients (EXPANSION_MATRIX_ELEMENT,
 * pot_comp.f:6465-8751) inside MATRIX_MAKE's expansion branch
 * (matrix_ini.f:1950-1956, then the same cutoff and spline as the
 * quadrature branch).
 *
 * TEST INFRASTRUCTURE ONLY: the parity tests compare the HIP library with
 * these loops; product code never links, loads or calls them.
 *
 * Term rows are A_TOP(1:5, t): type 1 (L), type 4 (L, NJU), type 0
 * (L1, NJU1, L2, NJU2, L); unused slots are 0.  Expansion coefficients are
 * expansion_terms(n_r, n_terms) in the reference's column-major layout,
 * i.e. terms[t * n_r + ir] here.  Type 0 covers distinguishable
 * partners, IDENTICAL_PES=YES and identical partners (IDENTICAL=YES).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mqct_oracle.h"

#define PI 3.14159265358979323846

static int triang(int a, int b, int c) {
  /* TRIANG_RULE, pot_comp.f:8749-8756 */
  int m = a > b ? a : b;
  if (c > m) m = c;
  return 2 * m <= a + b + c;
}

static double sgn_pow(int e) { return (abs(e) % 2 == 0) ? 1.0 : -1.0; }

/* EXPANSION_TERM (pot_comp.f:3739-3896), cases 1, 4 and 0. */
void or_expansion_term(int coll_type, const int *t, double b1, double g1,
                       double a2, double b2, double g2, double *re,
                       double *im) {
  double wr = 0.0, wi = 0.0;
  if (coll_type == 1) {
    or_wf_diat_top(t[0], 0, b1, 0.0, &wr, &wi);
  } else if (coll_type == 4) {
    double dr, di;
    or_wf_diat_top(t[0], t[1], b1, g1, &dr, &di);
    wr = dr * sgn_pow(t[1]);
    wi = di * sgn_pow(t[1]);
  } else {
    int l1 = t[0], n1 = t[1], l2 = t[2], n2 = t[3], l = t[4];
    int mm = l1 < l2 ? l1 : l2;
    double s1 = sqrt((2.0 * l1 + 1.0) / 8.0 / (PI * PI));
    double s2 = sqrt((2.0 * l2 + 1.0) / 8.0 / (PI * PI));
    for (int m = -mm; m <= mm; m++) {
      double dr1, di1, dr2, di2;
      or_djkm(l1, n1, m, 0.0, b1, g1, &dr1, &di1);
      dr1 *= s1;
      di1 *= s1;
      or_djkm(l2, n2, -m, a2, b2, g2, &dr2, &di2);
      dr2 *= s2;
      di2 *= s2;
      double cg = or_cg(l1, l2, l, m, -m, 0);
      wr += cg * (dr1 * dr2 - di1 * di2);
      wi += cg * (dr1 * di2 + dr2 * di1);
    }
  }
  *re = wr;
  *im = wi;
}

/* Real part of EXPANSION_TERM at every point of the context's grid, in the
 * V-grid order CALC_EXPANSION_TERMS walks (type 4: beta, gamma; type 0:
 * alpha2, beta1, beta2, gamma1, gamma2), angles formed as it forms them. */
void or_expansion_term_grid(const or_matrix_ctx *c, const int *t, double *out) {
  const int nb = c->n_beta, ng = c->n_gamma;
  double wr, wi;
  if (c->coll_type == 1) {
    for (int i2 = 0; i2 < nb; i2++) {
      or_expansion_term(1, t, c->xbm + c->xbl * c->xb[i2], 0, 0, 0, 0, &wr, &wi);
      out[i2] = wr;
    }
    return;
  }
  if (c->coll_type == 4) {
    for (int i2 = 0; i2 < nb; i2++)
      for (int i3 = 0; i3 < ng; i3++) {
        double gamma = c->eq_gamma ? c->xg[i3] : c->xgm + c->xgl * c->xg[i3];
        or_expansion_term(4, t, c->xbm + c->xbl * c->xb[i2], gamma, 0, 0, 0, &wr, &wi);
        out[(size_t)i2 * ng + i3] = wr;
      }
    return;
  }
  const int na = c->n_alpha2, nbb = c->n_beta2, ngg = c->n_gamma2;
  size_t g = 0;
  for (int i2 = 0; i2 < na; i2++)
    for (int i3 = 0; i3 < nb; i3++)
      for (int i4 = 0; i4 < nbb; i4++)
        for (int i5 = 0; i5 < ng; i5++)
          for (int i6 = 0; i6 < ngg; i6++) {
            double gamma = c->eq_gamma ? c->xg[i5] : c->xgm + c->xgl * c->xg[i5];
            double aalpha = c->eq_gamma ? c->xaa[i2] : c->xaam + c->xaal * c->xaa[i2];
            double ggamma = c->eq_gamma ? c->xgg[i6] : c->xggm + c->xggl * c->xgg[i6];
            or_expansion_term(0, t, c->xbm + c->xbl * c->xb[i3], gamma, aalpha,
                              c->xbbm + c->xbbl * c->xbb[i4], ggamma, &wr, &wi);
            out[g++] = wr;
          }
}

/* CALC_EXPANSION_TERMS, the quadrature of each term against the V grid:
 * type 1 pot_comp.f:4969-4988, type 4 5054-5099, type 0 5557-5753 (the
 * real part, intgrlr, is what becomes expansion_terms). */
void or_calc_expansion_terms(const or_matrix_ctx *c, int n_terms,
                             const int *a_top, double *out) {
  const int nb = c->n_beta, ng = c->n_gamma;
  for (int it = 0; it < n_terms; it++) {
    const int *t = a_top + 5 * it;
    for (int ir = 0; ir < c->n_r; ir++) {
      double acc = 0.0, wr, wi;
      if (c->coll_type == 1) {
        for (int i2 = 0; i2 < nb; i2++) {
          double beta = c->xbm + c->xbl * c->xb[i2];
          or_expansion_term(1, t, beta, 0.0, 0.0, 0.0, 0.0, &wr, &wi);
          double V = c->V[(size_t)ir * nb + i2];
          acc += (wr * V) * c->wb[i2] * sin(beta) * c->xbl * 2.0 * PI;
        }
      } else if (c->coll_type == 4) {
        for (int i2 = 0; i2 < nb; i2++)
          for (int i3 = 0; i3 < ng; i3++) {
            double beta = c->xbm + c->xbl * c->xb[i2];
            double gamma = c->eq_gamma ? c->xg[i3] : c->xgm + c->xgl * c->xg[i3];
            or_expansion_term(4, t, beta, gamma, 0.0, 0.0, 0.0, &wr, &wi);
            double V = c->V[((size_t)ir * nb + i2) * ng + i3];
            acc += wr * V * c->wb[i2] * c->wg[i3] * sin(beta) * c->xbl * c->xgl;
          }
      } else {
        const int na = c->n_alpha2, nbb = c->n_beta2, ngg = c->n_gamma2;
        size_t G = (size_t)na * nb * nbb * ng * ngg;
        for (int i2 = 0; i2 < na; i2++)
          for (int i3 = 0; i3 < nb; i3++)
            for (int i4 = 0; i4 < nbb; i4++)
              for (int i5 = 0; i5 < ng; i5++)
                for (int i6 = 0; i6 < ngg; i6++) {
                  double beta = c->xbm + c->xbl * c->xb[i3];
                  double bbeta = c->xbbm + c->xbbl * c->xbb[i4];
                  double gamma, aalpha, ggamma;
                  if (c->eq_gamma) {
                    gamma = c->xg[i5];
                    aalpha = c->xaa[i2];
                    ggamma = c->xgg[i6];
                  } else {
                    gamma = c->xgm + c->xgl * c->xg[i5];
                    aalpha = c->xaam + c->xaal * c->xaa[i2];
                    ggamma = c->xggm + c->xggl * c->xgg[i6];
                  }
                  or_expansion_term(0, t, beta, gamma, aalpha, bbeta, ggamma, &wr, &wi);
                  size_t g = ((((size_t)i2 * nb + i3) * nbb + i4) * ng + i5) * ngg + i6;
                  double V = c->V[(size_t)ir * G + g];
                  double w = V * wr * c->wb[i3] * c->wg[i5] * sin(beta) * c->waa[i2] *
                             c->wbb[i4] * c->wgg[i6] * sin(bbeta);
                  acc += w * c->xbl * c->xgl * c->xaal * c->xbbl * c->xggl * 2 * PI;
                }
      }
      out[(size_t)it * c->n_r + ir] = acc;
    }
  }
}

static double eigc(const or_matrix_ctx *c, int which, int chan1, int j, int K) {
  /* M_VECTORS / M1_VECTORS / M2_VECTORS(chan, j+1+K); chan 1-based */
  const double *e = which == 2 ? c->eig2 : c->eig1;
  return e[(size_t)(chan1 - 1) * c->ldvec + j + K];
}

/* R-independent part of EXPANSION_MATRIX_ELEMENT for pair k: ang[t] and the
 * overall factor (coef1, or coef1*coef2 for type 0). */
static double angular_parts(const or_matrix_ctx *c, int n_terms,
                            const int *a_top, const int *ind_mat, int64_t k,
                            double *ang) {
  int stp = ind_mat[2 * k], stpp = ind_mat[2 * k + 1];
  int chp = c->indx_chann[stp - 1], chpp = c->indx_chann[stpp - 1];
  int m = c->m12[stp - 1];
  for (int i = 0; i < n_terms; i++) ang[i] = 0.0;
  if (c->coll_type == 1) {
    /* pot_comp.f:6551-6600 */
    int jp = c->j1_ch[chp - 1], jpp = c->j1_ch[chpp - 1];
    double coef1 = sqrt((2.0 * jp + 1.0) / (2.0 * jpp + 1.0));
    for (int i = 0; i < n_terms; i++) {
      int l = a_top[5 * i];
      if (!triang(jpp, l, jp)) continue;
      double coef2 = sqrt((2.0 * l + 1.0) / (4.0 * PI));
      ang[i] = coef2 * or_cg(jp, l, jpp, m, 0, m) * or_cg(jp, l, jpp, 0, 0, 0);
    }
    return coef1;
  }
  if (c->coll_type == 4) {
    /* pot_comp.f:6955-7048 */
    int jp = c->j1_ch[chp - 1], jpp = c->j1_ch[chpp - 1];
    double coef1 = sqrt((2.0 * jp + 1.0) / (2.0 * jpp + 1.0));
    for (int i = 0; i < n_terms; i++) {
      int l = a_top[5 * i], nju = a_top[5 * i + 1];
      if (!triang(jpp, l, jp)) continue;
      double coef2 = sqrt((2.0 * l + 1.0) / (4.0 * PI));
      double cg_m = or_cg(jp, l, jpp, m, 0, m);
      double s = 0.0;
      int nplanar = 2 - (nju == 0);
      for (int pl = 1; pl <= nplanar; pl++) {
        double phase = 1.0;
        if (pl == 2) {
          phase = sgn_pow(nju);
          nju = -nju;
        }
        for (int kp = -jp; kp <= jp; kp++) {
          int kpp = kp + nju;
          if (abs(kpp) > jpp) continue;
          double cp = eigc(c, 1, chp, jp, kp), cpp = eigc(c, 1, chpp, jpp, kpp);
          s += phase * sgn_pow(nju) * or_cg(jp, l, jpp, kp, nju, kpp) * cp * cpp;
        }
      }
      ang[i] = coef2 * cg_m * s;
    }
    return coef1;
  }
  /* type 0: pot_comp.f:8257-8716.  Symmetry 1 always; symmetry 2 with
   * IDENTICAL_PES (the swapped term l1 <-> l2); symmetries 3 and 4 with
   * identical partners (the swapped levels), weighted by the exchange
   * parity of the initial state. */
  int j12p = c->j12[stp - 1], j12pp = c->j12[stpp - 1];
  int j1p = c->j1_ch[chp - 1], j2p = c->j2_ch[chp - 1];
  int j1pp = c->j1_ch[chpp - 1], j2pp = c->j2_ch[chpp - 1];
  const int ident = c->identical, ipes = c->identical_pes || c->identical;
  double coef1 = sqrt((2.0 * j1pp + 1.0) / (2.0 * j1p + 1.0));
  double coef2 = sqrt((2.0 * j2pp + 1.0) / (2.0 * j2p + 1.0));
  int par_p = 1, bk1 = 1;
  if (ident) {
    par_p = c->parity_state[stp - 1];
    bk1 = c->parity_state_bk[stp - 1];
    if (stp == stpp && j1p == j2p && c->ka1_ch[chp - 1] == c->ka2_ch[chp - 1] &&
        c->kc1_ch[chp - 1] == c->kc2_ch[chp - 1])
      bk1 = (int)sgn_pow(j12p);
  }
  for (int i = 0; i < n_terms; i++) {
    const int *t = a_top + 5 * i;
    const int l1 = t[0], l2 = t[2], l = t[4];
    const int n1o = t[1], n2o = t[3];
    double coef3 = sqrt((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0)) / (8.0 * PI * PI);
    const int npl = 2 - (n1o == 0 && n2o == 0);
    int sym[4] = {0, 0, 0, 0};
    double cgjk[4] = {0, 0, 0, 0};
    /* the four K sums: (molecule of the p-side level, term index) per
     * symmetry: 1 (l1 on 1, l2 on 2), 2 (l2 on 1, l1 on 2), 3 and 4 the
     * same with the p-side levels of the two molecules exchanged */
    for (int s = 0; s < 4; s++) {
      if (s == 1 && !ipes) continue;
      if (s >= 2 && !ident) continue;
      if ((s == 1 || s == 3) && l1 == l2 && abs(n1o) == abs(n2o)) continue;
      const int la = (s == 0 || s == 2) ? l1 : l2;   /* on molecule 1 (pp side) */
      const int lb = (s == 0 || s == 2) ? l2 : l1;   /* on molecule 2 (pp side) */
      const int ja = s < 2 ? j1p : j2p;               /* p-side level paired with 1 */
      const int jb = s < 2 ? j2p : j1p;
      const int wa = s < 2 ? 1 : 2, wb = s < 2 ? 2 : 1;
      if (!(triang(j1pp, la, ja) && triang(j2pp, lb, jb))) continue;
      sym[s] = 1;
      for (int pl = 1; pl <= npl; pl++) {
        int n1 = n1o, n2 = n2o;
        double sc = 1.0;
        if (pl == 2) {
          sc = sgn_pow(l1 + l2 + l + n1o + n2o);
          n1 = -n1o;
          n2 = -n2o;
        }
        if (s == 1 || s == 3) sc *= sgn_pow(l1 + l2);
        const int na = (s == 0 || s == 2) ? n1 : n2;
        const int nb = (s == 0 || s == 2) ? n2 : n1;
        double db1 = 0.0, db2 = 0.0;
        for (int kp = -ja; kp <= ja; kp++) {
          int kpp = kp - na;
          if (abs(kpp) > j1pp) continue;
          db1 += or_cg(j1pp, la, ja, kpp, na, kp) * eigc(c, wa, chp, ja, kp) *
                 eigc(c, 1, chpp, j1pp, kpp);
        }
        for (int kp = -jb; kp <= jb; kp++) {
          int kpp = kp - nb;
          if (abs(kpp) > j2pp) continue;
          db2 += or_cg(j2pp, lb, jb, kpp, nb, kp) * eigc(c, wb, chp, jb, kp) *
                 eigc(c, 2, chpp, j2pp, kpp);
        }
        cgjk[s] += db1 * db2 * sc * (s >= 2 ? par_p * bk1 : 1);
      }
    }
    if (!(sym[0] || sym[1] || sym[2] || sym[3])) continue;
    if (!ident && !(sym[0] || sym[1])) continue;
    double m5 = 0.0;
    int mx = l1 < l2 ? l1 : l2;
    /* pot_comp.f:8589 runs mp over -j1'..j1' only; the exchanged pair
     * state |j2' mp; j1' m-mp> needs -j2'..j2' as well, and the reference's
     * own H2O+H2O IDENTICAL MTRX.DAT holds those terms, so with identical
     * partners mp covers -max(j1', j2')..max(j1', j2') (the direct terms
     * vanish for |mp| > j1' through their Clebsch-Gordan coefficients) */
    const int mpx = ident && j2p > j1p && !(c->plant & 1) ? j2p : j1p;
    for (int mp = -mpx; mp <= mpx; mp++) {
      double cg_p = 0.0, cg_x = 0.0;
      if (abs(m - mp) > j2p) {
        if (!ident) continue;
      } else {
        cg_p = or_cg(j1p, j2p, j12p, mp, m - mp, m);
      }
      if (ident) {
        if (abs(m - mp) > j1p && abs(m - mp) > j2p) continue;
        cg_x = or_cg(j2p, j1p, j12p, mp, m - mp, m);
      }
      double acc[4] = {0, 0, 0, 0};
      for (int me = -mx; me <= mx; me++) {
        int mpp = mp - me;
        if (abs(mpp) > j1pp) continue;
        if (abs(m - mpp) > j2pp) continue;
        double cg_pp = or_cg(j1pp, j2pp, j12pp, mpp, m - mpp, m);
        double cl = or_cg(l1, l2, l, me, -me, 0);
        if (sym[0])
          acc[0] += cl * cg_pp * or_cg(j1pp, l1, j1p, mpp, me, mp) * or_cg(j2pp, l2, j2p, m - mpp, -me, m - mp);
        if (sym[1])
          acc[1] += cl * cg_pp * or_cg(j1pp, l2, j1p, mpp, me, mp) * or_cg(j2pp, l1, j2p, m - mpp, -me, m - mp);
        if (abs(mp) <= j2p) {
          if (sym[2])
            acc[2] += cl * cg_pp * or_cg(j1pp, l1, j2p, mpp, me, mp) * or_cg(j2pp, l2, j1p, m - mpp, -me, m - mp);
          if (sym[3])
            acc[3] += cl * cg_pp * or_cg(j1pp, l2, j2p, mpp, me, mp) * or_cg(j2pp, l1, j1p, m - mpp, -me, m - mp);
        }
      }
      if (sym[0]) m5 += acc[0] * cg_p * cgjk[0];
      if (sym[1]) m5 += acc[1] * cg_p * cgjk[1];
      if (ident) {
        if (sym[2]) m5 += acc[2] * cg_x * cgjk[2];
        if (sym[3]) m5 += acc[3] * cg_x * cgjk[3];
      }
    }
    ang[i] = coef3 * m5;
  }
  double f = coef1 * coef2;
  if (ident) {
    /* pot_comp.f:8699-8708 */
    int dp = c->ka1_ch[chp - 1] == c->ka2_ch[chp - 1] && c->kc1_ch[chp - 1] == c->kc2_ch[chp - 1];
    int dpp = c->ka1_ch[chpp - 1] == c->ka2_ch[chpp - 1] && c->kc1_ch[chpp - 1] == c->kc2_ch[chpp - 1];
    f = f / sqrt(1.0 + (j1p == j2p) * dp) / sqrt(1.0 + (j1pp == j2pp) * dpp);
  }
  return f;
}

/* EXPANSION_MATRIX_ELEMENT for pair k at every R point. */
void or_expansion_matrix_element(const or_matrix_ctx *c, int n_terms,
                                 const int *a_top, const double *terms,
                                 const int *ind_mat, int64_t k, double *out) {
  double *ang = (double *)malloc(sizeof(double) * (n_terms > 0 ? n_terms : 1));
  double f = angular_parts(c, n_terms, a_top, ind_mat, k, ang);
  for (int ir = 0; ir < c->n_r; ir++) {
    double s = 0.0;
    for (int i = 0; i < n_terms; i++) s += terms[(size_t)i * c->n_r + ir] * ang[i];
    out[ir] = f * s;
  }
  free(ang);
}

/* MATRIX_MAKE with expansion_defined (matrix_ini.f:1950-1956 and the
 * cutoff and spline of 1975-2016). */
void or_expansion_matrix_make(const or_matrix_ctx *c, int n_terms,
                              const int *a_top, const double *terms,
                              int64_t total_size, const int *ind_mat,
                              const double *R, double conv_unit_e,
                              double mij_zero_cut_hartree, int ir_cut1,
                              int ir_cut2, double *mat_el, double *mat_el_der,
                              uint8_t *skipped) {
  int n = c->n_r;
  for (int64_t k = 0; k < total_size; k++) {
    double *col = mat_el + (size_t)k * n;
    double *der = mat_el_der + (size_t)k * n;
    or_expansion_matrix_element(c, n_terms, a_top, terms, ind_mat, k, col);
    for (int i = 0; i < n; i++) col[i] *= conv_unit_e;
    int skip = 0;
    double a1 = fabs(col[ir_cut1 - 1]), a2 = fabs(col[ir_cut2 - 1]);
    if ((a1 > a2 ? a1 : a2) < mij_zero_cut_hartree) {
      for (int i = 0; i < n; i++) { col[i] = 0.0; der[i] = 0.0; }
      skip = 1;
    }
    if (skipped) skipped[k] = (uint8_t)skip;
    if (!skip) {
      double d0 = (col[1] - col[0]) / (R[1] - R[0]);
      double d1 = (col[n - 1] - col[n - 2]) / (R[n - 1] - R[n - 2]);
      or_spline(R, col, n, d0, d1, der);
    }
  }
}
