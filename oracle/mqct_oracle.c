/*
 * mqct_oracle.c -- CPU FP64 restatement of MQCT's matrix-build hot path
 * (state tables, angular quadrature, cutoff, spline).  The dynamics half is
 * in mqct_oracle_dyn.c.
 *
 * TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the CPU
 * baseline of the benchmark compare against it.  Nothing in the product
 * (mqct_amd/, libmqct_hip.so) links or calls it.
 *
 * Pins (what each function is checked against, besides the GPU library):
 *   or_jacobi_poly / or_djkm : Wigner-d unitarity and closed forms d^1, d^2
 *   or_plgndr / or_wf_diat_top: Y_jm orthonormality on the Gauss grid
 *   or_w3js / or_cg          : closed-form 3j values and orthogonality
 *   or_gauleg                : numpy.polynomial.legendre.leggauss
 *   or_hamilt_kyro/eigen     : reference level energies and eigenvectors of
 *                              EXAMPLES_IO/SYS_TYPE_4/H2O_He_external_rot_wfn
 *   or_channels_* / or_ini_arrays : state tables and MIJ index lists of the
 *                              MTRX.DAT fixtures (bit-exact)
 *   or_integrator_temp       : closed-form Gaunt (type 1) and 3j (type 4)
 *                              matrix elements, brute-force sums (type 0)
 *   or_spline / or_splint    : exactness on cubics
 *   or_mij_shift / or_read_mij_post : last R point exactly 0, differences
 *                              kept exactly, the read-path spline equals
 *                              the unshifted spline minus the constant
 */
#include "mqct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846264338327950288;

/* ---------------------------------------------------------------------- */
/* Log-factorials: pot_comp.f:311-322 (INI_FACT).                          */
static double *g_logfact = NULL;
static int g_fact_max = -1;

void or_logfact_init(int fact_max) {
  if (fact_max < 99) fact_max = 99;
  if (fact_max <= g_fact_max) return;
  free(g_logfact);
  g_logfact = (double *)malloc(sizeof(double) * (size_t)(fact_max + 1));
  g_logfact[0] = 0.0;
  for (int i = 1; i <= fact_max; i++)
    g_logfact[i] = g_logfact[i - 1] + log((double)i);
  g_fact_max = fact_max;
}

double or_logfact(int n) {
  if (g_fact_max < 0) or_logfact_init(99);
  return g_logfact[n];
}

/* Binomial via log-factorials: pot_comp.f:324-331 (C_N_K). */
static double c_n_k(int n, int k) {
  return exp(or_logfact(n) - or_logfact(n - k) - or_logfact(k));
}

static double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

/* Three-term recursion for P_n^(a,b): suppl_routines.f:1-148. */
void or_jacobi_poly(int n, double alpha, double beta, double x, double *cx) {
  if (n < 0) return;
  cx[0] = 1.0;
  if (n == 0) return;
  cx[1] = (1.0 + 0.5 * (alpha + beta)) * x + 0.5 * (alpha - beta);
  for (int i = 2; i <= n; i++) {
    double ri = (double)i;
    double c1 = 2.0 * ri * (ri + alpha + beta) * (2.0 * ri - 2.0 + alpha + beta);
    double c2 = (2.0 * ri - 1.0 + alpha + beta) * (2.0 * ri + alpha + beta) *
                (2.0 * ri - 2.0 + alpha + beta);
    double c3 = (2.0 * ri - 1.0 + alpha + beta) * (alpha + beta) * (alpha - beta);
    double c4 = -2.0 * (ri - 1.0 + alpha) * (ri - 1.0 + beta) *
                (2.0 * ri + alpha + beta);
    cx[i] = ((c3 + c2 * x) * cx[i - 1] + c4 * cx[i - 2]) / c1;
  }
}

/* Wigner function as MQCT defines it: pot_comp.f:256-309 (Djkm).
 * d via a Jacobi polynomial, times exp(+i(k*gamma + m*alpha)).  The four
 * branch tests are applied in the reference's order, later ones winning. */
void or_djkm(int j, int k, int m, double alpha, double beta, double gamma,
             double *re, double *im) {
  double cx[256];
  int n = j + k;
  if (j - k < n) n = j - k;
  if (j + m < n) n = j + m;
  if (j - m < n) n = j - m;
  int a = 0, lambda = 0;
  if (n == j + m) { a = k - m; lambda = 0; }
  if (n == j - m) { a = m - k; lambda = k - m; }
  if (n == j + k) { a = m - k; lambda = k - m; }
  if (n == j - k) { a = k - m; lambda = 0; }
  int b = 2 * j - 2 * n - a;
  or_jacobi_poly(n, (double)a, (double)b, cos(beta), cx);
  double sgn = ((lambda % 2) == 0) ? 1.0 : -1.0;
  double dkm = cx[n] * sgn * sqrt(c_n_k(2 * j - n, n + a)) /
               sqrt(c_n_k(n + b, b)) * ipow(sin(beta / 2.0), a) *
               ipow(cos(beta / 2.0), b);
  *re = dkm * cos(k * gamma + m * alpha);
  *im = dkm * sin(k * gamma + m * alpha);
}

/* Associated Legendre with Condon-Shortley phase: suppl_routines.f:207-243. */
double or_plgndr(int l, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; i++) {
      pmm = -pmm * fact * somx2;
      fact += 2.0;
    }
  }
  if (l == m) return pmm;
  double pmmp1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.0;
  for (int ll = m + 2; ll <= l; ll++) {
    pll = (x * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

/* 3j symbol with doubled arguments (Racah sum): suppl_routines.f:264-310. */
double or_w3js(int J1, int J2, int J3, int M1, int M2, int M3) {
  if (M1 + M2 + M3 != 0) return 0.0;
  int IA = J1 + J2;
  if (J3 > IA) return 0.0;
  int IB = J1 - J2;
  if (J3 < abs(IB)) return 0.0;
  int JSUM = J3 + IA;
  int IC = J1 - M1;
  int ID = J2 - M2;
  if (JSUM % 2 != 0 || IC % 2 != 0 || ID % 2 != 0) return 0.0;
  if (abs(M1) > J1 || abs(M2) > J2 || abs(M3) > J3) return 0.0;
  int IE = J3 - J2 + M1;
  int IF = J3 - J1 - M2;
  int ZMIN = 0;
  if (-IE > ZMIN) ZMIN = -IE;
  if (-IF > ZMIN) ZMIN = -IF;
  int IG = IA - J3;
  int IH = J2 + M2;
  int ZMAX = IG;
  if (IH < ZMAX) ZMAX = IH;
  if (IC < ZMAX) ZMAX = IC;
  double CC = 0.0;
  for (int Z = ZMIN; Z <= ZMAX; Z += 2) {
    double DENOM = exp(or_logfact(Z / 2) + or_logfact((IG - Z) / 2) +
                       or_logfact((IC - Z) / 2) + or_logfact((IH - Z) / 2) +
                       or_logfact((IE + Z) / 2) + or_logfact((IF + Z) / 2));
    if (Z % 4 != 0) DENOM = -DENOM;
    CC += 1.0 / DENOM;
  }
  double CC1 = exp(or_logfact(IG / 2) + or_logfact((J3 + IB) / 2) +
                   or_logfact((J3 - IB) / 2) - or_logfact((JSUM + 2) / 2));
  double CC2 = exp(or_logfact((J1 + M1) / 2) + or_logfact(IC / 2) +
                   or_logfact(IH / 2) + or_logfact(ID / 2) +
                   or_logfact((J3 - M3) / 2) + or_logfact((J3 + M3) / 2));
  CC = CC * sqrt(CC1 * CC2);
  if ((IB - M3) % 4 != 0) CC = -CC;
  return CC;
}

/* Clebsch-Gordan from the 3j symbol: suppl_routines.f:246-252. */
double or_cg(int j1, int j2, int j3, int m1, int m2, int m3) {
  int e = -j1 + j2 - m3;
  double s = (e % 2 == 0) ? 1.0 : -1.0;
  return or_w3js(2 * j1, 2 * j2, 2 * j3, 2 * m1, 2 * m2, -2 * m3) *
         sqrt(2.0 * j3 + 1.0) * s;
}

/* Gauss-Legendre nodes/weights on [-1,1], Newton to 3e-15:
 * suppl_routines.f:157-204. */
void or_gauleg(int ngp, double *x, double *w) {
  const double EPS = 3.0e-15;
  int m = (ngp + 1) / 2;
  for (int i = 1; i <= m; i++) {
    double z = cos(PI * (i - 0.25) / (ngp + 0.5));
    double z1, pp, p1, p2, p3;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= ngp; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = ngp * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > EPS);
    x[i - 1] = -z;
    x[ngp - i] = z;
    w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
    w[ngp - i] = w[i - 1];
  }
}

/* ---------------------------------------------------------------------- */
/* Y_jm: basis_wave_functions.f:153-175 (WF_DIAT_TOP). */
void or_wf_diat_top(int j, int m, double theta, double phi, double *re,
                    double *im) {
  int am = abs(m);
  double p = or_plgndr(j, am, cos(theta));
  double nrm = sqrt((2.0 * j + 1.0) / 4.0 / PI *
                    exp(or_logfact(j - am) - or_logfact(j + am)));
  double r = p * nrm * cos(m * phi);
  double i = p * nrm * sin(m * phi);
  if (m < 0 && (m % 2 != 0)) { r = -r; i = -i; }
  *re = r;
  *im = i;
}

/* Asymmetric-top function sum_K c_K sqrt((2J+1)/8pi^2) D(J,K,M):
 * basis_wave_functions.f:83-116 (WF_ASYM_TOP). */
void or_wf_asym_top(int J, const double *coeff, int M, double alpha,
                    double beta, double gamma, double *re, double *im) {
  double wr = 0.0, wi = 0.0;
  for (int i = 1; i <= 2 * J + 1; i++) {
    int DK = i - J - 1;
    double dwr, dwi;
    or_djkm(J, DK, M, alpha, beta, gamma, &dwr, &dwi);
    dwr = dwr * sqrt((2.0 * J + 1.0) / 8.0 / (PI * PI));
    dwi = dwi * sqrt((2.0 * J + 1.0) / 8.0 / (PI * PI));
    wr += dwr * coeff[i - 1];
    wi += dwi * coeff[i - 1];
  }
  *re = wr;
  *im = wi;
}

/* Pair function of two asymmetric tops coupled to (j12, m12):
 * basis_wave_functions.f:1-41 (WF_ASYM_TOP_ASYM_TOP). */
static void wf_asym_asym(int j1, int j2, int j12, int m12, const double *c1,
                         const double *c2, double a1, double b1, double g1,
                         double a2, double b2, double g2, double *re,
                         double *im) {
  double wr = 0.0, wi = 0.0;
  for (int m1 = -j1; m1 <= j1; m1++) {
    int m2 = m12 - m1;
    if (abs(m2) > j2) continue;
    double dr1, di1, dr2, di2;
    or_wf_asym_top(j1, c1, m1, a1, b1, g1, &dr1, &di1);
    or_wf_asym_top(j2, c2, m2, a2, b2, g2, &dr2, &di2);
    double cg = or_cg(j1, j2, j12, m1, m2, m12);
    wr += cg * (dr1 * dr2 - di1 * di2);
    wi += cg * (dr1 * di2 + di1 * dr2);
  }
  *re = wr;
  *im = wi;
}

/* ---------------------------------------------------------------------- */
/* Rigid-rotor matrix element in |J K>: asym_top.f:103-137 (hamilt_kyro). */
double or_hamilt_kyro(int J1, int J2, int K1, int K2, double A, double B,
                      double C) {
  double E = 0.0;
  double jzz = (double)(K1 * K1);
  double jj = J1 * (J1 + 1.0);
  if (J1 == J2) {
    if (K1 != K2) {
      if (K1 == K2 - 2 || K1 == K2 + 2) {
        int Kav = (K1 + K2) / 2;
        double jxy = sqrt((J1 + 1.0) * J1 - Kav * K1) *
                     sqrt((J1 + 1.0) * J1 - Kav * K2) / 2.0;
        E = 0.5 * (A - B) * jxy;
      }
    } else {
      E = 0.5 * (A + B) * jj + (C - 0.5 * (A + B)) * jzz;
    }
  }
  return E;
}

/* Ka/Kc labels of the (J+DK+1)-th eigenvalue: asym_top.f:139-173 (STATE). */
int or_state(int J, int DK, int *KA, int *KC) {
  int n = 2 * J + 1;
  int KAa[256], KCa[256];
  KAa[0] = 0;
  for (int i = 2; i <= 2 * J; i += 2) {
    KAa[i - 1] = i / 2;
    KAa[i] = i / 2;
  }
  int ex = 0;
  for (int i = 1; i <= n; i++) {
    KCa[i - 1] = KAa[2 * J + 2 - i - 1];
    int dk = KAa[i - 1] - KCa[i - 1];
    if (DK == dk) {
      *KA = KAa[i - 1];
      *KC = KCa[i - 1];
      ex = 1;
    }
  }
  return ex;
}

/* Cyclic Jacobi diagonalisation of a small symmetric matrix (row-major).
 * Eigenvectors come back as columns of v. */
static void jacobi_eig(int n, double *a, double *d, double *v) {
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) v[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a[p * n + q];
        if (apq == 0.0) continue;
        double app = a[p * n + p], aqq = a[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) /
                   (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; i++) d[i] = a[i * n + i];
}

/* All eigenpairs of the (2J+1)x(2J+1) rigid-rotor matrix built from
 * hamilt_kyro (asym_top.f:258-301, EIGEN_VEC's matrix), ascending, with the
 * sign convention: the largest-|c| component (lowest K index on ties within
 * 1e-9) is positive.  LAPACK DSYEVX fixes signs in an implementation-defined
 * way, so callers compare eigenvectors only up to sign.  The matrix is
 * diagonalised in the Wang basis (K even/odd x symmetric/antisymmetric under
 * K -> -K), where it splits exactly into four blocks, so nearly degenerate
 * levels of different symmetry are never mixed.
 * z is column-major [i + n*col]: column col is the eigenvector of w[col]. */
int or_asym_top_eigen(double A, double B, double C, int J, double *w,
                      double *z) {
  int n = 2 * J + 1;
  double *H = (double *)calloc((size_t)n * n, sizeof(double));
  for (int K1 = -J; K1 <= J; K1++)
    for (int K2 = -J; K2 <= J; K2++)
      H[(K1 + J) * n + (K2 + J)] = or_hamilt_kyro(J, J, K1, K2, A, B, C);
  /* Wang basis vectors, grouped by block. */
  double *U = (double *)calloc((size_t)n * n, sizeof(double)); /* rows=basis */
  int blk_start[5], nb = 0, cnt = 0;
  for (int par = 0; par < 2; par++)
    for (int sym = 0; sym < 2; sym++) {
      blk_start[nb++] = cnt;
      for (int K = 0; K <= J; K++) {
        if (K % 2 != par) continue;
        if (K == 0) {
          if (sym == 0) { U[cnt * n + J] = 1.0; cnt++; }
          continue;
        }
        double s = 1.0 / sqrt(2.0);
        U[cnt * n + (J + K)] = s;
        U[cnt * n + (J - K)] = (sym == 0) ? s : -s;
        cnt++;
      }
    }
  blk_start[nb] = cnt;
  double *wall = (double *)malloc(sizeof(double) * n);
  double *vall = (double *)calloc((size_t)n * n, sizeof(double)); /* cols */
  for (int b = 0; b < 4; b++) {
    int s0 = blk_start[b], m = blk_start[b + 1] - s0;
    if (m == 0) continue;
    double *hb = (double *)malloc(sizeof(double) * m * m);
    double *vb = (double *)malloc(sizeof(double) * m * m);
    double *db = (double *)malloc(sizeof(double) * m);
    for (int p = 0; p < m; p++)
      for (int q = 0; q < m; q++) {
        double acc = 0.0;
        for (int i = 0; i < n; i++) {
          double ui = U[(s0 + p) * n + i];
          if (ui == 0.0) continue;
          for (int j = 0; j < n; j++)
            acc += ui * H[i * n + j] * U[(s0 + q) * n + j];
        }
        hb[p * m + q] = acc;
      }
    jacobi_eig(m, hb, db, vb);
    for (int c = 0; c < m; c++) {
      wall[s0 + c] = db[c];
      for (int i = 0; i < n; i++) {
        double acc = 0.0;
        for (int p = 0; p < m; p++) acc += U[(s0 + p) * n + i] * vb[p * m + c];
        vall[i + n * (s0 + c)] = acc;
      }
    }
    free(hb); free(vb); free(db);
  }
  /* sort ascending (stable selection) */
  int *ord = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (wall[ord[j]] < wall[ord[i]]) { int t = ord[i]; ord[i] = ord[j]; ord[j] = t; }
  for (int c = 0; c < n; c++) {
    int src = ord[c];
    w[c] = wall[src];
    double big = 0.0;
    for (int i = 0; i < n; i++) if (fabs(vall[i + n * src]) > big) big = fabs(vall[i + n * src]);
    int piv = 0;
    for (int i = 0; i < n; i++)
      if (fabs(vall[i + n * src]) >= big * (1.0 - 1e-9)) { piv = i; break; }
    double s = vall[piv + n * src] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) z[i + n * c] = s * vall[i + n * src];
  }
  free(ord); free(wall); free(vall); free(U); free(H);
  return 0;
}

/* Eigenvector and energy of level (J, KA, KC): asym_top.f:280-388
 * (EIGEN_VEC) and asym_top.f:1-101 (EIGEN_VAL): the (J+K1+1)-th eigenvalue
 * belongs to the K1 for which STATE gives (KA, KC). */
int or_eigen_vec(double A, double B, double C, int J, int KA, int KC,
                 double *E, double *zout) {
  int n = 2 * J + 1;
  double *w = (double *)malloc(sizeof(double) * n);
  double *z = (double *)malloc(sizeof(double) * n * n);
  or_asym_top_eigen(A, B, C, J, w, z);
  int found = 0;
  for (int K1 = -J; K1 <= J; K1++) {
    int ka, kc;
    or_state(J, K1, &ka, &kc);
    if (ka == KA && kc == KC) {
      int col = J + K1;
      if (E) *E = w[col];
      if (zout)
        for (int K2 = -J; K2 <= J; K2++) zout[K2 + J] = z[(J + K2) + n * col];
      found = 1;
    }
  }
  free(w); free(z);
  return found;
}

/* ---------------------------------------------------------------------- */
/* Channel lists from JMIN/JMAX. */

/* Type 1, iotest.f:5426-5440: E = Be j(j+1) - De (j(j+1))^2. */
int or_channels_type1(int jmin, int jmax, double Be, double De, int *j_ch,
                      double *E_ch) {
  if (jmin < 0) jmin = 0;
  int i = 0;
  for (int j = jmin; j <= jmax; j++) {
    j_ch[i] = j;
    E_ch[i] = Be * j * (j + 1.0) - De * (j * (j + 1.0)) * (j * (j + 1.0));
    i++;
  }
  return i;
}

/* Type 4, iotest.f:6600-6626: j ascending, then dk = -j..j through STATE. */
int or_channels_type4(int jmin, int jmax, double A, double B, double C,
                      int *j_ch, int *ka_ch, int *kc_ch, double *E_ch) {
  if (jmin < 0) jmin = 0;
  int i = 0;
  for (int j = jmin; j <= jmax; j++)
    for (int dk = -j; dk <= j; dk++) {
      int ka, kc;
      or_state(j, dk, &ka, &kc);
      j_ch[i] = j;
      ka_ch[i] = ka;
      kc_ch[i] = kc;
      or_eigen_vec(A, B, C, j, ka, kc, &E_ch[i], NULL);
      i++;
    }
  return i;
}

/* Type 0 (non-identical), iotest.f:10321-10411: product of the two
 * molecules' level lists, molecule 1 outermost. */
int or_channels_type0(int jmin1, int jmax1, int jmin2, int jmax2,
                      double A1, double B1, double C1, double A2, double B2,
                      double C2, int *j1, int *ka1, int *kc1, int *j2,
                      int *ka2, int *kc2, double *E_ch) {
  int n1 = (jmax1 - jmin1 + 1) * (jmax1 + jmin1 + 1);
  int n2 = (jmax2 - jmin2 + 1) * (jmax2 + jmin2 + 1);
  int *tj1 = (int *)malloc(sizeof(int) * n1 * 3);
  int *tj2 = (int *)malloc(sizeof(int) * n2 * 3);
  double *e1 = (double *)malloc(sizeof(double) * n1);
  double *e2 = (double *)malloc(sizeof(double) * n2);
  int i = 0;
  for (int j = jmin1; j <= jmax1; j++)
    for (int dk = -j; dk <= j; dk++) {
      int ka, kc;
      or_state(j, dk, &ka, &kc);
      tj1[3 * i] = j; tj1[3 * i + 1] = ka; tj1[3 * i + 2] = kc;
      or_eigen_vec(A1, B1, C1, j, ka, kc, &e1[i], NULL);
      i++;
    }
  i = 0;
  for (int j = jmin2; j <= jmax2; j++)
    for (int dk = -j; dk <= j; dk++) {
      int ka, kc;
      or_state(j, dk, &ka, &kc);
      tj2[3 * i] = j; tj2[3 * i + 1] = ka; tj2[3 * i + 2] = kc;
      or_eigen_vec(A2, B2, C2, j, ka, kc, &e2[i], NULL);
      i++;
    }
  int n = 0;
  for (int a = 0; a < n1; a++)
    for (int b = 0; b < n2; b++) {
      j1[n] = tj1[3 * a]; ka1[n] = tj1[3 * a + 1]; kc1[n] = tj1[3 * a + 2];
      j2[n] = tj2[3 * b]; ka2[n] = tj2[3 * b + 1]; kc2[n] = tj2[3 * b + 2];
      E_ch[n] = e1[a] + e2[b];
      n++;
    }
  free(tj1); free(tj2); free(e1); free(e2);
  return n;
}

/* State enumeration and the pair list: matrix_ini.f:58-255 (channel ->
 * first state), 260-545 (states), 553-630 (ind_mat: st1 >= st2, equal m12).
 * Types 1 and 4: m = -j..j per channel.  Type 0 (non-identical):
 * j12 = |j1-j2|..j1+j2, m12 = -j12..j12 per channel.
 * Returns total_size; writes at most cap_pairs pairs (ind_mat may be NULL to
 * only count).  chann_indx gets the reference's "channel -> state" index. */
int64_t or_ini_arrays(int coll_type, int n_ch, const int *j1_ch,
                      const int *j2_ch, int *j12, int *m12, int *indx_chann,
                      int *chann_indx, int64_t cap_pairs, int *ind_mat,
                      int *n_states_out) {
  int st = 0;
  for (int i = 0; i < n_ch; i++) {
    if (coll_type == 1 || coll_type == 4) {
      int j = j1_ch[i];
      for (int m = -j; m <= j; m++) {
        st++;
        if (m == 0 && chann_indx) chann_indx[i] = st;
        if (j12) { j12[st - 1] = j; m12[st - 1] = m; indx_chann[st - 1] = i + 1; }
      }
    } else { /* type 0 */
      int ja = j1_ch[i], jb = j2_ch[i];
      int lo = abs(ja - jb), hi = ja + jb;
      for (int js = lo; js <= hi; js++)
        for (int m = -js; m <= js; m++) {
          st++;
          if (m == 0 && js == lo && chann_indx) chann_indx[i] = st;
          if (j12) { j12[st - 1] = js; m12[st - 1] = m; indx_chann[st - 1] = i + 1; }
        }
    }
  }
  if (n_states_out) *n_states_out = st;
  if (!j12) return -1;
  int64_t total = 0;
  for (int s1 = 1; s1 <= st; s1++)
    for (int s2 = 1; s2 <= s1; s2++)
      if (m12[s1 - 1] == m12[s2 - 1]) {
        if (ind_mat && total < cap_pairs) {
          ind_mat[2 * total] = s1;
          ind_mat[2 * total + 1] = s2;
        }
        total++;
      }
  return total;
}

/* ---------------------------------------------------------------------- */
/* One matrix element at one R: pot_comp.f:1790-3737 (INTEGRATOR_TEMP),
 * CASE(1) 1947-1952, CASE(4) 2503-2560, CASE(0) 3660-3726 (V taken from the
 * stored grid, as with VGRID_FILE=YES; its values equal the on-the-fly PES
 * calls of the default path). */
static double beta_of(const or_matrix_ctx *c, int i) {
  double b = c->xbm + c->xbl * c->xb[i];
  if (b < 0.0) b = 0.0;
  if (b > PI) b = PI;
  return b;
}
static double gamma_of(const or_matrix_ctx *c, int i) {
  return c->eq_gamma ? c->xg[i] : c->xgm + c->xgl * c->xg[i];
}

double or_integrator_temp(const or_matrix_ctx *c, const int *ind_mat,
                          int64_t k, int i_r) {
  int st1 = ind_mat[2 * k] - 1, st2 = ind_mat[2 * k + 1] - 1;
  int ch1 = c->indx_chann[st1] - 1, ch2 = c->indx_chann[st2] - 1;
  double acc = 0.0;
  if (c->coll_type == 1) {
    int j1 = c->j1_ch[ch1], j2 = c->j1_ch[ch2];
    int m1 = c->m12[st1], m2 = c->m12[st2];
    for (int i2 = 0; i2 < c->n_beta; i2++) {
      double beta = beta_of(c, i2);
      double dr1, di1, dr2, di2;
      or_wf_diat_top(j1, m1, beta, 0.0, &dr1, &di1);
      or_wf_diat_top(j2, m2, beta, 0.0, &dr2, &di2);
      double V = c->V[(size_t)i_r * c->n_beta + i2];
      double integrant = (dr1 * dr2 + di1 * di2) * c->wb[i2] * sin(beta) * V;
      if (c->abs_mode) integrant = fabs(integrant);
      acc += integrant * c->xbl * 2.0 * PI;
    }
  } else if (c->coll_type == 4) {
    int j1 = c->j1_ch[ch1], j2 = c->j1_ch[ch2];
    int m1 = c->m12[st1], m2 = c->m12[st2];
    const double *c1 = c->eig1 + (size_t)ch1 * c->ldvec;
    const double *c2 = c->eig1 + (size_t)ch2 * c->ldvec;
    for (int i2 = 0; i2 < c->n_beta; i2++) {
      double beta = beta_of(c, i2);
      for (int i3 = 0; i3 < c->n_gamma; i3++) {
        double gamma = gamma_of(c, i3);
        double V = c->V[((size_t)i_r * c->n_beta + i2) * c->n_gamma + i3];
        double dr1, di1, dr2, di2;
        or_wf_asym_top(j1, c1, m1, 0.0, beta, gamma, &dr1, &di1);
        or_wf_asym_top(j2, c2, m2, 0.0, beta, gamma, &dr2, &di2);
        double integrant;
        if (!c->ref_xz)
          integrant = (dr1 * dr2 + di1 * di2 - dr1 * di2 + di1 * dr2);
        else
          integrant = (dr1 * dr2 + di1 * di2);
        integrant = integrant * c->wb[i2] * c->wg[i3] * sin(beta);
        if (c->abs_mode) integrant = fabs(integrant * V) / fabs(V == 0.0 ? 1.0 : V);
        acc += integrant * c->xbl * c->xgl * 2.0 * PI * (c->abs_mode ? fabs(V) : V);
      }
    }
  } else { /* type 0 */
    int j1a = c->j1_ch[ch1], j2a = c->j2_ch[ch1];
    int j1b = c->j1_ch[ch2], j2b = c->j2_ch[ch2];
    const double *c1a = c->eig1 + (size_t)ch1 * c->ldvec;
    const double *c2a = c->eig2 + (size_t)ch1 * c->ldvec;
    const double *c1b = c->eig1 + (size_t)ch2 * c->ldvec;
    const double *c2b = c->eig2 + (size_t)ch2 * c->ldvec;
    int js1 = c->j12[st1], ms1 = c->m12[st1];
    int js2 = c->j12[st2], ms2 = c->m12[st2];
    size_t G = (size_t)c->n_alpha2 * c->n_beta * c->n_beta2 * c->n_gamma *
               c->n_gamma2;
    const double *Vr = c->V + (size_t)i_r * G;
    size_t g = 0;
    for (int i2 = 0; i2 < c->n_alpha2; i2++)
      for (int i3 = 0; i3 < c->n_beta; i3++)
        for (int i4 = 0; i4 < c->n_beta2; i4++)
          for (int i5 = 0; i5 < c->n_gamma; i5++)
            for (int i6 = 0; i6 < c->n_gamma2; i6++, g++) {
              double beta = c->xbm + c->xbl * c->xb[i3];
              double bbeta = c->xbbm + c->xbbl * c->xbb[i4];
              double gamma, aalpha, ggamma;
              if (!c->eq_gamma) {
                gamma = c->xgm + c->xgl * c->xg[i5];
                aalpha = c->xaam + c->xaal * c->xaa[i2];
                ggamma = c->xggm + c->xggl * c->xgg[i6];
              } else {
                gamma = c->xg[i5];
                aalpha = c->xaa[i2];
                ggamma = c->xgg[i6];
              }
              double V = Vr[g];
              double dr1, di1, dr2, di2;
              wf_asym_asym(j1a, j2a, js1, ms1, c1a, c2a, 0.0, beta, gamma,
                           aalpha, bbeta, ggamma, &dr1, &di1);
              wf_asym_asym(j1b, j2b, js2, ms2, c1b, c2b, 0.0, beta, gamma,
                           aalpha, bbeta, ggamma, &dr2, &di2);
              double integrant = (dr1 * dr2 + di1 * di2) * c->wb[i3] *
                                 c->wg[i5] * sin(beta) * c->waa[i2] *
                                 c->wbb[i4] * c->wgg[i6] * sin(bbeta) * V;
              if (c->abs_mode) integrant = fabs(integrant);
              acc += integrant * c->xbl * c->xgl * c->xaal * c->xbbl *
                     c->xggl * 2 * PI;
            }
  }
  return acc;
}

/* Natural/clamped cubic spline second derivatives: suppl_routines.f:914-965. */
void or_spline(const double *x, const double *y, int n, double yp1,
               double ypn, double *y2) {
  double *u = (double *)malloc(sizeof(double) * n);
  if (yp1 > 0.99e30) {
    y2[0] = 0.0;
    u[0] = 0.0;
  } else {
    y2[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
  }
  for (int i = 1; i < n - 1; i++) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) -
                   (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                (x[i + 1] - x[i - 1]) -
            sig * u[i - 1]) /
           p;
  }
  double qn, un;
  if (ypn > 0.99e30) {
    qn = 0.0;
    un = 0.0;
  } else {
    qn = 0.5;
    un = (3.0 / (x[n - 1] - x[n - 2])) *
         (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
  }
  y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
  free(u);
}

/* Spline evaluation with derivative: suppl_routines.f:968-1002. */
void or_splint(const double *xa, const double *ya, const double *y2a, int n,
               double x, double *y, double *yprime) {
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    int k = (khi + klo) / 2; /* same midpoint as the 1-based (khi+klo)/2 */
    if (xa[k] > x) khi = k; else klo = k;
  }
  double h = xa[khi] - xa[klo];
  double a = (xa[khi] - x) / h;
  double b = (x - xa[klo]) / h;
  *y = a * ya[klo] + b * ya[khi] +
       ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.0;
  if (yprime)
    *yprime = -ya[klo] / h + ya[khi] / h +
              (-(3 * a * a - 1.0) * y2a[klo] / h +
               (3 * b * b - 1.0) * y2a[khi] / h) * (h * h) / 6.0;
}

/* minloc(|R - r0|), 1-based: matrix_ini.f:1969-1986. */
int or_nearest_r(const double *R, int n, double r0) {
  int best = 0;
  double bd = fabs(R[0] - r0);
  for (int i = 1; i < n; i++) {
    double d = fabs(R[i] - r0);
    if (d < bd) { bd = d; best = i; }
  }
  return best + 1;
}

/* The matrix-build loop: matrix_ini.f:1916-2016.  Element k, R-point i:
 * Mat_el(i,k) = INTEGRATOR_TEMP * conv_unit_e; column k zeroed when
 * max(|M(ir1)|, |M(ir2)|) < MIJ_ZERO_CUT (Hartree); then the spline with
 * finite-difference end slopes gives Mat_el_der.  mat_el / mat_el_der are
 * [T][n_r] (Fortran Mat_el(n_r, T)).  ir_cut1/2 are 1-based. */
void or_matrix_make(const or_matrix_ctx *c, int64_t total_size,
                    const int *ind_mat, const double *R, double conv_unit_e,
                    double mij_zero_cut_hartree, int ir_cut1, int ir_cut2,
                    double *mat_el, double *mat_el_der, uint8_t *skipped) {
  int n = c->n_r;
  for (int64_t k = 0; k < total_size; k++) {
    double *col = mat_el + (size_t)k * n;
    double *der = mat_el_der + (size_t)k * n;
    for (int i = 0; i < n; i++)
      col[i] = or_integrator_temp(c, ind_mat, k, i) * conv_unit_e;
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

/* MIJ_SHIFT in MATRIX_MAKE's compute branch (matrix_ini.f:2226-2230), run
 * after the spline: Mat_el(i,:) = Mat_el(i,:) - Mat_el(n_r_coll,:) for
 * i = 1..n_r_coll in that order, so rows i < n_r_coll subtract the
 * unshifted last row and the last row becomes 0.  Mat_el_der is left as
 * the unshifted values' second derivatives. */
void or_mij_shift(int64_t total_size, int n, double *mat_el) {
  for (int i = 0; i < n; i++)
    for (int64_t k = 0; k < total_size; k++)
      mat_el[(size_t)k * n + i] = mat_el[(size_t)k * n + i] - mat_el[(size_t)k * n + n - 1];
}

/* READ_MIJ_USER after each column k is read (matrix_ini.f:6603-6621):
 * with MIJ_SHIFT, Mat_el(i,k) - Mat_el(n_r_coll,k) for i in order; then
 * the clamped spline with finite-difference end slopes of the values as
 * they now are. */
void or_read_mij_post(const double *R, int n, int64_t total_size, int shift,
                      double *mat_el, double *mat_el_der) {
  for (int64_t k = 0; k < total_size; k++) {
    double *col = mat_el + (size_t)k * n;
    if (shift)
      for (int i = 0; i < n; i++) col[i] = col[i] - col[n - 1];
    double d0 = (col[1] - col[0]) / (R[1] - R[0]);
    double d1 = (col[n - 1] - col[n - 2]) / (R[n - 1] - R[n - 2]);
    or_spline(R, col, n, d0, d1, mat_el_der + (size_t)k * n);
  }
}
