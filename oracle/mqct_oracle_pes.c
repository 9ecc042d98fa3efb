/*
 * mqct_oracle_pes.c -- the three synthetic potentials of the workloads and
 * the V-grid fill of pot_comp.f INI_POT_STORAGE (pot_comp.f:333-1788).
 *
 * TEST INFRASTRUCTURE ONLY (the reference results of the tests); the product
 * evaluates the same potentials in its own HIP kernels.
 *
 * Potentials are in cm^-1, R in bohr, angles in radians.  They play the role
 * of USER_DEFINED_PES (EXAMPLES_IO/.../PES_H2O_He.f:1-10), and the grid fill
 * passes angles exactly as pes_sys_type.f does: type 1 (R, beta)
 * (pes_sys_type.f:2-9), type 4 (R, -beta, -gamma) (pes_sys_type.f:33-39),
 * type 0 (R, 0, beta, gamma, aalpha, bbeta, ggamma) (pes_sys_type.f:88-97).
 *
 * Parameter layouts (shared with the product's definition of the workloads):
 *  Legendre  : par = {L, A_0..A_{L-1}, b_0..b_{L-1}, C_0..C_{L-1}}
 *              V = sum_l (A_l e^{-b_l R} - C_l/R^6) P_l(cos theta)
 *  real-SH   : par = {lmax, {A_l, b_l, C_l, c_{l,-l..l}} for l=0..lmax}
 *              V = sum_l (A_l e^{-b_l R} - C_l/R^6) sum_m c_lm S_lm(beta,gamma)
 *  sites     : par = {alpha, n1, n2, n1 x {x,y,z,q,eps,r0}, n2 x {...}}
 *              Buckingham exp-6 with combined eps = sqrt(eps_i eps_j),
 *              r0 = (r0_i + r0_j)/2, plus q_i q_j / r (Hartree -> cm^-1).
 */
#include "mqct_oracle.h"
#include <math.h>
#include <stdlib.h>

static const double PI = 3.14159265358979323846264338327950288;
static const double HARTREE_CM = 27.21113845 * 8065.5448;

static double legendre_p(int l, double x) {
  if (l == 0) return 1.0;
  double p0 = 1.0, p1 = x;
  for (int n = 2; n <= l; n++) {
    double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

double or_pes_legendre(const double *par, double R, double theta) {
  int L = (int)par[0];
  const double *A = par + 1, *b = par + 1 + L, *C = par + 1 + 2 * L;
  double x = cos(theta), v = 0.0;
  double r6 = R * R * R * R * R * R;
  for (int l = 0; l < L; l++)
    v += (A[l] * exp(-b[l] * R) - C[l] / r6) * legendre_p(l, x);
  const double *w = par + 1 + 3 * L; /* isotropic core Aw exp(-bw (R - Rw)) */
  v += w[0] * exp(-w[1] * (R - w[2]));
  return v;
}

double or_pes_realsh(const double *par, double R, double beta, double gamma) {
  int lmax = (int)par[0];
  const double *p = par + 1;
  double x = cos(beta), v = 0.0;
  double r6 = R * R * R * R * R * R;
  for (int l = 0; l <= lmax; l++) {
    double rad = p[0] * exp(-p[1] * R) - p[2] / r6;
    const double *c = p + 3;
    double ang = 0.0;
    for (int m = -l; m <= l; m++) {
      int am = abs(m);
      double nrm = sqrt((2.0 * l + 1.0) / (4.0 * PI) *
                        exp(or_logfact(l - am) - or_logfact(l + am)));
      double plm = or_plgndr(l, am, x) * nrm;
      double s;
      if (m > 0) s = sqrt(2.0) * plm * cos(m * gamma);
      else if (m < 0) s = sqrt(2.0) * plm * sin(am * gamma);
      else s = plm;
      ang += c[m + l] * s;
    }
    v += rad * ang;
    p += 3 + 2 * l + 1;
  }
  v += p[0] * exp(-p[1] * (R - p[2])); /* isotropic core */
  return v;
}

/* active rotation Rz(a) Ry(b) Rz(g) applied to v */
static void rotate(double a, double b, double g, const double *v, double *o) {
  double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g),
         sg = sin(g);
  double x1 = cg * v[0] - sg * v[1], y1 = sg * v[0] + cg * v[1], z1 = v[2];
  double x2 = cb * x1 + sb * z1, y2 = y1, z2 = -sb * x1 + cb * z1;
  o[0] = ca * x2 - sa * y2;
  o[1] = sa * x2 + ca * y2;
  o[2] = z2;
}

double or_pes_sites(const double *par, double R, double b1, double g1,
                    double a2, double b2, double g2) {
  double alpha = par[0];
  int n1 = (int)par[1], n2 = (int)par[2];
  const double *s1 = par + 3, *s2 = par + 3 + 6 * n1;
  double v = 0.0;
  for (int i = 0; i < n1; i++) {
    double p1[3];
    rotate(0.0, b1, g1, s1 + 6 * i, p1);
    for (int j = 0; j < n2; j++) {
      double p2[3];
      rotate(a2, b2, g2, s2 + 6 * j, p2);
      p2[2] += R;
      double dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
      double r = sqrt(dx * dx + dy * dy + dz * dz);
      double eps = sqrt(s1[6 * i + 4] * s2[6 * j + 4]);
      double r0 = 0.5 * (s1[6 * i + 5] + s2[6 * j + 5]);
      double q = r0 / r, q2 = q * q, q6 = q2 * q2 * q2;
      v += eps * (6.0 / (alpha - 6.0) * exp(alpha * (1.0 - r / r0)) -
                  alpha / (alpha - 6.0) * q6);
      v += s1[6 * i + 3] * s2[6 * j + 3] / r * HARTREE_CM;
    }
  }
  return v;
}

/* V on the (R x angle) grid in the storage order of POT_STORE
 * (iotest.f:33-48; pot_comp.f:374-395 type 1, 612-640 type 4, type 0 grid
 * order of pot_comp.f:3663-3667). pes_kind: 0 Legendre, 1 real-SH, 2 sites. */
void or_vgrid(int coll_type, int pes_kind, const double *par,
              const or_matrix_ctx *c, const double *R, double *V) {
  size_t idx = 0;
  for (int ir = 0; ir < c->n_r; ir++) {
    double r = R[ir];
    if (coll_type == 1) {
      for (int i2 = 0; i2 < c->n_beta; i2++) {
        double beta = c->xbm + c->xbl * c->xb[i2];
        V[idx++] = or_pes_legendre(par, r, beta);
      }
    } else if (coll_type == 4) {
      for (int i2 = 0; i2 < c->n_beta; i2++)
        for (int i3 = 0; i3 < c->n_gamma; i3++) {
          double beta = c->xbm + c->xbl * c->xb[i2];
          double gamma = c->eq_gamma ? c->xg[i3] : c->xgm + c->xgl * c->xg[i3];
          V[idx++] = or_pes_realsh(par, r, -beta, -gamma);
        }
    } else {
      for (int i2 = 0; i2 < c->n_alpha2; i2++)
        for (int i3 = 0; i3 < c->n_beta; i3++)
          for (int i4 = 0; i4 < c->n_beta2; i4++)
            for (int i5 = 0; i5 < c->n_gamma; i5++)
              for (int i6 = 0; i6 < c->n_gamma2; i6++) {
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
                (void)pes_kind;
                V[idx++] = or_pes_sites(par, r, beta, gamma, aalpha, bbeta,
                                        ggamma);
              }
    }
  }
}
