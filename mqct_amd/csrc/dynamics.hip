// dynamics.hip -- propagation of trajectories on the GPU (plain FP64 form).
//
// One workgroup integrates one trajectory at a time and pulls the next one
// from a device work counter until all are done (persistent work queue).
// The right-hand side restates DERIVS_BK (mqct_eqtns.f:112-499) and
// DERIVS_BK_adia (mqct_eqtns.f:720-879) in rotated amplitudes
// v_s = exp(-i E_s t) a_s:
//   da_i/dt = -i exp(i E_i t) sum_j M_ij(R) v_j                (pair loop)
//   dP_R/dt = -sum_k d_f M'_k Re(conj(v_i) v_j) + centrifugal  (263-266)
//   dP_Phi/dt = sum_k M_k [Coriolis accumulators]               (270-370)
// plus the state-diagonal Coriolis coupling (434-466) and the classical
// equations (473-495).  M(R) comes from the clamped spline of Mat_el
// (potential_data_bk / splint_bk, mqct_eqtns.f:501-611) read from the
// R-major copy Mt[r][k] (the reference's bk_mat_splnt = transpose(Mat_el)).
// Integrators: rk4 (suppl_routines.f:1076-1113), rkck / rkqs
// (1229-1356, A4 is the REAL*4 literal .6), odeint (1362-1463, MAXSTP
// 10000, no wall-clock interrupt), stepping and termination of DOTRAJECT
// (propagator.f:3817-4050).
#include "mqct_internal.h"
#include "dyn_internal.h"

#include <cmath>
#include <cstring>

namespace {

constexpr int BS = 256;

struct Dev {
  int N;
  long long T;
  int n_r;
  const double *R;
  const double *Mt, *Y2t;   // [n_r][T]
  const int *pi, *pj;       // [T], 0-based
  const double *E;          // [N], Hartree
  const int *j12, *m12;     // [N]
  const double *misc;       // [T][4]
  double mu;
  int theta_on;
  // AT paths
  const double *prec;       // [nrec][8]
  const double *pder;       // [nrec][7] y'' of columns 1..7
  const long long *poff;    // [n_paths+1]
  // trajectories
  const mqct_traj_t *traj;
  int n_traj;
  int mode;
  double *q_out;            // [n_traj][2N+8]
  mqct_traj_stat_t *stat;
  double *scratch;          // [grid][SLOTS][2N+8]
  int *counter;
  // optional SAVE_TRAJECT records (propagator.f:3457-3462, 4330-4333)
  double *rec;              // [n_traj][rec_cap][8]
  int rec_cap;
  int *rec_n;               // [n_traj]
};

constexpr int SLOTS = 12;

struct Lds {
  double2 *v;    // [N]
  double2 *w;    // [N]
  double *red;   // [BS/64 * 4]
  double *bcast; // [16]
};

__device__ inline double block_sum(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = x;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < BS / 64; i++) s += red[i];
  __syncthreads();
  return s;
}

__device__ inline double block_max(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = x;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < BS / 64; i++) s = fmax(s, red[i]);
  __syncthreads();
  return s;
}

// Spline of a recorded path column at time t (splint, suppl_routines.f:968).
__device__ double path_at(const Dev &d, int path, int col, double t) {
  long long o0 = d.poff[path], o1 = d.poff[path + 1];
  int n = (int)(o1 - o0);
  const double *rec = d.prec + o0 * 8;
  const double *der = d.pder + o0 * 7;
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    int k = (khi + klo) / 2;
    if (rec[k * 8] > t) khi = k; else klo = k;
  }
  double h = rec[khi * 8] - rec[klo * 8];
  double a = (rec[khi * 8] - t) / h, b = (t - rec[klo * 8]) / h;
  double ylo = rec[klo * 8 + 1 + col], yhi = rec[khi * 8 + 1 + col];
  double y2lo = der[klo * 7 + col], y2hi = der[khi * 7 + col];
  return a * ylo + b * yhi + ((a * a * a - a) * y2lo + (b * b * b - b) * y2hi) * (h * h) / 6.0;
}

// Right-hand side.  q/dq in global memory, visible to the whole block.
__device__ void rhs(const Dev &d, const Lds &L, bool adia, int path, double t,
                    const double *q, double *dq) {
  const int N = d.N, tid = threadIdx.x;
  double Rrr, tet, ptet, pphi_c;
  if (adia) {
    if (tid == 0) {
      L.bcast[0] = path_at(d, path, 0, t);
      L.bcast[1] = path_at(d, path, 2, t);
      L.bcast[2] = path_at(d, path, 3, t);
      L.bcast[3] = path_at(d, path, 5, t);
    }
    __syncthreads();
    Rrr = L.bcast[0];
    tet = L.bcast[1];
    ptet = L.bcast[2];
    pphi_c = L.bcast[3];
  } else {
    Rrr = q[2 * N];
    tet = q[2 * N + 2];
    ptet = q[2 * N + 3];
    pphi_c = q[2 * N + 5];
  }
  (void)ptet;
  const double mu = d.mu;
  const double sin_tet = sin(tet), cos_tet = cos(tet);
  double dThetadT = 0.0;
  if (d.theta_on && !adia) dThetadT = q[2 * N + 3] / Rrr / Rrr / mu;
  const double dPhidT = pphi_c / Rrr / Rrr / mu / (sin_tet * sin_tet);
  // rotated amplitudes
  for (int s = tid; s < N; s += BS) {
    double c = cos(t * d.E[s]), sn = sin(t * d.E[s]);
    double x = q[s], y = q[s + N];
    L.v[s] = make_double2(x * c + y * sn, y * c - x * sn);
    L.w[s] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  // spline interval (splint_bk): R > R_n gives zero coupling
  const int n = d.n_r;
  const bool zero = Rrr > d.R[n - 1];
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    int k = (khi + klo) / 2;
    if (d.R[k] > Rrr) khi = k; else klo = k;
  }
  const double h = d.R[khi] - d.R[klo];
  const double aa = (d.R[khi] - Rrr) / h, bb = (Rrr - d.R[klo]) / h;
  const double cc1 = aa * aa * aa - aa, cc2 = bb * bb * bb - bb;
  const double cc3 = 3 * aa * aa - 1.0, cc4 = 3 * bb * bb - 1.0;
  const double cc5 = (h * h) / 6.0, cc6 = cc5 / h;
  const double *Ml = d.Mt + (long long)klo * d.T, *Mh = d.Mt + (long long)khi * d.T;
  const double *Yl = d.Y2t + (long long)klo * d.T, *Yh = d.Y2t + (long long)khi * d.T;
  double force = 0.0, pphi = 0.0, pthe = 0.0;
  if (!zero) {
    for (long long k = tid; k < d.T; k += BS) {
      int i = d.pi[k], j = d.pj[k];
      double yl = Ml[k], yh = Mh[k], y2l = Yl[k], y2h = Yh[k];
      double M = aa * yl + bb * yh + (cc1 * y2l + cc2 * y2h) * cc5;
      double2 vi = L.v[i], vj = L.v[j];
      atomicAdd(&L.w[i].x, M * vj.x);
      atomicAdd(&L.w[i].y, M * vj.y);
      if (i != j) {
        atomicAdd(&L.w[j].x, M * vi.x);
        atomicAdd(&L.w[j].y, M * vi.y);
      }
      if (!adia) {
        double dM = (yh - yl) / h + (-cc3 * y2l + cc4 * y2h) * cc6;
        double df = (i == j) ? 1.0 : 2.0;
        // Re(conj(v_i) v_j)
        force -= df * dM * (vi.x * vj.x + vi.y * vj.y);
        // Coriolis accumulators, neighbours m+1/m-1 of j and of i
        const double *mc = d.misc + 4 * k;
        int jj = d.j12[j], mi = d.m12[i];
        double2 z = make_double2(0.0, 0.0);
        if (abs(mi + 1) <= jj) { double2 u = L.v[j + 1]; z.x += mc[0] * u.x; z.y += mc[0] * u.y; }
        if (abs(mi - 1) <= jj) { double2 u = L.v[j - 1]; z.x += mc[1] * u.x; z.y += mc[1] * u.y; }
        // Im(conj(v_i) z)
        double p1 = vi.x * z.y - vi.y * z.x;
        int ji = d.j12[i], mj = d.m12[j];
        double2 u2 = make_double2(0.0, 0.0);
        if (abs(mj + 1) <= ji) { double2 u = L.v[i + 1]; u2.x += mc[2] * u.x; u2.y -= mc[2] * u.y; }
        if (abs(mj - 1) <= ji) { double2 u = L.v[i - 1]; u2.x += mc[3] * u.x; u2.y -= mc[3] * u.y; }
        // Im(conj(u) v_j) with u2 = conj(u) accumulated
        double p2 = u2.x * vj.y + u2.y * vj.x;
        pphi += M * (-sin_tet * df / 2.0 * p1 + sin_tet * df / 2.0 * p2);
        if (d.theta_on) {
          // mqct_eqtns.f:285-290, 340-359: d_f/2 Re(e^{-i phi} (...)) terms
          double r1 = 0.0, r2 = 0.0;
          if (abs(mi + 1) <= jj) { double2 u = L.v[j + 1]; r1 += mc[0] * (vi.x * u.x + vi.y * u.y); }
          if (abs(mi - 1) <= jj) { double2 u = L.v[j - 1]; r1 -= mc[1] * (vi.x * u.x + vi.y * u.y); }
          if (abs(mj + 1) <= ji) { double2 u = L.v[i + 1]; r2 += mc[2] * (u.x * vj.x + u.y * vj.y); }
          if (abs(mj - 1) <= ji) { double2 u = L.v[i - 1]; r2 -= mc[3] * (u.x * vj.x + u.y * vj.y); }
          pthe += M * (df / 2.0 * r1 + df / 2.0 * r2);
        }
      }
    }
  }
  if (!adia) {
    force = block_sum(force, L.red);
    pphi = block_sum(pphi, L.red);
    if (d.theta_on) pthe = block_sum(pthe, L.red);
  }
  __syncthreads();
  // quantum derivatives + state-diagonal Coriolis
  for (int s = tid; s < N; s += BS) {
    double c = cos(t * d.E[s]), sn = sin(t * d.E[s]);
    double2 w = L.w[s];
    // -i exp(i E t) w
    double re = sn * w.x + c * w.y;
    double im = -(c * w.x - sn * w.y);
    double j_12 = d.j12[s], m_12 = d.m12[s];
    double m1 = m_12 + 1.0, m0 = m_12 - 1.0;
    double sqrt1 = sqrt(j_12 * (j_12 + 1.0) - m0 * (m0 + 1.0));
    double sqrt2 = sqrt(j_12 * (j_12 + 1.0) - m1 * (m1 - 1.0));
    if (d.theta_on && !adia) {
      re += dPhidT * m_12 * cos_tet * q[s];
      im -= dPhidT * m_12 * cos_tet * q[s + N];
      if (s > 0) {
        re += (dPhidT * sin_tet * q[s - 1 + N] - dThetadT * q[s - 1]) * sqrt1 / 2.0;
        im -= (dPhidT * sin_tet * q[s - 1] - dThetadT * q[s - 1 + N]) * sqrt1 / 2.0;
      }
      if (s < N - 1) {
        re += (dPhidT * sin_tet * q[s + 1 + N] - dThetadT * q[s + 1]) * sqrt2 / 2.0;
        im -= (dPhidT * sin_tet * q[s + 1] - dThetadT * q[s + 1 + N]) * sqrt2 / 2.0;
      }
    } else {
      if (s > 0) {
        re += dPhidT / 2.0 * sqrt1 * q[s - 1 + N];
        im -= dPhidT / 2.0 * sqrt1 * q[s - 1];
      }
      if (s < N - 1) {
        re += dPhidT / 2.0 * sqrt2 * q[s + 1 + N];
        im -= dPhidT / 2.0 * sqrt2 * q[s + 1];
      }
    }
    dq[s] = re;
    dq[s + N] = im;
  }
  if (!adia && tid == 0) {
    const double *c = q + 2 * N;
    double *o = dq + 2 * N;
    o[0] = c[1] / mu;
    o[4] = c[5] / c[0] / c[0] / mu / (sin_tet * sin_tet);
    o[1] = force + c[3] / c[0] / c[0] / mu * c[3] / c[0] +
           c[5] / c[0] / c[0] / mu * c[5] / c[0] / (sin_tet * sin_tet);
    o[5] = pphi;
    o[2] = 0.0;
    o[3] = 0.0;
    if (d.theta_on) {
      o[2] = c[3] / c[0] / c[0] / mu;
      o[3] = pthe + c[5] / c[0] / c[0] / mu * c[5] / (sin_tet * sin_tet * sin_tet) * cos_tet;
    }
    o[6] = 0.0;
    o[7] = 0.0;
  }
  __syncthreads();
}

// Mean-field energy sum_k d_f M_k Re(conj(v_i) v_j) = Re(v^H M v) from the
// LDS state the last rhs call left (propagator.f:3383-3425).
__device__ double pot_from_lds(const Dev &d, const Lds &L) {
  double s = 0.0;
  for (int i = threadIdx.x; i < d.N; i += BS) {
    double2 v = L.v[i], w = L.w[i];
    s += v.x * w.x + v.y * w.y;
  }
  return block_sum(s, L.red);
}

__device__ void put_record(const Dev &d, int it, int &nrec, double t,
                           const double *c, double pot) {
  if (!d.rec) return;
  if (threadIdx.x == 0 && nrec < d.rec_cap) {
    double *r = d.rec + ((long long)it * d.rec_cap + nrec) * 8;
    r[0] = t;
    for (int i = 0; i < 6; i++) r[1 + i] = c[i];
    r[7] = pot;
  }
  nrec++;
}

__device__ inline void vsync() { __syncthreads(); }

// rk4 (suppl_routines.f:1076-1113); yout may alias y.
__device__ void rk4(const Dev &d, const Lds &L, bool adia, int path, double *y,
                    const double *dydx, int n, double x, double h, double *yout,
                    double *yt, double *dyt, double *dym, int &nrhs) {
  const int tid = threadIdx.x;
  double hh = h * 0.5, h6 = h / 6.0, xh = x + hh;
  for (int i = tid; i < n; i += BS) yt[i] = y[i] + hh * dydx[i];
  vsync();
  rhs(d, L, adia, path, xh, yt, dyt);
  for (int i = tid; i < n; i += BS) yt[i] = y[i] + hh * dyt[i];
  vsync();
  rhs(d, L, adia, path, xh, yt, dym);
  for (int i = tid; i < n; i += BS) {
    yt[i] = y[i] + h * dym[i];
    dym[i] = dyt[i] + dym[i];
  }
  vsync();
  rhs(d, L, adia, path, x + h, yt, dyt);
  for (int i = tid; i < n; i += BS) yout[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
  vsync();
  nrhs += 3;
}

struct Stage {
  double *ak2, *ak3, *ak4, *ak5, *ak6, *yt, *yerr, *ytemp;
};

// Cash-Karp step (suppl_routines.f:1289-1356).
__device__ void rkck(const Dev &d, const Lds &L, bool adia, int path,
                     const double *y, const double *dydx, int n, double x,
                     double h, double *yout, double *yerr, const Stage &s) {
  const int tid = threadIdx.x;
  const double A2 = .2, A3 = .3, A4 = (double)0.6f, A5 = 1.0, A6 = .875;
  const double B21 = .2, B31 = 3.0 / 40.0, B32 = 9.0 / 40.0, B41 = .3, B42 = -.9,
               B43 = 1.2, B51 = -11.0 / 54.0, B52 = 2.5, B53 = -70.0 / 27.0,
               B54 = 35.0 / 27.0, B61 = 1631.0 / 55296.0, B62 = 175.0 / 512.0,
               B63 = 575.0 / 13824.0, B64 = 44275.0 / 110592.0,
               B65 = 253.0 / 4096.0, C1 = 37.0 / 378.0, C3 = 250.0 / 621.0,
               C4 = 125.0 / 594.0, C6 = 512.0 / 1771.0;
  const double DC1 = C1 - 2825.0 / 27648.0, DC3 = C3 - 18575.0 / 48384.0,
               DC4 = C4 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0, DC6 = C6 - .25;
  double *yt = s.yt;
  for (int i = tid; i < n; i += BS) yt[i] = y[i] + B21 * h * dydx[i];
  vsync();
  rhs(d, L, adia, path, x + A2 * h, yt, s.ak2);
  for (int i = tid; i < n; i += BS) yt[i] = y[i] + h * (B31 * dydx[i] + B32 * s.ak2[i]);
  vsync();
  rhs(d, L, adia, path, x + A3 * h, yt, s.ak3);
  for (int i = tid; i < n; i += BS)
    yt[i] = y[i] + h * (B41 * dydx[i] + B42 * s.ak2[i] + B43 * s.ak3[i]);
  vsync();
  rhs(d, L, adia, path, x + A4 * h, yt, s.ak4);
  for (int i = tid; i < n; i += BS)
    yt[i] = y[i] + h * (B51 * dydx[i] + B52 * s.ak2[i] + B53 * s.ak3[i] + B54 * s.ak4[i]);
  vsync();
  rhs(d, L, adia, path, x + A5 * h, yt, s.ak5);
  for (int i = tid; i < n; i += BS)
    yt[i] = y[i] + h * (B61 * dydx[i] + B62 * s.ak2[i] + B63 * s.ak3[i] +
                        B64 * s.ak4[i] + B65 * s.ak5[i]);
  vsync();
  rhs(d, L, adia, path, x + A6 * h, yt, s.ak6);
  for (int i = tid; i < n; i += BS) {
    yout[i] = y[i] + h * (C1 * dydx[i] + C3 * s.ak3[i] + C4 * s.ak4[i] + C6 * s.ak6[i]);
    yerr[i] = h * (DC1 * dydx[i] + DC3 * s.ak3[i] + DC4 * s.ak4[i] + DC5 * s.ak5[i] +
                   DC6 * s.ak6[i]);
  }
  vsync();
}

// odeint (suppl_routines.f:1362-1463) with rkqs inlined (1229-1283).
// Returns 1 when MAXSTP is exhausted.
__device__ int odeint(const Dev &d, const Lds &L, bool adia, int path,
                      double *y, double *dydx, double *yscal, int nvar,
                      double x1, double x2, double eps, double h1, double hmin,
                      double &hnext_out, const Stage &s, int &nsteps,
                      int &nrhs) {
  const int tid = threadIdx.x;
  const double TINY = 1.e-30, SAFETY = 0.9, PGROW = -.2, PSHRNK = -.25,
               ERRCON = 1.89e-4;
  const int MAXSTP = 10000;
  double x = x1, h = copysign(h1, x2 - x1), hnext = h;
  for (int nstp = 1; nstp <= MAXSTP; nstp++) {
    rhs(d, L, adia, path, x, y, dydx);
    nrhs++;
    nsteps++;
    if (nstp == MAXSTP) {
      double hk4 = copysign(fabs(x2 - x), x2 - x1);
      rk4(d, L, adia, path, y, dydx, nvar, x, hk4, y, s.yt, s.ak2, s.ak3, nrhs);
      x = x2 + copysign(TINY, x2 - x1);
    } else {
      for (int i = tid; i < nvar; i += BS) yscal[i] = fabs(y[i]) + fabs(h * dydx[i]) + TINY;
      vsync();
      if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
      // rkqs
      double hh = h;
      for (;;) {
        rkck(d, L, adia, path, y, dydx, nvar, x, hh, s.ytemp, s.yerr, s);
        nrhs += 5;
        if (hh <= hmin) break;
        double em = 0.0;
        for (int i = tid; i < nvar; i += BS) em = fmax(em, fabs(s.yerr[i] / yscal[i]));
        em = block_max(em, L.red) / eps;
        if (em > 1.0) {
          double htemp = SAFETY * hh * pow(em, PSHRNK);
          double a = fabs(htemp), b = 0.1 * fabs(hh);
          hh = copysign(a > b ? a : b, hh);
          continue;
        }
        if (em > ERRCON) hnext = SAFETY * hh * pow(em, PGROW);
        else hnext = 5.0 * hh;
        break;
      }
      x = x + hh;
      for (int i = tid; i < nvar; i += BS) y[i] = s.ytemp[i];
      vsync();
    }
    if ((x - x2) * (x2 - x1) >= 0.0) {
      hnext_out = hnext;
      return 0;
    }
    if (fabs(hnext) < hmin) hnext = copysign(hmin, x2 - x1);
    h = hnext;
  }
  hnext_out = hnext;
  return 1;
}

__global__ __launch_bounds__(BS) void traj_kernel(Dev d) {
  extern __shared__ double lds[];
  const int N = d.N, tid = threadIdx.x;
  Lds L;
  L.v = reinterpret_cast<double2 *>(lds);
  L.w = L.v + N;
  L.red = reinterpret_cast<double *>(L.w + N);
  L.bcast = L.red + 16;
  __shared__ int s_traj;
  const int nv = 2 * N + 8;
  double *base = d.scratch + (long long)blockIdx.x * SLOTS * nv;
  double *y = base, *dydx = base + nv, *yscal = base + 2 * nv;
  Stage st;
  st.ak2 = base + 3 * nv;
  st.ak3 = base + 4 * nv;
  st.ak4 = base + 5 * nv;
  st.ak5 = base + 6 * nv;
  st.ak6 = base + 7 * nv;
  st.yt = base + 8 * nv;
  st.yerr = base + 9 * nv;
  st.ytemp = base + 10 * nv;
  const bool adia = (d.mode == MQCT_PROP_AT_RK4 || d.mode == MQCT_PROP_AT_ODEINT);
  for (;;) {
    if (tid == 0) s_traj = atomicAdd(d.counter, 1);
    __syncthreads();
    const int it = s_traj;
    __syncthreads();
    if (it >= d.n_traj) break;
    const mqct_traj_t p = d.traj[it];
    const int path = p.at_path;
    // initial conditions: propagator.f:2431-2566
    const double mu = d.mu;
    const double k_vec = sqrt(mu * 2.0 * p.E_coll);
    const double l_orb = sqrt(p.l_real * (p.l_real + 1.0));
    const double R_st = p.R_start;
    for (int i = tid; i < nv; i += BS) y[i] = 0.0;
    vsync();
    if (tid == 0) {
      y[p.s_st - 1] = 1.0;
      if (!adia) {
        y[2 * N] = R_st;
        y[2 * N + 1] = -sqrt(k_vec * k_vec - l_orb * l_orb / (R_st * R_st));
        y[2 * N + 2] = acos(0.0);
        y[2 * N + 3] = l_orb * sin(0.0);
        y[2 * N + 4] = 0.0;
        y[2 * N + 5] = l_orb * cos(0.0);
      }
    }
    vsync();
    const int nvar = adia ? 2 * N : nv;
    double tcur = 0.0, dt = p.dt, dt_step = p.dt, norm_err = 0.0;
    double tmp_R1 = 0.0, tmp_R2 = 0.0;
    int vib = 0, counter = 0, fail = 0, nsteps = 0, nrhs = 0;
    bool run = true;
    double t_end = 0.0;
    if (adia) {
      long long o1 = d.poff[path + 1];
      t_end = d.prec[(o1 - 1) * 8];
    }
    int nrec = 0;
    while (run) {
      counter++;
      if (d.mode == MQCT_PROP_CC_RK4) {
        rhs(d, L, false, path, tcur, y, dydx);
        if (d.rec) {
          double pot = pot_from_lds(d, L);
          put_record(d, it, nrec, tcur, y + 2 * N, pot);
        }
        rk4(d, L, false, path, y, dydx, nv, tcur, dt, y, st.yt, st.ak2, st.ak3, nrhs);
        nrhs += 1;
        nsteps++;
      } else if (d.mode == MQCT_PROP_CC_ODEINT) {
        double hn;
        fail = odeint(d, L, false, path, y, dydx, yscal, nv, tcur, tcur + dt, p.eps,
                      dt_step, p.hmin, hn, st, nsteps, nrhs);
        dt_step = hn;
      } else if (d.mode == MQCT_PROP_AT_RK4) {
        if (tcur + dt >= t_end) { dt = t_end - tcur; run = false; }
        rhs(d, L, true, path, tcur, y, dydx);
        rk4(d, L, true, path, y, dydx, 2 * N, tcur, dt, y, st.yt, st.ak2, st.ak3, nrhs);
        nrhs += 1;
        nsteps++;
      } else {
        double hn;
        fail = odeint(d, L, true, path, y, dydx, yscal, 2 * N, tcur, t_end, p.eps,
                      dt_step, p.hmin, hn, st, nsteps, nrhs);
        run = false;
        tcur = t_end;
      }
      double nrm = 0.0;
      for (int i = tid; i < N; i += BS) nrm += y[i] * y[i] + y[i + N] * y[i + N];
      nrm = block_sum(nrm, L.red);
      norm_err = fmax(norm_err, fabs(1.0 - nrm));
      if (adia) {
        if (d.mode == MQCT_PROP_AT_RK4) tcur += dt;
        if (fail) break;
        if (p.max_steps > 0 && nsteps >= p.max_steps && run) { run = false; fail = 2; }
        vsync();
        continue;
      }
      double uR = y[2 * N], uphi = y[2 * N + 4];
      if (fail) break;
      if (tcur > dt && counter > 2)
        if ((tmp_R2 - tmp_R1) < 0.0 && (tmp_R1 - uR) > 0.0) vib++;
      tmp_R2 = tmp_R1;
      tmp_R1 = uR;
      tcur += dt;
      int period_cnt = (int)(fabs(uphi) / 2.0 / MQCT_PI);
      if (p.numb_orb_prds <= period_cnt) run = false;
      if (vib >= p.numb_oscl_prds) run = false;
      if (p.time_lim > 0.0 && tcur >= p.time_lim) run = false;
      if (uR > R_st) run = false;
      if (p.max_steps > 0 && nsteps >= p.max_steps) { run = false; fail = 2; }
      vsync();
    }
    (void)nvar;
    if (d.rec && !adia) {
      // final record at tcur + dt, as the reference writes it
      rhs(d, L, false, path, tcur, y, dydx);
      double pot = pot_from_lds(d, L);
      put_record(d, it, nrec, tcur + dt, y + 2 * N, pot);
      if (tid == 0) d.rec_n[it] = nrec < d.rec_cap ? nrec : d.rec_cap;
    }
    double *qo = d.q_out + (long long)it * nv;
    for (int i = tid; i < nv; i += BS) qo[i] = (adia && i >= 2 * N) ? 0.0 : y[i];
    if (tid == 0 && d.stat) {
      mqct_traj_stat_t s;
      s.t_end = tcur;
      s.n_steps = nsteps;
      s.n_rhs = nrhs;
      s.fail = fail;
      s.norm_err = norm_err;
      d.stat[it] = s;
    }
    vsync();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
int mqct_dyn_run_plain(mqct_hip *h, int mode, int n_traj,
                       const mqct_traj_t *traj, double *q_out,
                       mqct_traj_stat_t *stat_out, int rec_cap, double *rec_out,
                       int *rec_n_out) {
  DynState *ds = h->dyn;
  const int N = ds->N;
  const int nv = 2 * N + 8;
  size_t lds = (size_t)N * 2 * sizeof(double2) + 32 * sizeof(double);
  // gfx950 gives a workgroup up to 160 KiB of LDS; above 64 KiB the kernel
  // has to be told so before the launch
  int lds_max = 0;
  HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
  if (lds + 64 > (size_t)lds_max)
    return mqct_fail(MQCT_ERR_ARG, "propagate: the plain per-trajectory form holds 2N complex amplitudes in LDS; N too large (use batched=1)");
  if (lds > 64 * 1024)
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&traj_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
  int grid = n_traj < 2 * cus ? n_traj : 2 * cus;
  if (ds->d_scratch.ensure((size_t)grid * SLOTS * nv * 8) ||
      ds->d_trajp.ensure((size_t)n_traj * sizeof(mqct_traj_t)) ||
      ds->d_q.ensure((size_t)n_traj * nv * 8) ||
      ds->d_stat.ensure((size_t)n_traj * sizeof(mqct_traj_stat_t)) ||
      ds->d_counter.ensure(64))
    return mqct_fail(MQCT_ERR_NOMEM, "propagate: scratch");
  HIP_TRY(hipMemcpyAsync(ds->d_trajp.p, traj, n_traj * sizeof(mqct_traj_t),
                         hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemsetAsync(ds->d_counter.p, 0, 64, h->stream));
  Dev d;
  d.N = N;
  d.T = ds->T;
  d.n_r = ds->n_r;
  d.R = ds->d_R.as<double>();
  d.Mt = ds->d_Mt.as<double>();
  d.Y2t = ds->d_Y2t.as<double>();
  d.pi = ds->d_pi.as<int>();
  d.pj = ds->d_pj.as<int>();
  d.E = ds->d_E.as<double>();
  d.j12 = ds->d_j12.as<int>();
  d.m12 = ds->d_m12.as<int>();
  d.misc = ds->d_misc.as<double>();
  d.mu = ds->mu;
  d.theta_on = ds->theta_on;
  d.prec = ds->d_prec.as<double>();
  d.pder = ds->d_pder.as<double>();
  d.poff = ds->d_poff.as<long long>();
  d.traj = ds->d_trajp.as<mqct_traj_t>();
  d.n_traj = n_traj;
  d.mode = mode;
  d.q_out = ds->d_q.as<double>();
  d.stat = ds->d_stat.as<mqct_traj_stat_t>();
  d.scratch = ds->d_scratch.as<double>();
  d.counter = ds->d_counter.as<int>();
  d.rec = nullptr;
  d.rec_cap = 0;
  d.rec_n = nullptr;
  DevBuf drec, drecn;
  if (rec_cap > 0 && mode == MQCT_PROP_CC_RK4) {
    if (drec.ensure((size_t)n_traj * rec_cap * 8 * 8) || drecn.ensure((size_t)n_traj * 4))
      return mqct_fail(MQCT_ERR_NOMEM, "propagate: record buffer");
    d.rec = drec.as<double>();
    d.rec_cap = rec_cap;
    d.rec_n = drecn.as<int>();
  }
  HIP_TRY(hipEventRecord(h->ev[5], h->stream));
  traj_kernel<<<grid, BS, lds, h->stream>>>(d);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev[6], h->stream));
  HIP_TRY(hipEventSynchronize(h->ev[6]));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev[5], h->ev[6]));
  ds->last_ms = ms;
  if (q_out)
    HIP_TRY(hipMemcpy(q_out, ds->d_q.p, (size_t)n_traj * nv * 8, hipMemcpyDeviceToHost));
  if (stat_out)
    HIP_TRY(hipMemcpy(stat_out, ds->d_stat.p, n_traj * sizeof(mqct_traj_stat_t),
                      hipMemcpyDeviceToHost));
  if (d.rec && rec_out)
    HIP_TRY(hipMemcpy(rec_out, drec.p, (size_t)n_traj * rec_cap * 64, hipMemcpyDeviceToHost));
  if (d.rec && rec_n_out)
    HIP_TRY(hipMemcpy(rec_n_out, drecn.p, (size_t)n_traj * 4, hipMemcpyDeviceToHost));
  return MQCT_OK;
}
