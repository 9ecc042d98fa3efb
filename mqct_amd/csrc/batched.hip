// batched.hip -- the interval-bucketed batched form of the AT-MQCT
// propagation (DERIVS_BK_adia, mqct_eqtns.f:720-879, integrated by odeint /
// rkqs / rkck or rk4 as in propagator.f:3873-3908).
//
// All active trajectories advance together, one right-hand-side evaluation
// each per iteration, every trajectory at its own stage of its own step:
//   prep   : stage vector of each trajectory, its time, the path values
//            (R, Theta, P_Phi) and the spline knots klo/khi of R; the
//            rotated amplitudes v = exp(-iEt) y_stage (block order); the
//            trajectory is appended to the buckets of its two knots.
//   gemm   : for every knot r and m-block b, one dense product over the
//            whole bucket: W = [Y_r | Y2_r] [w_Y v ; w_Y2 v] on FP64 MFMA,
//            where Y_r (Mat_el at knot r) and Y2_r (its spline second
//            derivatives) are the block's dense matrices and w_Y, w_Y2 the
//            trajectory's spline weights for that knot (splint_bk,
//            mqct_eqtns.f:575-607).  Each knot's matrices are read once per
//            iteration for all trajectories that need them.
//   finish : da/dt = -i exp(iEt) (W_lo + W_hi) + state-diagonal Coriolis
//            (mqct_eqtns.f:845-877) and the integrator's state machine
//            (Cash-Karp stages, error control and step-size rules of rkqs,
//            step loop and end condition of odeint; or classical rk4).
// The three launches of an iteration are captured once in a HIP graph and
// replayed; the host reads the number of running trajectories between
// replays.
#include "dyn_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int BS = 256;
constexpr int NSLOT = 8;  // y, k1..k6, ys
constexpr int MAXSTP = 10000;  // odeint (suppl_routines.f:1362)

struct Ctl {
  double x, x1, x2, h, hnext, hmin, eps;
  double tau, aa, bb, c1, c2, dphidt;  // current stage evaluation
  double norm_err;
  int stage, nstp, status, zero;       // status: 0 run, 1 done
  int klo, path, nrhs;
  int rk4;                             // 1: the stages are rk4's (AT rk4, or odeint's last step)
  int last;                            // rk4 stepping: the current step ends the path
  int pad;
};

struct BDev {
  int N, nblk, n_r, n_traj, mode;
  const double *R;
  const double *KM;           // [n_r][2][S2] dense block matrices Y, Y2
  long long S2;
  const int *blk_start;       // [nblk+1] positions
  const long long *blk_off;   // [nblk] offset of the block in S2
  const double *E;            // per position, Hartree
  const int *nb_lo, *nb_hi;   // per position: position of state-1 / state+1, -1 none
  const double *sq1, *sq2;    // per position: Coriolis factors
  const double *prec, *pder;  // paths
  const long long *poff;
  double *S;                  // [n_traj][NSLOT][2N]
  double2 *V;                 // [n_traj][N]
  double2 *W;                 // [n_traj][2][N]
  Ctl *ctl;
  int *cnt;                   // [n_r]
  int2 *bucket;               // [n_r][n_traj] (t, side)
  int *active;                // running trajectories
  unsigned long long *nbucketed;  // right-hand sides with R inside the grid
  double *scal;               // [n_traj][2N] yscal
};

__device__ double path_at(const BDev &d, int path, int col, double t) {
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
  return a * rec[klo * 8 + 1 + col] + b * rec[khi * 8 + 1 + col] +
         ((a * a * a - a) * der[klo * 7 + col] + (b * b * b - b) * der[khi * 7 + col]) * (h * h) / 6.0;
}

// Cash-Karp tableau (suppl_routines.f:1306-1321); A4 = REAL*4 .6
__constant__ double CK_A[6] = {0.0, .2, .3, (double)0.6f, 1.0, .875};
__constant__ double CK_B[6][5] = {
    {0, 0, 0, 0, 0},
    {.2, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
    {.3, -.9, 1.2, 0, 0},
    {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};

__global__ __launch_bounds__(BS) void prep_kernel(BDev d) {
  const int t = blockIdx.x, tid = threadIdx.x, N = d.N, n2 = 2 * N;
  Ctl c = d.ctl[t];
  if (c.status != 0) return;
  double *S = d.S + (long long)t * NSLOT * n2;
  const double *y = S;
  double *ys = S + 7 * n2;
  const int s = c.stage;
  double tau;
  if (!c.rk4) {
    tau = c.x + CK_A[s] * c.h;
    for (int i = tid; i < n2; i += BS) {
      double v = y[i];
      if (s > 0) {
        double acc = 0.0;
        for (int q = 0; q < s; q++) acc += CK_B[s][q] * S[(1 + q) * n2 + i];
        // rkck forms y + h*(sum b k) (and y + B21*h*k1 for the first stage)
        v = (s == 1) ? y[i] + CK_B[1][0] * c.h * S[n2 + i] : y[i] + c.h * acc;
      }
      ys[i] = v;
    }
  } else {  // rk4: x, x+h/2, x+h/2, x+h
    double hh = c.h * 0.5;
    tau = (s == 0) ? c.x : (s == 3 ? c.x + c.h : c.x + hh);
    for (int i = tid; i < n2; i += BS) {
      double v = y[i];
      if (s == 1) v = y[i] + hh * S[n2 + i];
      else if (s == 2) v = y[i] + hh * S[2 * n2 + i];
      else if (s == 3) v = y[i] + c.h * S[3 * n2 + i];
      ys[i] = v;
    }
  }
  __syncthreads();
  __shared__ double sh[8];
  if (tid == 0) {
    double R = path_at(d, c.path, 0, tau);
    double tet = path_at(d, c.path, 2, tau);
    double pphi = path_at(d, c.path, 5, tau);
    double st = sin(tet);
    double mu = d.R[-1];  // stored just before the R grid
    sh[0] = R;
    sh[1] = pphi / R / R / mu / (st * st);
  }
  __syncthreads();
  const double R = sh[0];
  if (tid == 0) {
    const int n = d.n_r;
    c.tau = tau;
    c.dphidt = sh[1];
    c.zero = R > d.R[n - 1];
    int klo = 0, khi = n - 1;
    while (khi - klo > 1) {
      int k = (khi + klo) / 2;
      if (d.R[k] > R) khi = k; else klo = k;
    }
    double h = d.R[khi] - d.R[klo];
    double aa = (d.R[khi] - R) / h, bb = (R - d.R[klo]) / h;
    double cc5 = (h * h) / 6.0;
    c.klo = klo;
    c.aa = aa;
    c.bb = bb;
    c.c1 = (aa * aa * aa - aa) * cc5;
    c.c2 = (bb * bb * bb - bb) * cc5;
    if (!c.zero) {
      atomicAdd(d.nbucketed, 1ULL);
      int p0 = atomicAdd(&d.cnt[klo], 1);
      d.bucket[(long long)klo * d.n_traj + p0] = make_int2(t, 0);
      int p1 = atomicAdd(&d.cnt[khi], 1);
      d.bucket[(long long)khi * d.n_traj + p1] = make_int2(t, 1);
    }
    d.ctl[t] = c;
  }
  // rotated amplitudes
  double2 *v = d.V + (long long)t * N;
  for (int i = tid; i < N; i += BS) {
    double cs = cos(tau * d.E[i]), sn = sin(tau * d.E[i]);
    double xr = ys[i], xi = ys[i + N];
    v[i] = make_double2(xr * cs + xi * sn, xi * cs - xr * sn);
  }
}

// Interval GEMM.  One workgroup per (knot r, m-block b, 64-row group):
//   W[t][side][rows] = [Y_r | Y2_r] [w_Y v_t ; w_Y2 v_t]
// for every bucket entry (t, side) of knot r, 64 entries (128 real columns:
// Re and Im of each v) per column super-tile.  Each wave owns a 16-row tile
// and keeps 8 column tiles of f64 16x16x4 MFMA accumulators; the weighted
// vectors are staged through LDS in slices of GK rows of K = 2 nb, the
// matrix fragments come straight from HBM/L2 (the blocks are symmetric, so
// lanes of one k read 16 consecutive rows).
constexpr int GE = 64;            // entries per column super-tile
constexpr int GC = 2 * GE / 16;   // MFMA column tiles
constexpr int GK = 16;            // K rows per LDS slice
constexpr int GLD = 2 * GE + 16;  // LDS row stride in doubles (half-waves hit disjoint banks)

__global__ __launch_bounds__(BS) void gemm_kernel(BDev d) {
  const int r = blockIdx.x, b = blockIdx.y;
  const int ne = d.cnt[r];
  if (ne == 0) return;
  const int p0 = d.blk_start[b], nb = d.blk_start[b + 1] - p0;
  const int row0 = blockIdx.z * 64;
  if (row0 >= nb) return;
  const int K = 2 * nb;
  const double *Y = d.KM + ((long long)r * 2) * d.S2 + d.blk_off[b];
  const double *Y2 = d.KM + ((long long)r * 2 + 1) * d.S2 + d.blk_off[b];
  __shared__ double Bs[GK][GLD];
  __shared__ int s_t[GE], s_side[GE];
  __shared__ double s_wy[GE], s_wy2[GE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rt0 = row0 + wave * 16;
  const bool wave_rows = rt0 < nb;
  const int rr = rt0 + (lane & 15), kq = lane >> 4;
  const int2 *bk = d.bucket + (long long)r * d.n_traj;
  for (int e0 = 0; e0 < ne; e0 += GE) {
    const int nE = min(GE, ne - e0);
    const int nct = (2 * nE + 15) / 16;
    __syncthreads();
    if (tid < GE) {
      int t = -1, side = 0;
      double wy = 0.0, wy2 = 0.0;
      if (tid < nE) {
        int2 ent = bk[e0 + tid];
        const Ctl &c = d.ctl[ent.x];
        t = ent.x;
        side = ent.y;
        wy = side ? c.bb : c.aa;
        wy2 = side ? c.c2 : c.c1;
      }
      s_t[tid] = t;
      s_side[tid] = side;
      s_wy[tid] = wy;
      s_wy2[tid] = wy2;
    }
    dbl4 acc[GC];
#pragma unroll
    for (int ct = 0; ct < GC; ct++) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += GK) {
      __syncthreads();
      for (int idx = tid; idx < GK * GE; idx += BS) {
        const int kk = idx % GK, e = idx / GK, k = k0 + kk;
        double vx = 0.0, vy = 0.0;
        if (e < nE && k < K) {
          const bool lo = k < nb;
          const double w = lo ? s_wy[e] : s_wy2[e];
          const double2 x = d.V[(long long)s_t[e] * d.N + p0 + (lo ? k : k - nb)];
          vx = w * x.x;
          vy = w * x.y;
        }
        Bs[kk][2 * e] = vx;
        Bs[kk][2 * e + 1] = vy;
      }
      double af[GK / 4];
#pragma unroll
      for (int s4 = 0; s4 < GK / 4; s4++) {
        const int k = k0 + 4 * s4 + kq;
        double a = 0.0;
        if (rr < nb && k < K) a = (k < nb) ? Y[(long long)k * nb + rr] : Y2[(long long)(k - nb) * nb + rr];
        af[s4] = a;
      }
      __syncthreads();
      if (wave_rows) {
#pragma unroll
        for (int s4 = 0; s4 < GK / 4; s4++) {
          const double *brow = &Bs[4 * s4 + kq][lane & 15];
#pragma unroll
          for (int ct = 0; ct < GC; ct++)
            if (ct < nct) acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s4], brow[16 * ct], acc[ct], 0, 0, 0);
        }
      }
    }
    // f64 16x16x4 output: row lane/16 + 4i, column lane%16 of each tile
    if (wave_rows) {
#pragma unroll
      for (int ct = 0; ct < GC; ct++) {
        if (ct >= nct) break;
        const int col = 16 * ct + (lane & 15), e = col >> 1;
        if (e >= nE) continue;
        double *wb = reinterpret_cast<double *>(d.W + ((long long)s_t[e] * 2 + s_side[e]) * d.N + p0) + (col & 1);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int orow = rt0 + (lane >> 4) + 4 * i;
          if (orow < nb) wb[2 * orow] = acc[ct][i];
        }
      }
    }
  }
}

__device__ inline double block_sum(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < BS / 64; i++) s += red[i];
  __syncthreads();
  return s;
}

__device__ inline double block_max(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < BS / 64; i++) s = fmax(s, red[i]);
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(BS) void finish_kernel(BDev d) {
  const int t = blockIdx.x, tid = threadIdx.x, N = d.N, n2 = 2 * N;
  __shared__ double red[8];
  if (t == 0)
    for (int i = tid; i < d.n_r; i += BS) d.cnt[i] = 0;  // the gemm has consumed the buckets
  Ctl c = d.ctl[t];
  if (c.status != 0) return;
  double *S = d.S + (long long)t * NSLOT * n2;
  double *y = S, *ys = S + 7 * n2;
  const int s = c.stage;
  double *k = S + (1 + s) * n2;  // k1 .. k6 (rk4 uses k1..k4)
  const double2 *W0 = d.W + ((long long)t * 2) * N, *W1 = W0 + N;
  for (int i = tid; i < N; i += BS) {
    double wx = 0.0, wy = 0.0;
    if (!c.zero) {
      wx = W0[i].x + W1[i].x;
      wy = W0[i].y + W1[i].y;
    }
    double cs = cos(c.tau * d.E[i]), sn = sin(c.tau * d.E[i]);
    double re = sn * wx + cs * wy;
    double im = -(cs * wx - sn * wy);
    int lo = d.nb_lo[i], hi = d.nb_hi[i];
    if (lo >= 0) {
      re += c.dphidt / 2.0 * d.sq1[i] * ys[lo + N];
      im -= c.dphidt / 2.0 * d.sq1[i] * ys[lo];
    }
    if (hi >= 0) {
      re += c.dphidt / 2.0 * d.sq2[i] * ys[hi + N];
      im -= c.dphidt / 2.0 * d.sq2[i] * ys[hi];
    }
    k[i] = re;
    k[i + N] = im;
  }
  __syncthreads();
  c.nrhs++;
  const double TINY = 1.e-30;
  if (c.rk4) {
    if (s < 3) {
      c.stage = s + 1;
    } else {
      double h6 = c.h / 6.0;
      const double *k1 = S + n2, *k2 = S + 2 * n2, *k3 = S + 3 * n2, *k4 = S + 4 * n2;
      for (int i = tid; i < n2; i += BS) y[i] = y[i] + h6 * (k1[i] + k4[i] + 2.0 * (k2[i] + k3[i]));
      c.nstp++;
      c.stage = 0;
      if (d.mode == MQCT_PROP_AT_ODEINT) {
        // odeint's MAXSTP-th step: one rk4 step to x2 (suppl_routines.f:1385-1391)
        c.x = c.x2 + copysign(TINY, c.x2 - c.x1);
        c.status = 1;
      } else {
        // DOTRAJECT's AT rk4 loop (propagator.f:3889-3908): a step that
        // would reach the path end is shortened to end there and is the last
        c.x += c.h;
        if (c.last) c.status = 1;
        else if (c.x + c.h >= c.x2) { c.h = c.x2 - c.x; c.last = 1; }
      }
      __syncthreads();
      double nrm = 0.0;
      for (int i = tid; i < N; i += BS) nrm += y[i] * y[i] + y[i + N] * y[i + N];
      nrm = block_sum(nrm, red);
      c.norm_err = fmax(c.norm_err, fabs(1.0 - nrm));
    }
  } else if (s == 0) {
    if (c.nstp == MAXSTP - 1) {
      // continue as rk4 with k1 = dydx at x
      c.rk4 = 1;
      c.h = copysign(fabs(c.x2 - c.x), c.x2 - c.x1);
      c.stage = 1;
    } else {
      // odeint: yscal from dydx at the start of the step, then clamp h
      double *ysc = d.scal + (long long)t * n2;
      for (int i = tid; i < n2; i += BS) ysc[i] = fabs(y[i]) + fabs(c.h * k[i]) + TINY;
      if ((c.x + c.h - c.x2) * (c.x + c.h - c.x1) > 0.0) c.h = c.x2 - c.x;
      c.stage = 1;
    }
  } else if (s < 5) {
    c.stage = s + 1;
  } else {
    const double C1 = 37.0 / 378.0, C3 = 250.0 / 621.0, C4 = 125.0 / 594.0, C6 = 512.0 / 1771.0;
    const double DC1 = C1 - 2825.0 / 27648.0, DC3 = C3 - 18575.0 / 48384.0,
                 DC4 = C4 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0, DC6 = C6 - .25;
    const double *k1 = S + n2, *k3 = S + 3 * n2, *k4 = S + 4 * n2, *k5 = S + 5 * n2, *k6 = S + 6 * n2;
    const double *ysc = d.scal + (long long)t * n2;
    double em = 0.0;
    for (int i = tid; i < n2; i += BS) {
      double yerr = c.h * (DC1 * k1[i] + DC3 * k3[i] + DC4 * k4[i] + DC5 * k5[i] + DC6 * k6[i]);
      em = fmax(em, fabs(yerr / ysc[i]));
    }
    em = block_max(em, red) / c.eps;
    // rkqs (suppl_routines.f:1229-1283): a step at or below hmin is taken as is
    bool accept = (c.h <= c.hmin) || !(em > 1.0);
    if (!accept) {
      double htemp = 0.9 * c.h * pow(em, -.25);
      double a = fabs(htemp), b = 0.1 * fabs(c.h);
      c.h = copysign(a > b ? a : b, c.h);
      c.stage = 1;  // retry with the same k1
    } else {
      if (c.h > c.hmin) c.hnext = (em > 1.89e-4) ? 0.9 * c.h * pow(em, -.2) : 5.0 * c.h;
      for (int i = tid; i < n2; i += BS)
        y[i] = y[i] + c.h * (C1 * k1[i] + C3 * k3[i] + C4 * k4[i] + C6 * k6[i]);
      c.x += c.h;
      c.nstp++;
      c.stage = 0;
      if ((c.x - c.x2) * (c.x2 - c.x1) >= 0.0) {
        c.status = 1;
      } else {
        if (fabs(c.hnext) < c.hmin) c.hnext = copysign(c.hmin, c.x2 - c.x1);
        c.h = c.hnext;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (c.status != 0) atomicSub(d.active, 1);
    d.ctl[t] = c;
  }
}

// dense block matrices of every knot from the R-major pair arrays
__global__ void knot_matrices_kernel(long long T, int n_r, const double *__restrict__ Mt,
                                     const double *__restrict__ Y2t, const int *__restrict__ pi,
                                     const int *__restrict__ pj, const int *__restrict__ state_pos,
                                     const int *__restrict__ pos_blk, const int *__restrict__ blk_start,
                                     const long long *__restrict__ blk_off, long long S2,
                                     double *__restrict__ KM) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= T * n_r) return;
  int r = (int)(idx / T);
  long long k = idx % T;
  int a = state_pos[pi[k]], b = state_pos[pj[k]];
  int blk = pos_blk[a];
  int nb = blk_start[blk + 1] - blk_start[blk];
  int la = a - blk_start[blk], lb = b - blk_start[blk];
  double m = Mt[(long long)r * T + k], m2 = Y2t[(long long)r * T + k];
  double *Yr = KM + ((long long)r * 2) * S2 + blk_off[blk];
  double *Y2r = KM + ((long long)r * 2 + 1) * S2 + blk_off[blk];
  Yr[(long long)la * nb + lb] = m;
  Yr[(long long)lb * nb + la] = m;
  Y2r[(long long)la * nb + lb] = m2;
  Y2r[(long long)lb * nb + la] = m2;
}

__global__ void init_kernel(BDev d, const mqct_traj_t *__restrict__ tr,
                            const int *__restrict__ state_pos) {
  const int t = blockIdx.x, tid = threadIdx.x, n2 = 2 * d.N;
  double *S = d.S + (long long)t * NSLOT * n2;
  for (int i = tid; i < n2; i += BS) S[i] = 0.0;
  __syncthreads();
  if (tid == 0) {
    const mqct_traj_t p = tr[t];
    S[state_pos[p.s_st - 1]] = 1.0;
    Ctl c{};
    long long o1 = d.poff[p.at_path + 1];
    c.x = 0.0;
    c.x1 = 0.0;
    c.x2 = d.prec[(o1 - 1) * 8];
    c.rk4 = d.mode == MQCT_PROP_AT_RK4;
    c.h = p.dt;
    if (c.rk4 && c.x + c.h >= c.x2) { c.h = c.x2 - c.x; c.last = 1; }
    c.hnext = c.h;
    c.hmin = p.hmin;
    c.eps = p.eps;
    c.path = p.at_path;
    c.status = 0;
    c.stage = 0;
    d.ctl[t] = c;
  }
}

__global__ void out_kernel(BDev d, const int *__restrict__ pos_state, double *__restrict__ q,
                           mqct_traj_stat_t *__restrict__ st) {
  const int t = blockIdx.x, tid = threadIdx.x, N = d.N, n2 = 2 * N;
  const double *y = d.S + (long long)t * NSLOT * n2;
  double *o = q + (long long)t * (n2 + 8);
  for (int p = tid; p < N; p += BS) {
    int s = pos_state[p];
    o[s] = y[p];
    o[s + N] = y[p + N];
  }
  for (int i = tid; i < 8; i += BS) o[n2 + i] = 0.0;
  __shared__ double red[8];
  double nrm = 0.0;
  for (int i = tid; i < N; i += BS) nrm += y[i] * y[i] + y[i + N] * y[i + N];
  nrm = block_sum(nrm, red);
  if (tid == 0) {
    const Ctl &c = d.ctl[t];
    mqct_traj_stat_t s;
    s.t_end = (d.mode == MQCT_PROP_AT_ODEINT) ? c.x2 : c.x;
    s.n_steps = c.nstp;
    s.n_rhs = c.nrhs;
    s.fail = 0;
    s.norm_err = (d.mode == MQCT_PROP_AT_ODEINT) ? fabs(1.0 - nrm) : c.norm_err;
    st[t] = s;
  }
}

}  // namespace

int mqct_dyn_run_batched(mqct_hip *h, int mode, int n_traj, const mqct_traj_t *traj,
                         double *q_out, mqct_traj_stat_t *stat_out) {
  DynState *ds = h->dyn;
  const int N = ds->N, n_r = ds->n_r;
  const long long T = ds->T;
  (void)hipGetLastError();  // a failed earlier call must not be reported here
  // block order of the states (m12 ascending, states ascending)
  std::vector<int> pos_state, state_pos(N), blk_start, pos_blk(N);
  std::vector<long long> blk_off;
  {
    std::vector<std::pair<int, int>> key(N);
    for (int s = 0; s < N; s++) key[s] = {ds->m12[s], s};
    std::sort(key.begin(), key.end());
    long long off = 0;
    for (int p = 0; p < N; p++) {
      pos_state.push_back(key[p].second);
      if (p == 0 || key[p].first != key[p - 1].first) blk_start.push_back(p);
    }
    blk_start.push_back(N);
    for (size_t b = 0; b + 1 < blk_start.size(); b++) {
      long long nb = blk_start[b + 1] - blk_start[b];
      blk_off.push_back(off);
      off += nb * nb;
      for (int p = blk_start[b]; p < blk_start[b + 1]; p++) pos_blk[p] = (int)b;
    }
    ds->bat_S2 = off;
  }
  for (int p = 0; p < N; p++) state_pos[pos_state[p]] = p;
  const int nblk = (int)blk_start.size() - 1;
  int maxnb = 0;
  for (int b = 0; b < nblk; b++) maxnb = std::max(maxnb, blk_start[b + 1] - blk_start[b]);
  if (maxnb > 64 * 1024)
    return mqct_fail(MQCT_ERR_ARG, "propagate: an m-block is too large for the batched form");
  // per position: energies, Coriolis neighbours (state-1 / state+1) and factors
  std::vector<double> E(N), sq1(N), sq2(N);
  std::vector<int> nlo(N), nhi(N);
  for (int p = 0; p < N; p++) {
    int s = pos_state[p];
    E[p] = ds->E[s];
    double j = ds->j12[s], m = ds->m12[s];
    double m1 = m + 1.0, m0 = m - 1.0;
    sq1[p] = std::sqrt(j * (j + 1.0) - m0 * (m0 + 1.0));
    sq2[p] = std::sqrt(j * (j + 1.0) - m1 * (m1 - 1.0));
    nlo[p] = s > 0 ? state_pos[s - 1] : -1;
    nhi[p] = s < N - 1 ? state_pos[s + 1] : -1;
  }
  const long long S2 = ds->bat_S2;
  const int n2 = 2 * N;
  // device buffers
  DevBuf &B = ds->d_bat;
  size_t need_km = (size_t)n_r * 2 * S2 * 8;
  if (ds->d_KM.ensure(need_km)) return mqct_fail(MQCT_ERR_NOMEM, "batched: knot matrices");
  DevBuf d_ps, d_sp, d_bs, d_bo, d_pb, d_E, d_nlo, d_nhi, d_sq1, d_sq2, d_R, d_tr;
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
    if (b.ensure(bytes)) return 1;
    return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, h->stream) != hipSuccess;
  };
  std::vector<double> Rmu(n_r + 1);
  Rmu[0] = ds->mu;
  for (int i = 0; i < n_r; i++) Rmu[i + 1] = ds->R[i];
  if (up(d_ps, pos_state.data(), N * 4) || up(d_sp, state_pos.data(), N * 4) ||
      up(d_bs, blk_start.data(), blk_start.size() * 4) || up(d_bo, blk_off.data(), blk_off.size() * 8) ||
      up(d_pb, pos_blk.data(), N * 4) || up(d_E, E.data(), N * 8) || up(d_nlo, nlo.data(), N * 4) ||
      up(d_nhi, nhi.data(), N * 4) || up(d_sq1, sq1.data(), N * 8) || up(d_sq2, sq2.data(), N * 8) ||
      up(d_R, Rmu.data(), (n_r + 1) * 8) || up(d_tr, traj, (size_t)n_traj * sizeof(mqct_traj_t)))
    return mqct_fail(MQCT_ERR_NOMEM, "batched: tables");
  if (!ds->km_valid) {
    HIP_TRY(hipMemsetAsync(ds->d_KM.p, 0, need_km, h->stream));
    knot_matrices_kernel<<<(unsigned)((T * n_r + 255) / 256), 256, 0, h->stream>>>(
        T, n_r, ds->d_Mt.as<double>(), ds->d_Y2t.as<double>(), ds->d_pi.as<int>(), ds->d_pj.as<int>(),
        d_sp.as<int>(), d_pb.as<int>(), d_bs.as<int>(), d_bo.as<long long>(), S2, ds->d_KM.as<double>());
    HIP_TRY(hipGetLastError());
    ds->km_valid = true;
  }
  size_t off_S = 0, off_V = off_S + (size_t)n_traj * NSLOT * n2 * 8;
  size_t off_W = off_V + (size_t)n_traj * N * 16;
  size_t off_C = off_W + (size_t)n_traj * 2 * N * 16;
  size_t off_cnt = off_C + (size_t)n_traj * sizeof(Ctl);
  size_t off_bk = off_cnt + ((size_t)n_r * 4 + 255) / 256 * 256;
  size_t off_act = off_bk + (size_t)n_r * n_traj * 8;
  size_t off_scal = off_act + 256;
  size_t total = off_scal + (size_t)n_traj * n2 * 8;
  if (B.ensure(total)) return mqct_fail(MQCT_ERR_NOMEM, "batched: trajectory state");
  char *base = B.as<char>();
  BDev d;
  d.N = N;
  d.nblk = nblk;
  d.n_r = n_r;
  d.n_traj = n_traj;
  d.mode = mode;
  d.R = d_R.as<double>() + 1;
  d.KM = ds->d_KM.as<double>();
  d.S2 = S2;
  d.blk_start = d_bs.as<int>();
  d.blk_off = d_bo.as<long long>();
  d.E = d_E.as<double>();
  d.nb_lo = d_nlo.as<int>();
  d.nb_hi = d_nhi.as<int>();
  d.sq1 = d_sq1.as<double>();
  d.sq2 = d_sq2.as<double>();
  d.prec = ds->d_prec.as<double>();
  d.pder = ds->d_pder.as<double>();
  d.poff = ds->d_poff.as<long long>();
  d.S = reinterpret_cast<double *>(base + off_S);
  d.V = reinterpret_cast<double2 *>(base + off_V);
  d.W = reinterpret_cast<double2 *>(base + off_W);
  d.ctl = reinterpret_cast<Ctl *>(base + off_C);
  d.cnt = reinterpret_cast<int *>(base + off_cnt);
  d.bucket = reinterpret_cast<int2 *>(base + off_bk);
  d.active = reinterpret_cast<int *>(base + off_act);
  d.nbucketed = reinterpret_cast<unsigned long long *>(base + off_act + 8);
  d.scal = reinterpret_cast<double *>(base + off_scal);
  HIP_TRY(hipMemsetAsync(d.cnt, 0, (size_t)n_r * 4, h->stream));
  HIP_TRY(hipMemsetAsync(d.active, 0, 16, h->stream));
  HIP_TRY(hipMemcpyAsync(d.active, &n_traj, 4, hipMemcpyHostToDevice, h->stream));
  init_kernel<<<n_traj, BS, 0, h->stream>>>(d, d_tr.as<mqct_traj_t>(), d_sp.as<int>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  const int rgroups = (maxnb + 63) / 64;
  // capture ITER iterations in one graph
  const int ITER = 16;
  const bool prof = ds->prof_on;
  int active = n_traj;
  // every trajectory ends after at most MAXSTP accepted steps and a bounded
  // number of step-size reductions; the cap only guards the host loop
  const int max_rounds = 1 << 17;
  HIP_TRY(hipEventRecord(h->ev[5], h->stream));
  if (!prof) {
    // ITER iterations (prep, GEMM, finish) captured once, replayed until
    // every trajectory has ended
    struct GraphGuard {
      hipGraph_t g = nullptr;
      hipGraphExec_t e = nullptr;
      ~GraphGuard() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
      }
    } gg;
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < ITER; it++) {
      prep_kernel<<<n_traj, BS, 0, h->stream>>>(d);
      gemm_kernel<<<dim3(n_r, nblk, rgroups), BS, 0, h->stream>>>(d);
      finish_kernel<<<n_traj, BS, 0, h->stream>>>(d);
    }
    hipError_t cap = hipStreamEndCapture(h->stream, &gg.g);
    HIP_TRY(cap);
    HIP_TRY(hipGraphInstantiate(&gg.e, gg.g, nullptr, nullptr, 0));
    for (int rep = 0; rep < max_rounds && active > 0; rep++) {
      HIP_TRY(hipGraphLaunch(gg.e, h->stream));
      HIP_TRY(hipMemcpyAsync(&active, d.active, 4, hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
    }
  } else {
    // the same iterations launched one by one with HIP events on this
    // stream around every GEMM launch (kernel timing)
    if (ds->prof_ev.empty()) {
      ds->prof_ev.resize(2 * ITER);
      for (auto &e : ds->prof_ev) HIP_TRY(hipEventCreate(&e));
    }
    for (int rep = 0; rep < max_rounds && active > 0; rep++) {
      for (int it = 0; it < ITER; it++) {
        prep_kernel<<<n_traj, BS, 0, h->stream>>>(d);
        HIP_TRY(hipEventRecord(ds->prof_ev[2 * it], h->stream));
        gemm_kernel<<<dim3(n_r, nblk, rgroups), BS, 0, h->stream>>>(d);
        HIP_TRY(hipEventRecord(ds->prof_ev[2 * it + 1], h->stream));
        finish_kernel<<<n_traj, BS, 0, h->stream>>>(d);
      }
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(&active, d.active, 4, hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
      for (int it = 0; it < ITER; it++) {
        float ems = 0.f;
        HIP_TRY(hipEventElapsedTime(&ems, ds->prof_ev[2 * it], ds->prof_ev[2 * it + 1]));
        ds->prof_ms += ems;
      }
      ds->prof_launches += ITER;
    }
  }
  HIP_TRY(hipEventRecord(h->ev[6], h->stream));
  HIP_TRY(hipEventSynchronize(h->ev[6]));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev[5], h->ev[6]));
  ds->last_ms = ms;
  if (active > 0) return mqct_fail(MQCT_ERR_HIP, "batched: trajectories did not finish");
  if (prof) {
    unsigned long long nbk = 0;
    HIP_TRY(hipMemcpy(&nbk, d.nbucketed, 8, hipMemcpyDeviceToHost));
    ds->prof_flops += 16.0 * (double)S2 * (double)nbk;
  }
  if (ds->d_q.ensure((size_t)n_traj * (n2 + 8) * 8) || ds->d_stat.ensure((size_t)n_traj * sizeof(mqct_traj_stat_t)))
    return mqct_fail(MQCT_ERR_NOMEM, "batched: output");
  out_kernel<<<n_traj, BS, 0, h->stream>>>(d, d_ps.as<int>(), ds->d_q.as<double>(),
                                           ds->d_stat.as<mqct_traj_stat_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (q_out) HIP_TRY(hipMemcpy(q_out, ds->d_q.p, (size_t)n_traj * (n2 + 8) * 8, hipMemcpyDeviceToHost));
  if (stat_out)
    HIP_TRY(hipMemcpy(stat_out, ds->d_stat.p, n_traj * sizeof(mqct_traj_stat_t), hipMemcpyDeviceToHost));
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_profile(mqct_hip_t *h, int on) {
  if (!h || !h->dyn) return mqct_fail(MQCT_ERR_STATE, "batched_profile: no matrix loaded");
  DynState *ds = h->dyn;
  ds->prof_on = on != 0;
  ds->prof_launches = 0;
  ds->prof_ms = 0.0;
  ds->prof_flops = 0.0;
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_kernel_stats(mqct_hip_t *h, int64_t *launches, double *gemm_ms,
                                             double *gemm_flops) {
  if (!h || !h->dyn || !launches || !gemm_ms || !gemm_flops)
    return mqct_fail(MQCT_ERR_ARG, "batched_kernel_stats: bad argument");
  *launches = h->dyn->prof_launches;
  *gemm_ms = h->dyn->prof_ms;
  *gemm_flops = h->dyn->prof_flops;
  return MQCT_OK;
}
