// batched.hip -- the interval-bucketed batched form of the AT-MQCT
// propagation (DERIVS_BK_adia, mqct_eqtns.f:720-879, integrated by odeint /
// rkqs / rkck or rk4 as in propagator.f:3873-3908).
//
// All running trajectories advance together, one right-hand-side
// evaluation each per iteration, every trajectory at its own stage of its
// own step.  An iteration is two launches:
//   gemm : for every knot r and m-block b, one dense product over the
//          knot's bucket: W = [Y_r | Y2_r] [w_Y v ; w_Y2 v] on FP64 MFMA,
//          where Y_r (Mat_el at knot r) and Y2_r (its spline second
//          derivatives) are the block's dense matrices and w_Y, w_Y2 the
//          trajectory's spline weights for that knot (splint_bk,
//          mqct_eqtns.f:575-607).  Each knot's matrices are read once per
//          iteration for all trajectories that need them.
//   step : per trajectory, completes the stage, k = -i exp(iEt) (W_lo +
//          W_hi) + state-diagonal Coriolis (mqct_eqtns.f:845-877), runs the
//          integrator's state machine (Cash-Karp stages, error control and
//          step-size rules of rkqs, step loop and end condition of odeint;
//          or classical rk4), then prepares the next stage: its time, the
//          path values (R, Theta, P_Phi) at that time, the knots and spline
//          weights of R, the stage vector, its rotated amplitudes v =
//          exp(-iEt) y_stage (block order) and its Coriolis part, and
//          appends the trajectory to the buckets of its two knots.
// Buckets are double-buffered by iteration parity.  The iterations are
// captured once in a HIP graph and replayed; the host reads the number of
// running trajectories between replays.  The knot matrices are kept in
// FP64 or, as the FP32 coefficient form, in FP32 (converted to FP64 in the
// MFMA operand).
//
// The CC modes (DERIVS_BK, mqct_eqtns.f:112-499) use the same GEMM with
// two more entries per knot whose weights are the spline's R-derivatives,
// so each trajectory gets M v and dM/dR v; their own step kernel forms the
// mean-field force and the Coriolis sums from them, integrates the
// classical variables beside the amplitudes and runs DOTRAJECT's dt
// intervals with their end checks (propagator.f:3873-4050).
#include "dyn_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
constexpr int BS = 256;
constexpr int NSLOT = 6;       // y, k1..k5 (k6 and rk4's k4 stay in registers)
constexpr int MAXSTP = 10000;  // odeint (suppl_routines.f:1362)
constexpr int MAXPT = 16;      // states per thread in the step kernel (N <= 4096)
static_assert(MAXPT * BS * 16 >= 64 * 1024, "N allowed by the LDS limit must fit MAXPT per thread");

struct Ctl {
  double x, x1, x2, h, hnext, hmin, eps;
  double tau, aa, bb, c1, c2;  // stage being evaluated: time, spline weights
  double dphidt;               // ... and dPhi/dt of the path there
  double h0;                   // odeint: h when the step began (error scale)
  double norm_err;
  int stage, nstp, status, zero;  // status: 0 running, 1 done
  int klo, path, nrhs, rk4;       // rk4: the stages are rk4's (AT rk4, or odeint's last step)
  int last, rec;                  // last: rk4 step that ends the path; rec: path-record hint
  // the bounds of the current path-record and knot intervals, so a stage
  // inside them finds its intervals without loads (zero: not yet known)
  double rec_t0, rec_t1, knot_r0, knot_r1;
};

// A bucket entry: trajectory, product slot (0: the knot below R, 1: above;
// CC also 2, 3: the R-derivative weights at those knots) and its two
// spline weights for that knot, so the GEMM reads one record per entry.
struct Entry {
  int t, side;
  double wy, wy2;
};

// CC-MQCT (DERIVS_BK, mqct_eqtns.f:205-500; DOTRAJECT's dt intervals,
// propagator.f:3873-4050): per trajectory, the classical part of the stage
// being evaluated and the interval loop's state.  The classical parts of y
// and of the k slots follow the 2N amplitudes in each slot.
struct CtlCC {
  double xc[8];           // stage's R, P_R, Theta, P_Theta, Phi, P_Phi, 0, 0
  double tcur;            // start of the current dt interval
  double dt_step;         // odeint: first trial step of the next interval
  double tmp_R1, tmp_R2;  // R at the ends of the last two intervals
  int vib, counter, nsteps, fail;
};

struct BDev {
  int N, nblk, n_r, n_traj, mode;
  int sw;                     // product vectors per trajectory: 2 (AT), 4 (CC: also dM/dR v)
  int bcap;                   // bucket entries per knot: n_traj (AT), 2 n_traj (CC)
  int theta_on;               // CC: the Theta equations are integrated
  const mqct_traj_t *traj;    // CC: the trajectories' end conditions
  CtlCC *cc;                  // CC: classical state per trajectory
  const double *R;            // knots; R[-1] holds the reduced mass
  const void *KM;             // [n_r][2][S2] dense block matrices Y, Y2 (double or float)
  long long S2;
  const int *blk_start;       // [nblk+1] positions
  const long long *blk_off;   // [nblk] offset of the block in S2
  const void *KT;             // [n_r][KS] 16-row tiles of [Y | Y2] in fragment order (double or float)
  long long KS;               // elements per knot in KT
  int ntile;                  // 16-row tiles over all blocks
  const int *tile_blk;        // [ntile] block of the tile
  const int *tile_r0;         // [ntile] first row of the tile in its block
  const int *blk_tile0;       // [nblk] first tile of the block
  const int *blk_order;       // [npg] row groups of the pipelined forms by decreasing cost
  const int *pg_blk;          // [npg] block of the row group (at most 128 rows of one block)
  const int *pg_t0;           // [npg] its first 16-row tile inside the block
  int npg;
  int *qctr;                  // item counter of the dynamically scheduled GEMM
  const long long *tile_off;  // [ntile] offset of the tile in a knot's KT
  int nrg;                    // 64-row groups over all blocks
  const int *rg_blk;          // [nrg] block of the row group
  const int *rg_row0;         // [nrg] first row of the group in its block
  const double *E;            // per position, Hartree
  const int2 *nb;             // per position: positions of state-1 / state+1, -1 none
  const int *m_pos;           // per position: m12 (CC Coriolis with Theta)
  const double2 *sq;          // per position: their Coriolis factors
  const double *prec, *pder;  // paths
  const long long *poff;
  double *S;                  // [n_traj][NSLOT][2N] (CC: [n_traj][NSLOT][2N + 8])
  double2 *V;                 // [n_traj][N]
  double2 *W;                 // [n_traj][sw][N] (float2 [n_traj][sw][N] after FP32 products)
  Ctl *ctl;
  int *cnt;                   // [2][n_r]
  Entry *bucket;              // [2][n_r][bcap]
  int *active;                // running trajectories
  int *alist;                 // [n_list] the trajectories the step kernel's workgroups take (compacted launches)
  int *klist;                 // [2][n_r] the knots with bucket entries, by bucket parity
  int *kcnt;                  // [2] their counts
  int *alist_tmp;             // [n_traj] compaction space
  unsigned long long *nbucketed;  // right-hand sides with R inside the grid
  unsigned long long *ntouched;   // (GEMM launch, knot) pairs with a non-empty bucket
};

// Cash-Karp tableau (suppl_routines.f:1306-1321); A4 = REAL*4 .6
__constant__ double CK_A[6] = {0.0, .2, .3, (double)0.6f, 1.0, .875};
__constant__ double CK_B[6][5] = {
    {0, 0, 0, 0, 0},
    {.2, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
    {.3, -.9, 1.2, 0, 0},
    {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};

template <int TB = BS>
__device__ inline double block_sum(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < TB / 64; i++) s += red[i];
  __syncthreads();
  return s;
}

template <int TB = BS>
__device__ inline double block_max(double x, double *red) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < TB / 64; i++) s = fmax(s, red[i]);
  __syncthreads();
  return s;
}

// Interval of a sorted table containing x, as splint's bisection finds it
// (suppl_routines.f:968-1002): the largest lo <= n-2 with xa[lo] <= x (0 if
// x < xa[0]).  Walks from a hint, which gives the same interval.
__device__ inline int interval(const double *xa, int stride, int n, double x, int hint) {
  int lo = hint < 0 ? 0 : (hint > n - 2 ? n - 2 : hint);
  while (lo > 0 && xa[(long long)lo * stride] > x) lo--;
  while (lo < n - 2 && xa[(long long)(lo + 1) * stride] <= x) lo++;
  return lo;
}

// The same with the bounds of the hinted interval cached in (b0, b1): an x
// with b0 <= x < b1 lies in it without a load.  The bounds kept are those
// the interval's conditions test (-inf for lo = 0, +inf for lo = n-2), so
// the cached answer is the one the walk would give.
__device__ inline int interval_cached(const double *xa, int stride, int n, double x, int &hint, double &b0,
                                      double &b1) {
  if (b0 <= x && x < b1) return hint;
  const int lo = interval(xa, stride, n, x, hint);
  hint = lo;
  b0 = lo == 0 ? -INFINITY : xa[(long long)lo * stride];
  b1 = lo == n - 2 ? INFINITY : xa[(long long)(lo + 1) * stride];
  return lo;
}

// Run by one workgroup of the GEMM with bucket parity par: clears the bucket
// counts that the following step kernel (parity par ^ 1) appends to.
__device__ inline void clear_next_counts(const BDev &d, int par) {
  for (int i = threadIdx.x; i < d.n_r; i += blockDim.x) d.cnt[(par ^ 1) * d.n_r + i] = 0;
  if (threadIdx.x == 0) d.kcnt[par ^ 1] = 0;
}

// The time of stage c.stage (Cash-Karp abscissae for odeint; x, x+h/2,
// x+h/2, x+h for rk4).
__device__ inline double stage_time(const Ctl &c) {
  if (!c.rk4) return c.x + CK_A[c.stage] * c.h;
  const int s = c.stage;
  return (s == 0) ? c.x : (s == 3 ? c.x + c.h : c.x + 0.5 * c.h);
}

// Prepares stage c.stage from its vector x (the thread's states in
// registers: position i = tid + j BS, Re at i, Im at i + N; Ej their
// energies): writes v = exp(-i E tau) x to V; thread 0 evaluates the
// recorded path at tau
// (dPhi/dt for the Coriolis term), finds the knots and spline weights,
// records them in its c and appends the trajectory to the buckets of
// parity `par`.
// Thread 0's part of preparing stage c.stage: evaluates the recorded path at
// the stage time (dPhi/dt for the Coriolis term), finds the knots and spline
// weights, records them in c and appends the trajectory to the buckets of
// parity `par`.
__device__ void stage_chain(const BDev &d, int t, Ctl &c, int par) {
  const double tau = stage_time(c);
  // recorded path at tau: R, Theta, P_Phi (DERIVS_BK_adia, mqct_eqtns.f:748-759)
  const long long o0 = d.poff[c.path];
  const int n = (int)(d.poff[c.path + 1] - o0);
  const double *rec = d.prec + o0 * 8;
  const double *der = d.pder + o0 * 7;
  const int lo = interval_cached(rec, 8, n, tau, c.rec, c.rec_t0, c.rec_t1);
  const int hi = lo + 1;
  const double hr = rec[hi * 8] - rec[lo * 8];
  const double a = (rec[hi * 8] - tau) / hr, b = (tau - rec[lo * 8]) / hr;
  const double ca = a * a * a - a, cb = b * b * b - b;
  double pv[3];
  const int col[3] = {0, 2, 5};
  for (int q = 0; q < 3; q++)
    pv[q] = a * rec[lo * 8 + 1 + col[q]] + b * rec[hi * 8 + 1 + col[q]] +
            (ca * der[lo * 7 + col[q]] + cb * der[hi * 7 + col[q]]) * (hr * hr) / 6.0;
  const double R = pv[0], st = sin(pv[1]);
  c.dphidt = pv[2] / R / R / d.R[-1] / (st * st);
  // knots of R (splint_bk, mqct_eqtns.f:575-607); R > R_n: no coupling
  const int nr = d.n_r;
  c.tau = tau;
  c.zero = R > d.R[nr - 1];
  const int klo = interval_cached(d.R, 1, nr, R, c.klo, c.knot_r0, c.knot_r1), khi = klo + 1;
  const double hk = d.R[khi] - d.R[klo];
  const double aa = (d.R[khi] - R) / hk, bb = (R - d.R[klo]) / hk;
  const double cc5 = (hk * hk) / 6.0;
  c.aa = aa;
  c.bb = bb;
  c.c1 = (aa * aa * aa - aa) * cc5;
  c.c2 = (bb * bb * bb - bb) * cc5;
  if (!c.zero) {
    if (d.nbucketed) atomicAdd(d.nbucketed, 1ULL);  // counted only in profiled runs
    int *cnt = d.cnt + par * nr;
    Entry *bk = d.bucket + (long long)par * nr * d.bcap;
    const int p0 = atomicAdd(&cnt[klo], 1), p1 = atomicAdd(&cnt[khi], 1);
    bk[(long long)klo * d.bcap + p0] = Entry{t, 0, aa, c.c1};
    bk[(long long)khi * d.bcap + p1] = Entry{t, 1, bb, c.c2};
    // a knot's first entry lists the knot for the knot-list GEMM form
    if (p0 == 0) d.klist[par * nr + atomicAdd(&d.kcnt[par], 1)] = klo;
    if (p1 == 0) d.klist[par * nr + atomicAdd(&d.kcnt[par], 1)] = khi;
  }
}

// The fields stage_chain sets in a control record.
__device__ inline void copy_chain(Ctl &dst, const Ctl &src) {
  dst.tau = src.tau;
  dst.dphidt = src.dphidt;
  dst.zero = src.zero;
  dst.aa = src.aa;
  dst.bb = src.bb;
  dst.c1 = src.c1;
  dst.c2 = src.c2;
  dst.klo = src.klo;
  dst.rec = src.rec;
  dst.rec_t0 = src.rec_t0;
  dst.rec_t1 = src.rec_t1;
  dst.knot_r0 = src.knot_r0;
  dst.knot_r1 = src.knot_r1;
}

// Prepares stage c.stage from its vector x (the thread's states in
// registers: position i = sid + j TS, Re at i, Im at i + N; Ej their
// energies): writes v = exp(-i E tau) x to V, and thread 0 runs stage_chain
// unless the control wave has run it already (chain false).
template <int PT, int TS>
__device__ void prepare_stage(const BDev &d, int t, Ctl &c, const double *xr, const double *xi, int par,
                              const double *Ej, int sid, bool chain) {
  const int N = d.N;
  const double tau = stage_time(c);
  double2 *v = d.V + (long long)t * N;
#pragma unroll
  for (int j = 0; j < PT; j++) {
    const int i = sid + j * TS;
    if (i < N) {
      double sn, cs;
      sincos(tau * Ej[j], &sn, &cs);
      v[i] = make_double2(xr[j] * cs + xi[j] * sn, xi[j] * cs - xr[j] * sn);
    }
  }
  if (chain && threadIdx.x == 0) stage_chain(d, t, c, par);
}

// Interval GEMM.  One workgroup per (knot r, m-block b, row group):
//   W[t][side][rows] = [Y_r | Y2_r] [w_Y v_t ; w_Y2 v_t]
// for every bucket entry (t, side) of knot r, GE entries (2 GE real
// columns: Re and Im of each v) per column super-tile.  Each wave owns GRT
// 16-row tiles with 2 GE / 16 column tiles of f64 16x16x4 MFMA
// accumulators each; the weighted vectors are staged through LDS in slices
// of GK rows of K = 2 nb, the matrix fragments come straight from HBM/L2
// (the blocks are symmetric, so lanes of one k read 16 consecutive rows).
// Block (0,0,0) also clears the other parity's bucket counts for the step
// kernel that follows.  The tile shape is a template parameter; the runner
// times the shapes on the first iterations and keeps the fastest.
// Shared memory of one LDS-staged super-tile.
template <int GE, int GK>
struct LdsTile {
  double Bs[GK][2 * GE + 16];  // row stride 2 GE + 16 doubles: half-waves hit disjoint banks
  int t[GE], side[GE];
  double wy[GE], wy2[GE];
};

// One super-tile of the LDS-staged product (whole workgroup): rows
// row0 .. row0 + 64 GRT of block b at knot r times bucket entries
// e0 .. e0 + nE (nE <= GE) of parity `par`.
template <typename KT, int GE, int GK, int GRT>
__device__ void lds_super_tile(const BDev &d, int par, int r, int b, int row0, int e0, int nE,
                               LdsTile<GE, GK> &sh) {
  constexpr int GC = 2 * GE / 16;  // MFMA column tiles
  const int p0 = d.blk_start[b], nb = d.blk_start[b + 1] - p0;
  const int K = 2 * nb;
  const KT *Y = static_cast<const KT *>(d.KM) + ((long long)r * 2) * d.S2 + d.blk_off[b];
  const KT *Y2 = static_cast<const KT *>(d.KM) + ((long long)r * 2 + 1) * d.S2 + d.blk_off[b];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kq = lane >> 4;
  int rr[GRT];
  bool tile_on[GRT];
#pragma unroll
  for (int q = 0; q < GRT; q++) {
    const int rt0 = row0 + (wave * GRT + q) * 16;
    tile_on[q] = rt0 < nb;
    rr[q] = rt0 + (lane & 15);
  }
  const Entry *bk = d.bucket + ((long long)par * d.n_r + r) * d.bcap;
  const int nct = (2 * nE + 15) / 16;
  const int nEpad = nct * 8;  // entries staged (zero columns up to the tile edge)
  __syncthreads();
  if (tid < GE) {
    int t = 0, side = 0;
    double wy = 0.0, wy2 = 0.0;
    if (tid < nE) {
      const Entry ent = bk[e0 + tid];
      t = ent.t;
      side = ent.side;
      wy = ent.wy;
      wy2 = ent.wy2;
    }
    sh.t[tid] = t;
    sh.side[tid] = side;
    sh.wy[tid] = wy;
    sh.wy2[tid] = wy2;
  }
  dbl4 acc[GRT][GC];
#pragma unroll
  for (int q = 0; q < GRT; q++)
#pragma unroll
    for (int ct = 0; ct < GC; ct++) acc[q][ct] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += GK) {
    __syncthreads();
    // kk fastest: a group of threads reads 16 consecutive amplitudes
    for (int idx = tid; idx < GK * nEpad; idx += BS) {
      const int kk = idx % GK, e = idx / GK, k = k0 + kk;
      double vx = 0.0, vy = 0.0;
      if (e < nE && k < K) {
        const bool lo = k < nb;
        const double w = lo ? sh.wy[e] : sh.wy2[e];
        const double2 x = d.V[(long long)sh.t[e] * d.N + p0 + (lo ? k : k - nb)];
        vx = w * x.x;
        vy = w * x.y;
      }
      sh.Bs[kk][2 * e] = vx;
      sh.Bs[kk][2 * e + 1] = vy;
    }
    double af[GRT][GK / 4];
#pragma unroll
    for (int q = 0; q < GRT; q++)
#pragma unroll
      for (int s4 = 0; s4 < GK / 4; s4++) {
        const int k = k0 + 4 * s4 + kq;
        double a = 0.0;
        if (rr[q] < nb && k < K)
          a = (k < nb) ? (double)Y[(long long)k * nb + rr[q]] : (double)Y2[(long long)(k - nb) * nb + rr[q]];
        af[q][s4] = a;
      }
    __syncthreads();
#pragma unroll
    for (int s4 = 0; s4 < GK / 4; s4++) {
      const double *brow = &sh.Bs[4 * s4 + kq][lane & 15];
#pragma unroll
      for (int q = 0; q < GRT; q++) {
        if (!tile_on[q]) continue;
#pragma unroll
        for (int ct = 0; ct < GC; ct++)
          if (ct < nct) acc[q][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[q][s4], brow[16 * ct], acc[q][ct], 0, 0, 0);
      }
    }
  }
  // f64 16x16x4 output: row lane/16 + 4i, column lane%16 of each tile
#pragma unroll
  for (int q = 0; q < GRT; q++) {
    if (!tile_on[q]) continue;
    const int rt0 = row0 + (wave * GRT + q) * 16;
#pragma unroll
    for (int ct = 0; ct < GC; ct++) {
      if (ct >= nct) break;
      const int col = 16 * ct + (lane & 15), e = col >> 1;
      if (e >= nE) continue;
      double *wb = reinterpret_cast<double *>(d.W + ((long long)sh.t[e] * d.sw + sh.side[e]) * d.N + p0) + (col & 1);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int orow = rt0 + (lane >> 4) + 4 * i;
        if (orow < nb) wb[2 * orow] = acc[q][ct][i];
      }
    }
  }
}

template <typename KT, int GE, int GK, int GRT>
__global__ __launch_bounds__(BS) void gemm_kernel(BDev d, int par) {
  constexpr int GROWS = 4 * GRT * 16;  // rows per workgroup
  const int r = blockIdx.x, b = blockIdx.y;
  if (r == 0 && b == 0 && blockIdx.z == 0) clear_next_counts(d, par);
  const int ne = d.cnt[par * d.n_r + r];
  if (ne == 0) return;
  const int nb = d.blk_start[b + 1] - d.blk_start[b];
  const int row0 = blockIdx.z * GROWS;
  if (row0 >= nb) return;
  if (b == 0 && blockIdx.z == 0 && threadIdx.x == 0) atomicAdd(d.ntouched, 1ULL);
  __shared__ LdsTile<GE, GK> sh;
  for (int e0 = 0; e0 < ne; e0 += GE) lds_super_tile<KT, GE, GK, GRT>(d, par, r, b, row0, e0, min(GE, ne - e0), sh);
}

// The same product without LDS or barriers: every wave owns one 16-row
// tile and 8 NCT bucket entries (NCT column tiles) and streams its operands
// straight into registers -- the matrix fragment (16 consecutive rows of one
// k per lane group) and the weighted amplitude component its lane needs --
// so waves run independently and occupancy hides the load latency.  Waves
// loop over further groups of 8 NCT entries.
template <typename KT, int NCT>
__global__ __launch_bounds__(BS) void gemm_direct_kernel(BDev d, int par) {
  const int r = blockIdx.x, b = blockIdx.y;
  if (r == 0 && b == 0 && blockIdx.z == 0) clear_next_counts(d, par);
  const int ne = d.cnt[par * d.n_r + r];
  if (ne == 0) return;
  const int p0 = d.blk_start[b], nb = d.blk_start[b + 1] - p0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rt0 = (blockIdx.z * 4 + wave) * 16;
  if (rt0 >= nb) return;
  if (b == 0 && blockIdx.z == 0 && threadIdx.x == 0) atomicAdd(d.ntouched, 1ULL);
  const int K = 2 * nb, kq = lane >> 4, col = lane & 15, comp = col & 1;
  const int rr = rt0 + (lane & 15);
  const KT *Y = static_cast<const KT *>(d.KM) + ((long long)r * 2) * d.S2 + d.blk_off[b];
  const KT *Y2 = static_cast<const KT *>(d.KM) + ((long long)r * 2 + 1) * d.S2 + d.blk_off[b];
  const Entry *bk = d.bucket + ((long long)par * d.n_r + r) * d.bcap;
  for (int e0 = 0; e0 < ne; e0 += 8 * NCT) {
    const double *vp[NCT];
    double wy[NCT], wy2[NCT];
    int tt[NCT], sd[NCT];
    bool ok[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const int e = e0 + ct * 8 + (col >> 1);
      ok[ct] = e < ne;
      const Entry ent = ok[ct] ? bk[e] : Entry{0, 0, 0.0, 0.0};
      tt[ct] = ent.t;
      sd[ct] = ent.side;
      wy[ct] = ent.wy;
      wy2[ct] = ent.wy2;
      vp[ct] = reinterpret_cast<const double *>(d.V + (long long)ent.t * d.N + p0) + comp;
    }
    dbl4 acc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k0 = 0; k0 < K; k0 += 4) {
      const int k = k0 + kq;
      const bool lo = k < nb, in = k < K;
      const int kk = lo ? k : k - nb;
      double a = 0.0;
      if (rr < nb && in) a = lo ? (double)Y[(long long)k * nb + rr] : (double)Y2[(long long)kk * nb + rr];
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) {
        const double bv = in ? (lo ? wy[ct] : wy2[ct]) * vp[ct][2 * kk] : 0.0;
        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[ct], 0, 0, 0);
      }
    }
    // f64 16x16x4 output: row lane/16 + 4i, column lane%16 of each tile
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      if (!ok[ct]) continue;
      double *wb = reinterpret_cast<double *>(d.W + ((long long)tt[ct] * d.sw + sd[ct]) * d.N + p0) + comp;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int orow = rt0 + (lane >> 4) + 4 * i;
        if (orow < nb) wb[2 * orow] = acc[ct][i];
      }
    }
  }
}

// 16-byte fragment loads of the tile layout: U consecutive MFMA steps of
// one lane
template <typename KT>
struct Frag;
template <>
struct Frag<double> {
  using V = double2;
  static constexpr int U = 2;
  __device__ static double get(const V &v, int u) { return u == 0 ? v.x : v.y; }
};
template <>
struct Frag<float> {
  using V = float4;
  static constexpr int U = 4;
  __device__ static double get(const V &v, int u) {
    return (double)(u == 0 ? v.x : (u == 1 ? v.y : (u == 2 ? v.z : v.w)));
  }
  __device__ static float getf(const V &v, int u) { return u == 0 ? v.x : (u == 1 ? v.y : (u == 2 ? v.z : v.w)); }
};

// The same product over the fragment-ordered tiles: each wave owns one
// 16-row tile of one block at knot r and groups of 8 NCT bucket entries.  A
// lane's 16-byte load holds its matrix fragments for U consecutive MFMA steps
// (U = 2 doubles or 4 floats), so one load instruction moves 1 KB per wave
// with no LDS and no barriers.  Inside a load the contraction order is
// permuted (lane group q holds k = 4U s + U q + u for step u) and the
// weighted amplitudes follow the same order.  Tile rows past the block and
// columns past nb are zero; amplitude reads clamp k to the block.
// KLIST: blockIdx.x indexes the list of knots with entries (the grid's x
// extent bounds the list: two knots per running trajectory), not the knots,
// so few running trajectories launch few workgroups.
template <typename KT, int NCT, bool KLIST = false>
__global__ __launch_bounds__(BS) void gemm_stream_kernel(BDev d, int par) {
  using F = Frag<KT>;
  using FV = typename F::V;
  constexpr int U = F::U, KB = 4 * U;
  if (blockIdx.x == 0 && blockIdx.y == 0) clear_next_counts(d, par);
  int r = blockIdx.x;
  if (KLIST) {
    if ((int)blockIdx.x >= __builtin_amdgcn_readfirstlane(d.kcnt[par])) return;
    r = __builtin_amdgcn_readfirstlane(d.klist[par * d.n_r + blockIdx.x]);
  }
  const int ne = d.cnt[par * d.n_r + r];
  if (ne == 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = blockIdx.y * 4 + wave;
  if (g >= d.ntile) return;
  if (g == 0 && lane == 0) atomicAdd(d.ntouched, 1ULL);
  const int b = d.tile_blk[g], rt0 = d.tile_r0[g];
  const int p0 = d.blk_start[b], nb = d.blk_start[b + 1] - p0;
  const int ns = (nb + KB - 1) / KB;  // loads per half ([Y | Y2])
  const int q = lane >> 4, col = lane & 15, comp = col & 1;
  const FV *A = reinterpret_cast<const FV *>(static_cast<const KT *>(d.KT) + (long long)r * d.KS + d.tile_off[g]) +
                lane;
  const Entry *bk = d.bucket + ((long long)par * d.n_r + r) * d.bcap;
  for (int e0 = 0; e0 < ne; e0 += 8 * NCT) {
    const double *vp[NCT];
    double wy[NCT], wy2[NCT];
    int tt[NCT], sd[NCT];
    bool ok[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const int e = e0 + ct * 8 + (col >> 1);
      ok[ct] = e < ne;
      const Entry ent = ok[ct] ? bk[e] : Entry{0, 0, 0.0, 0.0};
      tt[ct] = ent.t;
      sd[ct] = ent.side;
      wy[ct] = ent.wy;
      wy2[ct] = ent.wy2;
      vp[ct] = reinterpret_cast<const double *>(d.V + (long long)ent.t * d.N + p0) + comp;
    }
    dbl4 acc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int hf = 0; hf < 2; hf++) {
      double w[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) w[ct] = hf ? wy2[ct] : wy[ct];
      const FV *Ah = A + hf * ns * 64;
#pragma unroll 2
      for (int s = 0; s < ns; s++) {
        const FV av = Ah[s * 64];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int kc = min(KB * s + U * q + u, nb - 1);
          const double a = F::get(av, u);
#pragma unroll
          for (int ct = 0; ct < NCT; ct++)
            acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, w[ct] * vp[ct][2 * kc], acc[ct], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      if (!ok[ct]) continue;
      double *wb = reinterpret_cast<double *>(d.W + ((long long)tt[ct] * d.sw + sd[ct]) * d.N + p0) + comp;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int orow = rt0 + (lane >> 4) + 4 * i;
        if (orow < nb) wb[2 * orow] = acc[ct][i];
      }
    }
  }
}

// The fragment-ordered product as a work queue.  Trajectories bunch in R
// (all start at the same R), so a few knots can hold thousands of bucket
// entries while most hold none; here the entries of every knot are cut into
// chunks of 8 NCT and the work items (knot, group of four 16-row tiles,
// chunk), in that order, are shared out over a fixed grid of persistent
// workgroups.  Each XCD (workgroup index mod 8) takes a contiguous range of
// items, so the chunks of one tile group meet in the same L2.  Every
// workgroup forms the chunk prefix over the knots from the bucket counts in
// LDS (n_r + 1 ints).
constexpr int QGRID = 2048;  // persistent workgroups (a multiple of 8)
constexpr int DGRID = 1024;  // workgroups of the dynamically scheduled form (four per CU)

// The work-queue preamble (whole workgroup): workgroup 0 clears the other
// parity's bucket counts and counts the touched knots; every workgroup forms
// in pre[0..n_r] the exclusive prefix over knots of the number of chunks of
// CH bucket entries (a contiguous run of knots per thread, then a block scan
// of the run totals).  Returns the total number of chunks.
__device__ int chunk_prefix(const BDev &d, int par, int CH, int *pre, int *wtot) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nr = d.n_r;
  const int *cnt = d.cnt + par * nr;
  if (blockIdx.x == 0) clear_next_counts(d, par);
  const int per = (nr + BS - 1) / BS, ra = min(nr, tid * per), rb = min(nr, ra + per);
  int own = 0, busy = 0;
  for (int r = ra; r < rb; r++) {
    const int c = cnt[r];
    own += (c + CH - 1) / CH;
    busy += c > 0;
  }
  int inc = own;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wtot[wave] = inc;
  __syncthreads();
  int run = inc - own;
  for (int w = 0; w < wave; w++) run += wtot[w];
  for (int r = ra; r < rb; r++) {
    pre[r] = run;
    run += (cnt[r] + CH - 1) / CH;
  }
  if (tid == BS - 1) pre[nr] = run;
  if (blockIdx.x == 0 && busy) atomicAdd(d.ntouched, (unsigned long long)busy);
  __syncthreads();
  return pre[nr];
}

// The largest knot r with pre[r] <= q (pre non-decreasing, pre[0] = 0 <= q
// < pre[n_r]): the knot whose chunks hold chunk index q.
__device__ inline int chunk_knot(const int *pre, int nr, int q) {
  int a = 0, b = nr - 1;
  while (a < b) {
    const int m = (a + b + 1) >> 1;
    if (pre[m] <= q) a = m;
    else b = m - 1;
  }
  return a;
}

template <typename KT, int NCT>
__global__ __launch_bounds__(BS) void gemm_queue_kernel(BDev d, int par) {
  using F = Frag<KT>;
  using FV = typename F::V;
  constexpr int U = F::U, KB = 4 * U, CH = 8 * NCT;
  extern __shared__ int pre[];  // [n_r + 1] chunk prefix
  __shared__ int wtot[BS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nr = d.n_r, TG = (d.ntile + 3) / 4;
  const int *cnt = d.cnt + par * nr;
  chunk_prefix(d, par, CH, pre, wtot);
  const long long M = (long long)pre[nr] * TG;
  const int X = 8, xcd = blockIdx.x % X, slot = blockIdx.x / X, Gx = gridDim.x / X;
  const long long lo = M * xcd / X, hi = M * (xcd + 1) / X;
  const int q = lane >> 4, col = lane & 15, comp = col & 1;
  for (long long it = lo + slot; it < hi; it += Gx) {
    // item -> (knot r, tile group tg, chunk c): pre[r] <= it / TG < pre[r+1]
    const int r = chunk_knot(pre, nr, (int)(it / TG)), cr = pre[r + 1] - pre[r];
    const int loc = (int)(it - (long long)pre[r] * TG), tg = loc / cr, ch = loc % cr;
    const int g = tg * 4 + wave;
    if (g >= d.ntile) continue;
    const int ne = cnt[r], e0 = ch * CH;
    const int bl = d.tile_blk[g], rt0 = d.tile_r0[g];
    const int p0 = d.blk_start[bl], nb = d.blk_start[bl + 1] - p0;
    const int ns = (nb + KB - 1) / KB;
    const FV *A =
        reinterpret_cast<const FV *>(static_cast<const KT *>(d.KT) + (long long)r * d.KS + d.tile_off[g]) + lane;
    const Entry *bk = d.bucket + ((long long)par * nr + r) * d.bcap;
    const double *vp[NCT];
    double wy[NCT], wy2[NCT];
    int tt[NCT], sd[NCT];
    bool ok[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const int e = e0 + ct * 8 + (col >> 1);
      ok[ct] = e < ne;
      const Entry ent = ok[ct] ? bk[e] : Entry{0, 0, 0.0, 0.0};
      tt[ct] = ent.t;
      sd[ct] = ent.side;
      wy[ct] = ent.wy;
      wy2[ct] = ent.wy2;
      vp[ct] = reinterpret_cast<const double *>(d.V + (long long)ent.t * d.N + p0) + comp;
    }
    dbl4 acc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int hf = 0; hf < 2; hf++) {
      double w[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) w[ct] = hf ? wy2[ct] : wy[ct];
      const FV *Ah = A + hf * ns * 64;
#pragma unroll 2
      for (int s = 0; s < ns; s++) {
        const FV av = Ah[s * 64];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int kc = min(KB * s + U * q + u, nb - 1);
          const double av_u = F::get(av, u);
#pragma unroll
          for (int ct = 0; ct < NCT; ct++)
            acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av_u, w[ct] * vp[ct][2 * kc], acc[ct], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      if (!ok[ct]) continue;
      double *wb = reinterpret_cast<double *>(d.W + ((long long)tt[ct] * d.sw + sd[ct]) * d.N + p0) + comp;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int orow = rt0 + (lane >> 4) + 4 * i;
        if (orow < nb) wb[2 * orow] = acc[ct][i];
      }
    }
  }
}

// The LDS-staged product as a work queue: items (knot, group of 64 rows of
// one block, chunk of GE bucket entries), in that order, over persistent
// workgroups, each item one super-tile of the whole workgroup (the four
// waves share the staged amplitudes).
template <typename KT, int GE, int GK>
__global__ __launch_bounds__(BS) void gemm_lds_queue_kernel(BDev d, int par) {
  extern __shared__ int pre[];  // [n_r + 1] chunk prefix
  __shared__ int wtot[BS / 64];
  __shared__ LdsTile<GE, GK> sh;
  const int nr = d.n_r, RG = d.nrg;
  const int *cnt = d.cnt + par * nr;
  const long long M = (long long)chunk_prefix(d, par, GE, pre, wtot) * RG;
  const int X = 8, xcd = blockIdx.x % X, slot = blockIdx.x / X, Gx = gridDim.x / X;
  const long long lo = M * xcd / X, hi = M * (xcd + 1) / X;
  for (long long it = lo + slot; it < hi; it += Gx) {
    const int r = chunk_knot(pre, nr, (int)(it / RG)), cr = pre[r + 1] - pre[r];
    const int loc = (int)(it - (long long)pre[r] * RG), rg = loc / cr, ch = loc % cr;
    const int e0 = ch * GE;
    lds_super_tile<KT, GE, GK, 1>(d, par, r, d.rg_blk[rg], d.rg_row0[rg], e0, min(GE, cnt[r] - e0), sh);
  }
}

// The software-pipelined product: the whole workgroup takes an item (knot
// r, m-block b, chunk c of GE bucket entries).  Wave w owns the block's
// 16-row tiles w and w + 4 (up to 128 rows) and keeps their fragment-ordered
// [Y | Y2] operands in registers, one slice at a time; a slice is GK = 16
// (or 32) contraction rows of the Y half and the same rows of the Y2 half,
// so each amplitude of the chunk is loaded once and staged twice, weighted
// by w_Y and by w_Y2, into a double-buffered LDS slice.  The next slice's
// matrix fragments and amplitudes are loaded while the MFMAs of the current
// one run, so there is one barrier per slice and the loads overlap the math.
template <typename KT, int GE, int GKS = 16>
struct PipeShape {
  static constexpr int U = Frag<KT>::U, KB = 4 * U;
  static constexpr int GK = GKS, LPS = GK / KB;  // rows per slice and half, fragment loads per tile, half and slice
  static constexpr int GC = 2 * GE / 16;        // MFMA column tiles
  // LDS row stride (doubles): the rows lane groups q and q + 1 read at one
  // MFMA step are U apart; U BW = 16 (mod 32) puts them 32 banks apart
  static constexpr int BW = 2 * GE + (U == 4 ? 4 : 8);
  // the same with FP32 staging (FP32 products): lane groups q read rows
  // 4 apart, 4 BW32 = 16 (mod 64) puts them 16 banks apart
  static constexpr int BW32 = 2 * GE + 4;
  static constexpr int SPT = GK * GE / BS;      // (row, entry) amplitudes loaded per thread per slice
  static_assert(SPT * BS == GK * GE && LPS * KB == GK, "slice shape");
};

template <typename KT, int GE, int GKS = 16, bool M32 = false>
struct PipeShared {
  using BT = typename std::conditional<M32, float, double>::type;
  // rows 0 .. GK-1: w_Y v of the slice's rows; GK .. 2 GK-1: w_Y2 v
  BT Bs[2][2 * PipeShape<KT, GE, GKS>::GK][M32 ? PipeShape<KT, GE, GKS>::BW32 : PipeShape<KT, GE, GKS>::BW];
  int t[GE], side[GE];
  double wy[GE], wy2[GE];
};

// Blocks of more than 128 states are cut into row groups of up to 128
// rows (tiles tg .. tg + 7 of the block); every group of a block contracts
// over the whole block (K = 2 nb), so the group's items read the chunk's
// amplitudes once each.
template <typename KT, int GE, int GKS = 16, bool M32 = false>
__device__ void pipe_item(const BDev &d, int par, int r, int b, int tg, int c,
                          PipeShared<KT, GE, GKS, M32> &sh) {
  using F = Frag<KT>;
  using FV = typename F::V;
  using P = PipeShape<KT, GE, GKS>;
  // M32: FP32 products -- the FP32 knot fragments times the weighted
  // amplitudes rounded to FP32 on f32 16x16x4 MFMA with FP32 accumulation
  static_assert(!M32 || std::is_same<KT, float>::value, "FP32 products need FP32 knot matrices");
  using BT = typename PipeShared<KT, GE, GKS, M32>::BT;
  using AccT = typename std::conditional<M32, flt4, dbl4>::type;
  constexpr int U = P::U, KB = P::KB, GK = P::GK, LPS = P::LPS, GC = P::GC, SPT = P::SPT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane >> 4, col = lane & 15;
  const int nr = d.n_r;
  const int *cnt = d.cnt + par * nr;
  const int p0 = d.blk_start[b], nb = d.blk_start[b + 1] - p0;
  // each half ([Y | Y2]) of a tile's rows is nbp = nfh KB contraction rows;
  // the last slice may run past them (GK = 32 with KB = 16 or 8): its
  // fragments and amplitudes are zero there
  const int nbp = (nb + KB - 1) / KB * KB, nfh = nbp / KB, nsl = (nbp + GK - 1) / GK;
  const int e0 = c * GE, nE = min(GE, cnt[r] - e0), nct = (2 * nE + 15) / 16;
  const Entry *bk = d.bucket + ((long long)par * nr + r) * d.bcap;
  __syncthreads();  // the previous item is done with the entry table and Bs
  if (tid < GE) {
    int t = 0, side = 0;
    double wy = 0.0, wy2 = 0.0;
    if (tid < nE) {
      const Entry ent = bk[e0 + tid];
      t = ent.t;
      side = ent.side;
      wy = ent.wy;
      wy2 = ent.wy2;
    }
    sh.t[tid] = t;
    sh.side[tid] = side;
    sh.wy[tid] = wy;
    sh.wy2[tid] = wy2;
  }
  __syncthreads();
  // the wave's tiles (of the row group) and their fragment streams
  const int ntl = (nb + 15) / 16;
  bool on[2];
  const FV *A[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int tl = tg + wave + 4 * h;
    on[h] = tl < ntl;
    const int g = d.blk_tile0[b] + (on[h] ? tl : 0);
    A[h] = reinterpret_cast<const FV *>(static_cast<const KT *>(d.KT) + (long long)r * d.KS + d.tile_off[g]) +
           lane;
  }
  FV an[2][2][LPS];  // [tile][half][load]
  double2 bn[SPT];
  // loads of slice s into registers: fragments of the owned tiles for rows
  // s GK .. s GK + GK - 1 of both halves, and the chunk's amplitudes of those rows
  auto load = [&](int s) {
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int hf = 0; hf < 2; hf++)
#pragma unroll
        for (int l = 0; l < LPS; l++) {
          const int f = s * LPS + l;
          an[h][hf][l] = (on[h] && f < nfh) ? A[h][(hf * nfh + f) * 64] : FV{};
        }
#pragma unroll
    for (int j = 0; j < SPT; j++) {
      const int idx = tid + j * BS, kk = idx % GK, e = idx / GK, k = s * GK + kk;
      const bool ok = e < nE && k < nb;
      bn[j] = ok ? d.V[(long long)sh.t[e] * d.N + p0 + k] : make_double2(0.0, 0.0);
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int j = 0; j < SPT; j++) {
      const int idx = tid + j * BS, kk = idx % GK, e = idx / GK;
      const double wy = sh.wy[e], wy2 = sh.wy2[e];
      sh.Bs[buf][kk][2 * e] = (BT)(wy * bn[j].x);
      sh.Bs[buf][kk][2 * e + 1] = (BT)(wy * bn[j].y);
      sh.Bs[buf][GK + kk][2 * e] = (BT)(wy2 * bn[j].x);
      sh.Bs[buf][GK + kk][2 * e + 1] = (BT)(wy2 * bn[j].y);
    }
  };
  AccT acc[2][GC];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int ct = 0; ct < GC; ct++) acc[h][ct] = AccT{0, 0, 0, 0};
  load(0);
  stage(0);
  __syncthreads();
  for (int s = 0; s < nsl; s++) {
    const int buf = s & 1;
    FV ac[2][2][LPS];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int hf = 0; hf < 2; hf++)
#pragma unroll
        for (int l = 0; l < LPS; l++) ac[h][hf][l] = an[h][hf][l];
    const bool more = s + 1 < nsl;
    if (more) load(s + 1);
#pragma unroll
    for (int hf = 0; hf < 2; hf++)
#pragma unroll
      for (int l = 0; l < LPS; l++)
#pragma unroll
        for (int u = 0; u < U; u++) {
          // fragment order: lane group q holds row KB l + U q + u at step u
          const BT *brow = &sh.Bs[buf][hf * GK + KB * l + U * q + u][col];
          BT bf[GC];
#pragma unroll
          for (int ct = 0; ct < GC; ct++) bf[ct] = (ct < nct) ? brow[16 * ct] : (BT)0;
#pragma unroll
          for (int h = 0; h < 2; h++) {
            if (!on[h]) continue;
            if constexpr (M32) {
              const float a = F::getf(ac[h][hf][l], u);
#pragma unroll
              for (int ct = 0; ct < GC; ct++)
                if (ct < nct) acc[h][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bf[ct], acc[h][ct], 0, 0, 0);
            } else {
              const double a = F::get(ac[h][hf][l], u);
#pragma unroll
              for (int ct = 0; ct < GC; ct++)
                if (ct < nct) acc[h][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bf[ct], acc[h][ct], 0, 0, 0);
            }
          }
        }
    if (more) stage(buf ^ 1);
    __syncthreads();
  }
  // 16x16x4 output, column lane%16 of each tile: row lane/16 + 4i (f64),
  // 4 (lane/16) + i (f32)
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (!on[h]) continue;
    const int rt0 = (tg + wave + 4 * h) * 16;
#pragma unroll
    for (int ct = 0; ct < GC; ct++) {
      if (ct >= nct) break;
      const int cc = 16 * ct + col, e = cc >> 1;
      if (e >= nE) continue;
      const long long wo = ((long long)sh.t[e] * d.sw + sh.side[e]) * d.N + p0;
      // FP32 products are stored as they were accumulated, in FP32
      if constexpr (M32) {
        float *wb = reinterpret_cast<float *>(reinterpret_cast<float2 *>(d.W) + wo) + (cc & 1);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int orow = rt0 + 4 * q + i;
          if (orow < nb) wb[2 * orow] = acc[h][ct][i];
        }
      } else {
        double *wb = reinterpret_cast<double *>(d.W + wo) + (cc & 1);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int orow = rt0 + q + 4 * i;
          if (orow < nb) wb[2 * orow] = acc[h][ct][i];
        }
      }
    }
  }
}

// The pipelined product over a static share of the items: (knot, block,
// chunk) in that order, so the chunks of one block at one knot follow each
// other and meet its matrix in L2, with a contiguous item range per XCD.
template <typename KT, int GE, bool M32 = false>
__global__ __launch_bounds__(BS, GE >= 64 ? 1 : 2) void gemm_pipe_kernel(BDev d, int par) {
  extern __shared__ int pre[];  // [n_r + 1] chunk prefix
  __shared__ int wtot[BS / 64];
  __shared__ PipeShared<KT, GE, 16, M32> sh;
  const int nr = d.n_r, npg = d.npg;
  const long long M = (long long)chunk_prefix(d, par, GE, pre, wtot) * npg;
  const int X = 8, xcd = blockIdx.x % X, slot = blockIdx.x / X, Gx = gridDim.x / X;
  const long long lo = M * xcd / X, hi = M * (xcd + 1) / X;
  for (long long it = lo + slot; it < hi; it += Gx) {
    // item -> (knot r, row group g, chunk c)
    const int r = chunk_knot(pre, nr, (int)(it / npg)), cr = pre[r + 1] - pre[r];
    const int loc = (int)(it - (long long)pre[r] * npg), g = loc / cr, c = loc % cr;
    pipe_item<KT, GE, 16, M32>(d, par, r, d.pg_blk[g], d.pg_t0[g], c, sh);
  }
}

// The pipelined product with the items handed out one at a time from a
// device counter (cleared by the step kernel before every GEMM), the
// largest m-blocks first: the block rank in d.blk_order, then the chunk
// index over all knots.  Item costs differ by the square of the block size,
// so a static share leaves a long tail; taken largest first, the small
// items fill in at the end.  Each workgroup fetches its next item while it
// works on the current one.
template <typename KT, int GE, int GKS = 16, bool M32 = false>
__global__ __launch_bounds__(BS, GE >= 64 ? 1 : (GE >= 32 ? 2 : 3)) void gemm_pipe_dyn_kernel(BDev d, int par) {
  extern __shared__ int pre[];  // [n_r + 1] chunk prefix
  __shared__ int wtot[BS / 64];
  __shared__ PipeShared<KT, GE, GKS, M32> sh;
  __shared__ int s_next;
  const int tid = threadIdx.x, nr = d.n_r;
  const int C = chunk_prefix(d, par, GE, pre, wtot);
  const int M = C * d.npg;
  if (tid == 0) s_next = atomicAdd(d.qctr, 1);
  __syncthreads();
  int it = s_next;
  while (it < M) {
    int nxt = 0;
    if (tid == 0) nxt = atomicAdd(d.qctr, 1);
    const int g = d.blk_order[it / C], qc = it % C;
    const int r = chunk_knot(pre, nr, qc), c = qc - pre[r];
    pipe_item<KT, GE, GKS, M32>(d, par, r, d.pg_blk[g], d.pg_t0[g], c, sh);
    __syncthreads();  // everyone has read s_next
    if (tid == 0) s_next = nxt;
    __syncthreads();
    it = s_next;
  }
}

// The GEMM forms and tile shapes the runner chooses from: form, bucket
// entries per super-tile (per wave for the LDS-free forms), K rows per LDS
// slice, 16-row tiles per wave.
enum GemmForm { kLds = 0, kDirect = 1, kStream = 2, kQueue = 3, kLdsQueue = 4, kPipe = 5, kPipeDyn = 6, kStreamList = 7 };
struct GemmShape {
  int form, ge, gk, grt;
};
constexpr GemmShape kShapes[] = {
    {kLds, 64, 16, 1},      {kLds, 32, 16, 1},    {kLds, 32, 32, 1},    {kLds, 16, 32, 1},
    {kLds, 64, 16, 2},      {kDirect, 32, 0, 1},  {kDirect, 16, 0, 1},  {kDirect, 8, 0, 1},
    {kStream, 8, 0, 1},     {kStream, 16, 0, 1},  {kStream, 32, 0, 1},  {kQueue, 16, 0, 1},
    {kQueue, 32, 0, 1},     {kLdsQueue, 64, 16, 1}, {kLdsQueue, 32, 16, 1}, {kPipe, 64, 16, 2},
    {kPipe, 32, 16, 2},     {kPipeDyn, 64, 16, 2},  {kPipeDyn, 32, 16, 2},  {kPipeDyn, 16, 16, 2},
    {kPipeDyn, 32, 32, 2},  {kPipeDyn, 16, 32, 2},  {kStreamList, 8, 0, 1}};
constexpr int kNShapes = sizeof(kShapes) / sizeof(kShapes[0]);
// the work-queue forms keep the chunk prefix of all knots in LDS
constexpr bool uses_queue(int shape) {
  return kShapes[shape].form == kQueue || kShapes[shape].form == kLdsQueue || kShapes[shape].form == kPipe ||
         kShapes[shape].form == kPipeDyn;
}
// the software-pipelined forms (FP32 products with batched = 3)
constexpr bool pipelined(int shape) { return kShapes[shape].form == kPipe || kShapes[shape].form == kPipeDyn; }

template <typename KT>
void launch_gemm(int shape, dim3 knots_blocks, int maxnb, hipStream_t st, const BDev &d, int par, int kx,
                 bool m32 = false) {
  const size_t shq = (size_t)(d.n_r + 1) * 4;
  const dim3 gs(knots_blocks.x, (d.ntile + 3) / 4);
  // FP32 products: the pipelined forms on f32 MFMA (the other forms keep
  // FP64 products)
  if constexpr (std::is_same<KT, float>::value) {
    if (m32) switch (shape) {
        case 15: gemm_pipe_kernel<KT, 64, true><<<QGRID, BS, shq, st>>>(d, par); return;
        case 16: gemm_pipe_kernel<KT, 32, true><<<QGRID, BS, shq, st>>>(d, par); return;
        case 17: gemm_pipe_dyn_kernel<KT, 64, 16, true><<<DGRID, BS, shq, st>>>(d, par); return;
        case 18: gemm_pipe_dyn_kernel<KT, 32, 16, true><<<DGRID, BS, shq, st>>>(d, par); return;
        case 19: gemm_pipe_dyn_kernel<KT, 16, 16, true><<<DGRID, BS, shq, st>>>(d, par); return;
        case 20: gemm_pipe_dyn_kernel<KT, 32, 32, true><<<DGRID, BS, shq, st>>>(d, par); return;
        case 21: gemm_pipe_dyn_kernel<KT, 16, 32, true><<<DGRID, BS, shq, st>>>(d, par); return;
        default: break;
      }
  }
  switch (shape) {
    case 8: gemm_stream_kernel<KT, 1><<<gs, BS, 0, st>>>(d, par); return;
    case 22: gemm_stream_kernel<KT, 1, true><<<dim3(kx, gs.y), BS, 0, st>>>(d, par); return;
    case 9: gemm_stream_kernel<KT, 2><<<gs, BS, 0, st>>>(d, par); return;
    case 10: gemm_stream_kernel<KT, 4><<<gs, BS, 0, st>>>(d, par); return;
    case 11: gemm_queue_kernel<KT, 2><<<QGRID, BS, shq, st>>>(d, par); return;
    case 12: gemm_queue_kernel<KT, 4><<<QGRID, BS, shq, st>>>(d, par); return;
    case 13: gemm_lds_queue_kernel<KT, 64, 16><<<QGRID, BS, shq, st>>>(d, par); return;
    case 14: gemm_lds_queue_kernel<KT, 32, 16><<<QGRID, BS, shq, st>>>(d, par); return;
    case 15: gemm_pipe_kernel<KT, 64><<<QGRID, BS, shq, st>>>(d, par); return;
    case 16: gemm_pipe_kernel<KT, 32><<<QGRID, BS, shq, st>>>(d, par); return;
    case 17: gemm_pipe_dyn_kernel<KT, 64><<<DGRID, BS, shq, st>>>(d, par); return;
    case 18: gemm_pipe_dyn_kernel<KT, 32><<<DGRID, BS, shq, st>>>(d, par); return;
    case 19: gemm_pipe_dyn_kernel<KT, 16><<<DGRID, BS, shq, st>>>(d, par); return;
    case 20: gemm_pipe_dyn_kernel<KT, 32, 32><<<DGRID, BS, shq, st>>>(d, par); return;
    case 21: gemm_pipe_dyn_kernel<KT, 16, 32><<<DGRID, BS, shq, st>>>(d, par); return;
    default: break;
  }
  auto go = [&](auto kern, int grt) {
    const int rows = 4 * grt * 16;
    dim3 g(knots_blocks.x, knots_blocks.y, (maxnb + rows - 1) / rows);
    kern<<<g, BS, 0, st>>>(d, par);
  };
  switch (shape) {
    case 0: go(gemm_kernel<KT, 64, 16, 1>, 1); break;
    case 1: go(gemm_kernel<KT, 32, 16, 1>, 1); break;
    case 2: go(gemm_kernel<KT, 32, 32, 1>, 1); break;
    case 3: go(gemm_kernel<KT, 16, 32, 1>, 1); break;
    case 4: go(gemm_kernel<KT, 64, 16, 2>, 2); break;
    case 5: go(gemm_direct_kernel<KT, 4>, 1); break;
    case 6: go(gemm_direct_kernel<KT, 2>, 1); break;
    case 7: go(gemm_direct_kernel<KT, 1>, 1); break;
  }
}

// The control update after one evaluated stage: odeint/rkqs/rkck's step
// rules (suppl_routines.f:1229-1283, 1355-1391) and DOTRAJECT's AT rk4 loop
// (propagator.f:3862-3880).  em is rkqs's scaled error, read only at a
// Cash-Karp step's last stage with h above hmin.
__device__ inline void ctl_advance(Ctl &c, int mode, double em) {
  const double TINY = 1.e-30;
  const int s = c.stage;
  if (c.rk4) {
    if (s < 3) {
      c.stage = s + 1;
      return;
    }
    c.nstp++;
    c.stage = 0;
    if (mode == MQCT_PROP_AT_ODEINT) {
      // odeint's MAXSTP-th step: one rk4 step to x2 (suppl_routines.f:1385-1391)
      c.x = c.x2 + copysign(TINY, c.x2 - c.x1);
      c.status = 1;
    } else {
      // a step that would reach the path end is shortened to end there and
      // is the last
      c.x += c.h;
      if (c.last) c.status = 1;
      else if (c.x + c.h >= c.x2) {
        c.h = c.x2 - c.x;
        c.last = 1;
      }
    }
    return;
  }
  if (s == 0) {
    if (c.nstp == MAXSTP - 1) {
      // continue as rk4 with k1 = dydx at x
      c.rk4 = 1;
      c.h = copysign(fabs(c.x2 - c.x), c.x2 - c.x1);
    } else {
      // odeint keeps h at the start of the step (error scale), then clamps h
      c.h0 = c.h;
      if ((c.x + c.h - c.x2) * (c.x + c.h - c.x1) > 0.0) c.h = c.x2 - c.x;
    }
    c.stage = 1;
    return;
  }
  if (s < 5) {
    c.stage = s + 1;
    return;
  }
  // a step at or below hmin is taken as is
  if (!((c.h <= c.hmin) || !(em > 1.0))) {
    const double htemp = 0.9 * c.h * pow(em, -.25);
    const double a = fabs(htemp), b = 0.1 * fabs(c.h);
    c.h = copysign(a > b ? a : b, c.h);
    c.stage = 1;  // retry with the same k1
    return;
  }
  if (c.h > c.hmin) c.hnext = (em > 1.89e-4) ? 0.9 * c.h * pow(em, -.2) : 5.0 * c.h;
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

// Whether ctl_advance's result is known before the stage's states are: at
// every stage but a Cash-Karp step's last one with h above hmin, whose
// error decides it.
__device__ inline bool ctl_known(const Ctl &c) { return c.rk4 || c.stage < 5 || c.h <= c.hmin; }

// Completes the stage the GEMM evaluated, advances the integrator and
// prepares the next stage (into the buckets of parity `par`).  Every thread
// keeps the states i = sid + j TS (j < PT) in registers through the three
// phases -- the right-hand side k of the completed stage, the integrator
// update, the next stage vector -- so k is stored only when a later stage
// reads it, and a finished step's y feeds the next stage without a reload.
// LIST: workgroup b takes entry b of the compacted list of running
// trajectories (the grid is the list's length); otherwise trajectory b.
// CW: wave 0 holds no states.  When the next stage's control record is
// known from this one (ctl_known), wave 0 advances a copy of it and runs the
// next stage's stage_chain (path record, knots, bucket appends) at once,
// while the other waves load and update the states; thread 0 keeps the
// chain's fields in LDS for the final control record.  Otherwise thread 0
// runs the chain at the end, as without CW.
template <int PT, int TB = BS, bool LIST = false, bool CW = false>
__global__ __launch_bounds__(TB) void step_kernel(BDev d, int par, int w32) {
  extern __shared__ double sys[];  // 2N: the completed stage's vector (Coriolis neighbours)
  __shared__ double red[TB / 64];
  __shared__ Ctl cw_next;  // thread 0: the control wave's stage_chain fields
  constexpr int TS = CW ? TB - 64 : TB;  // threads that hold states
  const int tid = threadIdx.x, N = d.N, n2 = 2 * N;
  const int sid = CW ? (tid >= 64 ? tid - 64 : N) : tid;  // wave 0 of CW: every i >= N
  if (blockIdx.x == 0 && tid == 0) *d.qctr = 0;  // item counter of the next GEMM
  const int t = LIST ? __builtin_amdgcn_readfirstlane(d.alist[blockIdx.x]) : (int)blockIdx.x;
  Ctl c = d.ctl[t];
  if (c.status != 0) return;
  const bool early = CW && ctl_known(c);
  if (early && tid == 0) {
    Ctl n = c;
    ctl_advance(n, d.mode, 0.0);
    if (n.status == 0) {
      stage_chain(d, t, n, par);
      cw_next = n;
    }
  }
  double *S = d.S + (long long)t * NSLOT * n2;
  double *y = S;
  const int s = c.stage;
  // the earlier k slots this stage reads, loaded with the other loads below
  // (1024 threads, two states per thread): odeint stage s = 1..4 reads k1..ks,
  // stage 5 k1, k3, k4, k5; rk4's last stage k1, k2, k3
  constexpr bool PF = PT <= 2 && TB == 1024;
  int sl[4] = {1, 2, 3, 4}, nsl = 0;
  if (c.rk4) {
    nsl = s == 3 ? 3 : 0;
  } else if (s == 5) {
    sl[1] = 3;
    sl[2] = 4;
    sl[3] = 5;
    nsl = 4;
  } else {
    nsl = s;
  }
  double2 kp[PF ? 4 : 1][PF ? PT : 1];
  if constexpr (PF) {
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int j = 0; j < PT; j++) {
        const int i = sid + j * TS;
        kp[m][j] = (m < nsl && i < N) ? make_double2(S[sl[m] * n2 + i], S[sl[m] * n2 + i + N])
                                      : make_double2(0.0, 0.0);
      }
  }
  // component p (0 Re, 1 Im) of state sid + j TS of the m-th slot read (slot sl[m])
  auto K = [&](int m, int j, int p) -> double {
    if constexpr (PF) return p ? kp[m][j].y : kp[m][j].x;
    else return S[sl[m] * n2 + sid + j * TS + p * N];
  };
  // -- k = -i exp(i E tau) (W_lo + W_hi) + Coriolis(stage vector), the stage
  //    vector formed again from its rotated amplitudes
  double kr[PT], ki[PT];
  // every load of the thread's states first (energies, rotated stage vector,
  // both knot sides' products, y and the Coriolis tables), so their
  // latencies overlap instead of following one another state by state
  double Ej[PT], yr[PT], yi[PT];
  {
    const double2 *vv = d.V + (long long)t * N;
    // the GEMM's products: FP64, or FP32 after the FP32-product forms (w32)
    const double2 *W0 = d.W + ((long long)t * 2) * N, *W1 = W0 + N;
    const float2 *W0f = reinterpret_cast<const float2 *>(d.W) + ((long long)t * 2) * N, *W1f = W0f + N;
    double2 vj[PT], w0j[PT], w1j[PT], fj[PT];
    int2 nj[PT];
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = sid + j * TS;
      Ej[j] = yr[j] = yi[j] = 0.0;
      vj[j] = w0j[j] = w1j[j] = fj[j] = make_double2(0.0, 0.0);
      nj[j] = make_int2(-1, -1);
      if (i < N) {
        Ej[j] = d.E[i];
        vj[j] = vv[i];
        if (!c.zero) {
          if (w32) {
            const float2 a = W0f[i], b = W1f[i];
            w0j[j] = make_double2(a.x, a.y);
            w1j[j] = make_double2(b.x, b.y);
          } else {
            w0j[j] = W0[i];
            w1j[j] = W1[i];
          }
        }
        yr[j] = y[i];
        yi[j] = y[i + N];
        nj[j] = d.nb[i];
        fj[j] = d.sq[i];
      }
    }
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = sid + j * TS;
      kr[j] = ki[j] = 0.0;
      if (i < N) {
        double sn, cs;
        sincos(c.tau * Ej[j], &sn, &cs);
        const double2 vi = vj[j];
        sys[i] = vi.x * cs - vi.y * sn;
        sys[i + N] = vi.y * cs + vi.x * sn;
        if (!c.zero) {
          const double wx = w0j[j].x + w1j[j].x, wy = w0j[j].y + w1j[j].y;
          kr[j] = sn * wx + cs * wy;
          ki[j] = -(cs * wx - sn * wy);
        }
      }
    }
    __syncthreads();
    const double hd = c.dphidt / 2.0;
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = sid + j * TS;
      if (i < N) {
        // state-diagonal Coriolis coupling (mqct_eqtns.f:845-877)
        const int2 nbr = nj[j];
        const double2 f = fj[j];
        if (nbr.x >= 0) {
          kr[j] += hd * f.x * sys[nbr.x + N];
          ki[j] -= hd * f.x * sys[nbr.x];
        }
        if (nbr.y >= 0) {
          kr[j] += hd * f.y * sys[nbr.y + N];
          ki[j] -= hd * f.y * sys[nbr.y];
        }
      }
    }
  }
  c.nrhs++;
  const double TINY = 1.e-30;
  double xr[PT], xi[PT];  // the next stage's vector
  // k into its slot (k1 .. k6 at S + (1 + s) n2) for the stages that read it
  auto store_k = [&]() {
    double *k = S + (1 + s) * n2;
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = sid + j * TS;
      if (i < N) {
        k[i] = kr[j];
        k[i + N] = ki[j];
      }
    }
  };
  // x = y + f k (one earlier slot at most: rk4 stages, odeint stage 1)
  auto stage_from_k = [&](double f) {
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = sid + j * TS;
      if (i < N) {
        xr[j] = yr[j] + f * kr[j];
        xi[j] = yi[j] + f * ki[j];
      }
    }
  };
  // the integrator's vector part; ctl_advance makes the control update
  if (c.rk4) {
    if (s < 3) {
      store_k();
      stage_from_k(s < 2 ? 0.5 * c.h : c.h);
      ctl_advance(c, d.mode, 0.0);
    } else {
      const double h6 = c.h / 6.0;
      double nrm = 0.0;
#pragma unroll
      for (int j = 0; j < PT; j++) {
        const int i = sid + j * TS;
        if (i < N) {
          // k1, k2, k3 are reads 0, 1, 2
          xr[j] = yr[j] + h6 * (K(0, j, 0) + kr[j] + 2.0 * (K(1, j, 0) + K(2, j, 0)));
          xi[j] = yi[j] + h6 * (K(0, j, 1) + ki[j] + 2.0 * (K(1, j, 1) + K(2, j, 1)));
          y[i] = xr[j];
          y[i + N] = xi[j];
          nrm += xr[j] * xr[j] + xi[j] * xi[j];
        }
      }
      ctl_advance(c, d.mode, 0.0);
      nrm = block_sum<TB>(nrm, red);
      c.norm_err = fmax(c.norm_err, fabs(1.0 - nrm));
    }
  } else if (s == 0) {
    // odeint: yscal = |y| + |h dydx| + TINY with h at the start of the step
    // (formed again from y and k1 when the error is measured); the first
    // stage uses h as clamped, or rk4's half step after MAXSTP - 1 steps
    store_k();
    ctl_advance(c, d.mode, 0.0);
    stage_from_k((c.rk4 ? 0.5 : CK_B[1][0]) * c.h);
  } else if (s < 5) {
    // Cash-Karp stage s + 1 = y + h sum_q B[s+1][q] k_{q+1}, the last k in registers
    store_k();
    ctl_advance(c, d.mode, 0.0);
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = sid + j * TS;
      if (i < N) {
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          if (q < s) {  // k_{q+1} is read q
            ar += CK_B[s + 1][q] * K(q, j, 0);
            ai += CK_B[s + 1][q] * K(q, j, 1);
          }
        }
        ar += CK_B[s + 1][s] * kr[j];
        ai += CK_B[s + 1][s] * ki[j];
        xr[j] = yr[j] + c.h * ar;
        xi[j] = yi[j] + c.h * ai;
      }
    }
  } else {
    // rkck's error estimate and fifth-order solution in one pass (k6 in
    // registers), then rkqs's step control (suppl_routines.f:1229-1283)
    const double C1 = 37.0 / 378.0, C3 = 250.0 / 621.0, C4 = 125.0 / 594.0, C6 = 512.0 / 1771.0;
    const double DC1 = C1 - 2825.0 / 27648.0, DC3 = C3 - 18575.0 / 48384.0,
                 DC4 = C4 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0, DC6 = C6 - .25;
    double em = 0.0;
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = sid + j * TS;
      if (i < N) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
          // k1, k3, k4, k5 are reads 0..3
          const double k1 = K(0, j, p), k3 = K(1, j, p), k4 = K(2, j, p), k5 = K(3, j, p);
          const double k6 = p ? ki[j] : kr[j];
          const double yerr = c.h * (DC1 * k1 + DC3 * k3 + DC4 * k4 + DC5 * k5 + DC6 * k6);
          const double yv = p ? yi[j] : yr[j];
          const double ysc = fabs(yv) + fabs(c.h0 * k1) + TINY;
          em = fmax(em, fabs(yerr / ysc));
          const double yn = yv + c.h * (C1 * k1 + C3 * k3 + C4 * k4 + C6 * k6);
          if (p) xi[j] = yn;
          else xr[j] = yn;
        }
      }
    }
    em = block_max<TB>(em, red) / c.eps;
    // a step at or below hmin is taken as is
    const bool accept = (c.h <= c.hmin) || !(em > 1.0);
    ctl_advance(c, d.mode, em);  // a rejected step retries with the same k1 and a smaller h
    if (!accept) {
#pragma unroll
      for (int j = 0; j < PT; j++) {
        const int i = sid + j * TS;
        if (i < N) {
          xr[j] = yr[j] + CK_B[1][0] * c.h * K(0, j, 0);
          xi[j] = yi[j] + CK_B[1][0] * c.h * K(0, j, 1);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < PT; j++) {
        const int i = sid + j * TS;
        if (i < N) {
          y[i] = xr[j];
          y[i + N] = xi[j];
        }
      }
    }
  }
  if (c.status != 0) {
    if (tid == 0) {
      atomicSub(d.active, 1);
      d.ctl[t] = c;
    }
    return;
  }
  prepare_stage<PT, TS>(d, t, c, xr, xi, par, Ej, sid, !early);
  if (tid == 0) {
    if (early) copy_chain(c, cw_next);
    d.ctl[t] = c;
  }
}

// ---------------------------------------------------------------------------
// CC-MQCT in the batched form.  The GEMM gives per trajectory M(R) v from
// the two knots of R (product slots 0, 1) and dM/dR v (slots 2, 3: the same
// knot matrices with the R-derivatives of the spline weights, as splint_bk's
// derivative branch, mqct_eqtns.f:575-607).  With W = M v and W' = dM/dR v
// the pair sums of DERIVS_BK (mqct_eqtns.f:263-359) are sums over states:
// M is real symmetric and couples equal m only, so a pair sum over i >= j
// with d_f = 2 off the diagonal is the full double sum, and
//   mean-field force  -sum d_f dM_ij Re(v_i* v_j)          = -sum_j Re(v_j* W'_j)
//   dP_Phi/dt  sin(Th) sum d_f/2 M_ij (Im(v_j+-* v_i) terms) = -sin(Th) sum_j Im(W_j* (P v)_j)
//   P_Theta part      sum d_f/2 M_ij (Re terms)             =  sum_j Re(W_j* (Q v)_j)
// with (P v)_j = c+_j v_{j+1} + c-_j v_{j-1}, (Q v)_j = c+_j v_{j+1} - c-_j v_{j-1}
// and c+-_j = sqrt(j(j+1) - m(m +- 1)), zero at m = +-j.

// The next stage of a CC trajectory: its rotated amplitudes v = exp(-i E
// tau) x into V; thread 0 finds the knots of the stage's R (cc.xc[0]) and
// appends four entries: the spline weights at the knot below and above, and
// their R-derivatives -1/h, -(3A^2-1)h/6 and 1/h, (3B^2-1)h/6.
template <int PT, int TB = BS>
__device__ void prepare_stage_cc(const BDev &d, int t, Ctl &c, const CtlCC &cc, const double *xr,
                                 const double *xi, int par, const double *Ej) {
  const int tid = threadIdx.x, N = d.N;
  const double tau = stage_time(c);
  double2 *v = d.V + (long long)t * N;
#pragma unroll
  for (int j = 0; j < PT; j++) {
    const int i = tid + j * TB;
    if (i < N) {
      double sn, cs;
      sincos(tau * Ej[j], &sn, &cs);
      v[i] = make_double2(xr[j] * cs + xi[j] * sn, xi[j] * cs - xr[j] * sn);
    }
  }
  if (tid == 0) {
    const int nr = d.n_r;
    const double R = cc.xc[0];
    c.tau = tau;
    c.zero = R > d.R[nr - 1];
    const int klo = interval_cached(d.R, 1, nr, R, c.klo, c.knot_r0, c.knot_r1), khi = klo + 1;
    const double hk = d.R[khi] - d.R[klo];
    const double aa = (d.R[khi] - R) / hk, bb = (R - d.R[klo]) / hk;
    const double cc5 = (hk * hk) / 6.0, cc6 = cc5 / hk;
    c.aa = aa;
    c.bb = bb;
    c.c1 = (aa * aa * aa - aa) * cc5;
    c.c2 = (bb * bb * bb - bb) * cc5;
    if (!c.zero) {
      if (d.nbucketed) atomicAdd(d.nbucketed, 2ULL);  // twice an AT right-hand side's products
      int *cnt = d.cnt + par * nr;
      Entry *bk = d.bucket + (long long)par * nr * d.bcap;
      const int p0 = atomicAdd(&cnt[klo], 2), p1 = atomicAdd(&cnt[khi], 2);
      if (p0 == 0) d.klist[par * nr + atomicAdd(&d.kcnt[par], 1)] = klo;
      if (p1 == 0) d.klist[par * nr + atomicAdd(&d.kcnt[par], 1)] = khi;
      bk[(long long)klo * d.bcap + p0] = Entry{t, 0, aa, c.c1};
      bk[(long long)klo * d.bcap + p0 + 1] = Entry{t, 2, -1.0 / hk, -(3.0 * aa * aa - 1.0) * cc6};
      bk[(long long)khi * d.bcap + p1] = Entry{t, 1, bb, c.c2};
      bk[(long long)khi * d.bcap + p1 + 1] = Entry{t, 3, 1.0 / hk, (3.0 * bb * bb - 1.0) * cc6};
    }
  }
}

// The CC step: completes the stage the GEMM evaluated (amplitude and
// classical derivatives, DERIVS_BK), advances rk4 or odeint/rkqs/rkck over
// all 2N + 8 variables (the classical eight kept by thread 0), ends the dt
// interval of DOTRAJECT with its checks (propagator.f:3995-4050: turning
// points of R, orbiting periods of Phi, time limit, R past R_start, step
// cap), and prepares the next stage.
template <int PT, int TB = BS, bool LIST = false>
__global__ __launch_bounds__(TB) void cc_step_kernel(BDev d, int par, int w32) {
  extern __shared__ double sys[];  // 2N: the completed stage's vector (Coriolis neighbours)
  __shared__ double red[8];
  __shared__ Ctl s_ctl;            // the control after an interval's end, from thread 0
  const int tid = threadIdx.x, N = d.N, n2 = 2 * N, ns = n2 + 8;
  if (blockIdx.x == 0 && tid == 0) *d.qctr = 0;  // item counter of the next GEMM
  const int t = LIST ? __builtin_amdgcn_readfirstlane(d.alist[blockIdx.x]) : (int)blockIdx.x;  // as in step_kernel
  Ctl c = d.ctl[t];
  if (c.status != 0) return;
  CtlCC cc = d.cc[t];
  double *S = d.S + (long long)t * NSLOT * ns;
  double *y = S;
  const int s = c.stage;
  const double mu = d.R[-1];
  const double *xc = cc.xc;  // the completed stage's classical part
  const double sin_t = sin(xc[2]), cos_t = cos(xc[2]);
  const double dPhidT = xc[5] / xc[0] / xc[0] / mu / (sin_t * sin_t);
  const double dThetadT = d.theta_on ? xc[3] / xc[0] / xc[0] / mu : 0.0;
  double kr[PT], ki[PT], Ej[PT], yr[PT], yi[PT];
  double fsum = 0.0, psum = 0.0, tsum = 0.0;  // -Re(v* W'), Im(W* P v), Re(W* Q v) partial sums
  {
    const double2 *vv = d.V + (long long)t * N;
    const long long w0 = (long long)t * 4 * N;
    double2 vj[PT], wj[PT], dj[PT], fj[PT], uj[PT], lj[PT];
    int2 nj[PT];
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = tid + j * TB;
      Ej[j] = yr[j] = yi[j] = 0.0;
      vj[j] = wj[j] = dj[j] = fj[j] = uj[j] = lj[j] = make_double2(0.0, 0.0);
      nj[j] = make_int2(-1, -1);
      if (i < N) {
        Ej[j] = d.E[i];
        vj[j] = vv[i];
        nj[j] = d.nb[i];
        fj[j] = d.sq[i];
        if (!c.zero) {
          double2 p[4];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            if (w32) {
              const float2 a = reinterpret_cast<const float2 *>(d.W)[w0 + q * N + i];
              p[q] = make_double2(a.x, a.y);
            } else {
              p[q] = d.W[w0 + q * N + i];
            }
          }
          wj[j] = make_double2(p[0].x + p[1].x, p[0].y + p[1].y);
          dj[j] = make_double2(p[2].x + p[3].x, p[2].y + p[3].y);
          if (nj[j].x >= 0) lj[j] = vv[nj[j].x];
          if (nj[j].y >= 0) uj[j] = vv[nj[j].y];
        }
        yr[j] = y[i];
        yi[j] = y[i + N];
      }
    }
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = tid + j * TB;
      kr[j] = ki[j] = 0.0;
      if (i < N) {
        double sn, cs;
        sincos(c.tau * Ej[j], &sn, &cs);
        const double2 vi = vj[j];
        sys[i] = vi.x * cs - vi.y * sn;
        sys[i + N] = vi.y * cs + vi.x * sn;
        if (!c.zero) {
          const double2 w = wj[j], f = fj[j];
          kr[j] = sn * w.x + cs * w.y;
          ki[j] = -(cs * w.x - sn * w.y);
          fsum -= vi.x * dj[j].x + vi.y * dj[j].y;
          // f = (c-, c+): the factors of state-1 and state+1
          const double px = f.y * uj[j].x + f.x * lj[j].x, py = f.y * uj[j].y + f.x * lj[j].y;
          const double qx = f.y * uj[j].x - f.x * lj[j].x, qy = f.y * uj[j].y - f.x * lj[j].y;
          psum += w.x * py - w.y * px;
          tsum += w.x * qx + w.y * qy;
        }
      }
    }
    __syncthreads();
    // state-diagonal Coriolis coupling of the stage vector (mqct_eqtns.f:389-466)
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = tid + j * TB;
      if (i < N) {
        const int2 nbr = nj[j];
        const double2 f = fj[j];
        if (d.theta_on) {
          const double m = d.m_pos[i];
          kr[j] += dPhidT * m * cos_t * sys[i];
          ki[j] -= dPhidT * m * cos_t * sys[i + N];
          if (nbr.x >= 0) {
            kr[j] += (dPhidT * sin_t * sys[nbr.x + N] - dThetadT * sys[nbr.x]) * f.x / 2.0;
            ki[j] -= (dPhidT * sin_t * sys[nbr.x] - dThetadT * sys[nbr.x + N]) * f.x / 2.0;
          }
          if (nbr.y >= 0) {
            kr[j] += (dPhidT * sin_t * sys[nbr.y + N] - dThetadT * sys[nbr.y]) * f.y / 2.0;
            ki[j] -= (dPhidT * sin_t * sys[nbr.y] - dThetadT * sys[nbr.y + N]) * f.y / 2.0;
          }
        } else {
          const double hd = dPhidT / 2.0;
          if (nbr.x >= 0) {
            kr[j] += hd * f.x * sys[nbr.x + N];
            ki[j] -= hd * f.x * sys[nbr.x];
          }
          if (nbr.y >= 0) {
            kr[j] += hd * f.y * sys[nbr.y + N];
            ki[j] -= hd * f.y * sys[nbr.y];
          }
        }
      }
    }
  }
  if (!c.zero) {
    fsum = block_sum<TB>(fsum, red);
    psum = block_sum<TB>(psum, red);
    if (d.theta_on) tsum = block_sum<TB>(tsum, red);
  }
  // classical derivatives (mqct_eqtns.f:473-495), thread 0
  double kc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (tid == 0) {
    kc[0] = xc[1] / mu;
    kc[4] = xc[5] / xc[0] / xc[0] / mu / (sin_t * sin_t);
    kc[1] = fsum + xc[3] / xc[0] / xc[0] / mu * xc[3] / xc[0] +
            xc[5] / xc[0] / xc[0] / mu * xc[5] / xc[0] / (sin_t * sin_t);
    kc[5] = -sin_t * psum;
    if (d.theta_on) {
      kc[2] = xc[3] / xc[0] / xc[0] / mu;
      kc[3] = tsum + xc[5] / xc[0] / xc[0] / mu * xc[5] / (sin_t * sin_t * sin_t) * cos_t;
    }
  }
  c.nrhs++;
  const double TINY = 1.e-30;
  double xr[PT], xi[PT];            // the next stage's vector
  double ycl[8], xcn[8];            // thread 0: classical y and the next stage's classical part
  if (tid == 0)
    for (int q = 0; q < 8; q++) ycl[q] = y[n2 + q];
  auto store_k = [&]() {
    double *k = S + (1 + s) * ns;
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = tid + j * TB;
      if (i < N) {
        k[i] = kr[j];
        k[i + N] = ki[j];
      }
    }
    if (tid == 0)
      for (int q = 0; q < 8; q++) k[n2 + q] = kc[q];
  };
  auto stage_from_k = [&](double f) {
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = tid + j * TB;
      if (i < N) {
        xr[j] = yr[j] + f * kr[j];
        xi[j] = yi[j] + f * ki[j];
      }
    }
    if (tid == 0)
      for (int q = 0; q < 8; q++) xcn[q] = ycl[q] + f * kc[q];
  };
  bool iv_end = false;  // the dt interval ended with this stage (uniform over the block)
  if (c.rk4) {
    if (s < 3) {
      store_k();
      c.stage = s + 1;
      stage_from_k(s < 2 ? 0.5 * c.h : c.h);
    } else {
      const double h6 = c.h / 6.0;
      const double *k1 = S + ns, *k2 = S + 2 * ns, *k3 = S + 3 * ns;
#pragma unroll
      for (int j = 0; j < PT; j++) {
        const int i = tid + j * TB;
        if (i < N) {
          xr[j] = yr[j] + h6 * (k1[i] + kr[j] + 2.0 * (k2[i] + k3[i]));
          xi[j] = yi[j] + h6 * (k1[i + N] + ki[j] + 2.0 * (k2[i + N] + k3[i + N]));
          y[i] = xr[j];
          y[i + N] = xi[j];
        }
      }
      if (tid == 0)
        for (int q = 0; q < 8; q++) {
          const int ii = n2 + q;
          xcn[q] = ycl[q] + h6 * (k1[ii] + kc[q] + 2.0 * (k2[ii] + k3[ii]));
          y[ii] = xcn[q];
        }
      c.nstp++;
      c.stage = 0;
      // odeint's MAXSTP-th step ends the trajectory as failed (suppl_routines.f:1385-1391)
      if (d.mode == MQCT_PROP_CC_ODEINT && tid == 0) cc.fail = 1;
      iv_end = true;
    }
  } else if (s == 0) {
    store_k();
    if (c.nstp == MAXSTP - 1) {
      c.rk4 = 1;
      c.h = copysign(fabs(c.x2 - c.x), c.x2 - c.x1);
      c.stage = 1;
      stage_from_k(0.5 * c.h);
    } else {
      c.h0 = c.h;
      if ((c.x + c.h - c.x2) * (c.x + c.h - c.x1) > 0.0) c.h = c.x2 - c.x;
      c.stage = 1;
      stage_from_k(CK_B[1][0] * c.h);
    }
  } else if (s < 5) {
    store_k();
    c.stage = s + 1;
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = tid + j * TB;
      if (i < N) {
        double ar = 0.0, ai = 0.0;
        for (int q = 0; q < s; q++) {
          ar += CK_B[s + 1][q] * S[(1 + q) * ns + i];
          ai += CK_B[s + 1][q] * S[(1 + q) * ns + i + N];
        }
        ar += CK_B[s + 1][s] * kr[j];
        ai += CK_B[s + 1][s] * ki[j];
        xr[j] = yr[j] + c.h * ar;
        xi[j] = yi[j] + c.h * ai;
      }
    }
    if (tid == 0)
      for (int cv = 0; cv < 8; cv++) {
        double a = 0.0;
        for (int q = 0; q < s; q++) a += CK_B[s + 1][q] * S[(1 + q) * ns + n2 + cv];
        a += CK_B[s + 1][s] * kc[cv];
        xcn[cv] = ycl[cv] + c.h * a;
      }
  } else {
    const double C1 = 37.0 / 378.0, C3 = 250.0 / 621.0, C4 = 125.0 / 594.0, C6 = 512.0 / 1771.0;
    const double DC1 = C1 - 2825.0 / 27648.0, DC3 = C3 - 18575.0 / 48384.0,
                 DC4 = C4 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0, DC6 = C6 - .25;
    const double *k1 = S + ns, *k3 = S + 3 * ns, *k4 = S + 4 * ns, *k5 = S + 5 * ns;
    double em = 0.0;
    auto one = [&](int ii, double k6, double yv) {
      const double yerr = c.h * (DC1 * k1[ii] + DC3 * k3[ii] + DC4 * k4[ii] + DC5 * k5[ii] + DC6 * k6);
      const double ysc = fabs(yv) + fabs(c.h0 * k1[ii]) + TINY;
      em = fmax(em, fabs(yerr / ysc));
      return yv + c.h * (C1 * k1[ii] + C3 * k3[ii] + C4 * k4[ii] + C6 * k6);
    };
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int i = tid + j * TB;
      if (i < N) {
        xr[j] = one(i, kr[j], yr[j]);
        xi[j] = one(i + N, ki[j], yi[j]);
      }
    }
    if (tid == 0)
      for (int q = 0; q < 8; q++) xcn[q] = one(n2 + q, kc[q], ycl[q]);
    em = block_max<TB>(em, red) / c.eps;
    const bool accept = (c.h <= c.hmin) || !(em > 1.0);
    if (!accept) {
      const double htemp = 0.9 * c.h * pow(em, -.25);
      const double a = fabs(htemp), b = 0.1 * fabs(c.h);
      c.h = copysign(a > b ? a : b, c.h);
      c.stage = 1;  // retry with the same k1
#pragma unroll
      for (int j = 0; j < PT; j++) {
        const int i = tid + j * TB;
        if (i < N) {
          xr[j] = yr[j] + CK_B[1][0] * c.h * k1[i];
          xi[j] = yi[j] + CK_B[1][0] * c.h * k1[i + N];
        }
      }
      if (tid == 0)
        for (int q = 0; q < 8; q++) xcn[q] = ycl[q] + CK_B[1][0] * c.h * k1[n2 + q];
    } else {
      if (c.h > c.hmin) c.hnext = (em > 1.89e-4) ? 0.9 * c.h * pow(em, -.2) : 5.0 * c.h;
#pragma unroll
      for (int j = 0; j < PT; j++) {
        const int i = tid + j * TB;
        if (i < N) {
          y[i] = xr[j];
          y[i + N] = xi[j];
        }
      }
      if (tid == 0)
        for (int q = 0; q < 8; q++) y[n2 + q] = xcn[q];
      c.x += c.h;
      c.nstp++;
      c.stage = 0;
      if ((c.x - c.x2) * (c.x2 - c.x1) >= 0.0) {
        iv_end = true;  // odeint reached the interval's end; its hnext is the next first step
      } else {
        if (fabs(c.hnext) < c.hmin) c.hnext = copysign(c.hmin, c.x2 - c.x1);
        c.h = c.hnext;
      }
    }
  }
  if (iv_end) {
    // DOTRAJECT after odeint / rk4 over [tcur, tcur + dt]
    double nrm = 0.0;
#pragma unroll
    for (int j = 0; j < PT; j++)
      if (tid + j * TB < N) nrm += xr[j] * xr[j] + xi[j] * xi[j];
    nrm = block_sum<TB>(nrm, red);
    if (tid == 0) {
      const mqct_traj_t &p = d.traj[t];
      c.norm_err = fmax(c.norm_err, fabs(1.0 - nrm));
      cc.nsteps += c.nstp;
      if (d.mode == MQCT_PROP_CC_ODEINT) cc.dt_step = c.hnext;
      bool run = cc.fail == 0;
      if (run) {
        const double uR = xcn[0], uphi = xcn[4];
        if (cc.tcur > p.dt && cc.counter > 2)
          if ((cc.tmp_R2 - cc.tmp_R1) < 0.0 && (cc.tmp_R1 - uR) > 0.0) cc.vib++;
        cc.tmp_R2 = cc.tmp_R1;
        cc.tmp_R1 = uR;
        cc.tcur += p.dt;
        const int period_cnt = (int)(fabs(uphi) / 2.0 / MQCT_PI);
        if (p.numb_orb_prds <= period_cnt) run = false;
        if (cc.vib >= p.numb_oscl_prds) run = false;
        if (p.time_lim > 0.0 && cc.tcur >= p.time_lim) run = false;
        if (uR > p.R_start) run = false;
        if (p.max_steps > 0 && cc.nsteps >= p.max_steps) {
          run = false;
          cc.fail = 2;
        }
      }
      if (!run) {
        c.status = 1;
      } else {
        cc.counter++;
        c.x = c.x1 = cc.tcur;
        c.x2 = cc.tcur + p.dt;
        c.nstp = 0;
        c.stage = 0;
        c.rk4 = d.mode == MQCT_PROP_CC_RK4;
        c.h = c.rk4 ? p.dt : cc.dt_step;
        c.hnext = c.h;
      }
      s_ctl = c;
    }
    __syncthreads();
    c = s_ctl;
  }
  if (c.status != 0) {
    if (tid == 0) {
      atomicSub(d.active, 1);
      d.ctl[t] = c;
      d.cc[t] = cc;
    }
    return;
  }
  if (tid == 0)
    for (int q = 0; q < 8; q++) cc.xc[q] = xcn[q];
  prepare_stage_cc<PT, TB>(d, t, c, cc, xr, xi, par, Ej);
  if (tid == 0) {
    d.ctl[t] = c;
    d.cc[t] = cc;
  }
}

// dense block matrices of every knot from the R-major pair arrays
template <typename KT>
__global__ void knot_matrices_kernel(long long T, int n_r, const double *__restrict__ Mt,
                                     const double *__restrict__ Y2t, const int *__restrict__ pi,
                                     const int *__restrict__ pj, const int *__restrict__ state_pos,
                                     const int *__restrict__ pos_blk, const int *__restrict__ blk_start,
                                     const long long *__restrict__ blk_off, long long S2,
                                     KT *__restrict__ KM) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= T * n_r) return;
  const int r = (int)(idx / T);
  const long long k = idx % T;
  const int a = state_pos[pi[k]], b = state_pos[pj[k]];
  const int blk = pos_blk[a];
  const int nb = blk_start[blk + 1] - blk_start[blk];
  const int la = a - blk_start[blk], lb = b - blk_start[blk];
  const KT m = (KT)Mt[(long long)r * T + k], m2 = (KT)Y2t[(long long)r * T + k];
  KT *Yr = KM + ((long long)r * 2) * S2 + blk_off[blk];
  KT *Y2r = KM + ((long long)r * 2 + 1) * S2 + blk_off[blk];
  Yr[(long long)la * nb + lb] = m;
  Yr[(long long)lb * nb + la] = m;
  Y2r[(long long)la * nb + lb] = m2;
  Y2r[(long long)lb * nb + la] = m2;
}

// the same matrices in the fragment order of the tile layout: element
// (row i, column k) of block [Y | Y2] (k < 2 nbp, nbp = nb rounded up to
// 4U) lies in tile i / 16 at ((k / 4U) 64 + lane) U + k % U with lane =
// 16 ((k % 4U) / U) + i % 16
template <typename KT>
__global__ void knot_tiles_kernel(long long T, int n_r, const double *__restrict__ Mt,
                                  const double *__restrict__ Y2t, const int *__restrict__ pi,
                                  const int *__restrict__ pj, const int *__restrict__ state_pos,
                                  const int *__restrict__ pos_blk, const int *__restrict__ blk_start,
                                  const int *__restrict__ blk_tile0, const long long *__restrict__ tile_off,
                                  long long KS, KT *__restrict__ KT_) {
  constexpr int U = Frag<KT>::U, KB = 4 * U;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= T * n_r) return;
  const int r = (int)(idx / T);
  const long long k = idx % T;
  const int a = state_pos[pi[k]], b = state_pos[pj[k]];
  const int blk = pos_blk[a];
  const int nb = blk_start[blk + 1] - blk_start[blk];
  const int nbp = (nb + KB - 1) / KB * KB;
  const int la = a - blk_start[blk], lb = b - blk_start[blk];
  const KT m = (KT)Mt[(long long)r * T + k], m2 = (KT)Y2t[(long long)r * T + k];
  KT *base = KT_ + (long long)r * KS;
  auto put = [&](int row, int kc, KT v) {
    const int lane = 16 * ((kc % KB) / U) + (row & 15);
    base[tile_off[blk_tile0[blk] + row / 16] + ((long long)(kc / KB) * 64 + lane) * U + kc % U] = v;
  };
  put(la, lb, m);
  put(lb, la, m);
  put(la, nbp + lb, m2);
  put(lb, nbp + la, m2);
}

template <int PT>
__global__ __launch_bounds__(BS) void init_kernel(BDev d, const mqct_traj_t *__restrict__ tr,
                                                  const int *__restrict__ state_pos) {
  const int t = blockIdx.x, tid = threadIdx.x, N = d.N, n2 = 2 * N;
  double *S = d.S + (long long)t * NSLOT * n2;
  const mqct_traj_t p = tr[t];
  const int s0 = state_pos[p.s_st - 1];
  double xr[PT], xi[PT], Ej[PT];
#pragma unroll
  for (int j = 0; j < PT; j++) {
    const int i = tid + j * BS;
    xr[j] = xi[j] = Ej[j] = 0.0;
    if (i < N) {
      Ej[j] = d.E[i];
      xr[j] = (i == s0) ? 1.0 : 0.0;
      S[i] = xr[j];
      S[i + N] = 0.0;
    }
  }
  Ctl c{};
  const long long o1 = d.poff[p.at_path + 1];
  c.x = 0.0;
  c.x1 = 0.0;
  c.x2 = d.prec[(o1 - 1) * 8];
  c.rk4 = d.mode == MQCT_PROP_AT_RK4;
  c.h = p.dt;
  if (c.rk4 && c.x + c.h >= c.x2) {
    c.h = c.x2 - c.x;
    c.last = 1;
  }
  c.hnext = c.h;
  c.hmin = p.hmin;
  c.eps = p.eps;
  c.path = p.at_path;
  prepare_stage<PT, BS>(d, t, c, xr, xi, 0, Ej, tid, true);
  if (tid == 0) {
    d.ctl[t] = c;
    d.alist[t] = t;
  }
}

// CC initial conditions (propagator.f:2431-2566): the initial state, R =
// R_start, P_R = -sqrt(k^2 - l(l+1)/R^2), Theta = pi/2, P_Theta = 0, Phi =
// 0, P_Phi = sqrt(l(l+1)); the first interval [0, dt].
template <int PT>
__global__ __launch_bounds__(BS) void cc_init_kernel(BDev d, const int *__restrict__ state_pos) {
  const int t = blockIdx.x, tid = threadIdx.x, N = d.N, n2 = 2 * N, ns = n2 + 8;
  double *S = d.S + (long long)t * NSLOT * ns;
  const mqct_traj_t p = d.traj[t];
  const int s0 = state_pos[p.s_st - 1];
  double xr[PT], xi[PT], Ej[PT];
#pragma unroll
  for (int j = 0; j < PT; j++) {
    const int i = tid + j * BS;
    xr[j] = xi[j] = Ej[j] = 0.0;
    if (i < N) {
      Ej[j] = d.E[i];
      xr[j] = (i == s0) ? 1.0 : 0.0;
      S[i] = xr[j];
      S[i + N] = 0.0;
    }
  }
  const double mu = d.R[-1];
  const double k_vec = sqrt(mu * 2.0 * p.E_coll), l_orb = sqrt(p.l_real * (p.l_real + 1.0));
  CtlCC cc{};
  cc.xc[0] = p.R_start;
  cc.xc[1] = -sqrt(k_vec * k_vec - l_orb * l_orb / (p.R_start * p.R_start));
  cc.xc[2] = acos(0.0);
  cc.xc[3] = l_orb * sin(0.0);
  cc.xc[4] = 0.0;
  cc.xc[5] = l_orb * cos(0.0);
  if (tid < 8) S[n2 + tid] = cc.xc[tid];
  cc.dt_step = p.dt;
  cc.counter = 1;
  Ctl c{};
  c.x = c.x1 = 0.0;
  c.x2 = p.dt;
  c.rk4 = d.mode == MQCT_PROP_CC_RK4;
  c.h = p.dt;
  c.hnext = c.h;
  c.hmin = p.hmin;
  c.eps = p.eps;
  prepare_stage_cc<PT>(d, t, c, cc, xr, xi, 0, Ej);
  if (tid == 0) {
    d.ctl[t] = c;
    d.cc[t] = cc;
    d.alist[t] = t;
  }
}

// Order-preserving compaction of the step kernel's trajectory list (one
// workgroup): the entries of alist[0 .. n) whose trajectories still run
// move to its front, in order.
__global__ __launch_bounds__(1024) void compact_kernel(const Ctl *__restrict__ ctl, int *__restrict__ alist,
                                                       int *__restrict__ tmp, int n) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (n + 1023) / 1024, a = min(n, tid * per), b = min(n, a + per);
  int own = 0;
  for (int i = a; i < b; i++) own += ctl[alist[i]].status == 0;
  int inc = own;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int pos = inc - own, tot = 0;
  for (int w = 0; w < 16; w++) {
    if (w < wave) pos += wsum[w];
    tot += wsum[w];
  }
  for (int i = a; i < b; i++) {
    const int v = alist[i];
    if (ctl[v].status == 0) tmp[pos++] = v;
  }
  __syncthreads();
  for (int i = tid; i < tot; i += 1024) alist[i] = tmp[i];
}

__global__ __launch_bounds__(BS) void out_kernel(BDev d, const int *__restrict__ pos_state,
                                                 double *__restrict__ q, mqct_traj_stat_t *__restrict__ st) {
  __shared__ double red[8];
  const int t = blockIdx.x, tid = threadIdx.x, N = d.N, n2 = 2 * N;
  const int ns = d.cc ? n2 + 8 : n2;  // CC slots carry the classical eight
  const double *y = d.S + (long long)t * NSLOT * ns;
  double *o = q + (long long)t * (n2 + 8);
  for (int p = tid; p < N; p += BS) {
    const int s = pos_state[p];
    o[s] = y[p];
    o[s + N] = y[p + N];
  }
  for (int i = tid; i < 8; i += BS) o[n2 + i] = d.cc ? y[n2 + i] : 0.0;
  double nrm = 0.0;
  for (int i = tid; i < N; i += BS) nrm += y[i] * y[i] + y[i + N] * y[i + N];
  nrm = block_sum(nrm, red);
  if (tid == 0) {
    const Ctl &c = d.ctl[t];
    mqct_traj_stat_t s;
    if (d.cc) {
      const CtlCC &cc = d.cc[t];
      s.t_end = cc.tcur;
      s.n_steps = cc.nsteps;
      s.n_rhs = c.nrhs;
      s.fail = cc.fail;
      s.norm_err = c.norm_err;
    } else {
      s.t_end = (d.mode == MQCT_PROP_AT_ODEINT) ? c.x2 : c.x;
      s.n_steps = c.nstp;
      s.n_rhs = c.nrhs;
      s.fail = 0;
      s.norm_err = (d.mode == MQCT_PROP_AT_ODEINT) ? fabs(1.0 - nrm) : c.norm_err;
    }
    st[t] = s;
  }
}

}  // namespace

int mqct_dyn_run_batched(mqct_hip *h, int mode, int n_traj, const mqct_traj_t *traj,
                         double *q_out, mqct_traj_stat_t *stat_out, bool knots_fp32, bool prod_fp32) {
  DynState *ds = h->dyn;
  const int N = ds->N, n_r = ds->n_r;
  const long long T = ds->T;
  // CC modes: classical variables in the slots, four product vectors per
  // trajectory (M v and dM/dR v at two knots), two bucket entries per knot
  const bool ccm = mode == MQCT_PROP_CC_RK4 || mode == MQCT_PROP_CC_ODEINT;
  (void)hipGetLastError();  // a failed earlier call must not be reported here
  // block order of the states (m12 ascending, states ascending)
  std::vector<int> pos_state, state_pos(N), blk_start, pos_blk(N);
  std::vector<long long> blk_off;
  long long S2 = 0;
  {
    std::vector<std::pair<int, int>> key(N);
    for (int s = 0; s < N; s++) key[s] = {ds->m12[s], s};
    std::sort(key.begin(), key.end());
    for (int p = 0; p < N; p++) {
      pos_state.push_back(key[p].second);
      if (p == 0 || key[p].first != key[p - 1].first) blk_start.push_back(p);
    }
    blk_start.push_back(N);
    for (size_t b = 0; b + 1 < blk_start.size(); b++) {
      const long long nb = blk_start[b + 1] - blk_start[b];
      blk_off.push_back(S2);
      S2 += nb * nb;
      for (int p = blk_start[b]; p < blk_start[b + 1]; p++) pos_blk[p] = (int)b;
    }
  }
  for (int p = 0; p < N; p++) state_pos[pos_state[p]] = p;
  const int nblk = (int)blk_start.size() - 1;
  int maxnb = 0;
  for (int b = 0; b < nblk; b++) maxnb = std::max(maxnb, blk_start[b + 1] - blk_start[b]);
  const size_t lds_step = (size_t)2 * N * 8;
  if (lds_step > 64 * 1024)
    return mqct_fail(MQCT_ERR_ARG,
                     "propagate: the batched form keeps a stage vector of 2N doubles in LDS; N too large");
  // per position: energies, Coriolis neighbours (state-1 / state+1) and factors
  std::vector<double> E(N), sqv(2 * N);
  std::vector<int> nbv(2 * N), mpos(N);
  for (int p = 0; p < N; p++) {
    const int s = pos_state[p];
    E[p] = ds->E[s];
    mpos[p] = ds->m12[s];
    const double j = ds->j12[s], m = ds->m12[s];
    const double m1 = m + 1.0, m0 = m - 1.0;
    sqv[2 * p] = std::sqrt(j * (j + 1.0) - m0 * (m0 + 1.0));
    sqv[2 * p + 1] = std::sqrt(j * (j + 1.0) - m1 * (m1 - 1.0));
    nbv[2 * p] = s > 0 ? state_pos[s - 1] : -1;
    nbv[2 * p + 1] = s < N - 1 ? state_pos[s + 1] : -1;
  }
  const int n2 = 2 * N, ns = ccm ? n2 + 8 : n2, sw = ccm ? 4 : 2;
  const long long bcap = ccm ? 2LL * n_traj : n_traj;
  DevBuf d_ps, d_sp, d_bs, d_bo, d_pb, d_E, d_nb, d_sq, d_R, d_tr, d_mp;
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
    if (b.ensure(bytes)) return 1;
    return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, h->stream) != hipSuccess;
  };
  std::vector<double> Rmu(n_r + 1);
  Rmu[0] = ds->mu;
  for (int i = 0; i < n_r; i++) Rmu[i + 1] = ds->R[i];
  if (up(d_ps, pos_state.data(), N * 4) || up(d_sp, state_pos.data(), N * 4) ||
      up(d_bs, blk_start.data(), blk_start.size() * 4) || up(d_bo, blk_off.data(), blk_off.size() * 8) ||
      up(d_pb, pos_blk.data(), N * 4) || up(d_E, E.data(), N * 8) || up(d_nb, nbv.data(), N * 8) ||
      up(d_sq, sqv.data(), N * 16) || up(d_mp, mpos.data(), N * 4) ||
      up(d_R, Rmu.data(), (n_r + 1) * 8) || up(d_tr, traj, (size_t)n_traj * sizeof(mqct_traj_t)))
    return mqct_fail(MQCT_ERR_NOMEM, "batched: tables");
  // knot matrices: made once per matrix and storage precision
  const size_t need_km = (size_t)n_r * 2 * S2 * (knots_fp32 ? 4 : 8);
  if (!ds->km_valid || ds->km_fp32 != knots_fp32) {
    if (ds->d_KM.ensure(need_km)) return mqct_fail(MQCT_ERR_NOMEM, "batched: knot matrices");
    HIP_TRY(hipMemsetAsync(ds->d_KM.p, 0, need_km, h->stream));
    const unsigned g = (unsigned)((T * n_r + 255) / 256);
    if (knots_fp32)
      knot_matrices_kernel<float><<<g, 256, 0, h->stream>>>(
          T, n_r, ds->d_Mt.as<double>(), ds->d_Y2t.as<double>(), ds->d_pi.as<int>(), ds->d_pj.as<int>(),
          d_sp.as<int>(), d_pb.as<int>(), d_bs.as<int>(), d_bo.as<long long>(), S2, ds->d_KM.as<float>());
    else
      knot_matrices_kernel<double><<<g, 256, 0, h->stream>>>(
          T, n_r, ds->d_Mt.as<double>(), ds->d_Y2t.as<double>(), ds->d_pi.as<int>(), ds->d_pj.as<int>(),
          d_sp.as<int>(), d_pb.as<int>(), d_bs.as<int>(), d_bo.as<long long>(), S2, ds->d_KM.as<double>());
    HIP_TRY(hipGetLastError());
    ds->km_valid = true;
    ds->km_fp32 = knots_fp32;
  }
  ds->bat_S2 = S2;
  // the fragment-ordered tiles: 16 rows of a block each, [Y | Y2] with both
  // halves padded to a multiple of 4U columns
  const int KBh = knots_fp32 ? 16 : 8;
  std::vector<int> tile_blk, tile_r0, blk_tile0(nblk);
  std::vector<long long> tile_off;
  long long KS = 0;
  for (int b = 0; b < nblk; b++) {
    const int nb = blk_start[b + 1] - blk_start[b], nbp = (nb + KBh - 1) / KBh * KBh;
    blk_tile0[b] = (int)tile_blk.size();
    for (int r0 = 0; r0 < nb; r0 += 16) {
      tile_blk.push_back(b);
      tile_r0.push_back(r0);
      tile_off.push_back(KS);
      KS += 16LL * 2 * nbp;
    }
  }
  const int ntile = (int)tile_blk.size();
  // row groups of the pipelined forms: up to 128 rows (8 tiles) of a block;
  // handed out by decreasing cost rows x nb
  std::vector<int> pg_blk, pg_t0;
  std::vector<long long> pg_cost;
  for (int b = 0; b < nblk; b++) {
    const int nb = blk_start[b + 1] - blk_start[b];
    for (int t0 = 0; 16 * t0 < nb; t0 += 8) {
      pg_blk.push_back(b);
      pg_t0.push_back(t0);
      pg_cost.push_back((long long)std::min(128, nb - 16 * t0) * nb);
    }
  }
  const int npg = (int)pg_blk.size();
  std::vector<int> blk_order(npg);
  for (int g = 0; g < npg; g++) blk_order[g] = g;
  std::stable_sort(blk_order.begin(), blk_order.end(), [&](int x, int y) { return pg_cost[x] > pg_cost[y]; });
  // 64-row groups of the blocks (the LDS-staged work queue)
  std::vector<int> rg_blk, rg_row0;
  for (int b = 0; b < nblk; b++)
    for (int r0 = 0; r0 < blk_start[b + 1] - blk_start[b]; r0 += 64) {
      rg_blk.push_back(b);
      rg_row0.push_back(r0);
    }
  const int nrg = (int)rg_blk.size();
  DevBuf d_tb, d_tr0, d_to, d_bt0, d_rgb, d_rgr, d_bord, d_pgb, d_pgt;
  if (up(d_tb, tile_blk.data(), ntile * 4) || up(d_tr0, tile_r0.data(), ntile * 4) ||
      up(d_to, tile_off.data(), ntile * 8) || up(d_bt0, blk_tile0.data(), nblk * 4) ||
      up(d_rgb, rg_blk.data(), nrg * 4) || up(d_rgr, rg_row0.data(), nrg * 4) ||
      up(d_bord, blk_order.data(), npg * 4) || up(d_pgb, pg_blk.data(), npg * 4) ||
      up(d_pgt, pg_t0.data(), npg * 4))
    return mqct_fail(MQCT_ERR_NOMEM, "batched: tile tables");
  if (!ds->kt_valid || ds->kt_fp32 != knots_fp32) {
    const size_t need_kt = (size_t)n_r * KS * (knots_fp32 ? 4 : 8);
    if (ds->d_KT.ensure(need_kt)) return mqct_fail(MQCT_ERR_NOMEM, "batched: knot tiles");
    HIP_TRY(hipMemsetAsync(ds->d_KT.p, 0, need_kt, h->stream));
    const unsigned g = (unsigned)((T * n_r + 255) / 256);
    if (knots_fp32)
      knot_tiles_kernel<float><<<g, 256, 0, h->stream>>>(
          T, n_r, ds->d_Mt.as<double>(), ds->d_Y2t.as<double>(), ds->d_pi.as<int>(), ds->d_pj.as<int>(),
          d_sp.as<int>(), d_pb.as<int>(), d_bs.as<int>(), d_bt0.as<int>(), d_to.as<long long>(), KS,
          ds->d_KT.as<float>());
    else
      knot_tiles_kernel<double><<<g, 256, 0, h->stream>>>(
          T, n_r, ds->d_Mt.as<double>(), ds->d_Y2t.as<double>(), ds->d_pi.as<int>(), ds->d_pj.as<int>(),
          d_sp.as<int>(), d_pb.as<int>(), d_bs.as<int>(), d_bt0.as<int>(), d_to.as<long long>(), KS,
          ds->d_KT.as<double>());
    HIP_TRY(hipGetLastError());
    ds->kt_valid = true;
    ds->kt_fp32 = knots_fp32;
  }
  // trajectory state: slots, rotated amplitudes, GEMM outputs, controls,
  // double-buffered buckets, counters
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t off_V = al((size_t)n_traj * NSLOT * ns * 8);
  const size_t off_W = off_V + al((size_t)n_traj * N * 16);
  const size_t off_C = off_W + al((size_t)n_traj * sw * N * 16);
  const size_t off_CC = off_C + al((size_t)n_traj * sizeof(Ctl));
  const size_t off_cnt = off_CC + (ccm ? al((size_t)n_traj * sizeof(CtlCC)) : 0);
  const size_t off_bk = off_cnt + al((size_t)2 * n_r * 4);
  const size_t off_act = off_bk + al((size_t)2 * n_r * bcap * sizeof(Entry));
  const size_t off_al = off_act + 256;
  const size_t off_kl = off_al + 2 * al((size_t)n_traj * 4);
  const size_t total = off_kl + al((size_t)2 * n_r * 4);
  if (ds->d_bat.ensure(total)) return mqct_fail(MQCT_ERR_NOMEM, "batched: trajectory state");
  char *base = ds->d_bat.as<char>();
  BDev d;
  d.N = N;
  d.nblk = nblk;
  d.n_r = n_r;
  d.n_traj = n_traj;
  d.mode = mode;
  d.sw = sw;
  d.bcap = (int)bcap;
  d.theta_on = ds->theta_on;
  d.traj = d_tr.as<mqct_traj_t>();
  d.cc = ccm ? reinterpret_cast<CtlCC *>(base + off_CC) : nullptr;
  d.m_pos = d_mp.as<int>();
  d.R = d_R.as<double>() + 1;
  d.KM = ds->d_KM.p;
  d.S2 = S2;
  d.blk_start = d_bs.as<int>();
  d.blk_off = d_bo.as<long long>();
  d.KT = ds->d_KT.p;
  d.KS = KS;
  d.ntile = ntile;
  d.tile_blk = d_tb.as<int>();
  d.tile_r0 = d_tr0.as<int>();
  d.blk_tile0 = d_bt0.as<int>();
  d.blk_order = d_bord.as<int>();
  d.pg_blk = d_pgb.as<int>();
  d.pg_t0 = d_pgt.as<int>();
  d.npg = npg;
  d.tile_off = d_to.as<long long>();
  d.nrg = nrg;
  d.rg_blk = d_rgb.as<int>();
  d.rg_row0 = d_rgr.as<int>();
  d.E = d_E.as<double>();
  d.nb = d_nb.as<int2>();
  d.sq = d_sq.as<double2>();
  d.prec = ds->d_prec.as<double>();
  d.pder = ds->d_pder.as<double>();
  d.poff = ds->d_poff.as<long long>();
  d.S = reinterpret_cast<double *>(base);
  d.V = reinterpret_cast<double2 *>(base + off_V);
  d.W = reinterpret_cast<double2 *>(base + off_W);
  d.ctl = reinterpret_cast<Ctl *>(base + off_C);
  d.cnt = reinterpret_cast<int *>(base + off_cnt);
  d.bucket = reinterpret_cast<Entry *>(base + off_bk);
  d.active = reinterpret_cast<int *>(base + off_act);
  d.nbucketed = ds->prof_on ? reinterpret_cast<unsigned long long *>(base + off_act + 8) : nullptr;
  d.ntouched = reinterpret_cast<unsigned long long *>(base + off_act + 16);
  d.qctr = reinterpret_cast<int *>(base + off_act + 24);
  d.alist = reinterpret_cast<int *>(base + off_al);
  d.alist_tmp = reinterpret_cast<int *>(base + off_al + al((size_t)n_traj * 4));
  d.klist = reinterpret_cast<int *>(base + off_kl);
  d.kcnt = reinterpret_cast<int *>(base + off_act + 32);
  HIP_TRY(hipMemsetAsync(d.cnt, 0, (size_t)2 * n_r * 4, h->stream));
  HIP_TRY(hipMemsetAsync(d.active, 0, 40, h->stream));
  HIP_TRY(hipMemcpyAsync(d.active, &n_traj, 4, hipMemcpyHostToDevice, h->stream));
  // states per thread of the step kernel
  const int PT = N <= 4 * BS ? 4 : (N <= 8 * BS ? 8 : MAXPT);
  // step kernel: 256 threads per trajectory, or 512 (half the states per
  // thread, twice the loads in flight) when the states fill them
  const int stb = ds->step_threads > 0 ? ds->step_threads : (N > 4 * BS ? 512 : 256);
  if (stb != 256 && stb != 512 && stb != 1024)
    return mqct_fail(MQCT_ERR_ARG, "batched: step threads must be 256, 512 or 1024");
  if (stb == 1024 && N > 2 * 1024)
    return mqct_fail(MQCT_ERR_ARG, "batched: 1024 step threads take at most 2048 states");
  // the step kernel reads FP32 products after the FP32-product GEMM forms
  auto w32_of = [&](int sh) -> int { return prod_fp32 && sh >= 0 && pipelined(sh) ? 1 : 0; };
  // step launches: one workgroup per entry of the trajectory list; `grid` is
  // the list's length (the init kernel lists every trajectory, and the list
  // is compacted to the running ones when fewer than half of it still run)
  auto step = [&](int par, int w32, int grid) {
    const int sg = std::max(1, std::min(n_traj, grid));
    auto go = [&](auto listc) {
      constexpr bool LI = decltype(listc)::value;
      if (ccm) {
        if (stb >= 512) {  // the CC step kernel runs 512 threads at most
          if (N <= 4 * 512) cc_step_kernel<4, 512, LI><<<sg, 512, lds_step, h->stream>>>(d, par, w32);
          else cc_step_kernel<8, 512, LI><<<sg, 512, lds_step, h->stream>>>(d, par, w32);
        } else if (PT == 4) {
          cc_step_kernel<4, BS, LI><<<sg, BS, lds_step, h->stream>>>(d, par, w32);
        } else if (PT == 8) {
          cc_step_kernel<8, BS, LI><<<sg, BS, lds_step, h->stream>>>(d, par, w32);
        } else {
          cc_step_kernel<MAXPT, BS, LI><<<sg, BS, lds_step, h->stream>>>(d, par, w32);
        }
        return;
      }
      // 1024 threads (two states per thread) when asked for, or by the
      // automatic choice once fewer than 256 trajectories run (there it
      // shortens each workgroup's run; with a full grid 512 is faster)
      // (with 1024 threads wave 0 runs the control chain when the other
      // waves hold every state; at 512 the 448-thread state stride costs
      // registers that the full grid's occupancy cannot spare)
      if (stb == 1024 || (ds->step_threads == 0 && stb == 512 && N <= 2 * 1024 && grid < 256)) {
        if (N <= 2 * (1024 - 64)) step_kernel<2, 1024, LI, true><<<sg, 1024, lds_step, h->stream>>>(d, par, w32);
        else step_kernel<2, 1024, LI><<<sg, 1024, lds_step, h->stream>>>(d, par, w32);
        return;
      }
      if (stb == 512) {
        if (N <= 4 * 512) step_kernel<4, 512, LI><<<sg, 512, lds_step, h->stream>>>(d, par, w32);
        else step_kernel<8, 512, LI><<<sg, 512, lds_step, h->stream>>>(d, par, w32);
        return;
      }
      if (PT == 4) step_kernel<4, BS, LI><<<sg, BS, lds_step, h->stream>>>(d, par, w32);
      else if (PT == 8) step_kernel<8, BS, LI><<<sg, BS, lds_step, h->stream>>>(d, par, w32);
      else step_kernel<MAXPT, BS, LI><<<sg, BS, lds_step, h->stream>>>(d, par, w32);
    };
    if (grid >= n_traj) go(std::false_type{});  // never compacted: workgroup b takes trajectory b
    else go(std::true_type{});
  };

  if (ccm) {
    if (PT == 4) cc_init_kernel<4><<<n_traj, BS, 0, h->stream>>>(d, d_sp.as<int>());
    else if (PT == 8) cc_init_kernel<8><<<n_traj, BS, 0, h->stream>>>(d, d_sp.as<int>());
    else cc_init_kernel<MAXPT><<<n_traj, BS, 0, h->stream>>>(d, d_sp.as<int>());
  } else if (PT == 4) {
    init_kernel<4><<<n_traj, BS, 0, h->stream>>>(d, d_tr.as<mqct_traj_t>(), d_sp.as<int>());
  } else if (PT == 8) {
    init_kernel<8><<<n_traj, BS, 0, h->stream>>>(d, d_tr.as<mqct_traj_t>(), d_sp.as<int>());
  } else {
    init_kernel<MAXPT><<<n_traj, BS, 0, h->stream>>>(d, d_tr.as<mqct_traj_t>(), d_sp.as<int>());
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  const dim3 kb(n_r, nblk);
  // tile shapes are kept per product form, for AT (0..2) and CC (3..5)
  const int pi_ = (prod_fp32 ? 2 : (knots_fp32 ? 1 : 0)) + (ccm ? 3 : 0);
  ds->last_form = pi_;
  int n_list = n_traj;  // the step grid (the trajectory list's length; see compact below)
  // (the knot-list form launches two workgroup columns per trajectory of the
  // step grid: each running trajectory lists at most two knots)
  auto gemm = [&](int sh, int par) {
    const int kx = std::max(1, (int)std::min<long long>(n_r, 2LL * n_list));
    if (knots_fp32) launch_gemm<float>(sh, kb, maxnb, h->stream, d, par, kx, prod_fp32);
    else launch_gemm<double>(sh, kb, maxnb, h->stream, d, par, kx);
  };
  // iterations per graph replay (even: the bucket parity pattern repeats per
  // replay); more per replay once few trajectories run and an iteration is
  // short, so the host's read of the running count between replays costs
  // less (iterations after the last trajectory ends find empty lists)
  const int ITER = 16, ITER_LOW = 64, LOW_ACTIVE = 512;
  const bool prof = ds->prof_on;
  int active = n_traj;
  long long it_done = 0;  // iterations run so far (parity of the next one)
  if (ds->prof_ev.empty()) {
    ds->prof_ev.resize(3 * ITER);
    for (auto &e : ds->prof_ev) HIP_TRY(hipEventCreate(&e));
  }
  // n_it iterations launched one by one with HIP events on this stream
  // around every GEMM and step launch; returns the summed GEMM time.  In
  // profiled runs the algorithmic work of the chunk is added up from the
  // device counts of bucketed right-hand sides and touched knots:
  // per bucketed RHS 2 sides x 2 matrices x nb x nb x (Re, Im)
  // multiply-adds summed over the blocks = 16 S2 flops, its rotated
  // amplitudes read for both entries (2 x 16 N bytes) and both partial
  // products written (2 x 16 N bytes, or 2 x 8 N as FP32 products); each
  // touched knot's matrices are read once (2 S2 values)
  unsigned long long prof_seen[2] = {0, 0};
  // the step grid: every trajectory, one workgroup each, until fewer than
  // half of them run; then the list of the running ones, compacted again
  // whenever fewer than half of it still run (the list entry is a scalar
  // load, so a list pays from the first halving: the steps of 512 to 2047
  // running trajectories measured 8 to 16 % faster than with the full grid)
  const int COMPACT_BELOW = 1 << 30;
  auto compact = [&]() -> int {
    if (active > 0 && active < COMPACT_BELOW && 2 * active < n_list) {
      compact_kernel<<<1, 1024, 0, h->stream>>>(d.ctl, d.alist, d.alist_tmp, n_list);
      HIP_TRY(hipGetLastError());
      n_list = active;
    }
    return MQCT_OK;
  };
  auto timed_iters = [&](int sh, int n_it, double &gemm_ms) -> int {
    if (int rc = compact()) return rc;
    const int active0 = active;
    unsigned long long cnts[2] = {0, 0};
    for (int it = 0; it < n_it; it++) {
      const int par = (int)((it_done + it) & 1);
      HIP_TRY(hipEventRecord(ds->prof_ev[3 * it], h->stream));
      gemm(sh, par);
      HIP_TRY(hipEventRecord(ds->prof_ev[3 * it + 1], h->stream));
      step(par ^ 1, w32_of(sh), n_list);
      HIP_TRY(hipEventRecord(ds->prof_ev[3 * it + 2], h->stream));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&active, d.active, 4, hipMemcpyDeviceToHost, h->stream));
    if (prof) HIP_TRY(hipMemcpyAsync(cnts, base + off_act + 8, 16, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    gemm_ms = 0.0;
    double step_ms = 0.0;
    for (int it = 0; it < n_it; it++) {
      float ems = 0.f, sms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ems, ds->prof_ev[3 * it], ds->prof_ev[3 * it + 1]));
      HIP_TRY(hipEventElapsedTime(&sms, ds->prof_ev[3 * it + 1], ds->prof_ev[3 * it + 2]));
      gemm_ms += ems;
      step_ms += sms;
    }
    it_done += n_it;
    if (prof) {
      const double nb_rhs = (double)(cnts[0] - prof_seen[0]), n_knots = (double)(cnts[1] - prof_seen[1]);
      prof_seen[0] = cnts[0];
      prof_seen[1] = cnts[1];
      ds->prof_flops += 16.0 * (double)S2 * nb_rhs;
      ds->prof_bytes += (32.0 + (w32_of(sh) ? 16.0 : 32.0)) * N * nb_rhs +
                        2.0 * (double)S2 * (knots_fp32 ? 4 : 8) * n_knots;
      ds->prof_ms += gemm_ms;
      ds->prof_launches += n_it;
      if (ds->prof_trace.size() < (1u << 22))
        ds->prof_trace.insert(ds->prof_trace.end(),
                              {(double)active0, (double)n_it, gemm_ms, (double)sh, step_ms});
    }
    return MQCT_OK;
  };
  // the tile shape for the current number of running trajectories: the
  // bucket fill, and with it the best shape, changes as trajectories end, so
  // shapes are chosen per level floor(log2(active)) -- on the first
  // iterations at that level (real iterations of this run): three rounds of
  // two timed iterations per shape, the fastest total wins and is kept for
  // later runs on this matrix
  auto level_of = [](int a) {
    int l = 0;
    while (l + 1 < DynState::kLevels && (2LL << l) <= a) l++;
    return l;
  };
  // (the work-queue shapes keep the chunk prefix of all knots in LDS)
  const bool queue_ok = (size_t)(n_r + 1) * 4 <= 64 * 1024;
  if (ds->gemm_fixed >= 0 && uses_queue(ds->gemm_fixed) && !queue_ok)
    return mqct_fail(MQCT_ERR_ARG, "batched: the work-queue GEMM shapes take at most 16383 knots");
  // shapes set per level (the tuner's own choices always fit)
  for (int l = 0; l < DynState::kLevels; l++) {
    const int sh = ds->gemm_shape[pi_][l];
    if (sh >= 0 && uses_queue(sh) && !queue_ok)
      return mqct_fail(MQCT_ERR_ARG, "batched: a GEMM shape set for a level does not fit this matrix");
  }
  auto shape_for = [&](int &sh) -> int {
    for (;;) {
      if (ds->gemm_fixed >= 0) {
        sh = ds->gemm_fixed;
        return MQCT_OK;
      }
      const int lv = level_of(active);
      if (ds->gemm_shape[pi_][lv] >= 0 || active == 0) {
        sh = ds->gemm_shape[pi_][lv];
        return MQCT_OK;
      }
      // candidates more than 1.5x slower than the best after the first
      // round are not timed again
      const int ncand = kNShapes;
      double tot[kNShapes] = {};
      bool keep[kNShapes];
      for (int s = 0; s < kNShapes; s++)
        keep[s] = queue_ok || !uses_queue(s);
      for (int round = 0; round < 3 && active > 0; round++) {
        for (int s = 0; s < ncand && active > 0; s++) {
          if (!keep[s]) continue;
          double gm = 0.0;
          if (int rc = timed_iters(s, 2, gm)) return rc;
          tot[s] += gm;
        }
        double best_t = 1e300;
        for (int s = 0; s < ncand; s++)
          if (keep[s]) best_t = std::min(best_t, tot[s]);
        for (int s = 0; s < ncand; s++)
          if (keep[s] && tot[s] > 1.5 * best_t) keep[s] = false;
      }
      int best = -1;
      for (int s = 0; s < ncand; s++)
        if (keep[s] && (best < 0 || tot[s] < tot[best])) best = s;
      ds->gemm_shape[pi_][lv] = best < 0 ? 0 : best;
      // loop: the level is read again (trajectories ended while timing)
    }
  };
  // every trajectory ends after at most MAXSTP accepted steps and a bounded
  // number of step-size reductions; the cap only guards the host loop
  const int max_rounds = 1 << 20;  // a safety bound: at least 16.7 M iterations
  HIP_TRY(hipEventRecord(h->ev[5], h->stream));
  int shape = -1;
  if (int rc = shape_for(shape)) return rc;
  ds->gemm_first[pi_] = shape;
  // ITER (or ITER_LOW) iterations (GEMM, step) captured per shape, bucket parity and step grid,
  // replayed until every trajectory has ended; without graphs (profiling)
  // the same iterations are launched one by one between HIP events
  struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    int shape = -1, par = -1, grid = 0, iters = 0;
    void reset() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
      e = nullptr;
      g = nullptr;
      shape = par = -1;
    }
    ~GraphGuard() { reset(); }
  } gg;
  auto t_prog = std::chrono::steady_clock::now();
  const auto t_start = t_prog;
  for (int rep = 0; rep < max_rounds && active > 0; rep++) {
    if (h->progress_s > 0.0) {
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - t_prog).count() >= h->progress_s) {
        std::fprintf(stderr, "[mqct] batched propagation: %d of %d trajectories running after %lld iterations, %.0f s\n",
                     active, n_traj, (long long)it_done,
                     std::chrono::duration<double>(now - t_start).count());
        std::fflush(stderr);
        t_prog = now;
      }
    }
    if (int rc = shape_for(shape)) return rc;
    if (active == 0) break;
    if (prof) {
      double gm = 0.0;
      if (int rc = timed_iters(shape, ITER, gm)) return rc;
      continue;
    }
    const int p0 = (int)(it_done & 1);
    // captured again when the shape, parity, step grid or replay length changes
    const int n_it = active < LOW_ACTIVE ? ITER_LOW : ITER;
    if (int rc = compact()) return rc;
    if (gg.shape != shape || gg.par != p0 || gg.grid != n_list || gg.iters != n_it) {
      gg.reset();
      gg.grid = n_list;
      gg.iters = n_it;
      HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      for (int it = 0; it < n_it; it++) {
        const int par = (it + p0) & 1;
        gemm(shape, par);
        step(par ^ 1, w32_of(shape), gg.grid);
      }
      const hipError_t cap = hipStreamEndCapture(h->stream, &gg.g);
      HIP_TRY(cap);
      HIP_TRY(hipGraphInstantiate(&gg.e, gg.g, nullptr, nullptr, 0));
      gg.shape = shape;
      gg.par = p0;
    }
    HIP_TRY(hipGraphLaunch(gg.e, h->stream));
    it_done += n_it;
    HIP_TRY(hipMemcpyAsync(&active, d.active, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  HIP_TRY(hipEventRecord(h->ev[6], h->stream));
  HIP_TRY(hipEventSynchronize(h->ev[6]));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev[5], h->ev[6]));
  ds->last_ms = ms;
  if (active > 0) return mqct_fail(MQCT_ERR_HIP, "batched: trajectories did not finish");
  if (ds->d_q.ensure((size_t)n_traj * (n2 + 8) * 8) ||
      ds->d_stat.ensure((size_t)n_traj * sizeof(mqct_traj_stat_t)))
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
  ds->prof_trace.clear();
  ds->prof_launches = 0;
  ds->prof_ms = 0.0;
  ds->prof_flops = 0.0;
  ds->prof_bytes = 0.0;
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_prof_trace(mqct_hip_t *h, int cap, double *rows, int *n) {
  if (!h || !h->dyn || !n || cap < 0 || (cap > 0 && !rows)) return mqct_fail(MQCT_ERR_ARG, "batched_prof_trace: bad argument");
  const auto &tr = h->dyn->prof_trace;
  *n = (int)(tr.size() / 5);
  const size_t m = std::min((size_t)cap, tr.size() / 5) * 5;
  std::copy(tr.begin(), tr.begin() + m, rows);
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_set_step_threads(mqct_hip_t *h, int threads) {
  if (!h || !h->dyn) return mqct_fail(MQCT_ERR_STATE, "batched_set_step_threads: no matrix loaded");
  if (threads != 0 && threads != 256 && threads != 512 && threads != 1024)
    return mqct_fail(MQCT_ERR_ARG, "batched_set_step_threads: 0, 256, 512 or 1024");
  h->dyn->step_threads = threads;
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_set_tile(mqct_hip_t *h, int shape) {
  if (!h || !h->dyn) return mqct_fail(MQCT_ERR_STATE, "batched_set_tile: no matrix loaded");
  if (shape < -1 || shape >= kNShapes) return mqct_fail(MQCT_ERR_ARG, "batched_set_tile: unknown shape");
  DynState *ds = h->dyn;
  ds->gemm_fixed = shape;
  if (shape < 0)
    for (auto &p : ds->gemm_shape)
      for (int &v : p) v = -1;
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_tile(mqct_hip_t *h, int shape, int *form, int *entries, int *k_rows,
                                     int *row_tiles) {
  if (!h || !form || !entries || !k_rows || !row_tiles || shape < -1 || shape >= kNShapes)
    return mqct_fail(MQCT_ERR_ARG, "batched_tile: bad argument");
  int sh = shape;
  if (sh < 0) {
    if (!h->dyn) return mqct_fail(MQCT_ERR_STATE, "batched_tile: no matrix loaded");
    sh = h->dyn->gemm_first[h->dyn->last_form];
    if (sh < 0) return mqct_fail(MQCT_ERR_STATE, "batched_tile: no batched run yet");
  }
  *form = kShapes[sh].form;
  *entries = kShapes[sh].ge;
  *k_rows = kShapes[sh].gk;
  *row_tiles = kShapes[sh].grt;
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_tile_levels(mqct_hip_t *h, int fp32, int n, int *shapes) {
  if (!h || !h->dyn || !shapes || n < 0 || fp32 < 0 || fp32 >= DynState::kForms)
    return mqct_fail(MQCT_ERR_ARG, "batched_tile_levels: bad argument");
  for (int l = 0; l < n; l++) shapes[l] = l < DynState::kLevels ? h->dyn->gemm_shape[fp32][l] : -1;
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_set_tile_levels(mqct_hip_t *h, int fp32, int n, const int *shapes) {
  if (!h || !h->dyn) return mqct_fail(MQCT_ERR_STATE, "batched_set_tile_levels: no matrix loaded");
  if (!shapes || n < 0 || n > DynState::kLevels || fp32 < 0 || fp32 >= DynState::kForms)
    return mqct_fail(MQCT_ERR_ARG, "batched_set_tile_levels: bad argument");
  for (int l = 0; l < n; l++)
    if (shapes[l] < -1 || shapes[l] >= kNShapes) return mqct_fail(MQCT_ERR_ARG, "batched_set_tile_levels: unknown shape");
  for (int l = 0; l < n; l++) h->dyn->gemm_shape[fp32][l] = shapes[l];
  return MQCT_OK;
}

extern "C" int mqct_hip_batched_kernel_stats(mqct_hip_t *h, int64_t *launches, double *gemm_ms,
                                             double *gemm_flops, double *gemm_bytes) {
  if (!h || !h->dyn || !launches || !gemm_ms || !gemm_flops || !gemm_bytes)
    return mqct_fail(MQCT_ERR_ARG, "batched_kernel_stats: bad argument");
  *launches = h->dyn->prof_launches;
  *gemm_ms = h->dyn->prof_ms;
  *gemm_flops = h->dyn->prof_flops;
  *gemm_bytes = h->dyn->prof_bytes;
  return MQCT_OK;
}
