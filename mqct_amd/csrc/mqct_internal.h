// mqct_internal.h -- shared host/device definitions of libmqct_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/mqct_hip.h"

#define MQCT_PI 3.14159265358979323846264338327950288

// ---------------------------------------------------------------------------
// Error plumbing: every entry point returns a status and leaves the message
// in a thread-local buffer read by mqct_hip_last_error.
void mqct_set_error(const std::string &msg);
int mqct_fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                         \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return mqct_fail(MQCT_ERR_HIP, std::string(#expr) + ": " +              \
                                         hipGetErrorString(_e));              \
  } while (0)

// Device buffer owned by a handle.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  int ensure(size_t n);  // grows (re-allocates) when too small
  void release();
  template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Output tile of the contraction: a block of a-states x b-states of one
// m-block (positions in the block-ordered basis), plus the first R point.
struct TileDesc {
  int a0, na, b0, nb;
};

struct MlpState;  // learned potential (mlp.hip)
struct DynState;  // propagation (dynamics.hip)

struct mqct_hip {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};

  // ---- basis ----
  int coll_type = -1;
  int n_states = 0, n_channels = 0, ldvec = 0;
  std::vector<int> j12, m12, indx_chann, j1_ch, j2_ch;
  std::vector<double> eig1, eig2;
  std::vector<int> pos_state;        // position -> state (0-based)
  std::vector<int> state_pos;        // state -> position
  std::vector<int> block_start;      // per m-block start position (+end)
  std::vector<int64_t> kbase;        // state -> index of its first pair
  int64_t total_pairs = 0;
  bool basis_ok = false;

  // ---- grid ----
  int n_beta = 0, n_gamma = 1, n_alpha2 = 1, n_beta2 = 1, n_gamma2 = 1;
  std::vector<double> beta, gamma, alpha2, beta2, gamma2;  // angles
  std::vector<double> omega;         // quadrature weight per grid point
  int64_t G = 0;
  int ref_xz = 1;
  bool grid_ok = false;

  // ---- potential ----
  int n_r = 0;
  std::vector<double> R;
  bool v_ok = false;

  // ---- device data of the build ----
  DevBuf d_j12, d_m12, d_pos_state, d_kbase, d_eig1, d_eig2;
  DevBuf d_chan;                     // [j per state][channel per state]
  DevBuf d_beta, d_gamma, d_alpha2, d_beta2, d_gamma2, d_omega;
  DevBuf d_R, d_V, d_par;
  DevBuf d_psi_re, d_psi_im, d_psib_re, d_psib_im;  // FP64 [pos][G]
  DevBuf d_psi32_re, d_psi32_im, d_psib32_re, d_psib32_im;
  DevBuf d_W64, d_W32, d_Whi, d_Wlo, d_Wlo2;  // [n_r][G]; BF16 terms of W
  DevBuf d_tab1, d_tab2;             // type 0 per-level one-molecule tables
  DevBuf d_tiles, d_ctiles;          // contraction tiles; 64 x 64 tiles of the cutoff sums
  DevBuf d_mat, d_der, d_skip, d_spl;  // outputs [T][n_r]
  DevBuf d_Wcut, d_cutv, d_cutp;     // FP64 cutoff columns (mixed precision), per-slice partial sums
  int64_t built_T = 0;
  bool built = false;
  bool psi_fresh = false;
  double timing[5] = {0, 0, 0, 0, 0};
  // contraction variant per precision form: chosen by timing (-1 not yet,
  // re-timed when the basis or grid changes) or fixed by the caller
  int cvar[3] = {-1, -1, -1}, cvar_fixed[3] = {-1, -1, -1};
  double tune_ms = 0.0;  // time the last build spent choosing the variant
  // MIJ_SHIFT= keyword (default YES, iotest.f:3754): the propagation sees
  // M(R) - M(R_max) (matrix_ini.f:2226-2230, 6603-6608)
  bool mij_shift = true;
  // progress lines of long batched propagations on stderr (0: none)
  double progress_s = 0.0;

  // ---- identical partners (type 0): IDENTICAL= / IDENTICAL_PES= ----
  bool identical = false, identical_pes = false;
  std::vector<int> parity_state, parity_state_bk, same_level;
  DevBuf d_xpar;                     // [parity_state][parity_state_bk] per state
  DevBuf d_same;                     // per channel: both partners in one level

  // ---- PES expansion (expansion.hip) ----
  int n_terms = 0;
  std::vector<int> atop;             // A_TOP rows, [n_terms][5]
  DevBuf d_atop, d_expch, d_cgt;     // terms, per-channel j1/j2, CG table
  DevBuf d_terms;                    // expansion_terms, [n_terms][n_r]
  DevBuf d_expA;                     // angular coefficients, [T][n_terms (padded)]
  DevBuf d_expF, d_expT1, d_expT2, d_expcgl, d_exppart, d_exppairs;
  int terms_nr = 0;                  // R points of d_terms (0: none)
  double exp_timing[4] = {0, 0, 0, 0};

  MlpState *mlp = nullptr;
  DynState *dyn = nullptr;
};

// ---------------------------------------------------------------------------
// Device math shared by the basis kernels (restating the reference's own
// formulas: pot_comp.f:256-331, suppl_routines.f:1-148, 207-314,
// basis_wave_functions.f:83-175).

__host__ __device__ inline double mq_logfact(int n) {
  // log n! by summation; n <= 4*jmax+2 is small.
  double s = 0.0;
  for (int i = 2; i <= n; i++) s += log((double)i);
  return s;
}

__host__ __device__ inline double mq_ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

// Wigner d^j_{k m}(beta) in MQCT's Jacobi-polynomial form (pot_comp.f:256-309).
__host__ __device__ inline double mq_wigner_d(int j, int k, int m, double cb,
                                              double sh, double ch) {
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
  double al = (double)a, be = (double)b;
  double p0 = 1.0, p1 = 0.0, pn = 1.0;
  if (n >= 1) {
    p1 = (1.0 + 0.5 * (al + be)) * cb + 0.5 * (al - be);
    pn = p1;
    for (int i = 2; i <= n; i++) {
      double ri = (double)i;
      double c1 = 2.0 * ri * (ri + al + be) * (2.0 * ri - 2.0 + al + be);
      double c2 = (2.0 * ri - 1.0 + al + be) * (2.0 * ri + al + be) *
                  (2.0 * ri - 2.0 + al + be);
      double c3 = (2.0 * ri - 1.0 + al + be) * (al + be) * (al - be);
      double c4 = -2.0 * (ri - 1.0 + al) * (ri - 1.0 + be) * (2.0 * ri + al + be);
      double p2 = ((c3 + c2 * cb) * p1 + c4 * p0) / c1;
      p0 = p1;
      p1 = p2;
      pn = p2;
    }
  }
  double sgn = ((lambda % 2) == 0) ? 1.0 : -1.0;
  // C(2j-n, n+a) / C(n+b, b) through log-factorials as C_N_K does
  double c_a = exp(mq_logfact(2 * j - n) - mq_logfact(2 * j - n - (n + a)) -
                   mq_logfact(n + a));
  double c_b = exp(mq_logfact(n + b) - mq_logfact(n) - mq_logfact(b));
  return pn * sgn * sqrt(c_a) / sqrt(c_b) * mq_ipow(sh, a) * mq_ipow(ch, b);
}

// Associated Legendre P_l^m(x), Condon-Shortley phase (suppl_routines.f:207-243).
__host__ __device__ inline double mq_plgndr(int l, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; i++) { pmm = -pmm * fact * somx2; fact += 2.0; }
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

// Clebsch-Gordan <j1 m1 j2 m2 | j3 m3> (host only; built into tables).
double mq_cg(int j1, int j2, int j3, int m1, int m2, int m3);

// Asynchronous host-to-device copy into a handle-owned buffer (grown as
// needed) on the handle's stream.
template <class T>
inline int upload(mqct_hip *h, DevBuf &b, const T *src, size_t n) {
  if (b.ensure(n * sizeof(T))) return mqct_fail(MQCT_ERR_NOMEM, "device alloc");
  if (n)
    HIP_TRY(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice,
                           h->stream));
  return MQCT_OK;
}

// R grid of the handle (>= 2 increasing points), host copy and device copy.
int mqct_set_r(mqct_hip *h, int n_r, const double *R);

// States s1 >= s2 (0-based) form a pair of the matrix: equal m12 and, for
// identical partners, equal exchange parity (matrix_ini.f:553-630).
inline bool mqct_pair_ok(const mqct_hip *h, int s1, int s2) {
  return h->m12[s1] == h->m12[s2] &&
         (!h->identical || h->parity_state[s1] == h->parity_state[s2]);
}

// Kernel launchers (defined in the .hip files).
int mqct_build_psi(mqct_hip *h);
int mqct_build_weights(mqct_hip *h, int precision);
int mqct_run_contraction(mqct_hip *h, int precision, double conv);
int mqct_run_post(mqct_hip *h, double cut, int ir1, int ir2,
                  const double *cutv);
int mqct_run_cutoff_fp64(mqct_hip *h, int ir1, int ir2, double conv);
void mqct_weight_rows(mqct_hip *h, const double *V, const double *omega,
                      int ir1, int ir2, double *out);
int mqct_eval_pes_dev(mqct_hip *h, int pes_kind, int n_par);
// Number of doubles the synthetic potential pes_kind reads from par (its
// header values decide the rest), or -1 if the header itself is invalid or
// n_par is too short for it.
long long mqct_pes_par_len(int pes_kind, int n_par, const double *par);
