// dyn_internal.h -- device-resident state of the propagation.
#pragma once
#include "mqct_internal.h"

struct DynState {
  int N = 0;
  long long T = 0;
  int n_r = 0;
  double mu = 0.0;
  int theta_on = 0;
  std::vector<int> j12, m12, pi, pj;
  std::vector<double> E, R;
  DevBuf d_R, d_Mt, d_Y2t, d_pi, d_pj, d_E, d_j12, d_m12, d_misc;
  DevBuf d_prec, d_pder, d_poff;
  int n_paths = 0;
  DevBuf d_scratch, d_trajp, d_q, d_stat, d_counter;
  // batched form
  DevBuf d_bat;      // trajectory states, buckets, controls
  DevBuf d_KM;       // dense per-knot block matrices [n_r][2][S2]
  long long bat_S2 = 0;
  bool km_valid = false;  // d_KM holds the current matrix
  bool km_fp32 = false;   // ... stored in FP32
  DevBuf d_KT;       // the same matrices in MFMA-fragment order (16-row tiles)
  bool kt_valid = false, kt_fp32 = false;
  // interval-GEMM tile shape per product form (FP64 knots, FP32 knots, FP32
  // knots with FP32 products; 0..2 for the AT modes, 3..5 for CC) and
  // level of the running-trajectory count (floor(log2(active))), -1 not yet
  // chosen; a shape fixed by the caller (>= 0) overrides them
  static constexpr int kForms = 6;
  static constexpr int kLevels = 24;
  int gemm_shape[kForms][kLevels];
  int gemm_fixed = -1;
  int step_threads = 0;  // batched step kernel: 0 automatic, 256 or 512
  int gemm_first[kForms] = {-1, -1, -1, -1, -1, -1};  // shape the last run started with
  int last_form = 0;                  // product form of the last run
  DynState() {
    for (auto &p : gemm_shape)
      for (int &v : p) v = -1;
  }
  // batched-form kernel timing
  bool prof_on = false;
  long long prof_launches = 0;
  double prof_ms = 0.0, prof_flops = 0.0, prof_bytes = 0.0;
  std::vector<hipEvent_t> prof_ev;
  std::vector<double> prof_trace;  // per timed chunk: running trajectories, iterations, GEMM ms, shape, step ms
  ~DynState() {
    for (auto e : prof_ev) (void)hipEventDestroy(e);
  }
  double last_ms = 0.0;
};

int mqct_dyn_run_plain(mqct_hip *h, int mode, int n_traj,
                       const mqct_traj_t *traj, double *q_out,
                       mqct_traj_stat_t *stat_out, int rec_cap = 0,
                       double *rec_out = nullptr, int *rec_n_out = nullptr);
int mqct_dyn_run_batched(mqct_hip *h, int mode, int n_traj,
                         const mqct_traj_t *traj, double *q_out,
                         mqct_traj_stat_t *stat_out, bool knots_fp32, bool prod_fp32 = false);
