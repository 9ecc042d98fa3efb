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
  DevBuf d_bat;  // see batched.hip
  double last_ms = 0.0;
};

int mqct_dyn_run_plain(mqct_hip *h, int mode, int n_traj,
                       const mqct_traj_t *traj, double *q_out,
                       mqct_traj_stat_t *stat_out, int rec_cap = 0,
                       double *rec_out = nullptr, int *rec_n_out = nullptr);
int mqct_dyn_run_batched(mqct_hip *h, int mode, int n_traj,
                         const mqct_traj_t *traj, double *q_out,
                         mqct_traj_stat_t *stat_out);
