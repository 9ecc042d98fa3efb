// batched.hip -- placeholder for the interval-bucketed batched AT form.
#include "dyn_internal.h"
int mqct_dyn_run_batched(mqct_hip *h, int mode, int n_traj, const mqct_traj_t *traj,
                         double *q_out, mqct_traj_stat_t *stat_out) {
  (void)h; (void)mode; (void)n_traj; (void)traj; (void)q_out; (void)stat_out;
  return mqct_fail(MQCT_ERR_ARG, "propagate: batched form not available in this build");
}
