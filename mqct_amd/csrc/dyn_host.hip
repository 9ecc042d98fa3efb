// dyn_host.hip -- C ABI of the propagation: spline-coefficient residency,
// recorded AT paths, and the propagate call.
#include "dyn_internal.h"

#include <cmath>
#include <cstring>

namespace {

// Mt[r][k] = M[k][r]: the R-major copy the right-hand side reads
// (bk_mat_splnt = transpose(Mat_el), mqct_eqtns.f:104-105).  shift: the
// compute branch's MIJ_SHIFT, Mat_el(i,k) - Mat_el(n_r,k) for every i
// (matrix_ini.f:2226-2230; the last R point becomes exactly 0).
__global__ void transpose_kernel(long long T, int n_r, int shift,
                                 const double *__restrict__ M,
                                 double *__restrict__ Mt) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= T * n_r) return;
  long long k = idx / n_r;
  int r = (int)(idx % n_r);
  double v = M[idx];
  if (shift) v = v - M[k * n_r + n_r - 1];
  Mt[(long long)r * T + k] = v;
}

template <class T>
int up(mqct_hip *h, DevBuf &b, const T *src, size_t n) {
  if (b.ensure(n * sizeof(T))) return mqct_fail(MQCT_ERR_NOMEM, "device alloc");
  if (n)
    HIP_TRY(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
  return MQCT_OK;
}

// Tables of resize (mqct_eqtns.f:1-58): pair indices and Coriolis factors.
int set_tables(mqct_hip *h, DynState *ds, const int *ind_mat) {
  ds->km_valid = false;
  ds->kt_valid = false;
  for (auto &p : ds->gemm_shape)
    for (int &v : p) v = -1;
  const long long T = ds->T;
  ds->pi.resize(T);
  ds->pj.resize(T);
  std::vector<double> misc(4 * T);
  for (long long k = 0; k < T; k++) {
    int i = ind_mat[2 * k] - 1, j = ind_mat[2 * k + 1] - 1;
    if (i < 0 || j < 0 || i >= ds->N || j >= ds->N)
      return mqct_fail(MQCT_ERR_ARG, "load_matrix: ind_mat out of range");
    ds->pi[k] = i;
    ds->pj[k] = j;
    int j_12 = ds->j12[j], m_12 = ds->m12[i];
    int m1 = m_12 + 1, m0 = m_12 - 1;
    misc[4 * k + 0] = std::sqrt(j_12 * (j_12 + 1.0) - m1 * (m1 - 1.0));
    misc[4 * k + 1] = std::sqrt(j_12 * (j_12 + 1.0) - m0 * (m0 + 1.0));
    j_12 = ds->j12[i];
    m_12 = ds->m12[j];
    m1 = m_12 + 1;
    m0 = m_12 - 1;
    misc[4 * k + 2] = std::sqrt(j_12 * (j_12 + 1.0) - m1 * (m1 - 1.0));
    misc[4 * k + 3] = std::sqrt(j_12 * (j_12 + 1.0) - m0 * (m0 + 1.0));
  }
  int rc;
  if ((rc = up(h, ds->d_pi, ds->pi.data(), T))) return rc;
  if ((rc = up(h, ds->d_pj, ds->pj.data(), T))) return rc;
  if ((rc = up(h, ds->d_misc, misc.data(), misc.size()))) return rc;
  if ((rc = up(h, ds->d_E, ds->E.data(), ds->N))) return rc;
  if ((rc = up(h, ds->d_j12, ds->j12.data(), ds->N))) return rc;
  if ((rc = up(h, ds->d_m12, ds->m12.data(), ds->N))) return rc;
  if ((rc = up(h, ds->d_R, ds->R.data(), ds->n_r))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}

int transpose_into(mqct_hip *h, DynState *ds, const double *d_src, DevBuf &dst,
                   bool shift = false) {
  const long long n = ds->T * ds->n_r;
  if (dst.ensure((size_t)n * 8)) return mqct_fail(MQCT_ERR_NOMEM, "spline coefficients");
  transpose_kernel<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(
      ds->T, ds->n_r, shift ? 1 : 0, d_src, dst.as<double>());
  HIP_TRY(hipGetLastError());
  return MQCT_OK;
}

}  // namespace

void mqct_dyn_free(mqct_hip *h) {
  delete h->dyn;
  h->dyn = nullptr;
}

extern "C" int mqct_hip_load_matrix(mqct_hip_t *h, int n_states, const double *Ej,
                                    const int *j12, const int *m12,
                                    int64_t total_size, const int *ind_mat,
                                    int n_r, const double *R_com,
                                    const double *mat_el,
                                    const double *mat_el_der, double mass,
                                    int theta_on) {
  if (!h || !Ej || !j12 || !m12 || !ind_mat || !R_com || !mat_el || !mat_el_der)
    return mqct_fail(MQCT_ERR_ARG, "load_matrix: null argument");
  if (n_states <= 0 || total_size <= 0 || n_r < 2 || !(mass > 0))
    return mqct_fail(MQCT_ERR_ARG, "load_matrix: bad sizes");
  HIP_TRY(hipSetDevice(h->device));
  if (!h->dyn) h->dyn = new DynState();
  DynState *ds = h->dyn;
  ds->N = n_states;
  ds->T = total_size;
  ds->n_r = n_r;
  ds->mu = mass;
  ds->theta_on = theta_on ? 1 : 0;
  ds->E.assign(Ej, Ej + n_states);
  ds->j12.assign(j12, j12 + n_states);
  ds->m12.assign(m12, m12 + n_states);
  ds->R.assign(R_com, R_com + n_r);
  int rc = set_tables(h, ds, ind_mat);
  if (rc) return rc;
  DevBuf tmp;
  if (tmp.ensure((size_t)total_size * n_r * 8)) return mqct_fail(MQCT_ERR_NOMEM, "load_matrix");
  HIP_TRY(hipMemcpy(tmp.p, mat_el, (size_t)total_size * n_r * 8, hipMemcpyHostToDevice));
  if ((rc = transpose_into(h, ds, tmp.as<double>(), ds->d_Mt))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(tmp.p, mat_el_der, (size_t)total_size * n_r * 8, hipMemcpyHostToDevice));
  if ((rc = transpose_into(h, ds, tmp.as<double>(), ds->d_Y2t))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}

extern "C" int mqct_hip_use_built_matrix(mqct_hip_t *h, const double *Ej, double mass,
                                         int theta_on) {
  if (!h || !Ej) return mqct_fail(MQCT_ERR_ARG, "use_built_matrix: null argument");
  if (!h->built) return mqct_fail(MQCT_ERR_STATE, "use_built_matrix: nothing built");
  HIP_TRY(hipSetDevice(h->device));
  if (!h->dyn) h->dyn = new DynState();
  DynState *ds = h->dyn;
  ds->N = h->n_states;
  ds->T = h->built_T;
  ds->n_r = h->n_r;
  ds->mu = mass;
  ds->theta_on = theta_on ? 1 : 0;
  ds->E.assign(Ej, Ej + h->n_states);
  ds->j12 = h->j12;
  ds->m12 = h->m12;
  ds->R = h->R;
  std::vector<int> ind(2 * (size_t)ds->T);
  long long k = 0;
  for (int s1 = 0; s1 < h->n_states; s1++)
    for (int s2 = 0; s2 <= s1; s2++)
      if (mqct_pair_ok(h, s1, s2)) {
        ind[2 * k] = s1 + 1;
        ind[2 * k + 1] = s2 + 1;
        k++;
      }
  int rc = set_tables(h, ds, ind.data());
  if (rc) return rc;
  // MIJ_SHIFT acts after the spline in the compute branch: the values are
  // shifted, their second derivatives are those of the unshifted values
  if ((rc = transpose_into(h, ds, h->d_mat.as<double>(), ds->d_Mt, h->mij_shift))) return rc;
  if ((rc = transpose_into(h, ds, h->d_der.as<double>(), ds->d_Y2t))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}

// Spline second derivatives of each path column with finite-difference end
// slopes (propagator.f:3012-3031; spline, suppl_routines.f:914-965).
static void path_spline(const double *rec, int n, int col, double *der7) {
  std::vector<double> x(n), y(n), y2(n), u(n);
  for (int i = 0; i < n; i++) {
    x[i] = rec[8 * i];
    y[i] = rec[8 * i + 1 + col];
  }
  double yp1 = (y[1] - y[0]) / (x[1] - x[0]);
  double ypn = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
  y2[0] = -0.5;
  u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
  for (int i = 1; i < n - 1; i++) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                (x[i + 1] - x[i - 1]) -
            sig * u[i - 1]) /
           p;
  }
  double qn = 0.5;
  double un = (3.0 / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
  y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
  for (int i = 0; i < n; i++) der7[7 * i + col] = y2[i];
}

extern "C" int mqct_hip_load_paths(mqct_hip_t *h, int n_paths, const int64_t *path_off,
                                   const double *records) {
  if (!h || !path_off || !records || n_paths <= 0)
    return mqct_fail(MQCT_ERR_ARG, "load_paths: bad argument");
  if (!h->dyn) return mqct_fail(MQCT_ERR_STATE, "load_paths: load the matrix first");
  HIP_TRY(hipSetDevice(h->device));
  DynState *ds = h->dyn;
  int64_t nrec = path_off[n_paths];
  std::vector<double> der(7 * (size_t)nrec);
  for (int p = 0; p < n_paths; p++) {
    int64_t o0 = path_off[p], o1 = path_off[p + 1];
    if (o1 - o0 < 4) return mqct_fail(MQCT_ERR_ARG, "load_paths: a path needs >= 4 records");
    for (int64_t i = o0 + 1; i < o1; i++)
      if (!(records[8 * i] > records[8 * (i - 1)]))
        return mqct_fail(MQCT_ERR_ARG, "load_paths: path times must increase");
    for (int c = 0; c < 7; c++) path_spline(records + 8 * o0, (int)(o1 - o0), c, der.data() + 7 * o0);
  }
  std::vector<long long> off(path_off, path_off + n_paths + 1);
  int rc;
  if ((rc = up(h, ds->d_prec, records, 8 * (size_t)nrec))) return rc;
  if ((rc = up(h, ds->d_pder, der.data(), der.size()))) return rc;
  if ((rc = up(h, ds->d_poff, off.data(), off.size()))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  ds->n_paths = n_paths;
  return MQCT_OK;
}

extern "C" int mqct_hip_propagate(mqct_hip_t *h, int mode, int n_traj,
                                  const mqct_traj_t *traj, int batched,
                                  double *q_out, mqct_traj_stat_t *stat_out) {
  if (!h || !traj || n_traj <= 0) return mqct_fail(MQCT_ERR_ARG, "propagate: bad argument");
  if (!h->dyn) return mqct_fail(MQCT_ERR_STATE, "propagate: no matrix loaded");
  if (mode < MQCT_PROP_CC_RK4 || mode > MQCT_PROP_AT_ODEINT)
    return mqct_fail(MQCT_ERR_ARG, "propagate: unknown mode");
  DynState *ds = h->dyn;
  const bool at = (mode == MQCT_PROP_AT_RK4 || mode == MQCT_PROP_AT_ODEINT);
  if (at && ds->n_paths == 0)
    return mqct_fail(MQCT_ERR_STATE, "propagate: AT modes need recorded paths (load_paths)");
  for (int t = 0; t < n_traj; t++) {
    const mqct_traj_t &p = traj[t];
    if (p.s_st < 1 || p.s_st > ds->N)
      return mqct_fail(MQCT_ERR_ARG, "propagate: initial state out of range");
    if (at && (p.at_path < 0 || p.at_path >= ds->n_paths))
      return mqct_fail(MQCT_ERR_ARG, "propagate: path index out of range");
    if (!(p.dt > 0)) return mqct_fail(MQCT_ERR_ARG, "propagate: dt must be positive");
    if (!at) {
      double k2 = ds->mu * 2.0 * p.E_coll;
      double l2 = p.l_real * (p.l_real + 1.0) / (p.R_start * p.R_start);
      if (!(k2 > l2))
        return mqct_fail(MQCT_ERR_ARG, "propagate: impact parameter exceeds RMAX (trajectory not propagated by the reference)");
    }
    if ((mode == MQCT_PROP_CC_ODEINT || mode == MQCT_PROP_AT_ODEINT) && !(p.eps > 0 && p.hmin > 0))
      return mqct_fail(MQCT_ERR_ARG, "propagate: odeint needs eps > 0 and hmin > 0");
  }
  HIP_TRY(hipSetDevice(h->device));
  if (batched < 0 || batched > 3) return mqct_fail(MQCT_ERR_ARG, "propagate: batched must be 0, 1, 2 or 3");
  if (batched) return mqct_dyn_run_batched(h, mode, n_traj, traj, q_out, stat_out, batched >= 2, batched == 3);
  return mqct_dyn_run_plain(h, mode, n_traj, traj, q_out, stat_out);
}

extern "C" int mqct_hip_propagate_timing(mqct_hip_t *h, double *ms) {
  if (!h || !ms || !h->dyn) return mqct_fail(MQCT_ERR_ARG, "propagate_timing: bad argument");
  *ms = h->dyn->last_ms;
  return MQCT_OK;
}

extern "C" int mqct_hip_propagate_record(mqct_hip_t *h, int n_traj,
                                         const mqct_traj_t *traj, int rec_cap,
                                         double *records, int *rec_n, double *q_out,
                                         mqct_traj_stat_t *stat_out) {
  if (!h || !traj || n_traj <= 0 || rec_cap < 2 || !records || !rec_n)
    return mqct_fail(MQCT_ERR_ARG, "propagate_record: bad argument");
  if (!h->dyn) return mqct_fail(MQCT_ERR_STATE, "propagate_record: no matrix loaded");
  for (int t = 0; t < n_traj; t++) {
    const mqct_traj_t &p = traj[t];
    if (p.s_st < 1 || p.s_st > h->dyn->N)
      return mqct_fail(MQCT_ERR_ARG, "propagate_record: initial state out of range");
    double k2 = h->dyn->mu * 2.0 * p.E_coll;
    if (!(k2 > p.l_real * (p.l_real + 1.0) / (p.R_start * p.R_start)) || !(p.dt > 0))
      return mqct_fail(MQCT_ERR_ARG, "propagate_record: trajectory not propagatable");
  }
  HIP_TRY(hipSetDevice(h->device));
  return mqct_dyn_run_plain(h, MQCT_PROP_CC_RK4, n_traj, traj, q_out, stat_out, rec_cap,
                            records, rec_n);
}
