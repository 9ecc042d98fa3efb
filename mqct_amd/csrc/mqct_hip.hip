// mqct_hip.hip -- C ABI entry points of libmqct_hip.so for the matrix build,
// the post-processing kernel (cutoff + spline), and MTRX_UF.DAT I/O.
#include "mqct_internal.h"

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>

// ---------------------------------------------------------------------------
static thread_local std::string g_err;

void mqct_set_error(const std::string &msg) { g_err = msg; }
int mqct_fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int DevBuf::ensure(size_t n) {
  if (n == 0) n = 8;
  if (p && bytes >= n) return 0;
  release();
  if (hipMalloc(&p, n) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    return 1;
  }
  bytes = n;
  return 0;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// Clebsch-Gordan through the Racah 3j sum (restates suppl_routines.f:246-310
// in double-argument form).
static double lf(int n) { return mq_logfact(n); }
static double w3js(int J1, int J2, int J3, int M1, int M2, int M3) {
  if (M1 + M2 + M3 != 0) return 0.0;
  int IA = J1 + J2;
  if (J3 > IA) return 0.0;
  int IB = J1 - J2;
  if (J3 < std::abs(IB)) return 0.0;
  int JSUM = J3 + IA, IC = J1 - M1, ID = J2 - M2;
  if (JSUM % 2 || IC % 2 || ID % 2) return 0.0;
  if (std::abs(M1) > J1 || std::abs(M2) > J2 || std::abs(M3) > J3) return 0.0;
  int IE = J3 - J2 + M1, IF = J3 - J1 - M2;
  int ZMIN = std::max(0, std::max(-IE, -IF));
  int IG = IA - J3, IH = J2 + M2;
  int ZMAX = std::min(IG, std::min(IH, IC));
  double CC = 0.0;
  for (int Z = ZMIN; Z <= ZMAX; Z += 2) {
    double D = std::exp(lf(Z / 2) + lf((IG - Z) / 2) + lf((IC - Z) / 2) +
                        lf((IH - Z) / 2) + lf((IE + Z) / 2) + lf((IF + Z) / 2));
    if (Z % 4 != 0) D = -D;
    CC += 1.0 / D;
  }
  double CC1 = std::exp(lf(IG / 2) + lf((J3 + IB) / 2) + lf((J3 - IB) / 2) -
                        lf((JSUM + 2) / 2));
  double CC2 = std::exp(lf((J1 + M1) / 2) + lf(IC / 2) + lf(IH / 2) +
                        lf(ID / 2) + lf((J3 - M3) / 2) + lf((J3 + M3) / 2));
  CC *= std::sqrt(CC1 * CC2);
  if ((IB - M3) % 4 != 0) CC = -CC;
  return CC;
}
double mq_cg(int j1, int j2, int j3, int m1, int m2, int m3) {
  int e = -j1 + j2 - m3;
  double s = (e % 2 == 0) ? 1.0 : -1.0;
  return w3js(2 * j1, 2 * j2, 2 * j3, 2 * m1, 2 * m2, -2 * m3) *
         std::sqrt(2.0 * j3 + 1.0) * s;
}

// ---------------------------------------------------------------------------
extern "C" int mqct_hip_last_error(char *buf, int len) {
  if (!buf || len <= 0) return (int)g_err.size();
  std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return (int)g_err.size();
}

extern "C" int mqct_hip_version(void) { return 10000; }

void mqct_dyn_free(mqct_hip *h);
void mqct_mlp_free(mqct_hip *h);
int mqct_xsec_scratch_init(int device);

extern "C" int mqct_hip_create(int device, mqct_hip_t **out) {
  if (!out) return mqct_fail(MQCT_ERR_ARG, "create: null handle pointer");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return mqct_fail(MQCT_ERR_ARG, "create: no such device");
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return mqct_fail(MQCT_ERR_ARG, std::string("create: library is built for "
                                               "gfx950, device is ") +
                                       prop.gcnArchName);
  int src = mqct_xsec_scratch_init(device);
  if (src) return src;
  mqct_hip *h = new mqct_hip();
  h->device = device;
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  for (auto &e : h->ev) HIP_TRY(hipEventCreate(&e));
  *out = h;
  return MQCT_OK;
}

extern "C" int mqct_hip_destroy(mqct_hip_t *h) {
  if (!h) return MQCT_OK;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  mqct_dyn_free(h);
  mqct_mlp_free(h);
  for (auto &e : h->ev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return MQCT_OK;
}

// ---------------------------------------------------------------------------
static int build_blocks(mqct_hip *h);

extern "C" int mqct_hip_set_basis(mqct_hip_t *h, int coll_type, int n_states,
                                  const int *j12, const int *m12,
                                  const int *indx_chann, int n_channels,
                                  const int *j1_ch, const int *j2_ch,
                                  int ldvec, const double *eig1,
                                  const double *eig2) {
  if (!h) return mqct_fail(MQCT_ERR_ARG, "set_basis: null handle");
  if (coll_type != 1 && coll_type != 4 && coll_type != 0)
    return mqct_fail(MQCT_ERR_ARG, "set_basis: supported SYS_TYPE are 1, 4, 0");
  if (n_states <= 0 || n_channels <= 0 || !j12 || !m12 || !indx_chann || !j1_ch)
    return mqct_fail(MQCT_ERR_ARG, "set_basis: empty basis");
  if (coll_type != 1 && (!eig1 || ldvec <= 0))
    return mqct_fail(MQCT_ERR_ARG, "set_basis: eigenvectors required");
  if (coll_type == 0 && (!eig2 || !j2_ch))
    return mqct_fail(MQCT_ERR_ARG, "set_basis: second-molecule data required");
  HIP_TRY(hipSetDevice(h->device));
  h->coll_type = coll_type;
  h->n_states = n_states;
  h->n_channels = n_channels;
  h->ldvec = ldvec;
  h->j12.assign(j12, j12 + n_states);
  h->m12.assign(m12, m12 + n_states);
  h->indx_chann.assign(indx_chann, indx_chann + n_states);
  h->j1_ch.assign(j1_ch, j1_ch + n_channels);
  h->j2_ch.assign(n_channels, 0);
  if (j2_ch) h->j2_ch.assign(j2_ch, j2_ch + n_channels);
  for (int s = 0; s < n_states; s++) {
    int c = indx_chann[s];
    if (c < 1 || c > n_channels)
      return mqct_fail(MQCT_ERR_ARG, "set_basis: indx_chann out of range");
    int jj = h->j1_ch[c - 1];
    if (coll_type != 0 && (j12[s] != jj || std::abs(m12[s]) > jj))
      return mqct_fail(MQCT_ERR_ARG, "set_basis: j12/m12 inconsistent with channel");
    if (coll_type != 1 && 2 * jj + 1 > ldvec)
      return mqct_fail(MQCT_ERR_ARG, "set_basis: ldvec too small");
    if (coll_type == 0) {
      int j2 = h->j2_ch[c - 1];
      if (j12[s] < std::abs(jj - j2) || j12[s] > jj + j2 || std::abs(m12[s]) > j12[s])
        return mqct_fail(MQCT_ERR_ARG, "set_basis: j12 violates the triangle rule");
      if (2 * j2 + 1 > ldvec)
        return mqct_fail(MQCT_ERR_ARG, "set_basis: ldvec too small");
    }
  }
  if (eig1) h->eig1.assign(eig1, eig1 + (size_t)n_channels * ldvec);
  else h->eig1.clear();
  if (eig2) h->eig2.assign(eig2, eig2 + (size_t)n_channels * ldvec);
  else h->eig2.clear();
  h->identical = h->identical_pes = false;
  h->parity_state.clear();
  h->parity_state_bk.clear();
  h->same_level.clear();
  int rc = build_blocks(h);
  if (rc) return rc;
  if (!h->eig1.empty() && (rc = upload(h, h->d_eig1, h->eig1.data(), h->eig1.size()))) return rc;
  if (!h->eig2.empty() && (rc = upload(h, h->d_eig2, h->eig2.data(), h->eig2.size()))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->basis_ok = true;
  h->psi_fresh = false;
  h->built = false;
  h->n_terms = 0;
  h->terms_nr = 0;
  for (int &v : h->cvar) v = -1;
  return MQCT_OK;
}

// Blocks of the pair structure: states of equal m12 (and, for identical
// partners, equal exchange parity) in ascending m12 then parity, states
// ascending inside a block; kbase in INI_ARRAYS' pair order.
static int build_blocks(mqct_hip *h) {
  const int n_states = h->n_states;
  const int *m12 = h->m12.data();
  const int *indx_chann = h->indx_chann.data();
  std::map<std::pair<int, int>, std::vector<int>> blocks;
  for (int s = 0; s < n_states; s++)
    blocks[std::make_pair(m12[s], h->identical ? -h->parity_state[s] : 0)].push_back(s);
  h->pos_state.clear();
  h->block_start.clear();
  std::vector<int> rank(n_states), block_of;
  for (auto &kv : blocks) {
    h->block_start.push_back((int)h->pos_state.size());
    int r = 0;
    for (int s : kv.second) {
      rank[s] = r++;
      h->pos_state.push_back(s);
    }
  }
  h->block_start.push_back((int)h->pos_state.size());
  h->state_pos.assign(n_states, 0);
  for (int p = 0; p < n_states; p++) h->state_pos[h->pos_state[p]] = p;
  block_of.resize(n_states);
  for (size_t b = 0; b + 1 < h->block_start.size(); b++)
    for (int p = h->block_start[b]; p < h->block_start[b + 1]; p++)
      block_of[p] = h->block_start[b];
  // kbase: pairs before state s in ind_mat order (st1 outer, st2 <= st1)
  h->kbase.assign(n_states, 0);
  int64_t acc = 0;
  for (int s = 0; s < n_states; s++) {
    h->kbase[s] = acc;
    acc += rank[s] + 1;
  }
  h->total_pairs = acc;
  // device tables
  std::vector<int> chan(2 * (size_t)n_states);
  for (int s = 0; s < n_states; s++) {
    chan[s] = h->j1_ch[indx_chann[s] - 1];
    chan[n_states + s] = indx_chann[s] - 1;
  }
  std::vector<long long> kb(n_states + (n_states + 1) / 2 + 1);
  for (int s = 0; s < n_states; s++) kb[s] = h->kbase[s];
  std::memcpy(&kb[n_states], block_of.data(), sizeof(int) * n_states);
  int rc;
  if ((rc = upload(h, h->d_j12, h->j12.data(), n_states))) return rc;
  if ((rc = upload(h, h->d_m12, h->m12.data(), n_states))) return rc;
  if ((rc = upload(h, h->d_pos_state, h->pos_state.data(), n_states))) return rc;
  if ((rc = upload(h, h->d_kbase, kb.data(), kb.size()))) return rc;
  if ((rc = upload(h, h->d_chan, chan.data(), chan.size()))) return rc;
  return MQCT_OK;
}

extern "C" int mqct_hip_set_identical(mqct_hip_t *h, int identical_pes,
                                      const int *parity_state,
                                      const int *parity_state_bk,
                                      const int *same_level) {
  if (!h) return mqct_fail(MQCT_ERR_ARG, "set_identical: null handle");
  if (!h->basis_ok || h->coll_type != 0)
    return mqct_fail(MQCT_ERR_STATE, "ERROR: EXCHANGE SYMMETRY DEFINED FOR NON-IDENTICAL PARTICLES");
  const bool ident = parity_state != nullptr;
  if (ident && (!parity_state_bk || !same_level))
    return mqct_fail(MQCT_ERR_ARG, "set_identical: parity_state_bk and same_level required");
  if (ident && !identical_pes)
    return mqct_fail(MQCT_ERR_ARG, "ERROR: INDISTINGUISHABLE CALCULATIONS SHOULD BE DONE USING IDENTICAL PES");
  HIP_TRY(hipSetDevice(h->device));
  const int n = h->n_states;
  if (ident) {
    for (int s = 0; s < n; s++)
      if ((parity_state[s] != 1 && parity_state[s] != -1) ||
          (parity_state_bk[s] != 1 && parity_state_bk[s] != -1))
        return mqct_fail(MQCT_ERR_ARG, "set_identical: parities must be +1 or -1");
    h->parity_state.assign(parity_state, parity_state + n);
    h->parity_state_bk.assign(parity_state_bk, parity_state_bk + n);
    h->same_level.assign(same_level, same_level + h->n_channels);
  } else {
    h->parity_state.clear();
    h->parity_state_bk.clear();
    h->same_level.clear();
  }
  h->identical = ident;
  h->identical_pes = identical_pes != 0;
  int rc = build_blocks(h);
  if (rc) return rc;
  if (ident) {
    std::vector<int> par(2 * (size_t)n);
    for (int s = 0; s < n; s++) {
      par[s] = h->parity_state[s];
      par[n + s] = h->parity_state_bk[s];
    }
    if ((rc = upload(h, h->d_xpar, par.data(), par.size()))) return rc;
    if ((rc = upload(h, h->d_same, h->same_level.data(), h->same_level.size()))) return rc;
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->psi_fresh = false;
  h->built = false;
  h->n_terms = 0;
  h->terms_nr = 0;
  for (int &v : h->cvar) v = -1;
  return MQCT_OK;
}

extern "C" int mqct_hip_set_grid(mqct_hip_t *h, const mqct_grid_t *g) {
  if (!h || !g) return mqct_fail(MQCT_ERR_ARG, "set_grid: null argument");
  if (!h->basis_ok) return mqct_fail(MQCT_ERR_STATE, "set_grid: call set_basis first");
  HIP_TRY(hipSetDevice(h->device));
  const int t = h->coll_type;
  auto need = [](int n, const double *x, const double *w) {
    return n >= 1 && x && w;
  };
  if (!need(g->n_beta, g->xb, g->wb))
    return mqct_fail(MQCT_ERR_ARG, "set_grid: beta grid required");
  if (t != 1 && !need(g->n_gamma, g->xg, g->wg))
    return mqct_fail(MQCT_ERR_ARG, "set_grid: gamma grid required");
  if (t == 0 && !(need(g->n_alpha2, g->xaa, g->waa) &&
                  need(g->n_beta2, g->xbb, g->wbb) &&
                  need(g->n_gamma2, g->xgg, g->wgg)))
    return mqct_fail(MQCT_ERR_ARG, "set_grid: type-0 grids required");
  h->n_beta = g->n_beta;
  h->n_gamma = (t == 1) ? 1 : g->n_gamma;
  h->n_alpha2 = (t == 0) ? g->n_alpha2 : 1;
  h->n_beta2 = (t == 0) ? g->n_beta2 : 1;
  h->n_gamma2 = (t == 0) ? g->n_gamma2 : 1;
  h->ref_xz = (t == 4) ? (g->ref_xz ? 1 : 0) : 1;
  const double PI = MQCT_PI;
  // angles exactly as INTEGRATOR_TEMP forms them
  h->beta.resize(h->n_beta);
  for (int i = 0; i < h->n_beta; i++) {
    double b = g->xbm + g->xbl * g->xb[i];
    if (b < 0.0) b = 0.0;
    if (b > PI) b = PI;
    h->beta[i] = b;
  }
  auto map_eq = [&](int n, const double *x, double m, double l,
                    std::vector<double> &o) {
    o.resize(n);
    for (int i = 0; i < n; i++) o[i] = g->eq_gamma ? x[i] : m + l * x[i];
  };
  if (t != 1) map_eq(h->n_gamma, g->xg, g->xgm, g->xgl, h->gamma);
  else h->gamma.assign(1, 0.0);
  if (t == 0) {
    map_eq(h->n_alpha2, g->xaa, g->xaam, g->xaal, h->alpha2);
    h->beta2.resize(h->n_beta2);
    for (int i = 0; i < h->n_beta2; i++) h->beta2[i] = g->xbbm + g->xbbl * g->xbb[i];
    map_eq(h->n_gamma2, g->xgg, g->xggm, g->xggl, h->gamma2);
  } else {
    h->alpha2.assign(1, 0.0);
    h->beta2.assign(1, 0.0);
    h->gamma2.assign(1, 0.0);
  }
  h->G = (int64_t)h->n_beta * h->n_gamma * h->n_alpha2 * h->n_beta2 * h->n_gamma2;
  h->omega.resize(h->G);
  // quadrature weight per point: every factor of the integrand except the
  // two wavefunctions and V (pot_comp.f:1950-1951, 2554-2555, 3718-3722)
  int64_t gi = 0;
  if (t == 1) {
    for (int i = 0; i < h->n_beta; i++)
      h->omega[gi++] = g->wb[i] * std::sin(h->beta[i]) * g->xbl * 2.0 * PI;
  } else if (t == 4) {
    for (int i = 0; i < h->n_beta; i++)
      for (int k = 0; k < h->n_gamma; k++)
        h->omega[gi++] = g->wb[i] * g->wg[k] * std::sin(h->beta[i]) * g->xbl *
                         g->xgl * 2.0 * PI;
  } else {
    for (int i2 = 0; i2 < h->n_alpha2; i2++)
      for (int i3 = 0; i3 < h->n_beta; i3++)
        for (int i4 = 0; i4 < h->n_beta2; i4++)
          for (int i5 = 0; i5 < h->n_gamma; i5++)
            for (int i6 = 0; i6 < h->n_gamma2; i6++)
              h->omega[gi++] = g->wb[i3] * g->wg[i5] * std::sin(h->beta[i3]) *
                               g->waa[i2] * g->wbb[i4] * g->wgg[i6] *
                               std::sin(h->beta2[i4]) * g->xbl * g->xgl *
                               g->xaal * g->xbbl * g->xggl * 2 * PI;
  }
  int rc;
  if ((rc = upload(h, h->d_beta, h->beta.data(), h->beta.size()))) return rc;
  if ((rc = upload(h, h->d_gamma, h->gamma.data(), h->gamma.size()))) return rc;
  if ((rc = upload(h, h->d_alpha2, h->alpha2.data(), h->alpha2.size()))) return rc;
  if ((rc = upload(h, h->d_beta2, h->beta2.data(), h->beta2.size()))) return rc;
  if ((rc = upload(h, h->d_gamma2, h->gamma2.data(), h->gamma2.size()))) return rc;
  if ((rc = upload(h, h->d_omega, h->omega.data(), h->omega.size()))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->grid_ok = true;
  h->psi_fresh = false;
  h->v_ok = false;
  h->built = false;
  for (int &v : h->cvar) v = -1;
  return MQCT_OK;
}

static std::vector<double> spline_tables(const std::vector<double> &R);

int mqct_set_r(mqct_hip *h, int n_r, const double *R) {
  if (n_r < 2 || !R) return mqct_fail(MQCT_ERR_ARG, "R grid needs >= 2 points");
  for (int i = 1; i < n_r; i++)
    if (!(R[i] > R[i - 1]))
      return mqct_fail(MQCT_ERR_ARG, "R grid must increase");
  bool same = (n_r == h->n_r);
  for (int i = 0; same && i < n_r; i++) same = (R[i] == h->R[i]);
  if (!same) h->terms_nr = 0;  // expansion coefficients were on another R grid
  h->n_r = n_r;
  h->R.assign(R, R + n_r);
  int rc = upload(h, h->d_R, R, n_r);
  if (rc) return rc;
  // the spline's forward factors on this grid, for every build's post_kernel
  const std::vector<double> tab = spline_tables(h->R);
  return upload(h, h->d_spl, tab.data(), tab.size());
}

extern "C" int mqct_hip_set_vgrid(mqct_hip_t *h, int n_r, const double *R_com,
                                  const double *V) {
  if (!h || !V) return mqct_fail(MQCT_ERR_ARG, "set_vgrid: null argument");
  if (!h->grid_ok) return mqct_fail(MQCT_ERR_STATE, "set_vgrid: call set_grid first");
  HIP_TRY(hipSetDevice(h->device));
  int rc = mqct_set_r(h, n_r, R_com);
  if (rc) return rc;
  int64_t n = (int64_t)n_r * h->G;
  for (int64_t i = 0; i < n; i++)
    if (V[i] != V[i]) return mqct_fail(MQCT_ERR_ARG, "ERROR IN POTENTIAL ROUTINE (NaN in V grid)");
  if ((rc = upload(h, h->d_V, V, (size_t)n))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->v_ok = true;
  h->built = false;
  return MQCT_OK;
}

extern "C" int mqct_hip_eval_pes(mqct_hip_t *h, int n_r, const double *R_com,
                                 int pes_kind, int n_par, const double *par) {
  if (!h || !par || n_par <= 0) return mqct_fail(MQCT_ERR_ARG, "eval_pes: bad argument");
  if (!h->grid_ok) return mqct_fail(MQCT_ERR_STATE, "eval_pes: call set_grid first");
  int want = h->coll_type == 1 ? MQCT_PES_LEGENDRE
             : h->coll_type == 4 ? MQCT_PES_REALSH : MQCT_PES_SITES;
  if (pes_kind != want)
    return mqct_fail(MQCT_ERR_ARG, "eval_pes: potential does not match SYS_TYPE");
  if (mqct_pes_par_len(pes_kind, n_par, par) < 0)
    return mqct_fail(MQCT_ERR_ARG, "eval_pes: parameter array too short for its header");
  HIP_TRY(hipSetDevice(h->device));
  int rc = mqct_set_r(h, n_r, R_com);
  if (rc) return rc;
  if ((rc = upload(h, h->d_par, par, n_par))) return rc;
  if ((rc = mqct_eval_pes_dev(h, pes_kind, n_par))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->v_ok = true;
  h->built = false;
  return MQCT_OK;
}

// ---------------------------------------------------------------------------
// Clamped cubic spline of one element's n values y (matrix_ini.f:2005-2016,
// 6611-6621 with suppl_routines.f:914-965): end slopes from the first and
// last finite differences.  The forward factors y2f depend only on the R
// grid and come precomputed; the data-dependent u sweep is kept in d and
// back-substituted in place.
__device__ inline void spline_column(int n, const double *__restrict__ R,
                                     const double *__restrict__ sig,
                                     const double *__restrict__ pp,
                                     const double *__restrict__ y2f,
                                     const double *y, double *d) {
  double yp1 = (y[1] - y[0]) / (R[1] - R[0]);
  double ypn = (y[n - 1] - y[n - 2]) / (R[n - 1] - R[n - 2]);
  double u = (3.0 / (R[1] - R[0])) * ((y[1] - y[0]) / (R[1] - R[0]) - yp1);
  d[0] = u;
  double yim1 = y[0], yi = y[1];
  for (int i = 1; i < n - 1; i++) {
    double yip1 = y[i + 1];
    u = (6.0 * ((yip1 - yi) / (R[i + 1] - R[i]) - (yi - yim1) / (R[i] - R[i - 1])) /
             (R[i + 1] - R[i - 1]) -
         sig[i] * u) /
        pp[i];
    d[i] = u;
    yim1 = yi;
    yi = yip1;
  }
  double qn = 0.5;
  double un = (3.0 / (R[n - 1] - R[n - 2])) *
              (ypn - (y[n - 1] - y[n - 2]) / (R[n - 1] - R[n - 2]));
  double next = (un - qn * u) / (qn * y2f[n - 2] + 1.0);
  d[n - 1] = next;
  for (int i = n - 2; i >= 0; i--) {
    next = y2f[i] * next + d[i];
    d[i] = next;
  }
}

// Cutoff (matrix_ini.f:1991-1997) and clamped spline of the compute branch.
// One wave per element (column k, n R-contiguous values), PW per workgroup.
// spline_column's two sweeps are first-order linear recurrences,
//   forward   u_i = (X_i - sig_i u_{i-1}) / pp_i = X_i / pp_i + (-sig_i / pp_i) u_{i-1},
//   backward  d_i = y2f_i d_{i+1} + u_i,
// so each 64-point chunk is an inclusive scan over the lanes of the affine
// maps x -> A + B x (composition (A, B) after (A', B') = (A + B A', B B')),
// applied to the value carried from the previous chunk.  The end rows are
// the maps of spline_column's first and last points (B = 0 starts a chain).
// The sums are reassociated, so the values differ from the sequential
// sweeps in the last bits.  A column below the cutoff is zeroed.
constexpr int PW = 4;  // elements (waves) per workgroup
constexpr int PR = 8;  // 64-point chunks kept in registers (n <= 512)
__device__ inline void affine_scan_up(double &A, double &B, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double Ap = __shfl_up(A, o, 64), Bp = __shfl_up(B, o, 64);
    if (lane >= o) {
      A = A + B * Ap;
      B = B * Bp;
    }
  }
}
__device__ inline void affine_scan_down(double &A, double &B, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double An = __shfl_down(A, o, 64), Bn = __shfl_down(B, o, 64);
    if (lane + o < 64) {
      A = A + B * An;
      B = B * Bn;
    }
  }
}
__global__ __launch_bounds__(64 * PW) void post_kernel(int64_t T, int n, const double *__restrict__ R,
                                                       const double *__restrict__ sig,
                                                       const double *__restrict__ pp,
                                                       const double *__restrict__ y2f, double cut,
                                                       int ir1, int ir2,
                                                       const double *__restrict__ cutv,
                                                       double *__restrict__ M,
                                                       double *__restrict__ D,
                                                       unsigned char *__restrict__ skip) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * PW + (threadIdx.x >> 6);
  if (k >= T) return;  // the whole wave
  double *y = M + k * n, *d = D + k * n;
  // cutv (mixed precision): FP64 values at the two cutoff points
  const double a1 = cutv ? fabs(cutv[2 * k]) : fabs(y[ir1]);
  const double a2 = cutv ? fabs(cutv[2 * k + 1]) : fabs(y[ir2]);
  if ((a1 > a2 ? a1 : a2) < cut) {
    for (int i = lane; i < n; i += 64) {
      y[i] = 0.0;
      d[i] = 0.0;
    }
    if (lane == 0) skip[k] = 1;
    return;
  }
  if (lane == 0) skip[k] = 0;
  // the affine map of forward step i (spline_column's first and last
  // points at i = 0 and n - 1; the identity past the end)
  auto fwd_map = [&](int i, double &A, double &B) {
    A = 0.0;
    B = 1.0;
    if (i == 0) {
      const double y0 = y[0], y1 = y[1];
      const double yp1 = (y1 - y0) / (R[1] - R[0]);
      A = (3.0 / (R[1] - R[0])) * ((y1 - y0) / (R[1] - R[0]) - yp1);
      B = 0.0;
    } else if (i < n - 1) {
      const double ym = y[i - 1], yi = y[i], yp = y[i + 1];
      const double X = 6.0 * ((yp - yi) / (R[i + 1] - R[i]) - (yi - ym) / (R[i] - R[i - 1])) /
                       (R[i + 1] - R[i - 1]);
      A = X / pp[i];
      B = -sig[i] / pp[i];
    } else if (i == n - 1) {
      const double yl = y[n - 1], yk = y[n - 2], h = R[n - 1] - R[n - 2];
      const double ypn = (yl - yk) / h;
      const double qn = 0.5, un = (3.0 / h) * (ypn - (yl - yk) / h);
      const double den = qn * y2f[n - 2] + 1.0;
      A = un / den;
      B = -qn / den;
    }
  };
  if (n <= 64 * PR) {
    // u in registers, PR chunks per lane
    double ur[PR];
    double carry = 0.0;
#pragma unroll
    for (int c = 0; c < PR; c++) {
      ur[c] = 0.0;
      if (64 * c >= n) continue;
      double A, B;
      fwd_map(64 * c + lane, A, B);
      affine_scan_up(A, B, lane);
      ur[c] = A + B * carry;
      carry = __shfl(ur[c], 63, 64);
    }
    carry = 0.0;
#pragma unroll
    for (int c = PR - 1; c >= 0; c--) {
      if (64 * c >= n) continue;
      const int i = 64 * c + lane;
      double A = 0.0, B = 1.0;
      if (i < n) {
        A = ur[c];
        B = i < n - 1 ? y2f[i] : 0.0;  // d_{n-1} = u_{n-1}
      }
      affine_scan_down(A, B, lane);
      const double v = A + B * carry;
      if (i < n) d[i] = v;
      carry = __shfl(v, 0, 64);
    }
    return;
  }
  // longer grids: u_i goes to d[i] until the backward sweep replaces it
  double carry = 0.0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    double A, B;
    fwd_map(i, A, B);
    affine_scan_up(A, B, lane);
    const double u = A + B * carry;
    if (i < n) d[i] = u;
    carry = __shfl(u, 63, 64);
  }
  carry = 0.0;
  for (int i0 = (n - 1) / 64 * 64; i0 >= 0; i0 -= 64) {
    const int i = i0 + lane;
    double A = 0.0, B = 1.0;
    if (i < n) {
      A = d[i];
      B = i < n - 1 ? y2f[i] : 0.0;
    }
    affine_scan_down(A, B, lane);
    const double v = A + B * carry;
    if (i < n) d[i] = v;
    carry = __shfl(v, 0, 64);
  }
}

// READ_MIJ_USER after each column is read (matrix_ini.f:6603-6621): the
// MIJ_SHIFT subtraction Mat_el(i,k) - Mat_el(n_r,k), then the spline of the
// shifted values.
__global__ void read_post_kernel(int64_t T, int n, const double *__restrict__ R,
                                 const double *__restrict__ sig,
                                 const double *__restrict__ pp,
                                 const double *__restrict__ y2f, int shift,
                                 double *__restrict__ M,
                                 double *__restrict__ D) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  double *y = M + k * n, *d = D + k * n;
  if (shift) {
    double last = y[n - 1];
    for (int i = 0; i < n; i++) y[i] = y[i] - last;
  }
  spline_column(n, R, sig, pp, y2f, y, d);
}

// Forward factors of the spline's tridiagonal sweep for the grid R:
// tab = [sig | p | y2 forward] (suppl_routines.f:930-945).
static std::vector<double> spline_tables(const std::vector<double> &R) {
  const int n = (int)R.size();
  std::vector<double> tab(3 * (size_t)n, 0.0);
  double *sig = &tab[0], *pp = &tab[n], *y2f = &tab[2 * n];
  for (int i = 0; i < n; i++) pp[i] = 1.0;
  y2f[0] = -0.5;
  for (int i = 1; i < n - 1; i++) {
    sig[i] = (R[i] - R[i - 1]) / (R[i + 1] - R[i - 1]);
    pp[i] = sig[i] * y2f[i - 1] + 2.0;
    y2f[i] = (sig[i] - 1.0) / pp[i];
  }
  return tab;
}

int mqct_run_post(mqct_hip *h, double cut, int ir1, int ir2,
                  const double *cutv) {
  int n = h->n_r;
  const double *t = h->d_spl.as<double>();  // uploaded with the R grid
  int64_t T = h->built_T;
  post_kernel<<<(unsigned)((T + PW - 1) / PW), 64 * PW, 0, h->stream>>>(
      T, n, h->d_R.as<double>(), t, t + n, t + 2 * n, cut, ir1 - 1, ir2 - 1,
      cutv, h->d_mat.as<double>(), h->d_der.as<double>(),
      h->d_skip.as<unsigned char>());
  HIP_TRY(hipGetLastError());
  return MQCT_OK;
}

extern "C" int mqct_hip_matrix_build(mqct_hip_t *h, int precision,
                                     int64_t total_size, const int *ind_mat,
                                     double conv_unit_e, double mij_zero_cut,
                                     int ir_cut1, int ir_cut2, double *mat_el,
                                     double *mat_el_der, uint8_t *skipped) {
  if (!h) return mqct_fail(MQCT_ERR_ARG, "matrix_build: null handle");
  if (!h->basis_ok || !h->grid_ok || !h->v_ok)
    return mqct_fail(MQCT_ERR_STATE, "matrix_build: basis, grid and potential must be set");
  if (precision < MQCT_PREC_FP64 || precision > MQCT_PREC_SURROGATE)
    return mqct_fail(MQCT_ERR_ARG, "matrix_build: unknown precision");
  if (precision == MQCT_PREC_SURROGATE)
    return mqct_fail(MQCT_ERR_ARG, "matrix_build: surrogate form needs mqct_hip_surrogate_vgrid first; pass MQCT_PREC_BF16X3 after it");
  if (total_size != h->total_pairs)
    return mqct_fail(MQCT_ERR_ARG, "ERROR: total_size does not match the basis (INI_ARRAYS pair count)");
  if (ir_cut1 < 1 || ir_cut1 > h->n_r || ir_cut2 < 1 || ir_cut2 > h->n_r)
    return mqct_fail(MQCT_ERR_ARG, "matrix_build: cutoff R index out of range");
  if (h->identical)
    return mqct_fail(MQCT_ERR_ARG, "matrix_build: the quadrature build of identical partners is not implemented; "
                                   "build their matrix through the expansion (IDENTICAL_PES)");
  if (ind_mat) {
    // must be INI_ARRAYS' order: st1 ascending, st2 ascending, equal m12
    int64_t k = 0;
    for (int s1 = 0; s1 < h->n_states; s1++)
      for (int s2 = 0; s2 <= s1; s2++) {
        if (!mqct_pair_ok(h, s1, s2)) continue;
        if (ind_mat[2 * k] != s1 + 1 || ind_mat[2 * k + 1] != s2 + 1)
          return mqct_fail(MQCT_ERR_ARG, "matrix_build: ind_mat is not the INI_ARRAYS order");
        k++;
      }
  }
  HIP_TRY(hipSetDevice(h->device));
  const int64_t T = total_size;
  if (h->d_mat.ensure((size_t)T * h->n_r * 8) ||
      h->d_der.ensure((size_t)T * h->n_r * 8) || h->d_skip.ensure((size_t)T))
    return mqct_fail(MQCT_ERR_NOMEM, "matrix_build: output does not fit in HBM");
  h->built_T = T;
  int rc;
  HIP_TRY(hipEventRecord(h->ev[0], h->stream));
  if (!h->psi_fresh) {
    if ((rc = mqct_build_psi(h))) return rc;
  }
  HIP_TRY(hipEventRecord(h->ev[1], h->stream));
  if ((rc = mqct_build_weights(h, precision))) return rc;
  HIP_TRY(hipEventRecord(h->ev[2], h->stream));
  if ((rc = mqct_run_contraction(h, precision, conv_unit_e))) return rc;
  const double *cutv = nullptr;
  if (precision != MQCT_PREC_FP64) {
    if ((rc = mqct_run_cutoff_fp64(h, ir_cut1, ir_cut2, conv_unit_e))) return rc;
    cutv = h->d_cutv.as<double>();
  }
  HIP_TRY(hipEventRecord(h->ev[3], h->stream));
  if ((rc = mqct_run_post(h, mij_zero_cut, ir_cut1, ir_cut2, cutv))) return rc;
  HIP_TRY(hipEventRecord(h->ev[4], h->stream));
  HIP_TRY(hipEventSynchronize(h->ev[4]));
  float ms;
  for (int i = 0; i < 4; i++) {
    HIP_TRY(hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]));
    h->timing[i] = ms;
  }
  HIP_TRY(hipEventElapsedTime(&ms, h->ev[0], h->ev[4]));
  h->timing[4] = ms;
  // choosing the contraction variant (first build of a basis and grid in
  // this precision form) is left out of the build time
  h->timing[2] = std::max(0.0, h->timing[2] - h->tune_ms);
  h->timing[4] = std::max(0.0, h->timing[4] - h->tune_ms);
  h->built = true;
  if (mat_el || mat_el_der || skipped) {
    rc = mqct_hip_matrix_get(h, mat_el, mat_el_der);
    if (rc) return rc;
    if (skipped)
      HIP_TRY(hipMemcpy(skipped, h->d_skip.p, (size_t)T, hipMemcpyDeviceToHost));
  }
  return MQCT_OK;
}

extern "C" int mqct_hip_matrix_get(mqct_hip_t *h, double *mat_el,
                                   double *mat_el_der) {
  if (!h || !h->built) return mqct_fail(MQCT_ERR_STATE, "matrix_get: nothing built");
  HIP_TRY(hipSetDevice(h->device));
  size_t bytes = (size_t)h->built_T * h->n_r * 8;
  if (mat_el) HIP_TRY(hipMemcpy(mat_el, h->d_mat.p, bytes, hipMemcpyDeviceToHost));
  if (mat_el_der) HIP_TRY(hipMemcpy(mat_el_der, h->d_der.p, bytes, hipMemcpyDeviceToHost));
  return MQCT_OK;
}

extern "C" int mqct_hip_build_timing(mqct_hip_t *h, double *ms5) {
  if (!h || !ms5) return mqct_fail(MQCT_ERR_ARG, "build_timing: null argument");
  for (int i = 0; i < 5; i++) ms5[i] = h->timing[i];
  return MQCT_OK;
}

extern "C" int mqct_hip_set_progress(mqct_hip_t *h, double seconds) {
  if (!h || !(seconds >= 0.0)) return mqct_fail(MQCT_ERR_ARG, "set_progress: bad argument");
  h->progress_s = seconds;
  return MQCT_OK;
}

extern "C" int mqct_hip_set_mij_shift(mqct_hip_t *h, int on) {
  if (!h) return mqct_fail(MQCT_ERR_ARG, "set_mij_shift: null handle");
  h->mij_shift = on != 0;
  return MQCT_OK;
}

extern "C" int mqct_hip_mij_spline(mqct_hip_t *h, int n_r, const double *R_com,
                                   int64_t total_size, double *mat_el,
                                   double *mat_el_der) {
  if (!h || !R_com || !mat_el || !mat_el_der)
    return mqct_fail(MQCT_ERR_ARG, "mij_spline: null argument");
  if (n_r < 3 || total_size <= 0) return mqct_fail(MQCT_ERR_ARG, "mij_spline: bad sizes");
  for (int i = 1; i < n_r; i++)
    if (!(R_com[i] > R_com[i - 1]))
      return mqct_fail(MQCT_ERR_ARG, "mij_spline: R grid must increase");
  HIP_TRY(hipSetDevice(h->device));
  std::vector<double> R(R_com, R_com + n_r);
  std::vector<double> tab = spline_tables(R);
  const size_t bytes = (size_t)total_size * n_r * 8;
  DevBuf d_M, d_D, d_tab, d_Rg;
  if (d_M.ensure(bytes) || d_D.ensure(bytes) || d_tab.ensure(tab.size() * 8) ||
      d_Rg.ensure((size_t)n_r * 8))
    return mqct_fail(MQCT_ERR_NOMEM, "mij_spline: matrix does not fit in HBM");
  HIP_TRY(hipMemcpyAsync(d_M.p, mat_el, bytes, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(d_Rg.p, R.data(), (size_t)n_r * 8, hipMemcpyHostToDevice, h->stream));
  const double *t = d_tab.as<double>();
  read_post_kernel<<<(unsigned)((total_size + 255) / 256), 256, 0, h->stream>>>(
      total_size, n_r, d_Rg.as<double>(), t, t + n_r, t + 2 * n_r, h->mij_shift ? 1 : 0,
      d_M.as<double>(), d_D.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(mat_el, d_M.p, bytes, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipMemcpyAsync(mat_el_der, d_D.p, bytes, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}

// ---------------------------------------------------------------------------
// MTRX_UF.DAT: Fortran sequential unformatted records, each framed by a
// 4-byte little-endian byte count before and after (flang/gfortran).
// Layout of PRINT_MIJ_USER, matrix_ini.f:6658-6769.
namespace {
struct RecWriter {
  FILE *f;
  std::vector<unsigned char> buf;
  void i32(int v) { put(&v, 4); }
  void f64(double v) { put(&v, 8); }
  void put(const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    buf.insert(buf.end(), c, c + n);
  }
  bool flush() {
    uint32_t n = (uint32_t)buf.size();
    bool ok = fwrite(&n, 4, 1, f) == 1;
    if (!buf.empty()) ok = ok && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    ok = ok && fwrite(&n, 4, 1, f) == 1;
    buf.clear();
    return ok;
  }
};
struct RecReader {
  FILE *f;
  std::vector<unsigned char> buf;
  size_t pos = 0;
  bool next() {
    uint32_t n, n2;
    if (fread(&n, 4, 1, f) != 1) return false;
    buf.resize(n);
    if (n && fread(buf.data(), 1, n, f) != n) return false;
    if (fread(&n2, 4, 1, f) != 1 || n2 != n) return false;
    pos = 0;
    return true;
  }
  bool i32(int &v) { return get(&v, 4); }
  bool f64(double &v) { return get(&v, 8); }
  bool get(void *p, size_t n) {
    if (pos + n > buf.size()) return false;
    std::memcpy(p, buf.data() + pos, n);
    pos += n;
    return true;
  }
};
int qn_count(int coll_type) { return coll_type == 1 ? 1 : coll_type == 4 ? 3 : 6; }
}  // namespace

extern "C" int mqct_hip_write_mtrx_uf(const char *path, int coll_type,
                                      int n_channels, int n_states,
                                      int64_t total_size, int n_r,
                                      const int *indx_chann, const int *j12,
                                      const int *m12, const int *chan_qn,
                                      const int *ind_mat, const double *R_com,
                                      const double *mat_el) {
  if (!path || !indx_chann || !j12 || !m12 || !chan_qn || !ind_mat || !R_com || !mat_el)
    return mqct_fail(MQCT_ERR_ARG, "write_mtrx_uf: null argument");
  if (coll_type != 1 && coll_type != 4 && coll_type != 0)
    return mqct_fail(MQCT_ERR_ARG, "write_mtrx_uf: SYS_TYPE 1, 4, 0 only");
  FILE *f = std::fopen(path, "wb");
  if (!f) return mqct_fail(MQCT_ERR_ARG, std::string("write_mtrx_uf: cannot open ") + path);
  RecWriter w{f, {}};
  bool ok = true;
  w.i32(coll_type); ok &= w.flush();
  w.i32(n_channels); ok &= w.flush();
  w.i32(n_states); ok &= w.flush();
  w.i32((int)total_size); ok &= w.flush();
  w.i32(n_r); ok &= w.flush();
  const int nq = qn_count(coll_type);
  for (int s = 0; s < n_states && ok; s++) {
    int c = indx_chann[s];
    w.i32(s + 1);
    w.i32(c);
    w.i32(j12[s]);
    w.i32(m12[s]);
    for (int q = 0; q < nq; q++) w.i32(chan_qn[(size_t)(c - 1) * nq + q]);
    ok &= w.flush();
  }
  for (int64_t k = 0; k < total_size && ok; k++) {
    w.i32((int)(k + 1));
    w.i32(ind_mat[2 * k]);
    w.i32(ind_mat[2 * k + 1]);
    ok &= w.flush();
  }
  for (int i = 0; i < n_r && ok; i++) {
    w.i32(i + 1);
    w.f64(R_com[i]);
    ok &= w.flush();
  }
  for (int64_t k = 0; k < total_size && ok; k++) {
    w.i32((int)(k + 1));
    ok &= w.flush();
    w.put(mat_el + (size_t)k * n_r, (size_t)n_r * 8);
    ok &= w.flush();
  }
  ok &= std::fclose(f) == 0;
  if (!ok) return mqct_fail(MQCT_ERR_ARG, "write_mtrx_uf: write failed");
  return MQCT_OK;
}

extern "C" int mqct_hip_read_mtrx_uf(const char *path, int64_t *hdr,
                                     int *indx_chann, int *j12, int *m12,
                                     int *chan_qn, int *ind_mat,
                                     double *R_com, double *mat_el) {
  if (!path || !hdr) return mqct_fail(MQCT_ERR_ARG, "read_mtrx_uf: null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return mqct_fail(MQCT_ERR_ARG, std::string("read_mtrx_uf: cannot open ") + path);
  RecReader r{f, {}};
  int v[5];
  for (int i = 0; i < 5; i++) {
    if (!r.next() || !r.i32(v[i])) {
      std::fclose(f);
      return mqct_fail(MQCT_ERR_ARG, "read_mtrx_uf: bad header");
    }
  }
  for (int i = 0; i < 5; i++) hdr[i] = v[i];
  const int coll_type = v[0], n_states = v[2], n_r = v[4];
  const int64_t T = v[3];
  if (!indx_chann) {
    std::fclose(f);
    return MQCT_OK;
  }
  const int nq = qn_count(coll_type);
  bool ok = true;
  for (int s = 0; s < n_states && ok; s++) {
    int st, c, a, b;
    ok = r.next() && r.i32(st) && r.i32(c) && r.i32(a) && r.i32(b) && st == s + 1;
    if (!ok) break;
    indx_chann[s] = c;
    j12[s] = a;
    m12[s] = b;
    for (int q = 0; q < nq && ok; q++) {
      int x;
      ok = r.i32(x);
      if (chan_qn) chan_qn[(size_t)(c - 1) * nq + q] = x;
    }
  }
  for (int64_t k = 0; k < T && ok; k++) {
    int kk, a, b;
    ok = r.next() && r.i32(kk) && r.i32(a) && r.i32(b) && kk == k + 1;
    ind_mat[2 * k] = a;
    ind_mat[2 * k + 1] = b;
  }
  for (int i = 0; i < n_r && ok; i++) {
    int ii;
    double x;
    ok = r.next() && r.i32(ii) && r.f64(x) && ii == i + 1;
    R_com[i] = x;
  }
  for (int64_t k = 0; k < T && ok; k++) {
    int kk;
    ok = r.next() && r.i32(kk) && kk == k + 1 && r.next() &&
         r.buf.size() == (size_t)n_r * 8;
    if (ok) std::memcpy(mat_el + (size_t)k * n_r, r.buf.data(), (size_t)n_r * 8);
  }
  std::fclose(f);
  if (!ok) return mqct_fail(MQCT_ERR_ARG, "read_mtrx_uf: truncated or malformed file");
  return MQCT_OK;
}

// ---------------------------------------------------------------------------
// The classical path of the first AT run (SAVE_TRAJECT): in directory dir,
// AT_<j12>_<m12>_<L>_<parity> holds one Fortran unformatted record of 8
// REAL*8 (t, R, P_R, Theta, P_Theta, Phi, P_Phi, V) per step and
// n_AT_<...> one record with the number of records as INTEGER*4
// (propagator.f:2838-2852 names and opens them, 3457-3462 and 4330-4333
// write, 2994-3005 read back).
static std::string at_name(const char *dir, const char *pre, int j12, int m12, int l, int parity) {
  char nm[128];
  std::snprintf(nm, sizeof nm, "%sAT_%d_%d_%d_%d", pre, j12, m12, l, parity);
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  return d + nm;
}

extern "C" int mqct_hip_write_at_path(const char *dir, int j12, int m12, int l, int parity,
                                      int n_rec, const double *records) {
  if (!dir || !records || n_rec < 0) return mqct_fail(MQCT_ERR_ARG, "write_at_path: bad argument");
  const std::string p = at_name(dir, "", j12, m12, l, parity), pn = at_name(dir, "n_", j12, m12, l, parity);
  FILE *f = std::fopen(p.c_str(), "wb");
  if (!f) return mqct_fail(MQCT_ERR_ARG, "write_at_path: cannot open " + p);
  RecWriter w{f, {}};
  bool ok = true;
  for (int i = 0; i < n_rec && ok; i++) {
    for (int c = 0; c < 8; c++) w.f64(records[8 * (size_t)i + c]);
    ok &= w.flush();
  }
  ok &= std::fclose(f) == 0;
  FILE *g = std::fopen(pn.c_str(), "wb");
  if (!g) return mqct_fail(MQCT_ERR_ARG, "write_at_path: cannot open " + pn);
  RecWriter wn{g, {}};
  wn.i32(n_rec);
  ok &= wn.flush();
  ok &= std::fclose(g) == 0;
  if (!ok) return mqct_fail(MQCT_ERR_ARG, "write_at_path: write failed for " + p);
  return MQCT_OK;
}

extern "C" int mqct_hip_read_at_path(const char *dir, int j12, int m12, int l, int parity, int cap,
                                     double *records, int *n_rec) {
  if (!dir || !n_rec) return mqct_fail(MQCT_ERR_ARG, "read_at_path: bad argument");
  const std::string p = at_name(dir, "", j12, m12, l, parity), pn = at_name(dir, "n_", j12, m12, l, parity);
  FILE *g = std::fopen(pn.c_str(), "rb");
  if (!g) return mqct_fail(MQCT_ERR_ARG, "read_at_path: cannot open " + pn);
  RecReader rn{g, {}};
  int n = -1;
  bool ok = rn.next() && rn.buf.size() == 4 && rn.i32(n);
  std::fclose(g);
  if (!ok || n < 0) return mqct_fail(MQCT_ERR_ARG, "read_at_path: bad record count in " + pn);
  *n_rec = n;
  if (!records) return MQCT_OK;  // size query
  if (n > cap) return mqct_fail(MQCT_ERR_ARG, "read_at_path: more records than cap in " + p);
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return mqct_fail(MQCT_ERR_ARG, "read_at_path: cannot open " + p);
  RecReader r{f, {}};
  for (int i = 0; i < n && ok; i++) {
    ok = r.next() && r.buf.size() == 64;
    for (int c = 0; c < 8 && ok; c++) ok = r.f64(records[8 * (size_t)i + c]);
  }
  std::fclose(f);
  if (!ok) return mqct_fail(MQCT_ERR_ARG, "read_at_path: " + p + " is not 8 REAL*8 per record for the count");
  return MQCT_OK;
}
