// basis.hip -- basis functions and potential on the angular grid.
//
// Psi: for every state (in m-block order, "position") the wavefunction the
// reference evaluates inside INTEGRATOR_TEMP at each grid point, computed
// once and kept in HBM: type 1 WF_DIAT_TOP (basis_wave_functions.f:153-175,
// phi=0, purely real), type 4 WF_ASYM_TOP at alpha=0
// (basis_wave_functions.f:83-116), type 0 WF_ASYM_TOP_ASYM_TOP
// (basis_wave_functions.f:1-41) assembled from per-level one-molecule tables.
// Layout: re and im planes, [position][G] with G the flattened grid in the
// reference's loop order (type 4: beta outer, gamma inner; type 0: alpha2,
// beta1, beta2, gamma1, gamma2).
//
// W: w_g * V(R, g) per R point, [n_r][G], the weight containing every factor
// of the reference's integrand except the two wavefunctions.
#include "mqct_internal.h"

#include <cmath>
#include <cstring>

namespace {

__global__ void psi_type1_kernel(const int *__restrict__ pos_state,
                                 const int *__restrict__ m12,
                                 const int *__restrict__ jst,
                                 const double *__restrict__ beta, int npos,
                                 int G, double *__restrict__ re,
                                 double *__restrict__ im) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)npos * G) return;
  int pos = (int)(idx / G), g = (int)(idx % G);
  int s = pos_state[pos];
  int j = jst[s], m = m12[s];
  int am = m < 0 ? -m : m;
  double p = mq_plgndr(j, am, cos(beta[g]));
  double nrm = sqrt((2.0 * j + 1.0) / 4.0 / MQCT_PI *
                    exp(mq_logfact(j - am) - mq_logfact(j + am)));
  double r = p * nrm;  // cos(m*0) = 1, sin(m*0) = 0
  if (m < 0 && (m % 2 != 0)) r = -r;
  re[idx] = r;
  im[idx] = 0.0;
}

// One asymmetric-top function sum_K c_K sqrt((2J+1)/8pi^2) d^J_{K M}(beta)
// exp(i(K gamma + M alpha)).
__device__ inline void asym_top_wf(int J, const double *__restrict__ c, int M,
                                   double alpha, double b, double gam,
                                   double *re, double *im) {
  double cb = cos(b), sh = sin(b / 2.0), ch = cos(b / 2.0);
  double nrm = sqrt((2.0 * J + 1.0) / 8.0 / (MQCT_PI * MQCT_PI));
  double wr = 0.0, wi = 0.0;
  for (int K = -J; K <= J; K++) {
    double ck = c[K + J];
    if (ck == 0.0) continue;
    double d = mq_wigner_d(J, K, M, cb, sh, ch) * nrm;
    double ph = K * gam + M * alpha;
    wr += d * cos(ph) * ck;
    wi += d * sin(ph) * ck;
  }
  *re = wr;
  *im = wi;
}

__global__ void psi_type4_kernel(const int *__restrict__ pos_state,
                                 const int *__restrict__ m12,
                                 const int *__restrict__ jst,
                                 const int *__restrict__ chan,
                                 const double *__restrict__ eig, int ldvec,
                                 const double *__restrict__ beta,
                                 const double *__restrict__ gamma, int nb,
                                 int ng, int npos, double *__restrict__ re,
                                 double *__restrict__ im) {
  int64_t G = (int64_t)nb * ng;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)npos * G) return;
  int pos = (int)(idx / G);
  int g = (int)(idx % G);
  int s = pos_state[pos];
  int ib = g / ng, ig = g % ng;
  double wr, wi;
  asym_top_wf(jst[s], eig + (size_t)chan[s] * ldvec, m12[s], 0.0, beta[ib],
              gamma[ig], &wr, &wi);
  re[idx] = wr;
  im[idx] = wi;
}

// Type 0 one-molecule tables: tab[(lvl*(2jm+1) + m+jm) * Gm + g].
__global__ void tab_kernel(int nlvl, int jm, const int *__restrict__ lvl_j,
                           const double *__restrict__ lvl_eig, int ldvec,
                           const double *__restrict__ a_ang, int na,
                           const double *__restrict__ b_ang, int nb,
                           const double *__restrict__ g_ang, int ngm,
                           double2 *__restrict__ tab) {
  int64_t Gm = (int64_t)na * nb * ngm;
  int64_t nm = 2 * jm + 1;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nlvl * nm * Gm) return;
  int64_t g = idx % Gm;
  int m = (int)((idx / Gm) % nm) - jm;
  int l = (int)(idx / (Gm * nm));
  int J = lvl_j[l];
  double2 out = make_double2(0.0, 0.0);
  if (m >= -J && m <= J) {
    int ia = (int)(g / (nb * ngm));
    int ib = (int)((g / ngm) % nb);
    int igm = (int)(g % ngm);
    double wr, wi;
    asym_top_wf(J, lvl_eig + (size_t)l * ldvec, m, a_ang[ia], b_ang[ib],
                g_ang[igm], &wr, &wi);
    out = make_double2(wr, wi);
  }
  tab[idx] = out;
}

__global__ void psi_type0_kernel(
    const int *__restrict__ pos_state, const int *__restrict__ m12,
    const int *__restrict__ lvl,  // per state: lvl1, lvl2, j1, j2
    const double *__restrict__ cg, int ncg, int jm,
    const double2 *__restrict__ tab1, const double2 *__restrict__ tab2,
    int na2, int nb1, int nb2, int ng1, int ng2, int npos,
    double *__restrict__ re, double *__restrict__ im) {
  int64_t G = (int64_t)na2 * nb1 * nb2 * ng1 * ng2;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)npos * G) return;
  int pos = (int)(idx / G);
  int64_t g = idx % G;
  int ig2 = (int)(g % ng2);
  int ig1 = (int)((g / ng2) % ng1);
  int ib2 = (int)((g / ((int64_t)ng2 * ng1)) % nb2);
  int ib1 = (int)((g / ((int64_t)ng2 * ng1 * nb2)) % nb1);
  int ia2 = (int)(g / ((int64_t)ng2 * ng1 * nb2 * nb1));
  int s = pos_state[pos];
  int l1 = lvl[4 * s], l2 = lvl[4 * s + 1], j1 = lvl[4 * s + 2],
      j2 = lvl[4 * s + 3];
  int m12s = m12[s];
  int nm = 2 * jm + 1;
  int64_t G1 = (int64_t)nb1 * ng1, G2 = (int64_t)na2 * nb2 * ng2;
  int64_t g1 = (int64_t)ib1 * ng1 + ig1;
  int64_t g2 = ((int64_t)ia2 * nb2 + ib2) * ng2 + ig2;
  double wr = 0.0, wi = 0.0;
  for (int m1 = -j1; m1 <= j1; m1++) {
    int m2 = m12s - m1;
    if (m2 < -j2 || m2 > j2) continue;
    double c = cg[(size_t)s * ncg + (m1 + jm)];
    double2 a = tab1[((int64_t)l1 * nm + (m1 + jm)) * G1 + g1];
    double2 b = tab2[((int64_t)l2 * nm + (m2 + jm)) * G2 + g2];
    wr += c * (a.x * b.x - a.y * b.y);
    wi += c * (a.x * b.y + a.y * b.x);
  }
  re[idx] = wr;
  im[idx] = wi;
}

// REF_XZ=NO operand: (Re - Im, Im + Re), pot_comp.f:2544-2553.
__global__ void psi_refxz_kernel(int64_t n, const double *__restrict__ re,
                                 const double *__restrict__ im,
                                 double *__restrict__ ore,
                                 double *__restrict__ oim) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ore[i] = re[i] - im[i];
  oim[i] = im[i] + re[i];
}

__global__ void to_f32_kernel(int64_t n, const double *__restrict__ a,
                              float *__restrict__ o) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  o[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
// Synthetic potentials (same definitions as the workload descriptions).
__device__ double pes_legendre(const double *__restrict__ par, double R,
                               double theta) {
  int L = (int)par[0];
  double x = cos(theta), v = 0.0, r6 = R * R * R * R * R * R;
  double p0 = 1.0, p1 = x;
  for (int l = 0; l < L; l++) {
    double pl;
    if (l == 0) pl = 1.0;
    else if (l == 1) pl = x;
    else {
      double p2 = ((2.0 * l - 1.0) * x * p1 - (l - 1.0) * p0) / l;
      p0 = p1;
      p1 = p2;
      pl = p2;
    }
    v += (par[1 + l] * exp(-par[1 + L + l] * R) - par[1 + 2 * L + l] / r6) * pl;
  }
  const double *w = par + 1 + 3 * L;  // isotropic core Aw exp(-bw (R - Rw))
  v += w[0] * exp(-w[1] * (R - w[2]));
  return v;
}

__device__ double pes_realsh(const double *__restrict__ par, double R,
                             double beta, double gamma) {
  int lmax = (int)par[0];
  const double *p = par + 1;
  double x = cos(beta), v = 0.0, r6 = R * R * R * R * R * R;
  for (int l = 0; l <= lmax; l++) {
    double rad = p[0] * exp(-p[1] * R) - p[2] / r6;
    const double *c = p + 3;
    double ang = 0.0;
    for (int m = -l; m <= l; m++) {
      int am = m < 0 ? -m : m;
      double nrm = sqrt((2.0 * l + 1.0) / (4.0 * MQCT_PI) *
                        exp(mq_logfact(l - am) - mq_logfact(l + am)));
      double plm = mq_plgndr(l, am, x) * nrm;
      double s;
      if (m > 0) s = sqrt(2.0) * plm * cos(m * gamma);
      else if (m < 0) s = sqrt(2.0) * plm * sin(am * gamma);
      else s = plm;
      ang += c[m + l] * s;
    }
    v += rad * ang;
    p += 3 + 2 * l + 1;
  }
  v += p[0] * exp(-p[1] * (R - p[2]));  // isotropic core
  return v;
}

__device__ inline void rot_zyz(double ca, double sa, double cb, double sb,
                               double cg, double sg, const double *v,
                               double *o) {
  double x1 = cg * v[0] - sg * v[1], y1 = sg * v[0] + cg * v[1], z1 = v[2];
  double x2 = cb * x1 + sb * z1, y2 = y1, z2 = -sb * x1 + cb * z1;
  o[0] = ca * x2 - sa * y2;
  o[1] = sa * x2 + ca * y2;
  o[2] = z2;
}

__device__ double pes_sites(const double *__restrict__ par, double R,
                            double b1, double g1, double a2, double b2,
                            double g2) {
  const double HARTREE_CM = 27.21113845 * 8065.5448;
  double alpha = par[0];
  int n1 = (int)par[1], n2 = (int)par[2];
  const double *s1 = par + 3, *s2 = par + 3 + 6 * n1;
  double cb1 = cos(b1), sb1 = sin(b1), cg1 = cos(g1), sg1 = sin(g1);
  double ca2 = cos(a2), sa2 = sin(a2), cb2 = cos(b2), sb2 = sin(b2),
         cg2 = cos(g2), sg2 = sin(g2);
  double v = 0.0;
  for (int i = 0; i < n1; i++) {
    double p1[3];
    rot_zyz(1.0, 0.0, cb1, sb1, cg1, sg1, s1 + 6 * i, p1);
    for (int j = 0; j < n2; j++) {
      double p2[3];
      rot_zyz(ca2, sa2, cb2, sb2, cg2, sg2, s2 + 6 * j, p2);
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

__global__ void pes_kernel(int coll_type, int pes_kind,
                           const double *__restrict__ par,
                           const double *__restrict__ R, int n_r,
                           const double *__restrict__ beta,
                           const double *__restrict__ gamma,
                           const double *__restrict__ alpha2,
                           const double *__restrict__ beta2,
                           const double *__restrict__ gamma2, int nb, int ng,
                           int na2, int nb2, int ng2,
                           double *__restrict__ V) {
  int64_t G = (int64_t)nb * ng * na2 * nb2 * ng2;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_r * G) return;
  int ir = (int)(idx / G);
  int64_t g = idx % G;
  double r = R[ir], v = 0.0;
  if (coll_type == 1) {
    v = pes_legendre(par, r, beta[g]);
  } else if (coll_type == 4) {
    int ib = (int)(g / ng), ig = (int)(g % ng);
    v = pes_realsh(par, r, -beta[ib], -gamma[ig]);
  } else {
    int i6 = (int)(g % ng2);
    int i5 = (int)((g / ng2) % ng);
    int i4 = (int)((g / ((int64_t)ng2 * ng)) % nb2);
    int i3 = (int)((g / ((int64_t)ng2 * ng * nb2)) % nb);
    int i2 = (int)(g / ((int64_t)ng2 * ng * nb2 * nb));
    v = pes_sites(par, r, beta[i3], gamma[i5], alpha2[i2], beta2[i4],
                  gamma2[i6]);
  }
  (void)pes_kind;
  V[idx] = v;
}

__global__ void pes_points_kernel(int pes_kind, const double *__restrict__ par, int n,
                                  const double *__restrict__ geo, double *__restrict__ V) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *p = geo + 6 * (long long)i;
  double v;
  if (pes_kind == MQCT_PES_LEGENDRE) v = pes_legendre(par, p[0], p[1]);
  else if (pes_kind == MQCT_PES_REALSH) v = pes_realsh(par, p[0], p[1], p[2]);
  else v = pes_sites(par, p[0], p[1], p[2], p[3], p[4], p[5]);
  V[i] = v;
}

// W = omega * V in the requested operand precisions.
__global__ void weight_kernel(int64_t G, int n_r, const double *__restrict__ V,
                              const double *__restrict__ omega,
                              double *__restrict__ W64, float *__restrict__ W32,
                              __hip_bfloat16 *__restrict__ W0,
                              __hip_bfloat16 *__restrict__ W1,
                              __hip_bfloat16 *__restrict__ W2) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_r * G) return;
  double w = omega[idx % G] * V[idx];
  if (W64) W64[idx] = w;
  if (W32) W32[idx] = (float)w;
  if (W0) {
    // three BF16 terms, w ~ w0 + w1 + w2 to about 24 significant bits
    const __hip_bfloat16 b0 = __float2bfloat16((float)w);
    const double r1 = w - (double)__bfloat162float(b0);
    const __hip_bfloat16 b1 = __float2bfloat16((float)r1);
    const double r2 = r1 - (double)__bfloat162float(b1);
    W0[idx] = b0;
    W1[idx] = b1;
    W2[idx] = __float2bfloat16((float)r2);
  }
}

__global__ void weight_rows_kernel(int64_t G, const double *__restrict__ V,
                                   const double *__restrict__ omega, int ir1,
                                   int ir2, double *__restrict__ out) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 2 * G) return;
  int row = (int)(idx / G);
  int64_t g = idx % G;
  int ir = row == 0 ? ir1 : ir2;
  out[idx] = omega[g] * V[(int64_t)ir * G + g];
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------------------
int mqct_build_psi(mqct_hip *h) {
  const int npos = h->n_states;
  const int64_t G = h->G;
  const int64_t n = (int64_t)npos * G;
  if (h->d_psi_re.ensure(n * 8) || h->d_psi_im.ensure(n * 8))
    return mqct_fail(MQCT_ERR_NOMEM, "basis functions on the grid do not fit");
  double *re = h->d_psi_re.as<double>(), *im = h->d_psi_im.as<double>();
  if (h->coll_type == 1) {
    psi_type1_kernel<<<nblk(n, 256), 256, 0, h->stream>>>(
        h->d_pos_state.as<int>(), h->d_m12.as<int>(), h->d_chan.as<int>(),
        h->d_beta.as<double>(), npos, (int)G, re, im);
  } else if (h->coll_type == 4) {
    psi_type4_kernel<<<nblk(n, 256), 256, 0, h->stream>>>(
        h->d_pos_state.as<int>(), h->d_m12.as<int>(), h->d_chan.as<int>(),
        h->d_chan.as<int>() + h->n_states, h->d_eig1.as<double>(), h->ldvec,
        h->d_beta.as<double>(),
        h->d_gamma.as<double>(), h->n_beta, h->n_gamma, npos, re, im);
  } else {
    // one-molecule tables per level
    std::vector<int> lvl1_j, lvl2_j;
    std::vector<double> e1, e2;
    std::vector<int> chan_l1(h->n_channels), chan_l2(h->n_channels);
    {
      std::vector<std::vector<int>> keys1, keys2;
      for (int c = 0; c < h->n_channels; c++) {
        // level identity: j and the eigenvector itself
        auto find_or_add = [&](int j, const double *v, std::vector<int> &lj,
                               std::vector<double> &ev) {
          for (size_t l = 0; l < lj.size(); l++) {
            if (lj[l] != j) continue;
            bool same = true;
            for (int k = 0; k < 2 * j + 1; k++)
              if (ev[l * h->ldvec + k] != v[k]) { same = false; break; }
            if (same) return (int)l;
          }
          lj.push_back(j);
          ev.insert(ev.end(), v, v + h->ldvec);
          return (int)lj.size() - 1;
        };
        chan_l1[c] = find_or_add(h->j1_ch[c], &h->eig1[(size_t)c * h->ldvec],
                                 lvl1_j, e1);
        chan_l2[c] = find_or_add(h->j2_ch[c], &h->eig2[(size_t)c * h->ldvec],
                                 lvl2_j, e2);
      }
    }
    int jm = 0;
    for (int j : lvl1_j) jm = j > jm ? j : jm;
    for (int j : lvl2_j) jm = j > jm ? j : jm;
    const int nm = 2 * jm + 1;
    // per state: lvl1, lvl2, j1, j2 and the CG row
    std::vector<int> st(4 * (size_t)h->n_states);
    std::vector<double> cg((size_t)h->n_states * nm, 0.0);
    for (int s = 0; s < h->n_states; s++) {
      int c = h->indx_chann[s] - 1;
      int j1 = h->j1_ch[c], j2 = h->j2_ch[c];
      st[4 * s] = chan_l1[c];
      st[4 * s + 1] = chan_l2[c];
      st[4 * s + 2] = j1;
      st[4 * s + 3] = j2;
      for (int m1 = -j1; m1 <= j1; m1++) {
        int m2 = h->m12[s] - m1;
        if (m2 < -j2 || m2 > j2) continue;
        cg[(size_t)s * nm + m1 + jm] = mq_cg(j1, j2, h->j12[s], m1, m2, h->m12[s]);
      }
    }
    DevBuf d_st, d_cg, d_lj1, d_lj2, d_e1, d_e2, d_zero;
    if (d_st.ensure(st.size() * 4) || d_cg.ensure(cg.size() * 8) ||
        d_lj1.ensure(lvl1_j.size() * 4) || d_lj2.ensure(lvl2_j.size() * 4) ||
        d_e1.ensure(e1.size() * 8) || d_e2.ensure(e2.size() * 8) ||
        d_zero.ensure(8))
      return mqct_fail(MQCT_ERR_NOMEM, "type-0 tables");
    HIP_TRY(hipMemcpyAsync(d_st.p, st.data(), st.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(d_cg.p, cg.data(), cg.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(d_lj1.p, lvl1_j.data(), lvl1_j.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(d_lj2.p, lvl2_j.data(), lvl2_j.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(d_e1.p, e1.data(), e1.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(d_e2.p, e2.data(), e2.size() * 8, hipMemcpyHostToDevice, h->stream));
    const double zero = 0.0;
    HIP_TRY(hipMemcpyAsync(d_zero.p, &zero, 8, hipMemcpyHostToDevice, h->stream));
    int64_t G1 = (int64_t)h->n_beta * h->n_gamma;
    int64_t G2 = (int64_t)h->n_alpha2 * h->n_beta2 * h->n_gamma2;
    int64_t n1 = (int64_t)lvl1_j.size() * nm * G1, n2 = (int64_t)lvl2_j.size() * nm * G2;
    if (h->d_tab1.ensure(n1 * 16) || h->d_tab2.ensure(n2 * 16))
      return mqct_fail(MQCT_ERR_NOMEM, "type-0 one-molecule tables");
    tab_kernel<<<nblk(n1, 256), 256, 0, h->stream>>>(
        (int)lvl1_j.size(), jm, d_lj1.as<int>(), d_e1.as<double>(), h->ldvec,
        d_zero.as<double>(), 1, h->d_beta.as<double>(), h->n_beta,
        h->d_gamma.as<double>(), h->n_gamma, h->d_tab1.as<double2>());
    tab_kernel<<<nblk(n2, 256), 256, 0, h->stream>>>(
        (int)lvl2_j.size(), jm, d_lj2.as<int>(), d_e2.as<double>(), h->ldvec,
        h->d_alpha2.as<double>(), h->n_alpha2, h->d_beta2.as<double>(),
        h->n_beta2, h->d_gamma2.as<double>(), h->n_gamma2,
        h->d_tab2.as<double2>());
    psi_type0_kernel<<<nblk(n, 256), 256, 0, h->stream>>>(
        h->d_pos_state.as<int>(), h->d_m12.as<int>(), d_st.as<int>(),
        d_cg.as<double>(), nm, jm, h->d_tab1.as<double2>(),
        h->d_tab2.as<double2>(), h->n_alpha2, h->n_beta, h->n_beta2,
        h->n_gamma, h->n_gamma2, npos, re, im);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));  // temporaries go out of scope
  }
  HIP_TRY(hipGetLastError());
  if (!h->ref_xz) {
    if (h->d_psib_re.ensure(n * 8) || h->d_psib_im.ensure(n * 8))
      return mqct_fail(MQCT_ERR_NOMEM, "REF_XZ operand");
    psi_refxz_kernel<<<nblk(n, 256), 256, 0, h->stream>>>(
        n, re, im, h->d_psib_re.as<double>(), h->d_psib_im.as<double>());
    HIP_TRY(hipGetLastError());
  }
  h->psi_fresh = true;
  return MQCT_OK;
}

long long mqct_pes_par_len(int pes_kind, int n_par, const double *par) {
  if (n_par < 1) return -1;
  long long need;
  if (pes_kind == MQCT_PES_LEGENDRE) {
    double L = par[0];
    if (!(L >= 1 && L <= 64)) return -1;
    need = 1 + 3 * (long long)L + 3;  // A, b, C per l; core triple
  } else if (pes_kind == MQCT_PES_REALSH) {
    double lmax = par[0];
    if (!(lmax >= 0 && lmax <= 32)) return -1;
    need = 1 + 3;
    for (int l = 0; l <= (int)lmax; l++) need += 3 + 2 * l + 1;
  } else if (pes_kind == MQCT_PES_SITES) {
    if (n_par < 3) return -1;
    double n1 = par[1], n2 = par[2];
    if (!(n1 >= 1 && n1 <= 64 && n2 >= 1 && n2 <= 64)) return -1;
    need = 3 + 6 * ((long long)n1 + (long long)n2);
  } else {
    return -1;
  }
  return need <= n_par ? need : -1;
}

int mqct_eval_pes_dev(mqct_hip *h, int pes_kind, int n_par) {
  (void)n_par;
  int64_t n = (int64_t)h->n_r * h->G;
  if (h->d_V.ensure(n * 8)) return mqct_fail(MQCT_ERR_NOMEM, "V grid");
  pes_kernel<<<nblk(n, 256), 256, 0, h->stream>>>(
      h->coll_type, pes_kind, h->d_par.as<double>(), h->d_R.as<double>(),
      h->n_r, h->d_beta.as<double>(), h->d_gamma.as<double>(),
      h->d_alpha2.as<double>(), h->d_beta2.as<double>(),
      h->d_gamma2.as<double>(), h->n_beta, h->n_gamma, h->n_alpha2,
      h->n_beta2, h->n_gamma2, h->d_V.as<double>());
  HIP_TRY(hipGetLastError());
  return MQCT_OK;
}

int mqct_build_weights(mqct_hip *h, int precision) {
  int64_t n = (int64_t)h->n_r * h->G;
  double *W64 = nullptr;
  float *W32 = nullptr;
  __hip_bfloat16 *Whi = nullptr, *Wlo = nullptr, *Wlo2 = nullptr;
  if (precision == MQCT_PREC_FP64) {
    if (h->d_W64.ensure(n * 8)) return mqct_fail(MQCT_ERR_NOMEM, "W fp64");
    W64 = h->d_W64.as<double>();
  } else if (precision == MQCT_PREC_FP32) {
    if (h->d_W32.ensure(n * 4)) return mqct_fail(MQCT_ERR_NOMEM, "W fp32");
    W32 = h->d_W32.as<float>();
  } else {
    if (h->d_Whi.ensure(n * 2) || h->d_Wlo.ensure(n * 2) || h->d_Wlo2.ensure(n * 2))
      return mqct_fail(MQCT_ERR_NOMEM, "W bf16");
    Whi = h->d_Whi.as<__hip_bfloat16>();
    Wlo = h->d_Wlo.as<__hip_bfloat16>();
    Wlo2 = h->d_Wlo2.as<__hip_bfloat16>();
  }
  weight_kernel<<<nblk(n, 256), 256, 0, h->stream>>>(
      h->G, h->n_r, h->d_V.as<double>(), h->d_omega.as<double>(), W64, W32,
      Whi, Wlo, Wlo2);
  HIP_TRY(hipGetLastError());
  if (precision != MQCT_PREC_FP64) {
    int64_t np = (int64_t)h->n_states * h->G;
    if (h->d_psi32_re.ensure(np * 4) || h->d_psi32_im.ensure(np * 4))
      return mqct_fail(MQCT_ERR_NOMEM, "psi fp32");
    to_f32_kernel<<<nblk(np, 256), 256, 0, h->stream>>>(
        np, h->d_psi_re.as<double>(), h->d_psi32_re.as<float>());
    to_f32_kernel<<<nblk(np, 256), 256, 0, h->stream>>>(
        np, h->d_psi_im.as<double>(), h->d_psi32_im.as<float>());
    if (!h->ref_xz) {
      if (h->d_psib32_re.ensure(np * 4) || h->d_psib32_im.ensure(np * 4))
        return mqct_fail(MQCT_ERR_NOMEM, "psi fp32 (REF_XZ)");
      to_f32_kernel<<<nblk(np, 256), 256, 0, h->stream>>>(
          np, h->d_psib_re.as<double>(), h->d_psib32_re.as<float>());
      to_f32_kernel<<<nblk(np, 256), 256, 0, h->stream>>>(
          np, h->d_psib_im.as<double>(), h->d_psib32_im.as<float>());
    }
    HIP_TRY(hipGetLastError());
  }
  return MQCT_OK;
}

// W rows of the two cutoff R points (1-based indices), FP64.
void mqct_weight_rows(mqct_hip *h, const double *V, const double *omega,
                      int ir1, int ir2, double *out) {
  weight_rows_kernel<<<nblk(2 * h->G, 256), 256, 0, h->stream>>>(
      h->G, V, omega, ir1 - 1, ir2 - 1, out);
}

extern "C" int mqct_hip_pes_points(mqct_hip_t *h, int pes_kind, int n_par, const double *par,
                                   int n, const double *geo, double *V) {
  if (!h || !par || n_par <= 0 || n <= 0 || !geo || !V)
    return mqct_fail(MQCT_ERR_ARG, "pes_points: bad argument");
  if (pes_kind < MQCT_PES_LEGENDRE || pes_kind > MQCT_PES_SITES)
    return mqct_fail(MQCT_ERR_ARG, "pes_points: unknown potential");
  if (mqct_pes_par_len(pes_kind, n_par, par) < 0)
    return mqct_fail(MQCT_ERR_ARG, "pes_points: parameter array too short for its header");
  HIP_TRY(hipSetDevice(h->device));
  DevBuf dp, dg, dv;
  if (dp.ensure((size_t)n_par * 8) || dg.ensure((size_t)n * 48) || dv.ensure((size_t)n * 8))
    return mqct_fail(MQCT_ERR_NOMEM, "pes_points");
  HIP_TRY(hipMemcpy(dp.p, par, (size_t)n_par * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dg.p, geo, (size_t)n * 48, hipMemcpyHostToDevice));
  pes_points_kernel<<<nblk(n, 256), 256, 0, h->stream>>>(pes_kind, dp.as<double>(), n, dg.as<double>(),
                                                         dv.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(V, dv.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return MQCT_OK;
}
