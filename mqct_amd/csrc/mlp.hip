// mlp.hip -- the learned potential: a 12 -> 256 -> 256 -> 256 -> 256 -> 1
// SiLU network with BF16 weights and FP32 accumulation on MFMA, its
// backward pass and an Adam optimiser, all in the library's own kernels.
//
// Features of a geometry (R, beta1, gamma1, alpha2, beta2, gamma2):
//   4/R, exp(-(R-4)/4), cos and sin of the five angles.
// Output y; the potential is V = scale * sinh(y) + shift (cm^-1), which
// keeps relative accuracy on the repulsive wall and absolute accuracy at
// long range.  On a grid the plain potential is used below r_min.
#include "mqct_internal.h"

#include <cmath>
#include <cstring>
#include <random>

namespace {

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int NF = 12;     // features
constexpr int NFP = 32;    // features padded to one MFMA K step
constexpr int HW = 256;    // hidden width
constexpr int NH = 4;      // hidden layers
// parameter layout (fp32, row-major W[out][in]): W1 b1 W2 b2 W3 b3 W4 b4 W5 b5
constexpr int P_W1 = 0, P_B1 = P_W1 + HW * NF;
constexpr int P_W2 = P_B1 + HW, P_B2 = P_W2 + HW * HW;
constexpr int P_W3 = P_B2 + HW, P_B3 = P_W3 + HW * HW;
constexpr int P_W4 = P_B3 + HW, P_B4 = P_W4 + HW * HW;
constexpr int P_W5 = P_B4 + HW, P_B5 = P_W5 + HW;
constexpr int NPARAM = P_B5 + 1;
const int P_W[5] = {P_W1, P_W2, P_W3, P_W4, P_W5};
const int P_B[5] = {P_B1, P_B2, P_B3, P_B4, P_B5};

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float dsilu(float z) {
  float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------------------
// C[M][N] (+)= op(A)[M][K] op(B)[K][N] with bf16 operands, fp32 result.
// A(i,k) = A[i*a_r + k*a_c], B(k,j) = B[k*b_r + j*b_c].  64x64 tiles, K in
// steps of 32 staged through LDS, one 16x16x32 MFMA per (2x2 per wave).
// Epilogues: 0 store C; 1 C=acc+bias, Z=C, H=bf16(silu(C)); 2 atomic add
// into C (split-K for weight gradients); 3 store C and H=bf16(C*mask) where
// mask = dsilu(Zp) (backward through the activation of the layer below;
// C is not written).
struct GemmArgs {
  const __bf16 *A;
  long long a_r, a_c;
  const __bf16 *B;
  long long b_r, b_c;
  int M, N, K;
  float *C;
  long long ldc;
  const float *bias;
  __bf16 *H;
  long long ldh;
  const float *Zp;  // epilogue 3: pre-activation of the layer below
  int epi;
  int ksplit;       // K range per blockIdx.z
};

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[64][40];
  __shared__ __attribute__((aligned(16))) __bf16 sB[64][40];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int kbeg = blockIdx.z * g.ksplit;
  const int kend = min(g.K, kbeg + g.ksplit);
  flt4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = flt4{0, 0, 0, 0};
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    for (int e = tid; e < 64 * 32; e += 256) {
      int r = e / 32, c = e % 32;
      int i = m0 + r, k = k0 + c;
      sA[r][c] = (i < g.M && k < kend) ? g.A[(long long)i * g.a_r + (long long)k * g.a_c] : (__bf16)0.0f;
      int j = n0 + r;
      sB[r][c] = (j < g.N && k < kend) ? g.B[(long long)k * g.b_r + (long long)j * g.b_c] : (__bf16)0.0f;
    }
    __syncthreads();
    const int row = lane & 15, kq = lane >> 4;
    bf16x8 af[2], bfr[2];
    for (int a = 0; a < 2; a++)
      af[a] = *reinterpret_cast<const bf16x8 *>(&sA[wr + a * 16 + row][8 * kq]);
    for (int b = 0; b < 2; b++)
      bfr[b] = *reinterpret_cast<const bf16x8 *>(&sB[wc + b * 16 + row][8 * kq]);
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int i = 0; i < 4; i++) {
        int r = m0 + wr + a * 16 + 4 * (lane >> 4) + i;
        int c = n0 + wc + b * 16 + (lane & 15);
        if (r >= g.M || c >= g.N) continue;
        float v = acc[a][b][i];
        long long o = (long long)r * g.ldc + c;
        if (g.epi == 0) {
          g.C[o] = v;
        } else if (g.epi == 1) {
          v += g.bias[c];
          if (g.C) g.C[o] = v;
          g.H[(long long)r * g.ldh + c] = (__bf16)silu(v);
        } else if (g.epi == 2) {
          atomicAdd(&g.C[o], v);
        } else {
          g.H[(long long)r * g.ldh + c] = (__bf16)(v * dsilu(g.Zp[o]));
        }
      }
}

// features of n geometries given as [n][6] doubles (R, b1, g1, a2, b2, g2)
__global__ void features_kernel(long long n, const double *__restrict__ geo,
                                __bf16 *__restrict__ X) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *p = geo + 6 * i;
  float f[NFP];
  for (int k = 0; k < NFP; k++) f[k] = 0.0f;
  double R = p[0];
  f[0] = (float)(4.0 / R);
  f[1] = (float)exp(-(R - 4.0) / 4.0);
  for (int a = 0; a < 5; a++) {
    f[2 + 2 * a] = (float)cos(p[1 + a]);
    f[3 + 2 * a] = (float)sin(p[1 + a]);
  }
  for (int k = 0; k < NFP; k++) X[i * NFP + k] = (__bf16)f[k];
}

// features of grid points: chunk [g0, g0+n) of the flattened (R, angle)
// grid in V_3_3 order.
__global__ void grid_features_kernel(long long g0, long long n, long long G,
                                     const double *__restrict__ R,
                                     const double *__restrict__ beta,
                                     const double *__restrict__ gamma,
                                     const double *__restrict__ alpha2,
                                     const double *__restrict__ beta2,
                                     const double *__restrict__ gamma2, int nb,
                                     int ng, int nb2, int ng2,
                                     __bf16 *__restrict__ X) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long idx = g0 + i;
  int ir = (int)(idx / G);
  long long g = idx % G;
  int i6 = (int)(g % ng2);
  int i5 = (int)((g / ng2) % ng);
  int i4 = (int)((g / ((long long)ng2 * ng)) % nb2);
  int i3 = (int)((g / ((long long)ng2 * ng * nb2)) % nb);
  int i2 = (int)(g / ((long long)ng2 * ng * nb2 * nb));
  double ang[5] = {beta[i3], gamma[i5], alpha2[i2], beta2[i4], gamma2[i6]};
  double r = R[ir];
  float f[NFP];
  for (int k = 0; k < NFP; k++) f[k] = 0.0f;
  f[0] = (float)(4.0 / r);
  f[1] = (float)exp(-(r - 4.0) / 4.0);
  for (int a = 0; a < 5; a++) {
    f[2 + 2 * a] = (float)cos(ang[a]);
    f[3 + 2 * a] = (float)sin(ang[a]);
  }
  for (int k = 0; k < NFP; k++) X[i * NFP + k] = (__bf16)f[k];
}

// last layer: y = H4 . W5 + b5 (fp32 dot over 256)
__global__ void out_kernel(long long n, const __bf16 *__restrict__ H,
                           const float *__restrict__ w5, float b5,
                           float *__restrict__ y) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const __bf16 *h = H + i * HW;
  float s = b5;
  for (int k = 0; k < HW; k++) s += (float)h[k] * w5[k];
  y[i] = s;
}

// dL/dy for L = mean (y - t)^2, and dH4 = dy * W5 masked by dsilu(Z4)
__global__ void out_back_kernel(long long n, const float *__restrict__ y,
                                const float *__restrict__ t,
                                const float *__restrict__ w5,
                                const float *__restrict__ Z4,
                                __bf16 *__restrict__ dZ4, float *__restrict__ dy,
                                float *__restrict__ loss_acc) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float e = y[i] - t[i];
  float d = 2.0f * e / (float)n;
  dy[i] = d;
  atomicAdd(loss_acc, e * e / (float)n);
  for (int k = 0; k < HW; k++)
    dZ4[i * HW + k] = (__bf16)(d * w5[k] * dsilu(Z4[i * HW + k]));
}

// column sums (bias gradients) of an [n][w] bf16 matrix, and the W5/b5
// gradients from H4 and dy
__global__ void colsum_kernel(long long n, int w, const __bf16 *__restrict__ D,
                              float *__restrict__ out) {
  int c = blockIdx.x;
  float s = 0.0f;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) s += (float)D[i * w + c];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[c] = red[0] + red[1] + red[2] + red[3];
}

__global__ void w5_grad_kernel(long long n, const __bf16 *__restrict__ H,
                               const float *__restrict__ dy,
                               float *__restrict__ gw5, float *__restrict__ gb5) {
  int c = blockIdx.x;  // 0..HW (HW = bias)
  float s = 0.0f;
  for (long long i = threadIdx.x; i < n; i += blockDim.x)
    s += (c < HW ? (float)H[i * HW + c] : 1.0f) * dy[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float v = red[0] + red[1] + red[2] + red[3];
    if (c < HW) gw5[c] = v; else gb5[0] = v;
  }
}

__global__ void adam_kernel(int n, float *__restrict__ p, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 weight copies; W1 padded to [256][32]
__global__ void to_bf16_params(const float *__restrict__ p, __bf16 *__restrict__ w1p,
                               __bf16 *__restrict__ wrest) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < HW * NFP) {
    int r = i / NFP, c = i % NFP;
    w1p[i] = (__bf16)(c < NF ? p[P_W1 + r * NF + c] : 0.0f);
  }
  if (i < 3 * HW * HW) {
    int l = i / (HW * HW), o = i % (HW * HW);
    const int base[3] = {P_W2, P_W3, P_W4};
    wrest[i] = (__bf16)p[base[l] + o];
  }
}

__global__ void to_potential_kernel(long long n, const float *__restrict__ y, double scale,
                                    double shift, double *__restrict__ V) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V[i] = scale * sinh((double)y[i]) + shift;
}

__global__ void fallback_kernel(long long n, long long G, const double *__restrict__ R,
                                double r_min, const double *__restrict__ Vplain,
                                double *__restrict__ V) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (R[i / G] < r_min) V[i] = Vplain[i];
}

inline unsigned nb(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

struct MlpState {
  DevBuf p, m, v, g, w1p, wrest, loss;
  DevBuf X, Z[NH], H[NH], y, dy, D1, D2;
  DevBuf Xall, Tall, geo;            // training set: features, targets
  long long ndata = 0;
  const __bf16 *xin = nullptr;       // features of the current batch
  long long cap = 0;
  int step = 0;
  double scale = 100.0, shift = 0.0, r_min = 0.0;
};

void mqct_mlp_free(mqct_hip *h) {
  delete h->mlp;
  h->mlp = nullptr;
}

static int ensure_batch(MlpState *s, long long n, bool train) {
  if (n <= s->cap) return 0;
  int bad = s->X.ensure(n * NFP * 2) || s->y.ensure(n * 4) || s->dy.ensure(n * 4);
  for (int l = 0; l < NH; l++) {
    bad |= s->H[l].ensure(n * HW * 2);
    if (train) bad |= s->Z[l].ensure(n * HW * 4);
  }
  if (train) bad |= s->D1.ensure(n * HW * 2) || s->D2.ensure(n * HW * 2);
  if (bad) return 1;
  s->cap = train ? n : 0;
  return 0;
}

static void gemm(hipStream_t st, const GemmArgs &g) {
  dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, (g.K + g.ksplit - 1) / g.ksplit);
  gemm_bf16_kernel<<<grid, 256, 0, st>>>(g);
}

// forward of n samples whose features are in s->X; keeps Z (train) and H.
static void forward(mqct_hip *h, MlpState *s, long long n, bool keepZ) {
  const float *p = s->p.as<float>();
  for (int l = 0; l < NH; l++) {
    GemmArgs g{};
    if (l == 0) {
      g.A = s->xin; g.a_r = NFP; g.a_c = 1;
      g.B = s->w1p.as<__bf16>(); g.b_r = 1; g.b_c = NFP;  // B(k,j) = W1[j][k]
      g.K = NFP;
    } else {
      g.A = s->H[l - 1].as<__bf16>(); g.a_r = HW; g.a_c = 1;
      g.B = s->wrest.as<__bf16>() + (long long)(l - 1) * HW * HW; g.b_r = 1; g.b_c = HW;
      g.K = HW;
    }
    g.M = (int)n; g.N = HW;
    g.C = keepZ ? s->Z[l].as<float>() : nullptr; g.ldc = HW;
    g.bias = p + P_B[l];
    g.H = s->H[l].as<__bf16>(); g.ldh = HW;
    g.epi = 1;
    g.ksplit = g.K;
    gemm(h->stream, g);
  }
  out_kernel<<<nb(n, 256), 256, 0, h->stream>>>(n, s->H[NH - 1].as<__bf16>(), p + P_W5, 0.0f,
                                                 s->y.as<float>());
}

// y above lacks b5 (kept on device); add it in a tiny kernel
__global__ void add_b5(long long n, const float *__restrict__ p, float *__restrict__ y) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += p[P_B5];
}

static void forward_full(mqct_hip *h, MlpState *s, long long n, bool keepZ) {
  forward(h, s, n, keepZ);
  add_b5<<<nb(n, 256), 256, 0, h->stream>>>(n, s->p.as<float>(), s->y.as<float>());
}

static void sync_bf16(mqct_hip *h, MlpState *s) {
  to_bf16_params<<<nb(3 * HW * HW, 256), 256, 0, h->stream>>>(s->p.as<float>(), s->w1p.as<__bf16>(),
                                                             s->wrest.as<__bf16>());
}

// gradients of the mean squared error into s->g; loss into s->loss[0]
__global__ void zero_kernel(int n, float *__restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0f;
}

static void backward(mqct_hip *h, MlpState *s, long long n, const float *t) {
  float *gr = s->g.as<float>();
  const float *p = s->p.as<float>();
  zero_kernel<<<nb(NPARAM, 256), 256, 0, h->stream>>>(NPARAM, gr);
  zero_kernel<<<1, 64, 0, h->stream>>>(1, s->loss.as<float>());
  out_back_kernel<<<nb(n, 256), 256, 0, h->stream>>>(n, s->y.as<float>(), t, p + P_W5,
                                                     s->Z[NH - 1].as<float>(), s->D1.as<__bf16>(),
                                                     s->dy.as<float>(), s->loss.as<float>());
  w5_grad_kernel<<<HW + 1, 256, 0, h->stream>>>(n, s->H[NH - 1].as<__bf16>(), s->dy.as<float>(),
                                                gr + P_W5, gr + P_B5);
  __bf16 *dZ = s->D1.as<__bf16>(), *dZn = s->D2.as<__bf16>();
  for (int l = NH - 1; l >= 0; l--) {
    // bias gradient
    colsum_kernel<<<HW, 256, 0, h->stream>>>(n, HW, dZ, gr + P_B[l]);
    // weight gradient dW[o][i] = sum_s dZ[s][o] * Hprev[s][i]
    GemmArgs g{};
    g.A = dZ; g.a_r = 1; g.a_c = HW;  // A(o, s) = dZ[s][o]
    if (l == 0) {
      g.B = s->xin; g.b_r = NFP; g.b_c = 1; g.N = NF;
    } else {
      g.B = s->H[l - 1].as<__bf16>(); g.b_r = HW; g.b_c = 1; g.N = HW;
    }
    g.M = HW; g.K = (int)n;
    g.C = gr + P_W[l]; g.ldc = g.N;
    g.epi = 2;
    g.ksplit = 4096;
    gemm(h->stream, g);
    if (l > 0) {
      // dZ_{l-1} = (dZ_l W_l) * dsilu(Z_{l-1})
      GemmArgs d{};
      d.A = dZ; d.a_r = HW; d.a_c = 1;
      d.B = s->wrest.as<__bf16>() + (long long)(l - 1) * HW * HW; d.b_r = HW; d.b_c = 1;  // B(k=o, j=i) = W[o][i]
      d.M = (int)n; d.N = HW; d.K = HW;
      d.C = nullptr; d.ldc = HW;
      d.Zp = s->Z[l - 1].as<float>();
      d.H = dZn; d.ldh = HW;
      d.epi = 3;
      d.ksplit = HW;
      gemm(h->stream, d);
      __bf16 *t2 = dZ;
      dZ = dZn;
      dZn = t2;
    }
  }
  (void)t;
}

static int fail_if_no_mlp(mqct_hip *h) {
  if (!h) return mqct_fail(MQCT_ERR_ARG, "mlp: null handle");
  if (!h->mlp) return mqct_fail(MQCT_ERR_STATE, "mlp: call mqct_hip_mlp_init first");
  return 0;
}

extern "C" int mqct_hip_mlp_n_params(void) { return NPARAM; }

extern "C" int mqct_hip_mlp_init(mqct_hip_t *h, uint64_t seed) {
  if (!h) return mqct_fail(MQCT_ERR_ARG, "mlp_init: null handle");
  HIP_TRY(hipSetDevice(h->device));
  if (!h->mlp) h->mlp = new MlpState();
  MlpState *s = h->mlp;
  std::vector<float> p(NPARAM, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  const int fan_in[5] = {NF, HW, HW, HW, HW};
  const int fan_out[5] = {HW, HW, HW, HW, 1};
  for (int l = 0; l < 5; l++) {
    float sd = std::sqrt(2.0f / fan_in[l]);  // He init for SiLU
    for (int i = 0; i < fan_out[l] * fan_in[l]; i++) p[P_W[l] + i] = sd * nd(rng);
  }
  if (s->p.ensure(NPARAM * 4) || s->m.ensure(NPARAM * 4) || s->v.ensure(NPARAM * 4) ||
      s->g.ensure(NPARAM * 4) || s->w1p.ensure(HW * NFP * 2) || s->wrest.ensure(3 * HW * HW * 2) ||
      s->loss.ensure(16))
    return mqct_fail(MQCT_ERR_NOMEM, "mlp_init");
  HIP_TRY(hipMemcpy(s->p.p, p.data(), NPARAM * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(s->m.p, 0, NPARAM * 4));
  HIP_TRY(hipMemset(s->v.p, 0, NPARAM * 4));
  s->step = 0;
  sync_bf16(h, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}

extern "C" int mqct_hip_mlp_get_params(mqct_hip_t *h, float *out) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  if (!out) return mqct_fail(MQCT_ERR_ARG, "mlp_get_params: null");
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(out, h->mlp->p.p, NPARAM * 4, hipMemcpyDeviceToHost));
  return MQCT_OK;
}

extern "C" int mqct_hip_mlp_set_params(mqct_hip_t *h, const float *in) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  if (!in) return mqct_fail(MQCT_ERR_ARG, "mlp_set_params: null");
  HIP_TRY(hipMemcpy(h->mlp->p.p, in, NPARAM * 4, hipMemcpyHostToDevice));
  sync_bf16(h, h->mlp);
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}

extern "C" int mqct_hip_mlp_set_scaling(mqct_hip_t *h, double scale, double shift, double r_min) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  h->mlp->scale = scale;
  h->mlp->shift = shift;
  h->mlp->r_min = r_min;
  return MQCT_OK;
}

// Training set: n geometries [n][6] (R, beta1, gamma1, alpha2, beta2,
// gamma2) and targets y (host arrays); features are formed once on the GPU.
extern "C" int mqct_hip_mlp_set_data(mqct_hip_t *h, int n, const double *geo, const float *y) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  if (n <= 0 || !geo || !y) return mqct_fail(MQCT_ERR_ARG, "mlp_set_data: bad argument");
  MlpState *s = h->mlp;
  HIP_TRY(hipSetDevice(h->device));
  if (s->Xall.ensure((size_t)n * NFP * 2) || s->Tall.ensure((size_t)n * 4) ||
      s->geo.ensure((size_t)n * 48))
    return mqct_fail(MQCT_ERR_NOMEM, "mlp_set_data");
  HIP_TRY(hipMemcpy(s->geo.p, geo, (size_t)n * 48, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->Tall.p, y, (size_t)n * 4, hipMemcpyHostToDevice));
  features_kernel<<<nb(n, 256), 256, 0, h->stream>>>(n, s->geo.as<double>(), s->Xall.as<__bf16>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  s->ndata = n;
  return MQCT_OK;
}

static int batch_ok(mqct_hip *h, int start, int n) {
  MlpState *s = h->mlp;
  if (n <= 0 || start < 0 || (long long)start + n > s->ndata)
    return mqct_fail(MQCT_ERR_ARG, "mlp: batch outside the training set");
  if (ensure_batch(s, n, true)) return mqct_fail(MQCT_ERR_NOMEM, "mlp batch");
  s->xin = s->Xall.as<__bf16>() + (long long)start * NFP;
  return MQCT_OK;
}

extern "C" int mqct_hip_mlp_grad(mqct_hip_t *h, int start, int n, float *grad_host, float *loss) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  if (int rc = batch_ok(h, start, n)) return rc;
  MlpState *s = h->mlp;
  forward_full(h, s, n, true);
  backward(h, s, n, s->Tall.as<float>() + start);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (grad_host) HIP_TRY(hipMemcpy(grad_host, s->g.p, NPARAM * 4, hipMemcpyDeviceToHost));
  if (loss) HIP_TRY(hipMemcpy(loss, s->loss.p, 4, hipMemcpyDeviceToHost));
  return MQCT_OK;
}

extern "C" int mqct_hip_mlp_train_step(mqct_hip_t *h, int start, int n, float lr, float *loss) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  if (int rc = batch_ok(h, start, n)) return rc;
  MlpState *s = h->mlp;
  forward_full(h, s, n, true);
  backward(h, s, n, s->Tall.as<float>() + start);
  s->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.0f - std::pow(b1, (float)s->step), c2 = 1.0f - std::pow(b2, (float)s->step);
  adam_kernel<<<nb(NPARAM, 256), 256, 0, h->stream>>>(NPARAM, s->p.as<float>(), s->g.as<float>(),
                                                      s->m.as<float>(), s->v.as<float>(), lr, b1, b2,
                                                      c1, c2, 1e-8f);
  sync_bf16(h, s);
  HIP_TRY(hipGetLastError());
  if (loss) {
    HIP_TRY(hipMemcpyAsync(loss, s->loss.p, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  return MQCT_OK;
}

// Network output y for n geometries (host arrays).
extern "C" int mqct_hip_mlp_predict(mqct_hip_t *h, int n, const double *geo, float *y) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  if (n <= 0 || !geo || !y) return mqct_fail(MQCT_ERR_ARG, "mlp_predict: bad argument");
  MlpState *s = h->mlp;
  HIP_TRY(hipSetDevice(h->device));
  DevBuf dg;
  if (dg.ensure((size_t)n * 48) || ensure_batch(s, n, true)) return mqct_fail(MQCT_ERR_NOMEM, "mlp_predict");
  HIP_TRY(hipMemcpy(dg.p, geo, (size_t)n * 48, hipMemcpyHostToDevice));
  features_kernel<<<nb(n, 256), 256, 0, h->stream>>>(n, dg.as<double>(), s->X.as<__bf16>());
  s->xin = s->X.as<__bf16>();
  forward_full(h, s, n, false);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(y, s->y.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  return MQCT_OK;
}

// Potential on the whole type-0 grid from the network (the V grid of the
// surrogate form); points with R < r_min keep the plain potential.
extern "C" int mqct_hip_surrogate_vgrid(mqct_hip_t *h, int n_r, const double *R_com) {
  if (int rc = fail_if_no_mlp(h)) return rc;
  if (!h->grid_ok || h->coll_type != 0)
    return mqct_fail(MQCT_ERR_STATE, "surrogate_vgrid: needs a SYS_TYPE 0 grid");
  HIP_TRY(hipSetDevice(h->device));
  MlpState *s = h->mlp;
  bool need_plain = false;
  for (int i = 0; i < n_r; i++) need_plain |= (R_com[i] < s->r_min);
  if (need_plain && !h->v_ok)
    return mqct_fail(MQCT_ERR_STATE, "surrogate_vgrid: R below r_min needs the plain V grid first");
  if (n_r != h->n_r && h->v_ok)
    return mqct_fail(MQCT_ERR_ARG, "surrogate_vgrid: R grid differs from the plain one");
  DevBuf Vplain;
  const long long n = (long long)n_r * h->G;
  if (need_plain) {
    if (Vplain.ensure(n * 8)) return mqct_fail(MQCT_ERR_NOMEM, "surrogate_vgrid");
    HIP_TRY(hipMemcpyAsync(Vplain.p, h->d_V.p, n * 8, hipMemcpyDeviceToDevice, h->stream));
  }
  int rc = mqct_set_r(h, n_r, R_com);  // the grid, its device copy and spline factors
  if (rc) return rc;
  if (h->d_V.ensure(n * 8)) return mqct_fail(MQCT_ERR_NOMEM, "surrogate_vgrid");
  const long long CH = 1 << 20;
  if (ensure_batch(s, CH, true)) return mqct_fail(MQCT_ERR_NOMEM, "surrogate batch");
  for (long long g0 = 0; g0 < n; g0 += CH) {
    long long m = (n - g0) < CH ? (n - g0) : CH;
    grid_features_kernel<<<nb(m, 256), 256, 0, h->stream>>>(
        g0, m, h->G, h->d_R.as<double>(), h->d_beta.as<double>(), h->d_gamma.as<double>(),
        h->d_alpha2.as<double>(), h->d_beta2.as<double>(), h->d_gamma2.as<double>(), h->n_beta,
        h->n_gamma, h->n_beta2, h->n_gamma2, s->X.as<__bf16>());
    s->xin = s->X.as<__bf16>();
    forward_full(h, s, m, false);
    to_potential_kernel<<<nb(m, 256), 256, 0, h->stream>>>(m, s->y.as<float>(), s->scale, s->shift,
                                                          h->d_V.as<double>() + g0);
  }
  if (need_plain)
    fallback_kernel<<<nb(n, 256), 256, 0, h->stream>>>(n, h->G, h->d_R.as<double>(), s->r_min,
                                                       Vplain.as<double>(), h->d_V.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->v_ok = true;
  h->built = false;
  return MQCT_OK;
}
