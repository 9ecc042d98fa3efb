// xsec.hip -- cross sections from the final amplitudes on the GPU:
// INELAST_CALC (propagator.f:4752-4983) with the per-channel probabilities
// of DOTRAJECT (propagator.f:4364-4383).
//
//   P_n(t)     = sum over the states s of channel n of |a_{t,s}|^2
//   sigma[g,n] = sum_t W[g,t] P_n(t)
//
// W carries the INELAST_CALC weights (2J+1) pi / (k^2 (2j+1)) summed over
// the (J, l) pairs that map to trajectory t, one row per group: per
// collision energy for the sigma(U) table, or per J for the J-resolved
// table.  The host builds W from the L grid.
//
// Every sum runs in a fixed order (states ascending; trajectories in fixed
// chunks, chunk partial sums added in chunk order), so the result does not
// depend on the launch and repeats bit for bit.  HBM-bound: one pass over
// the amplitudes (16 bytes per state and trajectory), the rest is small.
#include "mqct_internal.h"

#include <algorithm>
#include <mutex>

namespace {

constexpr int kChunks = 64;  // trajectory chunks of the weighted sums

// First state of every channel (states are channel-major, as INI_ARRAYS
// orders them, matrix_ini.f:255-545); start[n_ch] = n_states.  bad = 1 if
// indx_chann is not 1..n_ch, non-decreasing and without gaps.
__global__ void chan_start_kernel(int N, int C, const int *__restrict__ indx,
                                  int *__restrict__ start, int *__restrict__ bad) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int c = indx[s];
  int prev = s > 0 ? indx[s - 1] : 0;
  if (c < 1 || c > C || c < prev || c > prev + 1) bad[0] = 1;
  if (s == N - 1 && c != C) bad[0] = 1;
  if (c >= 1 && c <= C && c != prev) start[c - 1] = s;
  if (s == N - 1) start[C] = N;
}

// P[t][n] = sum_{s in channel n} (Re a_s)^2 + (Im a_s)^2, q rows of ldq
// doubles [Re a (N) | Im a (N) | ...] as propagate returns them.
__global__ void opacity_kernel(long long T, int N, int ldq, const double *__restrict__ q,
                               int C, const int *__restrict__ start,
                               double *__restrict__ P) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= T * C) return;
  long long t = idx / C;
  int n = (int)(idx - t * C);
  const double *re = q + t * ldq, *im = re + N;
  double p = 0.0;
  for (int s = start[n]; s < start[n + 1]; s++) p += re[s] * re[s] + im[s] * im[s];
  P[idx] = p;
}

// partial[chunk][g][n] = sum over the chunk's trajectories t of W[g][t] P[t][n]
__global__ void weighted_partial_kernel(int T, int C, int G, int len,
                                        const double *__restrict__ W,
                                        const double *__restrict__ P,
                                        double *__restrict__ partial) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int g = blockIdx.y, chunk = blockIdx.z;
  if (n >= C) return;
  int t0 = chunk * len, t1 = min(T, t0 + len);
  const double *w = W + (long long)g * T;
  double acc = 0.0;
  for (int t = t0; t < t1; t++) acc += w[t] * P[(long long)t * C + n];
  partial[((long long)chunk * G + g) * C + n] = acc;
}

__global__ void chunk_reduce_kernel(int C, int G, int n_chunks,
                                    const double *__restrict__ partial,
                                    double *__restrict__ sigma) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= G * C) return;
  double s = 0.0;
  for (int c = 0; c < n_chunks; c++) s += partial[(long long)c * G * C + idx];
  sigma[idx] = s;
}

}  // namespace

// Doubles of working memory the kernels need beyond inputs and outputs:
// P (T x C) unless the caller takes the opacities, the chunk partial sums
// and the channel starts (as doubles, rounded up).
long long mqct_xsec_work_doubles(int T, int C, int G, bool opacity_out) {
  int n_chunks = std::min(kChunks, T);
  long long w = (long long)n_chunks * G * C + (C + 2 + 1) / 2 + 1;
  if (!opacity_out) w += (long long)T * C;
  return w;
}

// Runs the kernels on h->stream.  work: mqct_xsec_work_doubles(...)
// doubles; opacity (optional, device): P[t][n] out.  Returns after
// checking the channel table (one 4-byte copy back).
int mqct_xsec_run(mqct_hip *h, int T, int N, int ldq, const double *q, int C,
                  const int *indx, int G, const double *W, double *sigma,
                  double *opacity, double *work) {
  const int n_chunks = std::min(kChunks, T);
  const int len = (T + n_chunks - 1) / n_chunks;
  double *partial = work;
  int *start = reinterpret_cast<int *>(work + (long long)n_chunks * G * C);
  int *bad = start + C + 1;
  double *P = opacity ? opacity
                      : work + (long long)n_chunks * G * C + (C + 2 + 1) / 2 + 1;
  HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), h->stream));
  chan_start_kernel<<<(N + 255) / 256, 256, 0, h->stream>>>(N, C, indx, start, bad);
  HIP_TRY(hipGetLastError());
  int hbad = 0;
  HIP_TRY(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (hbad)
    return mqct_fail(MQCT_ERR_ARG, "cross_sections: indx_chann must number the channels 1..n_channels "
                                   "in INI_ARRAYS order (channel-major, every channel present)");
  long long TC = (long long)T * C;
  opacity_kernel<<<(unsigned)((TC + 255) / 256), 256, 0, h->stream>>>(T, N, ldq, q, C, start, P);
  HIP_TRY(hipGetLastError());
  dim3 grid((C + 255) / 256, G, n_chunks);
  weighted_partial_kernel<<<grid, 256, 0, h->stream>>>(T, C, G, len, W, P, partial);
  HIP_TRY(hipGetLastError());
  chunk_reduce_kernel<<<(G * C + 255) / 256, 256, 0, h->stream>>>(C, G, n_chunks, partial, sigma);
  HIP_TRY(hipGetLastError());
  return MQCT_OK;
}

static int xsec_args(int n_traj, int n_states, int ldq, int n_channels, int n_groups) {
  if (n_traj <= 0 || n_states <= 0 || n_channels <= 0 || n_groups <= 0)
    return mqct_fail(MQCT_ERR_ARG, "cross_sections: sizes must be positive");
  if (ldq < 2 * n_states)
    return mqct_fail(MQCT_ERR_ARG, "cross_sections: ldq must be at least 2*n_states");
  if (n_channels > n_states)
    return mqct_fail(MQCT_ERR_ARG, "cross_sections: more channels than states");
  return MQCT_OK;
}

extern "C" int mqct_hip_cross_sections(mqct_hip_t *h, int n_traj, int n_states, int ldq,
                                       const double *q, int n_channels,
                                       const int *indx_chann, int n_groups,
                                       const double *weight, double *sigma,
                                       double *opacity) {
  if (!h || !q || !indx_chann || !weight || !sigma)
    return mqct_fail(MQCT_ERR_ARG, "cross_sections: null argument");
  int rc = xsec_args(n_traj, n_states, ldq, n_channels, n_groups);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(h->device));
  const long long T = n_traj, C = n_channels, G = n_groups;
  DevBuf d_q, d_ind, d_W, d_sig, d_op, d_work;
  if (d_q.ensure((size_t)T * ldq * 8) || d_ind.ensure((size_t)n_states * 4) ||
      d_W.ensure((size_t)G * T * 8) || d_sig.ensure((size_t)G * C * 8) ||
      (opacity && d_op.ensure((size_t)T * C * 8)) ||
      d_work.ensure((size_t)mqct_xsec_work_doubles(n_traj, n_channels, n_groups, opacity != nullptr) * 8))
    return mqct_fail(MQCT_ERR_NOMEM, "cross_sections: device allocation");
  HIP_TRY(hipMemcpyAsync(d_q.p, q, (size_t)T * ldq * 8, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(d_ind.p, indx_chann, (size_t)n_states * 4, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(d_W.p, weight, (size_t)G * T * 8, hipMemcpyHostToDevice, h->stream));
  rc = mqct_xsec_run(h, n_traj, n_states, ldq, d_q.as<double>(), n_channels, d_ind.as<int>(), n_groups,
                     d_W.as<double>(), d_sig.as<double>(), opacity ? d_op.as<double>() : nullptr,
                     d_work.as<double>());
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(sigma, d_sig.p, (size_t)G * C * 8, hipMemcpyDeviceToHost, h->stream));
  if (opacity)
    HIP_TRY(hipMemcpyAsync(opacity, d_op.p, (size_t)T * C * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}

// ---------------------------------------------------------------------------
// The variant that allocates nothing (mqct_hip_cross_sections_noalloc):
// the working memory is this file-scope device array, one copy per GPU,
// held until the process ends.  Host code never passes its name to a
// kernel or copy: the per-device address comes from hipGetSymbolAddress,
// taken once per device at handle creation with that device current.
constexpr long long kScratchDoubles = 4LL << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

int mqct_xsec_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return mqct_fail(MQCT_ERR_ARG, "create: device index beyond the library's table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return MQCT_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return mqct_fail(MQCT_ERR_HIP, std::string("create: hipGetSymbolAddress(scratch): ") +
                                       hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<double *>(ptr);
  return MQCT_OK;
}

extern "C" int mqct_hip_cross_sections_noalloc(mqct_hip_t *h, int n_traj, int n_states, int ldq,
                                               const double *q_dev, int n_channels,
                                               const int *indx_chann_dev, int n_groups,
                                               const double *weight_dev, double *sigma_dev,
                                               double *opacity_dev) {
  if (!h || !q_dev || !indx_chann_dev || !weight_dev || !sigma_dev)
    return mqct_fail(MQCT_ERR_ARG, "cross_sections_noalloc: null argument");
  int rc = xsec_args(n_traj, n_states, ldq, n_channels, n_groups);
  if (rc) return rc;
  long long need = mqct_xsec_work_doubles(n_traj, n_channels, n_groups, opacity_dev != nullptr);
  if (need > kScratchDoubles)
    return mqct_fail(MQCT_ERR_ARG, "cross_sections_noalloc: needs " + std::to_string(need * 8) +
                                       " bytes of working memory, the fixed scratch holds " +
                                       std::to_string(kScratchDoubles * 8) +
                                       " (fewer trajectories, channels or groups per call, or "
                                       "mqct_hip_cross_sections)");
  double *work = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    work = (h->device >= 0 && h->device < kMaxDevices) ? g_scratch_ptr[h->device] : nullptr;
  }
  if (!work) return mqct_fail(MQCT_ERR_STATE, "cross_sections_noalloc: no scratch address for this device");
  HIP_TRY(hipSetDevice(h->device));
  rc = mqct_xsec_run(h, n_traj, n_states, ldq, q_dev, n_channels, indx_chann_dev, n_groups, weight_dev,
                     sigma_dev, opacity_dev, work);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return MQCT_OK;
}
