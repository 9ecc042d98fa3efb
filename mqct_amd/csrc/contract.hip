// contract.hip -- the quadrature contraction of the matrix build on MFMA.
//
// For every m-block (states with equal m12; INI_ARRAYS pairs only those,
// matrix_ini.f:553-630) and every R point:
//   M_ab(R) = sum_g W_g(R) * (Re psi_a(g) Re psi_b(g) + Im psi_a(g) Im psi_b(g))
// which is INTEGRATOR_TEMP's sum over the grid (pot_comp.f:1947-1952,
// 2535-2557, 3718-3722) with the wavefunctions precomputed.  As a GEMM:
// rows = pairs (a, b) of an (a-tile x b-tile) block, columns = R points,
// reduction = grid points; the A operand is formed on the fly from two
// basis-function rows staged in LDS (a Khatri-Rao product), the B operand is
// W(g, R).  Only the lower triangle a >= b is written (ind_mat order:
// k = kbase[state a] + rank of b in its block).
//
// Three operand precisions:
//   FP64 : v_mfma_f64_16x16x4f64, FP64 accumulation.
//   FP32 : v_mfma_f32_16x16x4f32, FP32 partial sums over each staged chunk
//          of grid points added into FP64 accumulators.
//   BF16X3: v_mfma_f32_16x16x32_bf16 on operands split into three BF16
//          terms each, the six cross products of order <= 2, FP32 partial
//          sums per chunk, FP64 accumulators.
#include "mqct_internal.h"

#include <algorithm>
#include <chrono>
#include <type_traits>

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int TA = 16;   // a-states per tile
constexpr int TB = 16;   // b-states per tile (MFMA rows)
constexpr int NR = 64;   // R points per workgroup (4 MFMA column tiles)
constexpr int NT = NR / 16;   // (bf16 kernel; the k4 kernel takes NRT)
constexpr int WAVES = 4; // each wave owns TA/WAVES a-states
constexpr int AW = TA / WAVES;

struct OutMap {
  const int *pos_state;
  const long long *kbase;  // per state
  const int *block_of;     // per position: block start position
  int n_r;                 // columns computed (R points)
  int ld;                  // row stride of the output per element
  double conv;
};

// Accumulator element i of a 16x16 MFMA output tile held by `lane`:
// column lane%16; row 4*(lane/16)+i for the FP32-accumulating instructions
// (f32 16x16x4, bf16 16x16x32) and lane/16+4*i for f64 16x16x4 (checked by
// mqct_hip_probe_mfma).
template <bool F64>
__device__ inline int c_row(int lane, int i) {
  return F64 ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
}
__device__ inline int c_col(int lane) { return lane & 15; }

template <bool F64>
__device__ inline void store_tile(const OutMap &om, const TileDesc &t,
                                  int a_local, int r0, int nt, int lane,
                                  const double (&v)[4], double *out) {
  int a = t.a0 + a_local;
  if (a_local >= t.na) return;
  int sa = om.pos_state[a];
  long long base = om.kbase[sa];
  int bstart = om.block_of[a];
  int col = c_col(lane);
  int r = r0 + nt * 16 + col;
  if (r >= om.n_r) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int bl = c_row<F64>(lane, i);
    if (bl >= t.nb) continue;
    int b = t.b0 + bl;
    if (b > a) continue;  // lower triangle only
    long long k = base + (b - bstart);
    out[k * om.ld + r] = v[i] * om.conv;
  }
}

// ---------------------------------------------------------------------------
// 16x16x4 family (FP64 / FP32 operands).  KC grid points per LDS stage.
template <typename T, int KC, int NRT>
__global__ __launch_bounds__(256) void contract_k4_kernel(
    const T *__restrict__ pa_re, const T *__restrict__ pa_im,
    const T *__restrict__ pb_re, const T *__restrict__ pb_im,
    const T *__restrict__ W, long long G, int n_r,
    const TileDesc *__restrict__ tiles, OutMap om, double *__restrict__ out) {
  __shared__ T sAr[TA][KC + 1], sAi[TA][KC + 1];
  __shared__ T sBr[TB][KC + 1], sBi[TB][KC + 1];
  constexpr int NTT = NRT / 16;
  __shared__ T sW[NRT][KC + 1];
  const TileDesc t = tiles[blockIdx.x];
  const int r0 = blockIdx.y * NRT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool has_im = (pa_im != nullptr);

  typedef typename std::conditional<sizeof(T) == 8, dbl4, flt4>::type accv;
  accv acc[AW][NTT];
  dbl4 acc64[AW][NTT];
#pragma unroll
  for (int a = 0; a < AW; a++)
#pragma unroll
    for (int n = 0; n < NTT; n++) {
      acc[a][n] = accv{0, 0, 0, 0};
      acc64[a][n] = dbl4{0, 0, 0, 0};
    }

  for (long long g0 = 0; g0 < G; g0 += KC) {
    // stage basis rows and weights (zero-padded at the edges)
    for (int e = tid; e < TA * KC; e += 256) {
      int r = e / KC, c = e % KC;
      long long g = g0 + c;
      bool ok = (g < G);
      bool oka = ok && r < t.na, okb = ok && r < t.nb;
      long long ia = (long long)(t.a0 + r) * G + g;
      long long ib = (long long)(t.b0 + r) * G + g;
      sAr[r][c] = oka ? pa_re[ia] : T(0);
      sBr[r][c] = okb ? pb_re[ib] : T(0);
      if (has_im) {
        sAi[r][c] = oka ? pa_im[ia] : T(0);
        sBi[r][c] = okb ? pb_im[ib] : T(0);
      }
    }
    for (int e = tid; e < NRT * KC; e += 256) {
      int r = e / KC, c = e % KC;
      long long g = g0 + c;
      int rr = r0 + r;
      sW[r][c] = (g < G && rr < n_r) ? W[(long long)rr * G + g] : T(0);
    }
    __syncthreads();
    const int row = lane & 15, kq = lane >> 4;
#pragma unroll 2
    for (int ks = 0; ks < KC; ks += 4) {
      int k = ks + kq;
      T br = sBr[row][k];
      T bi = has_im ? sBi[row][k] : T(0);
      T wv[NTT];
#pragma unroll
      for (int n = 0; n < NTT; n++) wv[n] = sW[n * 16 + row][k];
#pragma unroll
      for (int a = 0; a < AW; a++) {
        int al = wave * AW + a;
        T av = sAr[al][k] * br;
        if (has_im) av += sAi[al][k] * bi;
#pragma unroll
        for (int n = 0; n < NTT; n++) {
          if constexpr (sizeof(T) == 8)
            acc[a][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wv[n], acc[a][n], 0, 0, 0);
          else
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, wv[n], acc[a][n], 0, 0, 0);
        }
      }
    }
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int a = 0; a < AW; a++)
#pragma unroll
        for (int n = 0; n < NTT; n++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            acc64[a][n][i] += (double)acc[a][n][i];
            acc[a][n][i] = 0.0f;
          }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < AW; a++)
#pragma unroll
    for (int n = 0; n < NTT; n++) {
      double v[4];
#pragma unroll
      for (int i = 0; i < 4; i++)
        v[i] = (sizeof(T) == 8) ? (double)acc[a][n][i] : acc64[a][n][i];
      store_tile<sizeof(T) == 8>(om, t, wave * AW + a, r0, n, lane, v, out);
    }
}

// ---------------------------------------------------------------------------
// The same contraction with 16-byte staging and a double-buffered LDS: the
// next chunk's basis rows and weights are loaded into registers while the
// current chunk's MFMAs run.  Needs G to be a multiple of the vector width
// (16 / sizeof(T)); the launcher falls back to contract_k4_kernel otherwise.
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double2 type; };
template <> struct Vec16<float> { typedef float4 type; };

template <typename T, int KC, int NRT, int NW = WAVES>
__global__ __launch_bounds__(64 * NW) void contract_v_kernel(
    const T *__restrict__ pa_re, const T *__restrict__ pa_im,
    const T *__restrict__ pb_re, const T *__restrict__ pb_im,
    const T *__restrict__ W, long long G, int n_r,
    const TileDesc *__restrict__ tiles, OutMap om, double *__restrict__ out) {
  typedef typename Vec16<T>::type V;
  constexpr int VW = 16 / sizeof(T);                  // elements per vector
  constexpr int KV = KC / VW;                         // vectors per staged row
  constexpr int KP = KC + (sizeof(T) == 8 ? 2 : 4);   // padded row: 16-B aligned, conflict-free reads
  constexpr int NTT = NRT / 16;
  constexpr int BT = 64 * NW, AWn = TA / NW;        // threads; a-states per wave
  constexpr int PA = 2 * TA * KV / BT;               // A (and B) vectors per thread per chunk
  constexpr int PW = NRT * KV / BT;                  // W vectors per thread per chunk
  static_assert(PA * BT == 2 * TA * KV && PW * BT == NRT * KV && AWn * NW == TA, "staging must tile the workgroup");
  __shared__ __attribute__((aligned(16))) T sA[2][2][TA][KP];
  __shared__ __attribute__((aligned(16))) T sB[2][2][TB][KP];
  __shared__ __attribute__((aligned(16))) T sW[2][NRT][KP];
  const TileDesc t = tiles[blockIdx.x];
  const int r0 = blockIdx.y * NRT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool has_im = (pa_im != nullptr);
  // operand planes chosen by select, not by indexing a private array
  // (which would live in scratch memory)
  auto pa = [&](int plane) { return plane ? pa_im : pa_re; };
  auto pb = [&](int plane) { return plane ? pb_im : pb_re; };

  typedef typename std::conditional<sizeof(T) == 8, dbl4, flt4>::type accv;
  accv acc[AWn][NTT];
  dbl4 acc64[AWn][NTT];
#pragma unroll
  for (int a = 0; a < AWn; a++)
#pragma unroll
    for (int n = 0; n < NTT; n++) {
      acc[a][n] = accv{0, 0, 0, 0};
      acc64[a][n] = dbl4{0, 0, 0, 0};
    }
  V ra[PA], rb[PA], rw[PW];
  const V vz = {};
  auto load = [&](long long g0) {
#pragma unroll
    for (int j = 0; j < PA; j++) {
      const int idx = tid + BT * j, plane = idx / (TA * KV), rem = idx % (TA * KV);
      const int row = rem / KV, kv = rem % KV;
      const long long g = g0 + (long long)kv * VW;
      const bool okp = (plane == 0 || has_im) && g < G;
      ra[j] = (okp && row < t.na) ? *reinterpret_cast<const V *>(pa(plane) + (long long)(t.a0 + row) * G + g) : vz;
      rb[j] = (okp && row < t.nb) ? *reinterpret_cast<const V *>(pb(plane) + (long long)(t.b0 + row) * G + g) : vz;
    }
#pragma unroll
    for (int j = 0; j < PW; j++) {
      const int idx = tid + BT * j, row = idx / KV, kv = idx % KV;
      const long long g = g0 + (long long)kv * VW;
      const int rr = r0 + row;
      rw[j] = (g < G && rr < n_r) ? *reinterpret_cast<const V *>(W + (long long)rr * G + g) : vz;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < PA; j++) {
      const int idx = tid + BT * j, plane = idx / (TA * KV), rem = idx % (TA * KV);
      const int row = rem / KV, kv = rem % KV;
      *reinterpret_cast<V *>(&sA[buf][plane][row][kv * VW]) = ra[j];
      *reinterpret_cast<V *>(&sB[buf][plane][row][kv * VW]) = rb[j];
    }
#pragma unroll
    for (int j = 0; j < PW; j++) {
      const int idx = tid + BT * j, row = idx / KV, kv = idx % KV;
      *reinterpret_cast<V *>(&sW[buf][row][kv * VW]) = rw[j];
    }
  };
  const long long nch = (G + KC - 1) / KC;
  load(0);
  store(0);
  __syncthreads();
  const int row = lane & 15, kq = lane >> 4;
  for (long long ch = 0; ch < nch; ch++) {
    const int buf = (int)(ch & 1);
    const bool more = ch + 1 < nch;
    if (more) load((ch + 1) * KC);
#pragma unroll 2
    for (int ks = 0; ks < KC; ks += 4) {
      const int k = ks + kq;
      const T br = sB[buf][0][row][k];
      const T bi = has_im ? sB[buf][1][row][k] : T(0);
      T wv[NTT];
#pragma unroll
      for (int n = 0; n < NTT; n++) wv[n] = sW[buf][n * 16 + row][k];
#pragma unroll
      for (int a = 0; a < AWn; a++) {
        const int al = wave * AWn + a;
        T av = sA[buf][0][al][k] * br;
        if (has_im) av += sA[buf][1][al][k] * bi;
#pragma unroll
        for (int n = 0; n < NTT; n++) {
          if constexpr (sizeof(T) == 8)
            acc[a][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wv[n], acc[a][n], 0, 0, 0);
          else
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, wv[n], acc[a][n], 0, 0, 0);
        }
      }
    }
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int a = 0; a < AWn; a++)
#pragma unroll
        for (int n = 0; n < NTT; n++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            acc64[a][n][i] += (double)acc[a][n][i];
            acc[a][n][i] = 0.0f;
          }
    }
    if (more) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < AWn; a++)
#pragma unroll
    for (int n = 0; n < NTT; n++) {
      double v[4];
#pragma unroll
      for (int i = 0; i < 4; i++)
        v[i] = (sizeof(T) == 8) ? (double)acc[a][n][i] : acc64[a][n][i];
      store_tile<sizeof(T) == 8>(om, t, wave * AWn + a, r0, n, lane, v, out);
    }
}

// ---------------------------------------------------------------------------
// Eight FP32 values split by truncation into three BF16 terms each,
// x = x0 + x1 + x2 exactly: x0 keeps the top 8 significant bits of x (the
// upper half of its encoding), x1 the top 8 of the remainder, and x2 the
// rest, which has at most 8 significant bits left.  Bit masks and
// subtractions replace round-trip conversions, and one byte permute packs
// two upper halves into a dword of the MFMA operand.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ inline void split3_pack8(const float (&x)[8], bf16x8 &t0, bf16x8 &t1, bf16x8 &t2) {
  u32x4 w0, w1, w2;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    unsigned h[2], m[2], l[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const float v = x[2 * j + q];
      const unsigned hb = __float_as_uint(v) & 0xffff0000u;
      const float r1 = v - __uint_as_float(hb);
      const unsigned mb = __float_as_uint(r1) & 0xffff0000u;
      h[q] = hb;
      m[q] = mb;
      l[q] = __float_as_uint(r1 - __uint_as_float(mb));
    }
    // bytes 2-3 of the first operand's pair member, then of the second
    w0[j] = __builtin_amdgcn_perm(h[1], h[0], 0x07060302u);
    w1[j] = __builtin_amdgcn_perm(m[1], m[0], 0x07060302u);
    w2[j] = __builtin_amdgcn_perm(l[1], l[0], 0x07060302u);
  }
  t0 = __builtin_bit_cast(bf16x8, w0);
  t1 = __builtin_bit_cast(bf16x8, w1);
  t2 = __builtin_bit_cast(bf16x8, w2);
}

// ---------------------------------------------------------------------------
// Three-term split-BF16 kernel.  Each operand is carried as three BF16
// terms, x = x0 + x1 + x2 (about 24 significant bits): the weights come
// split from HBM, the pair product p = Re_a Re_b + Im_a Im_b (FP32) is split
// on the fly.  The six cross products of order <= 2,
//   p2 w0 + p1 w1 + p0 w2 + p1 w0 + p0 w1 + p0 w0,
// run on v_mfma_f32_16x16x32_bf16 into FP32 partial sums over each staged
// chunk of KC grid points, added into FP64 accumulators; the omitted terms
// are below 2^-24 relative.  (The weights are split by rounding to nearest
// where they are made; the pair products by truncation, split3_pack8.)

template <int KC>
__global__ __launch_bounds__(256) void contract_bf16x3_kernel(
    const float *__restrict__ pa_re, const float *__restrict__ pa_im,
    const float *__restrict__ pb_re, const float *__restrict__ pb_im,
    const __bf16 *__restrict__ W0, const __bf16 *__restrict__ W1, const __bf16 *__restrict__ W2,
    long long G, int n_r, const TileDesc *__restrict__ tiles, OutMap om,
    double *__restrict__ out) {
  __shared__ float sAr[TA][KC + 4], sAi[TA][KC + 4];
  __shared__ float sBr[TB][KC + 4], sBi[TB][KC + 4];
  __shared__ __attribute__((aligned(16))) __bf16 sW[3][NR][KC + 8];
  const TileDesc t = tiles[blockIdx.x];
  const int r0 = blockIdx.y * NR;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool has_im = (pa_im != nullptr);
  flt4 acc[AW][NT];
  dbl4 acc64[AW][NT];
#pragma unroll
  for (int a = 0; a < AW; a++)
#pragma unroll
    for (int n = 0; n < NT; n++) {
      acc[a][n] = flt4{0, 0, 0, 0};
      acc64[a][n] = dbl4{0, 0, 0, 0};
    }
  const __bf16 *Wp[3] = {W0, W1, W2};
  for (long long g0 = 0; g0 < G; g0 += KC) {
    for (int e = tid; e < TA * KC; e += 256) {
      int r = e / KC, c = e % KC;
      long long g = g0 + c;
      bool ok = (g < G);
      bool oka = ok && r < t.na, okb = ok && r < t.nb;
      long long ia = (long long)(t.a0 + r) * G + g;
      long long ib = (long long)(t.b0 + r) * G + g;
      sAr[r][c] = oka ? pa_re[ia] : 0.0f;
      sBr[r][c] = okb ? pb_re[ib] : 0.0f;
      if (has_im) {
        sAi[r][c] = oka ? pa_im[ia] : 0.0f;
        sBi[r][c] = okb ? pb_im[ib] : 0.0f;
      }
    }
    for (int e = tid; e < NR * KC; e += 256) {
      int r = e / KC, c = e % KC;
      long long g = g0 + c;
      int rr = r0 + r;
      bool ok = (g < G && rr < n_r);
#pragma unroll
      for (int q = 0; q < 3; q++) sW[q][r][c] = ok ? Wp[q][(long long)rr * G + g] : (__bf16)0.0f;
    }
    __syncthreads();
    const int row = lane & 15, kq = lane >> 4;
    for (int ks = 0; ks < KC; ks += 32) {
      int k0 = ks + 8 * kq;
      float br[8], bi[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        br[j] = sBr[row][k0 + j];
        bi[j] = has_im ? sBi[row][k0 + j] : 0.0f;
      }
#pragma unroll
      for (int a = 0; a < AW; a++) {
        int al = wave * AW + a;
        bf16x8 p0, p1, p2;
        float pv[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          pv[j] = sAr[al][k0 + j] * br[j];
          if (has_im) pv[j] += sAi[al][k0 + j] * bi[j];
        }
        split3_pack8(pv, p0, p1, p2);
#pragma unroll
        for (int n = 0; n < NT; n++) {
          const bf16x8 w0 = *reinterpret_cast<const bf16x8 *>(&sW[0][n * 16 + row][k0]);
          const bf16x8 w1 = *reinterpret_cast<const bf16x8 *>(&sW[1][n * 16 + row][k0]);
          const bf16x8 w2 = *reinterpret_cast<const bf16x8 *>(&sW[2][n * 16 + row][k0]);
          acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p2, w0, acc[a][n], 0, 0, 0);
          acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p1, w1, acc[a][n], 0, 0, 0);
          acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p0, w2, acc[a][n], 0, 0, 0);
          acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p1, w0, acc[a][n], 0, 0, 0);
          acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p0, w1, acc[a][n], 0, 0, 0);
          acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p0, w0, acc[a][n], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int a = 0; a < AW; a++)
#pragma unroll
      for (int n = 0; n < NT; n++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          acc64[a][n][i] += (double)acc[a][n][i];
          acc[a][n][i] = 0.0f;
        }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < AW; a++)
#pragma unroll
    for (int n = 0; n < NT; n++) {
      double v[4];
#pragma unroll
      for (int i = 0; i < 4; i++) v[i] = acc64[a][n][i];
      store_tile<false>(om, t, wave * AW + a, r0, n, lane, v, out);
    }
}

// The three-term split-BF16 contraction with 16-byte staging and a
// double-buffered LDS (needs G to be a multiple of 8).  IL: the six products
// are issued term by term over all of the wave's accumulators, so
// consecutive MFMAs write different accumulators instead of a chain of six
// into one (each accumulator still receives its terms in the same order, so
// the sums are the same).
template <int KC, int NW = WAVES, int NRB = NR, bool IL = false>
__global__ __launch_bounds__(64 * NW) void contract_bf16x3_v_kernel(
    const float *__restrict__ pa_re, const float *__restrict__ pa_im,
    const float *__restrict__ pb_re, const float *__restrict__ pb_im,
    const __bf16 *__restrict__ W0, const __bf16 *__restrict__ W1, const __bf16 *__restrict__ W2,
    long long G, int n_r, const TileDesc *__restrict__ tiles, OutMap om,
    double *__restrict__ out) {
  constexpr int KP = KC + 4;             // fp32 rows
  constexpr int KPW = KC + 8;            // bf16 rows
  constexpr int KV = KC / 4;             // float4 per fp32 row
  constexpr int KVW = KC / 8;            // 8 x bf16 per weight row
  constexpr int BT = 64 * NW, AWn = TA / NW;  // threads; a-states per wave
  constexpr int NTB = NRB / 16;               // R column tiles per workgroup
  constexpr int PA = 2 * TA * KV / BT;        // A (and B) vectors per thread
  constexpr int PW = 3 * NRB * KVW / BT;       // weight vectors per thread
  static_assert(PA * BT == 2 * TA * KV && PW * BT == 3 * NRB * KVW && AWn * NW == TA, "staging must tile the workgroup");
  __shared__ __attribute__((aligned(16))) float sA[2][2][TA][KP];
  __shared__ __attribute__((aligned(16))) float sB[2][2][TB][KP];
  __shared__ __attribute__((aligned(16))) __bf16 sW[2][3][NRB][KPW];
  const TileDesc t = tiles[blockIdx.x];
  const int r0 = blockIdx.y * NRB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool has_im = (pa_im != nullptr);
  auto pa = [&](int plane) { return plane ? pa_im : pa_re; };
  auto pb = [&](int plane) { return plane ? pb_im : pb_re; };
  auto Wp = [&](int q) { return q == 0 ? W0 : (q == 1 ? W1 : W2); };
  flt4 acc[AWn][NTB];
  dbl4 acc64[AWn][NTB];
#pragma unroll
  for (int a = 0; a < AWn; a++)
#pragma unroll
    for (int n = 0; n < NTB; n++) {
      acc[a][n] = flt4{0, 0, 0, 0};
      acc64[a][n] = dbl4{0, 0, 0, 0};
    }
  float4 ra[PA], rb[PA];
  uint4 rw[PW];
  auto load = [&](long long g0) {
#pragma unroll
    for (int j = 0; j < PA; j++) {
      const int idx = tid + BT * j, plane = idx / (TA * KV), rem = idx % (TA * KV);
      const int row = rem / KV, kv = rem % KV;
      const long long g = g0 + (long long)kv * 4;
      const bool okp = (plane == 0 || has_im) && g < G;
      ra[j] = (okp && row < t.na) ? *reinterpret_cast<const float4 *>(pa(plane) + (long long)(t.a0 + row) * G + g)
                                  : float4{0, 0, 0, 0};
      rb[j] = (okp && row < t.nb) ? *reinterpret_cast<const float4 *>(pb(plane) + (long long)(t.b0 + row) * G + g)
                                  : float4{0, 0, 0, 0};
    }
#pragma unroll
    for (int j = 0; j < PW; j++) {
      const int idx = tid + BT * j, q = idx / (NRB * KVW), rem = idx % (NRB * KVW);
      const int row = rem / KVW, kv = rem % KVW;
      const long long g = g0 + (long long)kv * 8;
      const int rr = r0 + row;
      rw[j] = (g < G && rr < n_r) ? *reinterpret_cast<const uint4 *>(Wp(q) + (long long)rr * G + g)
                                  : uint4{0, 0, 0, 0};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < PA; j++) {
      const int idx = tid + BT * j, plane = idx / (TA * KV), rem = idx % (TA * KV);
      const int row = rem / KV, kv = rem % KV;
      *reinterpret_cast<float4 *>(&sA[buf][plane][row][kv * 4]) = ra[j];
      *reinterpret_cast<float4 *>(&sB[buf][plane][row][kv * 4]) = rb[j];
    }
#pragma unroll
    for (int j = 0; j < PW; j++) {
      const int idx = tid + BT * j, q = idx / (NRB * KVW), rem = idx % (NRB * KVW);
      const int row = rem / KVW, kv = rem % KVW;
      *reinterpret_cast<uint4 *>(&sW[buf][q][row][kv * 8]) = rw[j];
    }
  };
  const long long nch = (G + KC - 1) / KC;
  load(0);
  store(0);
  __syncthreads();
  const int row = lane & 15, kq = lane >> 4;
  for (long long ch = 0; ch < nch; ch++) {
    const int buf = (int)(ch & 1);
    const bool more = ch + 1 < nch;
    if (more) load((ch + 1) * KC);
    for (int ks = 0; ks < KC; ks += 32) {
      const int k0 = ks + 8 * kq;
      // eight consecutive grid points per lane, read as two 16-byte vectors
      // (rows are 16-byte aligned and k0 is a multiple of 8)
      auto row8 = [&](const float *src, float (&dst)[8]) {
        const float4 u = reinterpret_cast<const float4 *>(src)[0], v = reinterpret_cast<const float4 *>(src)[1];
        dst[0] = u.x; dst[1] = u.y; dst[2] = u.z; dst[3] = u.w;
        dst[4] = v.x; dst[5] = v.y; dst[6] = v.z; dst[7] = v.w;
      };
      float br[8], bi[8] = {};
      row8(&sB[buf][0][row][k0], br);
      if (has_im) row8(&sB[buf][1][row][k0], bi);
      // the split pair products of all the wave's a-states first, then each
      // R tile's three weight fragments read from LDS once for all of them
      bf16x8 p0[AWn], p1[AWn], p2[AWn];
#pragma unroll
      for (int a = 0; a < AWn; a++) {
        const int al = wave * AWn + a;
        float ar[8], ai[8], pv[8];
        row8(&sA[buf][0][al][k0], ar);
        if (has_im) row8(&sA[buf][1][al][k0], ai);
#pragma unroll
        for (int j = 0; j < 8; j++) {
          pv[j] = ar[j] * br[j];
          if (has_im) pv[j] += ai[j] * bi[j];
        }
        split3_pack8(pv, p0[a], p1[a], p2[a]);
      }
      if constexpr (IL) {
        bf16x8 w[3][NTB];
#pragma unroll
        for (int n = 0; n < NTB; n++)
#pragma unroll
          for (int q = 0; q < 3; q++) w[q][n] = *reinterpret_cast<const bf16x8 *>(&sW[buf][q][n * 16 + row][k0]);
        // terms p2 w0, p1 w1, p0 w2, p1 w0, p0 w1, p0 w0, each over all accumulators
        auto term = [&](const bf16x8 (&p)[AWn], const bf16x8 (&wq)[NTB]) {
#pragma unroll
          for (int n = 0; n < NTB; n++)
#pragma unroll
            for (int a = 0; a < AWn; a++)
              acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p[a], wq[n], acc[a][n], 0, 0, 0);
        };
        term(p2, w[0]);
        term(p1, w[1]);
        term(p0, w[2]);
        term(p1, w[0]);
        term(p0, w[1]);
        term(p0, w[0]);
      } else {
#pragma unroll
        for (int n = 0; n < NTB; n++) {
          const bf16x8 w0 = *reinterpret_cast<const bf16x8 *>(&sW[buf][0][n * 16 + row][k0]);
          const bf16x8 w1 = *reinterpret_cast<const bf16x8 *>(&sW[buf][1][n * 16 + row][k0]);
          const bf16x8 w2 = *reinterpret_cast<const bf16x8 *>(&sW[buf][2][n * 16 + row][k0]);
#pragma unroll
          for (int a = 0; a < AWn; a++) {
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p2[a], w0, acc[a][n], 0, 0, 0);
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p1[a], w1, acc[a][n], 0, 0, 0);
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p0[a], w2, acc[a][n], 0, 0, 0);
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p1[a], w0, acc[a][n], 0, 0, 0);
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p0[a], w1, acc[a][n], 0, 0, 0);
            acc[a][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(p0[a], w0, acc[a][n], 0, 0, 0);
          }
        }
      }
    }
#pragma unroll
    for (int a = 0; a < AWn; a++)
#pragma unroll
      for (int n = 0; n < NTB; n++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          acc64[a][n][i] += (double)acc[a][n][i];
          acc[a][n][i] = 0.0f;
        }
    if (more) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < AWn; a++)
#pragma unroll
    for (int n = 0; n < NTB; n++) {
      double v[4];
#pragma unroll
      for (int i = 0; i < 4; i++) v[i] = acc64[a][n][i];
      store_tile<false>(om, t, wave * AWn + a, r0, n, lane, v, out);
    }
}

// ---------------------------------------------------------------------------
// FP64 sums of every pair at the two cutoff R points, on the vector ALU (a
// 16-column MFMA tile would be 7/8 empty).  With two output columns the
// work is bound by reading the basis rows, so the tiles here are 64 x 64
// pairs (CT): 128 rows read per 4096 pairs.  Each thread owns the 4 x 4
// pairs (ta + 16 i, tb + 16 j) of the tile, so one pass over LDS feeds
// sixteen pair products; basis rows and the two weight rows pass through
// LDS in chunks of CK grid points, two points per 16-byte read.  The grid
// is split into CS slices for enough workgroups; per-slice partial sums go
// to part[slice][pair][2] and cutoff_reduce_kernel adds them in slice
// order, so the result does not depend on scheduling.
constexpr int CS = 16;  // grid slices of the cutoff sums
constexpr int CT = 64;  // states per side of a cutoff tile
template <int CK>
__global__ __launch_bounds__(256) void cutoff_fp64_kernel(
    const double *__restrict__ pa_re, const double *__restrict__ pa_im,
    const double *__restrict__ pb_re, const double *__restrict__ pb_im,
    const double *__restrict__ Wc, long long G, const TileDesc *__restrict__ tiles, OutMap om, long long T,
    double *__restrict__ part) {
  constexpr int CP = CK + 2;  // row stride: 16-byte aligned rows
  __shared__ __attribute__((aligned(16))) double sA[2][CT][CP], sB[2][CT][CP], sW[2][CK];
  const TileDesc t = tiles[blockIdx.x];
  const int tid = threadIdx.x, ta = tid >> 4, tb = tid & 15;
  const bool has_im = pa_im != nullptr;
  const long long gs = (G + CS - 1) / CS, g_lo = blockIdx.y * gs, g_hi = min(G, g_lo + gs);
  double s[4][4][2] = {};  // [a][b][R point]
  for (long long g0 = g_lo; g0 < g_hi; g0 += CK) {
    for (int e = tid; e < CT * CK; e += 256) {
      const int r = e / CK, c = e % CK;
      const long long g = g0 + c;
      const bool ok = g < g_hi, oka = ok && r < t.na, okb = ok && r < t.nb;
      const long long ia = (long long)(t.a0 + r) * G + g, ib = (long long)(t.b0 + r) * G + g;
      sA[0][r][c] = oka ? pa_re[ia] : 0.0;
      sB[0][r][c] = okb ? pb_re[ib] : 0.0;
      sA[1][r][c] = (oka && has_im) ? pa_im[ia] : 0.0;
      sB[1][r][c] = (okb && has_im) ? pb_im[ib] : 0.0;
    }
    for (int e = tid; e < 2 * CK; e += 256) {
      const int q = e / CK, c = e % CK;
      const long long g = g0 + c;
      sW[q][c] = g < g_hi ? Wc[q * G + g] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < CK; c += 2) {
      double2 ar[4], ai[4], br[4], bi[4];
#pragma unroll
      for (int h = 0; h < 4; h++) {
        ar[h] = *reinterpret_cast<const double2 *>(&sA[0][ta + 16 * h][c]);
        ai[h] = *reinterpret_cast<const double2 *>(&sA[1][ta + 16 * h][c]);
        br[h] = *reinterpret_cast<const double2 *>(&sB[0][tb + 16 * h][c]);
        bi[h] = *reinterpret_cast<const double2 *>(&sB[1][tb + 16 * h][c]);
      }
      const double2 w0 = *reinterpret_cast<const double2 *>(&sW[0][c]);
      const double2 w1 = *reinterpret_cast<const double2 *>(&sW[1][c]);
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
          const double px = ar[x].x * br[y].x + ai[x].x * bi[y].x;
          const double py = ar[x].y * br[y].y + ai[x].y * bi[y].y;
          s[x][y][0] += w0.x * px;
          s[x][y][1] += w1.x * px;
          s[x][y][0] += w0.y * py;
          s[x][y][1] += w1.y * py;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int la = ta + 16 * x, lb = tb + 16 * y;
      if (la >= t.na || lb >= t.nb) continue;
      const int a = t.a0 + la, b = t.b0 + lb;
      if (b > a) continue;  // lower triangle only
      const long long k = om.kbase[om.pos_state[a]] + (b - om.block_of[a]);
      double *o = part + ((long long)blockIdx.y * T + k) * 2;
      o[0] = s[x][y][0];
      o[1] = s[x][y][1];
    }
}

// cutv[k][j] = conv * sum over slices of part[slice][k][j], slices in order
__global__ void cutoff_reduce_kernel(const double *__restrict__ part, long long T, double conv,
                                     double *__restrict__ cutv) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * T) return;
  double acc = 0.0;
  for (int sl = 0; sl < CS; sl++) acc += part[(long long)sl * 2 * T + i];
  cutv[i] = acc * conv;
}

// ---------------------------------------------------------------------------
// Single-instruction probes of the lane layouts assumed above.
__global__ void probe_kernel(int which, const double *A, const double *B,
                             double *C) {
  int lane = threadIdx.x;
  int row = lane & 15, kq = lane >> 4;
  if (which == 0) {
    dbl4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(A[row * 4 + kq], B[kq * 16 + row], c, 0, 0, 0);
    for (int i = 0; i < 4; i++) C[c_row<true>(lane, i) * 16 + c_col(lane)] = c[i];
  } else if (which == 1) {
    flt4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f32_16x16x4f32((float)A[row * 4 + kq], (float)B[kq * 16 + row], c, 0, 0, 0);
    for (int i = 0; i < 4; i++) C[c_row<false>(lane, i) * 16 + c_col(lane)] = c[i];
  } else {
    bf16x8 a, b;
    for (int j = 0; j < 8; j++) {
      a[j] = (__bf16)(float)A[row * 32 + 8 * kq + j];
      b[j] = (__bf16)(float)B[(8 * kq + j) * 16 + row];
    }
    flt4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    for (int i = 0; i < 4; i++) C[c_row<false>(lane, i) * 16 + c_col(lane)] = c[i];
  }
}

}  // namespace

// Tile list: all (a-tile, b-tile) blocks with b-tile <= a-tile inside each
// m-block.
static void make_tiles(const mqct_hip *h, std::vector<TileDesc> &tiles) {
  tiles.clear();
  for (size_t blk = 0; blk + 1 < h->block_start.size(); blk++) {
    int s0 = h->block_start[blk], s1 = h->block_start[blk + 1];
    for (int a0 = s0; a0 < s1; a0 += TA)
      for (int b0 = s0; b0 <= a0; b0 += TB) {
        TileDesc t;
        t.a0 = a0;
        t.na = (s1 - a0) < TA ? (s1 - a0) : TA;
        t.b0 = b0;
        t.nb = (s1 - b0) < TB ? (s1 - b0) : TB;
        tiles.push_back(t);
      }
  }
}

static int ensure_tiles(mqct_hip *h, int &ntiles) {
  std::vector<TileDesc> tiles;
  make_tiles(h, tiles);
  ntiles = (int)tiles.size();
  if (tiles.empty()) return MQCT_OK;
  if (h->d_tiles.ensure(tiles.size() * sizeof(TileDesc)))
    return mqct_fail(MQCT_ERR_NOMEM, "tiles");
  HIP_TRY(hipMemcpyAsync(h->d_tiles.p, tiles.data(),
                         tiles.size() * sizeof(TileDesc),
                         hipMemcpyHostToDevice, h->stream));
  return MQCT_OK;
}

static OutMap out_map(mqct_hip *h, int n_cols, int ld, double conv) {
  OutMap om;
  om.pos_state = h->d_pos_state.as<int>();
  om.kbase = h->d_kbase.as<long long>();
  // block_of per position is stored after kbase in d_kbase
  om.block_of = reinterpret_cast<const int *>(om.kbase + h->n_states);
  om.n_r = n_cols;
  om.ld = ld;
  om.conv = conv;
  return om;
}

// Contraction variants per precision form: grid points per LDS stage and
// waves per workgroup (each wave owns TA / waves a-states of the tile).
struct CVariant {
  int kc, nw, nr;  // grid points per stage, waves, R points per workgroup
};
constexpr CVariant kCVar[3][4] = {{{32, 4, 64}, {32, 8, 64}, {32, 8, 128}, {0, 0, 0}},  // FP64
                                  {{64, 4, 64}, {64, 8, 64}, {32, 4, 64}, {64, 8, 128}},  // FP32
                                  {{64, 4, 64}, {64, 8, 64}, {32, 4, 64}, {64, 8, 64}}};  // split-BF16 (3: term-interleaved)
constexpr int kNCVar[3] = {3, 4, 4};

// One launch of variant v of a precision form over tiles [0, ntiles).
static void launch_contraction(mqct_hip *h, int precision, int v, unsigned ntiles, const OutMap &om, double *out) {
  const int nrv = (v >= 0 && kCVar[precision][v].nr > 0) ? kCVar[precision][v].nr : NR;
  const dim3 grid(ntiles, (unsigned)((h->n_r + nrv - 1) / nrv));
  const dim3 grid64(ntiles, (unsigned)((h->n_r + NR - 1) / NR));  // the fallback kernels take 64 R
  const bool im = (h->coll_type != 1) || !h->ref_xz;
  hipStream_t st = h->stream;
  const TileDesc *tl = h->d_tiles.as<TileDesc>();
  if (precision == MQCT_PREC_FP64) {
    const double *ar = h->d_psi_re.as<double>(), *ai = im ? h->d_psi_im.as<double>() : nullptr;
    const double *br = h->ref_xz ? h->d_psi_re.as<double>() : h->d_psib_re.as<double>();
    const double *bi = im ? (h->ref_xz ? h->d_psi_im.as<double>() : h->d_psib_im.as<double>()) : nullptr;
    const double *W = h->d_W64.as<double>();
    if (h->G % 2 != 0) {
      contract_k4_kernel<double, 32, 64><<<grid64, 256, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
      return;
    }
    if (v == 1) contract_v_kernel<double, 32, 64, 8><<<grid, 512, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
    else if (v == 2) contract_v_kernel<double, 32, 128, 8><<<grid, 512, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
    else contract_v_kernel<double, 32, 64, 4><<<grid, 256, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
    return;
  }
  const float *ar = h->d_psi32_re.as<float>(), *ai = im ? h->d_psi32_im.as<float>() : nullptr;
  const float *br = h->ref_xz ? h->d_psi32_re.as<float>() : h->d_psib32_re.as<float>();
  const float *bi = im ? (h->ref_xz ? h->d_psi32_im.as<float>() : h->d_psib32_im.as<float>()) : nullptr;
  if (precision == MQCT_PREC_FP32) {
    const float *W = h->d_W32.as<float>();
    if (h->G % 4 != 0) {
      contract_k4_kernel<float, 64, 64><<<grid64, 256, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
      return;
    }
    if (v == 1) contract_v_kernel<float, 64, 64, 8><<<grid, 512, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
    else if (v == 2) contract_v_kernel<float, 32, 64, 4><<<grid, 256, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
    else if (v == 3) contract_v_kernel<float, 64, 128, 8><<<grid, 512, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
    else contract_v_kernel<float, 64, 64, 4><<<grid, 256, 0, st>>>(ar, ai, br, bi, W, h->G, h->n_r, tl, om, out);
    return;
  }
  const __bf16 *w0 = h->d_Whi.as<__bf16>(), *w1 = h->d_Wlo.as<__bf16>(), *w2 = h->d_Wlo2.as<__bf16>();
  if (h->G % 8 != 0) {
    contract_bf16x3_kernel<64><<<grid64, 256, 0, st>>>(ar, ai, br, bi, w0, w1, w2, h->G, h->n_r, tl, om, out);
    return;
  }
  if (v == 1) contract_bf16x3_v_kernel<64, 8><<<grid, 512, 0, st>>>(ar, ai, br, bi, w0, w1, w2, h->G, h->n_r, tl, om, out);
  else if (v == 2) contract_bf16x3_v_kernel<32, 4><<<grid, 256, 0, st>>>(ar, ai, br, bi, w0, w1, w2, h->G, h->n_r, tl, om, out);
  else if (v == 3) contract_bf16x3_v_kernel<64, 8, 64, true><<<grid, 512, 0, st>>>(ar, ai, br, bi, w0, w1, w2, h->G, h->n_r, tl, om, out);
  else contract_bf16x3_v_kernel<64, 4><<<grid, 256, 0, st>>>(ar, ai, br, bi, w0, w1, w2, h->G, h->n_r, tl, om, out);
}

int mqct_run_contraction(mqct_hip *h, int precision, double conv) {
  int ntiles = 0, rc;
  if ((rc = ensure_tiles(h, ntiles))) return rc;
  if (ntiles == 0) return MQCT_OK;
  if (precision < 0 || precision > 2) return mqct_fail(MQCT_ERR_ARG, "contraction: precision");
  OutMap om = out_map(h, h->n_r, h->n_r, conv);
  double *out = h->d_mat.as<double>();
  int v = h->cvar_fixed[precision];
  if (v < 0) v = h->cvar[precision];
  h->tune_ms = 0.0;
  if (v < 0) {
    const auto t0 = std::chrono::steady_clock::now();
    // time every candidate twice on a sample of the tiles (their outputs are
    // written again by the full run below) and keep the fastest
    const unsigned nsub = (unsigned)(ntiles < 512 ? ntiles : 512);
    float best = 0.f;
    for (int c = 0; c < kNCVar[precision]; c++) {
      float tmin = 0.f;
      for (int rep = 0; rep < 2; rep++) {
        HIP_TRY(hipEventRecord(h->ev[6], h->stream));
        launch_contraction(h, precision, c, nsub, om, out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(h->ev[7], h->stream));
        HIP_TRY(hipEventSynchronize(h->ev[7]));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev[6], h->ev[7]));
        if (rep == 0 || ms < tmin) tmin = ms;
      }
      if (v < 0 || tmin < best) {
        v = c;
        best = tmin;
      }
    }
    h->cvar[precision] = v;
    h->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  launch_contraction(h, precision, v, (unsigned)ntiles, om, out);
  HIP_TRY(hipGetLastError());
  return MQCT_OK;
}

extern "C" int mqct_hip_set_contract_variant(mqct_hip_t *h, int precision, int v) {
  if (!h || precision < 0 || precision > 2 || v < -1 || v >= kNCVar[precision < 0 || precision > 2 ? 0 : precision])
    return mqct_fail(MQCT_ERR_ARG, "set_contract_variant: bad argument");
  h->cvar_fixed[precision] = v;
  if (v < 0) h->cvar[precision] = -1;
  return MQCT_OK;
}

extern "C" int mqct_hip_contract_variant(mqct_hip_t *h, int precision, int *v, int *kc, int *waves,
                                         int *n_candidates) {
  if (!h || precision < 0 || precision > 2 || !v || !kc || !waves || !n_candidates)
    return mqct_fail(MQCT_ERR_ARG, "contract_variant: bad argument");
  const int c = h->cvar_fixed[precision] >= 0 ? h->cvar_fixed[precision] : h->cvar[precision];
  *v = c;
  *kc = c >= 0 ? kCVar[precision][c].kc : 0;
  *waves = c >= 0 ? kCVar[precision][c].nw : 0;
  *n_candidates = kNCVar[precision];
  return MQCT_OK;
}

// FP64 values of every element at the two cutoff R points, for the
// mixed-precision forms: the MIJ_ZERO_CUT decision (matrix_ini.f:1991-1997)
// is then taken from the same FP64 sums the FP64 build forms.  Output
// [T][2] in d_cutv; W64 rows of the two R points in d_Wcut.
int mqct_run_cutoff_fp64(mqct_hip *h, int ir1, int ir2, double conv) {
  // 64 x 64 tiles (b-tile <= a-tile) of every m-block
  std::vector<TileDesc> ct;
  for (size_t blk = 0; blk + 1 < h->block_start.size(); blk++) {
    const int s0 = h->block_start[blk], s1 = h->block_start[blk + 1];
    for (int a0 = s0; a0 < s1; a0 += CT)
      for (int b0 = s0; b0 <= a0; b0 += CT)
        ct.push_back(TileDesc{a0, std::min(CT, s1 - a0), b0, std::min(CT, s1 - b0)});
  }
  const int ntiles = (int)ct.size();
  if (ntiles == 0) return MQCT_OK;
  if (h->d_ctiles.ensure(ct.size() * sizeof(TileDesc))) return mqct_fail(MQCT_ERR_NOMEM, "cutoff tiles");
  HIP_TRY(hipMemcpyAsync(h->d_ctiles.p, ct.data(), ct.size() * sizeof(TileDesc), hipMemcpyHostToDevice, h->stream));
  if (h->d_Wcut.ensure((size_t)2 * h->G * 8) ||
      h->d_cutv.ensure((size_t)h->built_T * 2 * 8))
    return mqct_fail(MQCT_ERR_NOMEM, "cutoff columns");
  const double *V = h->d_V.as<double>();
  const double *om_w = h->d_omega.as<double>();
  mqct_weight_rows(h, V, om_w, ir1, ir2, h->d_Wcut.as<double>());
  OutMap om = out_map(h, 2, 2, conv);
  const bool im = (h->coll_type != 1) || !h->ref_xz;
  const double *pbr = h->ref_xz ? h->d_psi_re.as<double>() : h->d_psib_re.as<double>();
  const double *pbi = h->ref_xz ? h->d_psi_im.as<double>() : h->d_psib_im.as<double>();
  const long long T = h->built_T;
  if (h->d_cutp.ensure((size_t)CS * T * 2 * 8)) return mqct_fail(MQCT_ERR_NOMEM, "cutoff partial sums");
  cutoff_fp64_kernel<16><<<dim3((unsigned)ntiles, CS), 256, 0, h->stream>>>(
      h->d_psi_re.as<double>(), im ? h->d_psi_im.as<double>() : nullptr, pbr, im ? pbi : nullptr,
      h->d_Wcut.as<double>(), h->G, h->d_ctiles.as<TileDesc>(), om, T, h->d_cutp.as<double>());
  HIP_TRY(hipGetLastError());
  cutoff_reduce_kernel<<<(unsigned)((2 * T + 255) / 256), 256, 0, h->stream>>>(h->d_cutp.as<double>(), T, conv,
                                                                                h->d_cutv.as<double>());
  HIP_TRY(hipGetLastError());
  return MQCT_OK;
}

extern "C" int mqct_hip_probe_mfma(int which, const double *A, const double *B,
                                   double *C) {
  if (which < 0 || which > 2) return mqct_fail(MQCT_ERR_ARG, "probe: which");
  int K = (which == 2) ? 32 : 4;
  double *dA, *dB, *dC;
  HIP_TRY(hipMalloc(&dA, 16 * K * 8));
  HIP_TRY(hipMalloc(&dB, 16 * K * 8));
  HIP_TRY(hipMalloc(&dC, 256 * 8));
  HIP_TRY(hipMemcpy(dA, A, 16 * K * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dB, B, 16 * K * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(dC, 0, 256 * 8));
  probe_kernel<<<1, 64>>>(which, dA, dB, dC);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(C, dC, 256 * 8, hipMemcpyDeviceToHost));
  hipFree(dA);
  hipFree(dB);
  hipFree(dC);
  return MQCT_OK;
}
