// Full-data potential of many chains at once (potential.py:219-293 for every chain of
// a table that shares one dataset): U_c = (-sum_i lik(x_i . theta_c, y_i) - log p(theta_c))
// / T, each chain's sum over rows accumulated in FP64.
//
// The row dot products for all chains form the product Z = X Theta^T of the [N, D] rows
// and the [C, D] positions, so they run on the fp32 matrix cores
// (v_mfma_f32_32x32x2_f32) and X is read from HBM once for up to 64 chains instead of
// once per chain. Layout:
// - Theta^T is staged once per block in LDS in B-operand order: for k-block j (eight
//   columns), chain tile t (32 chains) and lane l, the float4
//   Theta[32 t + l % 32][8 j + 4 (l / 32) + 0..3] (128 KiB at D = 512 and 64 chains);
// - each wave takes 64-row tiles (two 32-row groups); lane l streams the float4
//   x[row + l % 32][8 j + 4 (l / 32) + 0..3] straight from HBM into the A operand, four
//   k-blocks ahead of the matrix instructions. Within the eight columns of a k-block
//   the lane halves split the sum as the two K entries of the instruction, so the sum
//   over columns runs in a fixed, permuted order;
// - the 32x32 accumulator of (row group, chain tile) holds z[m][n] at register r of
//   lane l with n = l % 32, m = 8 (r / 4) + 4 (l / 32) + r % 4; the likelihood is
//   applied there and summed per chain in FP64, then across lanes, waves and blocks
//   (blocks in order by the last block to arrive), so the result is deterministic.
#include <math.h>

#include "glm_common.h"
#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {
namespace {

constexpr int kFcThreads = 512;
constexpr int kFcWaves = kFcThreads / kWave;
constexpr int kFcLdsF4 = 8192;     // float4 entries of Theta operands (128 KiB)
constexpr int kFcAhead = 4;        // k-blocks of A loads in flight per row group
constexpr int kFcChainsMax = 64;

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kSpAhead = 2;     // 16-column k-blocks of A loads in flight (split kernel)
constexpr int kSpLdsE = 4096;   // 16-byte Theta operand entries per half (64 KiB each)

struct FcArgs {
  const float* X;
  const float* y;
  const float* thetas;   // chain c at thetas + c * ld_theta
  double* out_U;         // [nchains]
  double* partial;       // [gridDim.x][kFcChainsMax]
  unsigned* counter;
  int64_t ldx;
  int64_t ld_theta;
  int64_t N;
  int64_t tiles;         // ceil(N / 64)
  int D;
  int D4;                // ceil(D / 4): float4 columns of a row that may be read
  int D8r;               // ceil(D / 8) rounded up to kFcAhead
  int nchains;           // chains of this launch (<= 32 NT)
  int D16r;              // split kernel: ceil(D / 16) rounded up to kSpAhead
  int x_exp;             // split kernel: X is scaled by 2^x_exp before splitting
  Lik L;
  Prior pr;
  double temperature;
};

// Branch-free: the load always goes to a valid address (column 0 of a valid row when
// out of range) and the value is masked afterwards, so the loads stay in flight across
// loop iterations instead of being drained at every branch.
__device__ __forceinline__ v4f load_x4(const float* row, int c4, bool ok, int D4) {
  const bool in = ok && c4 < D4;
  const v4f v = *reinterpret_cast<const v4f*>(row + 4 * (in ? c4 : 0));
  const v4f z = {0.f, 0.f, 0.f, 0.f};
  return in ? v : z;
}

// The products of k-blocks jb .. jb + kFcAhead - 1 for both row groups and all chain
// tiles, A operands in registers and B operands from LDS.
template <int NT>
__device__ __forceinline__ void mma_group(v16f (&acc)[2][NT], const v4f (&cur0)[kFcAhead],
                                          const v4f (&cur1)[kFcAhead], const v4f* lds_b,
                                          int jb, int lane) {
  v4f b[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) b[t] = lds_b[(jb * NT + t) * kWave + lane];
#pragma unroll
  for (int u = 0; u < kFcAhead; ++u) {
    // B operands of the next k-block are read while this one's products run.
    v4f bn[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
      bn[t] = u + 1 < kFcAhead ? lds_b[((jb + u + 1) * NT + t) * kWave + lane] : b[t];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(cur0[u][tt], b[t][tt], acc[0][t], 0, 0,
                                                         0);
        acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(cur1[u][tt], b[t][tt], acc[1][t], 0, 0,
                                                         0);
      }
#pragma unroll
    for (int t = 0; t < NT; ++t) b[t] = bn[t];
  }
}

// Per-row likelihood of one family. The logistic softplus uses the hardware base-2
// exponential and logarithm (about 1 ulp each): log(1 + e^-|z|) = ln2 log2(1 + 2^(-|z|
// log2 e)); where 1 + e^-|z| rounds to 1 the dropped term is below 2^-24, so each row
// is within a few float32 ulps of the exact value while costing a tenth of the
// library routines.
template <int FAM>
__device__ __forceinline__ float fam_lik(const Lik& L, float z, float y) {
  if (FAM == SGMC_LIK_LINEAR_GAUSSIAN) {
    const float d = y - z;
    return (L.log_norm + (d * d) * L.inv_s2) / -2.0f;
  }
  constexpr float kLog2e = 1.44269504088896340736f;
  constexpr float kLn2 = 0.69314718055994530942f;
  const float e = __builtin_amdgcn_exp2f(-fabsf(z) * kLog2e);
  const float sp = fmaxf(z, 0.0f) + __builtin_amdgcn_logf(1.0f + e) * kLn2;
  return y * z - sp;
}

template <int NT, int FAM>
__global__ void __launch_bounds__(kFcThreads) glm_full_chains_kernel(FcArgs a) {
  __shared__ v4f lds_b[kFcLdsF4];
  __shared__ double lds_red[kFcWaves][kFcChainsMax];
  __shared__ int flag;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int rsub = lane & 31;
  const int q = lane >> 5;

  // Theta operands: entry e = (j * NT + t) * 64 + l. Every load of the thread is
  // issued before the first LDS store (clamped addresses, masked values).
  const int entries = a.D8r * NT * kWave;
  constexpr int kPer = kFcLdsF4 / kFcThreads;
  v4f stage[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int e = threadIdx.x + i * kFcThreads;
    const int l = e % kWave;
    const int t = (e / kWave) % NT;
    const int j = e / (kWave * NT);
    const int c = 32 * t + (l & 31);
    const int k0 = 8 * j + 4 * (l >> 5);
    const bool live = e < entries && c < a.nchains;
    const float* src = a.thetas + (int64_t)(live ? c : 0) * a.ld_theta;
    float v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool ok = live && k0 + u < a.D;
      const float x = src[ok ? k0 + u : 0];
      v[u] = ok ? x : 0.0f;
    }
    stage[i] = v4f{v[0], v[1], v[2], v[3]};
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int e = threadIdx.x + i * kFcThreads;
    if (e < entries) lds_b[e] = stage[i];
  }
  __syncthreads();

  const int64_t t0 = (int64_t)blockIdx.x * a.tiles / gridDim.x;
  const int64_t t1 = (int64_t)(blockIdx.x + 1) * a.tiles / gridDim.x;
  double sacc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) sacc[t] = 0.0;

  for (int64_t tile = t0 + w; tile < t1; tile += kFcWaves) {
    const int64_t r0 = tile * 64 + rsub;
    const int64_t r1 = r0 + 32;
    const bool ok0 = r0 < a.N, ok1 = r1 < a.N;
    const float* x0 = a.X + (ok0 ? r0 : 0) * a.ldx;
    const float* x1 = a.X + (ok1 ? r1 : 0) * a.ldx;
    v16f acc[2][NT];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[g][t] = v16f{};
    v4f cur0[kFcAhead], cur1[kFcAhead];
#pragma unroll
    for (int u = 0; u < kFcAhead; ++u) {
      cur0[u] = load_x4(x0, 2 * u + q, ok0, a.D4);
      cur1[u] = load_x4(x1, 2 * u + q, ok1, a.D4);
    }
    const int jlast = a.D8r - kFcAhead;
    for (int jb = 0; jb < jlast; jb += kFcAhead) {
      v4f nxt0[kFcAhead], nxt1[kFcAhead];
#pragma unroll
      for (int u = 0; u < kFcAhead; ++u) {
        const int c4 = 2 * (jb + kFcAhead + u) + q;
        nxt0[u] = load_x4(x0, c4, ok0, a.D4);
        nxt1[u] = load_x4(x1, c4, ok1, a.D4);
      }
      mma_group<NT>(acc, cur0, cur1, lds_b, jb, lane);
#pragma unroll
      for (int u = 0; u < kFcAhead; ++u) {
        cur0[u] = nxt0[u];
        cur1[u] = nxt1[u];
      }
    }
    // Labels of the lane's 32 accumulator rows, loaded while the last k-blocks run
    // (clamped index, masked below).
    float yv[2][16];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = tile * 64 + 32 * g + 8 * (r / 4) + 4 * q + (r % 4);
        yv[g][r] = a.y[row < a.N ? row : a.N - 1];
      }
    mma_group<NT>(acc, cur0, cur1, lds_b, jlast, lane);
    // Likelihood of every (row, chain) entry of the tile, summed per chain in FP64.
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = tile * 64 + 32 * g + 8 * (r / 4) + 4 * q + (r % 4);
        const bool ok = row < a.N;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const float lik = fam_lik<FAM>(a.L, acc[g][t][r], yv[g][r]);
          sacc[t] += ok ? (double)lik : 0.0;
        }
      }
  }

  // Lanes l and l ^ 32 hold the same chain; then waves in order; then blocks.
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    sacc[t] += __shfl_xor(sacc[t], 32, kWave);
    if (q == 0) lds_red[w][32 * t + rsub] = sacc[t];
  }
  __syncthreads();
  if (threadIdx.x < 32 * NT) {
    double s = 0.0;
    for (int v = 0; v < kFcWaves; ++v) s += lds_red[v][threadIdx.x];
    publish(a.partial + (int64_t)blockIdx.x * kFcChainsMax + threadIdx.x, s);
  }
  if (!arrive_last(a.counter, gridDim.x, &flag)) return;
  const int c = threadIdx.x;
  if (c >= a.nchains) return;
  double tot = 0.0;
  for (int b = 0; b < (int)gridDim.x; ++b) tot += consume(a.partial + (int64_t)b * kFcChainsMax + c);
  double lp = 0.0;
  if (a.pr.has_prior) {
    const double iv = (double)a.pr.inv_var;
    const double ln = (double)a.pr.log_norm;
    const float* th = a.thetas + (int64_t)c * a.ld_theta;
    for (int k = 0; k < a.D; ++k) {
      const double wk = (double)th[k];
      lp += (ln + wk * wk * iv) / -2.0;
    }
  }
  a.out_U[c] = (-tot - lp) / a.temperature;
}

// Split-fp16 variant. Each scaled value v = 2^e x (|v| < 2^14) becomes hi = f16(v) and
// lo = f16(v - hi), which hold v to 2^-22 relative; a product is taken as
// lo_x hi_t + hi_x lo_t + hi_x hi_t on the fp16 matrix cores
// (v_mfma_f32_32x32x16_f16, fp32 accumulation; every fp16 product is exact in fp32),
// dropping lo_x lo_t (below 2^-20 of the product). The row dot products are then good
// to about 1e-6 relative where float32 products give 6e-8, for a fifth of the
// matrix-core time of v_mfma_f32_32x32x2_f32. X is read as float32 and split in
// registers (splitting the dataset once into stored halves measured slower: the pass is
// bound by memory latency, not by the conversions); X is scaled by the caller's 2^x_exp
// (one exponent for the dataset), Theta per chain by 2^e_c from max |theta_c| over its
// finite entries; z = acc 2^-(x_exp + e_c) is exact. Layout as in the float32 kernel
// with 16-column k-blocks: lane l takes x[row + l % 32][16 j + 8 (l / 32) + 0..7] and
// the Theta entry (j, t, l) holds Theta[32 t + l % 32][16 j + 8 (l / 32) + 0..7] as hi
// and lo halves.
__device__ __forceinline__ void split8(v4f a, v4f b, float sc, f16x8& hi, f16x8& lo) {
  const float v[8] = {a.x * sc, a.y * sc, a.z * sc, a.w * sc,
                      b.x * sc, b.y * sc, b.z * sc, b.w * sc};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const _Float16 h = (_Float16)v[i];
    hi[i] = h;
    lo[i] = (_Float16)(v[i] - (float)h);
  }
}

template <int NT>
__device__ __forceinline__ void mma_split(v16f (&acc)[2][NT], const v4f (&cur0)[kSpAhead][2],
                                          const v4f (&cur1)[kSpAhead][2], const f16x8* lds_h,
                                          const f16x8* lds_l, int jb, int lane, float xs) {
#pragma unroll
  for (int u = 0; u < kSpAhead; ++u) {
    f16x8 ah0, al0, ah1, al1;
    split8(cur0[u][0], cur0[u][1], xs, ah0, al0);
    split8(cur1[u][0], cur1[u][1], xs, ah1, al1);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int e = ((jb + u) * NT + t) * kWave + lane;
      const f16x8 bh = lds_h[e];
      const f16x8 bl = lds_l[e];
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al0, bh, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al1, bh, acc[1][t], 0, 0, 0);
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bl, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bl, acc[1][t], 0, 0, 0);
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bh, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bh, acc[1][t], 0, 0, 0);
    }
  }
}

template <int NT, int FAM>
__global__ void __launch_bounds__(kFcThreads) glm_full_chains_split_kernel(FcArgs a) {
  __shared__ f16x8 lds_h[kSpLdsE];
  __shared__ f16x8 lds_l[kSpLdsE];
  __shared__ double lds_red[kFcWaves][kFcChainsMax];
  __shared__ float lds_sc[kFcChainsMax];
  __shared__ float lds_inv[kFcChainsMax];
  __shared__ int flag;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int rsub = lane & 31;
  const int q = lane >> 5;

  // Per-chain exponent: eight threads per chain find max |theta_c| over finite entries.
  {
    const int c = threadIdx.x >> 3, sub = threadIdx.x & 7;
    float m = 0.0f;
    if (c < a.nchains) {
      const float* th = a.thetas + (int64_t)c * a.ld_theta;
      for (int k = sub; k < a.D; k += 8) {
        const float v = fabsf(th[k]);
        if (v <= 3.0e38f) m = fmaxf(m, v);
      }
    }
    m = fmaxf(m, __shfl_xor(m, 1, kWave));
    m = fmaxf(m, __shfl_xor(m, 2, kWave));
    m = fmaxf(m, __shfl_xor(m, 4, kWave));
    if (sub == 0) {
      int E = 0;
      if (m > 0.0f) frexpf(m, &E);  // m < 2^E
      const int e = min(64, max(-64, 14 - E));
      lds_sc[c] = ldexpf(1.0f, e);
      lds_inv[c] = ldexpf(1.0f, -(e + a.x_exp));
    }
  }
  __syncthreads();

  // Theta operands, every load of the thread issued before the first LDS store.
  const int entries = a.D16r * NT * kWave;
  constexpr int kPer = kSpLdsE / kFcThreads;
  f16x8 sh[kPer], sl[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int e = threadIdx.x + i * kFcThreads;
    const int l = e % kWave;
    const int t = (e / kWave) % NT;
    const int j = e / (kWave * NT);
    const int c = 32 * t + (l & 31);
    const int k0 = 16 * j + 8 * (l >> 5);
    const bool live = e < entries && c < a.nchains;
    const float* src = a.thetas + (int64_t)(live ? c : 0) * a.ld_theta;
    const float sc = lds_sc[c];
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool ok = live && k0 + u < a.D;
      const float x = src[ok ? k0 + u : 0];
      v[u] = ok ? x * sc : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const _Float16 h = (_Float16)v[u];
      sh[i][u] = h;
      sl[i][u] = (_Float16)(v[u] - (float)h);
    }
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int e = threadIdx.x + i * kFcThreads;
    if (e < entries) {
      lds_h[e] = sh[i];
      lds_l[e] = sl[i];
    }
  }
  __syncthreads();

  const float xs = ldexpf(1.0f, a.x_exp);
  float inv[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) inv[t] = lds_inv[32 * t + rsub];
  const int64_t t0 = (int64_t)blockIdx.x * a.tiles / gridDim.x;
  const int64_t t1 = (int64_t)(blockIdx.x + 1) * a.tiles / gridDim.x;
  double sacc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) sacc[t] = 0.0;

  for (int64_t tile = t0 + w; tile < t1; tile += kFcWaves) {
    const int64_t r0 = tile * 64 + rsub;
    const int64_t r1 = r0 + 32;
    const bool ok0 = r0 < a.N, ok1 = r1 < a.N;
    const float* x0 = a.X + (ok0 ? r0 : 0) * a.ldx;
    const float* x1 = a.X + (ok1 ? r1 : 0) * a.ldx;
    v16f acc[2][NT];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[g][t] = v16f{};
    v4f cur0[kSpAhead][2], cur1[kSpAhead][2];
#pragma unroll
    for (int u = 0; u < kSpAhead; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        cur0[u][h] = load_x4(x0, 4 * u + 2 * q + h, ok0, a.D4);
        cur1[u][h] = load_x4(x1, 4 * u + 2 * q + h, ok1, a.D4);
      }
    const int jlast = a.D16r - kSpAhead;
    for (int jb = 0; jb < jlast; jb += kSpAhead) {
      v4f nxt0[kSpAhead][2], nxt1[kSpAhead][2];
#pragma unroll
      for (int u = 0; u < kSpAhead; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int c4 = 4 * (jb + kSpAhead + u) + 2 * q + h;
          nxt0[u][h] = load_x4(x0, c4, ok0, a.D4);
          nxt1[u][h] = load_x4(x1, c4, ok1, a.D4);
        }
      mma_split<NT>(acc, cur0, cur1, lds_h, lds_l, jb, lane, xs);
#pragma unroll
      for (int u = 0; u < kSpAhead; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          cur0[u][h] = nxt0[u][h];
          cur1[u][h] = nxt1[u][h];
        }
    }
    float yv[2][16];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = tile * 64 + 32 * g + 8 * (r / 4) + 4 * q + (r % 4);
        yv[g][r] = a.y[row < a.N ? row : a.N - 1];
      }
    mma_split<NT>(acc, cur0, cur1, lds_h, lds_l, jlast, lane, xs);
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = tile * 64 + 32 * g + 8 * (r / 4) + 4 * q + (r % 4);
        const bool ok = row < a.N;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const float lik = fam_lik<FAM>(a.L, acc[g][t][r] * inv[t], yv[g][r]);
          sacc[t] += ok ? (double)lik : 0.0;
        }
      }
  }

#pragma unroll
  for (int t = 0; t < NT; ++t) {
    sacc[t] += __shfl_xor(sacc[t], 32, kWave);
    if (q == 0) lds_red[w][32 * t + rsub] = sacc[t];
  }
  __syncthreads();
  if (threadIdx.x < 32 * NT) {
    double sum = 0.0;
    for (int v = 0; v < kFcWaves; ++v) sum += lds_red[v][threadIdx.x];
    publish(a.partial + (int64_t)blockIdx.x * kFcChainsMax + threadIdx.x, sum);
  }
  if (!arrive_last(a.counter, gridDim.x, &flag)) return;
  const int c = threadIdx.x;
  if (c >= a.nchains) return;
  double tot = 0.0;
  for (int b = 0; b < (int)gridDim.x; ++b) tot += consume(a.partial + (int64_t)b * kFcChainsMax + c);
  double lp = 0.0;
  if (a.pr.has_prior) {
    const double iv = (double)a.pr.inv_var;
    const double ln = (double)a.pr.log_norm;
    const float* th = a.thetas + (int64_t)c * a.ld_theta;
    for (int k = 0; k < a.D; ++k) {
      const double wk = (double)th[k];
      lp += (ln + wk * wk * iv) / -2.0;
    }
  }
  a.out_U[c] = (-tot - lp) / a.temperature;
}

}  // namespace
}  // namespace sgmc

using namespace sgmc;

namespace {

int launch_full_chains(const char* name, bool split, int x_exp, int family, const float* X,
                       int64_t ldx, const float* y, int64_t N, int D, const float* thetas,
                       int64_t ld_theta, int nchains, float sigma, int has_prior,
                       float prior_scale, double temperature, double* out_U, double* workspace,
                       int64_t workspace_doubles, int nblocks, void* stream) {
  SGMC_REQUIRE(X && y && thetas && out_U && workspace, SGMC_E_ARG, "%s: null pointer", name);
  SGMC_REQUIRE(family == SGMC_LIK_LINEAR_GAUSSIAN || family == SGMC_LIK_LOGISTIC, SGMC_E_ARG,
               "%s: unknown family %d", name, family);
  SGMC_REQUIRE(D >= 1 && D <= 1024, SGMC_E_ARG, "%s: D=%d outside [1, 1024]", name, D);
  SGMC_REQUIRE(ldx % 4 == 0 && ldx >= 4 * ((D + 3) / 4), SGMC_E_ARG,
               "%s: ldx must be a multiple of 4 and >= D rounded up", name);
  SGMC_REQUIRE(((uintptr_t)X & 15) == 0, SGMC_E_ARG, "%s: X not 16-byte aligned", name);
  SGMC_REQUIRE(ld_theta >= D, SGMC_E_ARG, "%s: ld_theta < D", name);
  SGMC_REQUIRE(N >= 1 && nchains >= 1 && nblocks >= 1, SGMC_E_ARG, "%s: bad size", name);
  SGMC_REQUIRE(workspace_doubles >= sgmc_glm_full_chains_workspace_doubles(nblocks), SGMC_E_ARG,
               "%s: workspace too small", name);
  SGMC_REQUIRE(temperature != 0.0, SGMC_E_ARG, "%s: temperature is zero", name);
  SGMC_REQUIRE(x_exp >= -64 && x_exp <= 64, SGMC_E_ARG, "%s: x_exp=%d outside [-64, 64]", name,
               x_exp);
  FcArgs a;
  a.X = X;
  a.y = y;
  a.ldx = ldx;
  a.ld_theta = ld_theta;
  a.N = N;
  a.tiles = (N + 63) / 64;
  a.D = D;
  a.D4 = (D + 3) / 4;
  const int D8 = (D + 7) / 8;
  a.D8r = (D8 + kFcAhead - 1) / kFcAhead * kFcAhead;
  const int D16 = (D + 15) / 16;
  a.D16r = (D16 + kSpAhead - 1) / kSpAhead * kSpAhead;
  a.x_exp = x_exp;
  a.L = make_lik(family, sigma);
  a.pr = make_prior(has_prior, prior_scale);
  a.temperature = temperature;
  a.partial = workspace;
  a.counter = reinterpret_cast<unsigned*>(workspace + (int64_t)nblocks * kFcChainsMax);
  // Two chain tiles per pass when their operands fit the LDS, else one.
  const int nt = split ? ((a.D16r * 2 * kWave <= kSpLdsE) ? 2 : 1)
                       : ((a.D8r * 2 * kWave <= kFcLdsF4) ? 2 : 1);
  hipStream_t s = as_stream(stream);
  const dim3 grid(nblocks), block(kFcThreads);
  for (int c0 = 0; c0 < nchains; c0 += 32 * nt) {
    a.thetas = thetas + (int64_t)c0 * ld_theta;
    a.out_U = out_U + c0;
    a.nchains = nchains - c0 < 32 * nt ? nchains - c0 : 32 * nt;
    const bool two = a.nchains > 32;
    const bool lg = family == SGMC_LIK_LOGISTIC;
    if (split) {
      if (lg) {
        if (two) hipLaunchKernelGGL((glm_full_chains_split_kernel<2, SGMC_LIK_LOGISTIC>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((glm_full_chains_split_kernel<1, SGMC_LIK_LOGISTIC>), grid, block, 0, s, a);
      } else {
        if (two) hipLaunchKernelGGL((glm_full_chains_split_kernel<2, SGMC_LIK_LINEAR_GAUSSIAN>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((glm_full_chains_split_kernel<1, SGMC_LIK_LINEAR_GAUSSIAN>), grid, block, 0, s, a);
      }
    } else {
      if (lg) {
        if (two) hipLaunchKernelGGL((glm_full_chains_kernel<2, SGMC_LIK_LOGISTIC>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((glm_full_chains_kernel<1, SGMC_LIK_LOGISTIC>), grid, block, 0, s, a);
      } else {
        if (two) hipLaunchKernelGGL((glm_full_chains_kernel<2, SGMC_LIK_LINEAR_GAUSSIAN>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((glm_full_chains_kernel<1, SGMC_LIK_LINEAR_GAUSSIAN>), grid, block, 0, s, a);
      }
    }
    const int rc = check_launch(name);
    if (rc != SGMC_OK) return rc;
  }
  return SGMC_OK;
}

}  // namespace

extern "C" {

int64_t sgmc_glm_full_chains_workspace_doubles(int nblocks) {
  return (int64_t)nblocks * kFcChainsMax + 2;
}

int sgmc_glm_full_potential_chains(int family, const float* X, int64_t ldx, const float* y,
                                   int64_t N, int D, const float* thetas, int64_t ld_theta,
                                   int nchains, float sigma, int has_prior, float prior_scale,
                                   double temperature, double* out_U, double* workspace,
                                   int64_t workspace_doubles, int nblocks, void* stream) {
  return launch_full_chains("sgmc_glm_full_potential_chains", false, 0, family, X, ldx, y, N, D,
                            thetas, ld_theta, nchains, sigma, has_prior, prior_scale,
                            temperature, out_U, workspace, workspace_doubles, nblocks, stream);
}

int sgmc_glm_full_potential_chains_split(int family, const float* X, int64_t ldx,
                                         const float* y, int64_t N, int D, const float* thetas,
                                         int64_t ld_theta, int nchains, float sigma,
                                         int has_prior, float prior_scale, double temperature,
                                         int x_exp, double* out_U, double* workspace,
                                         int64_t workspace_doubles, int nblocks, void* stream) {
  return launch_full_chains("sgmc_glm_full_potential_chains_split", true, x_exp, family, X, ldx,
                            y, N, D, thetas, ld_theta, nchains, sigma, has_prior, prior_scale,
                            temperature, out_U, workspace, workspace_doubles, nblocks, stream);
}

}  // extern "C"
