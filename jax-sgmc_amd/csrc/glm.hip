// Generalized-linear likelihood families on the device: stochastic potential and
// gradient (potential.py:159-214 under jax.value_and_grad, integrator.py:166, 593,
// 792), the full-data potential in FP64 (potential.py:219-293), and the fused
// multi-chain step that advances every chain of a chain table by one gradient
// evaluation per launch.
//
// Families, per observation (x_i, y_i), z_i = x_i . w:
//   linear-Gaussian  lik = (log(2 pi s^2) + (y - z)^2 / s^2) / (-2)   dlik/dz = (y - z) / s^2
//   logistic         lik = y z - log(1 + e^z)                          dlik/dz = y - sigmoid(z)
// Potential: U = (-N * mean(lik) - log_prior) / T;
//            grad U = ((-N / n) * sum_i dlik_i/dz x_i - grad log_prior) / T
// with the Gaussian prior log p = sum_j (log(2 pi s_p^2) + w_j^2 / s_p^2) / (-2).
//
// Row layout in HBM: X is row-major [N, ldx] float32 with ldx a multiple of 4 and
// 16-byte aligned rows; a row is read as float4 chunks, chunk c by lane c % LPR of a
// segment of LPR lanes, so a wavefront covers 64/LPR rows per load instruction and
// every load is a contiguous 16*LPR-byte run. Rows are reduced with xor-shuffles
// inside the segment; the per-lane gradient accumulators are reduced across segments,
// then across the block's wavefronts through LDS, then across blocks by the last
// block to arrive (deterministic order).
#include <math.h>

#include <mutex>

#include "glm_common.h"
#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {

constexpr int kNT = 256;
constexpr int kWaves = kNT / kWave;
constexpr int kMaxD = 1024;  // fused path: features per row
constexpr int kLdsRows = 4096;  // per-row likelihoods staged in LDS per block
constexpr int kResWindow = 512;  // resident loop: step sizes / save slots staged per window

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float dot4(float4 a, float4 b) {
  return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}

__device__ __forceinline__ float dot4(f4 a, float4 b) {
  return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}

__device__ __forceinline__ void axpy4(float4& acc, float r, float4 x) {
  acc.x += r * x.x;
  acc.y += r * x.y;
  acc.z += r * x.z;
  acc.w += r * x.w;
}

__device__ __forceinline__ void axpy4(float4& acc, float r, f4 x) {
  acc.x += r * x.x;
  acc.y += r * x.y;
  acc.z += r * x.z;
  acc.w += r * x.w;
}

// Source of the row index of minibatch row j.
struct RowSource {
  const int32_t* idx;  // explicit indices, or null
  int draw;            // nonzero: in-kernel randint with (k1, k2)
  Key k1, k2;
  uint32_t n, span, mult, magic;
  int64_t rows;        // rows addressable through X (bounds guard)
  // Row index of minibatch row j, or -1 if it falls outside [0, rows).
  __device__ __forceinline__ int64_t row(int j) const {
    int64_t r;
    if (draw) r = (int64_t)draw_index(k1, k2, n, (uint32_t)j, span, mult, magic);
    else if (idx) r = (int64_t)idx[j];
    else r = (int64_t)j;
    return (r >= 0 && r < rows) ? r : -1;
  }
};

// Sum over the LPR lanes of a segment (segments are aligned runs of LPR lanes); every
// lane of the segment gets the sum. Steps inside a 16-lane row use DPP (quad
// permutes, half-row and row mirrors), the xor-16 step a ds_swizzle, the last step
// two readlanes.
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF,
                                                        false));
}

template <int LPR>
__device__ __forceinline__ float seg_sum(float v) {
  if (LPR >= 2) v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
  if (LPR >= 4) v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
  if (LPR >= 8) v = dpp_add<0x141>(v);  // row_half_mirror
  if (LPR >= 16) v = dpp_add<0x140>(v); // row_mirror
  if (LPR >= 32) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
  if (LPR >= 64)
    v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)) +
        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  return v;
}

// Rows in flight per segment per pipeline stage: 4, also for D > 512 (eight 4 KB rows in
// flight per wavefront there; 2 rows per stage left the row loop latency-bound); a stage
// of RPW-row loads never exceeds one 64-row chunk.
template <int LPR, int CPL>
struct Unroll {
  static constexpr int U0 = 4;
  static constexpr int U = U0 < LPR ? U0 : LPR;
};

// Row ids of one 64-row chunk of a wavefront's range: lane l holds the id of row cb + l
// (-1 when outside [0, rows) or past the range end). Labels are loaded with the rows'
// features, so nothing here has to be waited on before the next loads are issued.
__device__ __forceinline__ int chunk_rows(const RowSource& src, int cb, int r1) {
  const int lane = threadIdx.x & (kWave - 1);
  int rid = -1;
  if (cb + lane < r1) {
    const int64_t r = src.row(cb + lane);
    if (r >= 0) rid = (int)r;
  }
  return rid;
}

template <int LPR, int CPL>
struct RowGroup {
  static constexpr int U = Unroll<LPR, CPL>::U;
  f4 x[U][CPL];
  float y[U];
  bool ok[U];
};

// Buffer resource word 3 for raw (unformatted) buffer loads on CDNA.
constexpr int kRawBufferFlags = 0x00020000;

// Issues the loads of the U row groups starting at group-local offset gl of a chunk.
// Rows of a whole wavefront (LPR = 64) are read through a buffer resource whose base is
// the row and whose size is the row's 4 D4 floats: the base lives in scalar registers,
// the chunks are one lane offset plus immediate offsets, and chunks past the row (or
// every chunk of a row outside the data, size 0) read as zero without a clamp.
template <int LPR, int CPL>
__device__ __forceinline__ void load_group(gcfloat* __restrict__ X, int64_t ldx,
                                           gcfloat* __restrict__ y, int D4, int gl, int rid,
                                           RowGroup<LPR, CPL>& g) {
  constexpr int RPW = kWave / LPR;
  constexpr int U = RowGroup<LPR, CPL>::U;
  const int lane = threadIdx.x & (kWave - 1);
  const int seg = lane / LPR;
  const int q = lane % LPR;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int sl = gl + u * RPW + seg;  // < 64
    int r;
    if (RPW == 1) r = __builtin_amdgcn_readlane(rid, sl);
    else r = __shfl(rid, sl, kWave);
    g.ok[u] = r >= 0;
    const int64_t rc = r >= 0 ? r : 0;
    g.y[u] = y[rc];
    if constexpr (RPW == 1) {
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(const float*)(X + rc * ldx), (short)0, r >= 0 ? 16 * D4 : 0, kRawBufferFlags);
#pragma unroll
      for (int k = 0; k < CPL; ++k)
        g.x[u][k] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(
                                               rs, 16 * (q + LPR * k), 0, 0));
    } else {
      SGMC_GLOBAL const f4* rp = reinterpret_cast<SGMC_GLOBAL const f4*>(X + rc * ldx);
      // every lane loads (clamped to a valid chunk); invalid chunks are zeroed at use
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const int c4 = q + LPR * k;
        g.x[u][k] = rp[c4 < D4 ? c4 : D4 - 1];
      }
    }

  }
}

// Dot product and accumulation of the row loop with fused multiply-adds.
__device__ __forceinline__ float dot4_fma(f4 a, float4 b, float t) {
  t = fmaf(a.x, b.x, t);
  t = fmaf(a.y, b.y, t);
  t = fmaf(a.z, b.z, t);
  return fmaf(a.w, b.w, t);
}

__device__ __forceinline__ void axpy4_fma(float4& acc, float r, f4 x) {
  acc.x = fmaf(r, x.x, acc.x);
  acc.y = fmaf(r, x.y, acc.y);
  acc.z = fmaf(r, x.z, acc.z);
  acc.w = fmaf(r, x.w, acc.w);
}

// One row group: z = x . w per row (segment sums), the likelihood and dlik/dz of row u
// evaluated once, by lane u of the row's segment, and broadcast; then acc += r x.
template <int LPR, int CPL>
__device__ __forceinline__ void process_group(RowGroup<LPR, CPL>& g, int row0, int D4,
                                              const float4 (&th)[CPL], const Lik& L,
                                              float4 (&acc)[CPL], float& lik_acc,
                                              float* lds_lik, int lds_row0) {
  constexpr int RPW = kWave / LPR;
  constexpr int U = RowGroup<LPR, CPL>::U;
  static_assert(U <= LPR, "one lane per row of the group");
  const int lane = threadIdx.x & (kWave - 1);
  const int seg = lane / LPR;
  const int q = lane % LPR;
  // chunks past D (clamped loads) and rows outside the data contribute zero; the
  // masking is skipped when the whole wavefront's group is valid, and always for rows
  // read through a buffer resource (those chunks read as zero)
  bool bad = RPW > 1 && D4 < LPR * CPL;
#pragma unroll
  for (int u = 0; u < U; ++u) bad = bad || (RPW > 1 && !g.ok[u]);
  if (__any(bad)) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const bool valid = g.ok[u] && (q + LPR * k) < D4;
        const f4 zero = {0.f, 0.f, 0.f, 0.f};
        g.x[u][k] = valid ? g.x[u][k] : zero;
      }
    }
  }
  float z[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    float t = 0.0f;
#pragma unroll
    for (int k = 0; k < CPL; ++k) t = dot4_fma(g.x[u][k], th[k], t);
    z[u] = seg_sum<LPR>(t);
  }
  // lane q < U of each segment takes row q (bitwise blends: a select chain on q would be
  // turned into an indexed table in scratch memory)
  uint32_t zb = __float_as_uint(z[0]), yb = __float_as_uint(g.y[0]);
  uint32_t okb = g.ok[0] ? 1u : 0u;
#pragma unroll
  for (int u = 1; u < U; ++u) {
    const uint32_t m = 0u - (uint32_t)(q == u);
    zb = (zb & ~m) | (__float_as_uint(z[u]) & m);
    yb = (yb & ~m) | (__float_as_uint(g.y[u]) & m);
    okb = (okb & ~m) | ((g.ok[u] ? 1u : 0u) & m);
  }
  const float zq = __uint_as_float(zb), yq = __uint_as_float(yb);
  const bool okq = okb != 0u;
  float likq, rq;
  row_lik(L, zq, yq, likq, rq);
  if (!okq) {
    likq = 0.0f;
    rq = 0.0f;
  }
  if (q < U && okq) {
    lik_acc += likq;
    lds_lik[row0 + q * RPW + seg - lds_row0] = likq;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    float r;
    if (RPW == 1) r = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rq), u));
    else r = __shfl(rq, seg * LPR + u, kWave);
#pragma unroll
    for (int k = 0; k < CPL; ++k) axpy4_fma(acc[k], r, g.x[u][k]);
  }
}

// Accumulates sum_j r_j x_j (this lane's chunks) and sum_j lik_j over minibatch rows
// [r0, r1). Loads run one row group ahead of the arithmetic (two register buffers), and
// the row ids of the next 64-row chunk are drawn before the current chunk is done.
template <int LPR, int CPL>
__device__ __forceinline__ void accumulate_rows(gcfloat* __restrict__ X, int64_t ldx,
                                                gcfloat* __restrict__ y, int D4,
                                                const RowSource& src, int r0, int r1,
                                                const float4 (&th)[CPL], const Lik& L,
                                                float4 (&acc)[CPL], float& lik_acc,
                                                float* lds_lik, int lds_row0) {
  constexpr int RPW = kWave / LPR;
  constexpr int G = RPW * RowGroup<LPR, CPL>::U;  // rows per group; divides 64
  static_assert(kWave % G == 0, "row groups must tile a 64-row chunk");
  if (r0 >= r1) return;
  int cb = r0;
  int rid = chunk_rows(src, cb, r1);
  int rid_n = chunk_rows(src, cb + kWave, r1);
  RowGroup<LPR, CPL> ga, gb;
  load_group<LPR, CPL>(X, ldx, y, D4, 0, rid, ga);
  int gl = 0;  // group offset inside the current chunk
  // The prefetch is issued on every path (rows past the range carry id -1 and load the
  // clamped row 0), so the number of loads in flight is the same on every path and the
  // wait before the arithmetic only covers the group being processed.
  for (;;) {
    int nl = gl + G;
    bool more = cb + nl < r1;
    bool next = nl >= kWave;
    load_group<LPR, CPL>(X, ldx, y, D4, next ? 0 : nl, next ? rid_n : rid, gb);
    process_group<LPR, CPL>(ga, cb + gl, D4, th, L, acc, lik_acc, lds_lik, lds_row0);
    if (!more) break;
    if (next) {
      cb += kWave;
      nl = 0;
      rid = rid_n;
      rid_n = chunk_rows(src, cb + kWave, r1);
    }
    gl = nl;
    nl = gl + G;
    more = cb + nl < r1;
    next = nl >= kWave;
    load_group<LPR, CPL>(X, ldx, y, D4, next ? 0 : nl, next ? rid_n : rid, ga);
    process_group<LPR, CPL>(gb, cb + gl, D4, th, L, acc, lik_acc, lds_lik, lds_row0);
    if (!more) break;
    if (next) {
      cb += kWave;
      nl = 0;
      rid = rid_n;
      rid_n = chunk_rows(src, cb + kWave, r1);
    }
    gl = nl;
  }
}

// Reduces acc over the segments of each wavefront and over the block's wavefronts;
// hands column sum c (0 <= c < 4 D4) to put(c, s) and returns the block's lik sum.
template <int LPR, int CPL, bool LDS_SYNC = false, class Put>
__device__ __forceinline__ float block_reduce(float4 (&acc)[CPL], float lik_acc, int D4,
                                              float* lds_cols, float* lds_small, Put put) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int q = lane % LPR;
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
#pragma unroll
    for (int m = LPR; m < kWave; m <<= 1) {
      acc[k].x += __shfl_xor(acc[k].x, m, kWave);
      acc[k].y += __shfl_xor(acc[k].y, m, kWave);
      acc[k].z += __shfl_xor(acc[k].z, m, kWave);
      acc[k].w += __shfl_xor(acc[k].w, m, kWave);
    }
  }
  if (LDS_SYNC) lds_sync();
  else __syncthreads();
  if (lane < LPR) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c4 = q + LPR * k;
      if (c4 < D4) reinterpret_cast<float4*>(lds_cols + w * 4 * D4)[c4] = acc[k];
    }
  }
  const float lsum = block_sum<kNT, LDS_SYNC>(lik_acc, lds_small);  // contains barriers
  for (int c = threadIdx.x; c < 4 * D4; c += kNT) {
    float s = 0.0f;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) s += lds_cols[ww * 4 * D4 + c];
    put(c, s);
  }
  return lsum;
}

// block_reduce with the column sums published to out[0..4*D4) for the chain's last block.
template <int LPR, int CPL>
__device__ __forceinline__ float block_reduce_write(float4 (&acc)[CPL], float lik_acc, int D4,
                                                    gfloat* __restrict__ out, float* lds_cols,
                                                    float* lds_small) {
  return block_reduce<LPR, CPL>(acc, lik_acc, D4, lds_cols, lds_small,
                                [&](int c, float s) { publish(out + c, s); });
}

// Splits rows [0, n) of a minibatch among the blocks of a chain and then among the
// wavefronts of a block (contiguous ranges).
__device__ __forceinline__ void wave_rows(int n, int nblocks, int b, int& r0, int& r1,
                                          int& b0, int& b1) {
  const int per_block = (n + nblocks - 1) / nblocks;
  b0 = min(n, b * per_block);
  b1 = min(n, b0 + per_block);
  const int per_wave = (b1 - b0 + kWaves - 1) / kWaves;
  const int w = threadIdx.x / kWave;
  r0 = min(b1, b0 + w * per_wave);
  r1 = min(b1, r0 + per_wave);
}

template <int LPR, int CPL>
__device__ __forceinline__ void load_theta(const float* __restrict__ w, int D, float4 (&th)[CPL]) {
  const int q = (threadIdx.x & (kWave - 1)) % LPR;
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = 4 * (q + LPR * k);
    th[k].x = c + 0 < D ? w[c + 0] : 0.f;
    th[k].y = c + 1 < D ? w[c + 1] : 0.f;
    th[k].z = c + 2 < D ? w[c + 2] : 0.f;
    th[k].w = c + 3 < D ? w[c + 3] : 0.f;
  }
}

// ---------------------------------------------------------------- plain path --

struct PotArgs {
  const float* X;
  const float* y;
  const int32_t* idx;
  const float* theta;
  float* out_U;
  float* out_grad;
  float* out_lik;
  float* out_var;
  float* partial;      // [nblocks][4*D4 + 2]
  unsigned* counter;
  int64_t ldx;
  int64_t N;
  int n, D, D4;
  Lik L;
  Prior pr;
  float temperature;
};

template <int LPR, int CPL>
__global__ void __launch_bounds__(kNT) glm_potential_kernel(PotArgs a) {
  __shared__ __attribute__((aligned(16))) float lds_cols[kWaves * kMaxD];
  __shared__ float lds_small[kWaves];
  __shared__ float lds_lik[kLdsRows];
  __shared__ int flag;
  float4 th[CPL];
  load_theta<LPR, CPL>(a.theta, a.D, th);
  float4 acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  float lik_acc = 0.0f;
  int r0, r1, b0, b1;
  wave_rows(a.n, gridDim.x, blockIdx.x, r0, r1, b0, b1);
  RowSource src{a.idx, 0, Key{0, 0}, Key{0, 0}, 0, 0, 0, 0, a.idx ? a.N : (int64_t)a.n};
  accumulate_rows<LPR, CPL>(to_global(a.X), a.ldx, to_global(a.y), a.D4, src, r0, r1, th, a.L,
                            acc, lik_acc, lds_lik, b0);
  // Block record: column sums, the block's lik sum and its lik sum of squared deviations
  // about the block mean; per-row likelihoods go straight to out_lik.
  const int stride = 4 * a.D4 + 2;
  gfloat* mine = to_global(a.partial) + (int64_t)blockIdx.x * stride;
  const float ls = block_reduce_write<LPR, CPL>(acc, lik_acc, a.D4, mine, lds_cols, lds_small);
  const int cnt = b1 - b0;
  float m2 = 0.0f;
  if (a.out_var) {
    const float mb = cnt > 0 ? ls / (float)cnt : 0.0f;
    float s2 = 0.0f;
    for (int j = threadIdx.x; j < cnt; j += kNT) {
      const float d = lds_lik[j] - mb;
      s2 += d * d;
    }
    m2 = block_sum<kNT>(s2, lds_small);
  }
  if (threadIdx.x == 0) {
    publish(mine + 4 * a.D4, ls);
    publish(mine + 4 * a.D4 + 1, m2);
  }
  if (a.out_lik) {
    gfloat* lo = to_global(a.out_lik);
    for (int j = threadIdx.x; j < cnt; j += kNT) lo[b0 + j] = lds_lik[j];
  }
  if (!arrive_last(a.counter, gridDim.x, &flag)) return;

  // Last block: combine the block records, add the prior, scale.
  gcfloat* part = to_global(a.partial);
  const int nb = gridDim.x;
  const float c = (float)(-a.N) / (float)a.n;
  float sl = 0.0f;
  for (int b = threadIdx.x; b < nb; b += kNT) sl += consume(part + (int64_t)b * stride + 4 * a.D4);
  const float sumlik = block_sum<kNT>(sl, lds_small);
  const float mean = sumlik / (float)a.n;
  float lp = 0.0f;
  for (int j = threadIdx.x; j < a.D; j += kNT) {
    float s, unused;
    sum_strided2(part + j, part + j, stride, nb, s, unused);
    const float wj = a.theta[j];
    float g = c * s;
    if (a.pr.has_prior) {
      g = g + wj * a.pr.inv_var;
      lp += (a.pr.log_norm + (wj * wj) * a.pr.inv_var) / -2.0f;
    }
    a.out_grad[j] = g / a.temperature;
  }
  lp = block_sum<kNT>(lp, lds_small);
  if (threadIdx.x == 0) a.out_U[0] = ((float)(-a.N) * mean - lp) / a.temperature;
  if (a.out_var) {
    // var(lik) = (sum_b M2_b + cnt_b (mean_b - mean)^2) / n  (pairwise/Chan combination)
    const int per_block = (a.n + nb - 1) / nb;
    float t = 0.0f;
    for (int b = threadIdx.x; b < nb; b += kNT) {
      const int lo = min(a.n, b * per_block);
      const int k = min(a.n, lo + per_block) - lo;
      if (k > 0) {
        const float d = consume(part + (int64_t)b * stride + 4 * a.D4) / (float)k - mean;
        t += consume(part + (int64_t)b * stride + 4 * a.D4 + 1) + (float)k * (d * d);
      }
    }
    t = block_sum<kNT>(t, lds_small);
    if (threadIdx.x == 0) a.out_var[0] = t / (float)a.n;
  }
}

// --------------------------------------------------------- full potential ----

struct FullArgs {
  const float* X;
  const float* y;
  const float* theta;
  double* out_U;
  double* partial;
  unsigned* counter;
  int64_t ldx;
  int64_t N;
  int D, D4;
  int rows_per_block;
  Lik L;
  Prior pr;
  double temperature;
};

template <int LPR, int CPL>
__global__ void __launch_bounds__(kNT) glm_full_kernel(FullArgs a) {
  __shared__ double lds_d[kWaves];
  __shared__ int flag;
  float4 th[CPL];
  load_theta<LPR, CPL>(a.theta, a.D, th);
  constexpr int RPW = kWave / LPR;
  const int lane = threadIdx.x & (kWave - 1);
  const int seg = lane / LPR;
  const int q = lane % LPR;
  const int w = threadIdx.x / kWave;
  const int64_t b0 = (int64_t)blockIdx.x * a.rows_per_block;
  const int64_t b1 = min<int64_t>(a.N, b0 + a.rows_per_block);
  double acc = 0.0;
  for (int64_t r = b0 + w * RPW + seg; r < b1; r += kWaves * RPW) {
    const float4* rp = reinterpret_cast<const float4*>(a.X + r * a.ldx);
    float z = 0.0f;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c4 = q + LPR * k;
      if (c4 < a.D4) z += dot4(rp[c4], th[k]);
    }
#pragma unroll
    for (int m = 1; m < LPR; m <<= 1) z += __shfl_xor(z, m, kWave);
    float lik, rr;
    row_lik(a.L, z, a.y[r], lik, rr);
    if (q == 0) acc += (double)lik;
  }
  const double bs = block_sum_d<kNT>(acc, lds_d);
  if (threadIdx.x == 0) publish(a.partial + blockIdx.x, bs);
  if (!arrive_last(a.counter, gridDim.x, &flag)) return;
  double tot = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kNT) tot += consume(a.partial + b);
  tot = block_sum_d<kNT>(tot, lds_d);
  double lp = 0.0;
  if (a.pr.has_prior) {
    const double iv = (double)a.pr.inv_var;
    const double ln = (double)a.pr.log_norm;
    for (int j = threadIdx.x; j < a.D; j += kNT) {
      const double wj = (double)a.theta[j];
      lp += (ln + wj * wj * iv) / -2.0;
    }
  }
  lp = block_sum_d<kNT>(lp, lds_d);
  if (threadIdx.x == 0) a.out_U[0] = (-tot - lp) / a.temperature;
}

// ------------------------------------------------------------ fused path ----

struct FusedArgs {
  const sgmc_glm_chain* chains;
  const int32_t* block_chain;   // grid block -> chain, or NULL with blocks_per_chain
  int blocks_per_chain;         // uniform table: chain = block / blocks_per_chain
  const float* eps_table;
  const int32_t* save_slot;
  int32_t* iter;
  unsigned* done;
  const float* inv_mass;   // shared diagonal mass (nullable = ones)
  const float* sqrt_mass;
  int table_len;
  int nchains;
  int advance;             // the launch's last chain advances the iteration counter;
                           // < 0: no counter bookkeeping (the loop kernel does it)
  int iter_offset;         // iteration of this evaluation = *iter + iter_offset
  int mode;                // SGMC_MODE_*
  int step;                // leapfrog step within the iteration (SGHMC)
  int steps;
  float friction;
  float four_friction;     // float32(4 friction) (AMAGOLD noise scale)
  float alpha, lmbd;       // rms_prop
  Lik L;
  Prior pr;
  float pot_temperature;
};

// Normal for element i of a single-leaf tree of size S under random_tree(sub):
// the leaf key is split(sub, 1)[0].
__device__ __forceinline__ float leaf_normal(Key leafkey, uint32_t S, uint32_t i) {
  return bits_to_normal(bits_word(leafkey, S, i));
}

// MODE: a SGMC_MODE_* fixed at compile time (its branches folded away, fewer registers),
// or -1 for the mode in a.mode.
template <int LPR, int CPL, int MODE>
__device__ __forceinline__ void fused_body(const FusedArgs& a, const sgmc_glm_chain& ch,
                                           int bchain, float* lds_cols, float* lds_small,
                                           float* lds_lik, int* flag) {
  const int mode = MODE >= 0 ? MODE : a.mode;
  const int it = max(0, min(a.iter[0] + a.iter_offset, a.table_len - 1));
  const float eps = a.eps_table[it];
  const int D = ch.D;
  const int D4 = (D + 3) >> 2;
  gfloat* g_theta = to_global(ch.theta);
  gfloat* g_mom = to_global(ch.mom);
  gfloat* g_v = to_global(ch.v);
  gfloat* g_partial = to_global(ch.partial);
  gfloat* g_lik = to_global(ch.lik);
  gfloat* g_samples = to_global(ch.samples);
  gcfloat* g_inv_mass = to_global(a.inv_mass);
  gcfloat* g_sqrt_mass = to_global(a.sqrt_mass);
  const Key key0 = load_key(ch.key);
  const bool hmc = mode == SGMC_MODE_SGHMC;
  const bool amg = mode == SGMC_MODE_AMAGOLD;
  const bool ob_mh = mode == SGMC_MODE_SGGMC;
  const bool ob = mode == SGMC_MODE_OBABO || ob_mh;
  const bool has_mom = hmc || amg || ob;
  const bool shift = hmc || amg;  // evaluation point theta + scale M^-1 p
  const bool resample = hmc && a.step == 0;
  // OBABO: two launches per leapfrog step, the even one O B A, the odd one B O.
  const int sub = a.step & 1;
  const int launches = ob ? 2 * a.steps : a.steps;
  // Position step before this evaluation: eps, or eps/2 at the first step of an
  // AMAGOLD iteration (integrator.py:521-523, 401-408).
  const float half = 0.5f * eps;
  const float pos_scale = (amg && a.step == 0) ? half : eps;

  // Momentum used by this step: resampled p = sqrt_mass * xi at the first step.
  Key key_noise = key0;  // key split for this step's noise
  Key resample_leaf{0, 0};
  if (resample) {
    Key k1, sub;
    split2(key0, k1, sub);
    resample_leaf = split_key(sub, 1, 0);
    key_noise = k1;
  }

  // Evaluation point: theta (SGLD) or theta + scale M^-1 p (SGHMC, AMAGOLD).
  float4 th[CPL];
  {
    const int q = (threadIdx.x & (kWave - 1)) % LPR;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      float e[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int c = 4 * (q + LPR * k) + t;
        float v = 0.0f;
        if (c < D) {
          v = g_theta[c];
          if (shift) {
            float p;
            if (resample) {
              const float xi = leaf_normal(resample_leaf, (uint32_t)D, (uint32_t)c);
              p = a.sqrt_mass ? g_sqrt_mass[c] * xi : xi;
            } else {
              p = g_mom[c];
            }
            const float s = a.inv_mass ? g_inv_mass[c] * p : p;
            v = v + pos_scale * s;
          }
        }
        e[t] = v;
      }
      th[k] = make_float4(e[0], e[1], e[2], e[3]);
    }
  }

  // Minibatch indices: (dkey', sub) = split(dkey); k1, k2 = split(sub).
  Key dnext, dsub;
  split2(load_key(ch.dkey), dnext, dsub);
  RowSource src;
  src.idx = nullptr;
  src.draw = 1;
  split2(dsub, src.k1, src.k2);
  src.n = (uint32_t)ch.n;
  src.span = (uint32_t)ch.N;
  src.mult = randint_mult(src.span);
  src.magic = randint_magic(src.span);
  src.rows = ch.N;
  // Noise key of this step (used by the chain's last block only; derived here, where
  // it overlaps the wait for the evaluation point, instead of after the arrival).
  // OBABO: (key', s1, s2) = split(key, 3) at both launches of a leapfrog step; the even
  // one draws from s1, the odd one from s2 and stores key'.
  Key knext, leaf;
  if (ob) {
    Key k0, s1, s2;
    split3(key0, k0, s1, s2);
    leaf = split_key(sub ? s2 : s1, 1, 0);
    knext = sub ? k0 : key0;
  } else {
    Key ksub;
    split2(key_noise, knext, ksub);
    leaf = split_key(ksub, 1, 0);
  }

  float4 acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  float lik_acc = 0.0f;
  int r0, r1, b0, b1;
  wave_rows(ch.n, ch.nblocks, bchain, r0, r1, b0, b1);
  accumulate_rows<LPR, CPL>(to_global(ch.X), ch.ldx, to_global(ch.y), D4, src, r0, r1, th, a.L,
                            acc, lik_acc, lds_lik, b0);
  // Block record: column sums, then the block's lik sum and its lik sum of squared
  // deviations about the block mean (combined into var(lik) by the last block).
  const int stride = 4 * D4 + 2;
  gfloat* mine = g_partial + (int64_t)bchain * stride;
  const float ls = block_reduce_write<LPR, CPL>(acc, lik_acc, D4, mine, lds_cols, lds_small);
  const int cnt = b1 - b0;
  float m2 = 0.0f;
  if (ch.var) {
    const float mb = cnt > 0 ? ls / (float)cnt : 0.0f;
    float s2 = 0.0f;
    for (int j = threadIdx.x; j < cnt; j += kNT) {
      const float d = lds_lik[j] - mb;
      s2 += d * d;
    }
    m2 = block_sum<kNT>(s2, lds_small);
  }
  if (threadIdx.x == 0) {
    publish(mine + 4 * D4, ls);
    publish(mine + 4 * D4 + 1, m2);
  }
  if (ch.lik)
    for (int j = threadIdx.x; j < cnt; j += kNT) g_lik[b0 + j] = lds_lik[j];
  // A chain of one block needs no arrival counter: a barrier makes its records
  // visible to its own threads.
  if (ch.nblocks == 1) {
    __syncthreads();
  } else if (!arrive_last(ch.counter, ch.nblocks, flag)) {
    return;
  }

  // ---- last block of this chain: potential, gradient and the state update ----
  // Loads first: each of the first nb threads takes its block's (lik sum, M2) record
  // now; the reductions that need them wait until after the parameter update, whose
  // own loads (column sums, theta, momentum) do not depend on them, so the whole
  // finalizer pays one memory round trip instead of one per phase.
  const int nb = ch.nblocks;
  const float c = (float)(-ch.N) / (float)ch.n;
  const int per_block = (ch.n + nb - 1) / nb;
  float rec_ls = 0.0f, rec_m2 = 0.0f;
  if (threadIdx.x < nb) {
    rec_ls = consume(g_partial + (int64_t)threadIdx.x * stride + 4 * D4);
    if (ch.var) rec_m2 = consume(g_partial + (int64_t)threadIdx.x * stride + 4 * D4 + 1);
  }

  // AMAGOLD and SGGMC samples are saved after the Metropolis-Hastings test, not here.
  const int slot = (a.save_slot && !amg && !ob_mh) ? a.save_slot[it] : -1;
  const bool last_step = !has_mom || a.step == launches - 1;
  const float Tn = ch.temperature;
  // obabo O-step coefficients (integrator.py:192-200): sqrt(exp(-f eps)) and
  // sqrt((1 - exp(-f eps)) T), float32.
  float ob_decay = 0.0f, ob_noise = 0.0f;
  if (ob) {
    const float perm = expf((-a.friction) * eps);
    ob_noise = sqrtf((1.0f - perm) * Tn);
    ob_decay = sqrtf(perm);
  }
  const bool rms = mode == SGMC_MODE_SGLD_RMS;
  // reversible_leapfrog momentum coefficients (integrator.py:420-437), float32.
  const float ef = eps * a.friction;
  const float am_decay = 1.0f - ef;
  const float am_rescale = 1.0f / (1.0f + ef);
  const float am_noise = sqrtf(a.four_friction * eps);
  gfloat* g_theta_old = to_global(ch.theta_old);
  gfloat* g_mom_old = to_global(ch.mom_old);
  float edot = 0.0f;

  float lp = 0.0f;
  const uint32_t S = (uint32_t)D;
  const uint32_t m = (S + 1) >> 1;
  for (uint32_t jp = threadIdx.x; jp < m; jp += kNT) {
    // Element jp and its partner jp + m share one Threefry block of the noise draw.
    uint32_t w0, w1;
    bits_pair(leaf, S, jp, w0, w1);
    const bool two = jp + m < S;
    const uint32_t jj[2] = {jp, two ? jp + m : jp};
    // every load of both columns is issued before the arithmetic
    float s[2];
    sum_strided2(g_partial + jj[0], g_partial + jj[1], stride, nb, s[0], s[1]);
    float wv[2], pv[2], imv[2], smv[2], vv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t j = jj[h];
      wv[h] = g_theta[j];
      pv[h] = (has_mom && !resample) ? g_mom[j] : 0.0f;
      imv[h] = (has_mom && a.inv_mass) ? g_inv_mass[j] : 1.0f;
      smv[h] = ((resample || amg || ob) && a.sqrt_mass) ? g_sqrt_mass[j] : 1.0f;
      vv[h] = rms ? g_v[j] : 0.0f;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h && !two) break;
      const uint32_t j = jj[h];
      const float xi = bits_to_normal(h ? w1 : w0);
      float wj = wv[h];
      float p = 0.0f, sm = 0.0f;
      if (has_mom) {
        if (resample) {
          const float xr = leaf_normal(resample_leaf, S, j);
          p = a.sqrt_mass ? smv[h] * xr : xr;
        } else {
          p = pv[h];
        }
        sm = a.inv_mass ? imv[h] * p : p;
        if (shift) wj = wj + pos_scale * sm;  // evaluation point
      }
      float g = c * s[h];
      if (a.pr.has_prior) {
        g = g + wj * a.pr.inv_var;
        lp += (a.pr.log_norm + (wj * wj) * a.pr.inv_var) / -2.0f;
      }
      g = g / a.pot_temperature;
      float wn;
      if (!has_mom) {
        const float sg = (-eps) * g;
        const float sn = sqrtf((2.0f * Tn) * eps) * xi;
        if (rms) {
          const float vn = a.alpha * vv[h] + (1.0f - a.alpha) * (g * g);
          g_v[j] = vn;
          const float G = 1.0f / (a.lmbd + sqrtf(vn));
          wn = wj + ((0.0f + G * sg) + sqrtf(G) * sn);
        } else {
          wn = wj + (sg + sn);
        }
      } else if (hmc) {
        const float C = a.friction;
        float pn = p + ((-eps) * C) * sm;
        pn = pn + (-eps) * g;
        pn = pn + C * (sqrtf(2.0f * eps) * xi);
        g_mom[j] = pn;
        wn = wj;
      } else if (ob) {
        if (sub == 0) {
          // O: p1 = decay p + noise (M^{1/2} xi); B: p2 = (-eps/2) g + p1;
          // A: theta += eps M^-1 p2 (integrator.py:214-238).
          const float p1 = ob_decay * p + ob_noise * (smv[h] * xi);
          edot += p1 * (imv[h] * p1);
          const float p2 = (-half) * g + p1;
          if (ob_mh && a.step == 0) {
            g_theta_old[j] = wv[h];
            g_mom_old[j] = p;
          }
          g_mom[j] = p2;
          wn = wj + eps * (imv[h] * p2);
        } else {
          // B: p3 = (-eps/2) g + p2; O with the second noise (integrator.py:244-262).
          const float p3 = (-half) * g + p;
          edot += p3 * (imv[h] * p3);
          g_mom[j] = ob_decay * p3 + ob_noise * (smv[h] * xi);
          wn = wj;
        }
      } else {
        // p' = (((1 - eps f) p + (-eps) g) + sqrt(4 f eps) (M^{1/2} xi)) / (1 + eps f);
        // energy term (p + p') . (M^-1 g); final half position step after the last.
        const float noise = am_noise * (smv[h] * xi);
        const float pn = am_rescale * ((am_decay * p + (-eps) * g) + noise);
        edot += (p + pn) * (imv[h] * g);
        if (a.step == 0) {
          g_theta_old[j] = wv[h];
          g_mom_old[j] = p;
        }
        g_mom[j] = pn;
        wn = last_step ? wj + half * (imv[h] * pn) : wj;
      }
      g_theta[j] = wn;
      if (slot >= 0 && last_step && ch.samples) g_samples[(int64_t)slot * D + j] = wn;
    }
  }
  // var(lik) = (sum_b M2_b + cnt_b (mean_b - mean)^2) / n  (pairwise/Chan combination)
  float sl = rec_ls;
  for (int b = threadIdx.x + kNT; b < nb; b += kNT)
    sl += consume(g_partial + (int64_t)b * stride + 4 * D4);
  const float sumlik = block_sum<kNT>(sl, lds_small);
  const float mean = sumlik / (float)ch.n;
  float var_sum = 0.0f;
  if (ch.var) {
    float t = 0.0f;
    for (int b = threadIdx.x; b < nb; b += kNT) {
      const int lo = min(ch.n, b * per_block);
      const int k = min(ch.n, lo + per_block) - lo;
      if (k > 0) {
        const bool held = b < kNT;
        const float ls_b = held ? rec_ls : consume(g_partial + (int64_t)b * stride + 4 * D4);
        const float m2_b = held ? rec_m2 : consume(g_partial + (int64_t)b * stride + 4 * D4 + 1);
        const float d = ls_b / (float)k - mean;
        t += m2_b + (float)k * (d * d);
      }
    }
    var_sum = block_sum<kNT>(t, lds_small);
  }
  lp = block_sum<kNT>(lp, lds_small);
  if (amg || ob) edot = block_sum<kNT>(edot, lds_small);
  if (threadIdx.x == 0) {
    if (amg) ch.energy[0] = half * edot + (a.step == 0 ? 0.0f : ch.energy[0]);
    if (ob) ch.energy[sub] = ch.energy[sub] + 0.5f * edot;
    const float Uc = ((float)(-ch.N) * mean - lp) / a.pot_temperature;
    // OBABO's potential is the mean of the step's two evaluations.
    float Uk = Uc;
    if (ob && sub) Uk = 0.5f * (ch.U[0] + Uc);
    if (ob_mh && a.step == 0) ch.energy[2] = ch.U[0];
    ch.U[0] = Uk;
    if (slot >= 0 && last_step && ch.samples_U) ch.samples_U[slot] = Uk;
    if (ch.var) ch.var[0] = var_sum / (float)ch.n;
    store_key(ch.key, knext);
    store_key(ch.dkey, dnext);
    // Last chain of the launch advances the iteration counter (read by the next launch).
    if (a.advance >= 0) {
      const unsigned prev = atomicAdd(a.done, 1u);
      if (prev == (unsigned)a.nchains - 1) {
        atomicExch(a.done, 0u);
        if (last_step && a.advance) atomicAdd(a.iter, 1);
      }
    }
  }
}

// ---- resident loop: one single-block SGLD chain, many iterations on chip ----
// For chains of D <= 4 LPR (one float4 chunk per lane of a row segment) and minibatches
// of at most 256 rows (one 64-row chunk per wavefront), the chain's state stays on chip
// for the whole launch: theta in LDS (and its owner threads' registers), the keys in
// registers, the wavefronts' column partials in LDS. The step's noise is drawn one
// iteration ahead, beside the previous update; the minibatch of the next evaluation (its
// indices depend only on the data key) is drawn and its rows loaded right after this
// evaluation's row groups are consumed, so their latency overlaps the reduction and the
// update. The arithmetic
// is fused_body's for a chain of one block, in the same order (row groups, reductions,
// the SGLD / rms_prop update, U, var), so both give the same bits.
template <int LPR, int MODE>
__device__ __forceinline__ void resident_loop(const FusedArgs& a, const sgmc_glm_chain& ch,
                                              int iterations, float* lds_cols, float* lds_small,
                                              float* lds_lik, float* th_lds, float* eps_lds,
                                              int* slot_lds, uint32_t* kp, float* xis) {
  constexpr int CPL = 1;
  using RG = RowGroup<LPR, CPL>;
  constexpr int G = (kWave / LPR) * RG::U;  // rows per group
  constexpr int NG = kWave / G;             // groups of a 64-row chunk
  constexpr int kSegs = kWave / LPR;         // row segments of a wavefront
  const int mode = MODE >= 0 ? MODE : a.mode;
  const bool rms = mode == SGMC_MODE_SGLD_RMS;
  const int D = ch.D, D4 = (D + 3) >> 2, n = ch.n;
  gfloat* g_theta = to_global(ch.theta);
  gfloat* g_v = to_global(ch.v);
  gfloat* g_samples = to_global(ch.samples);
  const int it0 = a.iter[0];
  const int wv = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  // Thread jp < m owns the noise pair (jp, jp + m), as in fused_body's update loop.
  const uint32_t S = (uint32_t)D, m = (S + 1) >> 1;
  const uint32_t jp = threadIdx.x;
  const bool own = jp < m;
  const bool two = own && jp + m < S;
  const uint32_t jj[2] = {jp, two ? jp + m : jp};
  float w[2] = {0.0f, 0.0f}, v[2] = {0.0f, 0.0f};
  if (own) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      w[h] = g_theta[jj[h]];
      if (rms) v[h] = g_v[jj[h]];
    }
  }
  for (int c = threadIdx.x; c < 4 * LPR; c += kNT) th_lds[c] = c < D ? g_theta[c] : 0.0f;
  int r0, r1, b0, b1;
  wave_rows(n, 1, 0, r0, r1, b0, b1);
  const int ng = r0 < r1 ? (r1 - r0 + G - 1) / G : 0;  // wavefront-uniform
  RowSource src;
  src.idx = nullptr;
  src.draw = 1;
  src.n = (uint32_t)n;
  src.span = (uint32_t)ch.N;
  src.mult = randint_mult(src.span);
  src.magic = randint_magic(src.span);
  src.rows = ch.N;
  RG buf[NG];
  // The theta-independent key work runs ahead of its use, after the reduction barrier
  // (beside wavefront 0's update), handed over through LDS (kp, xis; slot = iteration
  // parity) across the loop's barriers. With key_{i+1}, ks_i = split(key_i), leaf_i =
  // split_key(ks_i, 1, 0) (the noise of iteration i), d_{i+1}, dsub_i = split(d_i) and
  // (k1_i, k2_i) = split(dsub_i) (the minibatch of iteration i), at iteration t:
  // - wavefront 3 runs one Threefry block per lane for all key steps at once: lanes 0-1
  //   split(key_{t+3}), lane 2 split_key(ks_{t+2}), lanes 3-4 split(d_{t+3}), lanes 5-6
  //   split(dsub_{t+2}); leaf_{t+2} goes to kp[2 slot ..], (k1, k2)_{t+2} to
  //   kp[4 + 4 slot ..], the chain states stay in its registers;
  // - wavefront 2 draws the normals of iteration t + 1 (lane l: word l / 32 of noise
  //   pair l % 32) into xis[64 slot' + 2 jp + h].
  // kp[12 ..]: the keys after the launch's last iteration.
  auto put_key = [&](int at, Key k) { kp[at] = k.k0; kp[at + 1] = k.k1; };
  auto get_key = [&](int at) { return Key{kp[at], kp[at + 1]}; };
  // the wavefront's rows of a minibatch with randint keys (k1, k2) into buf; a wavefront
  // of at most 32 rows draws their two randint words on the two lane halves (lanes 0-31
  // the words of k1, lanes 32-63 those of k2), one Threefry block per lane
  const bool halves = r1 - r0 <= kWave / 2;
  auto draw_rows = [&](Key k1, Key k2) {
    src.k1 = k1;
    src.k2 = k2;
    int rid = -1;
    if (halves) {
      const int rr = lane & (kWave / 2 - 1);
      const Key kk = lane < kWave / 2 ? src.k1 : src.k2;
      const uint32_t word = bits_word(kk, src.n, (uint32_t)(r0 + rr));
      const uint32_t lo = (uint32_t)__shfl((int)word, rr + kWave / 2, kWave);
      if (lane < kWave / 2 && r0 + rr < r1) {
        const int64_t r = (int64_t)randint_offset(word, lo, src.span, src.mult, src.magic);
        if (r >= 0 && r < src.rows) rid = (int)r;
      }
    } else if (r0 < r1) {
      rid = chunk_rows(src, r0, r1);
    }
#pragma unroll
    for (int gi = 0; gi < NG; ++gi)
      if (gi < ng) load_group<LPR, CPL>(to_global(ch.X), ch.ldx, to_global(ch.y), D4, gi * G, rid,
                                        buf[gi]);
  };
  // wavefront 3's chain states at iteration t: key_{t+3}, ks_{t+2}, d_{t+3}, dsub_{t+2}
  Key kc, ksn, dc, dsn;
  {
    // keys of iterations 0 to 2 (every thread; uniform values)
    Key key1, key2, ks0, ks1, d1, d2, ds0, ds1, k1, k2, k1n, k2n;
    split2_uniform(load_key(ch.key), key1, ks0);
    split2_uniform(key1, key2, ks1);
    split2_uniform(key2, kc, ksn);
    const Key leaf0 = split_key(ks0, 1, 0);
    const Key leaf1 = split_key(ks1, 1, 0);
    split2_uniform(load_key(ch.dkey), d1, ds0);
    split2_uniform(d1, d2, ds1);
    split2_uniform(d2, dc, dsn);
    split2_uniform(ds0, k1, k2);
    split2_uniform(ds1, k1n, k2n);
    if (threadIdx.x == 0) {
      put_key(2, leaf1);
      put_key(8, k1n);
      put_key(10, k2n);
      // the keys after the launch (iterations >= 3: set again in the loop)
      put_key(12, iterations == 1 ? key1 : iterations == 2 ? key2 : kc);
      put_key(14, iterations == 1 ? d1 : iterations == 2 ? d2 : dc);
    }
    if (own) {  // the normals of iteration 0
      uint32_t wb0, wb1;
      bits_pair(leaf0, S, jp, wb0, wb1);
      xis[2 * jp] = bits_to_normal(wb0);
      if (two) xis[2 * jp + 1] = bits_to_normal(wb1);
    }
    draw_rows(k1, k2);
  }
  const float c = (float)(-ch.N) / (float)n;
  const float Tn = ch.temperature;
  float Uk = 0.0f;
  // Barriers inside the loop wait for LDS only (lds_sync): the next minibatch's row
  // loads stay in flight through the reduction and the update. The step sizes and save
  // slots come from LDS (a window of kResWindow iterations), so no global load has to be
  // waited on behind those row loads. An iteration has two barriers (the wavefronts'
  // reduction records, the new theta), four more where U or var(lik) is kept.
  for (int t = 0; t < iterations; ++t) {
    const int tw = t % kResWindow;
    if (tw == 0) {
      __syncthreads();
      for (int u = threadIdx.x; u < kResWindow; u += kNT) {
        const int iu = max(0, min(it0 + t + u, a.table_len - 1));
        eps_lds[u] = a.eps_table[iu];
        slot_lds[u] = a.save_slot ? a.save_slot[iu] : -1;
      }
      __syncthreads();
    }
    const float eps = eps_lds[tw];
    const int slot = slot_lds[tw];
    const int par = t & 1;
    float4 th[CPL];
    {
      const int q = (threadIdx.x & (kWave - 1)) % LPR;
      th[0] = make_float4(th_lds[4 * q], th_lds[4 * q + 1], th_lds[4 * q + 2], th_lds[4 * q + 3]);
    }
    float4 acc[CPL];
    acc[0] = make_float4(0.f, 0.f, 0.f, 0.f);
    float lik_acc = 0.0f;
#pragma unroll
    for (int gi = 0; gi < NG; ++gi)
      if (gi < ng) process_group<LPR, CPL>(buf[gi], r0 + gi * G, D4, th, a.L, acc, lik_acc, lds_lik, 0);
    if (t + 1 < iterations) draw_rows(get_key(4 + 4 * (par ^ 1)), get_key(6 + 4 * (par ^ 1)));
    // Every lane's column partials go to LDS unreduced (one barrier); each owner thread
    // then adds its two columns over the row segments in the order of block_reduce's
    // segment shuffles (pairs of segments, pairs of pairs, ...) and over the wavefronts
    // in block_reduce's order. The lik sum is only formed where U is kept.
    const bool want_u = (slot >= 0 && ch.samples_U) || t == iterations - 1;
    {
      reinterpret_cast<float4*>(lds_cols)[wv * kWave + lane] = acc[0];
      if (want_u) {
        const float lw = wave_sum(lik_acc);
        if (lane == 0) lds_small[wv] = lw;
      }
    }
    // the owners' terms that do not need the column sums, before the barrier: the prior
    // gradient and log density, the noise term
    float lp = 0.0f, pg[2] = {0.0f, 0.0f}, sn[2] = {0.0f, 0.0f};
    if (own) {
      const float nscale = sqrtf((2.0f * Tn) * eps);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h && !two) break;
        const float wj = w[h];
        if (a.pr.has_prior) {
          pg[h] = wj * a.pr.inv_var;
          lp += (a.pr.log_norm + (wj * wj) * a.pr.inv_var) / -2.0f;
        }
        sn[h] = nscale * xis[64 * par + 2 * jp + h];
      }
    }
    lds_sync();
    if (wv == 3) {  // every key step of this iteration in one Threefry block per lane
      Key k = kc;
      uint32_t x0 = lane & 1, x1 = (lane & 1) + 2u;  // lanes 0, 1: split(key_{t+3})
      if (lane == 2) {  // split_key(ks_{t+2}, 1, 0)
        k = ksn;
        x0 = 0u;
        x1 = 1u;
      } else if (lane == 3 || lane == 4) {  // split(d_{t+3})
        k = dc;
        x0 = lane - 3;
        x1 = lane - 1;
      } else if (lane >= 5) {  // lanes 5, 6: split(dsub_{t+2})
        k = dsn;
        x0 = (lane - 5) & 1;
        x1 = x0 + 2u;
      }
      threefry2x32(k.k0, k.k1, x0, x1);
      auto rl = [](uint32_t v, int src) {
        return (uint32_t)__builtin_amdgcn_readlane((int)v, src);
      };
      if (t + 3 == iterations && lane == 0) {
        put_key(12, kc);
        put_key(14, dc);
      }
      if (lane == 0) {
        put_key(2 * par, Key{rl(x0, 2), rl(x1, 2)});
        put_key(4 + 4 * par, Key{rl(x0, 5), rl(x0, 6)});
        put_key(6 + 4 * par, Key{rl(x1, 5), rl(x1, 6)});
      }
      kc = Key{rl(x0, 0), rl(x0, 1)};
      ksn = Key{rl(x1, 0), rl(x1, 1)};
      dc = Key{rl(x0, 3), rl(x0, 4)};
      dsn = Key{rl(x1, 3), rl(x1, 4)};
    } else if (wv == 2 && t + 1 < iterations) {  // normals of iteration t + 1
      const uint32_t q = (uint32_t)lane & 31u, h = (uint32_t)lane >> 5;
      if (q < m && (h == 0 || q + m < S)) {
        uint32_t wb0, wb1;
        bits_pair(get_key(2 * (par ^ 1)), S, q, wb0, wb1);
        xis[64 * (par ^ 1) + 2 * q + h] = bits_to_normal(h ? wb1 : wb0);
      }
    }
    if (own) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h && !two) break;
        const uint32_t j = jj[h];
        float sc = 0.0f;
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) {
          float sv[kSegs];
#pragma unroll
          for (int sg = 0; sg < kSegs; ++sg) sv[sg] = lds_cols[(ww * kWave + sg * LPR) * 4 + j];
#pragma unroll
          for (int w2 = 1; w2 < kSegs; w2 <<= 1)
#pragma unroll
            for (int i = 0; i < kSegs; i += 2 * w2) sv[i] = sv[i] + sv[i + w2];
          sc += sv[0];
        }
        const float wj = w[h];
        // one published record: the last block's sum over records starts from +0
        float g = c * (0.0f + sc);
        if (a.pr.has_prior) g = g + pg[h];
        if (a.pot_temperature != 1.0f) g = g / a.pot_temperature;  // x / 1 is x
        const float sg = (-eps) * g;
        float wn;
        if (rms) {
          const float vn = a.alpha * v[h] + (1.0f - a.alpha) * (g * g);
          v[h] = vn;
          const float Gs = 1.0f / (a.lmbd + sqrtf(vn));
          wn = wj + ((0.0f + Gs * sg) + sqrtf(Gs) * sn[h]);
        } else {
          wn = wj + (sg + sn[h]);
        }
        w[h] = wn;
        th_lds[j] = wn;
        if (slot >= 0 && ch.samples) g_samples[(int64_t)slot * D + j] = wn;
      }
    }
    // U (and var(lik)) only where they are kept: a save slot or the last iteration
    const bool last = t == iterations - 1;
    if (want_u) {
      float ls = 0.0f;  // read before block_sum's barrier lets lds_small be reused
#pragma unroll
      for (int i = 0; i < kWaves; ++i) ls += lds_small[i];
      lp = block_sum<kNT, true>(lp, lds_small);
      // the block's lik sum is the only record: the sum over records is ls + 0
      const float mean = (ls + 0.0f) / (float)n;
      Uk = ((float)(-ch.N) * mean - lp) / a.pot_temperature;
      if (threadIdx.x == 0 && slot >= 0 && ch.samples_U) ch.samples_U[slot] = Uk;
      if (ch.var && last) {
        // var(lik) of the last evaluation, as fused_body forms it from one record
        const float mb = n > 0 ? ls / (float)n : 0.0f;
        float s2 = 0.0f;
        for (int jr = threadIdx.x; jr < n; jr += kNT) {
          const float d = lds_lik[jr] - mb;
          s2 += d * d;
        }
        const float m2 = block_sum<kNT, true>(s2, lds_small);
        float tv = 0.0f;
        if (threadIdx.x == 0 && n > 0) {
          const float d = ls / (float)n - mean;
          tv += m2 + (float)n * (d * d);
        }
        const float var_sum = block_sum<kNT, true>(tv, lds_small);
        if (threadIdx.x == 0) ch.var[0] = var_sum / (float)n;
      }
    }
    lds_sync();
  }
  if (own) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h && !two) break;
      g_theta[jj[h]] = w[h];
      if (rms) g_v[jj[h]] = v[h];
    }
  }
  if (threadIdx.x == 0) {
    ch.U[0] = Uk;
    store_key(ch.key, get_key(12));
    store_key(ch.dkey, get_key(14));
  }
}

// Many iterations in one launch for tables whose chains have one block each: block c
// runs chain c's iterations back to back (every leapfrog step of every iteration),
// reading its parameters and keys back from memory after a barrier, so no launch and
// no arrival counter separates two evaluations. The last chain to finish advances the
// iteration counter by the number of iterations.
template <int DCLASS, int MODE>
__global__ void __launch_bounds__(kNT) glm_loop_kernel(FusedArgs a, int iterations) {
  __shared__ __attribute__((aligned(16))) float lds_cols[kWaves * kMaxD];
  __shared__ float lds_small[kWaves];
  __shared__ float lds_lik[kLdsRows];
  __shared__ float res_th[64];
  __shared__ float res_eps[kResWindow];
  __shared__ int res_slot[kResWindow];
  __shared__ uint32_t res_kp[16];
  __shared__ float res_xi[128];
  __shared__ int flag;
  // The grid has one block per chain and every chain one block (sgmc_glm_check_table).
  const int chain = blockIdx.x;
  const sgmc_glm_chain ch = a.chains[chain];
  const int dc = DCLASS >= 0 ? DCLASS : ch.dclass;
  FusedArgs b = a;
  b.advance = -1;
  const int mode = MODE >= 0 ? MODE : a.mode;
  const int launches = mode == SGMC_MODE_OBABO ? 2 * a.steps : a.steps;
  // SGLD chains of D <= 64 and minibatches of at most 256 rows without per-row
  // likelihood output run resident (block-uniform choice)
  bool resident = false;
  constexpr bool kResident = DCLASS == 0 || DCLASS == 5 || DCLASS == 1;
  if constexpr (kResident)
    resident = (mode == SGMC_MODE_SGLD || mode == SGMC_MODE_SGLD_RMS) && !ch.lik &&
               ch.n <= kWaves * kWave;
  if (resident) {
    if constexpr (kResident)
      resident_loop<DCLASS == 0 ? 4 : DCLASS == 5 ? 8 : 16, MODE>(
          a, ch, iterations, lds_cols, lds_small, lds_lik, res_th, res_eps, res_slot, res_kp,
          res_xi);
  }
  for (int t = 0; t < (resident ? 0 : iterations); ++t) {
    b.iter_offset = t;
    for (int st = 0; st < launches; ++st) {
      b.step = st;
      switch (dc) {
        case 0: fused_body<4, 1, MODE>(b, ch, 0, lds_cols, lds_small, lds_lik, &flag); break;
        case 1: fused_body<16, 1, MODE>(b, ch, 0, lds_cols, lds_small, lds_lik, &flag); break;
        case 5: fused_body<8, 1, MODE>(b, ch, 0, lds_cols, lds_small, lds_lik, &flag); break;
        case 2: fused_body<64, 1, MODE>(b, ch, 0, lds_cols, lds_small, lds_lik, &flag); break;
        case 3: fused_body<64, 2, MODE>(b, ch, 0, lds_cols, lds_small, lds_lik, &flag); break;
        default: fused_body<64, 4, MODE>(b, ch, 0, lds_cols, lds_small, lds_lik, &flag); break;
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    const unsigned prev = atomicAdd(a.done, 1u);
    if (prev == (unsigned)a.nchains - 1) {
      atomicExch(a.done, 0u);
      atomicAdd(a.iter, iterations);
    }
  }
}

template <int DCLASS, int MODE>
__global__ void __launch_bounds__(kNT) glm_fused_kernel(FusedArgs a) {
  __shared__ __attribute__((aligned(16))) float lds_cols[kWaves * kMaxD];
  __shared__ float lds_small[kWaves];
  __shared__ float lds_lik[kLdsRows];
  __shared__ int flag;
  // A uniform table maps blocks to chains arithmetically, which saves the dependent
  // load of the map before the chain descriptor can be read.
  // The block map and the chains' block ranges agree (sgmc_glm_check_table checks the
  // table on the host before it is uploaded), so every block belongs to a chain.
  const int chain = a.blocks_per_chain > 0 ? (int)blockIdx.x / a.blocks_per_chain
                                           : a.block_chain[blockIdx.x];
  // The chain descriptor is copied once: block-uniform, it lives in scalar registers
  // instead of being re-read after every barrier and atomic.
  const sgmc_glm_chain ch = a.chains[chain];
  const int bchain = blockIdx.x - ch.block_begin;
  const int dc = DCLASS >= 0 ? DCLASS : ch.dclass;
  switch (dc) {
    case 0: fused_body<4, 1, MODE>(a, ch, bchain, lds_cols, lds_small, lds_lik, &flag); break;
    case 1: fused_body<16, 1, MODE>(a, ch, bchain, lds_cols, lds_small, lds_lik, &flag); break;
    case 5: fused_body<8, 1, MODE>(a, ch, bchain, lds_cols, lds_small, lds_lik, &flag); break;
    case 2: fused_body<64, 1, MODE>(a, ch, bchain, lds_cols, lds_small, lds_lik, &flag); break;
    case 3: fused_body<64, 2, MODE>(a, ch, bchain, lds_cols, lds_small, lds_lik, &flag); break;
    default: fused_body<64, 4, MODE>(a, ch, bchain, lds_cols, lds_small, lds_lik, &flag); break;
  }
}

// The fused kernels of a table: one D class (or -1, mixed, with the class switch per
// block) and the SGLD mode compiled in, or any mode read at run time (MODE = -1).
template <int MODE>
void launch_fused(int dclass, int nblocks, hipStream_t s, const FusedArgs& a) {
  const dim3 grid(nblocks), block(kNT);
  switch (dclass) {
    case 0: hipLaunchKernelGGL((glm_fused_kernel<0, MODE>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((glm_fused_kernel<1, MODE>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((glm_fused_kernel<2, MODE>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((glm_fused_kernel<3, MODE>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((glm_fused_kernel<4, MODE>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((glm_fused_kernel<5, MODE>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((glm_fused_kernel<-1, MODE>), grid, block, 0, s, a); break;
  }
}

template <int MODE>
void launch_loop(int dclass, int nchains, int iterations, hipStream_t s, const FusedArgs& a) {
  const dim3 grid(nchains), block(kNT);
  switch (dclass) {
    case 0: hipLaunchKernelGGL((glm_loop_kernel<0, MODE>), grid, block, 0, s, a, iterations); break;
    case 1: hipLaunchKernelGGL((glm_loop_kernel<1, MODE>), grid, block, 0, s, a, iterations); break;
    case 2: hipLaunchKernelGGL((glm_loop_kernel<2, MODE>), grid, block, 0, s, a, iterations); break;
    case 3: hipLaunchKernelGGL((glm_loop_kernel<3, MODE>), grid, block, 0, s, a, iterations); break;
    case 4: hipLaunchKernelGGL((glm_loop_kernel<4, MODE>), grid, block, 0, s, a, iterations); break;
    case 5: hipLaunchKernelGGL((glm_loop_kernel<5, MODE>), grid, block, 0, s, a, iterations); break;
    default:
      hipLaunchKernelGGL((glm_loop_kernel<-1, MODE>), grid, block, 0, s, a, iterations);
      break;
  }
}

static int dclass_of(int D) {
  if (D <= 16) return 0;
  if (D <= 32) return 5;  // 8 lanes per row (added after 0-4: the numbers are labels)
  if (D <= 64) return 1;
  if (D <= 256) return 2;
  if (D <= 512) return 3;
  return 4;
}

}  // namespace sgmc

using namespace sgmc;

namespace {

// Validates the arguments of a minibatch potential and launches its kernel with the
// block records at `partial` ([nblocks][4 ceil(D/4) + 2] floats) and the arrival counter
// at `counter` (zero before the launch; the last block resets it).
int launch_potential(const char* who, int family, const float* X, int64_t ldx, const float* y,
                     const int32_t* idx, int n, int64_t N, int D, const float* theta,
                     float sigma, int has_prior, float prior_scale, float temperature,
                     float* out_U, float* out_grad, float* out_lik, float* out_var,
                     float* partial, unsigned* counter, int nblocks, void* stream) {
  SGMC_REQUIRE(X && y && theta && out_U && out_grad, SGMC_E_ARG, "%s: null pointer", who);
  SGMC_REQUIRE(family == SGMC_LIK_LINEAR_GAUSSIAN || family == SGMC_LIK_LOGISTIC, SGMC_E_ARG,
               "%s: unknown family %d", who, family);
  SGMC_REQUIRE(D >= 1 && D <= kMaxD, SGMC_E_ARG, "%s: D=%d outside [1, %d]", who, D, kMaxD);
  SGMC_REQUIRE(ldx % 4 == 0 && ldx >= 4 * ((D + 3) / 4), SGMC_E_ARG,
               "%s: ldx=%lld must be a multiple of 4 and >= D rounded up", who, (long long)ldx);
  SGMC_REQUIRE(((uintptr_t)X & 15) == 0, SGMC_E_ARG, "%s: X not 16B aligned", who);
  SGMC_REQUIRE(n >= 1 && N >= 1, SGMC_E_ARG, "%s: empty batch", who);
  SGMC_REQUIRE(nblocks >= 1 && (int64_t)nblocks * kLdsRows >= n, SGMC_E_ARG,
               "%s: nblocks must be >= 1 and >= n / %d", who, kLdsRows);
  SGMC_REQUIRE(temperature != 0.0f, SGMC_E_ARG, "%s: temperature is zero", who);
  PotArgs a;
  a.X = X;
  a.y = y;
  a.idx = idx;
  a.theta = theta;
  a.out_U = out_U;
  a.out_grad = out_grad;
  a.out_lik = out_lik;
  a.out_var = out_var;
  a.D = D;
  a.D4 = (D + 3) / 4;
  a.partial = partial;
  a.counter = counter;
  a.ldx = ldx;
  a.N = N;
  a.n = n;
  a.L = make_lik(family, sigma);
  a.pr = make_prior(has_prior, prior_scale);
  a.temperature = temperature;
  hipStream_t s = as_stream(stream);
  dim3 grid(nblocks), block(kNT);
  switch (dclass_of(D)) {
    case 0: hipLaunchKernelGGL((glm_potential_kernel<4, 1>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((glm_potential_kernel<16, 1>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((glm_potential_kernel<8, 1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((glm_potential_kernel<64, 1>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((glm_potential_kernel<64, 2>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((glm_potential_kernel<64, 4>), grid, block, 0, s, a); break;
  }
  return check_launch(who);
}

// ---- scratch of the no-allocation entry point (sgmc_glm_potential_noalloc) ----
// One 1 MiB array per GPU, held until the process ends: word 0 is the arrival counter
// (padded to 256 bytes, so it stays at one address whatever the call's size), the block
// records follow. Every call on a GPU shares that GPU's array; calls that overlap on one
// GPU race. The counter is zeroed when the array is first used on a device, and every
// call leaves it at zero (the potential's last block resets it); the records are
// written before they are read in every call.
constexpr int64_t kScratchFloats = 1 << 18;
constexpr int64_t kScratchCounterFloats = 64;
constexpr int64_t kScratchRecordFloats = kScratchFloats - kScratchCounterFloats;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) float scratch[kScratchFloats];

__global__ void __launch_bounds__(kNT) zero_floats_kernel(float* p, int n) {
  const int i = blockIdx.x * kNT + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

std::mutex g_scratch_mu;
float* g_scratch[kScratchDevices] = {};

// Device address of the current device's scratch (the host-side name `scratch` is only
// a stand-in for it): looked up once per device with hipGetSymbolAddress, its counter
// zeroed on `stream` at that first use.
int scratch_of_current_device(void* stream, float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  SGMC_REQUIRE(e == hipSuccess, SGMC_E_HIP, "sgmc_glm_potential_noalloc: hipGetDevice: %s",
               hipGetErrorString(e));
  SGMC_REQUIRE(dev >= 0 && dev < kScratchDevices, SGMC_E_ARG,
               "sgmc_glm_potential_noalloc: device %d beyond %d", dev, kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    float* p = nullptr;
    e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    SGMC_REQUIRE(e == hipSuccess && p, SGMC_E_HIP,
                 "sgmc_glm_potential_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(zero_floats_kernel, dim3(1), dim3(kNT), 0, as_stream(stream), p,
                       (int)kScratchCounterFloats);
    const int rc = check_launch("sgmc_glm_potential_noalloc: scratch set-up");
    if (rc != SGMC_OK) return rc;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return SGMC_OK;
}

}  // namespace

extern "C" {

int64_t sgmc_glm_workspace_floats(int n, int D, int nblocks) {
  const int D4 = (D + 3) / 4;
  (void)n;
  return (int64_t)nblocks * (4 * D4 + 2) + 4;
}

int sgmc_glm_blocks(int n, int D) {
  (void)D;
  const int rows = 256;
  int b = (n + rows - 1) / rows;
  return b < 1 ? 1 : b;
}

int sgmc_glm_potential(int family, const float* X, int64_t ldx, const float* y,
                       const int32_t* idx, int n, int64_t N, int D, const float* theta,
                       float sigma, int has_prior, float prior_scale, float temperature,
                       float* out_U, float* out_grad, float* out_lik, float* out_var,
                       float* workspace, int64_t workspace_floats, int nblocks, void* stream) {
  SGMC_REQUIRE(workspace, SGMC_E_ARG, "sgmc_glm_potential: null workspace");
  SGMC_REQUIRE(D >= 1 && D <= kMaxD && nblocks >= 1, SGMC_E_ARG,
               "sgmc_glm_potential: D=%d outside [1, %d] or nblocks=%d", D, kMaxD, nblocks);
  SGMC_REQUIRE(workspace_floats >= sgmc_glm_workspace_floats(n, D, nblocks), SGMC_E_ARG,
               "sgmc_glm_potential: workspace too small");
  const int64_t records = (int64_t)nblocks * (4 * ((D + 3) / 4) + 2);
  return launch_potential("sgmc_glm_potential", family, X, ldx, y, idx, n, N, D, theta, sigma,
                          has_prior, prior_scale, temperature, out_U, out_grad, out_lik,
                          out_var, workspace, reinterpret_cast<unsigned*>(workspace + records),
                          nblocks, stream);
}

int sgmc_glm_potential_noalloc(int family, const float* X, int64_t ldx, const float* y,
                               const int32_t* idx, int n, int64_t N, int D, const float* theta,
                               float sigma, int has_prior, float prior_scale, float temperature,
                               float* out_U, float* out_grad, float* out_lik, float* out_var,
                               int nblocks, void* stream) {
  SGMC_REQUIRE(D >= 1 && D <= kMaxD && nblocks >= 1, SGMC_E_ARG,
               "sgmc_glm_potential_noalloc: D=%d outside [1, %d] or nblocks=%d", D, kMaxD,
               nblocks);
  const int64_t records = (int64_t)nblocks * (4 * ((D + 3) / 4) + 2);
  SGMC_REQUIRE(records <= kScratchRecordFloats, SGMC_E_ARG,
               "sgmc_glm_potential_noalloc: %d blocks of D=%d need %lld record floats; the "
               "static scratch holds %lld (use sgmc_glm_potential with a workspace)",
               nblocks, D, (long long)records, (long long)kScratchRecordFloats);
  float* base = nullptr;
  const int rc = scratch_of_current_device(stream, &base);
  if (rc != SGMC_OK) return rc;
  return launch_potential("sgmc_glm_potential_noalloc", family, X, ldx, y, idx, n, N, D, theta,
                          sigma, has_prior, prior_scale, temperature, out_U, out_grad, out_lik,
                          out_var, base + kScratchCounterFloats,
                          reinterpret_cast<unsigned*>(base), nblocks, stream);
}

int64_t sgmc_glm_full_workspace_doubles(int64_t N, int rows_per_block) {
  return (N + rows_per_block - 1) / rows_per_block + 2;
}

int sgmc_glm_full_potential(int family, const float* X, int64_t ldx, const float* y, int64_t N,
                            int D, const float* theta, float sigma, int has_prior,
                            float prior_scale, double temperature, double* out_U,
                            double* workspace, int64_t workspace_doubles, int rows_per_block,
                            void* stream) {
  SGMC_REQUIRE(X && y && theta && out_U && workspace, SGMC_E_ARG,
               "sgmc_glm_full_potential: null pointer");
  SGMC_REQUIRE(family == SGMC_LIK_LINEAR_GAUSSIAN || family == SGMC_LIK_LOGISTIC, SGMC_E_ARG,
               "sgmc_glm_full_potential: unknown family %d", family);
  SGMC_REQUIRE(D >= 1 && D <= kMaxD, SGMC_E_ARG, "sgmc_glm_full_potential: D out of range");
  SGMC_REQUIRE(ldx % 4 == 0 && ldx >= 4 * ((D + 3) / 4), SGMC_E_ARG,
               "sgmc_glm_full_potential: bad ldx");
  SGMC_REQUIRE(((uintptr_t)X & 15) == 0, SGMC_E_ARG, "sgmc_glm_full_potential: X misaligned");
  SGMC_REQUIRE(N >= 1 && rows_per_block >= 1, SGMC_E_ARG, "sgmc_glm_full_potential: bad size");
  const int64_t nb = (N + rows_per_block - 1) / rows_per_block;
  SGMC_REQUIRE(nb <= 2147483647LL, SGMC_E_ARG, "sgmc_glm_full_potential: too many blocks");
  SGMC_REQUIRE(workspace_doubles >= sgmc_glm_full_workspace_doubles(N, rows_per_block),
               SGMC_E_ARG, "sgmc_glm_full_potential: workspace too small");
  SGMC_REQUIRE(temperature != 0.0, SGMC_E_ARG, "sgmc_glm_full_potential: temperature is zero");
  FullArgs a;
  a.X = X;
  a.y = y;
  a.theta = theta;
  a.out_U = out_U;
  a.partial = workspace;
  a.counter = reinterpret_cast<unsigned*>(workspace + nb);
  a.ldx = ldx;
  a.N = N;
  a.D = D;
  a.D4 = (D + 3) / 4;
  a.rows_per_block = rows_per_block;
  a.L = make_lik(family, sigma);
  a.pr = make_prior(has_prior, prior_scale);
  a.temperature = temperature;
  hipStream_t s = as_stream(stream);
  dim3 grid((unsigned)nb), block(kNT);
  switch (dclass_of(D)) {
    case 0: hipLaunchKernelGGL((glm_full_kernel<4, 1>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((glm_full_kernel<16, 1>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((glm_full_kernel<8, 1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((glm_full_kernel<64, 1>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((glm_full_kernel<64, 2>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((glm_full_kernel<64, 4>), grid, block, 0, s, a); break;
  }
  return check_launch("sgmc_glm_full_potential");
}

int sgmc_glm_dclass(int D) { return dclass_of(D); }

int sgmc_glm_check_table(const sgmc_glm_chain* chains, int nchains, const int32_t* block_chain,
                         int nblocks, int loop) {
  SGMC_REQUIRE(chains && nchains >= 1, SGMC_E_ARG, "sgmc_glm_check_table: no chains");
  if (loop) {
    SGMC_REQUIRE(nblocks == nchains, SGMC_E_ARG,
                 "sgmc_glm_check_table: a loop launch has one block per chain");
  } else {
    SGMC_REQUIRE(nblocks >= nchains, SGMC_E_ARG, "sgmc_glm_check_table: %d blocks for %d chains",
                 nblocks, nchains);
    SGMC_REQUIRE(block_chain || nblocks % nchains == 0, SGMC_E_ARG,
                 "sgmc_glm_check_table: without a block map nblocks must be a multiple of "
                 "nchains");
  }
  const int per = !block_chain ? nblocks / nchains : 1;
  int64_t covered = 0;
  for (int c = 0; c < nchains; ++c) {
    const sgmc_glm_chain& ch = chains[c];
    SGMC_REQUIRE(ch.X && ch.y && ch.theta && ch.key && ch.dkey && ch.U && ch.partial &&
                     ch.counter,
                 SGMC_E_ARG, "sgmc_glm_check_table: chain %d has a null pointer", c);
    SGMC_REQUIRE(ch.D >= 1 && ch.D <= kMaxD && ch.dclass == dclass_of(ch.D), SGMC_E_ARG,
                 "sgmc_glm_check_table: chain %d: D=%d or dclass %d", c, ch.D, ch.dclass);
    SGMC_REQUIRE(ch.n >= 1 && ch.N >= 1 && ch.N <= 2147483647LL, SGMC_E_ARG,
                 "sgmc_glm_check_table: chain %d: n=%d, N=%lld", c, ch.n, (long long)ch.N);
    SGMC_REQUIRE(ch.ldx % 4 == 0 && ch.ldx >= 4 * ((ch.D + 3) / 4) &&
                     ((uintptr_t)ch.X & 15) == 0,
                 SGMC_E_ARG, "sgmc_glm_check_table: chain %d: rows not 16-byte aligned", c);
    SGMC_REQUIRE(ch.nblocks >= 1 && ch.block_begin >= 0 &&
                     (int64_t)ch.block_begin + ch.nblocks <= nblocks,
                 SGMC_E_ARG, "sgmc_glm_check_table: chain %d: blocks [%d, %d) outside the grid",
                 c, ch.block_begin, ch.block_begin + ch.nblocks);
    // a block stages the likelihoods of its ceil(n / nblocks) rows in LDS
    SGMC_REQUIRE((ch.n + ch.nblocks - 1) / ch.nblocks <= kLdsRows, SGMC_E_ARG,
                 "sgmc_glm_check_table: chain %d: %d rows over %d blocks exceed %d rows per "
                 "block", c, ch.n, ch.nblocks, kLdsRows);
    if (loop) {
      // the loop launch runs chain c on block c (block_begin is not read)
      SGMC_REQUIRE(ch.nblocks == 1, SGMC_E_ARG,
                   "sgmc_glm_check_table: chain %d has %d blocks; a loop launch needs one", c,
                   ch.nblocks);
    } else if (!block_chain) {
      SGMC_REQUIRE(ch.nblocks == per && ch.block_begin == c * per, SGMC_E_ARG,
                   "sgmc_glm_check_table: chain %d must own blocks [%d, %d)", c, c * per,
                   (c + 1) * per);
    } else {
      for (int b = ch.block_begin; b < ch.block_begin + ch.nblocks; ++b)
        SGMC_REQUIRE(block_chain[b] == c, SGMC_E_ARG,
                     "sgmc_glm_check_table: block %d maps to chain %d, not %d", b,
                     block_chain[b], c);
    }
    covered += ch.nblocks;
  }
  // Every chain's range maps to it and the ranges add up to the grid: the ranges tile
  // the grid and every map entry names its owner.
  SGMC_REQUIRE(covered == nblocks, SGMC_E_ARG,
               "sgmc_glm_check_table: chains own %lld of %d blocks", (long long)covered, nblocks);
  return SGMC_OK;
}

int sgmc_glm_fused_step(const sgmc_glm_chain* chains, int nchains, const int32_t* block_chain,
                        int nblocks, int uniform_dclass, int family, float sigma, int has_prior,
                        float prior_scale, float pot_temperature, int mode, int step, int steps,
                        float friction, float alpha, float lmbd, const float* inv_mass,
                        const float* sqrt_mass, const float* eps_table, int table_len,
                        const int32_t* save_slot, int32_t* iter, unsigned* done,
                        int advance_iter, void* stream) {
  SGMC_REQUIRE(chains && eps_table && iter && done, SGMC_E_ARG,
               "sgmc_glm_fused_step: null pointer");
  SGMC_REQUIRE(block_chain || (nchains > 0 && nblocks % nchains == 0), SGMC_E_ARG,
               "sgmc_glm_fused_step: without a block map nblocks must be a multiple of "
               "nchains");
  SGMC_REQUIRE(nchains >= 1 && nblocks >= nchains && table_len >= 1, SGMC_E_ARG,
               "sgmc_glm_fused_step: bad sizes");
  SGMC_REQUIRE(family == SGMC_LIK_LINEAR_GAUSSIAN || family == SGMC_LIK_LOGISTIC, SGMC_E_ARG,
               "sgmc_glm_fused_step: unknown family %d", family);
  SGMC_REQUIRE(mode == SGMC_MODE_SGLD || mode == SGMC_MODE_SGLD_RMS ||
               mode == SGMC_MODE_SGHMC || mode == SGMC_MODE_AMAGOLD ||
               mode == SGMC_MODE_OBABO || mode == SGMC_MODE_SGGMC,
               SGMC_E_ARG, "sgmc_glm_fused_step: unknown mode %d", mode);
  const int launches =
      (mode == SGMC_MODE_OBABO || mode == SGMC_MODE_SGGMC) ? 2 * steps : steps;
  SGMC_REQUIRE(steps >= 1 && step >= 0 && step < launches, SGMC_E_ARG,
               "sgmc_glm_fused_step: launch %d of %d", step, launches);
  SGMC_REQUIRE(pot_temperature != 0.0f, SGMC_E_ARG, "sgmc_glm_fused_step: temperature is zero");
  FusedArgs a;
  a.chains = chains;
  a.block_chain = block_chain;
  a.blocks_per_chain = block_chain ? 0 : nblocks / nchains;
  a.eps_table = eps_table;
  a.save_slot = save_slot;
  a.iter = iter;
  a.done = done;
  a.inv_mass = inv_mass;
  a.sqrt_mass = sqrt_mass;
  a.table_len = table_len;
  a.nchains = nchains;
  a.mode = mode;
  a.step = step;
  a.steps = steps;
  a.friction = friction;
  a.four_friction = (float)(4.0 * (double)friction);
  a.alpha = alpha;
  a.lmbd = lmbd;
  a.L = make_lik(family, sigma);
  a.pr = make_prior(has_prior, prior_scale);
  a.pot_temperature = pot_temperature;
  a.advance = advance_iter ? 1 : 0;
  a.iter_offset = 0;
  hipStream_t s = as_stream(stream);
  if (mode == SGMC_MODE_SGLD)
    launch_fused<SGMC_MODE_SGLD>(uniform_dclass, nblocks, s, a);
  else
    launch_fused<-1>(uniform_dclass, nblocks, s, a);
  return check_launch("sgmc_glm_fused_step");
}

int sgmc_glm_fused_loop(const sgmc_glm_chain* chains, int nchains, int uniform_dclass,
                        int family, float sigma, int has_prior, float prior_scale,
                        float pot_temperature, int mode, int steps, float friction, float alpha,
                        float lmbd, const float* inv_mass, const float* sqrt_mass,
                        const float* eps_table, int table_len, const int32_t* save_slot,
                        int32_t* iter, unsigned* done, int iterations, void* stream) {
  SGMC_REQUIRE(chains && eps_table && iter && done, SGMC_E_ARG,
               "sgmc_glm_fused_loop: null pointer");
  SGMC_REQUIRE(nchains >= 1 && table_len >= 1 && iterations >= 0, SGMC_E_ARG,
               "sgmc_glm_fused_loop: bad sizes");
  SGMC_REQUIRE(family == SGMC_LIK_LINEAR_GAUSSIAN || family == SGMC_LIK_LOGISTIC, SGMC_E_ARG,
               "sgmc_glm_fused_loop: unknown family %d", family);
  SGMC_REQUIRE(mode == SGMC_MODE_SGLD || mode == SGMC_MODE_SGLD_RMS ||
                   mode == SGMC_MODE_SGHMC || mode == SGMC_MODE_OBABO,
               SGMC_E_ARG, "sgmc_glm_fused_loop: mode %d needs a test between iterations",
               mode);
  SGMC_REQUIRE(steps >= 1, SGMC_E_ARG, "sgmc_glm_fused_loop: steps %d", steps);
  SGMC_REQUIRE(pot_temperature != 0.0f, SGMC_E_ARG, "sgmc_glm_fused_loop: temperature is zero");
  if (iterations == 0) return SGMC_OK;
  FusedArgs a;
  a.chains = chains;
  a.block_chain = nullptr;
  a.blocks_per_chain = 1;
  a.eps_table = eps_table;
  a.save_slot = save_slot;
  a.iter = iter;
  a.done = done;
  a.inv_mass = inv_mass;
  a.sqrt_mass = sqrt_mass;
  a.table_len = table_len;
  a.nchains = nchains;
  a.mode = mode;
  a.step = 0;
  a.steps = (mode == SGMC_MODE_SGHMC || mode == SGMC_MODE_OBABO) ? steps : 1;
  a.friction = friction;
  a.four_friction = (float)(4.0 * (double)friction);
  a.alpha = alpha;
  a.lmbd = lmbd;
  a.L = make_lik(family, sigma);
  a.pr = make_prior(has_prior, prior_scale);
  a.pot_temperature = pot_temperature;
  a.advance = -1;
  a.iter_offset = 0;
  hipStream_t s = as_stream(stream);
  if (mode == SGMC_MODE_SGLD)
    launch_loop<SGMC_MODE_SGLD>(uniform_dclass, nchains, iterations, s, a);
  else
    launch_loop<-1>(uniform_dclass, nchains, iterations, s, a);
  return check_launch("sgmc_glm_fused_loop");
}

}  // extern "C"
