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

#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {

constexpr int kNT = 256;
constexpr int kWaves = kNT / kWave;
constexpr int kMaxD = 1024;  // fused path: features per row
constexpr int kUnroll = 2;   // row groups in flight per wavefront

struct Lik {
  int family;
  float inv_s2;    // 1 / sigma^2 (linear-Gaussian)
  float log_norm;  // log(2 pi sigma^2)
};

struct Prior {
  int has_prior;
  float inv_var;   // 1 / s_p^2
  float log_norm;  // log(2 pi s_p^2)
};

__device__ __forceinline__ void row_lik(const Lik& L, float z, float y, float& lik, float& r) {
  if (L.family == SGMC_LIK_LINEAR_GAUSSIAN) {
    const float d = y - z;
    lik = (L.log_norm + (d * d) * L.inv_s2) / -2.0f;
    r = d * L.inv_s2;
  } else {
    const float az = fabsf(z);
    const float e = expf(-az);
    const float sp = fmaxf(z, 0.0f) + log1pf(e);
    lik = y * z - sp;
    const float inv1pe = 1.0f / (1.0f + e);
    const float sig = z >= 0.0f ? inv1pe : e * inv1pe;
    r = y - sig;
  }
}

__device__ __forceinline__ float dot4(float4 a, float4 b) {
  return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}

__device__ __forceinline__ void axpy4(float4& acc, float r, float4 x) {
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
  uint32_t n, span, mult;
  int64_t rows;        // rows addressable through X (bounds guard)
  // Row index of minibatch row j, or -1 if it falls outside [0, rows).
  __device__ __forceinline__ int64_t row(int j) const {
    int64_t r;
    if (draw) r = (int64_t)draw_index(k1, k2, n, (uint32_t)j, span, mult);
    else if (idx) r = (int64_t)idx[j];
    else r = (int64_t)j;
    return (r >= 0 && r < rows) ? r : -1;
  }
};

// Accumulates sum_j r_j x_j (this lane's chunks) and sum_j lik_j over rows
// [r0, r1) of the minibatch. th holds the lane's chunks of w.
template <int LPR, int CPL>
__device__ __forceinline__ void accumulate_rows(const float* __restrict__ X, int64_t ldx,
                                                const float* __restrict__ y, int D4,
                                                const RowSource& src, int r0, int r1,
                                                const float4 (&th)[CPL], const Lik& L,
                                                float4 (&acc)[CPL], float& lik_acc,
                                                float* __restrict__ lik_out) {
  constexpr int RPW = kWave / LPR;
  const int lane = threadIdx.x & (kWave - 1);
  const int seg = lane / LPR;
  const int q = lane % LPR;
  for (int cb = r0; cb < r1; cb += kWave) {
    const int nrows = min(kWave, r1 - cb);
    int64_t my_row = -1;
    float my_y = 0.0f;
    if (lane < nrows) {
      my_row = src.row(cb + lane);
      if (my_row >= 0) my_y = y[my_row];
    }
    for (int g = 0; g < nrows; g += RPW * kUnroll) {
      float4 xv[kUnroll][CPL];
      float yv[kUnroll];
      bool ok[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int rl = g + u * RPW + seg;
        const int src_lane = rl < nrows ? rl : 0;
        const int64_t rid = __shfl((long long)my_row, src_lane, kWave);
        ok[u] = rl < nrows && rid >= 0;
        yv[u] = __shfl(my_y, src_lane, kWave);
        const float4* rp = reinterpret_cast<const float4*>(X + (ok[u] ? rid : 0) * ldx);
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          const int c4 = q + LPR * k;
          xv[u][k] = (ok[u] && c4 < D4) ? rp[c4] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        float z = 0.0f;
#pragma unroll
        for (int k = 0; k < CPL; ++k) z += dot4(xv[u][k], th[k]);
#pragma unroll
        for (int m = 1; m < LPR; m <<= 1) z += __shfl_xor(z, m, kWave);
        float lik, r;
        row_lik(L, z, yv[u], lik, r);
        if (!ok[u]) {
          lik = 0.0f;
          r = 0.0f;
        }
#pragma unroll
        for (int k = 0; k < CPL; ++k) axpy4(acc[k], r, xv[u][k]);
        if (q == 0 && ok[u]) {
          lik_acc += lik;
          if (lik_out) lik_out[cb + g + u * RPW + seg] = lik;
        }
      }
    }
  }
}

// Reduces acc over the segments of each wavefront and over the block's wavefronts;
// writes the block's column sums to out[0..4*D4) and returns the block's lik sum.
template <int LPR, int CPL>
__device__ __forceinline__ float block_reduce_write(float4 (&acc)[CPL], float lik_acc, int D4,
                                                    float* __restrict__ out, float* lds_cols,
                                                    float* lds_small) {
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
  __syncthreads();
  if (lane < LPR) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c4 = q + LPR * k;
      if (c4 < D4) reinterpret_cast<float4*>(lds_cols + w * 4 * D4)[c4] = acc[k];
    }
  }
  const float lsum = block_sum<kNT>(lik_acc, lds_small);  // contains __syncthreads
  for (int c = threadIdx.x; c < 4 * D4; c += kNT) {
    float s = 0.0f;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) s += lds_cols[ww * 4 * D4 + c];
    out[c] = s;
  }
  return lsum;
}

// Splits rows [0, n) of a minibatch among the blocks of a chain and then among the
// wavefronts of a block (contiguous ranges).
__device__ __forceinline__ void wave_rows(int n, int nblocks, int b, int& r0, int& r1) {
  const int per_block = (n + nblocks - 1) / nblocks;
  const int b0 = min(n, b * per_block);
  const int b1 = min(n, b0 + per_block);
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
  float* partial;      // [nblocks][4*D4 + 1]
  float* lik_scratch;  // [n]
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
  __shared__ int flag;
  float4 th[CPL];
  load_theta<LPR, CPL>(a.theta, a.D, th);
  float4 acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  float lik_acc = 0.0f;
  int r0, r1;
  wave_rows(a.n, gridDim.x, blockIdx.x, r0, r1);
  RowSource src{a.idx, 0, Key{0, 0}, Key{0, 0}, 0, 0, 0, a.idx ? a.N : (int64_t)a.n};
  accumulate_rows<LPR, CPL>(a.X, a.ldx, a.y, a.D4, src, r0, r1, th, a.L, acc, lik_acc,
                            a.lik_scratch);
  const int stride = 4 * a.D4 + 1;
  float* mine = a.partial + (int64_t)blockIdx.x * stride;
  const float ls = block_reduce_write<LPR, CPL>(acc, lik_acc, a.D4, mine, lds_cols, lds_small);
  if (threadIdx.x == 0) mine[4 * a.D4] = ls;
  if (!arrive_last(a.counter, gridDim.x, &flag)) return;

  // Last block: combine partials in block order, add the prior, scale.
  const float c = (float)(-a.N) / (float)a.n;
  float sumlik = 0.0f;
  for (int b = 0; b < (int)gridDim.x; ++b) sumlik += a.partial[(int64_t)b * stride + 4 * a.D4];
  float lp = 0.0f;
  for (int j = threadIdx.x; j < a.D; j += kNT) {
    float s = 0.0f;
    for (int b = 0; b < (int)gridDim.x; ++b) s += a.partial[(int64_t)b * stride + j];
    const float wj = a.theta[j];
    float g = c * s;
    if (a.pr.has_prior) {
      g = g + wj * a.pr.inv_var;
      lp += (a.pr.log_norm + (wj * wj) * a.pr.inv_var) / -2.0f;
    }
    a.out_grad[j] = g / a.temperature;
  }
  lp = block_sum<kNT>(lp, lds_small);
  const float mean = sumlik / (float)a.n;
  if (threadIdx.x == 0) a.out_U[0] = ((float)(-a.N) * mean - lp) / a.temperature;
  if (a.out_lik || a.out_var) {
    float vs = 0.0f;
    for (int j = threadIdx.x; j < a.n; j += kNT) {
      const float l = a.lik_scratch[j];
      if (a.out_lik) a.out_lik[j] = l;
      const float d = l - mean;
      vs += d * d;
    }
    vs = block_sum<kNT>(vs, lds_small);
    if (a.out_var && threadIdx.x == 0) a.out_var[0] = vs / (float)a.n;
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
  if (threadIdx.x == 0) a.partial[blockIdx.x] = bs;
  if (!arrive_last(a.counter, gridDim.x, &flag)) return;
  double tot = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kNT) tot += a.partial[b];
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
  const int32_t* block_chain;
  const float* eps_table;
  const int32_t* save_slot;
  int32_t* iter;
  unsigned* done;
  const float* inv_mass;   // shared diagonal mass (nullable = ones)
  const float* sqrt_mass;
  int table_len;
  int nchains;
  int mode;                // SGMC_MODE_*
  int step;                // leapfrog step within the iteration (SGHMC)
  int steps;
  float friction;
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

template <int LPR, int CPL>
__device__ __forceinline__ void fused_body(const FusedArgs& a, const sgmc_glm_chain& ch,
                                           int bchain, float* lds_cols, float* lds_small,
                                           int* flag) {
  const int it = max(0, min(a.iter[0], a.table_len - 1));
  const float eps = a.eps_table[it];
  const int D = ch.D;
  const int D4 = (D + 3) >> 2;
  const Key key0 = load_key(ch.key);
  const bool hmc = a.mode == SGMC_MODE_SGHMC;
  const bool resample = hmc && a.step == 0;

  // Momentum used by this step: resampled p = sqrt_mass * xi at the first step.
  Key key_noise = key0;  // key split for this step's noise
  Key resample_leaf{0, 0};
  if (resample) {
    Key k1, sub;
    split2(key0, k1, sub);
    resample_leaf = split_key(sub, 1, 0);
    key_noise = k1;
  }

  // Evaluation point: theta (SGLD) or theta + eps M^-1 p (SGHMC).
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
          v = ch.theta[c];
          if (hmc) {
            float p;
            if (resample) {
              const float xi = leaf_normal(resample_leaf, (uint32_t)D, (uint32_t)c);
              p = a.sqrt_mass ? a.sqrt_mass[c] * xi : xi;
            } else {
              p = ch.mom[c];
            }
            const float s = a.inv_mass ? a.inv_mass[c] * p : p;
            v = v + eps * s;
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
  src.rows = ch.N;

  float4 acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  float lik_acc = 0.0f;
  int r0, r1;
  wave_rows(ch.n, ch.nblocks, bchain, r0, r1);
  accumulate_rows<LPR, CPL>(ch.X, ch.ldx, ch.y, D4, src, r0, r1, th, a.L, acc, lik_acc,
                            ch.lik);
  const int stride = 4 * D4 + 1;
  float* mine = ch.partial + (int64_t)bchain * stride;
  const float ls = block_reduce_write<LPR, CPL>(acc, lik_acc, D4, mine, lds_cols, lds_small);
  if (threadIdx.x == 0) mine[4 * D4] = ls;
  if (!arrive_last(ch.counter, ch.nblocks, flag)) return;

  // ---- last block of this chain: potential, gradient and the state update ----
  const float c = (float)(-ch.N) / (float)ch.n;
  float sumlik = 0.0f;
  for (int b = 0; b < ch.nblocks; ++b) sumlik += ch.partial[(int64_t)b * stride + 4 * D4];
  const float mean = sumlik / (float)ch.n;

  // Noise key of this step.
  Key knext, ksub;
  split2(key_noise, knext, ksub);
  const Key leaf = split_key(ksub, 1, 0);
  const int slot = a.save_slot ? a.save_slot[it] : -1;
  const bool last_step = !hmc || a.step == a.steps - 1;
  const float Tn = ch.temperature;

  float lp = 0.0f;
  const uint32_t S = (uint32_t)D;
  const uint32_t m = (S + 1) >> 1;
  for (uint32_t jp = threadIdx.x; jp < m; jp += kNT) {
    uint32_t w0, w1;
    bits_pair(leaf, S, jp, w0, w1);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const uint32_t j = half ? jp + m : jp;
      if (half && j >= S) break;
      const float xi = bits_to_normal(half ? w1 : w0);
      float s = 0.0f;
      for (int b = 0; b < ch.nblocks; ++b) s += ch.partial[(int64_t)b * stride + j];
      float wj = ch.theta[j];
      float p = 0.0f, sm = 0.0f;
      if (hmc) {
        if (resample) {
          const float xr = leaf_normal(resample_leaf, S, j);
          p = a.sqrt_mass ? a.sqrt_mass[j] * xr : xr;
        } else {
          p = ch.mom[j];
        }
        sm = a.inv_mass ? a.inv_mass[j] * p : p;
        wj = wj + eps * sm;  // evaluation point
      }
      float g = c * s;
      if (a.pr.has_prior) {
        g = g + wj * a.pr.inv_var;
        lp += (a.pr.log_norm + (wj * wj) * a.pr.inv_var) / -2.0f;
      }
      g = g / a.pot_temperature;
      float wn;
      if (!hmc) {
        const float sg = (-eps) * g;
        const float sn = sqrtf((2.0f * Tn) * eps) * xi;
        if (a.mode == SGMC_MODE_SGLD_RMS) {
          const float vn = a.alpha * ch.v[j] + (1.0f - a.alpha) * (g * g);
          ch.v[j] = vn;
          const float G = 1.0f / (a.lmbd + sqrtf(vn));
          wn = wj + ((0.0f + G * sg) + sqrtf(G) * sn);
        } else {
          wn = wj + (sg + sn);
        }
      } else {
        const float C = a.friction;
        float pn = p + ((-eps) * C) * sm;
        pn = pn + (-eps) * g;
        pn = pn + C * (sqrtf(2.0f * eps) * xi);
        ch.mom[j] = pn;
        wn = wj;
      }
      ch.theta[j] = wn;
      if (slot >= 0 && last_step && ch.samples) ch.samples[(int64_t)slot * D + j] = wn;
    }
  }
  lp = block_sum<kNT>(lp, lds_small);
  float vs = 0.0f;
  for (int j = threadIdx.x; j < ch.n; j += kNT) {
    const float d = ch.lik[j] - mean;
    vs += d * d;
  }
  vs = block_sum<kNT>(vs, lds_small);
  if (threadIdx.x == 0) {
    const float Uc = ((float)(-ch.N) * mean - lp) / a.pot_temperature;
    ch.U[0] = Uc;
    if (slot >= 0 && last_step && ch.samples_U) ch.samples_U[slot] = Uc;
    if (ch.var) ch.var[0] = vs / (float)ch.n;
    store_key(ch.key, knext);
    store_key(ch.dkey, dnext);
    // Last chain of the launch advances the iteration counter.
    __threadfence();
    const unsigned prev = atomicAdd(a.done, 1u);
    if (prev == (unsigned)a.nchains - 1) {
      atomicExch(a.done, 0u);
      if (last_step) atomicAdd(a.iter, 1);
    }
  }
}

template <int DCLASS>
__global__ void __launch_bounds__(kNT) glm_fused_kernel(FusedArgs a) {
  __shared__ __attribute__((aligned(16))) float lds_cols[kWaves * kMaxD];
  __shared__ float lds_small[kWaves];
  __shared__ int flag;
  const int chain = a.block_chain[blockIdx.x];
  if (chain < 0 || chain >= a.nchains) return;
  const sgmc_glm_chain& ch = a.chains[chain];
  const int bchain = blockIdx.x - ch.block_begin;
  if (bchain < 0 || bchain >= ch.nblocks) return;
  const int dc = DCLASS >= 0 ? DCLASS : ch.dclass;
  switch (dc) {
    case 0: fused_body<4, 1>(a, ch, bchain, lds_cols, lds_small, &flag); break;
    case 1: fused_body<16, 1>(a, ch, bchain, lds_cols, lds_small, &flag); break;
    case 2: fused_body<64, 1>(a, ch, bchain, lds_cols, lds_small, &flag); break;
    case 3: fused_body<64, 2>(a, ch, bchain, lds_cols, lds_small, &flag); break;
    default: fused_body<64, 4>(a, ch, bchain, lds_cols, lds_small, &flag); break;
  }
}

static int dclass_of(int D) {
  if (D <= 16) return 0;
  if (D <= 64) return 1;
  if (D <= 256) return 2;
  if (D <= 512) return 3;
  return 4;
}

static Lik make_lik(int family, float sigma) {
  Lik L;
  L.family = family;
  const float s2 = sigma * sigma;
  L.inv_s2 = 1.0f / s2;
  L.log_norm = logf(2.0f * (float)M_PI * s2);
  return L;
}

static Prior make_prior(int has_prior, float scale) {
  Prior p;
  p.has_prior = has_prior;
  const float s2 = scale * scale;
  p.inv_var = has_prior ? 1.0f / s2 : 0.0f;
  p.log_norm = has_prior ? logf(2.0f * (float)M_PI * s2) : 0.0f;
  return p;
}

}  // namespace sgmc

using namespace sgmc;

extern "C" {

int64_t sgmc_glm_workspace_floats(int n, int D, int nblocks) {
  const int D4 = (D + 3) / 4;
  return (int64_t)nblocks * (4 * D4 + 1) + (int64_t)n + 4;
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
  SGMC_REQUIRE(X && y && theta && out_U && out_grad && workspace, SGMC_E_ARG,
               "sgmc_glm_potential: null pointer");
  SGMC_REQUIRE(family == SGMC_LIK_LINEAR_GAUSSIAN || family == SGMC_LIK_LOGISTIC, SGMC_E_ARG,
               "sgmc_glm_potential: unknown family %d", family);
  SGMC_REQUIRE(D >= 1 && D <= kMaxD, SGMC_E_ARG, "sgmc_glm_potential: D=%d outside [1, %d]", D,
               kMaxD);
  SGMC_REQUIRE(ldx % 4 == 0 && ldx >= 4 * ((D + 3) / 4), SGMC_E_ARG,
               "sgmc_glm_potential: ldx=%lld must be a multiple of 4 and >= D rounded up",
               (long long)ldx);
  SGMC_REQUIRE(((uintptr_t)X & 15) == 0, SGMC_E_ARG, "sgmc_glm_potential: X not 16B aligned");
  SGMC_REQUIRE(n >= 1 && N >= 1, SGMC_E_ARG, "sgmc_glm_potential: empty batch");
  SGMC_REQUIRE(nblocks >= 1, SGMC_E_ARG, "sgmc_glm_potential: nblocks < 1");
  SGMC_REQUIRE(workspace_floats >= sgmc_glm_workspace_floats(n, D, nblocks), SGMC_E_ARG,
               "sgmc_glm_potential: workspace too small");
  SGMC_REQUIRE(temperature != 0.0f, SGMC_E_ARG, "sgmc_glm_potential: temperature is zero");
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
  a.partial = workspace;
  a.lik_scratch = workspace + (int64_t)nblocks * (4 * a.D4 + 1);
  a.counter = reinterpret_cast<unsigned*>(a.lik_scratch + n);
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
    case 2: hipLaunchKernelGGL((glm_potential_kernel<64, 1>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((glm_potential_kernel<64, 2>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((glm_potential_kernel<64, 4>), grid, block, 0, s, a); break;
  }
  return check_launch("sgmc_glm_potential");
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
    case 2: hipLaunchKernelGGL((glm_full_kernel<64, 1>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((glm_full_kernel<64, 2>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((glm_full_kernel<64, 4>), grid, block, 0, s, a); break;
  }
  return check_launch("sgmc_glm_full_potential");
}

int sgmc_glm_dclass(int D) { return dclass_of(D); }

int sgmc_glm_fused_step(const sgmc_glm_chain* chains, int nchains, const int32_t* block_chain,
                        int nblocks, int uniform_dclass, int family, float sigma, int has_prior,
                        float prior_scale, float pot_temperature, int mode, int step, int steps,
                        float friction, float alpha, float lmbd, const float* inv_mass,
                        const float* sqrt_mass, const float* eps_table, int table_len,
                        const int32_t* save_slot, int32_t* iter, unsigned* done, void* stream) {
  SGMC_REQUIRE(chains && block_chain && eps_table && iter && done, SGMC_E_ARG,
               "sgmc_glm_fused_step: null pointer");
  SGMC_REQUIRE(nchains >= 1 && nblocks >= nchains && table_len >= 1, SGMC_E_ARG,
               "sgmc_glm_fused_step: bad sizes");
  SGMC_REQUIRE(family == SGMC_LIK_LINEAR_GAUSSIAN || family == SGMC_LIK_LOGISTIC, SGMC_E_ARG,
               "sgmc_glm_fused_step: unknown family %d", family);
  SGMC_REQUIRE(mode == SGMC_MODE_SGLD || mode == SGMC_MODE_SGLD_RMS || mode == SGMC_MODE_SGHMC,
               SGMC_E_ARG, "sgmc_glm_fused_step: unknown mode %d", mode);
  SGMC_REQUIRE(steps >= 1 && step >= 0 && step < steps, SGMC_E_ARG,
               "sgmc_glm_fused_step: step %d of %d", step, steps);
  SGMC_REQUIRE(pot_temperature != 0.0f, SGMC_E_ARG, "sgmc_glm_fused_step: temperature is zero");
  FusedArgs a;
  a.chains = chains;
  a.block_chain = block_chain;
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
  a.alpha = alpha;
  a.lmbd = lmbd;
  a.L = make_lik(family, sigma);
  a.pr = make_prior(has_prior, prior_scale);
  a.pot_temperature = pot_temperature;
  hipStream_t s = as_stream(stream);
  dim3 grid(nblocks), block(kNT);
  switch (uniform_dclass) {
    case 0: hipLaunchKernelGGL((glm_fused_kernel<0>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((glm_fused_kernel<1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((glm_fused_kernel<2>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((glm_fused_kernel<3>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((glm_fused_kernel<4>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((glm_fused_kernel<-1>), grid, block, 0, s, a); break;
  }
  return check_launch("sgmc_glm_fused_step");
}

}  // extern "C"
