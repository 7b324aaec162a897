// Device-side building blocks shared by every kernel of the library (gfx950).
//
// Random numbers follow jax.random's original (non-partitionable) Threefry-2x32
// derivation, which is what the reference draws through (jax 0.4.x):
//   split(key, num)[i]   = words (2i, 2i+1) of threefry_2x32(key, iota(2 num))
//   random_bits(key, S)  = threefry_2x32(key, iota(S)); with m = ceil(S/2), word i
//                          comes from counter pair (j, j+m) (j = i mod m; the pair
//                          (m-1, 0) when S is odd), output 0 if i < m else output 1
//   uniform              = bitcast((bits >> 9) | 0x3F800000) - 1, scaled
//   normal               = sqrt(2) * erfinv(uniform on [nextafter(-1,0), 1))
//   randint(key, n, lo, hi): k1, k2 = split(key); ((hi%s)*mult + lo%s) % s
// Reference call sites: integrator.py:119-135 (random_tree), integrator.py:208,
// 630, 736, 871 (key splits), data/numpy_loader.py:132-134 (index draw).
//
// The float32 elementary functions (log, log1p, erfinv) are written out op by op
// with floating-point contraction disabled, so a normal is a fixed function of its
// 32 input bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgmc {

constexpr int kWave = 64;

// Global-address-space pointers. Pointers read from device memory (a chain table)
// are generic to the compiler and become flat loads, which also count against the
// LDS/scalar wait counter and so serialize with every LDS or DPP wait; casting them
// to address space 1 yields global loads.
#define SGMC_GLOBAL __attribute__((address_space(1)))
typedef SGMC_GLOBAL float gfloat;
typedef SGMC_GLOBAL const float gcfloat;
typedef SGMC_GLOBAL uint32_t guint;
typedef SGMC_GLOBAL unsigned gunsigned;

template <class T>
__device__ __forceinline__ SGMC_GLOBAL T* to_global(T* p) {
  return (SGMC_GLOBAL T*)p;
}

struct Key {
  uint32_t k0, k1;
};

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) {
  return (x << r) | (x >> (32 - r));
}

// Threefry-2x32, 20 rounds, on one counter pair (x0, x1).
__device__ __forceinline__ void threefry2x32(uint32_t k0, uint32_t k1, uint32_t& x0,
                                             uint32_t& x1) {
  const uint32_t k2 = k0 ^ k1 ^ 0x1BD11BDAu;
  x0 += k0;
  x1 += k1;
#define SGMC_TF_ROUND(r) \
  x0 += x1;              \
  x1 = rotl32(x1, r);    \
  x1 ^= x0;
  SGMC_TF_ROUND(13) SGMC_TF_ROUND(15) SGMC_TF_ROUND(26) SGMC_TF_ROUND(6)
  x0 += k1; x1 += k2 + 1u;
  SGMC_TF_ROUND(17) SGMC_TF_ROUND(29) SGMC_TF_ROUND(16) SGMC_TF_ROUND(24)
  x0 += k2; x1 += k0 + 2u;
  SGMC_TF_ROUND(13) SGMC_TF_ROUND(15) SGMC_TF_ROUND(26) SGMC_TF_ROUND(6)
  x0 += k0; x1 += k1 + 3u;
  SGMC_TF_ROUND(17) SGMC_TF_ROUND(29) SGMC_TF_ROUND(16) SGMC_TF_ROUND(24)
  x0 += k1; x1 += k2 + 4u;
  SGMC_TF_ROUND(13) SGMC_TF_ROUND(15) SGMC_TF_ROUND(26) SGMC_TF_ROUND(6)
  x0 += k2; x1 += k0 + 5u;
#undef SGMC_TF_ROUND
}

// Word q (0 <= q < 2 num) of threefry_2x32(key, iota(2 num)).
__device__ __forceinline__ uint32_t split_word(Key k, uint32_t num, uint32_t q) {
  const uint32_t pair = q < num ? q : q - num;
  uint32_t x0 = pair, x1 = num + pair;
  threefry2x32(k.k0, k.k1, x0, x1);
  return q < num ? x0 : x1;
}

// split(key, num)[i].
__device__ __forceinline__ Key split_key(Key k, uint32_t num, uint32_t i) {
  if (num == 1) {
    uint32_t x0 = 0, x1 = 1;
    threefry2x32(k.k0, k.k1, x0, x1);
    return Key{x0, x1};
  }
  return Key{split_word(k, num, 2 * i), split_word(k, num, 2 * i + 1)};
}

// (split(key)[0], split(key)[1]) with two Threefry blocks.
__device__ __forceinline__ void split2(Key k, Key& a, Key& b) {
  uint32_t p0 = 0, q0 = 2, p1 = 1, q1 = 3;
  threefry2x32(k.k0, k.k1, p0, q0);
  threefry2x32(k.k0, k.k1, p1, q1);
  a = Key{p0, p1};
  b = Key{q0, q1};
}

// split2 for a key that is the same on every lane: lanes 0 and 1 each run one of the two
// Threefry blocks and the words are broadcast, so the split costs one block of latency.
// Every lane of the wavefront must be active.
__device__ __forceinline__ void split2_uniform(Key k, Key& a, Key& b) {
  const uint32_t l = __lane_id() & 1u;
  uint32_t p = l, q = l + 2u;
  threefry2x32(k.k0, k.k1, p, q);
  a = Key{(uint32_t)__builtin_amdgcn_readlane((int)p, 0), (uint32_t)__builtin_amdgcn_readlane((int)p, 1)};
  b = Key{(uint32_t)__builtin_amdgcn_readlane((int)q, 0), (uint32_t)__builtin_amdgcn_readlane((int)q, 1)};
}

// split(key, 3) with three Threefry blocks.
__device__ __forceinline__ void split3(Key k, Key& a, Key& b, Key& c) {
  uint32_t a0 = 0, b0 = 3, a1 = 1, b1 = 4, a2 = 2, b2 = 5;
  threefry2x32(k.k0, k.k1, a0, b0);
  threefry2x32(k.k0, k.k1, a1, b1);
  threefry2x32(k.k0, k.k1, a2, b2);
  a = Key{a0, a1};
  b = Key{a2, b0};
  c = Key{b1, b2};
}

// Counter pair j (0 <= j < m = ceil(S/2)) of random_bits(key, S): returns the two
// words that land at positions j and j + m.
__device__ __forceinline__ void bits_pair(Key k, uint32_t S, uint32_t j, uint32_t& w_lo,
                                          uint32_t& w_hi) {
  const uint32_t m = (S + 1) >> 1;
  uint32_t x0 = j, x1 = (j + m < S) ? j + m : 0u;
  threefry2x32(k.k0, k.k1, x0, x1);
  w_lo = x0;
  w_hi = x1;
}

// Word i of random_bits(key, S).
__device__ __forceinline__ uint32_t bits_word(Key k, uint32_t S, uint32_t i) {
  const uint32_t m = (S + 1) >> 1;
  const uint32_t j = i < m ? i : i - m;
  uint32_t a, b;
  bits_pair(k, S, j, a, b);
  return i < m ? a : b;
}

__device__ __forceinline__ float unit_float(uint32_t bits) {
  return __uint_as_float((bits >> 9) | 0x3F800000u) - 1.0f;
}

// ---- float32 elementary functions (contraction off) ---------------------------

// Cephes logf for positive normal x.
__device__ __forceinline__ float logf_cephes(float x) {
#pragma clang fp contract(off)
  const uint32_t xb = __float_as_uint(x);
  float e = (float)((int)(xb >> 23) - 126);
  const float m = __uint_as_float((xb & 0x807FFFFFu) | 0x3F000000u);
  float t = m - 1.0f;
  if (m < 0.707106781186547524f) {
    t = t + m;
    e = e - 1.0f;
  }
  const float x2 = t * t;
  const float x3 = x2 * t;
  float y = 7.0376836292e-2f * t + -1.1514610310e-1f;
  float y1 = -1.2420140846e-1f * t + 1.4249322787e-1f;
  float y2 = 2.0000714765e-1f * t + -2.4999993993e-1f;
  y = y * t + 1.1676998740e-1f;
  y1 = y1 * t + -1.6668057665e-1f;
  y2 = y2 * t + 3.3333331174e-1f;
  y = y * x3 + y1;
  y = y * x3 + y2;
  y = y * x3;
  y = y + -2.12194440e-4f * e;
  float r = t - 0.5f * x2;
  r = r + y;
  r = r + 0.693359375f * e;
  return r;
}

// log1p: Cephes rational approximation for |x| < sqrt(2) - 1, else log(1 + x).
__device__ __forceinline__ float log1p_cephes(float x) {
#pragma clang fp contract(off)
  if (!(fabsf(x) < 0.41421356237309504880f)) return logf_cephes(x + 1.0f);
  const float x2 = x * x;
  float num = 0.0f;
  num = num * x + 4.5270000862445199635215e-5f;
  num = num * x + 4.9854102823193375972212e-1f;
  num = num * x + 6.5787325942061044846969e0f;
  num = num * x + 2.9911919328553073277375e1f;
  num = num * x + 6.0949667980987787057556e1f;
  num = num * x + 5.7112963590585538103336e1f;
  num = num * x + 2.0039553499201281259648e1f;
  float den = 0.0f;
  den = den * x + 1.0f;
  den = den * x + 1.5062909083469192043167e1f;
  den = den * x + 8.3047565967967209469434e1f;
  den = den * x + 2.2176239823732856465394e2f;
  den = den * x + 3.0909872225312059774938e2f;
  den = den * x + 2.1642788614495947685003e2f;
  den = den * x + 6.0118660497603843919306e1f;
  const float ratio = num / den;
  return x + ((-0.5f * x2) + ((x * x2) * ratio));
}

// erfinv (Giles 2010 single-precision approximation).
__device__ __forceinline__ float erfinv_giles(float x) {
#pragma clang fp contract(off)
  float w = -log1p_cephes(x * (-x));
  float p;
  if (w < 5.0f) {
    w = w - 2.5f;
    p = 2.81022636e-08f;
    p = 3.43273939e-07f + p * w;
    p = -3.5233877e-06f + p * w;
    p = -4.39150654e-06f + p * w;
    p = 0.00021858087f + p * w;
    p = -0.00125372503f + p * w;
    p = -0.00417768164f + p * w;
    p = 0.246640727f + p * w;
    p = 1.50140941f + p * w;
  } else {
    w = sqrtf(w) - 3.0f;
    p = -0.000200214257f;
    p = 0.000100950558f + p * w;
    p = 0.00134934322f + p * w;
    p = -0.00367342844f + p * w;
    p = 0.00573950773f + p * w;
    p = -0.0076224613f + p * w;
    p = 0.00943887047f + p * w;
    p = 1.00167406f + p * w;
    p = 2.83297682f + p * w;
  }
  if (fabsf(x) == 1.0f) return x * __builtin_inff();
  return p * x;
}

// Standard normal from one 32-bit word (jax.random.normal, float32).
__device__ __forceinline__ float bits_to_normal(uint32_t bits) {
#pragma clang fp contract(off)
  const float lo = -0.99999994f;  // nextafter(-1, 0)
  const float f = unit_float(bits);
  float u = f * 2.0f + lo;        // (1 - lo) rounds to 2 in float32
  u = fmaxf(lo, u);
  return 1.41421354f * erfinv_giles(u);
}

// Standard normal for the MLP update epilogues, from the same word as bits_to_normal:
// the same Giles polynomial, with w = -log1p(-x^2) taken as -ln2 log2(1 - x^2) on the
// hardware log (x^2 rounded first, as log1p(-x * x) rounds it; 1 - x^2 is then exact for
// x^2 >= 1/2 and within 3e-8 below) and the polynomial in fused multiply-adds. |x| < 1
// always (the uniform lies in [-0.99999994, 0.99999982]). Within 3 ulp of
// bits_to_normal over 8.4M words, about a fifth of its instructions (no division, no
// divergent log branches).
__device__ __forceinline__ float bits_to_normal_fast(uint32_t bits) {
  const float lo = -0.99999994f;
  const float f = unit_float(bits);
  const float x = fmaxf(lo, f * 2.0f + lo);
  const float x2 = x * x;
  const float w = -0.693147182f * __builtin_amdgcn_logf(1.0f - x2);
  // central branch (w < 5) on every lane; the tail (|x| > 0.9966, about 1 lane in 300)
  // only when a lane of the wavefront needs it, a wave-uniform branch instead of both
  // polynomials on every lane
  const float wc = w - 2.5f;
  float p = 2.81022636e-08f;
  p = __builtin_fmaf(p, wc, 3.43273939e-07f);
  p = __builtin_fmaf(p, wc, -3.5233877e-06f);
  p = __builtin_fmaf(p, wc, -4.39150654e-06f);
  p = __builtin_fmaf(p, wc, 0.00021858087f);
  p = __builtin_fmaf(p, wc, -0.00125372503f);
  p = __builtin_fmaf(p, wc, -0.00417768164f);
  p = __builtin_fmaf(p, wc, 0.246640727f);
  p = __builtin_fmaf(p, wc, 1.50140941f);
  if (__builtin_expect(__any(w >= 5.0f), 0)) {
    const float wt = __builtin_amdgcn_sqrtf(w) - 3.0f;
    float pt = -0.000200214257f;
    pt = __builtin_fmaf(pt, wt, 0.000100950558f);
    pt = __builtin_fmaf(pt, wt, 0.00134934322f);
    pt = __builtin_fmaf(pt, wt, -0.00367342844f);
    pt = __builtin_fmaf(pt, wt, 0.00573950773f);
    pt = __builtin_fmaf(pt, wt, -0.0076224613f);
    pt = __builtin_fmaf(pt, wt, 0.00943887047f);
    pt = __builtin_fmaf(pt, wt, 1.00167406f);
    pt = __builtin_fmaf(pt, wt, 2.83297682f);
    p = w < 5.0f ? p : pt;
  }
  return 1.41421354f * (p * x);
}

// jax.random.uniform(key, (), lo, hi) from one word.
__device__ __forceinline__ float bits_to_uniform(uint32_t bits, float lo, float hi) {
#pragma clang fp contract(off)
  const float f = unit_float(bits);
  return fmaxf(lo, f * (hi - lo) + lo);
}

// randint offset: ((hi % span) * mult + lo % span) % span in wrapping uint32.
__host__ __device__ __forceinline__ uint32_t randint_mult(uint32_t span) {
  uint32_t m = 65536u % span;
  return (uint32_t)(m * m) % span;
}

// a % d without a division: magic = floor(2^32 / d) for d >= 2 (0 for d = 1). Since
// magic > 2^32 / d - 1, q' = umulhi(a, magic) is q or q - 1 (q = floor(a / d)), so
// r' = a - q' d lies in [0, 2d) (and r' <= a < 2^32): one conditional subtraction of d
// gives a % d.
__host__ __device__ __forceinline__ uint32_t randint_magic(uint32_t span) {
  return span >= 2u ? (uint32_t)(0x100000000ull / span) : 0u;
}
__device__ __forceinline__ uint32_t mod_magic(uint32_t a, uint32_t d, uint32_t magic) {
  if (d < 2u) return 0u;
  const uint32_t r = a - __umulhi(a, magic) * d;
  return r >= d ? r - d : r;
}

__device__ __forceinline__ uint32_t randint_offset(uint32_t hi, uint32_t lo, uint32_t span,
                                                   uint32_t mult, uint32_t magic) {
  return mod_magic(mod_magic(hi, span, magic) * mult + mod_magic(lo, span, magic), span, magic);
}

// Index j (0 <= j < n) of randint(sub, (n,), 0, N) given k1, k2 = split(sub); mult and
// magic: randint_mult(N), randint_magic(N).
__device__ __forceinline__ int32_t draw_index(Key k1, Key k2, uint32_t n, uint32_t j,
                                              uint32_t span, uint32_t mult, uint32_t magic) {
  const uint32_t hi = bits_word(k1, n, j);
  const uint32_t lo = bits_word(k2, n, j);
  return (int32_t)randint_offset(hi, lo, span, mult, magic);
}

__device__ __forceinline__ Key load_key(const uint32_t* p) { return Key{p[0], p[1]}; }
__device__ __forceinline__ void store_key(uint32_t* p, Key k) {
  p[0] = k.k0;
  p[1] = k.k1;
}

// ---- wavefront / block reductions ---------------------------------------------

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over a block of NT threads (NT a multiple of 64); every thread gets the total.
// Workgroup barrier for LDS hand-offs only: waits for this wavefront's LDS (and scalar)
// accesses, not for its global loads, which stay in flight across the barrier (their
// registers are waited on where they are first used).
__device__ __forceinline__ void lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// LDS_SYNC: barriers that leave global loads in flight (lds_sync) instead of
// __syncthreads; for blocks that share only LDS between the two barriers.
template <int NT, bool LDS_SYNC = false>
__device__ __forceinline__ float block_sum(float v, float* lds /* NT/64 floats */) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave;
  if (LDS_SYNC) lds_sync();
  else __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) lds[w] = v;
  if (LDS_SYNC) lds_sync();
  else __syncthreads();
  float t = 0.0f;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) t += lds[i];
  return t;
}

template <int NT>
__device__ __forceinline__ double block_sum_d(double v, double* lds) {
  v = wave_sum_d(v);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) t += lds[i];
  return t;
}

// ---- block-to-block hand-off inside one launch ----
// Values that blocks write for the last-arriving block of the same launch go through
// agent-scope relaxed atomic stores and loads (publish / consume). The L2 keeps such
// accesses coherent across the XCDs by itself (the stores write through, the loads do
// not hit lines another XCD may have changed), so the hand-off needs none of the
// whole-cache L2 writeback and invalidate that a device-scope fence performs.
__device__ __forceinline__ void publish(SGMC_GLOBAL float* p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish(float* p, float v) { publish(to_global(p), v); }
__device__ __forceinline__ float consume(const SGMC_GLOBAL float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float consume(const float* p) { return consume(to_global(p)); }
__device__ __forceinline__ void publish(double* p, double v) {
  __hip_atomic_store(to_global(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double consume(const double* p) {
  return __hip_atomic_load(to_global(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// s0 = sum over b in [0, nb) of p0[b * stride], s1 likewise for p1, each added in block
// order, from published records. The loads go out eight records at a time (missing
// records read as +0, which leaves the sums unchanged), so one memory latency covers
// eight records.
__device__ __forceinline__ void sum_strided2(const SGMC_GLOBAL float* p0,
                                             const SGMC_GLOBAL float* p1, int64_t stride,
                                             int nb, float& s0, float& s1) {
  float a0 = 0.0f, a1 = 0.0f;
  for (int b = 0; b < nb; b += 8) {
    float x0[8], x1[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const bool in = b + t < nb;
      x0[t] = in ? consume(p0 + (int64_t)(b + t) * stride) : 0.0f;
      x1[t] = in ? consume(p1 + (int64_t)(b + t) * stride) : 0.0f;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      a0 += x0[t];
      a1 += x1[t];
    }
  }
  s0 = a0;
  s1 = a1;
}

// Arrival counter for "last block finishes" reductions: every block calls this after
// publishing its partial results; exactly one (the last to arrive) gets true and can
// consume all of them. The wait makes every published store of the block complete
// before the block's arrival is counted. The last block resets the counter for the
// next launch.
__device__ __forceinline__ bool arrive_last(unsigned* counter, unsigned expected,
                                            int* lds_flag) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);  // vmcnt(0) expcnt(0) lgkmcnt(0)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool last = (prev == expected - 1);
    if (last) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *lds_flag = last ? 1 : 0;
  }
  __syncthreads();
  return *lds_flag != 0;
}

}  // namespace sgmc
