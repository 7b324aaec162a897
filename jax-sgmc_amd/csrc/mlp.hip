// Bayesian MLP classifier likelihood on MFMA (bf16 operands, fp32 accumulation), with
// a hand-written backward pass and the SGLD / rms_prop update fused into the weight-
// gradient GEMM epilogue. Model (the likelihood a user writes in the reference as
// likelihood(sample, obs), potential.py:128-137):
//   h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), logits = h2 W3 + b3,
//   lik = log_softmax(logits)[label]
// Parameters are one float32 vector per chain in tree_flatten order of
// {"W1","W2","W3","b1","b2","b3"} (sorted keys): W1 [D_in][H1], W2 [H1][H2],
// W3 [H2][10], b1, b2, b3. A bf16 shadow of the parameters (same layout) feeds the
// GEMMs and is rewritten by every update.
//
// One step for C chains (grid z = chain), potential U = (-N mean(lik) - log p)/T with
// a Gaussian prior, gradient by back-propagation:
//   gather   Xb = bf16(X[idx]) (index draw in-kernel, data/numpy_loader.py:132-134)
//   G1       H1 = relu(Xb W1 + b1)
//   G2       H2 = relu(H1 W2 + b2)
//   head     logits, softmax, lik, dL = c (onehot - p), dA2 = (dL W3^T) * (H2 > 0)
//   G3       dA1 = (dA2 W2^T) * (H1 > 0)
//   G4       gW2 = H1^T dA2  -> update W2 in the epilogue
//   G5       gW1 = Xb^T dA1  -> update W1 in the epilogue
//   small    W3, b1, b2, b3 gradients and updates; U; key and data-key advance
// c = -N / (n T) carries the potential's scaling into the backward pass.
//
// Precisions: bf16 GEMM operands (v_mfma_f32_32x32x16_bf16), OCP e4m3 with per-tensor
// scales (v_mfma_f32_32x32x16_fp8_fp8), or float32 (v_mfma_f32_32x32x2_f32: activations
// stored in float32, weights read from theta itself, no shadow); fp32 accumulation in
// all three.
//
// GEMM tiling (gfx950): a block computes 64 rows x 128
// columns; its 64 rows are two bands of 32 rows M/2 apart, so each lane's two
// accumulators hold the element pairs (i, i + size/2) that one Threefry block of
// random_bits serves (integrator.py:119-135 noise). Four wavefronts own 32 columns
// each. Operands are staged through LDS with k contiguous (padded rows); a source that
// is m- or n-contiguous is transposed while being written to LDS.
#include <math.h>

#include <type_traits>

#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {
namespace mlp {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// GEMM operand precision (template argument) and its element bytes.
enum Prec { P_BF16 = 0, P_FP8 = 1, P_F32 = 2 };
template <int PREC>
constexpr int elem_bytes() {
  return PREC == P_FP8 ? 1 : PREC == P_BF16 ? 2 : 4;
}

constexpr int GT = 256;
constexpr int BM = 64;
constexpr int BN = 128;
constexpr int NCLS = 10;
constexpr int NLEAF = 6;     // W1, W2, W3, b1, b2, b3
constexpr int kStrip = 32;   // rows per partial sum of the small leaves' gradients
constexpr int kXcds = 8;         // accelerator complex dies, each with its own L2
#ifndef SGMC_GATHER_BATCH
#define SGMC_GATHER_BATCH 12
#endif
constexpr int kGatherBatch = SGMC_GATHER_BATCH;  // float4 row loads in flight per lane
constexpr int kFinalBlocks = 64;
// Update epilogue tuning: parameters pairs per prefetch chunk (1, 2, 4 or 8 of a lane's
// 16 pairs) and the waves per SIMD the update GEMMs are compiled for (register budget).
#ifndef SGMC_UPD_CHUNK
#define SGMC_UPD_CHUNK 2
#endif
#ifndef SGMC_UPD_NT
#define SGMC_UPD_NT 0
#endif
#ifndef SGMC_UPD_WAVES
#define SGMC_UPD_WAVES 1
#endif
constexpr int kUpdChunk = SGMC_UPD_CHUNK;
static_assert(16 % kUpdChunk == 0, "chunks of the 16 pairs");
// The compiled-in rms_prop update at deep k-tiles (two waves per SIMD either way) has the
// registers for chunks of four pairs (pSGLD 3072-512-512 step 294.9 -> 292.7 us; the
// run-time rule of the OBABO updates loses a wave with them)
constexpr int kUpdChunkDeep = 4;
static_assert(16 % kUpdChunkDeep == 0, "chunks of the 16 pairs");

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// OCP fp8 e4m3 with a per-tensor scale: q = e4m3(clamp(v * scale, +-448)), v = q / scale.
__device__ __forceinline__ uint8_t f2q(float v) {
  v = fminf(fmaxf(v, -448.0f), 448.0f);
  return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(v, v, 0, false) & 0xFFu);
}
__device__ __forceinline__ float q2f(uint8_t b) { return __builtin_amdgcn_cvt_f32_fp8((int)b, 0); }

// fp8 tensors and their per-chain scale words: scales [0, 6), the scale the next weight
// shadows are written with [6, 8), amax bit patterns [8, 14).
enum Q8 { Q_X = 0, Q_H1 = 1, Q_DA2 = 2, Q_DA1 = 3, Q_W1 = 4, Q_W2 = 5 };
constexpr int kQWords = 16;
constexpr float kQTarget = 224.0f;  // scale = kQTarget / amax: half the e4m3 range as headroom

enum Epi { EPI_BIAS_RELU = 0, EPI_MASK = 1, EPI_UPDATE = 2, EPI_GRAD = 3, EPI_PARTIAL = 4 };

// Per-chain state and step parameters shared by the kernels.
struct Step {
  const float* X;          // dataset [N][ldx] float32
  const int32_t* labels;   // [N]
  int64_t N, ldx;
  int D_in, H1, H2, n, C;
  float* theta;            // [C][P]
  float* v;                // [C][P] or null (rms_prop)
  uint16_t* shadow;        // [C][P] bf16 (null for float32 GEMMs)
  int64_t P;
  int64_t off[NLEAF];      // leaf offsets
  int64_t size[NLEAF];     // leaf sizes
  uint32_t* key;           // [C][2]
  uint32_t* dkey;          // [C][2]
  uint32_t* lkey;          // [C][NLEAF][2] noise keys of this update, written by the gather
  uint32_t rmult, rmagic;  // randint_mult(N), randint_magic(N) of the index draw
  // activations: bf16; e4m3 except h2 (bf16) in fp8 mode; float32 in fp32 mode
  void* Xb;                // [C][n][D_in]
  int32_t* lab;            // [C][n]
  void* h1;                // [C][n][H1]
  void* h2;                // [C][n][H2]
  void* da2;               // [C][n][H2]
  void* da1;               // [C][n][H1]
  float* dL;               // [C][n][10]
  float* lik;              // [C][n]
  float* prior_part;       // [C][nparts]
  int nparts;
  float* U;                // [C]
  float* grad;             // [C][P] (gradient-only mode) or null
  float* samples;          // [C][n_slots][P] or null
  float* samples_U;        // [C][n_slots] or null
  int n_slots;
  const float* eps_table;
  const int32_t* save_slot;
  int32_t* iter;
  unsigned* counters;      // [C] arrival counters of the small kernel
  unsigned* done;
  int table_len;
  int rms;
  int update;              // 0: gradient only (autograd backward)
  float alpha, lmbd, temperature, noise_temperature;
  float prior_inv_var, prior_log_norm;
  int has_prior;
  // OBABO (update != 0, mode != MLP_SGLD)
  int mode;
  float* mom;              // [C][P] momentum
  float* ke_part;          // [C][nparts] kinetic-energy partials of one evaluation
  float* col_part;         // [C][n / kStrip][small leaves] row-strip gradient partials
  double* fin_part;        // [C][kFinalBlocks] log-prior partials of the full pass
  float* split_part;       // split-K partial products of the forward GEMMs
  double* ke;              // [C][2] accumulated kinetic energy at start / end
  float* U_half;           // [C] potential of the step's first evaluation
  const float* decay_table;  // [table_len] sqrt(exp(-friction eps))
  const float* noise_table;  // [table_len] sqrt((1 - exp(-friction eps)) T)
  // full-data forward pass (likelihoods of rows row0 .. row0 + n - 1 only)
  int full;
  int64_t row0;
  // fp8 GEMM operands (Xb, h1, da2, da1 and the W1 / W2 shadows in e4m3)
  int fp8;
  int fp32;                // float32 activations and GEMM operands (weights from theta)
  int plan[5];             // per GEMM (G1..G5): 16 k-tile + split-K factor
  float* q;                // [C][kQWords]
  uint8_t* shadow8;        // [C][P] (W1 and W2 leaves)
};

// Records max |v| of a wavefront for tensor t of chain c (every lane of the wavefront
// calls this). The atomic is skipped when the recorded maximum already covers the
// wavefront's: thousands of wavefronts updating one word otherwise serialize at the L2.
// A stale read is never above the true maximum, so no needed update is skipped.
__device__ __forceinline__ void q_note_amax(const Step& s, int c, int t, float a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) a = fmaxf(a, __shfl_xor(a, m, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0 && a > 0.0f) {
    unsigned* slot = reinterpret_cast<unsigned*>(s.q + (int64_t)c * kQWords + 8 + t);
    const unsigned cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__float_as_uint(a) > cur) atomicMax(slot, __float_as_uint(a));
  }
}

// Block-wide version (every thread of the block calls it): one check and at most one
// atomic per block, for kernels whose wavefronts are too short to carry their own.
__device__ __forceinline__ void q_note_amax_block(const Step& s, int c, int t, float a,
                                                  float* red) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) a = fmaxf(a, __shfl_xor(a, m, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = red[0];
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i) b = fmaxf(b, red[i]);
    if (b > 0.0f) {
      unsigned* slot = reinterpret_cast<unsigned*>(s.q + (int64_t)c * kQWords + 8 + t);
      const unsigned cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__float_as_uint(b) > cur) atomicMax(slot, __float_as_uint(b));
    }
  }
}

enum MlpMode { MLP_SGLD = 0, MLP_OBABO_BA = 1, MLP_OBABO_BO = 2 };

__device__ __forceinline__ int cur_iter(const Step& s) {
  return max(0, min(s.iter[0], s.table_len - 1));
}

// Gradient-only variant: grad = (gsum + prior) / T, no update.
__device__ __forceinline__ void grad_param(const Step& s, const float* th, float* gout,
                                           int64_t i, float gsum, float& lp) {
  const float w = th[i];
  float g = gsum;
  if (s.has_prior) {
    g = g + w * s.prior_inv_var;
    lp += (s.prior_log_norm + (w * w) * s.prior_inv_var) / -2.0f;
  }
  gout[i] = g / s.temperature;
}


// Parameter update of one leaf of chain c, applied to values so that a thread can
// issue the loads of all its parameters before the arithmetic. The leaf's arrays are
// addressed from uniform base pointers with 32-bit offsets (a leaf holds < 2^32
// entries). Update rules (prior gradient and temperature folded into g = (gsum + w /
// s^2) / T, written with fused multiply-adds: within an ulp or two of the reference's
// separate roundings, far inside the bf16 / e4m3 GEMM error):
//   SGLD (integrator.py:882-912): w += (-eps) g + sqrt(2 T eps) xi; with rms_prop
//     (adaption.py:254-291): v = alpha v + (1 - alpha) g^2, G = 1 / (lmbd + sqrt(v)),
//     w += G (-eps) g + sqrt(G) sqrt(2 T eps) xi (hardware reciprocal and square roots);
//   OBABO with the identity mass (integrator.py:181-273), in two halves:
//     BA: p1 = sqrt(pi) p + sqrt((1 - pi) T) xi(s1); KE_start += p1^2 / 2;
//         p2 = (-eps/2) g + p1; w += eps p2
//     BO: p3 = (-eps/2) g(w') + p2; KE_end += p3^2 / 2; p = sqrt(pi) p3 + ... xi(s2).
struct Upd {
  float* th;          // theta of the leaf
  float* aux;         // rms_prop v or OBABO momentum of the leaf, or null (plain SGLD)
  uint16_t* sh;       // bf16 shadow of the leaf, or null
  uint8_t* sh8;       // e4m3 shadow (W1, W2 in fp8 mode), or null
  float q8;           // scale the e4m3 shadow is written with
  int mode, rms, has_prior;
  float inv_var, log_norm, inv_T, eps, noise_scale, decay, noise, half, alpha, one_m_alpha,
      lmbd;
};

// Per-kernel step coefficients of the current iteration.
struct Coef {
  float eps, noise_scale, decay, noise;
};

__device__ __forceinline__ Coef coef_of(const Step& s) {
  const int it = cur_iter(s);
  Coef k;
  k.eps = s.eps_table[it];
  k.noise_scale = sqrtf((2.0f * s.noise_temperature) * k.eps);
  k.decay = s.decay_table ? s.decay_table[it] : 0.0f;
  k.noise = s.noise_table ? s.noise_table[it] : 0.0f;
  return k;
}

__device__ __forceinline__ Upd make_upd(const Step& s, int c, int leaf, int64_t base,
                                        const Coef& k) {
  Upd u;
  u.th = s.theta + base;
  u.aux = s.mode == MLP_SGLD ? (s.rms ? s.v + base : nullptr) : s.mom + base;
  u.sh = nullptr;
  u.sh8 = nullptr;
  u.q8 = 1.0f;
  if (s.fp8 && leaf < 2) {
    u.sh8 = s.shadow8 + base;
    u.q8 = s.q[(int64_t)c * kQWords + 6 + leaf];
  } else if (!s.fp32) {
    u.sh = s.shadow + base;
  }
  u.mode = s.mode;
  u.rms = s.rms;
  u.has_prior = s.has_prior;
  u.inv_var = s.prior_inv_var;
  u.log_norm = s.prior_log_norm;
  u.inv_T = 1.0f / s.temperature;
  u.eps = k.eps;
  u.noise_scale = k.noise_scale;
  u.decay = k.decay;
  u.noise = k.noise;
  u.half = 0.5f * k.eps;
  u.alpha = s.alpha;
  u.one_m_alpha = 1.0f - s.alpha;
  u.lmbd = s.lmbd;
  return u;
}

// Parameter state is streamed (read once, written once per launch): with SGMC_UPD_NT the
// accesses carry the non-temporal hint.
template <class T>
__device__ __forceinline__ T upd_ld(const T* p) {
  if (SGMC_UPD_NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <class T>
__device__ __forceinline__ void upd_st(T v, T* p) {
  if (SGMC_UPD_NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// UM: the update rule compiled in -- UM_RMS: SGLD with rms_prop (state v in aux);
// UM_OBA / UM_OBO: the O B A and B O halves of an OBABO step (momentum in aux); or
// UM_ANY: the rule of u.mode / u.rms at run time.
constexpr int UM_ANY = 0;
constexpr int UM_RMS = 1;
constexpr int UM_OBA = 2;
constexpr int UM_OBO = 3;

// the rule's mode, and whether it carries state in aux / writes theta
template <int UM>
__device__ __forceinline__ int um_mode(const Upd& u) {
  return UM == UM_RMS ? MLP_SGLD : UM == UM_OBA ? MLP_OBABO_BA : UM == UM_OBO ? MLP_OBABO_BO
                                                                               : u.mode;
}

template <int UM = UM_ANY>
__device__ __forceinline__ void upd_load(const Upd& u, uint32_t i, float& w, float& aux) {
  w = upd_ld(u.th + i);
  aux = (UM != UM_ANY || u.aux) ? upd_ld(u.aux + i) : 0.0f;
}

template <int UM = UM_ANY>
__device__ __forceinline__ void upd_step(const Upd& u, float& w, float& aux, float gsum,
                                         float xi, float& lp, float& ke) {
  float g = gsum;
  if (u.has_prior) {
    g = __builtin_fmaf(w, u.inv_var, g);
    lp = __builtin_fmaf(__builtin_fmaf(w * w, u.inv_var, u.log_norm), -0.5f, lp);
  }
  g = g * u.inv_T;
  const int mode = um_mode<UM>(u);
  if (mode == MLP_SGLD) {
    const float sg = (-u.eps) * g;
    const float sn = u.noise_scale * xi;
    if (UM == UM_RMS || u.rms) {
      const float vn = __builtin_fmaf(u.alpha, aux, u.one_m_alpha * (g * g));
      aux = vn;
      // G = 1 / (lambda + sqrt(v)) and sqrt(G) from one reciprocal square root
      const float rG = __builtin_amdgcn_rsqf(u.lmbd + __builtin_amdgcn_sqrtf(vn));
      w = w + __builtin_fmaf(rG, sn, (rG * rG) * sg);
    } else {
      w = w + (sg + sn);
    }
  } else if (mode == MLP_OBABO_BA) {
    const float p1 = __builtin_fmaf(u.decay, aux, u.noise * xi);
    ke = __builtin_fmaf(p1, p1, ke);
    const float p2 = __builtin_fmaf(-u.half, g, p1);
    aux = p2;
    w = __builtin_fmaf(u.eps, p2, w);
  } else {
    const float p3 = __builtin_fmaf(-u.half, g, aux);
    ke = __builtin_fmaf(p3, p3, ke);
    aux = __builtin_fmaf(u.decay, p3, u.noise * xi);
  }
}

// SH: the GEMM copy written -- SH_NONE, SH_BF16, SH_E4M3 (known from the operand
// precision in the parameter-update GEMMs), or SH_ANY: whichever pointer is set.
constexpr int SH_ANY = -1, SH_NONE = 0, SH_BF16 = 1, SH_E4M3 = 2;

template <int UM = UM_ANY, int SH = SH_ANY>
__device__ __forceinline__ void upd_store(const Upd& u, uint32_t i, float w, float aux,
                                          float& wmax) {
  if (um_mode<UM>(u) != MLP_OBABO_BO) {
    upd_st(w, u.th + i);
    if (SH == SH_E4M3 || (SH == SH_ANY && u.sh8)) {
      upd_st(f2q(w * u.q8), u.sh8 + i);
      wmax = fmaxf(wmax, fabsf(w));
    } else if (SH == SH_BF16 || (SH == SH_ANY && u.sh)) {
      upd_st(f2bf(w), u.sh + i);
    }
  }
  if (UM != UM_ANY || u.aux) upd_st(aux, u.aux + i);
}

// Buffer resources of one leaf's state (theta, v, GEMM copy) for the compiled-in
// rms_prop update: a lane addresses its parameters with one 32-bit offset and the
// per-element offsets are uniform (scalar offset field), so no per-element address
// arithmetic runs on the vector ALU.
constexpr int kRawBufferFlags = 0x00020000;  // word 3 of a raw buffer resource (CDNA)

struct UpdRes {
  __amdgpu_buffer_rsrc_t th, aux, sh;
};

template <int SH>
__device__ __forceinline__ UpdRes make_res(const Upd& u, uint32_t count) {
  UpdRes r;
  r.th = __builtin_amdgcn_make_buffer_rsrc(u.th, (short)0, (int)(4 * count), kRawBufferFlags);
  r.aux = __builtin_amdgcn_make_buffer_rsrc(u.aux, (short)0, (int)(4 * count), kRawBufferFlags);
  if (SH == SH_E4M3)
    r.sh = __builtin_amdgcn_make_buffer_rsrc(u.sh8, (short)0, (int)count, kRawBufferFlags);
  else
    r.sh = __builtin_amdgcn_make_buffer_rsrc(u.sh, (short)0, (int)(2 * count), kRawBufferFlags);
  return r;
}

// element vo + so (vo per lane, so uniform)
__device__ __forceinline__ void res_load(const UpdRes& r, uint32_t vo, uint32_t so, float& w,
                                         float& aux) {
  w = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r.th, 4 * vo, 4 * so, 0));
  aux = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r.aux, 4 * vo, 4 * so, 0));
}

// TH: the rule changes theta (and so its GEMM copy); the B O half of OBABO does not.
template <int SH, bool TH = true>
__device__ __forceinline__ void res_store(const Upd& u, const UpdRes& r, uint32_t vo,
                                          uint32_t so, float w, float aux, float& wmax) {
  if (TH) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(w), r.th, 4 * vo, 4 * so, 0);
    if (SH == SH_E4M3) {
      __builtin_amdgcn_raw_buffer_store_b8(f2q(w * u.q8), r.sh, vo, so, 0);
      wmax = fmaxf(wmax, fabsf(w));
    } else if (SH == SH_BF16) {
      __builtin_amdgcn_raw_buffer_store_b16(f2bf(w), r.sh, 2 * vo, 2 * so, 0);
    }
  }
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(aux), r.aux, 4 * vo, 4 * so, 0);
}

// Key of leaf `leaf` of the noise tree of this update: random_tree(sub) with
// (key, sub) = split(key) for SGLD; s1 / s2 of (key, s1, s2) = split(key, 3) for the
// two momentum refreshes of an OBABO step.
__device__ __forceinline__ Key derive_leaf_key(const Step& s, int chain, int leaf) {
  const Key k = load_key(s.key + 2 * chain);
  Key a, b, c;
  if (s.mode == MLP_SGLD) {
    split2(k, a, b);
    return split_key(b, NLEAF, (uint32_t)leaf);
  }
  split3(k, a, b, c);
  return split_key(s.mode == MLP_OBABO_BA ? b : c, NLEAF, (uint32_t)leaf);
}

// The same key as stored by the step's gather kernel (the chain key does not change
// until the step's last kernel), so the update epilogues load two words instead of
// running four to seven Threefry blocks per wave.
__device__ __forceinline__ Key leaf_key(const Step& s, int chain, int leaf) {
  return load_key(s.lkey + 2 * (chain * NLEAF + leaf));
}

// ------------------------------------------------------------------ gather ----
// Xb[c][j] = bf16(X[idx_j]), lab[c][j] = labels[idx_j]; one wavefront per row.
__global__ void __launch_bounds__(GT) gather_kernel(Step s) {
  const int c = blockIdx.y;
  if (s.update && blockIdx.x == 0 && threadIdx.x < NLEAF)
    store_key(s.lkey + 2 * (c * NLEAF + threadIdx.x), derive_leaf_key(s, c, threadIdx.x));
  const int j = blockIdx.x * (GT / kWave) + threadIdx.x / kWave;
  if (j >= s.n) return;
  const int lane = threadIdx.x & (kWave - 1);
  int64_t r;
  bool valid = true;
  if (s.full) {  // rows row0 + j of the full-data pass; rows past N are zero
    r = s.row0 + j;
    valid = r < s.N;
    if (!valid) r = 0;
  } else {
    Key next, sub, k1, k2;
    split2(load_key(s.dkey + 2 * c), next, sub);
    split2(sub, k1, k2);
    const uint32_t span = (uint32_t)s.N;
    r = (int64_t)draw_index(k1, k2, (uint32_t)s.n, (uint32_t)j, span, s.rmult, s.rmagic);
    if (r < 0 || r >= s.N) r = 0;
  }
  const float* __restrict__ src = s.X + r * s.ldx;
  const int64_t row = (int64_t)c * s.n + j;
  // kGatherBatch float4 loads per lane go out before their conversions and stores, so a
  // row costs D_in / (256 kGatherBatch) memory round trips instead of D_in / 256. The
  // loads are unconditional: offsets past the row are clamped to its last float4 (and
  // not stored), a row past N reads row 0 and is zeroed.
  constexpr int kStep = kWave * 4;
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  auto rowcopy = [&](auto&& put) {
    for (int f0 = lane * 4; f0 < s.D_in; f0 += kGatherBatch * kStep) {
      float4 xs[kGatherBatch];
#pragma unroll
      for (int u = 0; u < kGatherBatch; ++u)
        xs[u] = *reinterpret_cast<const float4*>(src + min(f0 + u * kStep, s.D_in - 4));
#pragma unroll
      for (int u = 0; u < kGatherBatch; ++u)
        if (f0 + u * kStep < s.D_in) put(f0 + u * kStep, valid ? xs[u] : zero4);
    }
  };
  if (s.fp8) {
    const float qs = s.q[(int64_t)c * kQWords + Q_X];
    uint8_t* __restrict__ dst = reinterpret_cast<uint8_t*>(s.Xb) + row * s.D_in;
    float amax = 0.f;
    rowcopy([&](int f, float4 x) {
      amax = fmaxf(amax, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))));
      const uint32_t v = (uint32_t)f2q(x.x * qs) | ((uint32_t)f2q(x.y * qs) << 8) |
                         ((uint32_t)f2q(x.z * qs) << 16) | ((uint32_t)f2q(x.w * qs) << 24);
      *reinterpret_cast<uint32_t*>(dst + f) = v;
    });
    q_note_amax(s, c, Q_X, amax);
  } else if (s.fp32) {
    float* __restrict__ dst = reinterpret_cast<float*>(s.Xb) + row * s.D_in;
    rowcopy([&](int f, float4 x) { *reinterpret_cast<float4*>(dst + f) = x; });
  } else {
    uint16_t* __restrict__ dst = reinterpret_cast<uint16_t*>(s.Xb) + row * s.D_in;
    rowcopy([&](int f, float4 x) {
      const uint32_t lo = (uint32_t)f2bf(x.x) | ((uint32_t)f2bf(x.y) << 16);
      const uint32_t hi = (uint32_t)f2bf(x.z) | ((uint32_t)f2bf(x.w) << 16);
      *reinterpret_cast<uint2*>(dst + f) = make_uint2(lo, hi);
    });
  }
  if (lane == 0) s.lab[row] = valid ? s.labels[r] : -1;
}

// ------------------------------------------------------------------ GEMM ------
// Operands are bf16, e4m3 (P_FP8) with per-tensor scales, or float32 (P_F32): element
// pointers are byte addresses of 2-, 1- or 4-byte elements, strides are in elements.
struct Gemm {
  const void* A;
  int64_t lda, sA;   // sA: chain stride in elements
  const void* B;
  int64_t ldb, sB;
  int M, N, K;
  void* out;         // EPI_BIAS_RELU / EPI_MASK
  int64_t ldo, sO;
  const float* bias; // chain-strided by sBias
  int64_t sBias;
  const void* mask;  // EPI_MASK: ReLU activations (same format as the operands)
  int64_t ldm, sM;
  int leaf;          // EPI_UPDATE / EPI_GRAD: parameter leaf of the [M][N] weight
  int part0;         // first prior-partial slot of this GEMM
  int splits;        // split-K factor (grid z = chains x splits)
  float* part;       // EPI_PARTIAL: [splits][C][M][N] float32 partial products
  int C;
  int ta, tb;        // P_FP8: scale slots of A and B
  int tout;          // P_FP8: scale slot of an e4m3 output, or -1 for a bf16 output
};

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// ---- operand tiles in LDS ----
// A tile: BM rows (m) x BK (k); B tile: BN rows (n) x BK. A row holds BK elements in
// 16-byte chunks (8 bf16 or 16 e4m3); chunk c of row r sits at position c ^ swz(r), so
// that the MFMA fragment reads (consecutive rows, one chunk) and the stores of
// transposed 8x8 blocks (row groups 8 apart) spread over the LDS banks.
template <int CH>
__device__ __forceinline__ int swz(int r) {
  return (r ^ (r >> 3)) & (CH - 1);
}

// 8x8 transpose of 16-bit elements: in[i] holds 8 consecutive columns of row i; out[j]
// holds rows 0..7 of column j.
__device__ __forceinline__ void transpose8x8(const u32x4 (&in)[8], u32x4 (&out)[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const u32x4& x = in[2 * d];
      const u32x4& y = in[2 * d + 1];
      const uint32_t a = (j >> 1) == 0 ? x.x : (j >> 1) == 1 ? x.y : (j >> 1) == 2 ? x.z : x.w;
      const uint32_t b = (j >> 1) == 0 ? y.x : (j >> 1) == 1 ? y.y : (j >> 1) == 2 ? y.z : y.w;
      const uint32_t v = (j & 1) ? ((a >> 16) | (b & 0xFFFF0000u)) : ((a & 0xFFFFu) | (b << 16));
      if (d == 0) out[j].x = v;
      else if (d == 1) out[j].y = v;
      else if (d == 2) out[j].z = v;
      else out[j].w = v;
    }
  }
}

// 4x4 transpose of 32-bit elements: in[i] holds 4 consecutive columns of row i.
__device__ __forceinline__ void transpose8x8(const u32x4 (&in)[4], u32x4 (&out)[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) out[j] = u32x4{in[0][j], in[1][j], in[2][j], in[3][j]};
}

// 8x8 transpose of bytes: in[i] holds 8 consecutive columns of row i.
__device__ __forceinline__ void transpose8x8(const u32x2 (&in)[8], u32x2 (&out)[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t w = (j < 4) ? in[i].x : in[i].y;
      const uint32_t byte = (w >> (8 * (j & 3))) & 0xFFu;
      if (i < 4) lo |= byte << (8 * i);
      else hi |= byte << (8 * (i - 4));
    }
    out[j].x = lo;
    out[j].y = hi;
  }
}

// Loads of one operand tile (ROWS x BK elements of EB bytes, at k0) into registers and
// their stores to LDS. KC: the source is k-contiguous (row r at grow(r) * ld);
// otherwise it is row-contiguous (k-row k at k * ld) and each thread moves one TBxTB
// block (8x8 of 1- or 2-byte elements, 4x4 of 4-byte ones), transposed in registers.
// grow maps tile rows to source rows (contiguous within a block).
template <int ROWS, int BK, bool KC, int EB>
struct Tile {
  static constexpr int EPC = 16 / EB;                 // elements per chunk
  static constexpr int CH = BK / EPC;                 // chunks per row
  static constexpr int ROWB = BK * EB;                // bytes per LDS row
  static constexpr int TB = EB == 4 ? 4 : 8;          // transposed block edge
  static constexpr int NR = KC ? ROWS * CH / GT : TB;  // registers per thread
  static constexpr int BLOCKS = ROWS * BK / (TB * TB);  // blocks (transposed case)
  using R = typename std::conditional<KC || EB >= 2, u32x4, u32x2>::type;
  static_assert(KC || BLOCKS <= GT, "transposed tile: one block per thread at most");
  static_assert(!KC || (ROWS * CH) % GT == 0, "k-contiguous tile: whole chunks per thread");
  R r[NR];

  __device__ __forceinline__ static int off(int row, int chunk) {
    return row * ROWB + ((chunk ^ swz<CH>(row)) << 4);
  }

  template <typename Grow>
  __device__ __forceinline__ void load(const uint8_t* src, int64_t ld, int k0, Grow grow) {
    const int t = threadIdx.x;
    if constexpr (KC) {
#pragma unroll
      for (int h = 0; h < NR; ++h) {
        const int u = t + h * GT;
        const int row = u / CH, ch = u % CH;
        r[h] = *reinterpret_cast<const u32x4*>(src + (grow(row) * ld + k0 + ch * EPC) * EB);
      }
    } else {
      if (!(BLOCKS >= GT || t < BLOCKS)) return;
      const int r8 = (t % (ROWS / TB)) * TB, k8 = (t / (ROWS / TB)) * TB;
#pragma unroll
      for (int i = 0; i < TB; ++i)
        r[i] = *reinterpret_cast<const R*>(src + ((int64_t)(k0 + k8 + i) * ld + grow(r8)) * EB);
    }
  }

  __device__ __forceinline__ void store(uint8_t* lds) const {
    const int t = threadIdx.x;
    if constexpr (KC) {
#pragma unroll
      for (int h = 0; h < NR; ++h) {
        const int u = t + h * GT;
        const int row = u / CH, ch = u % CH;
        *reinterpret_cast<u32x4*>(lds + off(row, ch)) = r[h];
      }
    } else {
      if (!(BLOCKS >= GT || t < BLOCKS)) return;
      const int r8 = (t % (ROWS / TB)) * TB, k8 = (t / (ROWS / TB)) * TB;
      R o[TB];
      transpose8x8(r, o);
      const int kb = k8 * EB;  // byte offset of the block's k range in a row
#pragma unroll
      for (int j = 0; j < TB; ++j)
        *reinterpret_cast<R*>(lds + off(r8 + j, kb >> 4) + (kb & 15)) = o[j];
    }
  }
};

// k-values one mfma_step consumes: 16 (one 32x32x16 instruction per band), or 8 for
// float32 (four 32x32x2 instructions per band).
template <int PREC>
constexpr int k_step() {
  return PREC == P_F32 ? 8 : 16;
}

// One MFMA step: a and b are read at chunk/offset of rows ra and rb. bf16 / e4m3: lanes
// 0-31 hold k .. k + 7 and lanes 32-63 k + 8 .. k + 15 of one 32x32x16 instruction.
// float32: each lane reads one 16-byte chunk (4 k-values; lanes 32-63 the next chunk)
// and feeds its elements to four 32x32x2 instructions, instruction u taking k + u from
// lanes 0-31 and k + 4 + u from lanes 32-63 (the same k for A and B, so every product
// of the 8 k-values enters once).
template <int BK, int PREC>
__device__ __forceinline__ void mfma_step(const uint8_t* As, const uint8_t* Bs, int kk, int lane,
                                          int w, f32x16& acc0, f32x16& acc1) {
  constexpr int EB = elem_bytes<PREC>();
  using TA = Tile<BM, BK, true, EB>;
  const int k = kk + (k_step<PREC>() / 2) * (lane >> 5);   // this lane's first k
  const int kb = k * EB;
  const int ra0 = lane & 31, ra1 = 32 + (lane & 31), rb0 = w * 32 + (lane & 31);
  const int o0 = TA::off(ra0, kb >> 4) + (kb & 15);
  const int o1 = TA::off(ra1, kb >> 4) + (kb & 15);
  const int ob = TA::off(rb0, kb >> 4) + (kb & 15);
  if constexpr (PREC == P_F32) {
    const f32x4 a0 = *reinterpret_cast<const f32x4*>(As + o0);
    const f32x4 a1 = *reinterpret_cast<const f32x4*>(As + o1);
    const f32x4 b = *reinterpret_cast<const f32x4*>(Bs + ob);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[u], b[u], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[u], b[u], acc1, 0, 0, 0);
    }
  } else if constexpr (PREC == P_FP8) {
    const long a0 = *reinterpret_cast<const long*>(As + o0);
    const long a1 = *reinterpret_cast<const long*>(As + o1);
    const long b = *reinterpret_cast<const long*>(Bs + ob);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a0, b, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a1, b, acc1, 0, 0, 0);
  } else {
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(As + o0);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(As + o1);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bs + ob);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc1, 0, 0, 0);
  }
}

// ReLU mask of an activation element (bf16, e4m3 or float32): positive and nonzero.
template <int PREC>
__device__ __forceinline__ bool mask_on(const void* m, int64_t i) {
  if constexpr (PREC == P_F32) {
    const uint32_t u = __float_as_uint(reinterpret_cast<const float*>(m)[i]);
    return u != 0 && !(u & 0x80000000u);
  } else if constexpr (PREC == P_FP8) {
    const uint8_t b = reinterpret_cast<const uint8_t*>(m)[i];
    return b != 0 && !(b & 0x80u);
  } else {
    const uint16_t h = reinterpret_cast<const uint16_t*>(m)[i];
    return h != 0 && !(h & 0x8000u);
  }
}

// Writes an activation: float32 (f32), e4m3 at scale qs (q8), or bf16.
__device__ __forceinline__ void put_act(void* out, int64_t i, float v, bool q8, float qs,
                                        bool f32) {
  if (f32) reinterpret_cast<float*>(out)[i] = v;
  else if (q8) reinterpret_cast<uint8_t*>(out)[i] = f2q(v * qs);
  else reinterpret_cast<uint16_t*>(out)[i] = f2bf(v);
}

// Bytes per element of an activation tensor written with (q8, f32).
__device__ __forceinline__ int act_bytes(bool q8, bool f32) { return f32 ? 4 : q8 ? 1 : 2; }

// h2 (bf16, or float32 in fp32 mode) and its ReLU mask.
__device__ __forceinline__ float h2_at(const Step& s, int64_t i) {
  if (s.fp32) return reinterpret_cast<const float*>(s.h2)[i];
  return bf2f(reinterpret_cast<const uint16_t*>(s.h2)[i]);
}
__device__ __forceinline__ bool h2_on(const Step& s, int64_t i) {
  if (s.fp32) return mask_on<P_F32>(s.h2, i);
  return mask_on<P_BF16>(s.h2, i);
}

// Reads an activation of tensor t (e4m3 tensors in fp8 mode, bf16 otherwise).
__device__ __forceinline__ float get_act(const Step& s, int c, int t, const void* p, int64_t i) {
  if (s.fp32) return reinterpret_cast<const float*>(p)[i];
  if (s.fp8) return q2f(reinterpret_cast<const uint8_t*>(p)[i]) / s.q[(int64_t)c * kQWords + t];
  return bf2f(reinterpret_cast<const uint16_t*>(p)[i]);
}

// One block's tile of GEMM g: block lin (linear id) of a gx x gy x gz grid.
template <int BK, int PREC, bool A_K, bool B_K, int EPI, int UM>
__device__ __forceinline__ void gemm_tile(const Gemm& g, const Step& s, uint32_t lin,
                                          uint32_t gx, uint32_t gy, uint32_t gz, uint8_t* As,
                                          uint8_t* Bs, float* red) {
  static_assert(BN == 2 * BM, "tile shapes");
  constexpr bool F8 = PREC == P_FP8;
  constexpr int EB = elem_bytes<PREC>();
  // Blocks go to the XCDs round-robin by linear id. In the GEMMs without a parameter
  // update, XCD x takes the contiguous range [x T / 8, (x + 1) T / 8) of logical tiles
  // (bm fastest, then bn, then chain and split), so the blocks that share an operand tile
  // read it through one L2 instead of fetching it from HBM once per XCD (Xb W1 of a
  // 3072-512 net: 390 -> 75 MB per launch). The update GEMMs keep the linear order, whose
  // state stream measured faster.
  const uint32_t T = gx * gy * gz;
  if (EPI != EPI_UPDATE && T % kXcds == 0) lin = (lin % kXcds) * (T / kXcds) + lin / kXcds;
  const int bm = (int)(lin % gx), bn = (int)((lin / gx) % gy);
  const int bz = (int)(lin / (gx * gy));
  const int c = bz / g.splits, sp = bz % g.splits;
  const int kspan = g.K / g.splits;
  const int kb = sp * kspan, ke = kb + kspan;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  f32x16 acc0, acc1;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    acc0[j] = 0.f;
    acc1[j] = 0.f;
  }
  const uint8_t* A = reinterpret_cast<const uint8_t*>(g.A) + (int64_t)c * g.sA * EB;
  const uint8_t* B = reinterpret_cast<const uint8_t*>(g.B) + (int64_t)c * g.sB * EB;
  const int hm = g.M / 2;
  auto growA = [&](int r) -> int64_t { return r < 32 ? bm * 32 + r : hm + bm * 32 + (r - 32); };
  auto growB = [&](int r) -> int64_t { return (int64_t)bn * BN + r; };
  Tile<BM, BK, A_K, EB> ta;
  Tile<BN, BK, B_K, EB> tb;
  ta.load(A, g.lda, kb, growA);
  tb.load(B, g.ldb, kb, growB);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    __syncthreads();
    ta.store(As);
    tb.store(Bs);
    __syncthreads();
    if (k0 + BK < ke) {
      ta.load(A, g.lda, k0 + BK, growA);
      tb.load(B, g.ldb, k0 + BK, growB);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += k_step<PREC>()) mfma_step<BK, PREC>(As, Bs, kk, lane, w, acc0, acc1);
  }

  // F8: products carry the operand scales; ab undoes them
  const float* qc = F8 ? s.q + (int64_t)c * kQWords : nullptr;
  const float ab = F8 ? 1.0f / (qc[g.ta] * qc[g.tb]) : 1.0f;
  const int half = g.M / 2;
  const int n = bn * BN + w * 32 + (lane & 31);
  if (EPI == EPI_PARTIAL) {
    float* pp = g.part + ((int64_t)sp * g.C + c) * g.M * g.N;
#pragma unroll
    for (int band = 0; band < 2; ++band) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int r = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        const int m = (band == 0 ? bm * 32 : half + bm * 32) + r;
        pp[(int64_t)m * g.N + n] = band == 0 ? acc0[j] : acc1[j];
      }
    }
  } else if (EPI == EPI_BIAS_RELU || EPI == EPI_MASK) {
    const bool q8 = F8 && g.tout >= 0;
    const bool f32 = PREC == P_F32;
    const float qs = q8 ? qc[g.tout] : 1.0f;
    void* out = reinterpret_cast<uint8_t*>(g.out) + (int64_t)c * g.sO * act_bytes(q8, f32);
    const float bias = (EPI == EPI_BIAS_RELU) ? g.bias[(int64_t)c * g.sBias + n] : 0.f;
    float amax = 0.f;
#pragma unroll
    for (int band = 0; band < 2; ++band) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int r = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        const int m = (band == 0 ? bm * 32 : half + bm * 32) + r;
        const float a = ab * (band == 0 ? acc0[j] : acc1[j]);
        float vv;
        if (EPI == EPI_BIAS_RELU) {
          vv = fmaxf(a + bias, 0.0f);
        } else {
          vv = mask_on<PREC>(g.mask, (int64_t)c * g.sM + (int64_t)m * g.ldm + n) ? a : 0.0f;
        }
        amax = fmaxf(amax, fabsf(vv));
        put_act(out, (int64_t)m * g.ldo + n, vv, q8, qs, f32);
      }
    }
    if (q8) q_note_amax(s, c, g.tout, amax);
  } else {
    // Parameter epilogue: element (m, n) of the [M][N] leaf at flat i = m N + n, its
    // noise pair i + M N / 2 sits in acc1 at the same lane and register.
    const Coef k = coef_of(s);
    const int64_t base = (int64_t)c * s.P + s.off[g.leaf];
    const float* th = s.theta + base;
    const uint32_t S = (uint32_t)((int64_t)g.M * g.N);
    float lp = 0.0f, ke = 0.0f, wmax = 0.0f;
    if (EPI == EPI_UPDATE) {
      const Key lk = leaf_key(s, c, g.leaf);
      const Upd up = make_upd(s, c, g.leaf, base, k);
      // 32-bit offsets of this lane's parameters: element j of band 0 at
      // o0 + j_row * N, its band-1 pair half * N further
      const uint32_t o0 = (uint32_t)(bm * 32 + 4 * (lane >> 5)) * (uint32_t)g.N + (uint32_t)n;
      const uint32_t oh = (uint32_t)half * (uint32_t)g.N;
      auto off = [&](int j) -> uint32_t {
        return o0 + (uint32_t)(8 * (j >> 2) + (j & 3)) * (uint32_t)g.N;
      };
      // Chunks of kUpdChunk pairs, software-pipelined: the next chunk's loads go out
      // before this chunk's noise arithmetic and stores (each parameter is read and
      // written by this thread only), so its memory time overlaps the arithmetic.
      constexpr int CJ = (BK >= 64 && UM == UM_RMS) ? kUpdChunkDeep : kUpdChunk;
      float w0[CJ], a0[CJ], w1[CJ], a1[CJ];
      constexpr int SH = PREC == P_BF16 ? SH_BF16 : PREC == P_FP8 ? SH_E4M3 : SH_NONE;
      // compiled-in rms_prop: buffer-resource accesses, lane offset o0 plus the uniform
      // row offset of element j (and oh for its band-1 pair)
      const UpdRes rs = UM != UM_ANY ? make_res<SH>(up, (uint32_t)s.size[g.leaf]) : UpdRes{};
      auto so = [&](int j) -> uint32_t { return (uint32_t)(8 * (j >> 2) + (j & 3)) * (uint32_t)g.N; };
      auto load2 = [&](int j, float& wa, float& aa, float& wb, float& ab_) {
        if constexpr (UM != UM_ANY) {
          res_load(rs, o0, so(j), wa, aa);
          res_load(rs, o0, so(j) + oh, wb, ab_);
        } else {
          upd_load<UM>(up, off(j), wa, aa);
          upd_load<UM>(up, off(j) + oh, wb, ab_);
        }
      };
#pragma unroll
      for (int u = 0; u < CJ; ++u) load2(u, w0[u], a0[u], w1[u], a1[u]);
#pragma unroll
      for (int j0 = 0; j0 < 16; j0 += CJ) {
        float nw0[CJ], na0[CJ], nw1[CJ], na1[CJ];
        if (j0 + CJ < 16) {
#pragma unroll
          for (int u = 0; u < CJ; ++u) load2(j0 + CJ + u, nw0[u], na0[u], nw1[u], na1[u]);
        }
#pragma unroll
        for (int u = 0; u < CJ; ++u) {
          const int j = j0 + u;
          const uint32_t i0 = off(j);
          uint32_t b0, b1;
          bits_pair(lk, S, i0, b0, b1);
          upd_step<UM>(up, w0[u], a0[u], ab * acc0[j], bits_to_normal_fast(b0), lp, ke);
          upd_step<UM>(up, w1[u], a1[u], ab * acc1[j], bits_to_normal_fast(b1), lp, ke);
          if constexpr (UM != UM_ANY) {
            res_store<SH, UM != UM_OBO>(up, rs, o0, so(j), w0[u], a0[u], wmax);
            res_store<SH, UM != UM_OBO>(up, rs, o0, so(j) + oh, w1[u], a1[u], wmax);
          } else {
            upd_store<UM, SH>(up, i0, w0[u], a0[u], wmax);
            upd_store<UM, SH>(up, i0 + oh, w1[u], a1[u], wmax);
          }
        }
        if (j0 + CJ < 16) {
#pragma unroll
          for (int u = 0; u < CJ; ++u) {
            w0[u] = nw0[u];
            a0[u] = na0[u];
            w1[u] = nw1[u];
            a1[u] = na1[u];
          }
        }
      }
    } else {
      float w0v[16], w1v[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int r = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        const int64_t i0 = (int64_t)(bm * 32 + r) * g.N + n;
        w0v[j] = th[i0];
        w1v[j] = th[i0 + (int64_t)half * g.N];
      }
      float* go = s.grad + base;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int r = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        const int64_t i0 = (int64_t)(bm * 32 + r) * g.N + n;
        const int64_t i1 = i0 + (int64_t)half * g.N;
        const float gs[2] = {ab * acc0[j], ab * acc1[j]};
        const float wv[2] = {w0v[j], w1v[j]};
        const int64_t ii[2] = {i0, i1};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          float gg = gs[h];
          if (s.has_prior) {
            gg = gg + wv[h] * s.prior_inv_var;
            lp += (s.prior_log_norm + (wv[h] * wv[h]) * s.prior_inv_var) / -2.0f;
          }
          go[ii[h]] = gg / s.temperature;
        }
      }
    }
    if (EPI == EPI_UPDATE && s.fp8 && g.leaf < 2) q_note_amax(s, c, Q_W1 + g.leaf, wmax);
    const int64_t slot = (int64_t)c * s.nparts + g.part0 + bm * gy + bn;
    const float tot = block_sum<GT>(lp, red);
    if (threadIdx.x == 0) s.prior_part[slot] = tot;
    if (EPI == EPI_UPDATE && UM != UM_RMS && s.mode != MLP_SGLD) {
      const float kt = block_sum<GT>(ke, red);
      if (threadIdx.x == 0) s.ke_part[slot] = kt;
    }
  }
}

template <int BK, int PREC, bool A_K, bool B_K, int EPI, int UM = UM_ANY>
__global__ void __launch_bounds__(GT, EPI == EPI_UPDATE ? SGMC_UPD_WAVES : 1)
    gemm_kernel(Gemm g, Step s) {
  constexpr int EB = elem_bytes<PREC>();
  __shared__ __attribute__((aligned(16))) uint8_t As[BM * BK * EB];
  __shared__ __attribute__((aligned(16))) uint8_t Bs[BN * BK * EB];
  __shared__ float red[GT / kWave];
  gemm_tile<BK, PREC, A_K, B_K, EPI, UM>(
      g, s, blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x,
      gridDim.y, gridDim.z, As, Bs, red);
}

// The two parameter GEMMs of an evaluation in one launch (one k-tile depth): a 1-D grid
// whose first ga.gx ga.gy C blocks take gW1's tiles (ga) in their linear order and the
// rest gW2's (gb), so gW2's blocks fill the last round of gW1's instead of a launch and
// a latency-bound round of their own.
template <int BK, int PREC, bool A_K, bool B_K, int EPI, int UM = UM_ANY>
__global__ void __launch_bounds__(GT, EPI == EPI_UPDATE ? SGMC_UPD_WAVES : 1)
    gemm_pair_kernel(Gemm ga, Gemm gb, Step s, uint32_t gxa, uint32_t gya, uint32_t gxb,
                     uint32_t gyb) {
  constexpr int EB = elem_bytes<PREC>();
  __shared__ __attribute__((aligned(16))) uint8_t As[BM * BK * EB];
  __shared__ __attribute__((aligned(16))) uint8_t Bs[BN * BK * EB];
  __shared__ float red[GT / kWave];
  const uint32_t na = gxa * gya * (uint32_t)s.C;
  if (blockIdx.x < na)
    gemm_tile<BK, PREC, A_K, B_K, EPI, UM>(ga, s, blockIdx.x, gxa, gya, s.C, As, Bs, red);
  else
    gemm_tile<BK, PREC, A_K, B_K, EPI, UM>(gb, s, blockIdx.x - na, gxb, gyb, s.C, As, Bs, red);
}

// Split-K finish: out = epilogue(ab * sum over splits of the partial products), the
// splits added in order; four consecutive columns per thread.
template <int EPI, int PREC>
__global__ void __launch_bounds__(GT) splitk_finish_kernel(Gemm g, Step s) {
  constexpr bool F8 = PREC == P_FP8;
  constexpr bool f32 = PREC == P_F32;
  __shared__ float red[GT / kWave];
  const int c = blockIdx.y;
  const int64_t e = ((int64_t)blockIdx.x * GT + threadIdx.x) * 4;
  const float* qc = F8 ? s.q + (int64_t)c * kQWords : nullptr;
  const bool q8 = F8 && g.tout >= 0;
  float amax = 0.f;
  if (e < (int64_t)g.M * g.N) {
    const int m = (int)(e / g.N), n = (int)(e % g.N);
    const float ab = F8 ? 1.0f / (qc[g.ta] * qc[g.tb]) : 1.0f;
    const float qs = q8 ? qc[g.tout] : 1.0f;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int sp = 0; sp < g.splits; ++sp) {
      const float4 v =
          *reinterpret_cast<const float4*>(g.part + ((int64_t)sp * g.C + c) * g.M * g.N + e);
      a.x += v.x;
      a.y += v.y;
      a.z += v.z;
      a.w += v.w;
    }
    const float r[4] = {ab * a.x, ab * a.y, ab * a.z, ab * a.w};
    void* out = reinterpret_cast<uint8_t*>(g.out) + (int64_t)c * g.sO * act_bytes(q8, f32);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float v;
      if (EPI == EPI_BIAS_RELU) {
        v = fmaxf(r[t] + g.bias[(int64_t)c * g.sBias + n + t], 0.0f);
      } else {
        v = mask_on<PREC>(g.mask, (int64_t)c * g.sM + (int64_t)m * g.ldm + n + t) ? r[t] : 0.0f;
      }
      amax = fmaxf(amax, fabsf(v));
      put_act(out, (int64_t)m * g.ldo + n + t, v, q8, qs, f32);
    }
  }
  if (q8) q_note_amax_block(s, c, g.tout, amax, red);
}

// ------------------------------------------------------------------ head ------
// logits = h2 W3 + b3; p = softmax; lik = logits[label] - lse; dL = c (onehot - p);
// dA2 = (dL W3^T) * (h2 > 0). A wavefront takes R rows of one chain. W3S: W3 is staged in
// LDS (row stride kW3Stride, so lanes reading consecutive rows of W3 hit distinct banks)
// and each W3 row read serves all of the wavefront's rows; otherwise W3 comes from global
// memory. For every row the sums run in the same order either way: lane-partial sums
// over the columns a = lane + 64 u in ascending u, then wave sums.
constexpr int kW3Stride = NCLS + 1;

template <bool W3S, int R>
__global__ void __launch_bounds__(GT) head_kernel(Step s) {
  extern __shared__ float w3s[];  // W3S: [H2][kW3Stride]
  const int c = blockIdx.y;
  const int lane = threadIdx.x & (kWave - 1);
  const float* th = s.theta + (int64_t)c * s.P;
  const float* W3 = th + s.off[2];
  const float* b3 = th + s.off[5];
  if (W3S) {
    // 16 loads per thread go out before their LDS stores (indices past W3 are clamped
    // to its last element and not stored)
    const int total = s.H2 * NCLS;
    for (int i0 = threadIdx.x; i0 < total; i0 += 16 * GT) {
      float v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = W3[min(i0 + u * GT, total - 1)];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int i = i0 + u * GT;
        if (i < total) w3s[(i / NCLS) * kW3Stride + i % NCLS] = v[u];
      }
    }
    __syncthreads();
  }
  auto w3at = [&](int a, int k) -> float {
    return W3S ? w3s[a * kW3Stride + k] : W3[(int64_t)a * NCLS + k];
  };
  const int j0 = (blockIdx.x * (GT / kWave) + threadIdx.x / kWave) * R;
  if (j0 >= s.n) return;  // s.n is a multiple of 4 R: rows j0 .. j0 + R - 1 exist
  const int64_t row0 = (int64_t)c * s.n + j0;
  float z[R][NCLS];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < NCLS; ++k) z[r][k] = 0.f;
  // columns in batches of four per lane (then pairs: H2 is a multiple of 128); the h2
  // loads of a batch go out together
  auto zbatch = [&](int a0, auto nb) {
    constexpr int NB = decltype(nb)::value;
    float h[R][NB], w[NB][NCLS];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
#pragma unroll
      for (int r = 0; r < R; ++r) h[r][u] = h2_at(s, (row0 + r) * s.H2 + a0 + u * kWave);
#pragma unroll
      for (int k = 0; k < NCLS; ++k) w[u][k] = w3at(a0 + u * kWave, k);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < NCLS; ++k) z[r][k] += h[r][u] * w[u][k];
  };
  {
    int a0 = lane;
    for (; a0 + 3 * kWave < s.H2; a0 += 4 * kWave) zbatch(a0, std::integral_constant<int, 4>{});
    for (; a0 < s.H2; a0 += 2 * kWave) zbatch(a0, std::integral_constant<int, 2>{});
  }
  const float cscale = ((float)(-s.N) / (float)s.n) / s.temperature;
  float dl[R][NCLS];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    float zr[NCLS];
#pragma unroll
    for (int k = 0; k < NCLS; ++k) zr[k] = wave_sum(z[r][k]) + b3[k];
    float mx = zr[0];
#pragma unroll
    for (int k = 1; k < NCLS; ++k) mx = fmaxf(mx, zr[k]);
    float se = 0.f;
#pragma unroll
    for (int k = 0; k < NCLS; ++k) se += expf(zr[k] - mx);
    const float lse = mx + logf(se);
    int lab = s.lab[row0 + r];
    if (s.full) {  // likelihood only; padding rows (label -1) contribute 0
      const int lb = lab < 0 ? 0 : (lab >= NCLS ? NCLS - 1 : lab);
      float zl = zr[0];
#pragma unroll
      for (int k = 1; k < NCLS; ++k)
        if (k == lb) zl = zr[k];
      if (lane == 0) s.lik[row0 + r] = lab < 0 ? 0.0f : zl - lse;
      continue;
    }
    lab = lab < 0 ? 0 : (lab >= NCLS ? NCLS - 1 : lab);
    float zl = 0.f;
#pragma unroll
    for (int k = 0; k < NCLS; ++k) {
      const float p = expf(zr[k] - lse);
      dl[r][k] = cscale * ((k == lab ? 1.0f : 0.0f) - p);
      if (k == lab) zl = zr[k];
    }
    if (lane == 0) {
      s.lik[row0 + r] = zl - lse;
#pragma unroll
      for (int k = 0; k < NCLS; ++k) s.dL[(row0 + r) * NCLS + k] = dl[r][k];
    }
  }
  if (s.full) return;
  const float qs = s.fp8 ? s.q[(int64_t)c * kQWords + Q_DA2] : 1.0f;
  float amax = 0.f;
  auto dbatch = [&](int a0, auto nb) {
    constexpr int NB = decltype(nb)::value;
    float w[NB][NCLS];
    bool on[R][NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
#pragma unroll
      for (int r = 0; r < R; ++r) on[r][u] = h2_on(s, (row0 + r) * s.H2 + a0 + u * kWave);
#pragma unroll
      for (int k = 0; k < NCLS; ++k) w[u][k] = w3at(a0 + u * kWave, k);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float d = 0.f;
#pragma unroll
        for (int k = 0; k < NCLS; ++k) d += dl[r][k] * w[u][k];
        const float v = on[r][u] ? d : 0.0f;
        amax = fmaxf(amax, fabsf(v));
        put_act(s.da2, (row0 + r) * s.H2 + a0 + u * kWave, v, s.fp8 != 0, qs, s.fp32 != 0);
      }
    }
  };
  {
    int a0 = lane;
    for (; a0 + 3 * kWave < s.H2; a0 += 4 * kWave) dbatch(a0, std::integral_constant<int, 4>{});
    for (; a0 < s.H2; a0 += 2 * kWave) dbatch(a0, std::integral_constant<int, 2>{});
  }
  if (s.fp8) q_note_amax(s, c, Q_DA2, amax);
}

// ------------------------------------------------------- small-leaf sums -----
// Element i of an activation tensor stored at precision PREC (e4m3 at scale qs).
template <int PREC>
__device__ __forceinline__ float act_of(const void* p, int64_t i, float qs) {
  if constexpr (PREC == P_F32) return reinterpret_cast<const float*>(p)[i];
  else if constexpr (PREC == P_FP8) return q2f(reinterpret_cast<const uint8_t*>(p)[i]) / qs;
  else return bf2f(reinterpret_cast<const uint16_t*>(p)[i]);
}

// Row-strip partial sums of the small leaves' gradients (flat order W3 [H2][10], b1,
// b2, b3): for strip t, part[t][.] = sum over its kStrip rows of h2^T dL, dA1, dA2, dL.
// Grid x: ceil(H2 / GT) blocks for W3 and b2, ceil(H1 / GT) for b1, one for b3.
// PREC: the evaluation's GEMM precision (h2 is bf16 in e4m3 mode). A thread loads its
// column's kStrip rows before summing them (in ascending row order), one memory round
// trip per strip; the strip's dL rows are staged in LDS.
template <int PREC>
__global__ void __launch_bounds__(GT) colsum_kernel(Step s) {
  constexpr int H2P = PREC == P_F32 ? P_F32 : P_BF16;
  __shared__ float dls[kStrip * NCLS];
  const int c = blockIdx.z, t = blockIdx.y;
  const int nb2 = (s.H2 + GT - 1) / GT, nb1 = (s.H1 + GT - 1) / GT;
  const int64_t nW3 = (int64_t)s.H2 * NCLS;
  const int64_t total = nW3 + s.H1 + s.H2 + NCLS;
  float* out = s.col_part + ((int64_t)c * (s.n / kStrip) + t) * total;
  const int64_t r0 = (int64_t)c * s.n + (int64_t)t * kStrip;
  const float* dL = s.dL + r0 * NCLS;
  const int bx = blockIdx.x;
  if (bx < nb2) {
    for (int q = threadIdx.x; q < kStrip * NCLS; q += GT) dls[q] = dL[q];
    __syncthreads();
    const int a = bx * GT + threadIdx.x;
    if (a >= s.H2) return;
    const float qs = PREC == P_FP8 ? s.q[(int64_t)c * kQWords + Q_DA2] : 1.0f;
    float h[kStrip], d2[kStrip];
#pragma unroll
    for (int r = 0; r < kStrip; ++r) {
      h[r] = act_of<H2P>(s.h2, (r0 + r) * s.H2 + a, 1.0f);
      d2[r] = act_of<PREC>(s.da2, (r0 + r) * s.H2 + a, qs);
    }
    float w3[NCLS];
#pragma unroll
    for (int k = 0; k < NCLS; ++k) w3[k] = 0.f;
    float b2 = 0.f;
#pragma unroll
    for (int r = 0; r < kStrip; ++r) {
      b2 += d2[r];
#pragma unroll
      for (int k = 0; k < NCLS; ++k) w3[k] += h[r] * dls[r * NCLS + k];
    }
#pragma unroll
    for (int k = 0; k < NCLS; ++k) out[(int64_t)a * NCLS + k] = w3[k];
    out[nW3 + s.H1 + a] = b2;
  } else if (bx < nb2 + nb1) {
    const int i = (bx - nb2) * GT + threadIdx.x;
    if (i >= s.H1) return;
    const float qs = PREC == P_FP8 ? s.q[(int64_t)c * kQWords + Q_DA1] : 1.0f;
    float d1[kStrip];
#pragma unroll
    for (int r = 0; r < kStrip; ++r) d1[r] = act_of<PREC>(s.da1, (r0 + r) * s.H1 + i, qs);
    float b1 = 0.f;
#pragma unroll
    for (int r = 0; r < kStrip; ++r) b1 += d1[r];
    out[nW3 + i] = b1;
  } else {
    const int k = threadIdx.x;
    if (k >= NCLS) return;
    float dk[kStrip];
#pragma unroll
    for (int r = 0; r < kStrip; ++r) dk[r] = dL[r * NCLS + k];
    float b3 = 0.f;
#pragma unroll
    for (int r = 0; r < kStrip; ++r) b3 += dk[r];
    out[nW3 + s.H1 + s.H2 + k] = b3;
  }
}

// ----------------------------------------------------------- small params -----
// W3, b1, b2, b3: gradient by a loop over the n rows, update, prior partial; the last
// block of a chain forms U and advances the chain's keys (and the iteration counter
// once every chain is done).
__global__ void __launch_bounds__(GT) small_kernel(Step s, int part0) {
  __shared__ float red[GT / kWave];
  __shared__ double red_d[GT / kWave];
  __shared__ int flag;
  const int c = blockIdx.y;
  const int64_t nW3 = (int64_t)s.H2 * NCLS;
  const int64_t total = nW3 + s.H1 + s.H2 + NCLS;
  const int64_t p = (int64_t)blockIdx.x * GT + threadIdx.x;
  const int it = cur_iter(s);
  const Coef cf = coef_of(s);
  float lp = 0.f, ke = 0.f;
  if (p < total) {
    int leaf;
    int64_t i;
    const int strips = s.n / kStrip;
    const float* cp = s.col_part + (int64_t)c * strips * total + p;
    float gsum = 0.f;
    for (int t = 0; t < strips; t += 8) {
      float x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = t + u < strips ? cp[(int64_t)(t + u) * total] : 0.f;
#pragma unroll
      for (int u = 0; u < 8; ++u) gsum += x[u];
    }
    if (p < nW3) {
      leaf = 2;
      i = p;
    } else if (p < nW3 + s.H1) {
      leaf = 3;
      i = p - nW3;
    } else if (p < nW3 + s.H1 + s.H2) {
      leaf = 4;
      i = p - nW3 - s.H1;
    } else {
      leaf = 5;
      i = p - nW3 - s.H1 - s.H2;
    }
    const int64_t base = (int64_t)c * s.P + s.off[leaf];
    if (s.update) {
      const Key lk = leaf_key(s, c, leaf);
      const float xi = bits_to_normal_fast(bits_word(lk, (uint32_t)s.size[leaf], (uint32_t)i));
      float wmax = 0.f, w, aux;
      const Upd up = make_upd(s, c, leaf, base, cf);
      upd_load(up, (uint32_t)i, w, aux);
      upd_step(up, w, aux, gsum, xi, lp, ke);
      upd_store(up, (uint32_t)i, w, aux, wmax);
    } else {
      grad_param(s, s.theta + base, s.grad + base, i, gsum, lp);
    }
  }
  const bool obabo = s.update && s.mode != MLP_SGLD;
  const int64_t slot = (int64_t)c * s.nparts + part0 + blockIdx.x;
  const float tot = block_sum<GT>(lp, red);
  if (threadIdx.x == 0) publish(s.prior_part + slot, tot);
  if (obabo) {
    const float kt = block_sum<GT>(ke, red);
    if (threadIdx.x == 0) publish(s.ke_part + slot, kt);
  }
  if (!arrive_last(s.counters + c, gridDim.x, &flag)) return;
  double kin = 0.0;
  if (obabo) {
    for (int q = threadIdx.x; q < s.nparts; q += GT)
      kin += (double)consume(s.ke_part + (int64_t)c * s.nparts + q);
    kin = block_sum_d<GT>(kin, red_d);
  }
  float lsum = 0.f;
  for (int r = threadIdx.x; r < s.n; r += GT) lsum += s.lik[(int64_t)c * s.n + r];
  lsum = block_sum<GT>(lsum, red);
  float prior = 0.f;
  for (int q = threadIdx.x; q < s.nparts; q += GT)
    prior += consume(s.prior_part + (int64_t)c * s.nparts + q);
  prior = block_sum<GT>(prior, red);
  if (threadIdx.x == 0) {
    const float mean = lsum / (float)s.n;
    const float Uc = ((float)(-s.N) * mean - prior) / s.temperature;
    if (obabo) {
      // ObaboState bookkeeping (integrator.py:203-273): the step's potential is the mean
      // of its two evaluations; the key advances once per step, the data key per draw.
      Key next, sub;
      split2(load_key(s.dkey + 2 * c), next, sub);
      store_key(s.dkey + 2 * c, next);
      if (s.mode == MLP_OBABO_BA) {
        s.U_half[c] = Uc;
        s.ke[2 * c] += 0.5 * kin;
      } else {
        s.U[c] = 0.5f * (s.U_half[c] + Uc);
        s.ke[2 * c + 1] += 0.5 * kin;
        Key a, b, cc;
        split3(load_key(s.key + 2 * c), a, b, cc);
        store_key(s.key + 2 * c, a);
      }
    } else {
      s.U[c] = Uc;
    }
    if (s.fp8) {
      // delayed scaling: this evaluation's amax sets the next scales; a weight shadow
      // written by this update carries the scale it was written with
      float* qc = s.q + (int64_t)c * kQWords;
      unsigned* am = reinterpret_cast<unsigned*>(qc + 8);
      for (int t = Q_X; t <= Q_DA1; ++t) {
        const float a = __uint_as_float(am[t]);
        if (a > 0.0f) qc[t] = kQTarget / a;
        am[t] = 0u;
      }
      if (s.update) {
        for (int l = 0; l < 2; ++l) {
          qc[Q_W1 + l] = qc[6 + l];
          const float a = __uint_as_float(am[Q_W1 + l]);
          if (a > 0.0f) qc[6 + l] = kQTarget / a;
          am[Q_W1 + l] = 0u;
        }
      }
    }
    if (s.update && !obabo) {
      const int ss = s.save_slot ? s.save_slot[it] : -1;
      if (ss >= 0 && ss < s.n_slots && s.samples_U) s.samples_U[(int64_t)c * s.n_slots + ss] = Uc;
      Key next, sub;
      split2(load_key(s.key + 2 * c), next, sub);
      store_key(s.key + 2 * c, next);
      split2(load_key(s.dkey + 2 * c), next, sub);
      store_key(s.dkey + 2 * c, next);
      // the last chain advances the iteration counter (read by the next launch)
      const unsigned prev = atomicAdd(s.done, 1u);
      if (prev == (unsigned)s.C - 1) {
        atomicExch(s.done, 0u);
        atomicAdd(s.iter, 1);
      }
    }
  }
}

// Copies theta into the sample buffer at kept iterations (after the step).
__global__ void __launch_bounds__(GT) save_kernel(Step s, int it_offset) {
  const int c = blockIdx.y;
  // the iteration that just completed is iter - 1
  const int it = max(0, min(s.iter[0] - it_offset, s.table_len - 1));
  const int slot = s.save_slot ? s.save_slot[it] : -1;
  if (slot < 0 || slot >= s.n_slots || !s.samples) return;
  for (int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x; i < s.P; i += (int64_t)gridDim.x * GT)
    s.samples[((int64_t)c * s.n_slots + slot) * s.P + i] = s.theta[(int64_t)c * s.P + i];
}

// theta -> bf16 shadow.
__global__ void __launch_bounds__(GT) shadow_kernel(const float* th, uint16_t* sh, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x; i < n; i += (int64_t)gridDim.x * GT)
    sh[i] = f2bf(th[i]);
}

// ------------------------------------------------------------ fp8 set-up ------
__global__ void __launch_bounds__(GT) q_zero_kernel(float* q, int n) {
  const int i = blockIdx.x * GT + threadIdx.x;
  if (i < n) q[i] = 0.0f;
}

// max |W1|, max |W2| of every chain into the amax words.
__global__ void __launch_bounds__(GT) w_amax_kernel(Step s) {
  const int c = blockIdx.y;
  const float* th = s.theta + (int64_t)c * s.P;
  float m0 = 0.f, m1 = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x; i < s.off[2];
       i += (int64_t)gridDim.x * GT) {
    const float a = fabsf(th[i]);
    if (i < s.off[1]) m0 = fmaxf(m0, a);
    else m1 = fmaxf(m1, a);
  }
  q_note_amax(s, c, Q_W1, m0);
  q_note_amax(s, c, Q_W2, m1);
}

// Weight scales from the amax words (current and next shadow), activation scales 1.
__global__ void __launch_bounds__(GT) q_init_kernel(Step s) {
  const int c = blockIdx.x * GT + threadIdx.x;
  if (c >= s.C) return;
  float* qc = s.q + (int64_t)c * kQWords;
  unsigned* am = reinterpret_cast<unsigned*>(qc + 8);
  for (int t = Q_X; t <= Q_DA1; ++t) qc[t] = 1.0f;
  for (int l = 0; l < 2; ++l) {
    const float a = __uint_as_float(am[Q_W1 + l]);
    const float sc = a > 0.0f ? kQTarget / a : 1.0f;
    qc[Q_W1 + l] = sc;
    qc[6 + l] = sc;
    am[Q_W1 + l] = 0u;
  }
}

__global__ void __launch_bounds__(GT) w_quant_kernel(Step s) {
  const int c = blockIdx.y;
  const float* th = s.theta + (int64_t)c * s.P;
  const float* qc = s.q + (int64_t)c * kQWords;
  for (int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x; i < s.off[2];
       i += (int64_t)gridDim.x * GT)
    s.shadow8[(int64_t)c * s.P + i] = f2q(th[i] * qc[i < s.off[1] ? Q_W1 : Q_W2]);
}

// ------------------------------------------------------ full-data potential ----
__global__ void __launch_bounds__(GT) zero_d_kernel(double* a, int n) {
  const int i = blockIdx.x * GT + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// acc[c] += sum of the pass's per-row likelihoods in float64 (fixed order).
__global__ void __launch_bounds__(GT) full_acc_kernel(Step s, double* acc) {
  __shared__ double red[GT / kWave];
  const int c = blockIdx.x;
  double t = 0.0;
  for (int r = threadIdx.x; r < s.n; r += GT) t += (double)s.lik[(int64_t)c * s.n + r];
  t = block_sum_d<GT>(t, red);
  if (threadIdx.x == 0) acc[c] += t;
}

// out[c] = (-acc[c] - log p(theta_c)) / T with the float64 Gaussian log prior; grid
// (kFinalBlocks, C), the chain's last-arriving block combines the block sums.
__global__ void __launch_bounds__(GT) full_final_kernel(Step s, const double* acc, double* out) {
  __shared__ double red[GT / kWave];
  __shared__ int flag;
  const int c = blockIdx.y;
  double lp = 0.0;
  if (s.has_prior) {
    const double iv = (double)s.prior_inv_var, ln = (double)s.prior_log_norm;
    const float* th = s.theta + (int64_t)c * s.P;
    for (int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x; i < s.P;
         i += (int64_t)gridDim.x * GT) {
      const double w = (double)th[i];
      lp += (ln + w * w * iv) / -2.0;
    }
  }
  lp = block_sum_d<GT>(lp, red);
  if (threadIdx.x == 0) publish(s.fin_part + (int64_t)c * gridDim.x + blockIdx.x, lp);
  if (!arrive_last(s.counters + c, gridDim.x, &flag)) return;
  double t = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += GT)
    t += consume(s.fin_part + (int64_t)c * gridDim.x + b);
  t = block_sum_d<GT>(t, red);
  if (threadIdx.x == 0) out[c] = (-acc[c] - t) / (double)s.temperature;
}

// ----------------------------------------------------------------- SGGMC MH ----
struct MH {
  double* U_full;
  const double* U_prop;
  float* theta_old;
  float* mom_old;
  float* U_old;
  uint32_t* key;
  float* stats;
  int32_t* accepted;
};

// Every block of chain c evaluates the same decision from inputs that only the chain's
// last-arriving block rewrites, applies it to its slice of the parameters, and saves
// the slice at a kept iteration.
__global__ void __launch_bounds__(GT) mh_kernel(Step s, MH m) {
  __shared__ int flag;
  const int c = blockIdx.y;
  const int it = cur_iter(s);
  const float T = s.noise_temperature;
  const double dE = (m.U_prop[c] - m.U_full[c]) + (s.ke[2 * c + 1] - s.ke[2 * c]);
  double log_alpha = -1.0 / (double)T * dE;
  log_alpha = log_alpha <= 0.0 ? log_alpha : 0.0;
  Key next, sub;
  split2(load_key(m.key + 2 * c), next, sub);
  const float u = bits_to_uniform(bits_word(sub, 1u, 0u), 0.0f, 1.0f);
  const bool accept = (double)logf_cephes(u) < log_alpha;
  const int slot = s.save_slot ? s.save_slot[it] : -1;
  const bool save = s.samples && slot >= 0 && slot < s.n_slots;
  float* th = s.theta + (int64_t)c * s.P;
  float* mo = s.mom + (int64_t)c * s.P;
  float* tho = m.theta_old + (int64_t)c * s.P;
  float* moo = m.mom_old + (int64_t)c * s.P;
  uint16_t* sh = s.fp32 ? nullptr : s.shadow + (int64_t)c * s.P;
  float* smp = save ? s.samples + ((int64_t)c * s.n_slots + slot) * s.P : nullptr;
  for (int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x; i < s.P;
       i += (int64_t)gridDim.x * GT) {
    float w;
    if (accept) {
      w = th[i];
      tho[i] = w;
      moo[i] = mo[i];
    } else {
      w = tho[i];
      th[i] = w;
      mo[i] = moo[i];
      if (s.fp8 && i < s.off[2]) {
        const int l = i < s.off[1] ? 0 : 1;
        s.shadow8[(int64_t)c * s.P + i] = f2q(w * s.q[(int64_t)c * kQWords + Q_W1 + l]);
      } else if (!s.fp32) {
        sh[i] = f2bf(w);
      }
    }
    if (save) smp[i] = w;
  }
  if (!arrive_last(s.counters + c, gridDim.x, &flag)) return;
  if (threadIdx.x == 0) {
    const float U = accept ? s.U[c] : m.U_old[c];
    s.U[c] = U;
    m.U_old[c] = U;
    if (accept) m.U_full[c] = m.U_prop[c];
    const double dk = s.ke[2 * c + 1] - s.ke[2 * c];
    s.ke[2 * c] = 0.0;
    s.ke[2 * c + 1] = 0.0;
    store_key(m.key + 2 * c, next);
    if (m.accepted && accept) m.accepted[c] += 1;
    if (save && s.samples_U) s.samples_U[(int64_t)c * s.n_slots + slot] = U;
    if (save && m.stats) {
      float* st = m.stats + ((int64_t)c * s.n_slots + slot) * 3;
      st[0] = (float)exp(log_alpha);
      st[1] = s.eps_table[it];
      st[2] = (float)dk;
    }
    const unsigned prev = atomicAdd(s.done, 1u);
    if (prev == (unsigned)s.C - 1) {
      atomicExch(s.done, 0u);
      atomicAdd(s.iter, 1);
    }
  }
}

}  // namespace mlp
}  // namespace sgmc

using namespace sgmc;
using namespace sgmc::mlp;

namespace {

struct Layout {
  int64_t xb, lab, h1, h2, da2, da1, dl, lik, lkey, prior, ke, facc, col, fin, split, total;
  int nparts, g4_blocks, g5_blocks, small_blocks;
};

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

Layout layout_of(int C, int n, int D_in, int H1, int H2) {
  Layout L;
  int64_t o = 0;
  // activations sized for float32 elements (the widest precision)
  L.xb = o; o = align256(o + (int64_t)C * n * D_in * 4);
  L.lab = o; o = align256(o + (int64_t)C * n * 4);
  L.h1 = o; o = align256(o + (int64_t)C * n * H1 * 4);
  L.h2 = o; o = align256(o + (int64_t)C * n * H2 * 4);
  L.da2 = o; o = align256(o + (int64_t)C * n * H2 * 4);
  L.da1 = o; o = align256(o + (int64_t)C * n * H1 * 4);
  L.dl = o; o = align256(o + (int64_t)C * n * NCLS * 4);
  L.lik = o; o = align256(o + (int64_t)C * n * 4);
  L.lkey = o; o = align256(o + (int64_t)C * NLEAF * 2 * 4);
  L.g4_blocks = (H1 / BM) * (H2 / BN);
  L.g5_blocks = (D_in / BM) * (H1 / BN);
  const int64_t small_total = (int64_t)H2 * NCLS + H1 + H2 + NCLS;
  L.small_blocks = (int)((small_total + GT - 1) / GT);
  L.nparts = L.g4_blocks + L.g5_blocks + L.small_blocks;
  L.prior = o; o = align256(o + (int64_t)C * L.nparts * 4);
  L.ke = o; o = align256(o + (int64_t)C * L.nparts * 4);
  L.facc = o; o = align256(o + (int64_t)C * 8);
  L.col = o; o = align256(o + (int64_t)C * (n / kStrip) * small_total * 4);
  L.fin = o; o = align256(o + (int64_t)C * kFinalBlocks * 8);
  L.split = o; o = align256(o + (int64_t)4 * C * n * std::max(H1, H2) * 4);
  L.total = o;
  return L;
}

}  // namespace

extern "C" {

int64_t sgmc_mlp_param_count(int D_in, int H1, int H2) {
  return (int64_t)D_in * H1 + (int64_t)H1 * H2 + (int64_t)H2 * NCLS + H1 + H2 + NCLS;
}

int64_t sgmc_mlp_workspace_bytes(int C, int n, int D_in, int H1, int H2) {
  return layout_of(C, n, D_in, H1, H2).total;
}

int sgmc_mlp_shadow(const float* theta, uint16_t* shadow, int64_t count, void* stream) {
  SGMC_REQUIRE(theta && shadow && count >= 0, SGMC_E_ARG, "sgmc_mlp_shadow: bad argument");
  if (count == 0) return SGMC_OK;
  const int64_t blocks = std::min<int64_t>((count + GT - 1) / GT, 65535);
  hipLaunchKernelGGL(shadow_kernel, dim3((unsigned)blocks), dim3(GT), 0, as_stream(stream),
                     theta, shadow, count);
  return check_launch("sgmc_mlp_shadow");
}

}  // extern "C"

namespace {

// Split-K factor of the heuristic; a tuned GEMM may take 4 (the partial-product
// workspace holds 4 C n max(H1, H2) floats).
constexpr int kMaxSplits = 2;

// GEMM call sites of one evaluation (the tune[] slots): G1 = Xb W1, G2 = H1 W2,
// G3 = dA2 W2^T, G4 = H1^T dA2 (gW2), G5 = Xb^T dA1 (gW1).
enum GemmSlot { SLOT_G1 = 0, SLOT_G2, SLOT_G3, SLOT_G4, SLOT_G5, kSlots };

template <int BK, int PREC, bool A_K, bool B_K, int EPI>
void launch_gemm_bk(dim3 grid, Gemm g, const Step& s, hipStream_t st, float* part,
                    int splits) {
  // parameter updates of SGLD with rms_prop and of the two OBABO halves run the kernel
  // with their rule compiled in
  if constexpr (EPI == EPI_UPDATE) {
    if (splits <= 1) {
      g.C = (int)grid.z;
      g.splits = 1;
      if (s.mode == MLP_SGLD && s.rms) {
        hipLaunchKernelGGL((gemm_kernel<BK, PREC, A_K, B_K, EPI, UM_RMS>), grid, dim3(GT), 0, st,
                           g, s);
        return;
      }
      if (s.mode == MLP_OBABO_BA) {
        hipLaunchKernelGGL((gemm_kernel<BK, PREC, A_K, B_K, EPI, UM_OBA>), grid, dim3(GT), 0, st,
                           g, s);
        return;
      }
      if (s.mode == MLP_OBABO_BO) {
        hipLaunchKernelGGL((gemm_kernel<BK, PREC, A_K, B_K, EPI, UM_OBO>), grid, dim3(GT), 0, st,
                           g, s);
        return;
      }
    }
  }
  g.C = (int)grid.z;
  g.splits = splits;
  if (g.splits > 1) {
    g.part = part;
    hipLaunchKernelGGL((gemm_kernel<BK, PREC, A_K, B_K, EPI_PARTIAL>),
                       dim3(grid.x, grid.y, grid.z * g.splits), dim3(GT), 0, st, g, s);
    const int64_t quads = (int64_t)g.M * g.N / 4;
    hipLaunchKernelGGL((splitk_finish_kernel<EPI, PREC>),
                       dim3((unsigned)((quads + GT - 1) / GT), g.C), dim3(GT), 0, st, g, s);
  } else {
    hipLaunchKernelGGL((gemm_kernel<BK, PREC, A_K, B_K, EPI>), grid, dim3(GT), 0, st, g, s);
  }
}

// Launches GEMM `slot` with the k-tile depth and split-K factor of s.plan[slot].
template <bool A_K, bool B_K, int EPI>
void launch_gemm(dim3 grid, const Gemm& g, const Step& s, hipStream_t st, int slot) {
  const int tv = s.plan[slot];
  const int splits = tv & 15;
  const bool deep = (tv >> 4) == 128;
  if (s.fp32) {
    launch_gemm_bk<32, P_F32, A_K, B_K, EPI>(grid, g, s, st, s.split_part, splits);
  } else if (s.fp8) {
    if (deep) launch_gemm_bk<128, P_FP8, A_K, B_K, EPI>(grid, g, s, st, s.split_part, splits);
    else launch_gemm_bk<64, P_FP8, A_K, B_K, EPI>(grid, g, s, st, s.split_part, splits);
  } else {
    if (deep) launch_gemm_bk<128, P_BF16, A_K, B_K, EPI>(grid, g, s, st, s.split_part, splits);
    else launch_gemm_bk<32, P_BF16, A_K, B_K, EPI>(grid, g, s, st, s.split_part, splits);
  }
}

// gW1 (ga, grid ra) and gW2 (gb, grid rb) in one launch at k-tile BK, with the update rule
// compiled in as launch_gemm_bk chooses it.
template <int BK, int PREC, int EPI>
void launch_pair_bk(Gemm ga, dim3 ra, Gemm gb, dim3 rb, const Step& s, hipStream_t st) {
  ga.C = gb.C = (int)ra.z;
  ga.splits = gb.splits = 1;
  const dim3 grid(ra.x * ra.y * ra.z + rb.x * rb.y * rb.z);
  if constexpr (EPI == EPI_UPDATE) {
    if (s.mode == MLP_SGLD && s.rms) {
      hipLaunchKernelGGL((gemm_pair_kernel<BK, PREC, false, false, EPI, UM_RMS>), grid, dim3(GT),
                         0, st, ga, gb, s, ra.x, ra.y, rb.x, rb.y);
      return;
    }
    if (s.mode == MLP_OBABO_BA) {
      hipLaunchKernelGGL((gemm_pair_kernel<BK, PREC, false, false, EPI, UM_OBA>), grid, dim3(GT),
                         0, st, ga, gb, s, ra.x, ra.y, rb.x, rb.y);
      return;
    }
    if (s.mode == MLP_OBABO_BO) {
      hipLaunchKernelGGL((gemm_pair_kernel<BK, PREC, false, false, EPI, UM_OBO>), grid, dim3(GT),
                         0, st, ga, gb, s, ra.x, ra.y, rb.x, rb.y);
      return;
    }
  }
  hipLaunchKernelGGL((gemm_pair_kernel<BK, PREC, false, false, EPI>), grid, dim3(GT), 0, st, ga,
                     gb, s, ra.x, ra.y, rb.x, rb.y);
}

// The parameter GEMMs of an evaluation: gW2 (g4) and gW1 (g5) in one launch when their
// plans agree (the same k-tile; they never split), else one launch each.
template <int EPI>
void launch_param_gemms(const Gemm& g4, dim3 r4, const Gemm& g5, dim3 r5, const Step& s,
                        hipStream_t st) {
  const int t4 = s.plan[SLOT_G4], t5 = s.plan[SLOT_G5];
  if (t4 != t5 || (t5 & 15) != 1) {
    launch_gemm<false, false, EPI>(r4, g4, s, st, SLOT_G4);
    launch_gemm<false, false, EPI>(r5, g5, s, st, SLOT_G5);
    return;
  }
  const int bk = t5 >> 4;
  if (s.fp32) {
    launch_pair_bk<32, P_F32, EPI>(g5, r5, g4, r4, s, st);
  } else if (s.fp8) {
    if (bk == 128) launch_pair_bk<128, P_FP8, EPI>(g5, r5, g4, r4, s, st);
    else launch_pair_bk<64, P_FP8, EPI>(g5, r5, g4, r4, s, st);
  } else {
    if (bk == 128) launch_pair_bk<128, P_BF16, EPI>(g5, r5, g4, r4, s, st);
    else launch_pair_bk<32, P_BF16, EPI>(g5, r5, g4, r4, s, st);
  }
}

// Checks tune[] against the shapes: each nonzero entry is 16 k + s with k a k-tile of
// the precision (bf16 32 / 128, e4m3 64 / 128, float32 32), s in {1, 2, 3, 4} (s > 1 for
// the forward GEMMs G1-G3 only) and the GEMM's K a multiple of k s.
int check_tune(int D_in, int H1, int H2, int n, int precision, const int32_t* tune) {
  if (!tune) return SGMC_OK;
  const int K[kSlots] = {D_in, H1, H2, n, n};
  for (int i = 0; i < kSlots; ++i) {
    const int tv = tune[i];
    if (tv == 0) continue;
    const int bk = tv >> 4, sp = tv & 15;
    const bool bk_ok = precision == 2 ? bk == 32 : precision == 1 ? (bk == 64 || bk == 128)
                                                                : (bk == 32 || bk == 128);
    SGMC_REQUIRE(bk_ok && sp >= 1 && sp <= 4 && (sp == 1 || i < SLOT_G4) &&
                     K[i] % (bk * sp) == 0,
                 SGMC_E_ARG, "sgmc_mlp: tune[%d] = %d (k-tile %d, split %d) does not fit K = %d",
                 i, tv, bk, sp, K[i]);
  }
  return SGMC_OK;
}

// The k-tile depth and split-K factor of each GEMM (16 k + s): tune[i] where set, else
// the built-in choice --
// - k-tiles of 128 when K allows (bf16, e4m3), so the register prefetch of the next tile
//   has a whole tile of matrix work to land behind; 32 (bf16) or 64 (e4m3) otherwise;
//   float32 always 32 (the transposed tiles of 4-byte elements fit one 4x4 block per
//   thread). The bf16 parameter-update GEMMs (G4, G5 with update) take 32: their short
//   K (the minibatch) gains little from deep tiles, and the smaller staging registers
//   give three waves per SIMD instead of two to the update epilogue (+3.4% on a
//   784-1024-1024 MLP). The k order of the matrix instructions is the same either way,
//   so the results are identical.
// - forward GEMMs G1-G3 split K in two while the grid holds at most about four blocks
//   per CU: they have few output tiles but long K. The parameter GEMMs never split.
void plan_gemms(int D_in, int H1, int H2, int n, int C, int precision, int update,
                const int32_t* tune, int* plan) {
  const int K[kSlots] = {D_in, H1, H2, n, n};
  const int64_t tiles[kSlots] = {(int64_t)(n / BM) * (H1 / BN), (int64_t)(n / BM) * (H2 / BN),
                                 (int64_t)(n / BM) * (H1 / BN), (int64_t)(H1 / BM) * (H2 / BN),
                                 (int64_t)(D_in / BM) * (H1 / BN)};
  for (int i = 0; i < kSlots; ++i) {
    if (tune && tune[i]) {
      plan[i] = tune[i];
      continue;
    }
    int bk;
    if (precision == 2) bk = 32;
    else if (precision == 1) bk = K[i] % 128 == 0 ? 128 : 64;
    else bk = K[i] % 128 == 0 && !(update && i >= SLOT_G4) ? 128 : 32;
    int sp = 1;
    if (i < SLOT_G4) {
      const int64_t blocks = tiles[i] * C;
      const int ktiles = K[i] / bk;
      while (sp < kMaxSplits && blocks * sp * 2 <= 1024 && ktiles % (sp * 2) == 0) sp *= 2;
    }
    plan[i] = bk * 16 + sp;
  }
}

// Operand / output formats and scale slots of one GEMM.
Gemm with_q(Gemm g, int ta, int tb, int tout) {
  g.ta = ta;
  g.tb = tb;
  g.tout = tout;
  return g;
}

// Validates the state and fills the kernels' view of it.
int make_step(const sgmc_mlp_state* st, int update, Layout& L, Step& s) {
  SGMC_REQUIRE(st, SGMC_E_ARG, "sgmc_mlp: null state");
  const sgmc_mlp_state& q = *st;
  SGMC_REQUIRE(q.precision >= 0 && q.precision <= 2, SGMC_E_ARG,
               "sgmc_mlp: precision %d is not 0 (bf16), 1 (fp8) or 2 (fp32)", q.precision);
  SGMC_REQUIRE(q.X && q.labels && q.theta && (q.shadow || q.precision == 2) && q.key &&
                   q.dkey && q.work && q.U &&
                   q.eps_table && q.iter && q.counters && q.done,
               SGMC_E_ARG, "sgmc_mlp: null pointer in state");
  SGMC_REQUIRE(update || q.grad, SGMC_E_ARG, "sgmc_mlp_step: gradient mode needs grad");
  SGMC_REQUIRE(q.C >= 1 && q.n >= 1 && q.N >= 1 && q.N <= 2147483647LL, SGMC_E_ARG,
               "sgmc_mlp: bad C/n/N");
  SGMC_REQUIRE(q.n % BM == 0 && q.D_in % BM == 0 && q.H1 % BM == 0, SGMC_E_ARG,
               "sgmc_mlp: n, D_in and H1 must be multiples of %d", BM);
  SGMC_REQUIRE(q.H1 % BN == 0 && q.H2 % BN == 0, SGMC_E_ARG,
               "sgmc_mlp: H1 and H2 must be multiples of %d", BN);
  SGMC_REQUIRE(q.ldx % 4 == 0 && q.ldx >= q.D_in && ((uintptr_t)q.X & 15) == 0, SGMC_E_ARG,
               "sgmc_mlp: X rows must be 16-byte aligned with ldx %% 4 == 0");
  SGMC_REQUIRE(q.P == sgmc_mlp_param_count(q.D_in, q.H1, q.H2), SGMC_E_ARG,
               "sgmc_mlp: P does not match the layer sizes");
  SGMC_REQUIRE(q.table_len >= 1 && q.temperature != 0.0f, SGMC_E_ARG,
               "sgmc_mlp: bad schedule");
  SGMC_REQUIRE((int64_t)q.D_in * q.H1 < 4294967295LL, SGMC_E_ARG, "sgmc_mlp: W1 too large");
  SGMC_REQUIRE(q.mode >= MLP_SGLD && q.mode <= MLP_OBABO_BO, SGMC_E_ARG, "sgmc_mlp: bad mode");
  if (update && q.mode != MLP_SGLD)
    SGMC_REQUIRE(q.mom && q.ke && q.U_half && q.decay_table && q.noise_table, SGMC_E_ARG,
                 "sgmc_mlp_step: OBABO needs mom, ke, U_half and the decay / noise tables");
  L = layout_of(q.C, q.n, q.D_in, q.H1, q.H2);
  SGMC_REQUIRE(q.work_bytes >= L.total, SGMC_E_ARG, "sgmc_mlp: workspace too small");
  uint8_t* w = static_cast<uint8_t*>(q.work);
  s.X = q.X; s.labels = q.labels; s.N = q.N; s.ldx = q.ldx;
  s.D_in = q.D_in; s.H1 = q.H1; s.H2 = q.H2; s.n = q.n; s.C = q.C;
  s.theta = q.theta; s.v = q.rms ? q.v : nullptr; s.shadow = q.shadow; s.P = q.P;
  const int64_t sizes[NLEAF] = {(int64_t)q.D_in * q.H1, (int64_t)q.H1 * q.H2,
                                (int64_t)q.H2 * NCLS, q.H1, q.H2, NCLS};
  int64_t o = 0;
  for (int l = 0; l < NLEAF; ++l) {
    s.off[l] = o;
    s.size[l] = sizes[l];
    o += sizes[l];
  }
  s.key = q.key; s.dkey = q.dkey;
  s.rmult = randint_mult((uint32_t)q.N);
  s.rmagic = randint_magic((uint32_t)q.N);
  s.Xb = w + L.xb;
  s.lab = reinterpret_cast<int32_t*>(w + L.lab);
  s.h1 = w + L.h1;
  s.h2 = w + L.h2;
  s.da2 = w + L.da2;
  s.da1 = w + L.da1;
  s.dL = reinterpret_cast<float*>(w + L.dl);
  s.lik = reinterpret_cast<float*>(w + L.lik);
  s.lkey = reinterpret_cast<uint32_t*>(w + L.lkey);
  s.prior_part = reinterpret_cast<float*>(w + L.prior);
  s.ke_part = reinterpret_cast<float*>(w + L.ke);
  s.col_part = reinterpret_cast<float*>(w + L.col);
  s.fin_part = reinterpret_cast<double*>(w + L.fin);
  s.split_part = reinterpret_cast<float*>(w + L.split);
  s.nparts = L.nparts;
  s.U = q.U; s.grad = q.grad;
  s.samples = q.samples; s.samples_U = q.samples_U; s.n_slots = q.n_slots;
  s.eps_table = q.eps_table; s.save_slot = q.save_slot; s.iter = q.iter;
  s.counters = q.counters; s.done = q.done; s.table_len = q.table_len;
  s.rms = q.rms && q.v ? 1 : 0;
  s.update = update ? 1 : 0;
  s.alpha = q.alpha; s.lmbd = q.lmbd; s.temperature = q.temperature;
  s.noise_temperature = q.noise_temperature;
  s.has_prior = q.has_prior;
  const float ps2 = q.prior_scale * q.prior_scale;
  s.prior_inv_var = q.has_prior ? 1.0f / ps2 : 0.0f;
  s.prior_log_norm = q.has_prior ? logf(2.0f * (float)M_PI * ps2) : 0.0f;
  s.mode = update ? q.mode : MLP_SGLD;
  s.mom = q.mom; s.ke = q.ke; s.U_half = q.U_half;
  s.decay_table = q.decay_table; s.noise_table = q.noise_table;
  s.full = 0;
  s.row0 = 0;
  s.fp8 = q.precision == 1 ? 1 : 0;
  s.fp32 = q.precision == 2 ? 1 : 0;
  const int trc = check_tune(q.D_in, q.H1, q.H2, q.n, q.precision, q.tune);
  if (trc != SGMC_OK) return trc;
  plan_gemms(q.D_in, q.H1, q.H2, q.n, q.C, q.precision, update, q.tune, s.plan);
  s.q = q.q;
  s.shadow8 = q.shadow8;
  if (s.fp8)
    SGMC_REQUIRE(q.q && q.shadow8, SGMC_E_ARG, "sgmc_mlp: fp8 precision needs q and shadow8");
  return SGMC_OK;
}

// GEMM operand of weight leaf l (W1, W2): the e4m3 or bf16 shadow, or theta itself.
const void* gemm_weight(const Step& s, const sgmc_mlp_state& q, int l) {
  if (s.fp32) return q.theta + s.off[l];
  if (s.fp8) return q.shadow8 + s.off[l];
  return q.shadow + s.off[l];
}

// The head over rows [0, n) of every chain: W3 through LDS while it fits in 64 KB; four
// rows per wavefront when that still leaves two blocks per CU (measured: 3072-1024-1024,
// 8192 rows, 46.8 against 55.4 us), else one (3072-512-512, 4096 rows: 17.3 against 31.1).
void launch_head(const Step& s, hipStream_t st) {
  const size_t lds = (size_t)s.H2 * kW3Stride * sizeof(float);
  const int rows_per_block4 = (GT / kWave) * 4;
  const bool four = (int64_t)s.C * s.n / rows_per_block4 >= 512;
  const int per_block = (GT / kWave) * (four ? 4 : 1);
  const dim3 grid(s.n / per_block, s.C);
  if (lds <= 65536) {
    if (four) hipLaunchKernelGGL((head_kernel<true, 4>), grid, dim3(GT), lds, st, s);
    else hipLaunchKernelGGL((head_kernel<true, 1>), grid, dim3(GT), lds, st, s);
  } else {
    if (four) hipLaunchKernelGGL((head_kernel<false, 4>), grid, dim3(GT), 0, st, s);
    else hipLaunchKernelGGL((head_kernel<false, 1>), grid, dim3(GT), 0, st, s);
  }
}

// gather, forward GEMMs and the head (likelihood, and its backward unless s.full).
void launch_forward(const Step& s, const sgmc_mlp_state& q, hipStream_t st_) {
  const int C = q.C, n = q.n;
  const dim3 rows_grid((n + 3) / 4, C);
  hipLaunchKernelGGL(gather_kernel, rows_grid, dim3(GT), 0, st_, s);
  const void* W1 = gemm_weight(s, q, 0);
  const void* W2 = gemm_weight(s, q, 1);
  Gemm g1{s.Xb, q.D_in, (int64_t)n * q.D_in, W1, q.H1, q.P, n, q.H1, q.D_in,
          s.h1, q.H1, (int64_t)n * q.H1, q.theta + s.off[3], q.P, nullptr, 0, 0, 0, 0};
  launch_gemm<true, false, EPI_BIAS_RELU>(dim3(n / BM, q.H1 / BN, C), with_q(g1, Q_X, Q_W1, Q_H1),
                                          s, st_, SLOT_G1);
  Gemm g2{s.h1, q.H1, (int64_t)n * q.H1, W2, q.H2, q.P, n, q.H2, q.H1,
          s.h2, q.H2, (int64_t)n * q.H2, q.theta + s.off[4], q.P, nullptr, 0, 0, 0, 0};
  launch_gemm<true, false, EPI_BIAS_RELU>(dim3(n / BM, q.H2 / BN, C), with_q(g2, Q_H1, Q_W2, -1),
                                          s, st_, SLOT_G2);
  launch_head(s, st_);
}

}  // namespace

extern "C" {

int sgmc_mlp_gemm_plan(int D_in, int H1, int H2, int n, int C, int precision, int update,
                       const int32_t* tune, int32_t* plan) {
  SGMC_REQUIRE(plan && C >= 1 && n >= BM && n % BM == 0 && D_in % BM == 0 && H1 % BN == 0 &&
                   H2 % BN == 0 && precision >= 0 && precision <= 2,
               SGMC_E_ARG, "sgmc_mlp_gemm_plan: bad shape or precision");
  const int rc = check_tune(D_in, H1, H2, n, precision, tune);
  if (rc != SGMC_OK) return rc;
  int p[kSlots];
  plan_gemms(D_in, H1, H2, n, C, precision, update, tune, p);
  for (int i = 0; i < kSlots; ++i) plan[i] = p[i];
  return SGMC_OK;
}

int sgmc_mlp_step(const sgmc_mlp_state* st, int update, void* stream) {
  Layout L;
  Step s;
  const int rc = make_step(st, update, L, s);
  if (rc != SGMC_OK) return rc;
  const sgmc_mlp_state& q = *st;
  hipStream_t st_ = as_stream(stream);
  const int C = q.C, n = q.n;
  launch_forward(s, q, st_);
  const void* W2 = gemm_weight(s, q, 1);
  // dA1 = (dA2 W2^T) * (H1 > 0): B[k=b][n=a] = W2[a][b], k-contiguous rows of W2.
  Gemm g3{s.da2, q.H2, (int64_t)n * q.H2, W2, q.H2, q.P, n, q.H1, q.H2,
          s.da1, q.H1, (int64_t)n * q.H1, nullptr, 0, s.h1, q.H1, (int64_t)n * q.H1, 0, 0};
  launch_gemm<true, true, EPI_MASK>(dim3(n / BM, q.H1 / BN, C), with_q(g3, Q_DA2, Q_W2, Q_DA1), s,
                                    st_, SLOT_G3);
  // gW2 = H1^T dA2: A[m=a][k=i] = H1[i][a] (m-contiguous), B[k=i][n=b] = dA2[i][b].
  Gemm g4{s.h1, q.H1, (int64_t)n * q.H1, s.da2, q.H2, (int64_t)n * q.H2, q.H1, q.H2, n,
          nullptr, 0, 0, nullptr, 0, nullptr, 0, 0, 1, 0};
  // gW1 = Xb^T dA1.
  Gemm g5{s.Xb, q.D_in, (int64_t)n * q.D_in, s.da1, q.H1, (int64_t)n * q.H1, q.D_in, q.H1, n,
          nullptr, 0, 0, nullptr, 0, nullptr, 0, 0, 0, L.g4_blocks};
  g4 = with_q(g4, Q_H1, Q_DA2, -1);
  g5 = with_q(g5, Q_X, Q_DA1, -1);
  const dim3 r4(q.H1 / BM, q.H2 / BN, C), r5(q.D_in / BM, q.H1 / BN, C);
  if (update) launch_param_gemms<EPI_UPDATE>(g4, r4, g5, r5, s, st_);
  else launch_param_gemms<EPI_GRAD>(g4, r4, g5, r5, s, st_);
  const dim3 col_grid((q.H2 + GT - 1) / GT + (q.H1 + GT - 1) / GT + 1, n / kStrip, C);
  if (s.fp32) hipLaunchKernelGGL(colsum_kernel<P_F32>, col_grid, dim3(GT), 0, st_, s);
  else if (s.fp8) hipLaunchKernelGGL(colsum_kernel<P_FP8>, col_grid, dim3(GT), 0, st_, s);
  else hipLaunchKernelGGL(colsum_kernel<P_BF16>, col_grid, dim3(GT), 0, st_, s);
  hipLaunchKernelGGL(small_kernel, dim3(L.small_blocks, C), dim3(GT), 0, st_, s,
                     L.g4_blocks + L.g5_blocks);
  if (update && s.mode == MLP_SGLD && q.samples && q.save_slot) {
    hipLaunchKernelGGL(save_kernel, dim3(256, C), dim3(GT), 0, st_, s, 1);
  }
  return check_launch("sgmc_mlp_step");
}

int sgmc_mlp_init_fp8(const sgmc_mlp_state* st, void* stream) {
  Layout L;
  Step s;
  const int rc = make_step(st, 1, L, s);
  if (rc != SGMC_OK) return rc;
  SGMC_REQUIRE(s.fp8, SGMC_E_ARG, "sgmc_mlp_init_fp8: state precision is not fp8");
  hipStream_t st_ = as_stream(stream);
  const int nq = st->C * kQWords;
  hipLaunchKernelGGL(q_zero_kernel, dim3((nq + GT - 1) / GT), dim3(GT), 0, st_, s.q, nq);
  hipLaunchKernelGGL(w_amax_kernel, dim3(512, st->C), dim3(GT), 0, st_, s);
  hipLaunchKernelGGL(q_init_kernel, dim3((st->C + GT - 1) / GT), dim3(GT), 0, st_, s);
  hipLaunchKernelGGL(w_quant_kernel, dim3(512, st->C), dim3(GT), 0, st_, s);
  return check_launch("sgmc_mlp_init_fp8");
}

int sgmc_mlp_full_potential(const sgmc_mlp_state* st, double* out, void* stream) {
  Layout L;
  Step s;
  const int rc = make_step(st, 1, L, s);
  if (rc != SGMC_OK) return rc;
  SGMC_REQUIRE(out, SGMC_E_ARG, "sgmc_mlp_full_potential: null output");
  const sgmc_mlp_state& q = *st;
  hipStream_t st_ = as_stream(stream);
  double* acc = reinterpret_cast<double*>(static_cast<uint8_t*>(q.work) + L.facc);
  hipLaunchKernelGGL(zero_d_kernel, dim3((q.C + GT - 1) / GT), dim3(GT), 0, st_, acc, q.C);
  s.full = 1;
  for (int64_t r0 = 0; r0 < q.N; r0 += q.n) {
    s.row0 = r0;
    launch_forward(s, q, st_);
    hipLaunchKernelGGL(full_acc_kernel, dim3(q.C), dim3(GT), 0, st_, s, acc);
  }
  hipLaunchKernelGGL(full_final_kernel, dim3(kFinalBlocks, q.C), dim3(GT), 0, st_, s, acc, out);
  return check_launch("sgmc_mlp_full_potential");
}

int sgmc_mlp_mh(const sgmc_mlp_state* st, void* stream) {
  Layout L;
  Step s;
  const int rc = make_step(st, 1, L, s);
  if (rc != SGMC_OK) return rc;
  const sgmc_mlp_state& q = *st;
  SGMC_REQUIRE(q.mom && q.ke && q.U_full && q.U_prop && q.theta_old && q.mom_old && q.U_old &&
                   q.mh_key,
               SGMC_E_ARG, "sgmc_mlp_mh: null pointer in state");
  SGMC_REQUIRE(q.noise_temperature != 0.0f, SGMC_E_ARG, "sgmc_mlp_mh: zero temperature");
  s.mom = q.mom;
  s.ke = q.ke;
  MH m{q.U_full, q.U_prop, q.theta_old, q.mom_old, q.U_old, q.mh_key, q.mh_stats, q.accepted};
  const int64_t blocks = std::min<int64_t>((q.P + GT - 1) / GT, 1024);
  hipLaunchKernelGGL(mh_kernel, dim3((unsigned)blocks, q.C), dim3(GT), 0, as_stream(stream), s,
                     m);
  return check_launch("sgmc_mlp_mh");
}

}  // extern "C"
