// jax.random (legacy Threefry-2x32 derivation) on the device, behind the C-ABI.
// Each kernel thread evaluates one Threefry counter pair and writes the two words
// (or values) that land at positions j and j + ceil(S/2).
#include <stdarg.h>
#include <string.h>

#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

constexpr int kNT = 256;

__global__ void __launch_bounds__(kNT) split_kernel(const uint32_t* __restrict__ key,
                                                    uint32_t* __restrict__ out, uint32_t num) {
  const uint32_t q = blockIdx.x * kNT + threadIdx.x;
  if (q >= 2 * num) return;
  out[q] = split_word(load_key(key), num, q);
}

enum class Dist { kBits, kUniform, kNormal };

template <Dist D>
__global__ void __launch_bounds__(kNT) draw_kernel(const uint32_t* __restrict__ key, void* out,
                                                   uint32_t S, float lo, float hi) {
  const uint32_t m = (S + 1) >> 1;
  const uint32_t j = blockIdx.x * kNT + threadIdx.x;
  if (j >= m) return;
  uint32_t w0, w1;
  bits_pair(load_key(key), S, j, w0, w1);
  if (D == Dist::kBits) {
    uint32_t* o = static_cast<uint32_t*>(out);
    o[j] = w0;
    if (j + m < S) o[j + m] = w1;
  } else if (D == Dist::kUniform) {
    float* o = static_cast<float*>(out);
    o[j] = bits_to_uniform(w0, lo, hi);
    if (j + m < S) o[j + m] = bits_to_uniform(w1, lo, hi);
  } else {
    float* o = static_cast<float*>(out);
    o[j] = bits_to_normal(w0);
    if (j + m < S) o[j + m] = bits_to_normal(w1);
  }
}

__global__ void __launch_bounds__(kNT) randint_kernel(const uint32_t* __restrict__ key,
                                                      int32_t* __restrict__ out, uint32_t S,
                                                      int32_t minval, uint32_t span, uint32_t mult,
                                                      uint32_t magic) {
  const uint32_t m = (S + 1) >> 1;
  const uint32_t j = blockIdx.x * kNT + threadIdx.x;
  if (j >= m) return;
  Key k1, k2;
  split2(load_key(key), k1, k2);
  uint32_t h0, h1, l0, l1;
  bits_pair(k1, S, j, h0, h1);
  bits_pair(k2, S, j, l0, l1);
  out[j] = minval + (int32_t)randint_offset(h0, l0, span, mult, magic);
  if (j + m < S) out[j + m] = minval + (int32_t)randint_offset(h1, l1, span, mult, magic);
}

// random_tree: keys = split(key, L); leaf l gets normal(keys[l], size_l).
// Table layout (int64): [L, B, (offset, size) x L, (leaf, first_pair) x B].
__global__ void __launch_bounds__(kNT) tree_normal_kernel(const uint32_t* __restrict__ key,
                                                          const int64_t* __restrict__ table,
                                                          float* __restrict__ out) {
  const int64_t L = table[0];
  const int64_t* leaves = table + 2;
  const int64_t* blocks = leaves + 2 * L;
  const int64_t leaf = blocks[2 * blockIdx.x];
  const int64_t p0 = blocks[2 * blockIdx.x + 1];
  const int64_t off = leaves[2 * leaf];
  const uint32_t S = (uint32_t)leaves[2 * leaf + 1];
  const uint32_t m = (S + 1) >> 1;
  const Key kl = split_key(load_key(key), (uint32_t)L, (uint32_t)leaf);
  const uint32_t pend = (uint32_t)min<int64_t>(p0 + SGMC_PAIRS_PER_BLOCK, m);
  for (uint32_t j = (uint32_t)p0 + threadIdx.x; j < pend; j += kNT) {
    uint32_t w0, w1;
    bits_pair(kl, S, j, w0, w1);
    out[off + j] = bits_to_normal(w0);
    if (j + m < S) out[off + j + m] = bits_to_normal(w1);
  }
}

}  // namespace sgmc

using namespace sgmc;

extern "C" {

const char* sgmc_last_error(void) { return g_err; }

int sgmc_abi_version(void) { return SGMC_ABI_VERSION; }

int sgmc_stream_create(void** stream) {
  SGMC_REQUIRE(stream, SGMC_E_ARG, "sgmc_stream_create: null pointer");
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  SGMC_REQUIRE(e == hipSuccess && cus > 0 && cus <= 1024, SGMC_E_HIP,
               "sgmc_stream_create: device query failed (%s)", hipGetErrorString(e));
  // every CU: the mask gives the stream a hardware queue of its own without
  // restricting where its kernels run
  uint32_t mask[32];
  const int words = (cus + 31) / 32;
  for (int i = 0; i < words; ++i) {
    const int bits = cus - 32 * i;
    mask[i] = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1u);
  }
  hipStream_t s = nullptr;
  e = hipExtStreamCreateWithCUMask(&s, (uint32_t)words, mask);
  SGMC_REQUIRE(e == hipSuccess, SGMC_E_HIP, "sgmc_stream_create: %s", hipGetErrorString(e));
  *stream = s;
  return SGMC_OK;
}

int sgmc_stream_destroy(void* stream) {
  SGMC_REQUIRE(stream, SGMC_E_ARG, "sgmc_stream_destroy: null stream");
  const hipError_t e = hipStreamDestroy(static_cast<hipStream_t>(stream));
  SGMC_REQUIRE(e == hipSuccess, SGMC_E_HIP, "sgmc_stream_destroy: %s", hipGetErrorString(e));
  return SGMC_OK;
}

int sgmc_split(const uint32_t* key, uint32_t* out, int num, void* stream) {
  SGMC_REQUIRE(key && out, SGMC_E_ARG, "sgmc_split: null pointer");
  SGMC_REQUIRE(num >= 1, SGMC_E_ARG, "sgmc_split: num must be >= 1, got %d", num);
  const uint32_t words = 2u * (uint32_t)num;
  hipLaunchKernelGGL(split_kernel, dim3((words + kNT - 1) / kNT), dim3(kNT), 0,
                     as_stream(stream), key, out, (uint32_t)num);
  return check_launch("sgmc_split");
}

static int draw_common(Dist d, const uint32_t* key, void* out, int64_t n, float lo, float hi,
                       void* stream, const char* name) {
  SGMC_REQUIRE(key && (out || n == 0), SGMC_E_ARG, "%s: null pointer", name);
  SGMC_REQUIRE(n >= 0 && n < 4294967295LL, SGMC_E_ARG, "%s: size %lld out of range", name,
               (long long)n);
  if (n == 0) return SGMC_OK;
  const uint32_t S = (uint32_t)n;
  const uint32_t m = (S + 1) >> 1;
  dim3 grid((m + kNT - 1) / kNT);
  hipStream_t s = as_stream(stream);
  switch (d) {
    case Dist::kBits:
      hipLaunchKernelGGL(draw_kernel<Dist::kBits>, grid, dim3(kNT), 0, s, key, out, S, lo, hi);
      break;
    case Dist::kUniform:
      hipLaunchKernelGGL(draw_kernel<Dist::kUniform>, grid, dim3(kNT), 0, s, key, out, S, lo, hi);
      break;
    case Dist::kNormal:
      hipLaunchKernelGGL(draw_kernel<Dist::kNormal>, grid, dim3(kNT), 0, s, key, out, S, lo, hi);
      break;
  }
  return check_launch(name);
}

int sgmc_random_bits(const uint32_t* key, uint32_t* out, int64_t n, void* stream) {
  return draw_common(Dist::kBits, key, out, n, 0.f, 1.f, stream, "sgmc_random_bits");
}

int sgmc_uniform(const uint32_t* key, float* out, int64_t n, float minval, float maxval,
                 void* stream) {
  return draw_common(Dist::kUniform, key, out, n, minval, maxval, stream, "sgmc_uniform");
}

int sgmc_normal(const uint32_t* key, float* out, int64_t n, void* stream) {
  return draw_common(Dist::kNormal, key, out, n, 0.f, 1.f, stream, "sgmc_normal");
}

int sgmc_randint(const uint32_t* key, int32_t* out, int64_t n, int32_t minval, int32_t maxval,
                 void* stream) {
  SGMC_REQUIRE(key && (out || n == 0), SGMC_E_ARG, "sgmc_randint: null pointer");
  SGMC_REQUIRE(n >= 0 && n < 4294967295LL, SGMC_E_ARG, "sgmc_randint: size out of range");
  if (n == 0) return SGMC_OK;
  const uint32_t span =
      maxval <= minval ? 1u : (uint32_t)((int64_t)maxval - (int64_t)minval);
  const uint32_t S = (uint32_t)n;
  const uint32_t m = (S + 1) >> 1;
  hipLaunchKernelGGL(randint_kernel, dim3((m + kNT - 1) / kNT), dim3(kNT), 0, as_stream(stream),
                     key, out, S, minval, span, randint_mult(span), randint_magic(span));
  return check_launch("sgmc_randint");
}

int sgmc_tree_normal(const uint32_t* key, const int64_t* leaf_table, int nblocks, float* out,
                     void* stream) {
  SGMC_REQUIRE(key && leaf_table && out, SGMC_E_ARG, "sgmc_tree_normal: null pointer");
  SGMC_REQUIRE(nblocks >= 0, SGMC_E_ARG, "sgmc_tree_normal: nblocks < 0");
  if (nblocks == 0) return SGMC_OK;
  hipLaunchKernelGGL(tree_normal_kernel, dim3(nblocks), dim3(kNT), 0, as_stream(stream), key,
                     leaf_table, out);
  return check_launch("sgmc_tree_normal");
}

}  // extern "C"
