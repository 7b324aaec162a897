// Minibatch index draw and row gather (DeviceNumpyDataLoader.get_random_data,
// data/numpy_loader.py:128-141, and tree_index, data/core.py:642-660).
#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {

constexpr int kNT = 256;

// (dkey_out, sub) = split(dkey_in); idx = randint(sub, (n,), 0, N).
__global__ void __launch_bounds__(kNT) draw_batch_kernel(const uint32_t* __restrict__ dkey_in,
                                                         uint32_t* __restrict__ dkey_out,
                                                         int32_t* __restrict__ idx, uint32_t n,
                                                         uint32_t span, uint32_t mult,
                                                         uint32_t magic) {
  Key next, sub, k1, k2;
  split2(load_key(dkey_in), next, sub);
  split2(sub, k1, k2);
  const uint32_t m = (n + 1) >> 1;
  const uint32_t j = blockIdx.x * kNT + threadIdx.x;
  if (j < m) {
    uint32_t h0, h1, l0, l1;
    bits_pair(k1, n, j, h0, h1);
    bits_pair(k2, n, j, l0, l1);
    idx[j] = (int32_t)randint_offset(h0, l0, span, mult, magic);
    if (j + m < n) idx[j + m] = (int32_t)randint_offset(h1, l1, span, mult, magic);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) store_key(dkey_out, next);
}

// One wavefront per row; rows of any byte length, 16-byte vectors when aligned.
__global__ void __launch_bounds__(kNT) gather16_kernel(const uint4* __restrict__ src,
                                                       const int32_t* __restrict__ idx,
                                                       uint4* __restrict__ dst, int n,
                                                       int64_t row_vecs, int64_t src_stride_vecs,
                                                       int64_t src_rows) {
  const int row = blockIdx.x * (kNT / kWave) + threadIdx.x / kWave;
  if (row >= n) return;
  if ((uint64_t)(int64_t)idx[row] >= (uint64_t)src_rows) return;
  const int lane = threadIdx.x & (kWave - 1);
  const uint4* s = src + (int64_t)idx[row] * src_stride_vecs;
  uint4* d = dst + (int64_t)row * row_vecs;
  for (int64_t c = lane; c < row_vecs; c += kWave) d[c] = s[c];
}

__global__ void __launch_bounds__(kNT) gather1_kernel(const uint8_t* __restrict__ src,
                                                      const int32_t* __restrict__ idx,
                                                      uint8_t* __restrict__ dst, int n,
                                                      int64_t row_bytes, int64_t src_stride,
                                                      int64_t src_rows) {
  const int row = blockIdx.x * (kNT / kWave) + threadIdx.x / kWave;
  if (row >= n) return;
  if ((uint64_t)(int64_t)idx[row] >= (uint64_t)src_rows) return;
  const int lane = threadIdx.x & (kWave - 1);
  const uint8_t* s = src + (int64_t)idx[row] * src_stride;
  uint8_t* d = dst + (int64_t)row * row_bytes;
  for (int64_t c = lane; c < row_bytes; c += kWave) d[c] = s[c];
}

}  // namespace sgmc

using namespace sgmc;

extern "C" {

int sgmc_draw_batch(const uint32_t* dkey_in, uint32_t* dkey_out, int32_t* idx, int n,
                    int64_t N, void* stream) {
  SGMC_REQUIRE(dkey_in && dkey_out && idx, SGMC_E_ARG, "sgmc_draw_batch: null pointer");
  SGMC_REQUIRE(dkey_in != dkey_out, SGMC_E_ARG,
               "sgmc_draw_batch: dkey_in and dkey_out must not alias");
  SGMC_REQUIRE(n >= 1, SGMC_E_ARG, "sgmc_draw_batch: batch size must be positive, got %d", n);
  SGMC_REQUIRE(N >= 1 && N <= 2147483647LL, SGMC_E_ARG,
               "sgmc_draw_batch: observation count %lld out of int32 range", (long long)N);
  const uint32_t span = (uint32_t)N;
  const uint32_t m = ((uint32_t)n + 1) >> 1;
  hipLaunchKernelGGL(draw_batch_kernel, dim3((m + kNT - 1) / kNT), dim3(kNT), 0,
                     as_stream(stream), dkey_in, dkey_out, idx, (uint32_t)n, span,
                     randint_mult(span), randint_magic(span));
  return check_launch("sgmc_draw_batch");
}

int sgmc_gather_rows(const void* src, int64_t src_rows, const int32_t* idx, void* dst, int n,
                     int64_t row_bytes, int64_t src_stride_bytes, void* stream) {
  SGMC_REQUIRE(src && idx && dst, SGMC_E_ARG, "sgmc_gather_rows: null pointer");
  SGMC_REQUIRE(n >= 0 && row_bytes >= 0 && src_stride_bytes >= row_bytes, SGMC_E_ARG,
               "sgmc_gather_rows: bad shape");
  if (n == 0 || row_bytes == 0) return SGMC_OK;
  const int rows_per_block = kNT / kWave;
  dim3 grid((n + rows_per_block - 1) / rows_per_block);
  const bool vec = ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 16 == 0) &&
                   (row_bytes % 16 == 0) && (src_stride_bytes % 16 == 0);
  if (vec) {
    hipLaunchKernelGGL(gather16_kernel, grid, dim3(kNT), 0, as_stream(stream),
                       (const uint4*)src, idx, (uint4*)dst, n, row_bytes / 16,
                       src_stride_bytes / 16, src_rows);
  } else {
    hipLaunchKernelGGL(gather1_kernel, grid, dim3(kNT), 0, as_stream(stream),
                       (const uint8_t*)src, idx, (uint8_t*)dst, n, row_bytes, src_stride_bytes, src_rows);
  }
  return check_launch("sgmc_gather_rows");
}

}  // extern "C"
