// Replica exchange between chains of a fused chain table: the swap step of
// parallel_tempering (solver.py:264-293), generalized to a ladder of adjacent pairs
// tested in alternating even / odd rounds.
#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {

constexpr int kSwapThreads = 256;

__global__ void __launch_bounds__(kSwapThreads) swap_kernel(const sgmc_glm_chain* __restrict__ chains,
                                                   int nchains,
                                                   const sgmc_swap_pair* __restrict__ pairs,
                                                   int npairs, float F, int swap_every,
                                                   const int32_t* __restrict__ save_slot,
                                                   int table_len,
                                                   const int32_t* __restrict__ iter) {
  __shared__ int s_swap;
  const int p = blockIdx.x;
  if (p >= npairs) return;
  const sgmc_swap_pair& pr = pairs[p];
  if (pr.lo < 0 || pr.lo >= nchains || pr.hi < 0 || pr.hi >= nchains) return;
  const sgmc_glm_chain& a = chains[pr.lo];
  const sgmc_glm_chain& b = chains[pr.hi];
  if (iter[0] < 1) return;  // no iteration has completed yet
  const int it = min(iter[0] - 1, table_len - 1);
  const int done = it + 1;
  const bool round = done % swap_every == 0;
  const int parity = (done / swap_every - 1) & 1;
  const bool tested = round && (pr.parity < 0 || pr.parity == parity);

  if (threadIdx.x == 0) {
    int sw = 0;
    if (tested) {
      // solver.py:275-277: step-size-1/step running average of the lo chain's var(lik)
      const int step = pr.step[0] + 1;
      pr.step[0] = step;
      const float lr = 1.0f / (float)step;
      const float ssq = (1.0f - lr) * pr.ssq[0] + lr * a.var[0];
      pr.ssq[0] = ssq;
      // solver.py:280-286
      const float temps = 1.0f / pr.t_lo - 1.0f / pr.t_hi;
      const float corr = temps * ssq / F;
      const float log_s = temps * ((a.U[0] - b.U[0]) - corr);
      Key next, sub;
      split2(load_key(pr.key), next, sub);
      store_key(pr.key, next);
      const float log_u = logf_cephes(bits_to_uniform(bits_word(sub, 1u, 0u), 0.0f, 1.0f));
      // solver.py:288-291: kept when log u < log s, exchanged otherwise
      sw = (log_u < log_s) ? 0 : 1;
      if (sw) {
        const float tU = a.U[0];
        a.U[0] = b.U[0];
        b.U[0] = tU;
        if (a.var && b.var) {
          const float tv = a.var[0];
          a.var[0] = b.var[0];
          b.var[0] = tv;
        }
        const Key ka = load_key(a.key), kb = load_key(b.key);
        store_key(a.key, kb);
        store_key(b.key, ka);
        const Key da = load_key(a.dkey), db = load_key(b.dkey);
        store_key(a.dkey, db);
        store_key(b.dkey, da);
        if (pr.swaps) pr.swaps[0] += 1;
      }
    }
    s_swap = sw;
  }
  __syncthreads();
  const bool sw = s_swap != 0;
  const int slot = (pr.samples && save_slot) ? save_slot[it] : -1;
  if (!sw && slot < 0) return;
  const int D = a.D;
  for (int j = threadIdx.x; j < D; j += kSwapThreads) {
    float x = a.theta[j];
    if (sw) {
      const float y = b.theta[j];
      a.theta[j] = y;
      b.theta[j] = x;
      x = y;
      if (a.mom && b.mom) {
        const float m = a.mom[j];
        a.mom[j] = b.mom[j];
        b.mom[j] = m;
      }
      if (a.v && b.v) {
        const float m = a.v[j];
        a.v[j] = b.v[j];
        b.v[j] = m;
      }
    }
    if (slot >= 0) pr.samples[(int64_t)slot * D + j] = x;
  }
  if (slot >= 0 && pr.samples_U && threadIdx.x == 0) pr.samples_U[slot] = a.U[0];
}

}  // namespace sgmc

using namespace sgmc;

extern "C" {

int sgmc_resgld_swap(const sgmc_glm_chain* chains, int nchains, const sgmc_swap_pair* pairs,
                     int npairs, float F, int swap_every, const int32_t* save_slot,
                     int table_len, const int32_t* iter, void* stream) {
  SGMC_REQUIRE(chains && pairs && iter, SGMC_E_ARG, "sgmc_resgld_swap: null pointer");
  SGMC_REQUIRE(nchains >= 2 && npairs >= 1, SGMC_E_ARG,
               "sgmc_resgld_swap: need >= 2 chains and >= 1 pair");
  SGMC_REQUIRE(swap_every >= 1 && table_len >= 1, SGMC_E_ARG,
               "sgmc_resgld_swap: swap_every and table_len must be >= 1");
  hipLaunchKernelGGL(swap_kernel, dim3(npairs), dim3(kSwapThreads), 0, as_stream(stream), chains,
                     nchains, pairs, npairs, F, swap_every, save_slot, table_len, iter);
  return check_launch("sgmc_resgld_swap");
}

}  // extern "C"
