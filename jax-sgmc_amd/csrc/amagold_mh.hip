// AMAGOLD's amortized Metropolis-Hastings test (solver.py:366-419) for the chains of a
// fused chain table, run after the last leapfrog launch of an iteration
// (it = iter[0] - 1) and after the full-data potential of every proposal:
//   (key', sub) = split(key); u = uniform(sub);
//   log a = min(0, U_full - U_prop + energy)      (FP64; NaN stays NaN and rejects)
//   accept iff log u < log a: U_full = U_prop and the kept energy becomes the
//   proposal's; otherwise theta and the momentum return to the iteration's start with
//   the momentum flipped (solver.py:390-400).
// The post-test position is saved at save_slot[it] with the kept energy and exp(log a).
//
// SGGMC's test (solver.py:501-567) for chains stepped in the OBABO proposal mode:
//   (key', sub) = split(key); u = uniform(sub);
//   log a = -1/T (U_prop - U_full + KE_end - KE_start), replaced by 0 unless <= 0
//   (FP64; NaN becomes 0, as jnp.where(log_alpha <= 0, log_alpha, 0.0) does);
//   accept iff log u < log a: U_full = U_prop; otherwise theta, the momentum and the
//   minibatch potential return to the iteration's start (the momentum not flipped).
// Both kinetic energies restart at zero, accepted or not (solver.py:539-549).
#include <math.h>

#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {

constexpr int kMhThreads = 256;

__global__ void __launch_bounds__(kMhThreads) amagold_mh_kernel(
    const sgmc_glm_chain* __restrict__ chains, int nchains, const double* __restrict__ U_prop,
    double* __restrict__ U_full, float* __restrict__ energy_kept, uint32_t* __restrict__ keys,
    float* __restrict__ ratio_samples, int n_slots, const int32_t* __restrict__ save_slot,
    int table_len, const int32_t* __restrict__ iter) {
  __shared__ int s_accept;
  const int c = blockIdx.x;
  if (c >= nchains) return;
  const sgmc_glm_chain& ch = chains[c];
  if (iter[0] < 1) return;
  const int it = min(iter[0] - 1, table_len - 1);
  const int slot = save_slot ? save_slot[it] : -1;
  if (threadIdx.x == 0) {
    Key next, sub;
    split2(load_key(keys + 2 * c), next, sub);
    store_key(keys + 2 * c, next);
    const float u = bits_to_uniform(bits_word(sub, 1u, 0u), 0.0f, 1.0f);
    double log_alpha = U_full[c] - U_prop[c] + (double)ch.energy[0];
    if (log_alpha > 0.0) log_alpha = 0.0;
    const bool acc = log((double)u) < log_alpha;
    if (acc) {
      U_full[c] = U_prop[c];
      energy_kept[c] = ch.energy[0];
    }
    if (slot >= 0 && slot < n_slots) {
      if (ratio_samples) ratio_samples[(int64_t)c * n_slots + slot] = (float)exp(log_alpha);
      if (ch.samples_U) ch.samples_U[slot] = energy_kept[c];
    }
    s_accept = acc ? 1 : 0;
  }
  __syncthreads();
  const bool acc = s_accept != 0;
  const bool save = slot >= 0 && slot < n_slots && ch.samples;
  for (int j = threadIdx.x; j < ch.D; j += kMhThreads) {
    float w = ch.theta[j];
    if (!acc) {
      w = ch.theta_old[j];
      ch.theta[j] = w;
      ch.mom[j] = -1.0f * ch.mom_old[j];
    }
    if (save) ch.samples[(int64_t)slot * ch.D + j] = w;
  }
}

__global__ void __launch_bounds__(kMhThreads) sggmc_mh_kernel(
    const sgmc_glm_chain* __restrict__ chains, int nchains, const double* __restrict__ U_prop,
    double* __restrict__ U_full, uint32_t* __restrict__ keys, float* __restrict__ ratio_samples,
    float* __restrict__ ke_samples, double* __restrict__ U_samples, int n_slots,
    const int32_t* __restrict__ save_slot, int table_len, const int32_t* __restrict__ iter) {
  __shared__ int s_accept;
  const int c = blockIdx.x;
  if (c >= nchains) return;
  const sgmc_glm_chain& ch = chains[c];
  if (iter[0] < 1) return;
  const int it = min(iter[0] - 1, table_len - 1);
  const int slot = save_slot ? save_slot[it] : -1;
  const bool in_slots = slot >= 0 && slot < n_slots;
  if (threadIdx.x == 0) {
    Key next, sub;
    split2(load_key(keys + 2 * c), next, sub);
    store_key(keys + 2 * c, next);
    const float u = bits_to_uniform(bits_word(sub, 1u, 0u), 0.0f, 1.0f);
    const float ks = ch.energy[0];
    const float ke = ch.energy[1];
    const double d = U_prop[c] - U_full[c] + (double)ke - (double)ks;
    double log_alpha = (-1.0 / (double)ch.temperature) * d;
    if (!(log_alpha <= 0.0)) log_alpha = 0.0;
    const bool acc = log((double)u) < log_alpha;
    if (acc) {
      U_full[c] = U_prop[c];
    } else {
      ch.U[0] = ch.energy[2];
    }
    ch.energy[0] = 0.0f;
    ch.energy[1] = 0.0f;
    if (in_slots) {
      const int64_t o = (int64_t)c * n_slots + slot;
      if (ratio_samples) ratio_samples[o] = (float)exp(log_alpha);
      if (ke_samples) ke_samples[o] = ke - ks;
      if (U_samples) U_samples[o] = U_full[c];
      if (ch.samples_U) ch.samples_U[slot] = ch.U[0];
    }
    s_accept = acc ? 1 : 0;
  }
  __syncthreads();
  const bool acc = s_accept != 0;
  const bool save = in_slots && ch.samples;
  for (int j = threadIdx.x; j < ch.D; j += kMhThreads) {
    float w = ch.theta[j];
    if (!acc) {
      w = ch.theta_old[j];
      ch.theta[j] = w;
      ch.mom[j] = ch.mom_old[j];
    }
    if (save) ch.samples[(int64_t)slot * ch.D + j] = w;
  }
}

}  // namespace sgmc

using namespace sgmc;

extern "C" {

int sgmc_glm_amagold_mh(const sgmc_glm_chain* chains, int nchains, const double* U_prop,
                        double* U_full, float* energy_kept, uint32_t* keys, float* ratio_samples,
                        int n_slots, const int32_t* save_slot, int table_len,
                        const int32_t* iter, void* stream) {
  SGMC_REQUIRE(chains && U_prop && U_full && energy_kept && keys && iter, SGMC_E_ARG,
               "sgmc_glm_amagold_mh: null pointer");
  SGMC_REQUIRE(nchains >= 1 && table_len >= 1 && n_slots >= 0, SGMC_E_ARG,
               "sgmc_glm_amagold_mh: bad sizes");
  hipLaunchKernelGGL(amagold_mh_kernel, dim3(nchains), dim3(kMhThreads), 0, as_stream(stream),
                     chains, nchains, U_prop, U_full, energy_kept, keys, ratio_samples, n_slots,
                     save_slot, table_len, iter);
  return check_launch("sgmc_glm_amagold_mh");
}

int sgmc_glm_sggmc_mh(const sgmc_glm_chain* chains, int nchains, const double* U_prop,
                      double* U_full, uint32_t* keys, float* ratio_samples, float* ke_samples,
                      double* U_samples, int n_slots, const int32_t* save_slot, int table_len,
                      const int32_t* iter, void* stream) {
  SGMC_REQUIRE(chains && U_prop && U_full && keys && iter, SGMC_E_ARG,
               "sgmc_glm_sggmc_mh: null pointer");
  SGMC_REQUIRE(nchains >= 1 && table_len >= 1 && n_slots >= 0, SGMC_E_ARG,
               "sgmc_glm_sggmc_mh: bad sizes");
  hipLaunchKernelGGL(sggmc_mh_kernel, dim3(nchains), dim3(kMhThreads), 0, as_stream(stream),
                     chains, nchains, U_prop, U_full, keys, ratio_samples, ke_samples, U_samples,
                     n_slots, save_slot, table_len, iter);
  return check_launch("sgmc_glm_sggmc_mh");
}

}  // extern "C"
