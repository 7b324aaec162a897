/*
 * sgmc.h — C-ABI of the MI355X (gfx950) SG-MCMC hot path.
 *
 * The reference (jax-sgmc 0.1.5) has no FFI: its hot path is Python over JAX. Each
 * entry point below replaces one piece of that path and cites the reference
 * function it stands in for. Conventions:
 *   - every pointer is device memory owned by the caller (PyTorch tensors in the
 *     package's host layer), unless a comment says "host";
 *   - `stream` is a hipStream_t passed as void* (torch's current stream);
 *   - return value 0 = SGMC_OK, otherwise an SGMC_E_* code; sgmc_last_error()
 *     gives the message (thread-local). Nothing aborts, nothing allocates;
 *   - keys are uint32[2] jax PRNG keys (legacy Threefry-2x32 derivation); an
 *     entry point that advances a key reads key_in and writes key_out, which must
 *     not alias (the reference is functional: state in, state out);
 *   - workspaces hold per-block partials and an arrival counter: the caller
 *     zero-fills a workspace once; every call leaves its counter at zero again, so a
 *     workspace can be reused across calls and in captured HIP graphs. Calls that
 *     share one workspace must be ordered on one stream.
 */
#ifndef SGMC_H_
#define SGMC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGMC_ABI_VERSION 2

#define SGMC_OK 0
#define SGMC_E_ARG 1 /* invalid argument: null pointer, shape, alignment, size */
#define SGMC_E_HIP 2 /* HIP runtime error at launch */

/* Counter pairs (two values) covered by one block of a leaf-table kernel. */
#define SGMC_PAIRS_PER_BLOCK 1024

#define SGMC_LIK_LINEAR_GAUSSIAN 0
#define SGMC_LIK_LOGISTIC 1

#define SGMC_MODE_SGLD 0
#define SGMC_MODE_SGLD_RMS 1
#define SGMC_MODE_SGHMC 2
#define SGMC_MODE_AMAGOLD 3
#define SGMC_MODE_OBABO 4
#define SGMC_MODE_SGGMC 5

const char* sgmc_last_error(void);
int sgmc_abi_version(void);

/* A stream of the current device with a hardware queue of its own: created with a CU
 * mask of every CU (hipExtStreamCreateWithCUMask), since streams created otherwise share
 * the process's few hardware queues (GPU_MAX_HW_QUEUES, 4 by default), and kernels of
 * two streams on one queue run one after the other. Default priority. For chain tables
 * run side by side (fused.TableGroup). No counterpart in the reference (XLA schedules
 * its own streams). */
int sgmc_stream_create(void** stream);
int sgmc_stream_destroy(void* stream);

/* ---------------------------------------------------------------- RNG ----
 * jax.random with the original (non-partitionable) Threefry-2x32 derivation, the
 * generator behind every random number of the reference's hot path:
 * integrator.py:119-135 (random_tree), integrator.py:208, 630, 736, 871 (splits),
 * data/numpy_loader.py:132-134 (randint), solver.py:283-284, 528 (uniform). */

/* jax.random.split(key, num): out is uint32[num][2]. */
int sgmc_split(const uint32_t* key, uint32_t* out, int num, void* stream);
/* jax.random.bits(key, (n,)) for 32-bit words (threefry_2x32(key, iota(n))). */
int sgmc_random_bits(const uint32_t* key, uint32_t* out, int64_t n, void* stream);
/* jax.random.uniform(key, (n,), float32, minval, maxval). */
int sgmc_uniform(const uint32_t* key, float* out, int64_t n, float minval, float maxval,
                 void* stream);
/* jax.random.normal(key, (n,), float32). */
int sgmc_normal(const uint32_t* key, float* out, int64_t n, void* stream);
/* jax.random.randint(key, (n,), minval, maxval) as int32. */
int sgmc_randint(const uint32_t* key, int32_t* out, int64_t n, int32_t minval, int32_t maxval,
                 void* stream);
/* random_tree(key, tree) (integrator.py:119-135) over a flattened pytree.
 * leaf_table (device int64): [L, B, (offset, size) x L, (leaf, first_pair) x B], with
 * leaves in tree_flatten order; block b covers SGMC_PAIRS_PER_BLOCK counter pairs of
 * one leaf; nblocks == B. */
int sgmc_tree_normal(const uint32_t* key, const int64_t* leaf_table, int nblocks, float* out,
                     void* stream);

/* --------------------------------------------------------------- data ----
 * DeviceNumpyDataLoader.get_random_data (data/numpy_loader.py:128-141):
 * (dkey_out, sub) = split(dkey_in); idx = randint(sub, (n,), 0, N). */
int sgmc_draw_batch(const uint32_t* dkey_in, uint32_t* dkey_out, int32_t* idx, int n,
                    int64_t N, void* stream);
/* tree_index (data/core.py:642-660) for one leaf: dst[r] = src[idx[r]] for rows of
 * row_bytes bytes, source rows src_stride_bytes apart; rows whose index lies
 * outside [0, src_rows) are left unwritten. */
int sgmc_gather_rows(const void* src, int64_t src_rows, const int32_t* idx, void* dst, int n,
                     int64_t row_bytes, int64_t src_stride_bytes, void* stream);

/* ------------------------------------------------------ GLM potential ----
 * minibatch_potential(prior, likelihood) (potential.py:94-216) evaluated with
 * jax.value_and_grad (integrator.py:166, 593, 792) for the linear-Gaussian and
 * logistic families, rows X[idx[j]] (idx == NULL: rows 0..n-1 of X, i.e. an already
 * gathered batch). X is float32 [N, ldx] with ldx % 4 == 0, ldx >= D rounded up to 4,
 * 16-byte aligned; D <= 1024. Outputs: U (scalar), grad [D], optional lik [n] (the
 * `likelihoods=True` output, potential.py:208-211) and var(lik) (integrator.py:880).
 * workspace: sgmc_glm_workspace_floats(n, D, nblocks) floats; nblocks from
 * sgmc_glm_blocks(n, D) or any value >= 1 with nblocks * 4096 >= n. */
int sgmc_glm_blocks(int n, int D);
int64_t sgmc_glm_workspace_floats(int n, int D, int nblocks);
int sgmc_glm_potential(int family, const float* X, int64_t ldx, const float* y,
                       const int32_t* idx, int n, int64_t N, int D, const float* theta,
                       float sigma, int has_prior, float prior_scale, float temperature,
                       float* out_U, float* out_grad, float* out_lik, float* out_var,
                       float* workspace, int64_t workspace_floats, int nblocks, void* stream);

/* sgmc_glm_potential without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). It has no counterpart in the reference, whose potentials are jax
 * functions; the arguments and results are sgmc_glm_potential's. The block records
 * and the arrival counter live in a static 1 MiB device array per GPU, so nblocks *
 * (4 ceil(D/4) + 2) must not exceed 262,080 floats (D = 1024: 255 blocks; D = 512: 509);
 * larger calls return SGMC_E_ARG. No allocation happens (the first call on a device
 * looks the array up and zeroes its counter on `stream`). Calls on one GPU share the
 * array: calls that overlap in time on one GPU (other streams, other threads) race. */
int sgmc_glm_potential_noalloc(int family, const float* X, int64_t ldx, const float* y,
                               const int32_t* idx, int n, int64_t N, int D, const float* theta,
                               float sigma, int has_prior, float prior_scale, float temperature,
                               float* out_U, float* out_grad, float* out_lik, float* out_var,
                               int nblocks, void* stream);

/* full_potential (potential.py:219-293) over all N rows with FP64 accumulation of the
 * per-row likelihoods and the prior; out_U is one double. workspace:
 * sgmc_glm_full_workspace_doubles(N, rows_per_block) doubles. */
int64_t sgmc_glm_full_workspace_doubles(int64_t N, int rows_per_block);
int sgmc_glm_full_potential(int family, const float* X, int64_t ldx, const float* y, int64_t N,
                            int D, const float* theta, float sigma, int has_prior,
                            float prior_scale, double temperature, double* out_U,
                            double* workspace, int64_t workspace_doubles, int rows_per_block,
                            void* stream);

/* full_potential (potential.py:219-293) of nchains chains sharing one dataset, chain c's
 * position at thetas + c * ld_theta: out_U[c] = (-sum_i lik_i(theta_c) - log p) / T,
 * each sum over rows accumulated in FP64. The row dot products of all chains run as one
 * fp32 matrix product, reading X once per 64 chains. nblocks: grid size (one block per
 * compute unit is the intended shape); workspace:
 * sgmc_glm_full_chains_workspace_doubles(nblocks) doubles whose counter word starts at
 * zero (left at zero). */
int64_t sgmc_glm_full_chains_workspace_doubles(int nblocks);
int sgmc_glm_full_potential_chains(int family, const float* X, int64_t ldx, const float* y,
                                   int64_t N, int D, const float* thetas, int64_t ld_theta,
                                   int nchains, float sigma, int has_prior, float prior_scale,
                                   double temperature, double* out_U, double* workspace,
                                   int64_t workspace_doubles, int nblocks, void* stream);
/* The same on the fp16 matrix cores. Each scaled value becomes a hi + lo pair of fp16
 * halves (2^-22 relative) and a product is lo hi + hi lo + hi hi with fp32
 * accumulation, so the row dot products hold about 1e-6 relative (float32 products:
 * 6e-8) for a fifth of the matrix time; sums stay FP64. X is scaled by 2^x_exp, chosen
 * once for the dataset so that max |X| 2^x_exp < 2^14 (for example 14 - the frexp
 * exponent of max |X|; -64 <= x_exp <= 64); each chain's position is scaled by its own
 * power of two from its largest finite entry. */
int sgmc_glm_full_potential_chains_split(int family, const float* X, int64_t ldx,
                                         const float* y, int64_t N, int D, const float* thetas,
                                         int64_t ld_theta, int nchains, float sigma,
                                         int has_prior, float prior_scale, double temperature,
                                         int x_exp, double* out_U, double* workspace,
                                         int64_t workspace_doubles, int nblocks, void* stream);

/* ------------------------------------------------------ step updates ----
 * Elementwise integrator updates over a flattened pytree; noise drawn in registers
 * with random_tree semantics (leaf_table as for sgmc_tree_normal). */

/* langevin_diffusion.update_fn (integrator.py:871-912) after the gradient:
 * (key_out, sub) = split(key_in); xi = random_tree(sub); theta += -eps g +
 * sqrt(2 T eps) xi. With v != NULL, rms_prop (adaption.py:254-291) first updates
 * v = alpha v + (1 - alpha) g^2 and scales the drift by G = 1/(lmbd + sqrt(v)) and the
 * noise by sqrt(G). In place on theta (and v). */
int sgmc_sgld_step(float* theta, const float* grad, float* v, int64_t P,
                   const int64_t* leaf_table, int nblocks, const uint32_t* key_in,
                   uint32_t* key_out, float eps, float temperature, float alpha, float lmbd,
                   void* stream);
/* friction_leapfrog.integrate momentum resample (integrator.py:736-738):
 * (key_out, sub) = split(key_in); p = sqrt_mass * random_tree(sub). sqrt_mass may be
 * NULL (unit mass). */
int sgmc_sghmc_resample(float* p, const float* sqrt_mass, const int64_t* leaf_table,
                        int nblocks, const uint32_t* key_in, uint32_t* key_out, void* stream);
/* friction_leapfrog._body_fun position part (integrator.py:609-618):
 * s = inv_mass p; theta += eps s; p += (-eps C) s. friction_arr (per element) may be
 * NULL, then the scalar friction applies. */
int sgmc_sghmc_position(float* theta, float* p, const float* inv_mass, const float* friction_arr,
                        float friction, int64_t P, float eps, void* stream);
/* friction_leapfrog._body_fun momentum part without noise model (integrator.py:626-655):
 * p += -eps g; (key_out, sub) = split(key_in); p += C * (sqrt(2 eps) xi). */
int sgmc_sghmc_momentum(float* p, const float* grad, const float* friction_arr, float friction,
                        const int64_t* leaf_table, int nblocks, const uint32_t* key_in,
                        uint32_t* key_out, float eps, void* stream);
/* friction_leapfrog._body_fun momentum part with a noise model (integrator.py:626-650,
 * adaption.fisher_information): p += -eps g; (key_out, sub) = split(key_in);
 * p += sqrt(2 eps) (cb_diff_sqrt xi), cb_diff_sqrt per element. */
int sgmc_sghmc_momentum_noise_model(float* p, const float* grad, const float* cb_diff_sqrt,
                                    const int64_t* leaf_table, int nblocks,
                                    const uint32_t* key_in, uint32_t* key_out, float eps,
                                    void* stream);
/* obabo O-step (_momentum_resampling, integrator.py:192-200) with noise key
 * noise_key (one of split(key, 3)), and the kinetic energy 0.5 p^T M^-1 p
 * (integrator.py:177-179) of the input momentum (ke_of_input != 0, used for
 * kinetic_energy_end, :261) or of the output momentum (kinetic_energy_start, :222),
 * accumulated as *ke_out = *ke_in + KE. workspace: sgmc_obabo_workspace_floats(nblocks)
 * floats. */
int64_t sgmc_obabo_workspace_floats(int nblocks);
int sgmc_obabo_o(float* p, const float* sqrt_mass, const float* inv_mass,
                 const int64_t* leaf_table, int nblocks, const uint32_t* noise_key,
                 float friction, float eps, float temperature, int ke_of_input,
                 const float* ke_in, float* ke_out, float* workspace, void* stream);
/* obabo B then A (integrator.py:230-240): p = (-eps/2) g + p; theta += eps M^-1 p. */
int sgmc_obabo_ba(float* theta, float* p, const float* grad, const float* inv_mass, int64_t P,
                  float eps, void* stream);
/* obabo second B (integrator.py:248-251): p = (-eps/2) g + p. */
int sgmc_obabo_b(float* p, const float* grad, int64_t P, float eps, void* stream);
/* reversible_leapfrog position step (_position_update, integrator.py:383-388):
 * theta += scale (inv_mass p); scale is eps/2 for the leading half step. */
int sgmc_amagold_position(float* theta, const float* p, const float* inv_mass, int64_t P,
                          float scale, void* stream);
/* reversible_leapfrog._body_fun after the step's gradient (integrator.py:412-453):
 * (key_out, sub) = split(key_in); noise = sqrt(4 f eps) (sqrt_mass xi(sub));
 * p' = (((1 - eps f) p + (-eps) g) + noise) / (1 + eps f);
 * *energy_out = 0.5 eps ((p + p') . (inv_mass g)) + *energy_in.
 * With theta != NULL the position step that follows is applied too:
 * theta += next_scale (inv_mass p') (eps before the next step, eps/2 after the last,
 * integrator.py:401-408, 540-541). workspace: sgmc_obabo_workspace_floats(nblocks)
 * floats whose counter word starts at zero (the kernel leaves it at zero). */
int sgmc_amagold_momentum(float* p, const float* grad, const float* inv_mass,
                          const float* sqrt_mass, const int64_t* leaf_table, int nblocks,
                          const uint32_t* key_in, uint32_t* key_out, float eps, float friction,
                          const float* energy_in, float* energy_out, float* theta,
                          float next_scale, float* workspace, void* stream);

/* --------------------------------------------------- fused multi-chain ----
 * One launch advances every chain of a chain table by one gradient evaluation:
 * in-kernel index draw (data/numpy_loader.py:132-134), row gather, GLM forward and
 * backward, cross-block reduction, and the integrator update with in-register noise:
 *   SGMC_MODE_SGLD / _SGLD_RMS: one launch per iteration (integrator.py:859-922);
 *   SGMC_MODE_SGHMC: `steps` launches per iteration, step = 0..steps-1; step 0 also
 *   resamples the momentum (integrator.py:715-757 with _body_fun 603-666);
 *   SGMC_MODE_AMAGOLD: `steps` launches of reversible_leapfrog (integrator.py:503-553,
 *   _body_fun 397-453): the evaluation point is theta + eps/2 M^-1 p at step 0 and
 *   theta + eps M^-1 p after; the momentum is carried (drawn once at init), energy
 *   accumulates 0.5 eps (p + p') . M^-1 g, step 0 keeps theta_old / mom_old and the
 *   last step adds the final half position step. Samples are saved by
 *   sgmc_glm_amagold_mh.
 *   SGMC_MODE_OBABO: 2 `steps` launches per iteration (obabo.integrate,
 *   integrator.py:318-338, over _leapfrog_steps 203-273), `steps` counting leapfrog
 *   steps; with (key', s1, s2) = split(key, 3), even launch 2k evaluates at theta and
 *   applies O (noise of s1, energy[0] += 0.5 p.M^-1 p of the refreshed momentum), B
 *   and A; odd launch 2k+1 applies B and O (noise of s2, energy[1] += 0.5 p.M^-1 p
 *   before the refresh) and stores key'. U holds U0 after the even launch and
 *   0.5 (U0 + U1) after the odd one. The noise scale uses the chain's temperature.
 *   SGMC_MODE_SGGMC: OBABO proposals for sgmc_glm_sggmc_mh (solver.py:434-567): the
 *   first launch also keeps theta_old, mom_old and energy[2] = U of the iteration's
 *   start; samples are saved by the test.
 * It replaces solver.mcmc's per-chain sequential loop (solver.py:165-171, strategy
 * 'map') for chains that share a likelihood family. The step size of iteration t is
 * eps_table[t] with t = *iter (a device counter the launch advances after the last
 * step of an iteration), so a captured graph of launches replays without host input.
 * save_slot (nullable, int32[table_len]): where save_slot[t] >= 0 the chain's
 * post-step theta is written to samples[save_slot[t]][D] and the potential of the
 * step (the value the integrator state stores) to samples_U[save_slot[t]]. */
typedef struct sgmc_glm_chain {
  const float* X;      /* [N, ldx] rows of this chain's dataset */
  const float* y;      /* [N] */
  float* theta;        /* [D] position, updated in place */
  float* mom;          /* [D] momentum (SGHMC) or NULL */
  float* v;            /* [D] rms_prop second moment or NULL */
  uint32_t* key;       /* [2] integrator key, advanced in place */
  uint32_t* dkey;      /* [2] data-loader key, advanced in place */
  float* U;            /* [1] potential at the evaluation point */
  float* var;          /* [1] var(lik) of the minibatch, or NULL */
  float* lik;          /* [n] per-row likelihoods of the last evaluation, or NULL */
  float* partial;      /* [nblocks][4 ceil(D/4) + 2] scratch */
  unsigned* counter;   /* [1] arrival counter, zero between launches */
  float* samples;      /* [slots][D] or NULL */
  float* samples_U;    /* [slots] potential of each saved sample, or NULL */
  int64_t N;           /* observations */
  int64_t ldx;         /* row stride in floats */
  int32_t D;           /* features (<= 1024) */
  int32_t n;           /* minibatch size */
  int32_t block_begin; /* first grid block of this chain */
  int32_t nblocks;     /* grid blocks of this chain */
  int32_t dclass;      /* sgmc_glm_dclass(D) */
  float temperature;   /* schedule temperature of this chain (SGLD noise) */
  float* energy;       /* AMAGOLD: [1] energy accumulated over the iteration; OBABO: [2]
                          kinetic energy start, end; SGGMC: [3] those and the start U;
                          NULL in the other modes */
  float* theta_old;    /* [D] AMAGOLD, SGGMC: position at the iteration's start, or NULL */
  float* mom_old;      /* [D] AMAGOLD, SGGMC: momentum at the iteration's start, or NULL */
} sgmc_glm_chain;

/* Row-width class of a chain of D features (the kernel instantiation that reads its rows:
 * lanes per row and float4 chunks per lane). A label, not an order: compare classes for
 * equality only. No reference counterpart. */
int sgmc_glm_dclass(int D);
/* Host-side check of a chain table before it is uploaded (no reference counterpart:
 * solver.py:162-171 runs chains one by one and has no table). chains and block_chain
 * are HOST copies. Checks every chain's pointers, D, dclass, n, N and row alignment,
 * that its block range [block_begin, block_begin + nblocks) lies in the grid and holds
 * at most 4096 rows per block, that block_chain maps each range to its chain (or,
 * without a map, that chain c owns blocks [c nblocks / nchains, ...)), and that the
 * ranges add up to nblocks. loop != 0 checks a table for sgmc_glm_fused_loop (one
 * block per chain; the loop kernel runs chain c on block c and ignores block_begin). sgmc_glm_fused_step and sgmc_glm_fused_loop
 * assume a table that passed: their kernels read the map without bounds checks.
 * Returns SGMC_E_ARG with the first violation in sgmc_last_error(). */
int sgmc_glm_check_table(const sgmc_glm_chain* chains, int nchains, const int32_t* block_chain,
                         int nblocks, int loop);
/* chains, block_chain (int32[nblocks], chain of each grid block), eps_table,
 * save_slot, iter and done (a zeroed unsigned) are device memory. block_chain may be
 * NULL for a uniform table: every chain then has nblocks / nchains blocks and chain c
 * owns blocks [c nblocks / nchains, (c + 1) nblocks / nchains) (block_begin must say
 * so). uniform_dclass is
 * the common dclass of all chains, or -1 for a mixed table. With advance_iter != 0 the
 * last chain to finish the launch's final leapfrog step (step == steps - 1, or every
 * SGLD launch) advances iter; a table split into several launches per step (one per
 * dclass) passes advance_iter = 1 on the last of them only. */
int sgmc_glm_fused_step(const sgmc_glm_chain* chains, int nchains, const int32_t* block_chain,
                        int nblocks, int uniform_dclass, int family, float sigma, int has_prior,
                        float prior_scale, float pot_temperature, int mode, int step, int steps,
                        float friction, float alpha, float lmbd, const float* inv_mass,
                        const float* sqrt_mass, const float* eps_table, int table_len,
                        const int32_t* save_slot, int32_t* iter, unsigned* done,
                        int advance_iter, void* stream);

/* AMAGOLD's amortized Metropolis-Hastings test (solver.py:366-419) for every chain of
 * a fused chain table stepped with SGMC_MODE_AMAGOLD, after the iteration's last step
 * (it = iter[0] - 1) and after U_prop[c], the full potential of chain c's proposal
 * (sgmc_glm_full_potential_chains): (key', sub) = split(keys[c]); u = uniform(sub);
 * log a = min(0, U_full[c] - U_prop[c] + energy) in FP64; accept iff log u < log a.
 * Accepted: U_full[c] = U_prop[c], energy_kept[c] = energy. Rejected: theta and mom
 * return to theta_old and -mom_old. At save_slot[it] >= 0 the chain's samples get the
 * post-test theta, samples_U the kept energy and ratio_samples[c][slot] exp(log a)
 * (ratio_samples may be NULL). */
int sgmc_glm_amagold_mh(const sgmc_glm_chain* chains, int nchains, const double* U_prop,
                        double* U_full, float* energy_kept, uint32_t* keys, float* ratio_samples,
                        int n_slots, const int32_t* save_slot, int table_len,
                        const int32_t* iter, void* stream);

/* SGGMC's Metropolis-Hastings test (solver.py:501-567) for every chain of a fused
 * chain table stepped with SGMC_MODE_SGGMC, after the iteration's last launch
 * (it = iter[0] - 1) and after U_prop[c], the full potential of chain c's proposal:
 * (key', sub) = split(keys[c]); u = uniform(sub);
 * log a = -1/T (U_prop - U_full + energy[1] - energy[0]) in FP64, T the chain's
 * temperature, set to 0 unless log a <= 0 (so NaN accepts, as jnp.where does);
 * accept iff log u < log a. Accepted: U_full[c] = U_prop[c]. Rejected: theta, mom and
 * U return to theta_old, mom_old and energy[2]. The kinetic energies restart at zero
 * either way. At save_slot[it] >= 0 the chain's samples get the kept theta and
 * samples_U the kept U; ratio_samples[c][slot] = exp(log a), ke_samples[c][slot] =
 * energy[1] - energy[0] of the proposal and U_samples[c][slot] the kept U_full
 * (each of these three may be NULL). */
int sgmc_glm_sggmc_mh(const sgmc_glm_chain* chains, int nchains, const double* U_prop,
                      double* U_full, uint32_t* keys, float* ratio_samples, float* ke_samples,
                      double* U_samples, int n_slots, const int32_t* save_slot, int table_len,
                      const int32_t* iter, void* stream);

/* `iterations` iterations of a table whose chains all have one block (nblocks == 1,
 * block c is chain c) in one launch: each block runs its chain's evaluations back to
 * back (every leapfrog step of every iteration) with a barrier in between, no arrival
 * counter and no launch per evaluation; step sizes, save slots and keys advance as
 * with sgmc_glm_fused_step, and iter advances by `iterations`. Modes SGLD, SGLD_RMS,
 * SGHMC and OBABO (AMAGOLD and SGGMC test every iteration between launches). */
int sgmc_glm_fused_loop(const sgmc_glm_chain* chains, int nchains, int uniform_dclass,
                        int family, float sigma, int has_prior, float prior_scale,
                        float pot_temperature, int mode, int steps, float friction, float alpha,
                        float lmbd, const float* inv_mass, const float* sqrt_mass,
                        const float* eps_table, int table_len, const int32_t* save_slot,
                        int32_t* iter, unsigned* done, int iterations, void* stream);

/* ---------------------------------------------------- replica exchange ----
 * parallel_tempering's swap step (solver.py:264-293, alias.py:122-207) for pairs of
 * chains of a fused chain table, run after the chains' update of iteration
 * it = iter[0] - 1 (iter as advanced by sgmc_glm_fused_step). For each tested pair:
 *   step += 1; ssq = (1 - 1/step) ssq + (1/step) var(lo);
 *   temps = 1/t_lo - 1/t_hi; log_s = temps ((U(lo) - U(hi)) - temps ssq / F);
 *   (key, sub) = split(key); the pair is kept when log(uniform(sub)) < log_s and
 *   otherwise the two chains exchange their states (theta, momentum, rms_prop state,
 *   integrator and data-loader keys, U, var); the slot temperatures stay.
 * A pair is tested after iteration it when (it + 1) % swap_every == 0 and its parity
 * is -1 or equals ((it + 1) / swap_every - 1) % 2 (alternating even / odd rungs of a
 * ladder; swap_every = 1 with parity -1 is the reference's two-chain solver). Tested
 * pairs of one call must not share a chain. After the test, a pair with samples !=
 * NULL saves the lo chain's theta and U at save_slot[it] when that is >= 0 (the
 * reference saves the normal chain after the exchange). pairs, save_slot and iter are
 * device memory; lo and hi must have the same D. */
typedef struct sgmc_swap_pair {
  int32_t lo, hi;      /* chain indices: lo runs the colder schedule */
  float t_lo, t_hi;    /* schedule temperatures of the two slots */
  int32_t parity;      /* -1, 0 or 1 (see above) */
  int32_t reserved;
  float* ssq;          /* [1] running estimate of var(lik) of the lo chain, init 0 */
  int32_t* step;       /* [1] tests so far, init 0 */
  uint32_t* key;       /* [2] swap key, advanced in place */
  int32_t* swaps;      /* [1] exchanges so far, or NULL */
  float* samples;      /* [slots][D] saved lo-chain samples, or NULL */
  float* samples_U;    /* [slots] their potentials, or NULL */
} sgmc_swap_pair;

int sgmc_resgld_swap(const sgmc_glm_chain* chains, int nchains, const sgmc_swap_pair* pairs,
                     int npairs, float F, int swap_every, const int32_t* save_slot,
                     int table_len, const int32_t* iter, void* stream);

/* ------------------------------------------------------------ MLP (MFMA) ----
 * Bayesian MLP classifier x -> relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3 with a
 * 10-class softmax cross-entropy likelihood: the potential and its gradient
 * (potential.py:159-214 under value_and_grad) by a hand-written backward pass on bf16
 * MFMA GEMMs with fp32 accumulation, and the langevin / rms_prop update
 * (integrator.py:859-922, adaption.py:254-291) fused into the weight-gradient GEMMs.
 * Parameters per chain: float32 [P] in sorted-key order W1, W2, W3, b1, b2, b3, and a
 * bf16 shadow of the same layout (sgmc_mlp_shadow initializes it). One call runs one
 * iteration for C chains (update != 0), or writes U and the gradient without updating
 * (update == 0, grad [C][P], the backward of a torch.autograd.Function). Keys, data
 * keys, the iteration counter and arrival counters live on the device as for the
 * fused GLM step. Shapes: n, D_in, H1 multiples of 64; H1, H2 multiples of 128. */
typedef struct sgmc_mlp_state {
  const float* X;          /* [N][ldx] inputs */
  const int32_t* labels;   /* [N] class labels 0..9 */
  int64_t N, ldx;
  int32_t D_in, H1, H2, n, C;
  float* theta;            /* [C][P] */
  float* v;                /* [C][P] rms_prop second moment, or NULL */
  uint16_t* shadow;        /* [C][P] bf16 copy of theta (NULL allowed for precision 2) */
  int64_t P;
  uint32_t* key;           /* [C][2] */
  uint32_t* dkey;          /* [C][2] */
  void* work;              /* sgmc_mlp_workspace_bytes(C, n, D_in, H1, H2) bytes */
  int64_t work_bytes;
  float* U;                /* [C] */
  float* grad;             /* [C][P] for update == 0, else NULL */
  float* samples;          /* [C][n_slots][P] or NULL */
  float* samples_U;        /* [C][n_slots] or NULL */
  int32_t n_slots;
  const float* eps_table;  /* [table_len] */
  const int32_t* save_slot;/* [table_len] or NULL */
  int32_t* iter;           /* [1] */
  unsigned* counters;      /* [C] zeroed once */
  unsigned* done;          /* [1] zeroed once */
  int32_t table_len;
  int32_t rms;
  float alpha, lmbd, temperature, noise_temperature;
  int32_t has_prior;
  float prior_scale;
  /* OBABO (obabo.integrate, integrator.py:203-338, identity mass): mode 1 runs the
   * first half of one step (O: momentum refresh with xi(s1); B with the gradient at
   * theta; A), mode 2 the second half (B with the gradient at the moved theta; O with
   * xi(s2)); each half draws its own minibatch. mode 0 is the SGLD / rms_prop step. */
  int32_t mode;
  float* mom;              /* [C][P] momentum (modes 1, 2) */
  double* ke;              /* [C][2] kinetic energy accumulated at start / end */
  float* U_half;           /* [C] potential of the first evaluation of the step */
  const float* decay_table;  /* [table_len] sqrt(exp(-friction eps)) */
  const float* noise_table;  /* [table_len] sqrt((1 - exp(-friction eps)) T) */
  /* SGGMC Metropolis-Hastings (solver.py:500-567), see sgmc_mlp_mh */
  double* U_full;          /* [C] full-data potential of the current state */
  double* U_prop;          /* [C] full-data potential of the proposal */
  float* theta_old;        /* [C][P] state the proposal started from */
  float* mom_old;          /* [C][P] */
  float* U_old;            /* [C] its potential (ObaboState.potential) */
  uint32_t* mh_key;        /* [C][2] the solver's key */
  float* mh_stats;         /* [C][n_slots][3] acceptance ratio, step size, KE_end - KE_start
                              at the saved iterations, or NULL */
  int32_t* accepted;       /* [C] accepted proposals, or NULL */
  /* precision 1: the GEMMs run on e4m3 (OCP fp8) operands with per-tensor delayed
   * scaling -- inputs, first hidden layer and the two backward activations, and the
   * W1 / W2 shadows -- each evaluation's amax setting the next scales (q[c][0..5]
   * current scales of X, h1, dA2, dA1, W1, W2; q[c][6..7] the scales of the next W1 /
   * W2 shadows; q[c][8..13] amax bits). precision 0: bf16. precision 2: float32
   * operands on the fp32 matrix instruction (v_mfma_f32_32x32x2_f32), float32
   * activations, the weights read from theta (shadow may be NULL). */
  int32_t precision;
  uint8_t* shadow8;        /* [C][P] e4m3 copy of W1, W2 (precision 1) */
  float* q;                /* [C][16] scale words (precision 1) */
  /* GEMM tiling per call site -- G1 = Xb W1, G2 = h1 W2, G3 = dA2 W2^T, G4 = h1^T dA2
   * (gW2), G5 = Xb^T dA1 (gW1): 0 keeps the built-in choice, else 16 k + s with k the
   * k-tile depth (bf16 32 or 128, e4m3 64 or 128, float32 32) and s the split-K factor
   * (1, 2 or 4; 1 for G4, G5), the GEMM's K a multiple of k s. The k order of the
   * matrix instructions changes with the split only for the forward GEMMs, whose
   * partial products are added in split order. Set by FusedMLP.autotune. */
  int32_t tune[5];
} sgmc_mlp_state;

int64_t sgmc_mlp_param_count(int D_in, int H1, int H2);
int64_t sgmc_mlp_workspace_bytes(int C, int n, int D_in, int H1, int H2);
int sgmc_mlp_shadow(const float* theta, uint16_t* shadow, int64_t count, void* stream);
/* precision 1: sets the weight scales from max |W1|, max |W2| of every chain, writes the
 * e4m3 shadows and sets the activation scales to 1 (an evaluation in gradient mode then
 * calibrates them). */
int sgmc_mlp_init_fp8(const sgmc_mlp_state* st, void* stream);
/* The tiling sgmc_mlp_step launches (update = its update flag) for these shapes and a
 * tune[5] array (see sgmc_mlp_state; NULL = all zero): plan[i] = 16 k + s of GEMM i,
 * the built-in choice where tune[i] is 0. Validates tune like sgmc_mlp_step (SGMC_E_ARG
 * with the message in sgmc_last_error). Host only, no GPU work. No counterpart in the
 * reference (XLA picks its own GEMM tiling under jit). */
int sgmc_mlp_gemm_plan(int D_in, int H1, int H2, int n, int C, int precision, int update,
                       const int32_t* tune, int32_t* plan);
int sgmc_mlp_step(const sgmc_mlp_state* st, int update, void* stream);
/* full_potential (potential.py:219-293) of every chain's current theta: the
 * likelihood of all N rows, n rows per pass, summed in float64 with the float64 log
 * prior: out[c] = (-sum_i lik_i - log p(theta_c)) / temperature. */
int sgmc_mlp_full_potential(const sgmc_mlp_state* st, double* out, void* stream);
/* sggmc's MH step after an OBABO proposal (solver.py:500-567), U_prop holding the
 * proposal's full potential: log a = min(0, -(U_prop - U_full + KE_end - KE_start) / T);
 * (mh_key, sub) = split(mh_key); accepted iff log(uniform(sub)) < log a. A rejected
 * proposal restores theta, momentum and the step potential from the *_old buffers
 * (keys and data keys keep their advanced values); then *_old take the chain's state,
 * the kinetic energies reset to 0, the sample is saved at save_slot[it], and the
 * iteration counter advances. */
int sgmc_mlp_mh(const sgmc_mlp_state* st, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SGMC_H_ */
