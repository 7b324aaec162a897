// Elementwise integrator updates on a flattened pytree (one pass over the
// parameters, noise generated in registers). Operation order follows the
// reference's tree algebra (util/tree_util.py:38-133):
//   langevin_diffusion.update_fn          integrator.py:871-912
//   rms_prop update / get                 adaption.py:254-291
//   friction_leapfrog momentum resample   integrator.py:736-738
//   friction_leapfrog _body_fun           integrator.py:609-655
//   obabo _momentum_resampling / B / A    integrator.py:177-271
//
// Noise follows random_tree(key, tree) (integrator.py:119-135): the key is split
// into one key per leaf and each leaf draws normal(key_leaf, leaf.shape). The leaf
// table (int64) is [L, B, (offset, size) x L, (leaf, first_pair) x B]; block b
// covers SGMC_PAIRS_PER_BLOCK counter pairs of one leaf.
#include <math.h>

#include "sgmc_common.h"
#include "sgmc_device.h"

namespace sgmc {

constexpr int kNT = 256;

struct LeafBlock {
  int64_t off;   // offset of the leaf in the flat vector
  uint32_t S;    // leaf size
  uint32_t m;    // counter pairs of the leaf
  uint32_t j0;   // first pair of this block
  uint32_t j1;   // one past the last pair of this block
  Key key;       // normal key of the leaf
};

__device__ __forceinline__ LeafBlock leaf_block(const int64_t* __restrict__ table, Key sub) {
  const int64_t L = table[0];
  const int64_t* leaves = table + 2;
  const int64_t* blocks = leaves + 2 * L;
  LeafBlock lb;
  const int64_t leaf = blocks[2 * blockIdx.x];
  const int64_t p0 = blocks[2 * blockIdx.x + 1];
  lb.off = leaves[2 * leaf];
  lb.S = (uint32_t)leaves[2 * leaf + 1];
  lb.m = (lb.S + 1) >> 1;
  lb.j0 = (uint32_t)p0;
  lb.j1 = (uint32_t)min<int64_t>(p0 + SGMC_PAIRS_PER_BLOCK, lb.m);
  lb.key = split_key(sub, (uint32_t)L, (uint32_t)leaf);
  return lb;
}

// Calls f(flat_index, xi) for every element of this block.
template <class F>
__device__ __forceinline__ void for_each_noise(const LeafBlock& lb, F&& f) {
  for (uint32_t j = lb.j0 + threadIdx.x; j < lb.j1; j += kNT) {
    uint32_t w0, w1;
    bits_pair(lb.key, lb.S, j, w0, w1);
    f(lb.off + j, bits_to_normal(w0));
    if (j + lb.m < lb.S) f(lb.off + j + lb.m, bits_to_normal(w1));
  }
}

// langevin_diffusion update, optionally with rms_prop:
//   (key_out, sub) = split(key_in); xi = random_tree(sub, theta)
//   sg = (-eps) g; sn = sqrt(2 T eps) xi
//   plain:  theta += sg + sn
//   rms:    v = alpha v + (1 - alpha) g^2; G = 1/(lambda + sqrt(v))
//           theta += (eps * 0 + G sg) + sqrt(G) sn
template <bool RMS>
__global__ void __launch_bounds__(kNT) sgld_kernel(float* __restrict__ theta,
                                                   const float* __restrict__ grad,
                                                   float* __restrict__ v,
                                                   const int64_t* __restrict__ table,
                                                   const uint32_t* __restrict__ key_in,
                                                   uint32_t* __restrict__ key_out, float neg_eps,
                                                   float noise_scale, float alpha,
                                                   float one_minus_alpha, float lmbd) {
  Key next, sub;
  split2(load_key(key_in), next, sub);
  if (blockIdx.x == 0 && threadIdx.x == 0) store_key(key_out, next);
  const LeafBlock lb = leaf_block(table, sub);
  for_each_noise(lb, [&](int64_t i, float xi) {
    const float sg = neg_eps * grad[i];
    const float sn = noise_scale * xi;
    if (RMS) {
      const float g = grad[i];
      const float vn = alpha * v[i] + one_minus_alpha * (g * g);
      v[i] = vn;
      const float G = 1.0f / (lmbd + sqrtf(vn));
      theta[i] = theta[i] + ((0.0f + G * sg) + sqrtf(G) * sn);
    } else {
      theta[i] = theta[i] + (sg + sn);
    }
  });
}

// friction_leapfrog momentum resample: (key_out, sub) = split(key_in);
// p = sqrt_mass * random_tree(sub, p).
__global__ void __launch_bounds__(kNT) resample_kernel(float* __restrict__ p,
                                                       const float* __restrict__ sqrt_mass,
                                                       const int64_t* __restrict__ table,
                                                       const uint32_t* __restrict__ key_in,
                                                       uint32_t* __restrict__ key_out) {
  Key next, sub;
  split2(load_key(key_in), next, sub);
  if (blockIdx.x == 0 && threadIdx.x == 0) store_key(key_out, next);
  const LeafBlock lb = leaf_block(table, sub);
  for_each_noise(lb, [&](int64_t i, float xi) { p[i] = sqrt_mass ? sqrt_mass[i] * xi : xi; });
}

// friction_leapfrog position step: s = M^-1 p; theta += eps s; p += (-eps C) s.
__global__ void __launch_bounds__(kNT) sghmc_position_kernel(float* __restrict__ theta,
                                                             float* __restrict__ p,
                                                             const float* __restrict__ inv_mass,
                                                             const float* __restrict__ fric_arr,
                                                             float fric, int64_t P, float eps) {
  const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (i >= P) return;
  const float pi = p[i];
  const float s = inv_mass ? inv_mass[i] * pi : pi;
  theta[i] = theta[i] + eps * s;
  const float c = fric_arr ? fric_arr[i] : fric;
  p[i] = pi + ((-eps) * c) * s;
}

// friction_leapfrog momentum step with a noise model (integrator.py:632-650):
// p += (-eps) g; (key_out, sub) = split(key_in); p += sqrt(2 eps) (cb_diff_sqrt xi).
__global__ void __launch_bounds__(kNT) sghmc_momentum_nm_kernel(float* __restrict__ p,
                                                                const float* __restrict__ grad,
                                                                const float* __restrict__ cb,
                                                                const int64_t* __restrict__ table,
                                                                const uint32_t* __restrict__ key_in,
                                                                uint32_t* __restrict__ key_out,
                                                                float neg_eps, float noise_scale) {
  Key next, sub;
  split2(load_key(key_in), next, sub);
  if (blockIdx.x == 0 && threadIdx.x == 0) store_key(key_out, next);
  const LeafBlock lb = leaf_block(table, sub);
  for_each_noise(lb, [&](int64_t i, float xi) {
    const float pn = p[i] + neg_eps * grad[i];
    p[i] = pn + noise_scale * (cb[i] * xi);
  });
}

// friction_leapfrog momentum step: p += (-eps) g; (key_out, sub) = split(key_in);
// p += C * (sqrt(2 eps) xi).
__global__ void __launch_bounds__(kNT) sghmc_momentum_kernel(float* __restrict__ p,
                                                             const float* __restrict__ grad,
                                                             const float* __restrict__ fric_arr,
                                                             float fric,
                                                             const int64_t* __restrict__ table,
                                                             const uint32_t* __restrict__ key_in,
                                                             uint32_t* __restrict__ key_out,
                                                             float neg_eps, float noise_scale) {
  Key next, sub;
  split2(load_key(key_in), next, sub);
  if (blockIdx.x == 0 && threadIdx.x == 0) store_key(key_out, next);
  const LeafBlock lb = leaf_block(table, sub);
  for_each_noise(lb, [&](int64_t i, float xi) {
    const float c = fric_arr ? fric_arr[i] : fric;
    const float pn = p[i] + neg_eps * grad[i];
    p[i] = pn + c * (noise_scale * xi);
  });
}

// obabo O-step with the kinetic energy 0.5 p^T M^-1 p of the input (ke_of_input) or
// of the output momentum, accumulated as ke_out = ke_in + KE:
//   p' = sqrt(pi) p + sqrt((1 - pi) T) (M^{1/2} xi(key))
__global__ void __launch_bounds__(kNT) obabo_o_kernel(float* __restrict__ p,
                                                      const float* __restrict__ sqrt_mass,
                                                      const float* __restrict__ inv_mass,
                                                      const int64_t* __restrict__ table,
                                                      const uint32_t* __restrict__ key,
                                                      float a_decay, float a_noise,
                                                      int ke_of_input,
                                                      const float* __restrict__ ke_in,
                                                      float* __restrict__ ke_out,
                                                      float* __restrict__ partial,
                                                      unsigned* __restrict__ counter) {
  __shared__ float lds[kNT / kWave];
  __shared__ int flag;
  const LeafBlock lb = leaf_block(table, load_key(key));
  float ke = 0.0f;
  for_each_noise(lb, [&](int64_t i, float xi) {
    const float pi = p[i];
    const float sm = sqrt_mass ? sqrt_mass[i] : 1.0f;
    const float im = inv_mass ? inv_mass[i] : 1.0f;
    const float pn = a_decay * pi + a_noise * (sm * xi);
    p[i] = pn;
    const float q = ke_of_input ? pi : pn;
    ke += q * (im * q);
  });
  const float tot = block_sum<kNT>(ke, lds);
  if (threadIdx.x == 0) publish(partial + blockIdx.x, tot);
  if (arrive_last(counter, gridDim.x, &flag)) {
    float s = 0.0f;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kNT) s += consume(partial + b);
    s = block_sum<kNT>(s, lds);
    if (threadIdx.x == 0) ke_out[0] = ke_in[0] + 0.5f * s;
  }
}

// obabo first B then A: p = (-eps/2) g + p; theta = theta + eps (M^-1 p).
__global__ void __launch_bounds__(kNT) obabo_ba_kernel(float* __restrict__ theta,
                                                       float* __restrict__ p,
                                                       const float* __restrict__ grad,
                                                       const float* __restrict__ inv_mass,
                                                       int64_t P, float neg_half_eps, float eps) {
  const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (i >= P) return;
  const float pn = neg_half_eps * grad[i] + p[i];
  p[i] = pn;
  const float s = inv_mass ? inv_mass[i] * pn : pn;
  theta[i] = theta[i] + eps * s;
}

// obabo second B: p = (-eps/2) g + p.
__global__ void __launch_bounds__(kNT) obabo_b_kernel(float* __restrict__ p,
                                                      const float* __restrict__ grad, int64_t P,
                                                      float neg_half_eps) {
  const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (i >= P) return;
  p[i] = neg_half_eps * grad[i] + p[i];
}

// reversible_leapfrog position step (integrator.py:383-388): theta += scale (M^-1 p).
__global__ void __launch_bounds__(kNT) amagold_position_kernel(float* __restrict__ theta,
                                                               const float* __restrict__ p,
                                                               const float* __restrict__ inv_mass,
                                                               int64_t P, float scale) {
  const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (i >= P) return;
  const float pi = p[i];
  const float s = inv_mass ? inv_mass[i] * pi : pi;
  theta[i] = theta[i] + scale * s;
}

// reversible_leapfrog momentum step (integrator.py:412-453) after the gradient g of
// this step:
//   (key_out, sub) = split(key_in); noise = c_n (M^{1/2} xi(sub))
//   p' = r (((1 - eps f) p + (-eps) g) + noise),  r = 1 / (1 + eps f)
//   energy_out = 0.5 eps ((p + p') . (M^-1 g)) + energy_in
// and, when theta is given, the position step that follows it:
//   theta += next_scale (M^-1 p')  (eps before the next step, eps/2 after the last).
__global__ void __launch_bounds__(kNT) amagold_momentum_kernel(
    float* __restrict__ p, const float* __restrict__ grad, const float* __restrict__ inv_mass,
    const float* __restrict__ sqrt_mass, const int64_t* __restrict__ table,
    const uint32_t* __restrict__ key_in, uint32_t* __restrict__ key_out, float decay,
    float neg_eps, float noise_scale, float rescale, float half_eps,
    const float* __restrict__ energy_in, float* __restrict__ energy_out,
    float* __restrict__ theta, float next_scale, float* __restrict__ partial,
    unsigned* __restrict__ counter) {
  __shared__ float lds[kNT / kWave];
  __shared__ int flag;
  Key next, sub;
  split2(load_key(key_in), next, sub);
  if (blockIdx.x == 0 && threadIdx.x == 0) store_key(key_out, next);
  const LeafBlock lb = leaf_block(table, sub);
  float dot = 0.0f;
  for_each_noise(lb, [&](int64_t i, float xi) {
    const float pi = p[i];
    const float gi = grad[i];
    const float sm = sqrt_mass ? sqrt_mass[i] : 1.0f;
    const float im = inv_mass ? inv_mass[i] : 1.0f;
    const float noise = noise_scale * (sm * xi);
    const float pn = rescale * ((decay * pi + neg_eps * gi) + noise);
    p[i] = pn;
    dot += (pi + pn) * (im * gi);
    if (theta) theta[i] = theta[i] + next_scale * (im * pn);
  });
  const float tot = block_sum<kNT>(dot, lds);
  if (threadIdx.x == 0) publish(partial + blockIdx.x, tot);
  if (arrive_last(counter, gridDim.x, &flag)) {
    float s = 0.0f;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kNT) s += consume(partial + b);
    s = block_sum<kNT>(s, lds);
    if (threadIdx.x == 0) energy_out[0] = half_eps * s + energy_in[0];
  }
}

static inline dim3 flat_grid(int64_t P) { return dim3((unsigned)((P + kNT - 1) / kNT)); }

}  // namespace sgmc

using namespace sgmc;

extern "C" {

int sgmc_sgld_step(float* theta, const float* grad, float* v, int64_t P,
                   const int64_t* leaf_table, int nblocks, const uint32_t* key_in,
                   uint32_t* key_out, float eps, float temperature, float alpha, float lmbd,
                   void* stream) {
  SGMC_REQUIRE(theta && grad && leaf_table && key_in && key_out, SGMC_E_ARG,
               "sgmc_sgld_step: null pointer");
  SGMC_REQUIRE(key_in != key_out, SGMC_E_ARG, "sgmc_sgld_step: keys must not alias");
  SGMC_REQUIRE(P >= 0 && nblocks >= 0, SGMC_E_ARG, "sgmc_sgld_step: bad size");
  if (nblocks == 0) return SGMC_OK;
  const float noise_scale = sqrtf((2.0f * temperature) * eps);
  hipStream_t s = as_stream(stream);
  if (v) {
    hipLaunchKernelGGL(sgld_kernel<true>, dim3(nblocks), dim3(kNT), 0, s, theta, grad, v,
                       leaf_table, key_in, key_out, -eps, noise_scale, alpha, 1.0f - alpha,
                       lmbd);
  } else {
    hipLaunchKernelGGL(sgld_kernel<false>, dim3(nblocks), dim3(kNT), 0, s, theta, grad, v,
                       leaf_table, key_in, key_out, -eps, noise_scale, alpha, 1.0f - alpha,
                       lmbd);
  }
  return check_launch("sgmc_sgld_step");
}

int sgmc_sghmc_resample(float* p, const float* sqrt_mass, const int64_t* leaf_table,
                        int nblocks, const uint32_t* key_in, uint32_t* key_out, void* stream) {
  SGMC_REQUIRE(p && leaf_table && key_in && key_out, SGMC_E_ARG,
               "sgmc_sghmc_resample: null pointer");
  SGMC_REQUIRE(key_in != key_out, SGMC_E_ARG, "sgmc_sghmc_resample: keys must not alias");
  if (nblocks <= 0) return SGMC_OK;
  hipLaunchKernelGGL(resample_kernel, dim3(nblocks), dim3(kNT), 0, as_stream(stream), p,
                     sqrt_mass, leaf_table, key_in, key_out);
  return check_launch("sgmc_sghmc_resample");
}

int sgmc_sghmc_position(float* theta, float* p, const float* inv_mass, const float* friction_arr,
                        float friction, int64_t P, float eps, void* stream) {
  SGMC_REQUIRE(theta && p, SGMC_E_ARG, "sgmc_sghmc_position: null pointer");
  if (P <= 0) return SGMC_OK;
  hipLaunchKernelGGL(sghmc_position_kernel, flat_grid(P), dim3(kNT), 0, as_stream(stream), theta,
                     p, inv_mass, friction_arr, friction, P, eps);
  return check_launch("sgmc_sghmc_position");
}

int sgmc_sghmc_momentum(float* p, const float* grad, const float* friction_arr, float friction,
                        const int64_t* leaf_table, int nblocks, const uint32_t* key_in,
                        uint32_t* key_out, float eps, void* stream) {
  SGMC_REQUIRE(p && grad && leaf_table && key_in && key_out, SGMC_E_ARG,
               "sgmc_sghmc_momentum: null pointer");
  SGMC_REQUIRE(key_in != key_out, SGMC_E_ARG, "sgmc_sghmc_momentum: keys must not alias");
  if (nblocks <= 0) return SGMC_OK;
  const float noise_scale = sqrtf(2.0f * eps);
  hipLaunchKernelGGL(sghmc_momentum_kernel, dim3(nblocks), dim3(kNT), 0, as_stream(stream), p,
                     grad, friction_arr, friction, leaf_table, key_in, key_out, -eps,
                     noise_scale);
  return check_launch("sgmc_sghmc_momentum");
}

int sgmc_sghmc_momentum_noise_model(float* p, const float* grad, const float* cb_diff_sqrt,
                                    const int64_t* leaf_table, int nblocks,
                                    const uint32_t* key_in, uint32_t* key_out, float eps,
                                    void* stream) {
  SGMC_REQUIRE(p && grad && cb_diff_sqrt && leaf_table && key_in && key_out, SGMC_E_ARG,
               "sgmc_sghmc_momentum_noise_model: null pointer");
  SGMC_REQUIRE(key_in != key_out, SGMC_E_ARG,
               "sgmc_sghmc_momentum_noise_model: keys must not alias");
  if (nblocks <= 0) return SGMC_OK;
  const float noise_scale = sqrtf(2.0f * eps);
  hipLaunchKernelGGL(sghmc_momentum_nm_kernel, dim3(nblocks), dim3(kNT), 0, as_stream(stream),
                     p, grad, cb_diff_sqrt, leaf_table, key_in, key_out, -eps, noise_scale);
  return check_launch("sgmc_sghmc_momentum_noise_model");
}

int sgmc_obabo_o(float* p, const float* sqrt_mass, const float* inv_mass,
                 const int64_t* leaf_table, int nblocks, const uint32_t* noise_key,
                 float friction, float eps, float temperature, int ke_of_input,
                 const float* ke_in, float* ke_out, float* workspace, void* stream) {
  SGMC_REQUIRE(p && leaf_table && noise_key && ke_in && ke_out && workspace, SGMC_E_ARG,
               "sgmc_obabo_o: null pointer");
  if (nblocks <= 0) return SGMC_OK;
  const float perm = expf((-friction) * eps);
  const float a_noise = sqrtf((1.0f - perm) * temperature);
  const float a_decay = sqrtf(perm);
  float* partial = workspace;
  unsigned* counter = reinterpret_cast<unsigned*>(workspace + nblocks);
  hipLaunchKernelGGL(obabo_o_kernel, dim3(nblocks), dim3(kNT), 0, as_stream(stream), p,
                     sqrt_mass, inv_mass, leaf_table, noise_key, a_decay, a_noise, ke_of_input,
                     ke_in, ke_out, partial, counter);
  return check_launch("sgmc_obabo_o");
}

int64_t sgmc_obabo_workspace_floats(int nblocks) { return (int64_t)nblocks + 4; }

int sgmc_obabo_ba(float* theta, float* p, const float* grad, const float* inv_mass, int64_t P,
                  float eps, void* stream) {
  SGMC_REQUIRE(theta && p && grad, SGMC_E_ARG, "sgmc_obabo_ba: null pointer");
  if (P <= 0) return SGMC_OK;
  const float half = 0.5f * eps;
  hipLaunchKernelGGL(obabo_ba_kernel, flat_grid(P), dim3(kNT), 0, as_stream(stream), theta, p,
                     grad, inv_mass, P, -1.0f * half, eps);
  return check_launch("sgmc_obabo_ba");
}

int sgmc_obabo_b(float* p, const float* grad, int64_t P, float eps, void* stream) {
  SGMC_REQUIRE(p && grad, SGMC_E_ARG, "sgmc_obabo_b: null pointer");
  if (P <= 0) return SGMC_OK;
  const float half = 0.5f * eps;
  hipLaunchKernelGGL(obabo_b_kernel, flat_grid(P), dim3(kNT), 0, as_stream(stream), p, grad, P,
                     -1.0f * half);
  return check_launch("sgmc_obabo_b");
}

int sgmc_amagold_position(float* theta, const float* p, const float* inv_mass, int64_t P,
                          float scale, void* stream) {
  SGMC_REQUIRE(theta && p, SGMC_E_ARG, "sgmc_amagold_position: null pointer");
  if (P <= 0) return SGMC_OK;
  hipLaunchKernelGGL(amagold_position_kernel, flat_grid(P), dim3(kNT), 0, as_stream(stream),
                     theta, p, inv_mass, P, scale);
  return check_launch("sgmc_amagold_position");
}

int sgmc_amagold_momentum(float* p, const float* grad, const float* inv_mass,
                          const float* sqrt_mass, const int64_t* leaf_table, int nblocks,
                          const uint32_t* key_in, uint32_t* key_out, float eps, float friction,
                          const float* energy_in, float* energy_out, float* theta,
                          float next_scale, float* workspace, void* stream) {
  SGMC_REQUIRE(p && grad && leaf_table && key_in && key_out && energy_in && energy_out &&
               workspace, SGMC_E_ARG, "sgmc_amagold_momentum: null pointer");
  SGMC_REQUIRE(key_in != key_out, SGMC_E_ARG, "sgmc_amagold_momentum: keys must not alias");
  SGMC_REQUIRE(energy_in != energy_out, SGMC_E_ARG,
               "sgmc_amagold_momentum: energies must not alias");
  if (nblocks <= 0) return SGMC_OK;
  // Coefficients in float32, each product rounded as the reference rounds it.
  const float ef = eps * friction;
  const float decay = 1.0f - ef;
  const float rescale = 1.0f / (1.0f + ef);
  const float four_f = (float)(4.0 * (double)friction);
  const float noise_scale = sqrtf(four_f * eps);
  const float half_eps = 0.5f * eps;
  float* partial = workspace;
  unsigned* counter = reinterpret_cast<unsigned*>(workspace + nblocks);
  hipLaunchKernelGGL(amagold_momentum_kernel, dim3(nblocks), dim3(kNT), 0, as_stream(stream), p,
                     grad, inv_mass, sqrt_mass, leaf_table, key_in, key_out, decay, -1.0f * eps,
                     noise_scale, rescale, half_eps, energy_in, energy_out, theta, next_scale,
                     partial, counter);
  return check_launch("sgmc_amagold_momentum");
}

}  // extern "C"
