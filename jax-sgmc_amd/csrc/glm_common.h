// Likelihood families and the Gaussian prior shared by the GLM kernels.
//   linear-Gaussian  lik = (log(2 pi s^2) + (y - z)^2 / s^2) / (-2)   dlik/dz = (y - z) / s^2
//   logistic         lik = y z - log(1 + e^z)                          dlik/dz = y - sigmoid(z)
//   prior            log p = sum_j (log(2 pi s_p^2) + w_j^2 / s_p^2) / (-2)
#pragma once
#include <math.h>

#include <hip/hip_runtime.h>

#include "sgmc_common.h"

namespace sgmc {

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

static inline Lik make_lik(int family, float sigma) {
  Lik L;
  L.family = family;
  const float s2 = sigma * sigma;
  L.inv_s2 = 1.0f / s2;
  L.log_norm = logf(2.0f * (float)M_PI * s2);
  return L;
}

static inline Prior make_prior(int has_prior, float scale) {
  Prior p;
  p.has_prior = has_prior;
  const float s2 = scale * scale;
  p.inv_var = has_prior ? 1.0f / s2 : 0.0f;
  p.log_norm = has_prior ? logf(2.0f * (float)M_PI * s2) : 0.0f;
  return p;
}

}  // namespace sgmc
