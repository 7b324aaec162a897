/* C port of the multi-chain SGHMC / SGLD loop on GLM likelihoods (oracle only).
 *
 * TEST INFRASTRUCTURE: the timed CPU baseline of bench.py (kind "port") and a second,
 * compiled restatement that tests/ cross-check against the numpy oracle. It follows
 * the reference's op sequence:
 *   jax.random legacy Threefry-2x32 (split, random_bits, normal, randint) as used at
 *     integrator.py:119-135, 630, 736 and data/numpy_loader.py:132-134;
 *   friction_leapfrog.integrate / _body_fun (integrator.py:715-757, 603-666);
 *   langevin_diffusion.update_fn (integrator.py:859-922);
 *   reversible_leapfrog.integrate (integrator.py:503-553) with solver.amagold's
 *     full-data Metropolis-Hastings test (solver.py:366-419);
 *   minibatch potential and gradient (potential.py:159-214) for the logistic and
 *   linear-Gaussian families with a Gaussian prior.
 * Chains run in parallel under OpenMP; within a chain the arithmetic is float32 with
 * float64 reductions.
 *
 * Built twice (Makefile): the oracle build keeps every expression's written order and
 * float64 gradient accumulators (what tests/ compare); the baseline build (CPORT_FAST)
 * lets the compiler vectorize the row dot products, the gradient accumulation and the
 * noise loops with float32 reassociated sums and FMA contraction, as XLA:CPU compiles
 * a dot product and an elementwise fusion. Both compute the same algorithm on the same
 * keys; the baseline's results differ from the oracle build's by float32 rounding.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CP_STR(x) #x
#ifdef CPORT_FAST
#define CP_SIMD_SUM(var) _Pragma(CP_STR(omp simd reduction(+ : var)))
#define CP_SIMD _Pragma("omp simd")
typedef float acc_t;
#else
#define CP_SIMD_SUM(var)
#define CP_SIMD
typedef double acc_t;
#endif

typedef struct {
  uint32_t k0, k1;
} key_t2;

static inline uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static void tf(uint32_t k0, uint32_t k1, uint32_t* x0p, uint32_t* x1p) {
  static const int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  uint32_t x0 = *x0p + ks[0], x1 = *x1p + ks[1];
#pragma GCC unroll 8
  for (int g = 0; g < 5; ++g) {
#pragma GCC unroll 8
    for (int r = 0; r < 4; ++r) {
      x0 += x1;
      x1 = rotl(x1, R[(g & 1) * 4 + r]);
      x1 ^= x0;
    }
    x0 += ks[(g + 1) % 3];
    x1 += ks[(g + 2) % 3] + (uint32_t)(g + 1);
  }
  *x0p = x0;
  *x1p = x1;
}

static uint32_t split_word(key_t2 k, uint32_t num, uint32_t q) {
  uint32_t pair = q < num ? q : q - num, a = pair, b = num + pair;
  tf(k.k0, k.k1, &a, &b);
  return q < num ? a : b;
}

static key_t2 split_key(key_t2 k, uint32_t num, uint32_t i) {
  key_t2 r;
  if (num == 1) {
    uint32_t a = 0, b = 1;
    tf(k.k0, k.k1, &a, &b);
    r.k0 = a;
    r.k1 = b;
    return r;
  }
  r.k0 = split_word(k, num, 2 * i);
  r.k1 = split_word(k, num, 2 * i + 1);
  return r;
}

static uint32_t bits_word(key_t2 k, uint32_t S, uint32_t i) {
  uint32_t m = (S + 1) / 2, j = i < m ? i : i - m;
  uint32_t a = j, b = (j + m < S) ? j + m : 0;
  tf(k.k0, k.k1, &a, &b);
  return i < m ? a : b;
}

/* float32 Cephes logf / log1p and Giles erfinv, contraction-free (see Makefile). */
static float logf_c(float x) {
  uint32_t xb;
  memcpy(&xb, &x, 4);
  float e = (float)((int)(xb >> 23) - 126);
  uint32_t mb = (xb & 0x807FFFFFu) | 0x3F000000u;
  float m;
  memcpy(&m, &mb, 4);
  float t = m - 1.0f;
  if (m < 0.707106781186547524f) {
    t = t + m;
    e = e - 1.0f;
  }
  float x2 = t * t, x3 = x2 * t;
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
  return r + 0.693359375f * e;
}

static float log1p_c(float x) {
  if (!(fabsf(x) < 0.41421356237309504880f)) return logf_c(x + 1.0f);
  static const float NUM[7] = {4.5270000862445199635215e-5f, 4.9854102823193375972212e-1f,
                               6.5787325942061044846969e0f, 2.9911919328553073277375e1f,
                               6.0949667980987787057556e1f, 5.7112963590585538103336e1f,
                               2.0039553499201281259648e1f};
  static const float DEN[7] = {1.0f, 1.5062909083469192043167e1f, 8.3047565967967209469434e1f,
                               2.2176239823732856465394e2f, 3.0909872225312059774938e2f,
                               2.1642788614495947685003e2f, 6.0118660497603843919306e1f};
  float x2 = x * x, num = 0.0f, den = 0.0f;
  for (int i = 0; i < 7; ++i) num = num * x + NUM[i];
  for (int i = 0; i < 7; ++i) den = den * x + DEN[i];
  return x + ((-0.5f * x2) + ((x * x2) * (num / den)));
}

static float erfinv_c(float x) {
  static const float A[9] = {2.81022636e-08f,  3.43273939e-07f, -3.5233877e-06f,
                             -4.39150654e-06f, 0.00021858087f,  -0.00125372503f,
                             -0.00417768164f,  0.246640727f,    1.50140941f};
  static const float B[9] = {-0.000200214257f, 0.000100950558f, 0.00134934322f,
                             -0.00367342844f,  0.00573950773f,  -0.0076224613f,
                             0.00943887047f,   1.00167406f,     2.83297682f};
  float w = -log1p_c(x * (-x));
  const float* c = w < 5.0f ? A : B;
  w = w < 5.0f ? w - 2.5f : sqrtf(w) - 3.0f;
  float p = c[0];
  for (int i = 1; i < 9; ++i) p = c[i] + p * w;
  if (fabsf(x) == 1.0f) return x * INFINITY;
  return p * x;
}

static float normal_of(uint32_t bits) {
  uint32_t fb = (bits >> 9) | 0x3F800000u;
  float f;
  memcpy(&f, &fb, 4);
  f = f - 1.0f;
  float u = f * 2.0f + -0.99999994f;
  if (u < -0.99999994f) u = -0.99999994f;
  return 1.41421354f * erfinv_c(u);
}

#ifdef CPORT_FAST
/* Branch-free forms for the baseline build: both sides of every select are computed,
 * so that gcc vectorizes the noise loops. Same values as the branching forms up to the
 * build's float32 contraction. */
static inline float logf_bf(float x) {
  uint32_t xb;
  memcpy(&xb, &x, 4);
  float e = (float)((int)(xb >> 23) - 126);
  uint32_t mb = (xb & 0x807FFFFFu) | 0x3F000000u;
  float m;
  memcpy(&m, &mb, 4);
  float t = m - 1.0f;
  const int lt = m < 0.707106781186547524f;
  t = lt ? t + m : t;
  e = lt ? e - 1.0f : e;
  float x2 = t * t, x3 = x2 * t;
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
  return r + 0.693359375f * e;
}

static inline float log1p_bf(float x) {
  static const float NUM[7] = {4.5270000862445199635215e-5f, 4.9854102823193375972212e-1f,
                               6.5787325942061044846969e0f, 2.9911919328553073277375e1f,
                               6.0949667980987787057556e1f, 5.7112963590585538103336e1f,
                               2.0039553499201281259648e1f};
  static const float DEN[7] = {1.0f, 1.5062909083469192043167e1f, 8.3047565967967209469434e1f,
                               2.2176239823732856465394e2f, 3.0909872225312059774938e2f,
                               2.1642788614495947685003e2f, 6.0118660497603843919306e1f};
  float x2 = x * x, num = 0.0f, den = 0.0f;
#pragma GCC unroll 8
  for (int i = 0; i < 7; ++i) num = num * x + NUM[i];
#pragma GCC unroll 8
  for (int i = 0; i < 7; ++i) den = den * x + DEN[i];
  const float small = x + ((-0.5f * x2) + ((x * x2) * (num / den)));
  const float big = logf_bf(x + 1.0f);
  return fabsf(x) < 0.41421356237309504880f ? small : big;
}

static inline float normal_bf(uint32_t bits) {
  static const float A[9] = {2.81022636e-08f,  3.43273939e-07f, -3.5233877e-06f,
                             -4.39150654e-06f, 0.00021858087f,  -0.00125372503f,
                             -0.00417768164f,  0.246640727f,    1.50140941f};
  static const float B[9] = {-0.000200214257f, 0.000100950558f, 0.00134934322f,
                             -0.00367342844f,  0.00573950773f,  -0.0076224613f,
                             0.00943887047f,   1.00167406f,     2.83297682f};
  uint32_t fb = (bits >> 9) | 0x3F800000u;
  float f;
  memcpy(&f, &fb, 4);
  f = f - 1.0f;
  float x = f * 2.0f + -0.99999994f;
  x = x < -0.99999994f ? -0.99999994f : x;
  const float w = -log1p_bf(x * (-x));
  const int lo = w < 5.0f;
  const float ww = lo ? w - 2.5f : sqrtf(w) - 3.0f;
  float p = lo ? A[0] : B[0];
#pragma GCC unroll 16
  for (int i = 1; i < 9; ++i) p = (lo ? A[i] : B[i]) + p * ww;
  return 1.41421354f * (p * x);
}

/* normal(key, (S,)) into out, one Threefry block per pair of outputs (j, j + ceil(S/2)),
 * vectorized; threads > 1 splits the pairs over OpenMP threads. */
static void normals_fast(key_t2 lk, uint32_t S, float* o, int threads) {
  const uint32_t m = (S + 1) / 2, k0 = lk.k0, k1 = lk.k1;
#pragma omp parallel for simd num_threads(threads) schedule(static) if (threads > 1) \
    firstprivate(m, k0, k1)
  for (uint32_t j = 0; j < S / 2; ++j) {
    uint32_t a = j, b = j + m;
    tf(k0, k1, &a, &b);
    o[j] = normal_bf(a);
    o[j + m] = normal_bf(b);
  }
  if (S & 1u) o[m - 1] = normal_of(bits_word(lk, S, m - 1));
}
#endif

/* normal(split(sub, 1)[0], (S,)) into out. */
static void leaf_normals(key_t2 sub, uint32_t S, float* out) {
  key_t2 lk = split_key(sub, 1, 0);
#ifdef CPORT_FAST
  normals_fast(lk, S, out, 1);
#else
  for (uint32_t i = 0; i < S; ++i) out[i] = normal_of(bits_word(lk, S, i));
#endif
}

static void split2(key_t2 k, key_t2* a, key_t2* b) {
  *a = split_key(k, 2, 0);
  *b = split_key(k, 2, 1);
}

static void draw_indices(key_t2* dkey, uint32_t n, uint32_t N, int32_t* idx) {
  key_t2 nk, sub, k1, k2;
  split2(*dkey, &nk, &sub);
  split2(sub, &k1, &k2);
  uint32_t m = 65536u % N;
  m = (uint32_t)(m * m) % N;
  for (uint32_t j = 0; j < n; ++j) {
    uint32_t hi = bits_word(k1, n, j), lo = bits_word(k2, n, j);
    idx[j] = (int32_t)(((hi % N) * m + (lo % N)) % N);
  }
  *dkey = nk;
}

/* grad U and U at w for rows idx; family 0 linear-Gaussian, 1 logistic. */
static float potential(int family, const float* X, int64_t ldx, const float* y, int64_t N,
                       int D, int n, const int32_t* idx, const float* w, float sigma,
                       float prior_scale, float* g, acc_t* acc) {
  const float inv_s2 = 1.0f / (sigma * sigma);
  const float ln = logf(2.0f * 3.14159265358979f * sigma * sigma);
  memset(acc, 0, sizeof(acc_t) * (size_t)D);
  double sumlik = 0.0;
  for (int i = 0; i < n; ++i) {
    const float* x = X + (int64_t)idx[i] * ldx;
    float z = 0.0f;
    CP_SIMD_SUM(z)
    for (int j = 0; j < D; ++j) z += x[j] * w[j];
    float lik, r;
    if (family == 0) {
      float d = y[idx[i]] - z;
      lik = (ln + d * d * inv_s2) / -2.0f;
      r = d * inv_s2;
    } else {
      float e = expf(-fabsf(z));
      lik = y[idx[i]] * z - (fmaxf(z, 0.0f) + log1pf(e));
      float s = z >= 0 ? 1.0f / (1.0f + e) : e / (1.0f + e);
      r = y[idx[i]] - s;
    }
    sumlik += lik;
    CP_SIMD
    for (int j = 0; j < D; ++j) acc[j] += (acc_t)(r * x[j]);
  }
  const float c = (float)(-N) / (float)n;
  const float iv = 1.0f / (prior_scale * prior_scale);
  const float pln = logf(2.0f * 3.14159265358979f * prior_scale * prior_scale);
  double lp = 0.0;
  for (int j = 0; j < D; ++j) {
    g[j] = c * (float)acc[j] + w[j] * iv;
    lp += (pln + w[j] * w[j] * iv) / -2.0;
  }
  return (float)((double)(-N) * sumlik / n - lp);
}

/* Runs `iters` SGHMC iterations (friction_leapfrog, `steps` leapfrog steps, unit mass)
 * for C chains: theta [C][D], keys/dkeys [C][2] advanced in place. */
int cport_sghmc(int family, const float* X, int64_t ldx, const float* y, int64_t N, int D,
                int n, int C, float* theta, uint32_t* keys, uint32_t* dkeys, const float* eps,
                int iters, int steps, float friction, float sigma, float prior_scale,
                float* U_out, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int c = 0; c < C; ++c) {
    float* w = theta + (int64_t)c * D;
    float* p = (float*)malloc(sizeof(float) * D);
    float* g = (float*)malloc(sizeof(float) * D);
    float* xi = (float*)malloc(sizeof(float) * D);
    acc_t* acc = (acc_t*)malloc(sizeof(acc_t) * D);
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * n);
    key_t2 key = {keys[2 * c], keys[2 * c + 1]}, dkey = {dkeys[2 * c], dkeys[2 * c + 1]};
    float U = 0.0f;
    for (int t = 0; t < iters; ++t) {
      const float e = eps[t];
      key_t2 nk, sub;
      split2(key, &nk, &sub);
      key = nk;
      leaf_normals(sub, (uint32_t)D, p);
      const float neg_eC = (-e) * friction;
      const float ns = sqrtf(2.0f * e);
      for (int s = 0; s < steps; ++s) {
        for (int j = 0; j < D; ++j) {
          float sm = p[j];
          w[j] = w[j] + e * sm;
          p[j] = p[j] + neg_eC * sm;
        }
        draw_indices(&dkey, (uint32_t)n, (uint32_t)N, idx);
        U = potential(family, X, ldx, y, N, D, n, idx, w, sigma, prior_scale, g, acc);
        split2(key, &nk, &sub);
        key = nk;
        leaf_normals(sub, (uint32_t)D, xi);
        for (int j = 0; j < D; ++j) {
          float pn = p[j] + (-e) * g[j];
          p[j] = pn + friction * (ns * xi[j]);
        }
      }
    }
    keys[2 * c] = key.k0;
    keys[2 * c + 1] = key.k1;
    dkeys[2 * c] = dkey.k0;
    dkeys[2 * c + 1] = dkey.k1;
    if (U_out) U_out[c] = U;
    free(p);
    free(g);
    free(xi);
    free(acc);
    free(idx);
  }
  return 0;
}

/* SGLD (langevin_diffusion without adaption), single-leaf theta. */
int cport_sgld(int family, const float* X, int64_t ldx, const float* y, int64_t N, int D, int n,
               int C, float* theta, uint32_t* keys, uint32_t* dkeys, const float* eps,
               float temperature, int iters, float sigma, float prior_scale, float* U_out,
               int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int c = 0; c < C; ++c) {
    float* w = theta + (int64_t)c * D;
    float* g = (float*)malloc(sizeof(float) * D);
    float* xi = (float*)malloc(sizeof(float) * D);
    acc_t* acc = (acc_t*)malloc(sizeof(acc_t) * D);
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * n);
    key_t2 key = {keys[2 * c], keys[2 * c + 1]}, dkey = {dkeys[2 * c], dkeys[2 * c + 1]};
    float U = 0.0f;
    for (int t = 0; t < iters; ++t) {
      key_t2 nk, sub;
      split2(key, &nk, &sub);
      key = nk;
      draw_indices(&dkey, (uint32_t)n, (uint32_t)N, idx);
      leaf_normals(sub, (uint32_t)D, xi);
      U = potential(family, X, ldx, y, N, D, n, idx, w, sigma, prior_scale, g, acc);
      const float e = eps[t];
      const float ns = sqrtf((2.0f * temperature) * e);
      for (int j = 0; j < D; ++j) w[j] = w[j] + ((-e) * g[j] + ns * xi[j]);
    }
    keys[2 * c] = key.k0;
    keys[2 * c + 1] = key.k1;
    dkeys[2 * c] = dkey.k0;
    dkeys[2 * c + 1] = dkey.k1;
    if (U_out) U_out[c] = U;
    free(g);
    free(xi);
    free(acc);
    free(idx);
  }
  return 0;
}

/* Full-data potential (potential.py:219-293): float32 rows, float64 sum and prior. */
static double full_potential_c(int family, const float* X, int64_t ldx, const float* y,
                               int64_t N, int D, const float* w, float sigma,
                               float prior_scale) {
  const float inv_s2 = 1.0f / (sigma * sigma);
  const float ln = logf(2.0f * 3.14159265358979f * sigma * sigma);
  double s = 0.0;
  for (int64_t i = 0; i < N; ++i) {
    const float* x = X + i * ldx;
    float z = 0.0f;
    CP_SIMD_SUM(z)
    for (int j = 0; j < D; ++j) z += x[j] * w[j];
    float lik;
    if (family == 0) {
      float d = y[i] - z;
      lik = (ln + d * d * inv_s2) / -2.0f;
    } else {
      lik = y[i] * z - (fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))));
    }
    s += lik;
  }
  const double iv = 1.0 / ((double)prior_scale * prior_scale);
  const double pln = log(2.0 * 3.14159265358979 * (double)prior_scale * prior_scale);
  double lp = 0.0;
  for (int j = 0; j < D; ++j) lp += (pln + (double)w[j] * w[j] * iv) / -2.0;
  return -s - lp;
}

/* AMAGOLD (unit mass): per iteration split the MH key, run reversible_leapfrog from
 * (theta, p), evaluate the full potential of the proposal and accept iff
 * log u < min(0, U_full - U' + energy); a rejection keeps theta and flips the
 * momentum. theta/mom [C][D], keys/mh_keys/dkeys [C][2] and U_full [C] advance in place;
 * accepts [C][iters] (nullable) records the decisions. */
int cport_amagold(int family, const float* X, int64_t ldx, const float* y, int64_t N, int D,
                  int n, int C, float* theta, float* mom, uint32_t* keys, uint32_t* mh_keys,
                  uint32_t* dkeys, double* U_full, const float* eps, int iters, int steps, float friction,
                  float sigma, float prior_scale, int8_t* accepts, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int c = 0; c < C; ++c) {
    float* w = theta + (int64_t)c * D;
    float* p = mom + (int64_t)c * D;
    float* th = (float*)malloc(sizeof(float) * D);
    float* p_old = (float*)malloc(sizeof(float) * D);
    float* g = (float*)malloc(sizeof(float) * D);
    float* xi = (float*)malloc(sizeof(float) * D);
    acc_t* acc = (acc_t*)malloc(sizeof(acc_t) * D);
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * n);
    key_t2 key = {keys[2 * c], keys[2 * c + 1]}, mkey = {mh_keys[2 * c], mh_keys[2 * c + 1]};
    key_t2 dkey = {dkeys[2 * c], dkeys[2 * c + 1]};
    for (int t = 0; t < iters; ++t) {
      key_t2 nk, sub, mnext, msub;
      split2(mkey, &mnext, &msub);
      mkey = mnext;
      const float e = eps[t];
      const float half = 0.5f * e;
      const float ef = e * friction;
      const float decay = 1.0f - ef;
      const float rescale = 1.0f / (1.0f + ef);
      const float ns = sqrtf((float)(4.0 * (double)friction) * e);
      memcpy(p_old, p, sizeof(float) * D);
      for (int j = 0; j < D; ++j) th[j] = w[j] + half * p[j];
      float energy = 0.0f;
      for (int s = 0; s < steps; ++s) {
        if (s > 0)
          for (int j = 0; j < D; ++j) th[j] = th[j] + e * p[j];
        split2(key, &nk, &sub);
        key = nk;
        leaf_normals(sub, (uint32_t)D, xi);
        draw_indices(&dkey, (uint32_t)n, (uint32_t)N, idx);
        potential(family, X, ldx, y, N, D, n, idx, th, sigma, prior_scale, g, acc);
        double dot = 0.0;
        for (int j = 0; j < D; ++j) {
          const float pn = rescale * ((decay * p[j] + (-e) * g[j]) + ns * xi[j]);
          dot += (double)((p[j] + pn) * g[j]);
          p[j] = pn;
        }
        energy = half * (float)dot + energy;
      }
      for (int j = 0; j < D; ++j) th[j] = th[j] + half * p[j];
      const double U_new = full_potential_c(family, X, ldx, y, N, D, th, sigma, prior_scale);
      double log_alpha = U_full[c] - U_new + (double)energy;
      if (log_alpha > 0.0) log_alpha = 0.0;
      uint32_t fb = (bits_word(msub, 1u, 0u) >> 9) | 0x3F800000u;
      float u;
      memcpy(&u, &fb, 4);
      u = u - 1.0f;
      if (u < 0.0f) u = 0.0f;
      const int ok = log((double)u) < log_alpha;
      if (ok) {
        memcpy(w, th, sizeof(float) * D);
        U_full[c] = U_new;
      } else {
        for (int j = 0; j < D; ++j) p[j] = -1.0f * p_old[j];
      }
      if (accepts) accepts[(int64_t)c * iters + t] = (int8_t)ok;
    }
    keys[2 * c] = key.k0;
    keys[2 * c + 1] = key.k1;
    mh_keys[2 * c] = mkey.k0;
    mh_keys[2 * c + 1] = mkey.k1;
    dkeys[2 * c] = dkey.k0;
    dkeys[2 * c + 1] = dkey.k1;
    free(th);
    free(p_old);
    free(g);
    free(xi);
    free(acc);
    free(idx);
  }
  return 0;
}

/* SGGMC (unit mass, solver.py:434-567 over obabo, integrator.py:138-346): per
 * iteration split the MH key, run `steps` OBABO steps from (theta, p) with both kinetic
 * energies starting at zero, evaluate the full potential of the proposal;
 * log a = -(U' - U_full + KE_end - KE_start) / T, replaced by 0 unless <= 0; accept iff
 * log u < log a, otherwise theta and p keep their values. theta/mom [C][D],
 * keys/mh_keys/dkeys [C][2] and U_full [C] advance in place; accepts [C][iters]
 * (nullable) records the decisions. */
int cport_sggmc(int family, const float* X, int64_t ldx, const float* y, int64_t N, int D,
                int n, int C, float* theta, float* mom, uint32_t* keys, uint32_t* mh_keys,
                uint32_t* dkeys, double* U_full, const float* eps, int iters, int steps,
                float friction, float temperature, float sigma, float prior_scale,
                int8_t* accepts, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int c = 0; c < C; ++c) {
    float* w = theta + (int64_t)c * D;
    float* p = mom + (int64_t)c * D;
    float* th = (float*)malloc(sizeof(float) * D);
    float* q = (float*)malloc(sizeof(float) * D);
    float* g = (float*)malloc(sizeof(float) * D);
    float* xi = (float*)malloc(sizeof(float) * D);
    acc_t* acc = (acc_t*)malloc(sizeof(acc_t) * D);
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * n);
    key_t2 key = {keys[2 * c], keys[2 * c + 1]}, mkey = {mh_keys[2 * c], mh_keys[2 * c + 1]};
    key_t2 dkey = {dkeys[2 * c], dkeys[2 * c + 1]};
    for (int t = 0; t < iters; ++t) {
      key_t2 mnext, msub;
      split2(mkey, &mnext, &msub);
      mkey = mnext;
      const float e = eps[t];
      const float half = 0.5f * e;
      const float perm = expf((-friction) * e);
      const float a_noise = sqrtf((1.0f - perm) * temperature);
      const float a_decay = sqrtf(perm);
      memcpy(th, w, sizeof(float) * D);
      memcpy(q, p, sizeof(float) * D);
      float ks = 0.0f, ke = 0.0f;
      for (int s = 0; s < steps; ++s) {
        const key_t2 k0 = split_key(key, 3, 0), s1 = split_key(key, 3, 1),
                     s2 = split_key(key, 3, 2);
        key = k0;
        leaf_normals(s1, (uint32_t)D, xi);
        double k = 0.0;
        for (int j = 0; j < D; ++j) {
          q[j] = a_decay * q[j] + a_noise * xi[j];
          k += (double)q[j] * q[j];
        }
        ks = ks + 0.5f * (float)k;
        draw_indices(&dkey, (uint32_t)n, (uint32_t)N, idx);
        potential(family, X, ldx, y, N, D, n, idx, th, sigma, prior_scale, g, acc);
        for (int j = 0; j < D; ++j) {
          q[j] = (-half) * g[j] + q[j];
          th[j] = th[j] + e * q[j];
        }
        draw_indices(&dkey, (uint32_t)n, (uint32_t)N, idx);
        potential(family, X, ldx, y, N, D, n, idx, th, sigma, prior_scale, g, acc);
        k = 0.0;
        for (int j = 0; j < D; ++j) {
          q[j] = (-half) * g[j] + q[j];
          k += (double)q[j] * q[j];
        }
        ke = ke + 0.5f * (float)k;
        leaf_normals(s2, (uint32_t)D, xi);
        for (int j = 0; j < D; ++j) q[j] = a_decay * q[j] + a_noise * xi[j];
      }
      const double U_new = full_potential_c(family, X, ldx, y, N, D, th, sigma, prior_scale);
      double log_alpha = (-1.0 / (double)temperature) * (U_new - U_full[c] + (double)ke -
                                                        (double)ks);
      if (!(log_alpha <= 0.0)) log_alpha = 0.0;
      uint32_t fb = (bits_word(msub, 1u, 0u) >> 9) | 0x3F800000u;
      float u;
      memcpy(&u, &fb, 4);
      u = u - 1.0f;
      if (u < 0.0f) u = 0.0f;
      const int ok = log((double)u) < log_alpha;
      if (ok) {
        memcpy(w, th, sizeof(float) * D);
        memcpy(p, q, sizeof(float) * D);
        U_full[c] = U_new;
      }
      if (accepts) accepts[(int64_t)c * iters + t] = (int8_t)ok;
    }
    keys[2 * c] = key.k0;
    keys[2 * c + 1] = key.k1;
    mh_keys[2 * c] = mkey.k0;
    mh_keys[2 * c + 1] = mkey.k1;
    dkeys[2 * c] = dkey.k0;
    dkeys[2 * c + 1] = dkey.k1;
    free(th);
    free(q);
    free(g);
    free(xi);
    free(acc);
    free(idx);
  }
  return 0;
}

double cport_full_potential(int family, const float* X, int64_t ldx, const float* y, int64_t N,
                            int D, const float* w, float sigma, float prior_scale) {
  return full_potential_c(family, X, ldx, y, N, D, w, sigma, prior_scale);
}

/* Raw jax.random words and normals for cross-checks. */
void cport_normal(const uint32_t* key, uint32_t S, float* out) {
  key_t2 k = {key[0], key[1]};
  for (uint32_t i = 0; i < S; ++i) out[i] = normal_of(bits_word(k, S, i));
}

/* ---- helpers of the float32 MLP port (oracle/mlp_port.py) ----
 * random_tree(key) over a flat vector whose leaves (tree_flatten order) have the given
 * sizes (integrator.py:119-135): leaf i draws normal(split(key, nleaves)[i], size_i). */
void cport_tree_normals(const uint32_t* key, const int64_t* sizes, int nleaves, float* out,
                        int nthreads) {
  key_t2 k = {key[0], key[1]};
  int64_t off = 0;
  for (int l = 0; l < nleaves; ++l) {
    const key_t2 lk = nleaves == 1 ? split_key(k, 1, 0) : split_key(k, (uint32_t)nleaves,
                                                                     (uint32_t)l);
    const uint32_t S = (uint32_t)sizes[l];
    float* o = out + off;
#ifdef CPORT_FAST
    normals_fast(lk, S, o, nthreads);
#else
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (uint32_t i = 0; i < S; ++i) o[i] = normal_of(bits_word(lk, S, i));
#endif
    off += sizes[l];
  }
}

/* (dkey', sub) = split(dkey); idx = randint(sub, (n,), 0, N) (numpy_loader.py:132-134). */
void cport_draw_indices(uint32_t* dkey, uint32_t n, uint32_t N, int32_t* idx) {
  key_t2 k = {dkey[0], dkey[1]};
  draw_indices(&k, n, N, idx);
  dkey[0] = k.k0;
  dkey[1] = k.k1;
}

/* langevin_diffusion's update (integrator.py:882-912) with optional rms_prop
 * (adaption.py:254-291): theta += (-eps) g + noise_scale xi, or with
 * v = alpha v + (1 - alpha) g^2, G = 1 / (lmbd + sqrt(v)): theta += G (-eps) g +
 * sqrt(G) noise_scale xi. */
void cport_sgld_update(float* theta, float* v, const float* g, const float* xi, int64_t P,
                       float eps, float noise_scale, float alpha, float lmbd, int rms,
                       int nthreads) {
#pragma omp parallel for simd num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < P; ++i) {
    const float sg = (-eps) * g[i];
    const float sn = noise_scale * xi[i];
    if (rms) {
      const float vn = alpha * v[i] + (1.0f - alpha) * (g[i] * g[i]);
      v[i] = vn;
      const float G = 1.0f / (lmbd + sqrtf(vn));
      theta[i] = theta[i] + ((0.0f + G * sg) + sqrtf(G) * sn);
    } else {
      theta[i] = theta[i] + (sg + sn);
    }
  }
}

/* The two halves of an OBABO step with identity mass (integrator.py:203-273):
 * first: p1 = decay p + noise xi; returns sum p1^2; p = (-eps/2) g + p1;
 *        theta += eps p;
 * second: p3 = (-eps/2) g + p; returns sum p3^2; p = decay p3 + noise xi. */
double cport_obabo_first(float* theta, float* p, const float* g, const float* xi, int64_t P,
                         float eps, float decay, float noise, int nthreads) {
  const float half = 0.5f * eps;
  double ke = 0.0;
#pragma omp parallel for simd num_threads(nthreads) schedule(static) reduction(+ : ke)
  for (int64_t i = 0; i < P; ++i) {
    const float p1 = decay * p[i] + noise * xi[i];
    ke += (double)(p1 * p1);
    const float p2 = (-half) * g[i] + p1;
    p[i] = p2;
    theta[i] = theta[i] + eps * p2;
  }
  return ke;
}

double cport_obabo_second(float* p, const float* g, const float* xi, int64_t P, float eps,
                          float decay, float noise, int nthreads) {
  const float half = 0.5f * eps;
  double ke = 0.0;
#pragma omp parallel for simd num_threads(nthreads) schedule(static) reduction(+ : ke)
  for (int64_t i = 0; i < P; ++i) {
    const float p3 = (-half) * g[i] + p[i];
    ke += (double)(p3 * p3);
    p[i] = decay * p3 + noise * xi[i];
  }
  return ke;
}
