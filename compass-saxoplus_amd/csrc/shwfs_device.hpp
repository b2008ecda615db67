// Device-side building blocks shared by the SH WFS kernels (gfx950).
//
//  * Philox4x32-10 counter RNG (Salmon et al., SC'11) and the float64
//    Poisson / Box-Muller samplers the noise stage draws with.
//  * Phase -> (cos, sin) with a compensated Cody-Waite reduction: the float
//    product phase*scale keeps its exact rounding error (FMA two-product), the
//    multiple of 2*pi is removed with a two-constant split, then the HIP
//    sincosf runs on a small argument.
//  * In-register radix-2 FFT of compile-time size with compile-time twiddles.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shwfs {

// ---------------------------------------------------------------- constexpr trig
// Twiddle tables are generated at compile time from these (Taylor series after
// reduction to [-pi/4, pi/4]; accurate to double rounding for the k/N angles
// used here).
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin_small(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_small(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// cos/sin of 2*pi*num/den, exact octant handling (no accumulated range error).
constexpr double cx_cos2pi(long num, long den) {
  long n = ((num % den) + den) % den;  // angle = 2*pi*n/den in [0, 2pi)
  // reduce to octant: 8n/den
  long oct = (8 * n) / den;
  double r = 2.0 * kPi * (double)(8 * n - oct * den) / (8.0 * (double)den);  // in [0, pi/4)
  switch (oct) {
    case 0: return cx_cos_small(r);
    case 1: return cx_sin_small(kPi / 4 - r);
    case 2: return -cx_sin_small(r);
    case 3: return -cx_cos_small(kPi / 4 - r);
    case 4: return -cx_cos_small(r);
    case 5: return -cx_sin_small(kPi / 4 - r);
    case 6: return cx_sin_small(r);
    default: return cx_cos_small(kPi / 4 - r);
  }
}
constexpr double cx_sin2pi(long num, long den) {
  return cx_cos2pi(4 * num - den, 4 * den);  // sin(a) = cos(a - pi/2)
}

// Forward twiddles W_N^k = exp(-2*pi*i*k/N), k = 0..N-1.
template <int N>
struct Twiddle {
  float c[N];
  float s[N];
  constexpr Twiddle() : c(), s() {
    for (int k = 0; k < N; ++k) {
      c[k] = (float)cx_cos2pi(k, N);
      s[k] = (float)(-cx_sin2pi(k, N));
    }
  }
};
template <int N>
__device__ constexpr Twiddle<N> kTw{};

// ---------------------------------------------------------------- complex helpers
// out = a * W_N^k (compile-time k), with the trivial angles special-cased.
template <int N, int K>
__device__ __forceinline__ void twiddle_mul(float& re, float& im) {
  constexpr int k = ((K % N) + N) % N;
  if constexpr (k == 0) {
    return;
  } else if constexpr (4 * k == N) {  // -i
    float t = re;
    re = im;
    im = -t;
  } else if constexpr (2 * k == N) {  // -1
    re = -re;
    im = -im;
  } else if constexpr (4 * k == 3 * N) {  // +i
    float t = re;
    re = -im;
    im = t;
  } else {
    constexpr float wr = kTw<N>.c[k];
    constexpr float wi = kTw<N>.s[k];
    float r = re * wr - im * wi;
    float i = re * wi + im * wr;
    re = r;
    im = i;
  }
}

template <int N, int I>
struct TwAt {
  __device__ __forceinline__ static void apply(float& re, float& im) { twiddle_mul<N, I>(re, im); }
};

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int bitrev(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

// ---- compile-time unrolled radix-2 DIT butterflies via templates
template <int M, int LEN, int I, int J>
struct Bfly {
  __device__ __forceinline__ static void run(float* re, float* im) {
    constexpr int a = I + J, b = I + J + LEN / 2;
    float vr = re[b], vi = im[b];
    twiddle_mul<LEN, J>(vr, vi);
    float ur = re[a], ui = im[a];
    re[a] = ur + vr;
    im[a] = ui + vi;
    re[b] = ur - vr;
    im[b] = ui - vi;
    if constexpr (J + 1 < LEN / 2) Bfly<M, LEN, I, J + 1>::run(re, im);
  }
};
template <int M, int LEN, int I>
struct Group {
  __device__ __forceinline__ static void run(float* re, float* im) {
    Bfly<M, LEN, I, 0>::run(re, im);
    if constexpr (I + LEN < M) Group<M, LEN, I + LEN>::run(re, im);
  }
};
template <int M, int LEN>
struct Stage {
  __device__ __forceinline__ static void run(float* re, float* im) {
    Group<M, LEN, 0>::run(re, im);
    if constexpr (LEN * 2 <= M) Stage<M, LEN * 2>::run(re, im);
  }
};
template <int M, int K>
struct BitRev {
  __device__ __forceinline__ static void run(const float* ir, const float* ii, float* re, float* im) {
    constexpr int r = bitrev(K, ilog2(M));
    re[r] = ir[K];
    im[r] = ii[K];
    if constexpr (K + 1 < M) BitRev<M, K + 1>::run(ir, ii, re, im);
  }
};

// Forward DFT of M points held in registers, natural order in and out.
template <int M>
__device__ __forceinline__ void fft_reg(float (&re)[M], float (&im)[M]) {
  if constexpr (M == 1) return;
  float tr[M], ti[M];
  BitRev<M, 0>::run(re, im, tr, ti);
  Stage<M, 2>::run(tr, ti);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    re[k] = tr[k];
    im[k] = ti[k];
  }
}

// Multiplies element x of a register array by W_N^(x*r) for a runtime r in {0,1}
// (r = 1 applies W_N^x).  Unrolled so every twiddle is a constant.
template <int N, int M, int X>
struct PreTw {
  __device__ __forceinline__ static void run(float* re, float* im) {
    twiddle_mul<N, X>(re[X], im[X]);
    if constexpr (X + 1 < M) PreTw<N, M, X + 1>::run(re, im);
  }
};

// ---------------------------------------------------------------- phase -> field
// sin / cos of theta = phase*scale - offset, accurate for |phase*scale| up to
// ~1e4 rad:
//  * the float product keeps its exact rounding error e (FMA two-product);
//  * q = nearest integer to theta / (pi/2); the multiple q*pi/2 is removed
//    from the product with a two-constant Cody-Waite split of pi/2
//    (H1 = float(pi/2), H2 = float(pi/2 - H1)), then the offset and e are
//    applied, leaving |t| <= pi/4 (+ rounding);
//  * sin t, cos t by their Taylor series to t^11 / t^10 (truncation < 3e-9 on
//    [-pi/4, pi/4]), then the quadrant rotation by q.
__device__ __forceinline__ void sincos_reduced(float t, int qi, float& s, float& c) {
  const float t2 = t * t;
  float sp = __builtin_fmaf(t2, 2.75573192e-06f, -1.98412698e-04f);  // 1/9!, -1/7!
  sp = __builtin_fmaf(t2, sp, 8.33333333e-03f);                      // 1/5!
  sp = __builtin_fmaf(t2, sp, -1.66666667e-01f);                     // -1/3!
  const float sv = __builtin_fmaf(t * t2, sp, t);
  float cp = __builtin_fmaf(t2, -2.75573192e-07f, 2.48015873e-05f);  // -1/10!, 1/8!
  cp = __builtin_fmaf(t2, cp, -1.38888889e-03f);                     // -1/6!
  cp = __builtin_fmaf(t2, cp, 4.16666667e-02f);                      // 1/4!
  cp = __builtin_fmaf(t2, cp, -0.5f);
  const float cv = __builtin_fmaf(t2, cp, 1.0f);
  const float a = (qi & 1) ? cv : sv;
  const float b = (qi & 1) ? sv : cv;
  s = (qi & 2) ? -a : a;
  c = ((qi + 1) & 2) ? -b : b;
}

__device__ __forceinline__ void phase_sincos(float phase, float scale, float offset, float& s,
                                             float& c) {
  constexpr float kInvHalfPi = 0.636619772367581343f;
  constexpr float kH1 = 1.57079637050628662f;      // float nearest pi/2
  constexpr float kH2 = -4.37113900018624283e-08f;  // pi/2 - kH1
  const float p = phase * scale;
  const float e = __builtin_fmaf(phase, scale, -p);  // exact rounding error of the product
  const float q = __builtin_rintf((p - offset) * kInvHalfPi);
  float t = __builtin_fmaf(-q, kH1, p);
  t = __builtin_fmaf(-q, kH2, t);
  t = (t - offset) + e;
  sincos_reduced(t, (int)q, s, c);
}

// ---------------------------------------------------------------- Philox4x32-10
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    // each product as one 64-bit multiply (v_mad_u64_u32) rather than a
    // high and a low 32-bit multiply (two quarter-rate instructions)
    const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
    c = u32x4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
              (uint32_t)p0};
  }
  return c;
}

__device__ __forceinline__ u32x4 rng_block(uint32_t elem, uint32_t j, uint32_t call, uint32_t domain,
                                           uint64_t seed) {
  return philox4x32_10(u32x4{elem, j, call, domain}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ __forceinline__ double u01(uint32_t x) {
  return ((double)x + 0.5) * (1.0 / 4294967296.0);
}

// Box-Muller on words 0, 1 of block 0 (float64 arithmetic).
__device__ __forceinline__ double rng_normal(uint32_t elem, uint32_t call, uint32_t domain,
                                             uint64_t seed) {
  u32x4 w = rng_block(elem, 0, call, domain, seed);
  double u0 = u01(w.x), u1 = u01(w.y);
  return sqrt(-2.0 * log(u0)) * cos(2.0 * 3.14159265358979323846 * u1);
}

// Poisson(lam): inversion for lam < 10, Hoermann's PTRS (1993) otherwise.
__device__ __forceinline__ double rng_poisson(double lam, uint32_t elem, uint32_t call,
                                              uint32_t domain, uint64_t seed) {
  if (!(lam > 0.0)) return 0.0;
  if (lam < 10.0) {
    u32x4 w = rng_block(elem, 0, call, domain, seed);
    double u = u01(w.x);
    double p = exp(-lam);
    double F = p;
    int k = 0;
    while (u > F && k < 200) {
      k += 1;
      p = p * lam / (double)k;
      F = F + p;
    }
    return (double)k;
  }
  double slam = sqrt(lam);
  double loglam = log(lam);
  double b = 0.931 + 2.53 * slam;
  double a = -0.059 + 0.02483 * b;
  double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  double vr = 0.9277 - 3.6224 / (b - 2.0);
  u32x4 w{0, 0, 0, 0};
  // Bounded loop: the acceptance rate of PTRS is > 0.87 per attempt, 4096
  // attempts cannot all be rejected in practice; the bound keeps every wave finite.
  for (int t = 0; t < 4096; ++t) {
    if ((t & 1) == 0) w = rng_block(elem, (uint32_t)(t >> 1), call, domain, seed);
    uint32_t wu = (t & 1) ? w.z : w.x;
    uint32_t wv = (t & 1) ? w.w : w.y;
    double U = u01(wu) - 0.5;
    double V = u01(wv);
    double us = 0.5 - fabs(U);
    double k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
      return k;
  }
  return floor(lam);
}

// log(k!) for the integers 0 <= k < 10, as a select chain (no indexed table)
__device__ __forceinline__ float log_factorial_small(float k) {
  float r = 12.8018274800814696f;               // 9
  r = (k < 8.5f) ? 10.6046029027452502f : r;  // 8
  r = (k < 7.5f) ? 8.52516136106541430f : r;
  r = (k < 6.5f) ? 6.57925121201010100f : r;
  r = (k < 5.5f) ? 4.78749174278204599f : r;
  r = (k < 4.5f) ? 3.17805383034794562f : r;
  r = (k < 3.5f) ? 1.79175946922805500f : r;
  r = (k < 2.5f) ? 0.693147180559945309f : r;
  r = (k < 1.5f) ? 0.0f : r;
  return r;
}

// Poisson(lam) of a float lam, always equal to rng_poisson((double)lam, ...):
// the same counters, draws and decisions, taken in float32 and integers.
//   * The PTRS uniforms are formed from the Philox words: with m = min(wu,
//     2^32 - 1 - wu), rng_poisson's us = 0.5 - |U| is (m + 1/2) 2^-32
//     exactly, so its tests us >= 0.07, us < 0.013 and V > us are the integer
//     tests m >= 300647711, m <= 55834574 and wv > m.
//   * Every other decision (u > F of the inversion, the floor of the PTRS
//     candidate, V <= vr, the acceptance test) is checked against a bound on
//     the float value's error: 2^-20 (k + 1) for the inversion's running sum,
//     2^-18 or 2^-19 times the sum of the magnitudes of the terms elsewhere,
//     8 to 16 times the rounding of every operation plus 1-ulp logf /
//     log1pf / expf / sqrtf / reciprocals.  A lane whose float value falls
//     within that bound of its threshold draws in float64 (rng_poisson), so
//     both give one value for every counter.
//   * The candidate floor(lam + y) (y = (2a/us + b) U + 0.43) is taken as
//     floor(lam) + floor(frac(lam) + y), and the acceptance's right-hand side
//     k log(lam) - lam - log(k!) as (k - lam) - k log1p((k - lam) / lam) -
//     log(2 pi k) / 2 - 1/(12 k) + 1/(360 k^3) for k >= 10 (Stirling's
//     series, truncation below 1e-8) and from log(k!) itself below: terms of
//     size sqrt(lam), not lam log(lam), so the bounds stay narrow at high flux.
// Both branches read block 0 of the element's stream, drawn once.
__device__ __forceinline__ float rng_poisson_f(float lam, uint32_t elem, uint32_t call,
                                               uint32_t domain, uint64_t seed) {
  if (!(lam > 0.0f)) return 0.0f;
  constexpr float k2m32 = 0x1p-32f;
  u32x4 w = rng_block(elem, 0, call, domain, seed);
  bool exact = false;
  float k = 0.0f;
  if (lam < 10.0f) {
    const float u = ((float)w.x + 0.5f) * k2m32;
    float p = expf(-lam);
    float F = p;
    int kk = 0;
    while (kk < 200) {
      if (fabsf(u - F) <= 0x1p-20f * (float)(kk + 1)) {
        exact = true;
        break;
      }
      if (!(u > F)) break;
      kk += 1;
      p = (p * lam) * __builtin_amdgcn_rcpf((float)kk);
      F = F + p;
    }
    k = (float)kk;
  } else if (lam < 1.0e6f) {
    const float slam = sqrtf(lam);
    const float b = 0.931f + 2.53f * slam;
    const float a = -0.059f + 0.02483f * b;
    const float invalpha = 1.1239f + 1.1328f / (b - 3.4f);
    const float vr = 0.9277f - 3.6224f / (b - 2.0f);
    const float loglam = logf(lam), linv = logf(invalpha);
    const float rlam = __builtin_amdgcn_rcpf(lam);
    const float L = floorf(lam), c0 = 0.43f + (lam - L);
    bool done = false;
    for (int t = 0; t < 4096; ++t) {
      if (t > 0 && (t & 1) == 0) w = rng_block(elem, (uint32_t)(t >> 1), call, domain, seed);
      const uint32_t wu = (t & 1) ? w.z : w.x;
      const uint32_t wv = (t & 1) ? w.w : w.y;
      const uint32_t m = min(wu, ~wu);
      const float U = ((float)(int32_t)(wu ^ 0x80000000u) + 0.5f) * k2m32;
      const float V = ((float)wv + 0.5f) * k2m32;
      const float us = ((float)m + 0.5f) * k2m32;
      const float rus = __builtin_amdgcn_rcpf(us);
      const float qU = (2.0f * a * rus) * U, bU = b * U;
      const float y = (qU + bU) + c0;
      const float fy = floorf(y);
      const float tk = 0x1p-18f * (fabsf(qU) + fabsf(bU) + 2.0f);
      if (y - fy <= tk || fy + 1.0f - y <= tk) {
        exact = true;
        break;
      }
      const float kf = L + fy;
      if (m >= 300647711u) {  // us >= 0.07
        if (fabsf(V - vr) <= 0x1p-19f) {
          exact = true;
          break;
        }
        if (V <= vr) {
          k = kf;
          done = true;
          break;
        }
      }
      if (kf < 0.0f || (m <= 55834574u && wv > m)) continue;  // us < 0.013 && V > us
      const float lV = logf(V), lY = logf(a * (rus * rus) + b);
      const float lhs = lV + linv - lY;
      float rhs, mag;
      if (kf < 10.0f) {
        const float lf = log_factorial_small(kf), kl = kf * loglam;
        rhs = (kl - lam) - lf;
        mag = fabsf(kl) + lam + lf;
      } else {
        const float d = kf - lam;
        const float m1 = kf * log1pf(d * rlam);
        const float h = 0.5f * logf(6.28318530717958648f * kf);
        const float ik = __builtin_amdgcn_rcpf(kf);
        rhs = (d - m1) - (h + ik * (1.0f / 12.0f - ik * ik * (1.0f / 360.0f)));
        mag = fabsf(d) + fabsf(m1) + h + 1.0f;
      }
      const float ts = 0x1p-19f * (1.0f + fabsf(lV) + fabsf(linv) + fabsf(lY) + mag);
      if (fabsf(lhs - rhs) <= ts) {
        exact = true;
        break;
      }
      if (lhs <= rhs) {
        k = kf;
        done = true;
        break;
      }
    }
    if (!exact && !done) k = L;
  } else {
    exact = true;
  }
  if (exact) return (float)rng_poisson((double)lam, elem, call, domain, seed);
  return k;
}

// ---------------------------------------------------------------- reductions
// Sum over groups of G consecutive lanes (G a power of two <= 64).
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

// Whole-wave sums with DPP moves (no LDS crossbar): quad exchanges, half-row
// and row mirrors give every lane its 16-lane row sum; row_bcast:15 / :31
// carry the row sums up to lane 63, which is read back as a scalar.
__device__ __forceinline__ int dpp_int(int v, int ctrl_sel) {
  switch (ctrl_sel) {
    case 0: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 1: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 2: return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    case 3: return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // row_mirror
    case 4: return __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    default: return __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false); // row_bcast:31
  }
}

// In-row permutation moves without an `old` operand: with full row and bank
// masks every lane receives a value (ctrl_sel as dpp_int, 0..3).
__device__ __forceinline__ int dpp_row_perm(int v, int ctrl_sel) {
  switch (ctrl_sel) {
    case 0: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);
    case 1: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);
    case 2: return __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);
    default: return __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true);
  }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int c = 0; c < 6; ++c) v += __int_as_float(dpp_int(__float_as_int(v), c));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const long long b = __double_as_longlong(v);
    const int lo = c < 4 ? dpp_row_perm((int)(b & 0xffffffffLL), c)
                         : dpp_int((int)(b & 0xffffffffLL), c);
    const int hi = c < 4 ? dpp_row_perm((int)(b >> 32), c) : dpp_int((int)(b >> 32), c);
    v += __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
  }
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Sum over each 16-lane row, held by every lane of the row.
__device__ __forceinline__ float row_sum(float v) {
#pragma unroll
  for (int c = 0; c < 4; ++c) v += __int_as_float(dpp_int(__float_as_int(v), c));
  return v;
}
__device__ __forceinline__ double row_sum(double v) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_row_perm((int)(b & 0xffffffffLL), c);
    const int hi = dpp_row_perm((int)(b >> 32), c);
    v += __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
  }
  return v;
}

__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Orders the LDS accesses of one wave around a hand-over between its own
// lanes: LDS operations of a wave complete in order, and the fences keep the
// compiler from moving accesses across.  No workgroup barrier: other waves of
// the workgroup are not waited for.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whole-wave totals of three per-lane partial sums, through 192 entries of
// LDS private to the wave: every lane stores its three partials, lane
// 16 q + r (q < 3) adds entries r, r + 16, r + 32, r + 48 of quantity q in
// that order, the rows reduce in registers and the totals are read from
// lanes 0, 16 and 32.  (The order is that of 16 lanes each accumulating the
// values of lanes r, r + 16, ... in turn, then one row sum.)
// Sums of G <= 16 groups of 16 doubles in LDS (red[16 g + r]): four partial
// sums in turn, s_k = ((x_4k + x_4k+1) + x_4k+2) + x_4k+3, then
// (s_0 + s_1) + (s_2 + s_3) through two quad DPP steps; lane 4 g returns
// group g's sum.  The one-sub-aperture and four-sub-aperture tails both sum
// their float64 moments this way (so they agree bit for bit), with three
// instructions per DPP step of a double against twelve for a 16-lane tree.
__device__ __forceinline__ double sum16_quads(const double* red, int lane, int G) {
  double s = 0.0;
  if (lane < 4 * G) {
    const double* p = red + 4 * lane;  // group lane / 4, quarter lane % 4
    s = ((p[0] + p[1]) + p[2]) + p[3];
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {  // lane ^ 1, then lane ^ 2
    const long long b = __double_as_longlong(s);
    const int lo = dpp_row_perm((int)(b & 0xffffffffLL), c);
    const int hi = dpp_row_perm((int)(b >> 32), c);
    s += __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
  }
  return s;
}

template <typename T>
__device__ __forceinline__ void wave_sum3(T* red, int lane, T& a, T& b, T& c) {
  red[lane] = a;
  red[64 + lane] = b;
  red[128 + lane] = c;
  wave_lds_sync();
  T v = T(0);
  if (lane < 48) {
    const T* p = red + (lane >> 4) * 64 + (lane & 15);
    v = ((p[0] + p[16]) + p[32]) + p[48];
  }
  if constexpr (std::is_same<T, double>::value) {
    // the 3 x 16 partials (lane 16 q + r) through sum16_quads
    wave_lds_sync();
    if (lane < 48) red[lane] = v;
    wave_lds_sync();
    v = sum16_quads(red, lane, 3);
    a = lane_value(v, 0);
    b = lane_value(v, 4);
    c = lane_value(v, 8);
  } else {
    v = row_sum(v);
    a = lane_value(v, 0);
    b = lane_value(v, 16);
    c = lane_value(v, 32);
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int c = 0; c < 6; ++c) v = fmaxf(v, __int_as_float(dpp_int(__float_as_int(v), c)));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Sum over groups of G lanes; G == 64 uses the DPP wave sum.
// Sum over groups of G lanes: G == 64 uses the DPP wave sum, G == 16 (one
// row) the DPP row sum, other sizes lane shuffles.
template <int G, typename T>
__device__ __forceinline__ T lane_group_sum(T v) {
  if constexpr (G == 64)
    return wave_sum(v);
  else if constexpr (G == 16)
    return row_sum(v);
  else
    return group_sum<G>(v);
}

}  // namespace shwfs
