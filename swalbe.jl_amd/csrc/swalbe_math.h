// swalbe_math.h -- host/device arithmetic shared by every kernel of
// libswalbe_hip: D2Q9 lattice constants, counter-based Philox4x32-10
// (Salmon et al., SC'11), Box-Muller normals built from a portable
// atanh-series logarithm and a Taylor-series sin/cos of pi*r.
//
// The transcendental pieces are written out op by op (no libm, no device
// intrinsics) and the library is compiled with -ffp-contract=off, so a
// normal drawn on the GPU is the same double as the one drawn on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SW_HD __host__ __device__ __forceinline__

namespace swalbe {

// Population k is the reference's plane k+1 (src/collide.jl:270-282);
// velocities from the circshift! shifts of src/collide.jl:92-100.
constexpr int kQ = 9;
__host__ __device__ constexpr int cx(int k) {
  return k == 1 || k == 5 || k == 8 ? 1 : (k == 3 || k == 6 || k == 7 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int k) {
  return k == 2 || k == 5 || k == 6 ? 1 : (k == 4 || k == 7 || k == 8 ? -1 : 0);
}

// ---------------------------------------------------------------- Philox
SW_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}

// Stream tags of the counter's second word.
constexpr uint32_t kTagThermal = 0u;
constexpr uint32_t kTagInitNormal = 1u;
constexpr uint32_t kTagInitUniform = 2u;

// Philox words for counter {c0, tag, step_lo, step_hi}, key {seed_lo, seed_hi}.
SW_HD void philox_ctr(uint64_t seed, uint64_t step, uint32_t c0, uint32_t tag, uint32_t r[4]) {
  r[0] = c0;
  r[1] = tag;
  r[2] = (uint32_t)step;
  r[3] = (uint32_t)(step >> 32);
  philox4x32_10(r, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 53-bit uniform in (0, 1): (floor(w / 2^11) + 1/2) 2^-53, w = hi:lo.
SW_HD double u53(uint32_t hi, uint32_t lo) {
  const uint64_t v = ((((uint64_t)hi) << 32) | lo) >> 11;
  return ((double)v + 0.5) * 1.1102230246251565e-16;
}

// 32-bit uniform in (0, 1): (w + 1/2) 2^-32 (exact in double).
SW_HD double u32(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

// ---------------------------------------------------------------- log
// x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1),
// series sum_k 2 s^(2k+1) / (2k+1) to k = 11 (|s| <= 0.1716).
SW_HD double log_pos(double x) {
  const double C0 = 2.0, C1 = 0.6666666666666666, C2 = 0.4, C3 = 0.2857142857142857,
               C4 = 0.2222222222222222, C5 = 0.18181818181818182, C6 = 0.15384615384615385,
               C7 = 0.13333333333333333, C8 = 0.11764705882352941, C9 = 0.10526315789473684,
               C10 = 0.09523809523809523, C11 = 0.08695652173913043;
  const uint64_t b = (uint64_t)__builtin_bit_cast(uint64_t, x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double q = C11;
  q = C10 + z * q;
  q = C9 + z * q;
  q = C8 + z * q;
  q = C7 + z * q;
  q = C6 + z * q;
  q = C5 + z * q;
  q = C4 + z * q;
  q = C3 + z * q;
  q = C2 + z * q;
  q = C1 + z * q;
  const double lnm = s * (C0 + z * q);
  const double de = (double)e;
  return de * 0.6931471805599453 + (lnm + de * 2.3190468138462996e-17);
}

// ---------------------------------------------------------------- sin/cos(pi x)
// Quarter-turn reduction to r in [-1/4, 1/4], then the Taylor series of
// sin(pi r) (to r^19) and cos(pi r) (to r^20).
SW_HD void sincospi(double x, double* so, double* co) {
  double ax = x < 0 ? -x : x;
  if (ax >= 2.0) ax = ax - 2.0 * floor(ax * 0.5);
  const double k = floor(2.0 * ax + 0.5);
  const double r = ax - 0.5 * k;
  const double z = r * r;
  double sp = -2.2948428997269873e-08;
  sp = 7.952054001475513e-07 + z * sp;
  sp = -2.1915353447830217e-05 + z * sp;
  sp = 0.00046630280576761255 + z * sp;
  sp = -0.0073704309457143504 + z * sp;
  sp = 0.08214588661112823 + z * sp;
  sp = -0.5992645293207921 + z * sp;
  sp = 2.5501640398773455 + z * sp;
  sp = -5.16771278004997 + z * sp;
  sp = 3.141592653589793 + z * sp;
  sp = r * sp;
  double cp = 3.604730797462501e-09;
  cp = -1.3878952462213771e-07 + z * cp;
  cp = 4.303069587032947e-06 + z * cp;
  cp = -0.0001046381049248457 + z * cp;
  cp = 0.0019295743094039231 + z * cp;
  cp = -0.02580689139001406 + z * cp;
  cp = 0.2353306303588932 + z * cp;
  cp = -1.3352627688545895 + z * cp;
  cp = 4.0587121264167685 + z * cp;
  cp = -4.934802200544679 + z * cp;
  cp = 1.0 + z * cp;
  const int q = ((int)k) & 3;
  double s, c;
  if (q == 0) {
    s = sp;
    c = cp;
  } else if (q == 1) {
    s = cp;
    c = -sp;
  } else if (q == 2) {
    s = -sp;
    c = -cp;
  } else {
    s = -cp;
    c = sp;
  }
  *so = x < 0 ? -s : s;
  *co = c;
}

SW_HD double cospi(double x) {
  double s, c;
  sincospi(x, &s, &c);
  return c;
}

// Box-Muller on two uniforms: z0 = r cos(2 pi u2), z1 = r sin(2 pi u2),
// r = sqrt(-2 ln u1).
SW_HD void box_muller(double u1, double u2, double* z0, double* z1) {
  const double rad = sqrt(-2.0 * log_pos(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  *z0 = rad * c;
  *z1 = rad * s;
}

// Stream index of lattice node (x, y) for the normals: 2 (x + Lx floor(y/2))
// + y mod 2, so the vertical node pair (x, 2m), (x, 2m+1) shares one call.
SW_HD uint32_t noise_index(int x, int y, int Lx) {
  return 2u * ((uint32_t)x + (uint32_t)Lx * (uint32_t)(y >> 1)) + (uint32_t)(y & 1);
}

// Two standard normals for (seed, step, stream index s, tag).  One Philox
// call serves the index pair {2q, 2q+1}, q = s / 2: parity b takes words
// 2b, 2b+1.
SW_HD void normal_pair(uint64_t seed, uint64_t step, uint32_t node, uint32_t tag, double* z0,
                       double* z1) {
  uint32_t r[4];
  philox_ctr(seed, step, node >> 1, tag, r);
  const int b = (int)(node & 1u);
  box_muller(u32(r[2 * b]), u32(r[2 * b + 1]), z0, z1);
}

// Uniform of node (seed, step, node, tag): 53 bits of the node's own counter.
SW_HD double uniform(uint64_t seed, uint64_t step, uint32_t node, uint32_t tag) {
  uint32_t r[4];
  philox_ctr(seed, step, node, tag, r);
  return u53(r[0], r[1]);
}

// FP32 Box-Muller on the same two words with the gfx950 transcendental
// instructions (v_log_f32 = log2, v_sin/v_cos_f32 take revolutions): the
// uniform is the top 24 bits of w, so the result is the FP64 normal to
// within FP32 rounding.
__device__ __forceinline__ void box_muller_f32(uint32_t w1, uint32_t w2, float* z0, float* z1) {
  const float u1 = ((float)(w1 >> 8) + 0.5f) * 5.9604644775390625e-08f;
  const float u2 = ((float)(w2 >> 8) + 0.5f) * 5.9604644775390625e-08f;
  const float rad = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
  *z0 = rad * __builtin_amdgcn_cosf(u2);
  *z1 = rad * __builtin_amdgcn_sinf(u2);
}

// Disjoining-pressure constant κ(θ) = (1-cos πθ)(n-1)(m-1)/((n-m) hmin),
// evaluated left to right as the reference broadcast (src/pressure.jl:141-147).
SW_HD double kappa(double theta, int n, int m, double hmin) {
  return (((1.0 - cospi(theta)) * (double)(n - 1)) * (double)(m - 1)) /
         ((double)(n - m) * hmin);
}

}  // namespace swalbe
