// d1q3.hip -- the one-dimensional thin-film model (D1Q3: c = 0, +1, -1).
//
// State_1D and its kinds (src/initialize.jl:276-341, allocator :587-616) on
// the GPU, one kernel per reference op, and two fused time loops
// (src/simulate.jl:98-205, run_gamma :518-560):
//   * k1_step: one launch per step, any L; h on a 3-node halo in LDS.
//   * k1_resident: the whole film in LDS and every step in one launch, one
//     workgroup per film.  A film of L nodes needs (6 + 6 [τ != 1]) L words
//     of LDS.  The same kernel steps an ensemble of films (one workgroup
//     each) in one launch.
// FP64 evaluates every expression in the reference's order (the library is
// built with -ffp-contract=off); FP32 stores plain values.
#include <atomic>
#include <mutex>
#include <string.h>

#include <algorithm>
#include <vector>

#include "swalbe_internal.h"

using swalbe::dkey;
using swalbe::dkey_inv;
using swalbe::hip_fail;
using swalbe::kTagThermal;
using swalbe::normal_pair;
using swalbe::set_error;

namespace {

constexpr int kNT1 = 256;  // output nodes per workgroup of k1_step and the per-op kernels
constexpr int kRTMax = 1024;  // most threads per film of k1_resident
// k1_resident's static LDS: its reduction slots, 4 doubles per wave, at the
// most threads (512 B); the film's planes take what is left
constexpr int64_t kResidentStaticLds = 4 * (kRTMax / 64) * (int64_t)sizeof(double);

inline int blocks1d(int64_t n) {
  int64_t b = (n + kNT1 - 1) / kNT1;
  return (int)(b < 65535 ? (b < 1 ? 1 : b) : 65535);
}

__device__ __forceinline__ int wrapL(int x, int L) {
  x %= L;
  return x < 0 ? x + L : x;
}

template <typename T>
__device__ __forceinline__ T powbroad(T a, int n) {
  T r = T(1);
  for (int i = 0; i < n; ++i) r = r * a;
  return r;
}

// the (n, m) law: power_broad (state form) or fast_93 / fast_32 (raw form),
// src/pressure.jl:363-422
template <typename T>
__device__ __forceinline__ T disjoin(T a, int n, int m, int raw) {
  if (raw) {
    if (n == 9) {
      const T a3 = (a * a) * a;
      return ((a3 * a3) * a3) - a3;
    }
    return ((a * a) * a) - (a * a);
  }
  return powbroad(a, n) - powbroad(a, m);
}

// p = -γ κ [x^n - x^m] - γ (h[x-1] - 2h + h[x+1]),  x = hmin / (h + hcrit)
// (src/pressure.jl:231-255, γ per node as in :284-315)
template <typename T>
__device__ __forceinline__ T pressure_node(T hm, T hc, T hp, T gam, T kap, int n, int m, T hmin,
                                           T hcrit, int raw) {
  const T a = hmin / (hc + hcrit);
  const T v = -gam * (kap * disjoin(a, n, m, raw));
  return v - gam * ((hm - T(2) * hc) + hp);
}

// equilibrium!(feq, h, u, g) for one node (src/equilibrium.jl:169-180)
template <typename T>
__device__ __forceinline__ void feq_node(T h, T u, T g, T& e0, T& e1, T& e2) {
  const T uu = u * u;
  e0 = h * ((T(1) - (T(0.5) * g) * h) - uu);
  e1 = h * (((T(0.25) * g) * h + T(0.5) * u) + T(0.5) * uu);
  e2 = h * (((T(0.25) * g) * h - T(0.5) * u) + T(0.5) * uu);
}

// slippage!(slip, h, u, δ, μ) (src/forcing.jl:68-71)
template <typename T>
__device__ __forceinline__ T slip_node(T h, T u, T c6mu, T c6d, T c3dd) {
  return (c6mu * h * u) / ((T(2) * (h * h) + c6d * h) + c3dd);
}

// thermal! amplitude sqrt(2 kbT μ 6 h / (2h² + 6hδ + 3δ²)) (src/forcing.jl:322-332)
template <typename T>
__device__ __forceinline__ T noise_amp(T hh, T kb26, T delta) {
  return sqrt(kb26 * hh /
              ((((T(2) * hh) * hh) + ((T(6) * hh) * delta)) + ((T(3) * delta) * delta)));
}

// randn! of thermal! replaced by the counter-based normal of node x
template <typename T>
__device__ __forceinline__ T noise_node(uint64_t seed, uint64_t step, int x) {
  double z0, z1;
  normal_pair(seed, step, (uint32_t)x, kTagThermal, &z0, &z1);
  return (T)z0;
}

// ---------------------------------------------------------------- per-op kernels
#define FOR_NODES(x, L)                                                \
  for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < (L); \
       x += gridDim.x * blockDim.x)

template <typename T>
__global__ void k1_equilibrium(T* __restrict__ feq, const T* __restrict__ h,
                               const T* __restrict__ u, T g, int L) {
  FOR_NODES(x, L) feq_node(h[x], u[x], g, feq[x], feq[L + x], feq[2 * L + x]);
}

// BGKandStream! (src/collide.jl:179-205) in two passes: collide every node
// into fo, then pull-stream (with the bounce-back of :214-249 when the
// border masks are given) into ftemp and fout
template <typename T>
__global__ void k1_collide(T* __restrict__ fo, const T* __restrict__ feq,
                           const T* __restrict__ ft, const T* __restrict__ F, T omeg, T itau,
                           int L) {
  FOR_NODES(x, L) {
    fo[x] = omeg * ft[x] + itau * feq[x];
    fo[L + x] = (omeg * ft[L + x] + itau * feq[L + x]) + T(0.5) * F[x];
    fo[2 * L + x] = (omeg * ft[2 * L + x] + itau * feq[2 * L + x]) - T(0.5) * F[x];
  }
}

template <typename T>
__global__ void k1_stream(T* __restrict__ ftemp, T* __restrict__ fout, const T* __restrict__ fo,
                          const T* __restrict__ b1, const T* __restrict__ b2, int L) {
  FOR_NODES(x, L) {
    const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
    T f0 = fo[x], f1 = fo[L + xm], f2 = fo[2 * L + xp];
    if (b1) {
      // fb1 = fo1 b1 and fb2 = fo2 b2 stay at their node, crossed
      f1 = (f1 - f1 * b1[xm]) + fo[2 * L + x] * b2[x];
      f2 = (f2 - f2 * b2[xp]) + fo[L + x] * b1[x];
    }
    ftemp[x] = f0;
    ftemp[L + x] = f1;
    ftemp[2 * L + x] = f2;
    fout[x] = f0;
    fout[L + x] = f1;
    fout[2 * L + x] = f2;
  }
}

template <typename T>
__global__ void k1_moments(T* __restrict__ h, T* __restrict__ u, const T* __restrict__ f, int L) {
  FOR_NODES(x, L) {
    h[x] = (f[x] + f[L + x]) + f[2 * L + x];
    u[x] = (f[L + x] - f[2 * L + x]) / h[x];
  }
}

// κ(θ) per node, in FP64 and the reference's order, stored in T
template <typename T>
__global__ void k1_kappa(T* __restrict__ kap, const T* __restrict__ theta, int n, int m,
                         double hmin, int L) {
  FOR_NODES(x, L) kap[x] = (T)swalbe::kappa((double)theta[x], n, m, hmin);
}

template <typename T>
__global__ void k1_filmpressure(T* __restrict__ p, const T* __restrict__ h, T gamma,
                                const T* __restrict__ gam_f, T kap_s, const T* __restrict__ kap_f,
                                int n, int m, T hmin, T hcrit, int raw, int L) {
  FOR_NODES(x, L) {
    const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
    p[x] = pressure_node(h[xm], h[x], h[xp], gam_f ? gam_f[x] : gamma, kap_f ? kap_f[x] : kap_s,
                         n, m, hmin, hcrit, raw);
  }
}

template <typename T>
__global__ void k1_grad(T* __restrict__ out, const T* __restrict__ f, const T* __restrict__ a,
                        int L) {
  FOR_NODES(x, L) {
    const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
    out[x] = a ? (a[x] * T(-0.5)) * (f[xm] - f[xp]) : T(-0.5) * (f[xm] - f[xp]);
  }
}

template <typename T>
__global__ void k1_laplacian(T* __restrict__ out, const T* __restrict__ f, int L) {
  FOR_NODES(x, L) {
    const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
    out[x] = (f[xm] - T(2) * f[x]) + f[xp];
  }
}

// ∇γ!(state) (form 0, src/forcing.jl:423-432) and ∇γ!(state, sys) (form 1, :434-447)
template <typename T>
__global__ void k1_grad_gamma(T* __restrict__ out, const T* __restrict__ gam,
                              const T* __restrict__ h, T delta, int form, int L) {
  FOR_NODES(x, L) {
    const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
    const T d = (gam[xm] - gam[xp]) / T(2.0);
    if (form == 0) {
      out[x] = T(-1.5) * d;
    } else {
      const T hh = h[x];
      const T a = ((T(2) * (hh * hh)) + ((T(6) * delta) * hh)) + T(3) * (delta * delta);
      out[x] = (((a / (T(6) * hh)) * hh) / T(2)) * d;
    }
  }
}

template <typename T>
__global__ void k1_hgradp(T* __restrict__ out, const T* __restrict__ p, const T* __restrict__ h,
                          int L) {
  FOR_NODES(x, L) {
    const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
    out[x] = (h[x] * T(-0.5)) * (p[xm] - p[xp]);
  }
}

template <typename T>
__global__ void k1_slippage(T* __restrict__ s, const T* __restrict__ h, const T* __restrict__ u,
                            T c6mu, T c6d, T c3dd, int L) {
  FOR_NODES(x, L) s[x] = slip_node(h[x], u[x], c6mu, c6d, c3dd);
}

template <typename T>
__global__ void k1_thermal(T* __restrict__ out, const T* __restrict__ h, T kb26, T delta,
                           uint64_t seed, uint64_t step, int L) {
  FOR_NODES(x, L) out[x] = noise_node<T>(seed, step, x) * noise_amp(h[x], kb26, delta);
}

template <typename T>
__global__ void k1_inclination(T* __restrict__ F, const T* __restrict__ h, T alpha, T fac, int L) {
  FOR_NODES(x, L) F[x] = F[x] + (h[x] * alpha) * fac;
}

// F .= -h∇p .- slip [.- kbt] [∓ ∇γ] (src/simulate.jl:111, 548)
template <typename T>
__global__ void k1_force(T* __restrict__ F, const T* __restrict__ hgp, const T* __restrict__ s,
                         const T* __restrict__ kbt, const T* __restrict__ gg, int gg_sign, int L) {
  FOR_NODES(x, L) {
    T v = -hgp[x] - s[x];
    if (kbt) v = v - kbt[x];
    if (gg) v = gg_sign < 0 ? v - gg[x] : v + gg[x];
    F[x] = v;
  }
}

template <typename T>
__global__ void k1_fill(T* __restrict__ p, T v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------- fused steps
// Everything a fused 1D step needs, for one film.  Pointers are device
// pointers of the working type (series: double); NULL switches a term off.
struct Film1 {
  const void* h_in;
  const void* u_in;
  const void* ft;        // f_t of the previous step (read when τ ≠ 1)
  const void* kap;       // κ(θ) field, or NULL for the scalar kap_s
  const void* gam;       // γ field, or NULL for the scalar gamma
  const void* gg;        // ∇γ field (force term), or NULL
  const void* b1;        // bounce-back masks, or NULL
  const void* b2;
  const void* interior;  // wall interior mask (mass series), or NULL
  void* h_out;
  void* u_out;
  void* f_out;
  void* p_out;
  void* h_last;  // k1_resident: receives the last step's input moments
  void* u_last;
  double* dh;    // k1_resident series: [nsteps], [nsteps], [nmass], [nsnap * L]
  double* wet;
  double* mass;
  double* snap;
  double gamma, kap_s, hmin, hcrit, c6mu, c6d, c3dd, kb26, delta, omeg, itau, g, alpha, fac;
  double wet_thresh;
  uint64_t seed;
  int64_t t0, tdump, snap_every;
  int64_t mass_phase, snap_phase;  // steps before the first mass / snapshot step
  int L, n, m, thermal, body, read_ft, gg_sign;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double a = __shfl_xor(v, o);
    v = a > v ? a : v;
  }
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double a = __shfl_xor(v, o);
    v = a < v ? a : v;
  }
  return v;
}

// Nodes per thread of k1_step: two up to 512 threads (LDS at most 49 KB).
// Measured at L = 2^24 with 256 threads: 69.4 GLUPS at 2, 52.5 at 4 (3 waves
// per SIMD), 62.2 at 1.
constexpr int k1_npt(int nb) { return nb <= 512 ? 2 : 1; }

// One workgroup: output nodes x0 .. x0 + NB*NPT - 1 (NPT per thread); h on a
// 3-halo, u on 1, p on 2, and the post-collision populations of the nodes
// x0 - 1 .. x0 + W, collided once at their source (each node's equilibrium
// is built once, not once per population that leaves it).  Odd steps take
// the blocks in reverse, so a step starts where the previous one stopped
// writing (those inputs are still in the caches).  Nodes are independent:
// neither order nor grouping changes a result.

template <typename T, int NB>
__global__ __launch_bounds__(NB) void k1_step(Film1 S, uint64_t step,
                                              unsigned long long* dh_keys,
                                              unsigned long long* wet) {
  constexpr int NPT = k1_npt(NB), W = NB * NPT;
  __shared__ T sh[W + 6], su[W + 2], sp[W + 4], s0[W + 2], s1[W + 2], s2[W + 2];
  const int t = threadIdx.x, L = S.L;
  const int bb = (step & 1) ? (int)gridDim.x - 1 - (int)blockIdx.x : (int)blockIdx.x;
  const int x0 = bb * W;
  const T* h_in = (const T*)S.h_in;
  const T* u_in = (const T*)S.u_in;
  const T* kap_f = (const T*)S.kap;
  const T* gam_f = (const T*)S.gam;
  // a workgroup whose 3-halo wraps nowhere indexes directly (no modulo)
  const bool inner = x0 >= 3 && x0 + W + 3 <= L;
  auto wr = [&](int y) { return inner ? y : wrapL(y, L); };
  __builtin_amdgcn_s_setprio(2);  // the loads ahead of other waves' arithmetic
  for (int i = t; i < W + 6; i += NB) sh[i] = h_in[wr(x0 - 3 + i)];
  for (int i = t; i < W + 2; i += NB) su[i] = u_in[wr(x0 - 1 + i)];
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  for (int i = t; i < W + 4; i += NB) {  // p at node x0 - 2 + i
    const int y = wr(x0 - 2 + i);
    sp[i] = pressure_node(sh[i], sh[i + 1], sh[i + 2], gam_f ? gam_f[y] : (T)S.gamma,
                          kap_f ? kap_f[y] : (T)S.kap_s, S.n, S.m, (T)S.hmin, (T)S.hcrit, 0);
  }
  __syncthreads();
  const T g = (T)S.g, omeg = (T)S.omeg, itau = (T)S.itau;
  const T* ft = (const T*)S.ft;
  for (int i = t; i < W + 2; i += NB) {  // force and collision of node x0 - 1 + i
    const int y = wr(x0 - 1 + i);
    const T hh = sh[i + 2];
    const T hgp = (hh * T(-0.5)) * (sp[i] - sp[i + 2]);
    T F = -hgp - slip_node(hh, su[i], (T)S.c6mu, (T)S.c6d, (T)S.c3dd);
    if (S.thermal) F = F - noise_node<T>(S.seed, step, y) * noise_amp(hh, (T)S.kb26, (T)S.delta);
    if (S.gg) F = S.gg_sign < 0 ? F - ((const T*)S.gg)[y] : F + ((const T*)S.gg)[y];
    if (S.body) F = F + (hh * (T)S.alpha) * (T)S.fac;
    T e0, e1, e2;
    feq_node(hh, su[i], g, e0, e1, e2);
    const T t0 = S.read_ft ? ft[y] : T(0), t1 = S.read_ft ? ft[L + y] : T(0),
            t2 = S.read_ft ? ft[2 * L + y] : T(0);
    s0[i] = omeg * t0 + itau * e0;
    s1[i] = (omeg * t1 + itau * e1) + T(0.5) * F;
    s2[i] = (omeg * t2 + itau * e2) - T(0.5) * F;
  }
  __syncthreads();
  const T* b1 = (const T*)S.b1;
  const T* b2 = (const T*)S.b2;
  T* fout = (T*)S.f_out;
  unsigned long long kmax = 0ull, kmin = ~0ull, cnt = 0ull;
#pragma unroll
  for (int q = 0; q < NPT; ++q) {
    const int j = t + q * NB;  // output node x0 + j
    const int x = x0 + j;
    if (x >= L) break;
    if (dh_keys || wet) {
      const double hv = (double)sh[j + 3];
      const unsigned long long k = dkey(hv);
      kmax = k > kmax ? k : kmax;
      kmin = k < kmin ? k : kmin;
      cnt += hv > S.wet_thresh ? 1ull : 0ull;
    }
    const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
    const T f0 = s0[j + 1];
    T f1 = s1[j], f2 = s2[j + 2];
    if (b1) {
      f1 = (f1 - f1 * b1[xm]) + s2[j + 1] * b2[x];
      f2 = (f2 - f2 * b2[xp]) + s1[j + 1] * b1[x];
    }
    // f and p stream past the caches (no step of this kernel reads them at
    // τ = 1, and a τ ≠ 1 step reads f_t once); h and u stay cached, the next
    // step's inputs.  Streaming took a long film from 69.5 to 82.3 GLUPS.
    __builtin_nontemporal_store(f0, &fout[x]);
    __builtin_nontemporal_store(f1, &fout[L + x]);
    __builtin_nontemporal_store(f2, &fout[2 * L + x]);
    const T hn = (f0 + f1) + f2;
    ((T*)S.h_out)[x] = hn;
    ((T*)S.u_out)[x] = (f1 - f2) / hn;
    __builtin_nontemporal_store(sp[j + 2], &((T*)S.p_out)[x]);
  }
  if (dh_keys || wet) {  // one atomic per wave
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a = __shfl_xor(kmax, o), b = __shfl_xor(kmin, o);
      kmax = a > kmax ? a : kmax;
      kmin = b < kmin ? b : kmin;
      cnt += __shfl_xor(cnt, o);
    }
    if ((t & 63) == 0) {
      if (dh_keys) {
        atomicMax(&dh_keys[0], kmax);
        atomicMin(&dh_keys[1], kmin);
      }
      if (wet && cnt) atomicAdd(wet, cnt);
    }
  }
}

// Pointers read from the film table are generic to the compiler, which then
// emits flat loads and stores (they wait on the LDS counter as well).  All of
// them address global memory: say so.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* glb(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// Whole films in LDS: workgroup b steps film table[b] through nsteps steps.
// LDS: h, u, h', u', p, F (and f_t, f_t' when τ ≠ 1), L words each (3L for f);
// within a step the collision at the source overwrites h, u and F with the
// three post-collision populations, which the stream phase gathers.
// Only what every step uses is held in registers; the series pointers,
// last-step outputs and noise constants are read from the table where they
// are used (the table is not restrict: the loop's global stores keep those
// loads in place), and the mass / snapshot schedule runs on countdowns.
template <typename T, int kRT>
__global__ __launch_bounds__(kRT) __attribute__((amdgpu_waves_per_eu(kRT == 512 ? 4 : 1)))
void k1_resident(const Film1* table, int64_t nsteps) {
  extern __shared__ __align__(16) unsigned char smem1[];
  __shared__ double red[4][kRT / 64];
  static_assert(kRT <= kRTMax && sizeof(red) <= kResidentStaticLds, "static LDS budget");
  const auto* F1 = glb(table + blockIdx.x);
  const int L = F1->L, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int read_ft = F1->read_ft;
  T* h = (T*)smem1;
  T* u = h + L;
  T* hn = u + L;
  T* un = hn + L;
  T* p = un + L;
  T* Fo = p + L;
  T* ft = Fo + L;
  T* ftn = ft + 3 * L;
  {
    const auto* h_in = glb((const T*)F1->h_in);
    const auto* u_in = glb((const T*)F1->u_in);
    for (int x = tid; x < L; x += kRT) {
      h[x] = h_in[x];
      u[x] = u_in[x];
    }
    if (read_ft) {
      const auto* ft_in = glb((const T*)F1->ft);
      for (int i = tid; i < 3 * L; i += kRT) ft[i] = ft_in[i];
    }
  }
  __syncthreads();
  const int n = F1->n, m = F1->m, gg_sign = F1->gg_sign, thermal = F1->thermal, body = F1->body;
  const T gam_s = (T)F1->gamma, kap_s = (T)F1->kap_s, hmin = (T)F1->hmin, hcrit = (T)F1->hcrit;
  const T c6mu = (T)F1->c6mu, c6d = (T)F1->c6d, c3dd = (T)F1->c3dd;
  const T g = (T)F1->g, omeg = (T)F1->omeg, itau = (T)F1->itau;
  const auto* kap_f = glb((const T*)F1->kap);
  const auto* gam_f = glb((const T*)F1->gam);
  const auto* gg = glb((const T*)F1->gg);
  const auto* b1 = glb((const T*)F1->b1);
  const auto* b2 = glb((const T*)F1->b2);
  const bool series = F1->dh || F1->wet || F1->mass;
  // The per-node constants every step reads (γ, ∇γ and the wall masks
  // around x) are held in registers for each thread's first kC nodes;
  // further nodes, and a θ field, are read from global memory each step.
  constexpr int kC = kRT <= 512 ? 2 : 0;
  constexpr int kCs = kC > 0 ? kC : 1;
  T c_gam[kCs], c_gg[kCs], c_b1m[kCs], c_b1[kCs], c_b2[kCs], c_b2p[kCs];
#pragma unroll
  for (int j = 0; j < kC; ++j) {
    const int x = tid + j * kRT;
    if (x < L) {
      const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
      c_gam[j] = gam_f ? gam_f[x] : gam_s;
      c_gg[j] = gg ? gg[x] : T(0);
      c_b1m[j] = b1 ? b1[xm] : T(0);
      c_b1[j] = b1 ? b1[x] : T(0);
      c_b2[j] = b2 ? b2[x] : T(0);
      c_b2p[j] = b2 ? b2[xp] : T(0);
    }
  }
  // body(x, j) for every node of this thread; j is the register slot of a
  // cached node (a constant after unrolling) or -1
  auto for_nodes = [&](auto&& body) {
#pragma unroll
    for (int j = 0; j < kC; ++j) {
      const int x = tid + j * kRT;
      if (x < L) body(x, j);
    }
    for (int x = tid + kC * kRT; x < L; x += kRT) body(x, -1);
  };
  auto pressure_phase = [&]() {
    for_nodes([&](int x, int j) {
      const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
      const T gm = j >= 0 ? c_gam[j] : (gam_f ? gam_f[x] : gam_s);
      const T kp = kap_f ? kap_f[x] : kap_s;
      p[x] = pressure_node(h[xm], h[x], h[xp], gm, kp, n, m, hmin, hcrit, 0);
    });
  };
  // A film without walls computes the next step's pressure in the stream
  // phase: each thread also streams the heights of its nodes' two
  // neighbours (the same sums, so the same bits), which saves the barrier
  // between the moments and the pressure.  Walled films keep three phases,
  // and so do 1024-thread films: one node per thread, run when a few films
  // hold the GPU, where the longer stream phase costs more than the barrier.
  const bool fuse_p = kRT <= 512 && b1 == nullptr;
  if (fuse_p) {
    pressure_phase();
    __syncthreads();
  }
  // countdowns to the next mass (t % tdump == 0) and snapshot step
  int64_t cm = F1->mass ? F1->mass_phase : -1, cs = F1->snap ? F1->snap_phase : -1;
  int64_t km = 0, ks = 0;
  for (int64_t it = 1; it <= nsteps; ++it) {
    // the table pointer is opaque from here on: cold fields are reloaded
    // (scalar-cache hits) instead of occupying scalar registers all loop
    asm volatile("" : "+s"(F1));
    const bool last = it == nsteps;
    const bool do_mass = cm == 0;
    if (series) {  // series over the step's input height
      const auto* inter = glb((const T*)F1->interior);
      const double wthr = F1->wet_thresh;
      double s = 0.0, mx = -__builtin_inf(), mn = __builtin_inf(), c = 0.0;
      for (int x = tid; x < L; x += kRT) {
        const double hv = (double)h[x];
        s += inter ? (double)(h[x] - inter[x] * h[x]) : hv;
        mx = hv > mx ? hv : mx;
        mn = hv < mn ? hv : mn;
        c += hv > wthr ? 1.0 : 0.0;
      }
      s = wave_sum(s);
      mx = wave_max(mx);
      mn = wave_min(mn);
      c = wave_sum(c);
      if (lane == 0) {
        red[0][wv] = s;
        red[1][wv] = mx;
        red[2][wv] = mn;
        red[3][wv] = c;
      }
      __syncthreads();
      if (tid == 0) {
        double s0 = red[0][0], m0 = red[1][0], n0 = red[2][0], c0 = red[3][0];
        for (int w = 1; w < kRT / 64; ++w) {
          s0 += red[0][w];
          m0 = red[1][w] > m0 ? red[1][w] : m0;
          n0 = red[2][w] < n0 ? red[2][w] : n0;
          c0 += red[3][w];
        }
        if (do_mass) glb(F1->mass)[km] = s0;
        if (F1->dh) glb(F1->dh)[it - 1] = m0 - n0;
        if (F1->wet) glb(F1->wet)[it - 1] = c0;
      }
    }
    if (do_mass) {
      ++km;
      cm = F1->tdump;
    }
    if (cm > 0) --cm;
    if (!fuse_p) {
      pressure_phase();
      __syncthreads();
    }
    // forces, then the collision at the source: node x's post-collision
    // populations replace its own h, u (no other node reads them in this
    // phase) and F: o0 -> h[x], o1 -> u[x], o2 -> Fo[x]
    for_nodes([&](int x, int j) {
      const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
      const T hh = h[x], uu = u[x];
      const T hgp = (hh * T(-0.5)) * (p[xm] - p[xp]);
      T F = -hgp - slip_node(hh, uu, c6mu, c6d, c3dd);
      if (thermal)
        F = F - noise_node<T>(F1->seed, (uint64_t)(F1->t0 + it), x) *
                    noise_amp(hh, (T)F1->kb26, (T)F1->delta);
      if (gg) {
        const T gv = j >= 0 ? c_gg[j] : gg[x];
        F = gg_sign < 0 ? F - gv : F + gv;
      }
      if (body) F = F + (hh * (T)F1->alpha) * (T)F1->fac;
      if (last) {  // the last step's input moments
        glb((T*)F1->h_last)[x] = hh;
        glb((T*)F1->u_last)[x] = uu;
      }
      T e0, e1, e2;
      feq_node(hh, uu, g, e0, e1, e2);
      const T t0 = read_ft ? ft[x] : T(0), t1 = read_ft ? ft[L + x] : T(0),
              t2 = read_ft ? ft[2 * L + x] : T(0);
      h[x] = omeg * t0 + itau * e0;
      u[x] = (omeg * t1 + itau * e1) + T(0.5) * F;
      Fo[x] = (omeg * t2 + itau * e2) - T(0.5) * F;
    });
    __syncthreads();
    // pull-stream (bounce-back at walls) and moments
    for_nodes([&](int x, int j) {
      const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
      const T f0 = h[x];
      T f1 = u[xm], f2 = Fo[xp];
      if (b1) {
        const T bm = j >= 0 ? c_b1m[j] : b1[xm], bx2 = j >= 0 ? c_b2[j] : b2[x];
        const T bp = j >= 0 ? c_b2p[j] : b2[xp], bx1 = j >= 0 ? c_b1[j] : b1[x];
        f1 = (f1 - f1 * bm) + Fo[x] * bx2;
        f2 = (f2 - f2 * bp) + u[x] * bx1;
      }
      const T hv = (f0 + f1) + f2;
      const T uv = (f1 - f2) / hv;
      hn[x] = hv;
      un[x] = uv;
      if (read_ft) {
        ftn[x] = f0;
        ftn[L + x] = f1;
        ftn[2 * L + x] = f2;
      }
      if (last) {
        auto* fout = glb((T*)F1->f_out);
        fout[x] = f0;
        fout[L + x] = f1;
        fout[2 * L + x] = f2;
        glb((T*)F1->h_out)[x] = hv;
        glb((T*)F1->u_out)[x] = uv;
        glb((T*)F1->p_out)[x] = p[x];
      } else if (fuse_p) {  // the next step's pressure (no walls here)
        const int xmm = xm == 0 ? L - 1 : xm - 1, xpp = xp == L - 1 ? 0 : xp + 1;
        const T hvm = (h[xm] + u[xmm]) + Fo[x];
        const T hvp = (h[xp] + u[x]) + Fo[xpp];
        const T gm = j >= 0 ? c_gam[j] : (gam_f ? gam_f[x] : gam_s);
        const T kp = kap_f ? kap_f[x] : kap_s;
        p[x] = pressure_node(hvm, hv, hvp, gm, kp, n, m, hmin, hcrit, 0);
      }
    });
    __syncthreads();
    T* tmp = h;
    h = hn;
    hn = tmp;
    tmp = u;
    u = un;
    un = tmp;
    tmp = ft;
    ft = ftn;
    ftn = tmp;
    if (cs == 0) {  // snapshot!(…, h, t) after the step
      auto* snap = glb(F1->snap + ks * L);
      for (int x = tid; x < L; x += kRT) snap[x] = (double)h[x];
      ++ks;
      cs = F1->snap_every;
    }
    if (cs > 0) --cs;
  }
}

template <typename T>
__global__ void k1_reduce(const T* __restrict__ h, const T* __restrict__ inter, int L,
                          double thresh, double* out) {
  // one workgroup, fixed order: {Σh (Σ(h - interior h) with a mask), max, min, count}
  __shared__ double red[4][256 / 64];
  double s = 0.0, mx = -__builtin_inf(), mn = __builtin_inf(), c = 0.0;
  for (int x = threadIdx.x; x < L; x += 256) {
    const double v = (double)h[x];
    s += inter ? (double)(h[x] - inter[x] * h[x]) : v;
    mx = v > mx ? v : mx;
    mn = v < mn ? v : mn;
    c += v > thresh ? 1.0 : 0.0;
  }
  s = wave_sum(s);
  mx = wave_max(mx);
  mn = wave_min(mn);
  c = wave_sum(c);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s;
    red[1][w] = mx;
    red[2][w] = mn;
    red[3][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; ++i) {
      red[0][0] += red[0][i];
      red[1][0] = red[1][i] > red[1][0] ? red[1][i] : red[1][0];
      red[2][0] = red[2][i] < red[2][0] ? red[2][i] : red[2][0];
      red[3][0] += red[3][i];
    }
    out[0] = red[0][0];
    out[1] = red[1][0];
    out[2] = red[2][0];
    out[3] = red[3][0];
  }
}

__global__ void k1_init_keys(unsigned long long* k, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    k[2 * i] = 0ull;
    k[2 * i + 1] = ~0ull;
  }
}

__global__ void k1_zero_u64(unsigned long long* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

}  // namespace

// ---------------------------------------------------------------- state
struct swalbe_film1d {
  int L = 0, dtype = SWALBE_F64, kind = SWALBE_SIMPLE, device = 0;
  size_t esz = 8;
  hipStream_t stream = nullptr;  // one of the device's shared 1D streams (never destroyed here)
  void* slab = nullptr;          // one allocation holding every fixed plane below and `red`
  void* f[2] = {nullptr, nullptr};  // distribution buffers (3 planes of L)
  int fout_buf = 0, ftemp_buf = 1;  // equal after a fused loop (fout .= ftemp)
  void* feq = nullptr;
  void* fo = nullptr;  // collision scratch of the per-op BGKandStream!
  void* hbuf[2] = {nullptr, nullptr};
  void* ubuf[2] = {nullptr, nullptr};
  int cur = 0;
  void *pressure = nullptr, *F = nullptr, *slip = nullptr, *hgp = nullptr, *kbt = nullptr;
  void *theta = nullptr, *kap = nullptr, *gamma = nullptr, *grad_gamma = nullptr;
  void *border[2] = {nullptr, nullptr}, *interior = nullptr;
  bool have_border = false;
  double* red = nullptr;
};

namespace {

#define SW_HIP1(call)                                 \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

int bad1(const char* m) {
  set_error(m);
  return SWALBE_EINVAL;
}

bool is_dist1(int f) { return f == SWALBE1D_FOUT || f == SWALBE1D_FTEMP || f == SWALBE1D_FEQ; }

// optional fields, allocated on first upload
void** opt_slot(swalbe_film1d* s, int f) {
  switch (f) {
    case SWALBE1D_THETA: return &s->theta;
    case SWALBE1D_GAMMA: return &s->gamma;
    case SWALBE1D_GRADGAMMA: return &s->grad_gamma;
    case SWALBE1D_BORDER1: return &s->border[0];
    case SWALBE1D_BORDER2: return &s->border[1];
    case SWALBE1D_INTERIOR: return &s->interior;
    default: return nullptr;
  }
}

void* field1(swalbe_film1d* s, int f) {
  switch (f) {
    case SWALBE1D_FOUT: return s->f[s->fout_buf];
    case SWALBE1D_FTEMP: return s->f[s->ftemp_buf];
    case SWALBE1D_FEQ: return s->feq;
    case SWALBE1D_HEIGHT: return s->hbuf[s->cur];
    case SWALBE1D_VEL: return s->ubuf[s->cur];
    case SWALBE1D_PRESSURE: return s->pressure;
    case SWALBE1D_F: return s->F;
    case SWALBE1D_SLIP: return s->slip;
    case SWALBE1D_HGRADP: return s->hgp;
    case SWALBE1D_KBT: return s->kbt;
    default: {
      void** o = opt_slot(s, f);
      return o ? *o : nullptr;
    }
  }
}

#define SW_BOTH1(s, CALL) \
  do {                    \
    if ((s)->esz == 8) {  \
      using T = double;   \
      CALL;               \
    } else {              \
      using T = float;    \
      CALL;               \
    }                     \
  } while (0)

int fill1(swalbe_film1d* s, void* p, double v, int64_t n) {
  SW_BOTH1(s, hipLaunchKernelGGL(k1_fill<T>, dim3(blocks1d(n)), dim3(kNT1), 0, s->stream, (T*)p,
                                 (T)v, n));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

// Streams of the 1D films: a few per device, created on first use and kept
// until the process ends.  A film takes one of them round-robin.  Films used
// to own a stream each; with thousands of live films (ensembles of scan
// films) that is thousands of HIP streams, each holding runtime queue
// resources (signals, kernel-argument pools) for as long as the film lives.
constexpr int kStreams1D = 4;
constexpr int kMaxDevices1D = 64;
std::mutex g_streams_mu;
hipStream_t g_streams[kMaxDevices1D][kStreams1D] = {};
unsigned g_stream_next[kMaxDevices1D] = {};

int film_stream(int dev, hipStream_t* out) {
  if (dev < 0 || dev >= kMaxDevices1D) return bad1("swalbe1d_create: device index out of range");
  std::lock_guard<std::mutex> lk(g_streams_mu);
  if (!g_streams[dev][0]) {
    for (int k = 0; k < kStreams1D; ++k) {
      const hipError_t e = hipStreamCreateWithFlags(&g_streams[dev][k], hipStreamNonBlocking);
      if (e != hipSuccess) {
        for (int j = 0; j < k; ++j) (void)hipStreamDestroy(g_streams[dev][j]);
        for (int j = 0; j < kStreams1D; ++j) g_streams[dev][j] = nullptr;
        return hip_fail(e, "swalbe1d_create: stream creation");
      }
    }
  }
  *out = g_streams[dev][g_stream_next[dev]++ % kStreams1D];
  return SWALBE_OK;
}

int alloc1(swalbe_film1d* s, void** p, int64_t n) {
  if (hipMalloc(p, (size_t)n * s->esz) != hipSuccess) {
    *p = nullptr;
    set_error("swalbe1d: device allocation failed");
    return SWALBE_ENOMEM;
  }
  return fill1(s, *p, 0.0, n);
}

// fout and ftemp share a buffer after a fused loop (they are equal); a
// write to one of them separates them again
void split_dists(swalbe_film1d* s, int field) {
  if ((field == SWALBE1D_FOUT || field == SWALBE1D_FTEMP) && s->fout_buf == s->ftemp_buf) {
    const int other = 1 - s->ftemp_buf;
    (void)hipMemcpyAsync(s->f[other], s->f[s->ftemp_buf], 3 * (size_t)s->L * s->esz,
                   hipMemcpyDeviceToDevice, s->stream);
    if (field == SWALBE1D_FOUT) s->fout_buf = other;
    else s->ftemp_buf = other;
  }
}

int64_t max_lds_bytes(int device) {
  // per-device cache; threads that race here store the same value
  static std::atomic<int64_t> cache[64];
  if (device < 0 || device >= 64) return 65536;
  int64_t c = cache[device].load(std::memory_order_relaxed);
  if (!c) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess ||
        v <= 0)
      v = 65536;
    c = v;
    cache[device].store(c, std::memory_order_relaxed);
  }
  return c;
}

int64_t resident_bytes(int L, size_t esz, bool read_ft) {
  return (int64_t)(read_ft ? 12 : 6) * L * (int64_t)esz;
}

bool fits_resident(const swalbe_film1d* s, const swalbe_params* p) {
  return resident_bytes(s->L, s->esz, p->tau != 1.0) + kResidentStaticLds <=
         max_lds_bytes(s->device);
}

int check_opts(swalbe_film1d* s, const swalbe_params* p, const swalbe1d_step_opts* o) {
  if (p->n < 0 || p->m < 0) return bad1("swalbe1d_time_loop: bad (n, m)");
  if (o->use_theta_field && !s->theta) return bad1("swalbe1d_time_loop: no θ field uploaded");
  if (o->use_gamma_field && !s->gamma) return bad1("swalbe1d_time_loop: no γ field uploaded");
  if (o->grad_gamma < -1 || o->grad_gamma > 1)
    return bad1("swalbe1d_time_loop: grad_gamma is -1, 0 or 1");
  if (o->grad_gamma && !s->grad_gamma)
    return bad1("swalbe1d_time_loop: no ∇γ field (upload it or call swalbe1d_grad_gamma)");
  if (o->bound && !s->have_border) return bad1("swalbe1d_time_loop: no border masks uploaded");
  if (o->variant < 0 || o->variant > 2) return bad1("swalbe1d_time_loop: variant is 0, 1 or 2");
  if (o->snap_every < 0) return bad1("swalbe1d_time_loop: snap_every < 0");
  if (o->threads != 0 && o->threads != 256 && o->threads != 512 && o->threads != 1024)
    return bad1("swalbe1d_time_loop: threads is 0, 256, 512 or 1024");
  return SWALBE_OK;
}

// κ(θ) field from the θ field for the given (n, m, hmin)
int refresh_kappa(swalbe_film1d* s, int n, int m, double hmin) {
  if (!s->kap)
    if (int rc = alloc1(s, &s->kap, s->L)) return rc;
  SW_BOTH1(s, hipLaunchKernelGGL(k1_kappa<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->kap, (const T*)s->theta, n, m, hmin, s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int64_t count_every(int64_t t0, int64_t nsteps, int64_t every) {
  if (every <= 0) return 0;
  int64_t k = 0;
  for (int64_t it = 1; it <= nsteps; ++it)
    if ((t0 + it) % every == 0) ++k;
  return k;
}

// the per-film description for the fused kernels; the loop's input moments
// are hbuf[cur], its output goes to hbuf[1 - cur]; f_t is f[ftemp_buf] and
// the output distributions go to the other buffer
Film1 describe(swalbe_film1d* s, const swalbe_params* p, const swalbe1d_step_opts* o) {
  Film1 S{};
  S.L = s->L;
  S.gamma = p->gamma;
  S.kap_s = swalbe::kappa(p->theta, p->n, p->m, p->hmin);
  S.hmin = p->hmin;
  S.hcrit = p->hcrit;
  S.c6mu = 6.0 * p->mu;
  S.c6d = 6.0 * p->delta;
  S.c3dd = 3.0 * (p->delta * p->delta);
  S.kb26 = ((2.0 * p->kbt) * p->mu) * 6.0;
  S.delta = p->delta;
  S.omeg = 1.0 - 1.0 / p->tau;
  S.itau = 1.0 / p->tau;
  S.g = p->g;
  S.alpha = o->alpha;
  S.fac = o->body_fac;
  S.n = p->n;
  S.m = p->m;
  S.thermal = o->thermal;
  S.body = o->body;
  S.read_ft = p->tau != 1.0;
  S.gg_sign = o->grad_gamma;
  S.seed = o->seed;
  S.t0 = o->t0;
  S.tdump = p->tdump;
  S.snap_every = o->snap_every;
  // step it (1-based) is a mass step when (t0 + it) % tdump == 0: the first is it = phase + 1
  S.mass_phase = p->tdump > 0 ? (p->tdump - (o->t0 + 1) % p->tdump) % p->tdump : -1;
  S.snap_phase = o->snap_every > 0 ? (o->snap_every - (o->t0 + 1) % o->snap_every) % o->snap_every : -1;
  S.wet_thresh = o->wet_thresh;
  S.kap = o->use_theta_field ? s->kap : nullptr;
  S.gam = o->use_gamma_field ? s->gamma : nullptr;
  S.gg = o->grad_gamma ? s->grad_gamma : nullptr;
  S.b1 = o->bound ? s->border[0] : nullptr;
  S.b2 = o->bound ? s->border[1] : nullptr;
  S.interior = o->bound ? s->interior : nullptr;
  S.h_in = s->hbuf[s->cur];
  S.u_in = s->ubuf[s->cur];
  S.h_out = s->hbuf[1 - s->cur];
  S.u_out = s->ubuf[1 - s->cur];
  S.h_last = s->hbuf[s->cur];
  S.u_last = s->ubuf[s->cur];
  S.ft = s->f[s->ftemp_buf];
  S.f_out = s->f[1 - s->ftemp_buf];
  S.p_out = s->pressure;
  return S;
}

}  // namespace

extern "C" {

int swalbe1d_create(int32_t L, int32_t dtype, int32_t kind, swalbe_film1d** out) {
  if (!out || L < 3 || (dtype != SWALBE_F64 && dtype != SWALBE_F32) ||
      (kind != SWALBE_SIMPLE && kind != SWALBE_THERMAL))
    return bad1("swalbe1d_create: bad arguments (L >= 3, dtype, kind)");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "swalbe1d_create: hipGetDevice");
  swalbe_film1d* s = new swalbe_film1d();
  s->L = L;
  s->dtype = dtype;
  s->kind = kind;
  s->esz = dtype == SWALBE_F64 ? 8 : 4;
  s->device = dev;
  int rc = film_stream(dev, &s->stream);
  // every fixed plane in one allocation, each plane 256-byte aligned
  void** planes3[] = {&s->f[0], &s->f[1], &s->feq, &s->fo};
  void** planes1[] = {&s->hbuf[0], &s->hbuf[1], &s->ubuf[0], &s->ubuf[1], &s->pressure,
                      &s->F,       &s->slip,    &s->hgp,     &s->kbt};
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b3 = up(3 * (size_t)L * s->esz), b1 = up((size_t)L * s->esz);
  const size_t total = 4 * b3 + 9 * b1 + up(sizeof(double) * 4);
  if (rc == SWALBE_OK) {
    e = hipMalloc(&s->slab, total);
    if (e != hipSuccess) {
      s->slab = nullptr;
      set_error(std::string("swalbe1d_create: device allocation failed: ") + hipGetErrorString(e));
      rc = SWALBE_ENOMEM;
    }
  }
  if (rc == SWALBE_OK) {
    char* q = static_cast<char*>(s->slab);
    for (void** p : planes3) {
      *p = q;
      q += b3;
    }
    for (void** p : planes1) {
      *p = q;
      q += b1;
    }
    s->red = reinterpret_cast<double*>(q);
    for (void** p : planes3)
      if (rc == SWALBE_OK) rc = fill1(s, *p, 0.0, 3 * (int64_t)L);
    for (void** p : planes1)
      if (rc == SWALBE_OK) rc = fill1(s, *p, 0.0, L);
  }
  if (rc == SWALBE_OK) rc = fill1(s, s->hbuf[0], 1.0, L);  // Sys: height = ones(L)
  if (rc == SWALBE_OK && hipStreamSynchronize(s->stream) != hipSuccess)
    rc = bad1("swalbe1d_create: initialisation failed");
  if (rc != SWALBE_OK) {
    swalbe1d_destroy(s);
    return rc;
  }
  *out = s;
  return SWALBE_OK;
}

int swalbe1d_destroy(swalbe_film1d* s) {
  if (!s) return SWALBE_OK;
  hipError_t e = hipSetDevice(s->device);
  // the film's work must be done before its memory goes (the stream is shared
  // with other films: this waits for theirs too)
  if (e == hipSuccess && s->stream) e = hipStreamSynchronize(s->stream);
  void* ptrs[] = {s->slab, s->theta, s->kap, s->gamma, s->grad_gamma, s->border[0],
                  s->border[1], s->interior};
  for (void* p : ptrs)
    if (p) {
      const hipError_t ef = hipFree(p);
      if (e == hipSuccess) e = ef;
    }
  delete s;
  if (e != hipSuccess) return hip_fail(e, "swalbe1d_destroy");
  return SWALBE_OK;
}

int swalbe1d_length(swalbe_film1d* s, int32_t* L) {
  if (!s || !L) return bad1("swalbe1d_length: null argument");
  *L = s->L;
  return SWALBE_OK;
}

int swalbe1d_upload(swalbe_film1d* s, int32_t field, const void* host, size_t bytes) {
  if (!s || !host || field < 0 || field >= SWALBE1D_NFIELDS)
    return bad1("swalbe1d_upload: bad arguments");
  (void)hipSetDevice(s->device);
  const int64_t n = is_dist1(field) ? 3 * (int64_t)s->L : s->L;
  const size_t hes = bytes == (size_t)n * 8 ? 8 : (bytes == (size_t)n * 4 ? 4 : 0);
  if (!hes) return bad1("swalbe1d_upload: byte count does not match the field size");
  if (field == SWALBE1D_BORDER1 || field == SWALBE1D_BORDER2) {
    for (int k = 0; k < 2; ++k)
      if (!s->border[k])
        if (int rc = alloc1(s, &s->border[k], s->L)) return rc;
    s->have_border = true;
  } else if (void** o = opt_slot(s, field)) {
    if (!*o)
      if (int rc = alloc1(s, o, s->L)) return rc;
  }
  split_dists(s, field);
  std::vector<unsigned char> tmp((size_t)n * s->esz);
  for (int64_t i = 0; i < n; ++i) {
    const double v = hes == 8 ? ((const double*)host)[i] : (double)((const float*)host)[i];
    if (s->esz == 8) ((double*)tmp.data())[i] = v;
    else ((float*)tmp.data())[i] = (float)v;
  }
  SW_HIP1(hipStreamSynchronize(s->stream));
  SW_HIP1(hipMemcpy(field1(s, field), tmp.data(), tmp.size(), hipMemcpyHostToDevice));
  return SWALBE_OK;
}

int swalbe1d_download(swalbe_film1d* s, int32_t field, void* host, size_t bytes) {
  if (!s || !host || field < 0 || field >= SWALBE1D_NFIELDS)
    return bad1("swalbe1d_download: bad arguments");
  (void)hipSetDevice(s->device);
  const void* src = field1(s, field);
  if (!src) return bad1("swalbe1d_download: field not allocated");
  const int64_t n = is_dist1(field) ? 3 * (int64_t)s->L : s->L;
  const size_t hes = bytes == (size_t)n * 8 ? 8 : (bytes == (size_t)n * 4 ? 4 : 0);
  if (!hes) return bad1("swalbe1d_download: byte count does not match the field size");
  std::vector<unsigned char> tmp((size_t)n * s->esz);
  SW_HIP1(hipStreamSynchronize(s->stream));
  SW_HIP1(hipMemcpy(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i) {
    const double v = s->esz == 8 ? ((double*)tmp.data())[i] : (double)((float*)tmp.data())[i];
    if (hes == 8) ((double*)host)[i] = v;
    else ((float*)host)[i] = (float)v;
  }
  return SWALBE_OK;
}

int swalbe1d_fill(swalbe_film1d* s, int32_t field, double value) {
  if (!s || field < 0 || field >= SWALBE1D_NFIELDS) return bad1("swalbe1d_fill: bad field");
  if (field == SWALBE1D_BORDER1 || field == SWALBE1D_BORDER2)
    return bad1("swalbe1d_fill: upload the border masks");
  (void)hipSetDevice(s->device);
  if (void** o = opt_slot(s, field))
    if (!*o)
      if (int rc = alloc1(s, o, s->L)) return rc;
  split_dists(s, field);
  return fill1(s, field1(s, field), value, is_dist1(field) ? 3 * (int64_t)s->L : s->L);
}

int swalbe1d_synchronize(swalbe_film1d* s) {
  if (!s) return bad1("null state");
  (void)hipSetDevice(s->device);
  SW_HIP1(hipStreamSynchronize(s->stream));
  return SWALBE_OK;
}

int swalbe1d_equilibrium(swalbe_film1d* s, double g) {
  if (!s) return bad1("null state");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_equilibrium<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->feq, (const T*)s->hbuf[s->cur], (const T*)s->ubuf[s->cur],
                                 (T)g, s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_bgk_and_stream(swalbe_film1d* s, double tau, int32_t bound) {
  if (!s) return bad1("null state");
  if (bound && !s->have_border) return bad1("swalbe1d_bgk_and_stream: no border masks uploaded");
  (void)hipSetDevice(s->device);
  if (s->fout_buf == s->ftemp_buf) s->fout_buf = 1 - s->ftemp_buf;  // fout is overwritten
  const double omeg = 1.0 - 1.0 / tau, itau = 1.0 / tau;
  SW_BOTH1(s, {
    hipLaunchKernelGGL(k1_collide<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream, (T*)s->fo,
                       (const T*)s->feq, (const T*)s->f[s->ftemp_buf], (const T*)s->F, (T)omeg,
                       (T)itau, s->L);
    hipLaunchKernelGGL(k1_stream<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                       (T*)s->f[s->ftemp_buf], (T*)s->f[s->fout_buf], (const T*)s->fo,
                       bound ? (const T*)s->border[0] : nullptr,
                       bound ? (const T*)s->border[1] : nullptr, s->L);
  });
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_moments(swalbe_film1d* s) {
  if (!s) return bad1("null state");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_moments<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->hbuf[s->cur], (T*)s->ubuf[s->cur],
                                 (const T*)s->f[s->fout_buf], s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_filmpressure(swalbe_film1d* s, double gamma, double theta, int32_t n, int32_t m,
                          double hmin, double hcrit, int32_t use_theta_field,
                          int32_t use_gamma_field, int32_t raw) {
  if (!s || n < 0 || m < 0) return bad1("swalbe1d_filmpressure: bad arguments");
  if (raw && !((n == 9 && m == 3) || (n == 3 && m == 2))) {
    set_error("DomainError: filmpressure! (raw form) supports (n, m) = (9, 3) or (3, 2)");
    return SWALBE_EDOMAIN;
  }
  if (use_theta_field && !s->theta) return bad1("swalbe1d_filmpressure: no θ field uploaded");
  if (use_gamma_field && !s->gamma) return bad1("swalbe1d_filmpressure: no γ field uploaded");
  (void)hipSetDevice(s->device);
  if (use_theta_field)
    if (int rc = refresh_kappa(s, n, m, hmin)) return rc;
  const double kap_s = swalbe::kappa(theta, n, m, hmin);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_filmpressure<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0,
                                 s->stream, (T*)s->pressure, (const T*)s->hbuf[s->cur], (T)gamma,
                                 use_gamma_field ? (const T*)s->gamma : nullptr, (T)kap_s,
                                 use_theta_field ? (const T*)s->kap : nullptr, n, m, (T)hmin,
                                 (T)hcrit, raw, s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_grad_gamma(swalbe_film1d* s, int32_t form, double delta) {
  if (!s || (form != 0 && form != 1)) return bad1("swalbe1d_grad_gamma: bad arguments");
  if (!s->gamma) return bad1("swalbe1d_grad_gamma: no γ field uploaded");
  (void)hipSetDevice(s->device);
  if (!s->grad_gamma)
    if (int rc = alloc1(s, &s->grad_gamma, s->L)) return rc;
  SW_BOTH1(s, hipLaunchKernelGGL(k1_grad_gamma<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->grad_gamma, (const T*)s->gamma,
                                 (const T*)s->hbuf[s->cur], (T)delta, form, s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_hgradp(swalbe_film1d* s) {
  if (!s) return bad1("null state");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_hgradp<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->hgp, (const T*)s->pressure, (const T*)s->hbuf[s->cur],
                                 s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_slippage(swalbe_film1d* s, double delta, double mu) {
  if (!s) return bad1("null state");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_slippage<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->slip, (const T*)s->hbuf[s->cur], (const T*)s->ubuf[s->cur],
                                 (T)(6.0 * mu), (T)(6.0 * delta), (T)(3.0 * (delta * delta)),
                                 s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_thermal(swalbe_film1d* s, double kbt, double mu, double delta, uint64_t seed,
                     uint64_t step) {
  if (!s) return bad1("null state");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_thermal<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->kbt, (const T*)s->hbuf[s->cur],
                                 (T)(((2.0 * kbt) * mu) * 6.0), (T)delta, seed, step, s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_inclination(swalbe_film1d* s, double alpha, double fac) {
  if (!s) return bad1("null state");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_inclination<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0,
                                 s->stream, (T*)s->F, (const T*)s->hbuf[s->cur], (T)alpha, (T)fac,
                                 s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_force_sum(swalbe_film1d* s, int32_t with_thermal, int32_t grad_gamma) {
  if (!s || grad_gamma < -1 || grad_gamma > 1) return bad1("swalbe1d_force_sum: bad arguments");
  if (grad_gamma && !s->grad_gamma) return bad1("swalbe1d_force_sum: no ∇γ field");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_force<T>, dim3(blocks1d(s->L)), dim3(kNT1), 0, s->stream,
                                 (T*)s->F, (const T*)s->hgp, (const T*)s->slip,
                                 with_thermal ? (const T*)s->kbt : nullptr,
                                 grad_gamma ? (const T*)s->grad_gamma : nullptr, grad_gamma,
                                 s->L));
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_reduce(swalbe_film1d* s, int32_t field, double thresh, int32_t use_interior,
                    double out[4]) {
  if (!s || !out || field < 0 || field >= SWALBE1D_NFIELDS || is_dist1(field) || !field1(s, field))
    return bad1("swalbe1d_reduce: bad arguments");
  if (use_interior && !s->interior) return bad1("swalbe1d_reduce: no interior mask uploaded");
  (void)hipSetDevice(s->device);
  SW_BOTH1(s, hipLaunchKernelGGL(k1_reduce<T>, dim3(1), dim3(256), 0, s->stream,
                                 (const T*)field1(s, field),
                                 use_interior ? (const T*)s->interior : nullptr, s->L, thresh,
                                 s->red));
  SW_HIP1(hipMemcpyAsync(out, s->red, sizeof(double) * 4, hipMemcpyDeviceToHost, s->stream));
  SW_HIP1(hipStreamSynchronize(s->stream));
  return SWALBE_OK;
}

int swalbe1d_resident_max_length(int32_t dtype, int32_t tau_is_one, int32_t* L) {
  if (!L || (dtype != SWALBE_F64 && dtype != SWALBE_F32))
    return bad1("swalbe1d_resident_max_length: bad arguments");
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int64_t per = resident_bytes(1, dtype == SWALBE_F64 ? 8 : 4, !tau_is_one);
  *L = (int32_t)((max_lds_bytes(dev) - kResidentStaticLds) / per);
  return SWALBE_OK;
}

}  // extern "C"

namespace {

// the intermediates of the last step (pressure, h∇p, slip, noise, F, feq)
// recomputed by the per-op kernels from its input moments, which the fused
// loops leave in the other moment buffers
void materialize(swalbe_film1d* s, const swalbe_params* p, const swalbe1d_step_opts* o,
                 int64_t nsteps) {
  const int keep = s->cur;
  s->cur = 1 - s->cur;
  swalbe1d_filmpressure(s, p->gamma, p->theta, p->n, p->m, p->hmin, p->hcrit,
                        o->use_theta_field, o->use_gamma_field, 0);
  swalbe1d_hgradp(s);
  swalbe1d_slippage(s, p->delta, p->mu);
  if (o->thermal)
    swalbe1d_thermal(s, p->kbt, p->mu, p->delta, o->seed, (uint64_t)(o->t0 + nsteps));
  swalbe1d_force_sum(s, o->thermal, o->grad_gamma);
  if (o->body) swalbe1d_inclination(s, o->alpha, o->body_fac);
  swalbe1d_equilibrium(s, p->g);
  s->cur = keep;
}

// Device buffers of one film's series (k1_resident writes them directly).
struct Series1 {
  double *dh = nullptr, *wet = nullptr, *mass = nullptr, *snap = nullptr;
  int64_t nsteps = 0, nmass = 0, nsnap = 0;
  int alloc(swalbe_film1d* s, const swalbe_params* p, const swalbe1d_step_opts* o, int64_t n) {
    nsteps = n;
    nmass = o->mass_series ? count_every(o->t0, n, p->tdump) : 0;
    nsnap = o->snap_series ? count_every(o->t0, n, o->snap_every) : 0;
    bool ok = true;
    if (o->dh_series && n) ok = ok && hipMalloc(&dh, 8 * n) == hipSuccess;
    if (o->wet_series && n) ok = ok && hipMalloc(&wet, 8 * n) == hipSuccess;
    if (nmass) ok = ok && hipMalloc(&mass, 8 * nmass) == hipSuccess;
    if (nsnap) ok = ok && hipMalloc(&snap, 8 * nsnap * (int64_t)s->L) == hipSuccess;
    if (!ok) {
      release();
      set_error("swalbe1d_time_loop: allocation of the series failed");
      return SWALBE_ENOMEM;
    }
    return SWALBE_OK;
  }
  hipError_t fetch(swalbe_film1d* s, const swalbe1d_step_opts* o, hipStream_t st) {
    hipError_t e = hipSuccess;
    std::vector<double> w;
    if (dh) e = hipMemcpyAsync(o->dh_series, dh, 8 * nsteps, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && wet) {
      w.resize(nsteps);
      e = hipMemcpyAsync(w.data(), wet, 8 * nsteps, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess && mass)
      e = hipMemcpyAsync(o->mass_series, mass, 8 * nmass, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && snap)
      e = hipMemcpyAsync(o->snap_series, snap, 8 * nsnap * (int64_t)s->L, hipMemcpyDeviceToHost,
                         st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (size_t i = 0; e == hipSuccess && i < w.size(); ++i) o->wet_series[i] = (int64_t)w[i];
    return e;
  }
  void release() {
    for (double* p : {dh, wet, mass, snap})
      if (p) (void)hipFree(p);
    dh = wet = mass = snap = nullptr;
  }
};

// after a fused loop of nsteps > 0: output moments current, fout = ftemp
void finish(swalbe_film1d* s, int64_t nsteps) {
  if (nsteps <= 0) return;
  const int out = 1 - s->ftemp_buf;
  // k1_step alternates the buffers every step; k1_resident writes once
  s->fout_buf = s->ftemp_buf = out;
  s->cur = 1 - s->cur;
}

template <typename T, int NT>
hipError_t launch_resident_t(const Film1* d_table, int nfilms, int64_t nsteps, int64_t lds,
                             hipStream_t st) {
  if (lds + kResidentStaticLds > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void*)k1_resident<T, NT>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k1_resident<T, NT>), dim3(nfilms), dim3(NT), (size_t)lds, st, d_table,
                     nsteps);
  return hipGetLastError();
}

// threads per film (0 = automatic).  Measured on MI355X, FP64, L = 1024:
// with at least one film per CU, 512 threads (2 films per CU: 8 waves each,
// up to 128 VGPRs for the per-node constants held in registers) step the
// ensemble fastest (155 against 137 GLUPS at 256 threads); a few films are
// latency bound and run fastest at one node per thread (1024).
int resident_threads(int requested, int maxL, int nfilms, int device) {
  if (requested == 256 || requested == 512 || requested == 1024) return requested;
  if (maxL <= 256) return 256;
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    cus = 256;
  if (maxL > 512 && nfilms < cus) return 1024;
  return 512;
}

template <typename T>
hipError_t launch_resident(const Film1* d_table, int nfilms, int64_t nsteps, int64_t lds, int nt,
                           hipStream_t st) {
  if (nt == 256) return launch_resident_t<T, 256>(d_table, nfilms, nsteps, lds, st);
  if (nt == 512) return launch_resident_t<T, 512>(d_table, nfilms, nsteps, lds, st);
  return launch_resident_t<T, 1024>(d_table, nfilms, nsteps, lds, st);
}

// one launch per step (any L): the moments and distributions ping-pong
// between the two buffers of each
int loop_per_step(swalbe_film1d* s, const swalbe_params* p, int64_t nsteps,
                  const swalbe1d_step_opts* o) {
  unsigned long long *keys = nullptr, *wet = nullptr;
  double* mass = nullptr;
  void* snap = nullptr;
  const int64_t nmass = o->mass_series ? count_every(o->t0, nsteps, p->tdump) : 0;
  const int64_t nsnap = o->snap_series ? count_every(o->t0, nsteps, o->snap_every) : 0;
  const int64_t ns = std::max<int64_t>(1, nsteps);
  bool ok = true;
  if (o->dh_series) ok = ok && hipMalloc(&keys, 16 * ns) == hipSuccess;
  if (o->wet_series) ok = ok && hipMalloc(&wet, 8 * ns) == hipSuccess;
  if (nmass) ok = ok && hipMalloc(&mass, 32 * nmass) == hipSuccess;
  if (nsnap) ok = ok && hipMalloc(&snap, (size_t)nsnap * s->L * s->esz) == hipSuccess;
  hipError_t e = hipSuccess;
  if (ok) {
    if (keys) hipLaunchKernelGGL(k1_init_keys, dim3(64), dim3(256), 0, s->stream, keys, ns);
    if (wet) hipLaunchKernelGGL(k1_zero_u64, dim3(64), dim3(256), 0, s->stream, wet, ns);
    Film1 S = describe(s, p, o);
    int a = s->ftemp_buf, c = s->cur;
    int64_t km = 0, ks = 0;
    // nodes (threads) per workgroup of k1_step: opts.threads, else 256
    const int nbt = o->threads ? o->threads : 256;
    const int nb = (s->L + nbt * k1_npt(nbt) - 1) / (nbt * k1_npt(nbt));
    for (int64_t it = 1; it <= nsteps; ++it) {
      const int64_t t = o->t0 + it;
      S.h_in = s->hbuf[c];
      S.u_in = s->ubuf[c];
      S.h_out = s->hbuf[1 - c];
      S.u_out = s->ubuf[1 - c];
      S.ft = s->f[a];
      S.f_out = s->f[1 - a];
      if (mass && t % p->tdump == 0) {
        SW_BOTH1(s, hipLaunchKernelGGL(k1_reduce<T>, dim3(1), dim3(256), 0, s->stream,
                                       (const T*)S.h_in, (const T*)S.interior, s->L, 0.0,
                                       mass + 4 * km));
        ++km;
      }
      unsigned long long* kk = keys ? keys + 2 * (it - 1) : nullptr;
      unsigned long long* ww = wet ? wet + (it - 1) : nullptr;
      if (nbt == 1024)
        SW_BOTH1(s, hipLaunchKernelGGL((k1_step<T, 1024>), dim3(nb), dim3(1024), 0, s->stream, S,
                                       (uint64_t)t, kk, ww));
      else if (nbt == 512)
        SW_BOTH1(s, hipLaunchKernelGGL((k1_step<T, 512>), dim3(nb), dim3(512), 0, s->stream, S,
                                       (uint64_t)t, kk, ww));
      else
        SW_BOTH1(s, hipLaunchKernelGGL((k1_step<T, 256>), dim3(nb), dim3(256), 0, s->stream, S,
                                       (uint64_t)t, kk, ww));
      if (snap && t % o->snap_every == 0) {
        (void)hipMemcpyAsync((char*)snap + ks * s->L * s->esz, S.h_out, s->L * s->esz,
                       hipMemcpyDeviceToDevice, s->stream);
        ++ks;
      }
      a = 1 - a;
      c = 1 - c;
    }
    if (nsteps > 0) {
      s->fout_buf = s->ftemp_buf = a;
      s->cur = c;
    }
    e = hipGetLastError();
    if (e == hipSuccess && keys) {
      std::vector<unsigned long long> k(2 * nsteps);
      e = hipMemcpyAsync(k.data(), keys, k.size() * 8, hipMemcpyDeviceToHost, s->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
      for (int64_t i = 0; i < nsteps && e == hipSuccess; ++i)
        o->dh_series[i] = dkey_inv(k[2 * i]) - dkey_inv(k[2 * i + 1]);
    }
    if (e == hipSuccess && wet) {
      std::vector<unsigned long long> w(nsteps);
      e = hipMemcpyAsync(w.data(), wet, w.size() * 8, hipMemcpyDeviceToHost, s->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
      for (int64_t i = 0; i < nsteps && e == hipSuccess; ++i) o->wet_series[i] = (int64_t)w[i];
    }
    if (e == hipSuccess && mass) {
      std::vector<double> mm(4 * nmass);
      e = hipMemcpyAsync(mm.data(), mass, mm.size() * 8, hipMemcpyDeviceToHost, s->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
      for (int64_t i = 0; i < nmass && e == hipSuccess; ++i) o->mass_series[i] = mm[4 * i];
    }
    if (e == hipSuccess && snap) {
      std::vector<unsigned char> sb((size_t)nsnap * s->L * s->esz);
      e = hipMemcpyAsync(sb.data(), snap, sb.size(), hipMemcpyDeviceToHost, s->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
      for (int64_t i = 0; i < nsnap * s->L && e == hipSuccess; ++i)
        o->snap_series[i] =
            s->esz == 8 ? ((double*)sb.data())[i] : (double)((float*)sb.data())[i];
    }
  } else {
    set_error("swalbe1d_time_loop: allocation of the series failed");
  }
  (void)hipStreamSynchronize(s->stream);
  if (keys) (void)hipFree(keys);
  if (wet) (void)hipFree(wet);
  if (mass) (void)hipFree(mass);
  if (snap) (void)hipFree(snap);
  if (!ok) return SWALBE_ENOMEM;
  if (e != hipSuccess) return hip_fail(e, "swalbe1d_time_loop");
  return SWALBE_OK;
}

// k1_resident over a list of films (one dtype, one device), launched on st
int loop_resident(swalbe_film1d* const* films, int nfilms, const swalbe_params* params,
                  int64_t nsteps, const swalbe1d_step_opts* const* opts, hipStream_t st,
                  double* ms_out) {
  std::vector<Film1> table(nfilms);
  std::vector<Series1> series(nfilms);
  int64_t lds = 0;
  int maxL = 0;
  int rc = SWALBE_OK;
  for (int i = 0; i < nfilms && rc == SWALBE_OK; ++i) {
    swalbe_film1d* s = films[i];
    const swalbe1d_step_opts* o = opts[i];
    if (o->use_theta_field) rc = refresh_kappa(s, params[i].n, params[i].m, params[i].hmin);
    if (rc == SWALBE_OK) rc = series[i].alloc(s, &params[i], o, nsteps);
    if (rc == SWALBE_OK && hipStreamSynchronize(s->stream) != hipSuccess)
      rc = bad1("swalbe1d: stream error before the resident loop");
    if (rc != SWALBE_OK) break;
    table[i] = describe(s, &params[i], o);
    table[i].dh = series[i].dh;
    table[i].wet = series[i].wet;
    table[i].mass = series[i].mass;
    table[i].snap = series[i].snap;
    lds = std::max(lds, resident_bytes(s->L, s->esz, table[i].read_ft));
    maxL = std::max(maxL, s->L);
  }
  Film1* d_table = nullptr;
  hipError_t e = hipSuccess;
  if (rc == SWALBE_OK && nsteps > 0) {
    e = hipMalloc(&d_table, sizeof(Film1) * nfilms);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_table, table.data(), sizeof(Film1) * nfilms, hipMemcpyHostToDevice, st);
    const int nt = resident_threads(opts[0]->threads, maxL, nfilms, films[0]->device);
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (e == hipSuccess && ms_out) {
      e = hipEventCreate(&ev[0]);
      if (e == hipSuccess) e = hipEventCreate(&ev[1]);
      if (e == hipSuccess) e = hipEventRecord(ev[0], st);
    }
    if (e == hipSuccess)
      e = films[0]->esz == 8 ? launch_resident<double>(d_table, nfilms, nsteps, lds, nt, st)
                             : launch_resident<float>(d_table, nfilms, nsteps, lds, nt, st);
    if (e == hipSuccess && ms_out) e = hipEventRecord(ev[1], st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && ms_out) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, ev[0], ev[1]);
      *ms_out = ms;
    }
    for (hipEvent_t v : ev)
      if (v) (void)hipEventDestroy(v);
    for (int i = 0; i < nfilms && e == hipSuccess; ++i) {
      finish(films[i], nsteps);
      e = series[i].fetch(films[i], opts[i], st);
    }
  }
  for (auto& sr : series) sr.release();
  if (d_table) (void)hipFree(d_table);
  if (rc != SWALBE_OK) return rc;
  if (e != hipSuccess) return hip_fail(e, "swalbe1d resident loop");
  return SWALBE_OK;
}

}  // namespace

extern "C" {

int swalbe1d_time_loop(swalbe_film1d* s, const swalbe_params* p, int64_t nsteps,
                       const swalbe1d_step_opts* o) {
  if (!s || !p || nsteps < 0) return bad1("swalbe1d_time_loop: bad arguments");
  swalbe1d_step_opts zero{};
  if (!o) o = &zero;
  if (int rc = check_opts(s, p, o)) return rc;
  (void)hipSetDevice(s->device);
  if (o->variant == 2 && !fits_resident(s, p))
    return bad1("swalbe1d_time_loop: the film does not fit in LDS (variant 2)");
  const bool resident = o->variant == 2 || (o->variant == 0 && fits_resident(s, p));
  int rc;
  if (resident) {
    rc = loop_resident(&s, 1, p, nsteps, &o, s->stream, nullptr);
  } else {
    if (o->use_theta_field)
      if ((rc = refresh_kappa(s, p->n, p->m, p->hmin))) return rc;
    rc = loop_per_step(s, p, nsteps, o);
  }
  if (rc == SWALBE_OK && o->materialize && nsteps > 0) {
    materialize(s, p, o, nsteps);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, "swalbe1d_time_loop (materialize)");
  }
  return rc;
}

int swalbe1d_ensemble_time_loop_timed(swalbe_film1d* const* films, int32_t nfilms,
                                      const swalbe_params* params, int64_t nsteps,
                                      const swalbe1d_step_opts* opts, double* ms_out) {
  if (!films || nfilms < 1 || !params || nsteps < 0)
    return bad1("swalbe1d_ensemble_time_loop: bad arguments");
  swalbe1d_step_opts zero{};
  std::vector<const swalbe1d_step_opts*> po(nfilms);
  for (int i = 0; i < nfilms; ++i) {
    if (!films[i]) return bad1("swalbe1d_ensemble_time_loop: null film");
    if (films[i]->esz != films[0]->esz || films[i]->device != films[0]->device)
      return bad1("swalbe1d_ensemble_time_loop: films differ in dtype or device");
    po[i] = opts ? &opts[i] : &zero;
    if (int rc = check_opts(films[i], &params[i], po[i])) return rc;
    if (po[i]->variant == 1)
      return bad1("swalbe1d_ensemble_time_loop: ensembles run resident (variant 0 or 2)");
    if (!fits_resident(films[i], &params[i]))
      return bad1("swalbe1d_ensemble_time_loop: a film does not fit in LDS");
  }
  {
    std::vector<swalbe_film1d*> sorted(films, films + nfilms);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return bad1("swalbe1d_ensemble_time_loop: a film is listed twice");
  }
  (void)hipSetDevice(films[0]->device);
  int rc = loop_resident(films, nfilms, params, nsteps, po.data(), films[0]->stream, ms_out);
  for (int i = 0; i < nfilms && rc == SWALBE_OK; ++i)
    if (po[i]->materialize && nsteps > 0) {
      materialize(films[i], &params[i], po[i], nsteps);
      if (hipStreamSynchronize(films[i]->stream) != hipSuccess)
        rc = bad1("swalbe1d_ensemble_time_loop: materialize failed");
    }
  return rc;
}

int swalbe1d_ensemble_time_loop(swalbe_film1d* const* films, int32_t nfilms,
                                const swalbe_params* params, int64_t nsteps,
                                const swalbe1d_step_opts* opts) {
  return swalbe1d_ensemble_time_loop_timed(films, nfilms, params, nsteps, opts, nullptr);
}

int swalbe1d_time_loop_timed(swalbe_film1d* s, const swalbe_params* p, int64_t nsteps,
                             const swalbe1d_step_opts* o, double* ms_out) {
  if (!s || !ms_out) return bad1("swalbe1d_time_loop_timed: null argument");
  (void)hipSetDevice(s->device);
  hipEvent_t a, b;
  SW_HIP1(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return bad1("swalbe1d_time_loop_timed: event creation failed");
  }
  (void)hipEventRecord(a, s->stream);
  const int rc = swalbe1d_time_loop(s, p, nsteps, o);
  (void)hipEventRecord(b, s->stream);
  (void)hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  *ms_out = ms;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

int swalbe1d_grad_ptr(int32_t dtype, void* out, const void* f, const void* a, int32_t L,
                      void* stream) {
  if (!out || !f || L < 1 || (dtype != SWALBE_F64 && dtype != SWALBE_F32))
    return bad1("swalbe1d_grad_ptr: bad arguments");
  if (dtype == SWALBE_F64)
    hipLaunchKernelGGL(k1_grad<double>, dim3(blocks1d(L)), dim3(kNT1), 0, (hipStream_t)stream,
                       (double*)out, (const double*)f, (const double*)a, L);
  else
    hipLaunchKernelGGL(k1_grad<float>, dim3(blocks1d(L)), dim3(kNT1), 0, (hipStream_t)stream,
                       (float*)out, (const float*)f, (const float*)a, L);
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

int swalbe1d_laplacian_ptr(int32_t dtype, void* out, const void* f, int32_t L, void* stream) {
  if (!out || !f || L < 1 || (dtype != SWALBE_F64 && dtype != SWALBE_F32))
    return bad1("swalbe1d_laplacian_ptr: bad arguments");
  if (dtype == SWALBE_F64)
    hipLaunchKernelGGL(k1_laplacian<double>, dim3(blocks1d(L)), dim3(kNT1), 0, (hipStream_t)stream,
                       (double*)out, (const double*)f, L);
  else
    hipLaunchKernelGGL(k1_laplacian<float>, dim3(blocks1d(L)), dim3(kNT1), 0, (hipStream_t)stream,
                       (float*)out, (const float*)f, L);
  SW_HIP1(hipGetLastError());
  return SWALBE_OK;
}

}  // extern "C"
