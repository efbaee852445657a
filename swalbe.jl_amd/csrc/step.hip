// step.hip -- the fused D2Q9 thin-film time step (one kernel per step).
//
// One reference step (src/simulate.jl:15-22) is
//   p = filmpressure(h)          3x3 stencil of h
//   F = -h∇p - slip(h,u) [-kbt] [+body]   3x3 stencil of p
//   fo = (1-1/τ) ftemp + feq(h,u)/τ + S(F)
//   f_{t+1,k}(x) = fo_k(x - c_k)  (pull stream)
//   h, u = moments(f_{t+1})
// so an output node needs F and u on a 1-node halo, p on 2 and h on 3.
// A workgroup owns a 64 x TY tile: it stages h (3-halo) and u (1-halo) in
// LDS, computes p on the 2-halo and F on the 1-halo in LDS, then every
// thread collides the nine source populations of its output nodes, writes
// the streamed f and the new moments.  HBM traffic per lattice update is
// h, u in (+ f_t when τ ≠ 1) and f, h, u, p out.
//
// FP64: every value equals the unfused reference sequence bit for bit (same
// expressions, same order, no contraction).  FP32: height and distributions
// are handled as deviations from a rest film of height href (Enc).
#include "swalbe_internal.h"

namespace swalbe {

namespace {

constexpr int TX = 64;        // tile width = one wavefront
constexpr int NT = 256;       // threads per workgroup (4 wavefronts)

__device__ __forceinline__ int wrap1(int a, int L) {
  return a < 0 ? a + L : (a >= L ? a - L : a);
}
__device__ __forceinline__ int wrapm(int a, int L) {
  a %= L;
  return a < 0 ? a + L : a;
}

template <typename T>
__device__ __forceinline__ T powlaw(T a, int n, int m, int mode) {
  if (mode == 1) {
    const T t = (a * a) * a;
    return ((t * t) * t) - t;
  }
  if (mode == 2) return ((a * a) * a) - (a * a);
  T tn = T(1), tm = T(1);
  for (int i = 0; i < n; ++i) tn *= a;
  for (int i = 0; i < m; ++i) tm *= a;
  return tn - tm;
}

// Equilibrium population k at a node (src/equilibrium.jl:71-114), plain values.
template <typename T>
__device__ __forceinline__ T feq_k(int k, T h, T u, T v, T usq, T c56g, T g0) {
  const T w1 = T(1.0 / 9.0), w5 = T(1.0 / 36.0);
  T a;
  switch (k) {
    case 0: return h * ((T(1) - c56g * h) - T(2.0 / 3.0) * usq);
    case 1: return w1 * h * (((g0 * h + T(3) * u) + T(4.5) * (u * u)) - T(1.5) * usq);
    case 2: return w1 * h * (((g0 * h + T(3) * v) + T(4.5) * (v * v)) - T(1.5) * usq);
    case 3: return w1 * h * (((g0 * h - T(3) * u) + T(4.5) * (u * u)) - T(1.5) * usq);
    case 4: return w1 * h * (((g0 * h - T(3) * v) + T(4.5) * (v * v)) - T(1.5) * usq);
    case 5: a = u + v; return w5 * h * (((g0 * h + T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
    case 6: a = v - u; return w5 * h * (((g0 * h + T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
    case 7: a = u + v; return w5 * h * (((g0 * h - T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
    default: a = u - v; return w5 * h * (((g0 * h + T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
  }
}

// Deviation feq_k(h, u) - feq_k(href, 0) with h = href + dh, free of the
// cancellation of the plain form (FP32 path).
__device__ __forceinline__ float feq_dev(int k, float dh, float h, float href, float u, float v,
                                         float usq, float c56g, float g0) {
  const float gg = dh * (h + href);  // h^2 - href^2
  float a, s;
  switch (k) {
    case 0: return (dh - c56g * gg) - (2.0f / 3.0f) * h * usq;
    case 1: a = u; s = 1.f; break;
    case 2: a = v; s = 1.f; break;
    case 3: a = u; s = -1.f; break;
    case 4: a = v; s = -1.f; break;
    case 5: a = u + v; s = 1.f; break;
    case 6: a = v - u; s = 1.f; break;
    case 7: a = u + v; s = -1.f; break;
    default: a = u - v; s = 1.f; break;
  }
  const float w = k <= 4 ? (1.0f / 9.0f) : (1.0f / 36.0f);
  return w * (g0 * gg + h * ((s * 3.0f * a + 4.5f * (a * a)) - 1.5f * usq));
}

template <typename T>
__device__ __forceinline__ T force_term(int k, T Fx, T Fy) {
  const T c3 = T(1.0 / 3.0), c24 = T(1.0 / 24.0);
  switch (k) {
    case 1: return c3 * Fx;
    case 2: return c3 * Fy;
    case 3: return -(c3 * Fx);
    case 4: return -(c3 * Fy);
    case 5: return c24 * (Fx + Fy);
    case 6: return c24 * (Fy - Fx);
    case 7: return -(c24 * (Fx + Fy));
    case 8: return c24 * (Fx - Fy);
    default: return T(0);
  }
}

// The force on one node.
template <typename T>
struct Fpair {
  T x, y;
};

// A store that streams past the caches (non-temporal) when NT.  The choice is
// a template argument: behind a run-time branch the compiler merges the two
// forms into one plain store.
template <bool NT, typename T>
__device__ __forceinline__ void store_t(T* p, T v) {
  using G = __attribute__((address_space(1))) T;
  if constexpr (NT) __builtin_nontemporal_store(v, (G*)p);
  else *(G*)p = v;
}

// p, a pointer uniform over the wave, as the compiler's scalar registers hold
// it: an address formed from it plus a 32-bit lane offset is a store's scalar
// base and offset, with no 64-bit vector add per store.
__device__ __forceinline__ __attribute__((address_space(1))) char* scalar_base(const void* p) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return (__attribute__((address_space(1))) char*)(((uint64_t)hi << 32) | lo);
}

// NTS: 0 every output cached, 1 every output streaming, 2 the distributions
// and p streaming while h, ux, uy (the next step's input) stay cached.  AL:
// the (Fx, Fy) planes overwrite p once every wave has read it (a barrier
// between; the forces wait in registers), so p and F share one LDS buffer:
// 29 KB instead of 35.5 KB for a 64 x 8 tile, 5 workgroups per CU instead of 4.
template <typename T, int TY, bool READ_FT, bool THERMAL, int NTS, bool AL>
__global__ __launch_bounds__(NT) void k_step(StepIO<T> io, StepParams sp, StepSeries ser,
                                             uint64_t step, int Lx, int Ly, int ntx, int ntiles,
                                             int small, int64_t FS, Enc enc) {
  constexpr bool DEV = sizeof(T) == 4;
  constexpr int HC = TX + 6, HR = TY + 6;  // h tile, 3-halo
  constexpr int PC = TX + 4, PR = TY + 4;  // p tile, 2-halo
  constexpr int UC = TX + 2, UR = TY + 2;  // u and F tiles, 1-halo
  constexpr int UN = UR * UC, PN = PR * PC;
  constexpr int PFN = AL ? (PN > 2 * UN ? PN : 2 * UN) : PN + 2 * UN;
  __shared__ T sh_h[HR * HC];
  __shared__ T sh_pf[PFN];  // p, and (Fx, Fy): after p when !AL, over it when AL
  __shared__ T sh_ux[UR * UC], sh_uy[UR * UC];
  T* const sh_p = sh_pf;
  T* const sh_fx = AL ? sh_pf : sh_pf + PN;
  T* const sh_fy = sh_fx + UN;
  __shared__ unsigned long long red_k[2 * (NT / 64)];
  __shared__ unsigned int red_c[NT / 64];

  const int tile = tile_of_block_f64(ntx, ntiles, step, 3 * (int64_t)sizeof(T) * Lx * Ly);
  const int x0 = (tile % ntx) * TX;
  const int y0 = (tile / ntx) * TY;
  const int tid = threadIdx.x;
  const int64_t N = (int64_t)Lx * Ly;

  // ---- phase A: stage h (3-halo) and u (1-halo).  Every load is issued
  // into registers before the first LDS store (memory-level parallelism);
  // interior tiles (no wrap) address by 32-bit byte offsets from the halo's
  // corner.
  {
    constexpr int HN = HR * HC, UN = UR * UC;
    constexpr int NHI = (HN + NT - 1) / NT, NUI = (UN + NT - 1) / NT;
    T hreg[NHI], ureg[NUI], vreg[NUI];
    __builtin_amdgcn_s_setprio(2);  // issue the tile's loads ahead of other waves' arithmetic
    if (!small && x0 >= 3 && y0 >= 3 && x0 + TX + 3 <= Lx && y0 + TY + 3 <= Ly &&
        Lx <= (1 << 23)) {
      const char* hb = (const char*)(io.h_in + ((x0 - 3) + (int64_t)Lx * (y0 - 3)));
      const char* ub = (const char*)(io.ux_in + ((x0 - 1) + (int64_t)Lx * (y0 - 1)));
      const char* vb = (const char*)(io.uy_in + ((x0 - 1) + (int64_t)Lx * (y0 - 1)));
#pragma unroll
      for (int i = 0; i < NHI; ++i) {
        const int e = tid + i * NT;
        if (e < HN) {
          const int r = e / HC, c = e - r * HC;
          hreg[i] = *(const T*)(hb + (uint32_t)((int)sizeof(T) * (c + Lx * r)));
        }
      }
#pragma unroll
      for (int i = 0; i < NUI; ++i) {
        const int e = tid + i * NT;
        if (e < UN) {
          const int r = e / UC, c = e - r * UC;
          const uint32_t o = (uint32_t)((int)sizeof(T) * (c + Lx * r));
          ureg[i] = *(const T*)(ub + o);
          vreg[i] = *(const T*)(vb + o);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < NHI; ++i) {
        const int e = tid + i * NT;
        if (e < HN) {
          const int r = e / HC, c = e - r * HC;
          int gx = x0 - 3 + c, gy = y0 - 3 + r;
          gx = small ? wrapm(gx, Lx) : wrap1(gx, Lx);
          gy = small ? wrapm(gy, Ly) : wrap1(gy, Ly);
          hreg[i] = io.h_in[gx + (int64_t)Lx * gy];
        }
      }
#pragma unroll
      for (int i = 0; i < NUI; ++i) {
        const int e = tid + i * NT;
        if (e < UN) {
          const int r = e / UC, c = e - r * UC;
          int gx = x0 - 1 + c, gy = y0 - 1 + r;
          gx = small ? wrapm(gx, Lx) : wrap1(gx, Lx);
          gy = small ? wrapm(gy, Ly) : wrap1(gy, Ly);
          const int64_t g = gx + (int64_t)Lx * gy;
          ureg[i] = io.ux_in[g];
          vreg[i] = io.uy_in[g];
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int i = 0; i < NHI; ++i) {
      const int e = tid + i * NT;
      if (e < HN) sh_h[e] = hreg[i];
    }
#pragma unroll
    for (int i = 0; i < NUI; ++i) {
      const int e = tid + i * NT;
      if (e < UN) {
        sh_ux[e] = ureg[i];
        sh_uy[e] = vreg[i];
      }
    }
  }
  __syncthreads();

  const T gamma = (T)sp.gamma, hmin = (T)sp.hmin, hcrit = (T)sp.hcrit;
  const T href = DEV ? (T)enc.href : T(0);

  // ---- phase B: film pressure on the 2-halo (src/pressure.jl:119-155)
  for (int idx = tid; idx < PR * PC; idx += NT) {
    const int r = idx / PC, c = idx - r * PC;
    const T* hc = &sh_h[(r + 1) * HC + (c + 1)];
    const T hip = hc[-1], him = hc[1], hjp = hc[-HC], hjm = hc[HC];
    const T hipjp = hc[-HC - 1], himjp = hc[-HC + 1], himjm = hc[HC + 1], hipjm = hc[HC - 1];
    const T L = (T(2.0 / 3.0) * (((hjp + hip) + him) + hjm) +
                 T(1.0 / 6.0) * (((hipjp + himjp) + himjm) + hipjm)) -
                T(10.0 / 3.0) * hc[0];
    const T hh = DEV ? href + hc[0] : hc[0];
    T kap = (T)sp.kappa;
    if (io.kappa) {
      int gx = x0 - 2 + c, gy = y0 - 2 + r;
      gx = small ? wrapm(gx, Lx) : wrap1(gx, Lx);
      gy = small ? wrapm(gy, Ly) : wrap1(gy, Ly);
      kap = io.kappa[gx + (int64_t)Lx * gy];
    }
    const T a = hmin / (hh + hcrit);
    const T p0 = -gamma * (kap * powlaw(a, sp.n, sp.m, sp.powmode));
    sh_p[idx] = p0 - gamma * L;
  }
  __syncthreads();

  // ---- phase C: forces on the 1-halo (src/forcing.jl:42-46, 168-187, 297-311, 363-368).
  // With AL, p at this thread's phase D nodes waits in registers and is
  // stored after the block's last barrier.
  constexpr int NRW = TY / (NT / TX);  // phase D rows per thread
  T pown[NRW];
  auto keep_p = [&]() {
#pragma unroll
    for (int j = 0; j < NRW; ++j)
      pown[j] = sh_p[(tid / TX + j * (NT / TX) + 2) * PC + (tid % TX + 2)];
  };
  auto put = [&](int idx, const Fpair<T>& F) {
    sh_fx[idx] = F.x;
    sh_fy[idx] = F.y;
  };
  {
    const T c6mu = (T)sp.c6mu, c6d = (T)sp.c6d, c3dd = (T)sp.c3dd;
    auto fnode = [&](int r, int c, double z0, double z1) {
      const int idx = r * UC + c;
      const T* pc = &sh_p[(r + 1) * PC + (c + 1)];
      const T fip = pc[-1], fim = pc[1], fjp = pc[-PC], fjm = pc[PC];
      const T fipjp = pc[-PC - 1], fimjp = pc[-PC + 1], fimjm = pc[PC + 1], fipjm = pc[PC - 1];
      const T hs = sh_h[(r + 2) * HC + (c + 2)];
      const T hh = DEV ? href + hs : hs;
      const T gx = T(-1.0 / 3.0) * (fip - fim) - T(1.0 / 12.0) * (((fipjp - fimjp) - fimjm) + fipjm);
      const T gy = T(-1.0 / 3.0) * (fjp - fjm) - T(1.0 / 12.0) * (((fipjp + fimjp) - fimjm) - fipjm);
      const T hgx = hh * gx, hgy = hh * gy;
      const T den = (T(2) * (hh * hh) + c6d * hh) + c3dd;
      const T sx = (c6mu * hh * sh_ux[idx]) / den;
      const T sy = (c6mu * hh * sh_uy[idx]) / den;
      T Fx = -hgx - sx;
      T Fy = -hgy - sy;
      if constexpr (THERMAL) {
        const T amp = sqrt((T)sp.kb26 * hh /
                           (((T(2) * hh) * hh + (T(6) * hh) * (T)sp.delta) + (T)sp.d3dd));
        Fx = Fx - (T)z0 * amp;
        Fy = Fy - (T)z1 * amp;
      }
      if (sp.body) {
        Fx = Fx + (hh * (T)sp.bax) * (T)sp.bfac;
        Fy = Fy + (hh * (T)sp.bay) * (T)sp.bfac;
      }
      if constexpr (!AL) {
        sh_fx[idx] = Fx;
        sh_fy[idx] = Fy;
      }
      return Fpair<T>{Fx, Fy};
    };
    if (!small && (Ly & 1) == 0) {
      // one Philox call per vertical node pair (x, 2m), (x, 2m+1): pair row
      // q covers halo rows 2q-1 (even y) and 2q (odd y); y0 is even
      constexpr int NPR = TY / 2 + 2, NIT = (NPR * UC + NT - 1) / NT;
      auto pair = [&](int t, Fpair<T>* Fp) {
        const int q = t / UC, c = t - q * UC;
        const int re = 2 * q - 1, ro = 2 * q;
        double z[4] = {0.0, 0.0, 0.0, 0.0};
        if constexpr (THERMAL) {
          const int gxx = wrap1(x0 - 1 + c, Lx);
          const int ye = wrap1(y0 - 2 + 2 * q, Ly);
          uint32_t w[4];
          philox_ctr(sp.seed, step, (uint32_t)gxx + (uint32_t)Lx * (uint32_t)(ye >> 1), kTagThermal,
                     w);
          box_muller(u32(w[0]), u32(w[1]), &z[0], &z[1]);
          box_muller(u32(w[2]), u32(w[3]), &z[2], &z[3]);
        }
        if (re >= 0) Fp[0] = fnode(re, c, z[0], z[1]);
        if (ro < UR) Fp[1] = fnode(ro, c, z[2], z[3]);
      };
      if constexpr (AL) {
        Fpair<T> Fq[NIT][2];
#pragma unroll
        for (int i = 0; i < NIT; ++i)
          if (tid + i * NT < NPR * UC) pair(tid + i * NT, Fq[i]);
        keep_p();
        __syncthreads();  // every wave has read p
#pragma unroll
        for (int i = 0; i < NIT; ++i) {
          const int t = tid + i * NT;
          if (t < NPR * UC) {
            const int q = t / UC, c = t - q * UC, re = 2 * q - 1, ro = 2 * q;
            if (re >= 0) put(re * UC + c, Fq[i][0]);
            if (ro < UR) put(ro * UC + c, Fq[i][1]);
          }
        }
      } else {
        for (int t = tid; t < NPR * UC; t += NT) {
          Fpair<T> Fp[2];  // fnode has stored them
          pair(t, Fp);
        }
      }
    } else {
      auto node = [&](int idx) {
        const int r = idx / UC, c = idx - r * UC;
        double z0 = 0.0, z1 = 0.0;
        if constexpr (THERMAL) {
          int gxx = x0 - 1 + c, gyy = y0 - 1 + r;
          gxx = small ? wrapm(gxx, Lx) : wrap1(gxx, Lx);
          gyy = small ? wrapm(gyy, Ly) : wrap1(gyy, Ly);
          normal_pair(sp.seed, step, noise_index(gxx, gyy, Lx), kTagThermal, &z0, &z1);
        }
        return fnode(r, c, z0, z1);
      };
      if constexpr (AL) {
        constexpr int NIU = (UN + NT - 1) / NT;
        Fpair<T> Fn[NIU];
#pragma unroll
        for (int i = 0; i < NIU; ++i)
          if (tid + i * NT < UN) Fn[i] = node(tid + i * NT);
        keep_p();
        __syncthreads();  // every wave has read p
#pragma unroll
        for (int i = 0; i < NIU; ++i)
          if (tid + i * NT < UN) put(tid + i * NT, Fn[i]);
      } else {
        for (int idx = tid; idx < UN; idx += NT) node(idx);
      }
    }
  }
  __syncthreads();

  __builtin_amdgcn_s_setprio(1);  // the stores ahead of other waves' pressure and forces
  // ---- phase D: collide the source populations, pull-stream, moments
  const int tx = tid & (TX - 1);
  const int ty = __builtin_amdgcn_readfirstlane(tid >> 6);  // the wave's rows: uniform
  const T omeg = (T)sp.omeg, itau = (T)sp.itau;
  const T c56g = (T)((5.0 / 6.0) * sp.g), g0 = (T)(1.5 * sp.g);
  double lmax = -__builtin_inf(), lmin = __builtin_inf();
  unsigned int lcnt = 0;
  const int gxo = x0 + tx;
  if constexpr (AL) {
    if (io.p_out) {
#pragma unroll
      for (int j = 0; j < NRW; ++j) {
        const int gyo = y0 + ty + j * (NT / TX);
        if (gxo < Lx && gyo < Ly) {
          uint32_t lo = (uint32_t)sizeof(T) * (uint32_t)tx;  // a 32-bit store offset
          asm volatile("" : "+v"(lo));
          store_t<NTS != 0>((T*)(scalar_base(io.p_out + ((int64_t)Lx * gyo + x0)) + lo), pown[j]);
        }
      }
    }
  }
#pragma unroll 1
  for (int row = ty; row < TY; row += NT / TX) {
    const int gyo = y0 + row;
    if (gxo >= Lx || gyo >= Ly) continue;
    // stores: the output row's first node is a scalar base (uniform over the
    // wave), the lane's byte offset a 32-bit value made opaque beside each
    // group of stores (hoisted, it becomes a 64-bit vector operand)
    const int64_t rowo = (int64_t)Lx * gyo + x0;
    auto lane_off = [&]() {
      uint32_t lo = (uint32_t)sizeof(T) * (uint32_t)tx;
      asm volatile("" : "+v"(lo));
      return lo;
    };
    auto at = [&](T* plane, uint32_t lo) { return (T*)(scalar_base(plane + rowo) + lo); };
    T fo[kQ];
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
      const int su = (row + 1 - cy(k)) * UC + (tx + 1 - cx(k));
      const T hs = sh_h[(row + 3 - cy(k)) * HC + (tx + 3 - cx(k))];
      const T us = sh_ux[su], vs = sh_uy[su];
      const T usq = us * us + vs * vs;
      T f;
      if constexpr (DEV) {
        f = itau * feq_dev(k, hs, href + hs, href, us, vs, usq, c56g, g0);
      } else {
        f = itau * feq_k<T>(k, hs, us, vs, usq, c56g, g0);
      }
      if constexpr (READ_FT) {
        int sx = gxo - cx(k), sy = gyo - cy(k);
        sx = small ? wrapm(sx, Lx) : wrap1(sx, Lx);
        sy = small ? wrapm(sy, Ly) : wrap1(sy, Ly);
        f = omeg * io.ft[k * FS + sx + (int64_t)Lx * sy] + f;
      }
      if (k > 0) f = f + force_term<T>(k, sh_fx[su], sh_fy[su]);
      fo[k] = f;
    }
    if (io.f_out) {
      const uint32_t lo = lane_off();
#pragma unroll
      for (int k = 0; k < kQ; ++k) store_t<NTS != 0>(at(io.f_out + k * FS, lo), fo[k]);
    }
    T s = fo[0];
#pragma unroll
    for (int k = 1; k < kQ; ++k) s = s + fo[k];
    const T hnew = DEV ? href + s : s;
    {
      const uint32_t lo = lane_off();
      store_t<NTS == 1>(at(io.h_out, lo), s);
      store_t<NTS == 1>(at(io.ux_out, lo), (((((fo[1] - fo[3]) + fo[5]) - fo[6]) - fo[7]) + fo[8]) / hnew);
      store_t<NTS == 1>(at(io.uy_out, lo), (((((fo[2] - fo[4]) + fo[5]) + fo[6]) - fo[7]) - fo[8]) / hnew);
    }
    if constexpr (!AL) {
      if (io.p_out) store_t<NTS != 0>(at(io.p_out, lane_off()), sh_p[(row + 2) * PC + (tx + 2)]);
    }
    if (ser.dh_keys || ser.wet) {
      const T hc = sh_h[(row + 3) * HC + (tx + 3)];
      const double hv = DEV ? (double)href + (double)hc : (double)hc;
      lmax = hv > lmax ? hv : lmax;
      lmin = hv < lmin ? hv : lmin;
      lcnt += hv > sp.wet_thresh ? 1u : 0u;
    }
  }

  __builtin_amdgcn_s_setprio(0);
  // ---- per-step series: max/min/count of h_t over the domain (exact)
  if (ser.dh_keys || ser.wet) {
    unsigned long long kmax = dkey(lmax), kmin = dkey(lmin);
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a = __shfl_xor(kmax, o), b = __shfl_xor(kmin, o);
      kmax = a > kmax ? a : kmax;
      kmin = b < kmin ? b : kmin;
      lcnt += __shfl_xor(lcnt, o);
    }
    if ((tid & 63) == 0) {
      red_k[2 * ty] = kmax;
      red_k[2 * ty + 1] = kmin;
      red_c[ty] = lcnt;
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < NT / 64; ++w) {
        kmax = red_k[2 * w] > kmax ? red_k[2 * w] : kmax;
        kmin = red_k[2 * w + 1] < kmin ? red_k[2 * w + 1] : kmin;
        lcnt += red_c[w];
      }
      if (ser.dh_keys) {
        atomicMax(&ser.dh_keys[0], kmax);
        atomicMin(&ser.dh_keys[1], kmin);
      }
      if (ser.wet) atomicAdd(&ser.wet[0], (unsigned long long)lcnt);
    }
  }
}

template <typename T, int TY>
void launch_ty(const StepIO<T>& io, const StepParams& sp, StepSeries ser, uint64_t step, int Lx,
               int Ly, int64_t FS, Enc enc, int nts, hipStream_t s) {
  const int ntx = (Lx + TX - 1) / TX, nty = (Ly + TY - 1) / TY;
  const int ntiles = ntx * nty;
  const int small = (Lx < TX + 6 || Ly < TY + 6) ? 1 : 0;
  dim3 grid(ntiles), block(NT);
  // The shared p / F buffer lifts a 64 x 8 tile from 4 to 5 workgroups per CU
  // (64 x 16: from 2 to 3) for one more barrier per tile: taken where the film
  // has more tiles than the CUs hold at once without it.
  constexpr int kResident = TY == 8 ? 4 : 2;
  const bool al = (int64_t)ntiles > kResident * (int64_t)device_cus();
#define SW_K(RF, TH, M)                                                                        \
  do {                                                                                         \
    if (al)                                                                                    \
      hipLaunchKernelGGL((k_step<T, TY, RF, TH, M, true>), grid, block, 0, s, io, sp, ser, step, \
                         Lx, Ly, ntx, ntiles, small, FS, enc);                                 \
    else                                                                                       \
      hipLaunchKernelGGL((k_step<T, TY, RF, TH, M, false>), grid, block, 0, s, io, sp, ser,     \
                         step, Lx, Ly, ntx, ntiles, small, FS, enc);                           \
  } while (0)
#define SW_M(RF, TH)          \
  do {                        \
    if (nts == 2) SW_K(RF, TH, 2); \
    else if (nts == 1) SW_K(RF, TH, 1); \
    else SW_K(RF, TH, 0);     \
  } while (0)
  if (sp.read_ft) {
    if (sp.thermal) SW_M(true, true);
    else SW_M(true, false);
  } else {
    if (sp.thermal) SW_M(false, true);
    else SW_M(false, false);
  }
#undef SW_M
#undef SW_K
}

// ---------------------------------------------------------------- reductions
constexpr int kRedBlock = 256;

template <typename T>
__global__ void k_reduce(const T* __restrict__ h, int64_t N, double href, double thresh,
                         double* __restrict__ partials) {
  // each block reduces a contiguous chunk in a fixed order: deterministic
  __shared__ double s_sum[kRedBlock / 64], s_max[kRedBlock / 64], s_min[kRedBlock / 64];
  __shared__ double s_cnt[kRedBlock / 64];
  const int64_t per = (N + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = blockIdx.x * per;
  const int64_t b1 = b0 + per < N ? b0 + per : N;
  double sum = 0.0, mx = -__builtin_inf(), mn = __builtin_inf(), cnt = 0.0;
  for (int64_t x = b0 + threadIdx.x; x < b1; x += kRedBlock) {
    const double v = href + (double)h[x];
    sum += v;
    mx = v > mx ? v : mx;
    mn = v < mn ? v : mn;
    cnt += v > thresh ? 1.0 : 0.0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    const double a = __shfl_xor(mx, o), b = __shfl_xor(mn, o);
    mx = a > mx ? a : mx;
    mn = b < mn ? b : mn;
    cnt += __shfl_xor(cnt, o);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_sum[w] = sum;
    s_max[w] = mx;
    s_min[w] = mn;
    s_cnt[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kRedBlock / 64; ++i) {
      sum += s_sum[i];
      mx = s_max[i] > mx ? s_max[i] : mx;
      mn = s_min[i] < mn ? s_min[i] : mn;
      cnt += s_cnt[i];
    }
    partials[4 * blockIdx.x + 0] = sum;
    partials[4 * blockIdx.x + 1] = mx;
    partials[4 * blockIdx.x + 2] = mn;
    partials[4 * blockIdx.x + 3] = cnt;
  }
}

__global__ void k_reduce_finish(const double* __restrict__ partials, int nparts,
                                double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double sum = 0.0, mx = -__builtin_inf(), mn = __builtin_inf(), cnt = 0.0;
  for (int i = 0; i < nparts; ++i) {
    sum += partials[4 * i];
    mx = partials[4 * i + 1] > mx ? partials[4 * i + 1] : mx;
    mn = partials[4 * i + 2] < mn ? partials[4 * i + 2] : mn;
    cnt += partials[4 * i + 3];
  }
  out[0] = sum;
  out[1] = mx;
  out[2] = mn;
  out[3] = cnt;
}

}  // namespace

template <typename T>
void launch_step(const StepIO<T>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                 int Lx, int Ly, int64_t fstride, Enc enc, int tile_y, int nts, hipStream_t s) {
  switch (tile_y) {
    case 8: launch_ty<T, 8>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s); break;
    default: launch_ty<T, 16>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s); break;
  }
}

template <typename T>
void launch_reduce(const T* h, int64_t N, double href, double thresh, double* partials,
                   int nparts, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce<T>, dim3(nparts), dim3(kRedBlock), 0, s, h, N, href, thresh,
                     partials);
}

void launch_reduce_finish(const double* partials, int nparts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(64), 0, s, partials, nparts, out);
}

template void launch_step<double>(const StepIO<double>&, const StepParams&, StepSeries, uint64_t,
                                  int, int, int64_t, Enc, int, int, hipStream_t);
template void launch_reduce<double>(const double*, int64_t, double, double, double*, int,
                                    hipStream_t);
template void launch_reduce<float>(const float*, int64_t, double, double, double*, int,
                                   hipStream_t);

}  // namespace swalbe
