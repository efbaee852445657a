// step_f32.hip -- the fused thin-film step for FP32 films (deviation
// storage), tuned for gfx950: the single-film kernel and the batched kernel
// that steps many independent films of different sizes in one launch.
//
// Same step and tile decomposition as the FP64 kernel of step.hip (64 x TY
// tile, h on a 3-halo, p on 2, F on 1), with the FP32 liberties the
// tolerance-checked path allows:
//   * FMA contraction and v_rcp_f32 reciprocals instead of IEEE division;
//   * Box-Muller with v_log/v_sin/v_cos/v_sqrt_f32, one Philox4x32-10 call
//     per pair of x-adjacent nodes (the pair's four words);
//   * halo rows loaded by whole wavefronts (one row offset per row), flat
//     thread mappings for the pressure and force phases;
//   * SRC = true: each F node collides all nine of its populations once and
//     parks them in LDS, so the stream is a gather; SRC = false keeps F in
//     LDS and collides at the destination.
// Height and distributions are deviations from a rest film of height href
// (Enc): dh = h - href, f' = f - feq(href, 0).
//
// One device function, tile_step, computes a tile; the single-film and the
// batched kernels both call it, so a film gives the same bits alone and in
// a batch.
#include <type_traits>

#include "swalbe_batch.h"

namespace swalbe {

namespace {

constexpr int TX = 64;
constexpr int NT = 256;
constexpr int NW = NT / 64;

__device__ __forceinline__ int wrap1(int a, int L) {
  return a < 0 ? a + L : (a >= L ? a - L : a);
}
__device__ __forceinline__ int wrapm(int a, int L) {
  a %= L;
  return a < 0 ? a + L : a;
}

__device__ __forceinline__ float powlaw_f(float a, int n, int m, int mode) {
  if (mode == 1) {
    const float t = (a * a) * a;
    return (t * t - 1.0f) * t;
  }
  if (mode == 2) return (a - 1.0f) * (a * a);
  float tn = 1.0f, tm = 1.0f;
  for (int i = 0; i < n; ++i) tn *= a;
  for (int i = 0; i < m; ++i) tm *= a;
  return tn - tm;
}

__device__ __forceinline__ double powlaw_d(double a, int n, int m, int mode) {
  if (mode == 1) {
    const double t = (a * a) * a;
    return (t * t - 1.0) * t;
  }
  if (mode == 2) return (a - 1.0) * (a * a);
  double tn = 1.0, tm = 1.0;
  for (int i = 0; i < n; ++i) tn *= a;
  for (int i = 0; i < m; ++i) tm *= a;
  return tn - tm;
}

// The disjoining part of the film pressure, -γ κ [a^n - a^m] with
// a = hmin / (h + hcrit), at a node of stored height deviation dh.  Where the
// film nears hmin (h < hguard) the term dominates the pressure and is
// evaluated in FP64 from the stored height with an IEEE division; elsewhere
// in FP32 with v_rcp_f32.  kap is κ (a θ field's value, or the film's).
__device__ __forceinline__ float disjoining_f(const F32K& K, float dh, float kap, bool kap_field) {
#pragma clang fp contract(fast)
  const float hh = K.href + dh;
  if (K.guard && hh < K.hguard) {
    const double hd = K.href_d + (double)dh;
    const double a = K.hmin_d / (hd + K.hcrit_d);
    const double kd = kap_field ? (double)kap : K.kap_d;
    return (float)(-K.gamma_d * (kd * powlaw_d(a, K.n, K.m, K.powmode)));
  }
  const float a = K.hmin * __builtin_amdgcn_rcpf(hh + K.hcrit);
  return -K.gamma * (kap * powlaw_f(a, K.n, K.m, K.powmode));
}

// p, a pointer uniform over the wave, as the compiler's scalar registers hold
// it: an address formed from it plus a 32-bit lane offset is a store's
// scalar base and offset, with no 64-bit vector add per store (left alone,
// the compiler adds the lane offset to the plane base first and the row
// offset to that per store, in 64-bit vector arithmetic).
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ gchar* scalar_ptr(float* p) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return (gchar*)(((uint64_t)hi << 32) | lo);
}

// Loads and stores of the tile code go through the global address space.
// The batch kernel takes its films' pointers from a table in memory, and
// through plain pointers the compiler would address them with flat
// instructions, which also count against the LDS counter: every wait for an
// LDS read would then wait for the outstanding stores too.
typedef __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ float gld(const float* p) { return *(const gfloat*)p; }

// A store of one output value, non-temporal (streaming past the caches) when
// NT.  The choice is a template argument: behind a run-time branch the
// compiler merges the two forms into one plain store.
template <bool NT>
__device__ __forceinline__ void store_f(float* p, float v) {
  gfloat* g = (gfloat*)p;
  if constexpr (NT) __builtin_nontemporal_store(v, g);
  else *g = v;
}

// f(std::integral_constant<int, m>) for the run-time store mode nts: 0 every
// output cached, 1 every output streaming, 2 the distributions and p
// streaming while h, ux, uy (the next step's input) stay cached.
template <class Fn>
__device__ __forceinline__ void with_store_mode(int nts, Fn&& f) {
  if (nts == 2) f(std::integral_constant<int, 2>{});
  else if (nts == 1) f(std::integral_constant<int, 1>{});
  else f(std::integral_constant<int, 0>{});
}

// LDS of one tile: h on the 3-halo, p on the 2-halo, (ux, uy) and the phase C
// results on the 1-halo.  AL (never with SRC): the (Fx, Fy) pairs overwrite
// p once every wave has read it (a barrier between), so the two share one
// buffer: 25.2 KB for a 64 x 16 tile (6 workgroups per CU) against 30.6 KB.
template <int TXW, int TY, bool SRC, bool ALIAS = !SRC>
struct Smem {
  static constexpr bool AL = ALIAS;
  static constexpr int HC = TXW + 6, HR = TY + 6, HN = HR * HC;
  static constexpr int PC = TXW + 4, PR = TY + 4, PN = PR * PC;
  static constexpr int UC = TXW + 2, UR = TY + 2, UN = UR * UC;
  static_assert(!(SRC && AL), "the collided planes never share p's buffer");
  static constexpr int AN = SRC ? 9 * UN : AL && PN > 2 * UN ? PN : 2 * UN;
  static_assert(HN % 2 == 0 && PN % 2 == 0, "float2 planes stay 8-byte aligned");
  float h[HN];
  float ps[AL ? 2 : PN];  // p, unless AL
  float2 uv[UN];          // (ux, uy) of the 1-halo nodes
  float a[AN];            // SRC: nine collided planes; else (Fx, Fy) pairs (AL: p first)
  __device__ __forceinline__ float* p() { return AL ? a : ps; }
};

struct TileIO {
  const float* h_in;
  const float* ux_in;
  const float* uy_in;
  const float* ft;
  float* f_out;
  float* h_out;
  float* ux_out;
  float* uy_out;
  float* p_out;
  const float* kappa;
};

struct TileRun {
  int Lx, Ly, small, pairs, nts, use_kappa, body;
  int64_t fstride;  // distance between distribution planes (elements)
  uint64_t seed, step;
  double wet_thresh;
};

struct SeriesAcc {
  double lmax, lmin;
  unsigned int cnt;
};

// Halo of one tile in registers: h on the 3-halo and (ux, uy) on the 1-halo,
// flat over the block (element e = tid + i * NTH).
template <int TXW, int NTH, int TY>
struct Halo {
  using SM = Smem<TXW, TY, false>;
  static constexpr int NHI = (SM::HN + NTH - 1) / NTH, NUI = (SM::UN + NTH - 1) / NTH;
  float h[NHI], u[NUI], v[NUI];
};

// ---- A1: issue every halo load of tile (x0, y0) (memory-level parallelism:
// nothing waits for them until halo_commit)
template <int TXW, int NTH, int TY>
__device__ __forceinline__ void halo_load(const TileIO& io, const TileRun& R, int x0, int y0,
                                          Halo<TXW, NTH, TY>& hr) {
  using SM = Smem<TXW, TY, false>;
  using HR = Halo<TXW, NTH, TY>;
  constexpr int HC = SM::HC, HN = SM::HN, UC = SM::UC, UN = SM::UN;
  const int tid = threadIdx.x;
  const int Lx = R.Lx, Ly = R.Ly;
  const int small = R.small;
  // Interior tiles (the 3-halo wraps nowhere): 32-bit byte offsets from the
  // halo's corner, so each load is a uniform base plus one VGPR offset
  if (!small && x0 >= 3 && y0 >= 3 && x0 + TXW + 3 <= Lx && y0 + TY + 3 <= Ly &&
      Lx <= (1 << 24)) {
    const char* hb = (const char*)(io.h_in + ((x0 - 3) + (int64_t)Lx * (y0 - 3)));
    const char* ub = (const char*)(io.ux_in + ((x0 - 1) + (int64_t)Lx * (y0 - 1)));
    const char* vb = (const char*)(io.uy_in + ((x0 - 1) + (int64_t)Lx * (y0 - 1)));
#pragma unroll
    for (int i = 0; i < HR::NHI; ++i) {
      const int e = tid + i * NTH;
      if (e < HN) {
        const int r = e / HC, c = e - r * HC;
        hr.h[i] = gld((const float*)(hb + (uint32_t)(4 * (c + Lx * r))));
      }
    }
#pragma unroll
    for (int i = 0; i < HR::NUI; ++i) {
      const int e = tid + i * NTH;
      if (e < UN) {
        const int r = e / UC, c = e - r * UC;
        const uint32_t o = (uint32_t)(4 * (c + Lx * r));
        hr.u[i] = gld((const float*)(ub + o));
        hr.v[i] = gld((const float*)(vb + o));
      }
    }
    return;
  }
  auto wx = [&](int a) { return small ? wrapm(a, Lx) : wrap1(a, Lx); };
  auto wy = [&](int a) { return small ? wrapm(a, Ly) : wrap1(a, Ly); };
#pragma unroll
  for (int i = 0; i < HR::NHI; ++i) {
    const int e = tid + i * NTH;
    if (e < HN) {
      const int r = e / HC, c = e - r * HC;
      hr.h[i] = gld(&io.h_in[wx(x0 - 3 + c) + (int64_t)Lx * wy(y0 - 3 + r)]);
    }
  }
#pragma unroll
  for (int i = 0; i < HR::NUI; ++i) {
    const int e = tid + i * NTH;
    if (e < UN) {
      const int r = e / UC, c = e - r * UC;
      const int64_t g = wx(x0 - 1 + c) + (int64_t)Lx * wy(y0 - 1 + r);
      hr.u[i] = gld(&io.ux_in[g]);
      hr.v[i] = gld(&io.uy_in[g]);
    }
  }
}

// ---- A2: park the halo in LDS (the block synchronises before phase B)
template <int TXW, int NTH, int TY, bool SRC, class SM>
__device__ __forceinline__ void halo_commit(const Halo<TXW, NTH, TY>& hr, SM& sm) {
  using HR = Halo<TXW, NTH, TY>;
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < HR::NHI; ++i) {
    const int e = tid + i * NTH;
    if (e < SM::HN) sm.h[e] = hr.h[i];
  }
#pragma unroll
  for (int i = 0; i < HR::NUI; ++i) {
    const int e = tid + i * NTH;
    if (e < SM::UN) sm.uv[e] = make_float2(hr.u[i], hr.v[i]);
  }
}

// Phase D of one tile: each wave takes whole 64-column rows of output nodes;
// per node it collides the nine populations at their source nodes (SRC:
// reads them, collided at the source in phase C), streams them (pull: f_k(x)
// = fo_k(x - c_k)), stores them (WF) and forms h, ux, uy (and p, WF, unless
// the layout shares p's buffer: then the caller stores p first).  Stores
// address each plane as a row base that is uniform across the wave plus the
// lane's column, so no per-store 64-bit address arithmetic is needed.  GRAV
// = false drops the gravity terms (g = 0: they are exactly zero).
template <int TXW, int NTH, int TY, bool READ_FT, bool SRC, bool SERIES, bool WF, bool GRAV,
          int NTS, class SM>
__device__ __forceinline__ void phase_d(const TileIO& io, const F32K& K, const TileRun& R, int x0,
                                        int y0, SM& sm, SeriesAcc& acc) {
#pragma clang fp contract(fast)
  constexpr int NW = NTH / 64;
  constexpr int HC = SM::HC, PC = SM::PC, UC = SM::UC, UN = SM::UN;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int Lx = R.Lx, Ly = R.Ly;
  const int64_t FS = R.fstride;
  const float w1 = 1.0f / 9.0f, w5 = 1.0f / 36.0f;
  const float2* F2 = reinterpret_cast<const float2*>(sm.a);
  for (int rc = wv; rc < TY * (TXW / 64); rc += NW) {
    const int row = rc / (TXW / 64), cb = 64 * (rc - row * (TXW / 64));
    const int col = lane + cb;
    const int gyo = y0 + row;
    if (gyo >= Ly) continue;
    if (x0 + col >= Lx) continue;
    // the output row's first node: uniform, so plane + rowo is a scalar base
    // and the lane's 32-bit byte offset the only per-lane part of an address
    const int64_t rowo = (int64_t)Lx * gyo + x0 + cb;
    // opaque per row, so the compiler keeps it a 32-bit offset beside the
    // stores (hoisted, it becomes a 64-bit vector operand of each address)
    uint32_t lo = 4u * (uint32_t)lane;
    asm volatile("" : "+v"(lo));
    auto at = [&](float* plane) { return (float*)(scalar_ptr(plane + rowo) + lo); };
    float fo[kQ];
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
      const int su = (row + 1 - cy(k)) * UC + (col + 1 - cx(k));
      float f;
      if constexpr (SRC) {
        f = sm.a[k * UN + su];
      } else {
        const float dh = sm.h[(row + 3 - cy(k)) * HC + (col + 3 - cx(k))];
        const float2 uvs = sm.uv[su];
        const float2 Fs = F2[su];
        const float u = uvs.x, v = uvs.y;
        const float hh = K.href + dh;
        const float usq = u * u + v * v;
        if (k == 0) {
          float fe = dh - (2.0f / 3.0f) * (hh * usq);
          if constexpr (GRAV) fe -= K.c56g * (dh * (hh + K.href));
          f = READ_FT ? K.itau * fe : fe;
        } else {
          // f_k = w (g0 gg + h ((4.5 a + 3) a - 1.5 |u|^2)) + cF c_k.F, a = c_k.u
          const float a = k == 1 ? u : k == 2 ? v : k == 3 ? -u : k == 4 ? -v
                        : k == 5 ? u + v : k == 6 ? v - u : k == 7 ? -(u + v) : u - v;
          const float w = k <= 4 ? w1 : w5;
          const float cF = k <= 4 ? (1.0f / 3.0f) : (1.0f / 24.0f);
          const float Fa = k == 1 ? Fs.x : k == 2 ? Fs.y : k == 3 ? -Fs.x : k == 4 ? -Fs.y
                         : k == 5 ? Fs.x + Fs.y : k == 6 ? Fs.y - Fs.x : k == 7 ? -(Fs.x + Fs.y)
                         : Fs.x - Fs.y;
          const float t = (4.5f * w) * a + 3.0f * w;
          float fe = hh * (t * a + (-1.5f * w) * usq);
          if constexpr (GRAV) fe += (w * K.g0) * (dh * (hh + K.href));
          f = READ_FT ? K.itau * fe + cF * Fa : fe + cF * Fa;
        }
        if constexpr (READ_FT) {
          const int sx = wrap1(x0 + col - cx(k), Lx), sy = wrap1(gyo - cy(k), Ly);
          f += K.omeg * gld(&io.ft[k * FS + sx + (int64_t)Lx * sy]);
        }
      }
      fo[k] = f;
      if constexpr (WF) store_f<NTS != 0>(at(io.f_out + k * FS), f);
    }
    const float s = (((((((fo[0] + fo[1]) + fo[2]) + fo[3]) + fo[4]) + fo[5]) + fo[6]) + fo[7]) + fo[8];
    const float ih = __builtin_amdgcn_rcpf(K.href + s);
    // mode 2: only f and p stream past the caches (h, u are the next step's input)
    store_f<NTS == 1>(at(io.h_out), s);
    store_f<NTS == 1>(at(io.ux_out), ((((fo[1] - fo[3]) + fo[5]) - fo[6]) - fo[7] + fo[8]) * ih);
    store_f<NTS == 1>(at(io.uy_out), ((((fo[2] - fo[4]) + fo[5]) + fo[6]) - fo[7] - fo[8]) * ih);
    if constexpr (WF && !SM::AL)
      store_f<NTS != 0>(at(io.p_out), sm.p()[(row + 2) * PC + (col + 2)]);
    if constexpr (SERIES) {
      const double hv = (double)K.href + (double)sm.h[(row + 3) * HC + (col + 3)];
      acc.lmax = hv > acc.lmax ? hv : acc.lmax;
      acc.lmin = hv < acc.lmin ? hv : acc.lmin;
      acc.cnt += hv > R.wet_thresh ? 1u : 0u;
    }
  }
}

// Phases B-D of one TXW x TY tile whose halo is in LDS: pressure, forces,
// collision, pull-stream and moments, computed by NTH threads.
template <int TXW, int NTH, int TY, bool READ_FT, bool THERMAL, bool SRC, bool SERIES, class SM>
__device__ __forceinline__ void tile_compute(const TileIO& io, const F32K& K, const TileRun& R,
                                             int x0, int y0, SM& sm, SeriesAcc& acc) {
#pragma clang fp contract(fast)
  static_assert(TXW % 64 == 0 && NTH % 64 == 0, "tiles are whole wavefronts wide");
  constexpr int NT = NTH, NW = NTH / 64;
  constexpr bool AL = SM::AL;
  constexpr int HC = SM::HC, PC = SM::PC;
  constexpr int UC = SM::UC, UR = SM::UR, UN = SM::UN;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // wave index made provably uniform: row offsets then live in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int Lx = R.Lx, Ly = R.Ly;
  const int64_t FS = R.fstride;
  const int small = R.small;
  auto wx = [&](int a) { return small ? wrapm(a, Lx) : wrap1(a, Lx); };
  auto wy = [&](int a) { return small ? wrapm(a, Ly) : wrap1(a, Ly); };

  // ---- B: film pressure on the 2-halo.  Each wave takes whole rows, one
  // column per lane, so each LDS access of a wave is one contiguous run of a
  // row (no bank conflicts where a flat index would wrap into the next row);
  // the PC - TXW columns beyond them follow, a share per wave.
  float* const P = sm.p();
  auto pnode = [&](int r, int c) {
    const float* hc = &sm.h[(r + 1) * HC + (c + 1)];
    const float hip = hc[-1], him = hc[1], hjp = hc[-HC], hjm = hc[HC];
    const float hipjp = hc[-HC - 1], himjp = hc[-HC + 1], himjm = hc[HC + 1], hipjm = hc[HC - 1];
    const float L = (2.0f / 3.0f) * (((hjp + hip) + him) + hjm) +
                    (1.0f / 6.0f) * (((hipjp + himjp) + himjm) + hipjm) - (10.0f / 3.0f) * hc[0];
    float kap = K.kap;
    if (R.use_kappa) kap = gld(&io.kappa[wx(x0 - 2 + c) + (int64_t)Lx * wy(y0 - 2 + r)]);
    P[r * PC + c] = disjoining_f(K, hc[0], kap, R.use_kappa) - K.gamma * L;
  };
  {
    constexpr int WPR = TXW / 64, PR = SM::PR, XB = PC - TXW, LB = XB * PR;
    constexpr int CB = (LB + NW - 1) / NW;
    static_assert(CB <= 64, "one pass of a wave covers its share of the edge columns");
    for (int rc = wv; rc < PR * WPR; rc += NW) {
      const int r = rc / WPR;
      pnode(r, lane + 64 * (rc - r * WPR));
    }
    const int k = wv * CB + lane;
    if (lane < CB && k < LB) pnode(k / XB, TXW + k % XB);
  }
  __syncthreads();

  // ---- C: forces (and with SRC the collision) on the 1-halo.  With AL the
  // (Fx, Fy) pairs stay in registers until every wave has read p, then
  // overwrite it; p at the wave's phase D nodes waits in registers too, and
  // is stored after the block's last barrier, beside phase D's stores (p
  // stored from phase B made the step 4.5 % slower).
  constexpr int NRW = (TY * (TXW / 64) + NW - 1) / NW;  // phase D rows per wave
  float pown[NRW];
  auto keep_p = [&]() {
#pragma unroll
    for (int i = 0; i < NRW; ++i) {
      const int rc = wv + i * NW, row = rc / (TXW / 64), col = lane + 64 * (rc % (TXW / 64));
      if (rc < TY * (TXW / 64)) pown[i] = P[(row + 2) * PC + (col + 2)];
    }
  };
  const float w1 = 1.0f / 9.0f, w5 = 1.0f / 36.0f;
  float2* const F2 = reinterpret_cast<float2*>(sm.a);
  auto fnode = [&](int r, int c, float z0, float z1) {
    const float* pc = &P[(r + 1) * PC + (c + 1)];
    const float fip = pc[-1], fim = pc[1], fjp = pc[-PC], fjm = pc[PC];
    const float fipjp = pc[-PC - 1], fimjp = pc[-PC + 1], fimjm = pc[PC + 1], fipjm = pc[PC - 1];
    const float dh = sm.h[(r + 2) * HC + (c + 2)];
    const float hh = K.href + dh;
    const float gx = (-1.0f / 3.0f) * (fip - fim) - (1.0f / 12.0f) * (((fipjp - fimjp) - fimjm) + fipjm);
    const float gy = (-1.0f / 3.0f) * (fjp - fjm) - (1.0f / 12.0f) * (((fipjp + fimjp) - fimjm) - fipjm);
    const int ui = r * UC + c;
    const float2 uvn = sm.uv[ui];
    const float u = uvn.x, v = uvn.y;
    // 2h^2 + 6δh + 3δ^2 is the denominator of both slip and noise amplitude
    const float iden = __builtin_amdgcn_rcpf((2.0f * hh + K.c6d) * hh + K.c3dd);
    const float sl = K.c6mu * hh * iden;
    float Fx = -hh * gx - sl * u;
    float Fy = -hh * gy - sl * v;
    if constexpr (THERMAL) {
      const float amp = __builtin_amdgcn_sqrtf(K.kb26 * hh * iden);
      Fx -= z0 * amp;
      Fy -= z1 * amp;
    }
    if (R.body) {
      Fx += hh * K.bfx;
      Fy += hh * K.bfy;
    }
    if constexpr (SRC) {
      const float usq = u * u + v * v;
      const float gg = dh * (hh + K.href);
      const float w1h = w1 * hh, w5h = w5 * hh;
      const float b1 = w1 * K.g0 * gg - 1.5f * w1h * usq;
      const float b5 = w5 * K.g0 * gg - 1.5f * w5h * usq;
      const float a5 = u + v, a6 = v - u;
      const float c3 = 1.0f / 3.0f, c24 = 1.0f / 24.0f;
      float fe[kQ], S[kQ];
      fe[0] = (dh - K.c56g * gg) - (2.0f / 3.0f) * hh * usq;
      fe[1] = b1 + w1h * ((4.5f * u + 3.0f) * u);
      fe[2] = b1 + w1h * ((4.5f * v + 3.0f) * v);
      fe[3] = b1 + w1h * ((4.5f * u - 3.0f) * u);
      fe[4] = b1 + w1h * ((4.5f * v - 3.0f) * v);
      fe[5] = b5 + w5h * ((4.5f * a5 + 3.0f) * a5);
      fe[6] = b5 + w5h * ((4.5f * a6 + 3.0f) * a6);
      fe[7] = b5 + w5h * ((4.5f * a5 - 3.0f) * a5);
      fe[8] = b5 + w5h * ((4.5f * a6 - 3.0f) * a6);
      S[0] = 0.0f;
      S[1] = c3 * Fx;
      S[2] = c3 * Fy;
      S[3] = -c3 * Fx;
      S[4] = -c3 * Fy;
      S[5] = c24 * (Fx + Fy);
      S[6] = c24 * (Fy - Fx);
      S[7] = -c24 * (Fx + Fy);
      S[8] = c24 * (Fx - Fy);
      int64_t gsrc = 0;
      if constexpr (READ_FT) gsrc = wx(x0 - 1 + c) + (int64_t)Lx * wy(y0 - 1 + r);
#pragma unroll
      for (int k = 0; k < kQ; ++k) {
        float f = READ_FT ? K.itau * fe[k] + S[k] : fe[k] + S[k];
        if constexpr (READ_FT) f += K.omeg * gld(&io.ft[k * FS + gsrc]);
        sm.a[k * UN + ui] = f;
      }
    } else if constexpr (!AL) {
      F2[ui] = make_float2(Fx, Fy);
    }
    return make_float2(Fx, Fy);
  };
  if (R.pairs) {
    // one Philox call per vertical node pair (x, 2m), (x, 2m+1): pair row q
    // covers halo rows 2q-1 (even y) and 2q (odd y); y0 is even
    // rows of pairs go to the waves whole, one column per lane (contiguous
    // LDS runs, as in phase B); the UC - TXW edge columns go to the last wave
    constexpr int NPR = TY / 2 + 2, WPR = TXW / 64, XC = UC - TXW, LC = XC * NPR;
    constexpr int NIT = (NPR * WPR + NW - 1) / NW;
    static_assert(LC <= 64, "one pass of the last wave covers the edge columns");
    auto pair = [&](int q, int c, float2* Fp) {
      const int re = 2 * q - 1, ro = 2 * q;
      float z[4] = {0.f, 0.f, 0.f, 0.f};
      if constexpr (THERMAL) {
        const int gxx = wrap1(x0 - 1 + c, Lx);
        const int ye = wrap1(y0 - 2 + 2 * q, Ly);
        uint32_t w[4];
        philox_ctr(R.seed, R.step, (uint32_t)gxx + (uint32_t)Lx * (uint32_t)(ye >> 1), kTagThermal,
                   w);
        box_muller_f32(w[0], w[1], &z[0], &z[1]);
        box_muller_f32(w[2], w[3], &z[2], &z[3]);
      }
      if (re >= 0) Fp[0] = fnode(re, c, z[0], z[1]);
      if (ro < UR) Fp[1] = fnode(ro, c, z[2], z[3]);
    };
    auto put = [&](int q, int c, const float2* Fp) {
      const int re = 2 * q - 1, ro = 2 * q;
      if (re >= 0) F2[re * UC + c] = Fp[0];
      if (ro < UR) F2[ro * UC + c] = Fp[1];
    };
    const bool edge = wv == NW - 1 && lane < LC;
    if constexpr (AL) {
      float2 Fq[NIT + 1][2];  // the last entry: the edge columns
#pragma unroll
      for (int i = 0; i < NIT; ++i) {
        const int t = wv + i * NW, q = t / WPR;
        if (t < NPR * WPR) pair(q, lane + 64 * (t - q * WPR), Fq[i]);
      }
      if (edge) pair(lane / XC, TXW + lane % XC, Fq[NIT]);
      keep_p();
      __syncthreads();  // every wave has read p
#pragma unroll
      for (int i = 0; i < NIT; ++i) {
        const int t = wv + i * NW, q = t / WPR;
        if (t < NPR * WPR) put(q, lane + 64 * (t - q * WPR), Fq[i]);
      }
      if (edge) put(lane / XC, TXW + lane % XC, Fq[NIT]);
    } else {
      float2 Fp[2];  // fnode has stored the pairs
      for (int t = wv; t < NPR * WPR; t += NW) {
        const int q = t / WPR;
        pair(q, lane + 64 * (t - q * WPR), Fp);
      }
      if (edge) pair(lane / XC, TXW + lane % XC, Fp);
    }
  } else {
    auto node = [&](int t) {
      const int r = t / UC, c = t - r * UC;
      float z0 = 0.f, z1 = 0.f;
      if constexpr (THERMAL) {
        const uint32_t s = noise_index(wx(x0 - 1 + c), wy(y0 - 1 + r), Lx);
        uint32_t w[4];
        philox_ctr(R.seed, R.step, s >> 1, kTagThermal, w);
        const int b = (int)(s & 1u);
        box_muller_f32(w[2 * b], w[2 * b + 1], &z0, &z1);
      }
      return fnode(r, c, z0, z1);
    };
    if constexpr (AL) {
      constexpr int NIU = (UN + NT - 1) / NT;
      float2 Fn[NIU];
#pragma unroll
      for (int i = 0; i < NIU; ++i)
        if (tid + i * NT < UN) Fn[i] = node(tid + i * NT);
      keep_p();
      __syncthreads();  // every wave has read p
#pragma unroll
      for (int i = 0; i < NIU; ++i)
        if (tid + i * NT < UN) F2[tid + i * NT] = Fn[i];
    } else {
      for (int t = tid; t < UN; t += NT) node(t);  // fnode stores the pairs
    }
  }
  __syncthreads();

  __builtin_amdgcn_s_setprio(1);  // the stores ahead of other waves' pressure and forces
  // ---- D: collide, stream (gather) and moments; one copy per store mode,
  // store set and gravity, chosen once per tile (uniform), so the loop body
  // has no branches
  with_store_mode(R.nts, [&](auto mode) {
    constexpr int NTS = decltype(mode)::value;
    if constexpr (AL) {
      if (io.p_out) {
#pragma unroll
        for (int i = 0; i < NRW; ++i) {
          const int rc = wv + i * NW, row = rc / (TXW / 64), cb = 64 * (rc % (TXW / 64));
          if (rc < TY * (TXW / 64) && y0 + row < Ly && x0 + cb + lane < Lx) {
            uint32_t lo = 4u * (uint32_t)lane;  // a 32-bit store offset, as in phase D
            asm volatile("" : "+v"(lo));
            store_f<NTS != 0>(
                (float*)(scalar_ptr(io.p_out + ((int64_t)Lx * (y0 + row) + x0 + cb)) + lo),
                pown[i]);
          }
        }
      }
    }
    if (io.f_out) {
      if (K.g0 != 0.0f)
        phase_d<TXW, NTH, TY, READ_FT, SRC, SERIES, true, true, NTS>(io, K, R, x0, y0, sm, acc);
      else
        phase_d<TXW, NTH, TY, READ_FT, SRC, SERIES, true, false, NTS>(io, K, R, x0, y0, sm, acc);
    } else {
      if (K.g0 != 0.0f)
        phase_d<TXW, NTH, TY, READ_FT, SRC, SERIES, false, true, NTS>(io, K, R, x0, y0, sm, acc);
      else
        phase_d<TXW, NTH, TY, READ_FT, SRC, SERIES, false, false, NTS>(io, K, R, x0, y0, sm, acc);
    }
  });
  __builtin_amdgcn_s_setprio(0);
}


// One TXW x TY tile of one film: stage the halo, then phases B-D.  The caller
// synchronises the block between two calls (the tile rewrites the LDS halo
// first).
template <int TXW, int NTH, int TY, bool READ_FT, bool THERMAL, bool SRC, bool SERIES, class SM>
__device__ __forceinline__ void tile_step(const TileIO& io, const F32K& K, const TileRun& R,
                                          int x0, int y0, SM& sm, SeriesAcc& acc) {
  Halo<TXW, NTH, TY> hr;
  __builtin_amdgcn_s_setprio(2);  // issue the tile's loads ahead of other waves' arithmetic
  halo_load<TXW, NTH, TY>(io, R, x0, y0, hr);
  __builtin_amdgcn_s_setprio(0);
  halo_commit<TXW, NTH, TY, SRC>(hr, sm);
  __syncthreads();
  tile_compute<TXW, NTH, TY, READ_FT, THERMAL, SRC, SERIES>(io, K, R, x0, y0, sm, acc);
}


// ---------------------------------------------------------------- strips
// LDS-free variant: one wavefront owns a strip of 64 columns (x0-3 .. x0+60,
// one per lane) and marches down a segment of rows.  Per row it loads h
// three rows ahead, computes p two rows ahead and the post-collision
// populations one row ahead, so the 3-row halo of h, p and f lives in
// registers, x-neighbours come from DPP wave shifts and no barrier is needed.
// The halo narrows by one column per stencil: 58 columns (lanes 3..60) are
// written per strip.
// Value held by the lane of column x-1 (wave_shr:1) / x+1 (wave_shl:1); the
// edge lanes keep their own value (they are halo lanes, never written).
__device__ __forceinline__ float at_xm1(float v) {
  const int i = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(i, i, 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float at_xp1(float v) {
  const int i = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(i, i, 0x130, 0xF, 0xF, false));
}

struct Row3 {
  float v, m, p;  // value at x, x-1, x+1
};

__device__ __forceinline__ Row3 row3(float v) { return Row3{v, at_xm1(v), at_xp1(v)}; }

// Rows [y0, y0 + S) of the strip whose first output column is x0.
template <bool READ_FT, bool THERMAL, bool SERIES, int NTS>
__device__ __forceinline__ void strip_run(const TileIO& io, const F32K& K, const TileRun& R,
                                          int x0, int y0, int S, SeriesAcc& acc) {
#pragma clang fp contract(fast)
  const int lane = threadIdx.x & 63;
  const int Lx = R.Lx, Ly = R.Ly;
  const int64_t FS = R.fstride;
  const int small = R.small;
  auto wy = [&](int a) { return small ? wrapm(a, Ly) : wrap1(a, Ly); };
  const int xc = x0 - 3 + lane;  // this lane's column (unwrapped)
  const int gx = small ? wrapm(xc, Lx) : wrap1(xc, Lx);
  const bool out_lane = lane >= 3 && lane < 3 + kStripOut && xc < Lx;
  const float w1 = 1.0f / 9.0f, w5 = 1.0f / 36.0f;
  const int jlast = y0 + S - 1;

  // prefetched inputs of the next iteration: h(j+3), u(j+1), κ(j+2), f_t(j+1)
  float nh, nu, nv, nk = 0.f, nft[kQ];
  auto prefetch = [&](int j) {
    const int64_t r3 = (int64_t)Lx * wy(j + 3) + gx, r1 = (int64_t)Lx * wy(j + 1) + gx;
    nh = io.h_in[r3];
    nu = io.ux_in[r1];
    nv = io.uy_in[r1];
    if (R.use_kappa) nk = io.kappa[(int64_t)Lx * wy(j + 2) + gx];
    if constexpr (READ_FT) {
#pragma unroll
      for (int k = 0; k < kQ; ++k) nft[k] = io.ft[k * FS + r1];
    }
  };

  uint32_t wc2 = 0, wc3 = 0;  // cached noise words of the odd row of a pair

  // one row: j is the output row; the arrays rotate roles between calls
  auto body = [&](int j, Row3& hN, Row3& h1, Row3& h2, Row3& pN, Row3& p0, Row3& p1, float* fN,
                  float* fm, float* f0) {
    if constexpr (SERIES) {
      if (j >= y0 && out_lane) {  // h(j) is the step's input height at the output row
        const double hv = (double)K.href + (double)hN.v;
        acc.lmax = hv > acc.lmax ? hv : acc.lmax;
        acc.lmin = hv < acc.lmin ? hv : acc.lmin;
        acc.cnt += hv > R.wet_thresh ? 1u : 0u;
      }
    }
    const float u = nu, v = nv, kap = R.use_kappa ? nk : K.kap;
    float ft[kQ];
    if constexpr (READ_FT) {
#pragma unroll
      for (int k = 0; k < kQ; ++k) ft[k] = nft[k];
    }
    hN = row3(nh);  // h(j+3)
    if (j < jlast) prefetch(j + 1);

    // p(j+2) from h rows j+1 .. j+3
    {
      const float L = (2.0f / 3.0f) * (((h1.v + h2.m) + h2.p) + hN.v) +
                      (1.0f / 6.0f) * (((h1.m + h1.p) + hN.p) + hN.m) - (10.0f / 3.0f) * h2.v;
      pN = row3(disjoining_f(K, h2.v, kap, R.use_kappa) - K.gamma * L);
    }
    if (j >= y0 - 2) {
      // F(j+1) from p rows j .. j+2, then the nine post-collision values
      const float dh = h1.v, hh = K.href + dh;
      const float gxp = (-1.0f / 3.0f) * (p1.m - p1.p) -
                        (1.0f / 12.0f) * (((p0.m - p0.p) - pN.p) + pN.m);
      const float gyp = (-1.0f / 3.0f) * (p0.v - pN.v) -
                        (1.0f / 12.0f) * (((p0.m + p0.p) - pN.p) - pN.m);
      const float iden = __builtin_amdgcn_rcpf((2.0f * hh + K.c6d) * hh + K.c3dd);
      const float sl = K.c6mu * hh * iden;
      float Fx = -hh * gxp - sl * u;
      float Fy = -hh * gyp - sl * v;
      if constexpr (THERMAL) {
        // one Philox call per vertical node pair (x, 2m), (x, 2m+1)
        const int rw = wy(j + 1);
        uint32_t wa, wb;
        if ((rw & 1) == 0 || j == y0 - 2) {
          uint32_t w[4];
          philox_ctr(R.seed, R.step, (uint32_t)gx + (uint32_t)Lx * (uint32_t)(rw >> 1),
                     kTagThermal, w);
          wa = (rw & 1) ? w[2] : w[0];
          wb = (rw & 1) ? w[3] : w[1];
          wc2 = w[2];
          wc3 = w[3];
        } else {
          wa = wc2;
          wb = wc3;
        }
        float z0, z1;
        box_muller_f32(wa, wb, &z0, &z1);
        const float amp = __builtin_amdgcn_sqrtf(K.kb26 * hh * iden);
        Fx -= z0 * amp;
        Fy -= z1 * amp;
      }
      if (R.body) {
        Fx += hh * K.bfx;
        Fy += hh * K.bfy;
      }
      const float usq = u * u + v * v;
      const float gg = dh * (hh + K.href);
      const float w1h = w1 * hh, w5h = w5 * hh;
      const float b1 = w1 * K.g0 * gg - 1.5f * w1h * usq;
      const float b5 = w5 * K.g0 * gg - 1.5f * w5h * usq;
      const float a5 = u + v, a6 = v - u;
      const float c3 = 1.0f / 3.0f, c24 = 1.0f / 24.0f;
      float fe[kQ], Sk[kQ];
      fe[0] = (dh - K.c56g * gg) - (2.0f / 3.0f) * hh * usq;
      fe[1] = b1 + w1h * ((4.5f * u + 3.0f) * u);
      fe[2] = b1 + w1h * ((4.5f * v + 3.0f) * v);
      fe[3] = b1 + w1h * ((4.5f * u - 3.0f) * u);
      fe[4] = b1 + w1h * ((4.5f * v - 3.0f) * v);
      fe[5] = b5 + w5h * ((4.5f * a5 + 3.0f) * a5);
      fe[6] = b5 + w5h * ((4.5f * a6 + 3.0f) * a6);
      fe[7] = b5 + w5h * ((4.5f * a5 - 3.0f) * a5);
      fe[8] = b5 + w5h * ((4.5f * a6 - 3.0f) * a6);
      Sk[0] = 0.0f;
      Sk[1] = c3 * Fx;
      Sk[2] = c3 * Fy;
      Sk[3] = -c3 * Fx;
      Sk[4] = -c3 * Fy;
      Sk[5] = c24 * (Fx + Fy);
      Sk[6] = c24 * (Fy - Fx);
      Sk[7] = -c24 * (Fx + Fy);
      Sk[8] = c24 * (Fx - Fy);
#pragma unroll
      for (int k = 0; k < kQ; ++k) {
        float f = READ_FT ? K.itau * fe[k] + Sk[k] : fe[k] + Sk[k];
        if constexpr (READ_FT) f += K.omeg * ft[k];
        fN[k] = f;
      }
    }
    if (j >= y0) {
      // pull stream: f_k(x, j) = fo_k(x - c_k) from rows j-1 (fm), j (f0), j+1 (fN)
      float fo[kQ];
      fo[0] = f0[0];
      fo[1] = at_xm1(f0[1]);
      fo[2] = fm[2];
      fo[3] = at_xp1(f0[3]);
      fo[4] = fN[4];
      fo[5] = at_xm1(fm[5]);
      fo[6] = at_xp1(fm[6]);
      fo[7] = at_xp1(fN[7]);
      fo[8] = at_xm1(fN[8]);
      if (out_lane) {
        const int64_t xo = (int64_t)Lx * j + xc;
        if (io.f_out) {
#pragma unroll
          for (int k = 0; k < kQ; ++k) store_f<NTS != 0>(&io.f_out[k * FS + xo], fo[k]);
        }
        const float s = (((((((fo[0] + fo[1]) + fo[2]) + fo[3]) + fo[4]) + fo[5]) + fo[6]) + fo[7]) +
                        fo[8];
        const float ih = __builtin_amdgcn_rcpf(K.href + s);
        // mode 2: h, u stay cacheable
        store_f<NTS == 1>(&io.h_out[xo], s);
        store_f<NTS == 1>(&io.ux_out[xo], ((((fo[1] - fo[3]) + fo[5]) - fo[6]) - fo[7] + fo[8]) * ih);
        store_f<NTS == 1>(&io.uy_out[xo], ((((fo[2] - fo[4]) + fo[5]) + fo[6]) - fo[7] - fo[8]) * ih);
        if (io.p_out) store_f<NTS != 0>(&io.p_out[xo], p0.v);
      }
    }
  };

  // prologue: h rows y0-3, y0-2 and the prefetch of iteration y0-4
  Row3 hA{0.f, 0.f, 0.f};  // role h(j) at j = y0-4: overwritten before use
  Row3 hB = row3(io.h_in[(int64_t)Lx * wy(y0 - 3) + gx]);
  Row3 hC = row3(io.h_in[(int64_t)Lx * wy(y0 - 2) + gx]);
  prefetch(y0 - 4);
  Row3 pA{0.f, 0.f, 0.f}, pB{0.f, 0.f, 0.f}, pC{0.f, 0.f, 0.f};
  float fA[kQ], fB[kQ], fC[kQ];
#pragma unroll
  for (int k = 0; k < kQ; ++k) fA[k] = fB[k] = fC[k] = 0.f;
  // roles per call: (hN, h1, h2) = (h(j), h(j+1), h(j+2)) on entry;
  // (pN, p0, p1) = (p(j-1), p(j), p(j+1)); (fN, fm, f0) = (free, fo(j-1), fo(j))
  int j = y0 - 4;
  for (; j + 2 <= jlast; j += 3) {
    body(j, hA, hB, hC, pA, pB, pC, fA, fB, fC);
    body(j + 1, hB, hC, hA, pB, pC, pA, fB, fC, fA);
    body(j + 2, hC, hA, hB, pC, pA, pB, fC, fA, fB);
  }
  if (j <= jlast) body(j, hA, hB, hC, pA, pB, pC, fA, fB, fC);
  if (j + 1 <= jlast) body(j + 1, hB, hC, hA, pB, pC, pA, fB, fC, fA);
}

// Single film: one wavefront per (strip, segment) work item, items in
// segment-major order, each XCD taking a contiguous share.
template <bool READ_FT, bool THERMAL>
__global__ __launch_bounds__(64) void k_strip_f32(TileIO io, F32K K, TileRun R, int nstrips,
                                                  int seg, int items, int chunk, StepSeries ser) {
  const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
  const int item = xcd * chunk + k;
  if (k >= chunk || item >= items) return;
  const int sg = item / nstrips, st = item - sg * nstrips;
  const int y0 = sg * seg;
  const int S = min(seg, R.Ly - y0);
  SeriesAcc acc{-__builtin_inf(), __builtin_inf(), 0u};
  const bool series = ser.dh_keys || ser.wet;
  with_store_mode(R.nts, [&](auto mode) {
    constexpr int NTS = decltype(mode)::value;
    if (series) strip_run<READ_FT, THERMAL, true, NTS>(io, K, R, st * kStripOut, y0, S, acc);
    else strip_run<READ_FT, THERMAL, false, NTS>(io, K, R, st * kStripOut, y0, S, acc);
  });
  if (!series) return;
  unsigned long long kmax = dkey(acc.lmax), kmin = dkey(acc.lmin);
  unsigned int lcnt = acc.cnt;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(kmax, o), b = __shfl_xor(kmin, o);
    kmax = a > kmax ? a : kmax;
    kmin = b < kmin ? b : kmin;
    lcnt += __shfl_xor(lcnt, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (ser.dh_keys) {
      atomicMax(&ser.dh_keys[0], kmax);
      atomicMin(&ser.dh_keys[1], kmin);
    }
    if (ser.wet) atomicAdd(&ser.wet[0], (unsigned long long)lcnt);
  }
}

// ---------------------------------------------------------------- single film
// per-step series: max/min/count of h_t over the block's tiles, folded into
// the domain totals (exact: integer keys and counts)
template <int NW>
__device__ __forceinline__ void series_flush(const SeriesAcc& acc, const StepSeries& ser,
                                             unsigned long long* red_k, unsigned int* red_c) {
  unsigned long long kmax = dkey(acc.lmax), kmin = dkey(acc.lmin);
  unsigned int lcnt = acc.cnt;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(kmax, o), b = __shfl_xor(kmin, o);
    kmax = a > kmax ? a : kmax;
    kmin = b < kmin ? b : kmin;
    lcnt += __shfl_xor(lcnt, o);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red_k[2 * wv] = kmax;
    red_k[2 * wv + 1] = kmin;
    red_c[wv] = lcnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NW; ++w) {
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

template <int TXW, int NTH, int TY, bool READ_FT, bool THERMAL, bool SRC, class SM>
__device__ __forceinline__ void single_film_step(const TileIO& io, const F32K& K, const TileRun& R,
                                                 int ntx, int ntiles, const StepSeries& ser, SM& sm,
                                                 unsigned long long* red_k, unsigned int* red_c) {
  constexpr int NW = NTH / 64;
  const int tile = tile_of_block(ntx, ntiles, R.step);
  const int x0 = (tile % ntx) * TXW, y0 = (tile / ntx) * TY;
  SeriesAcc acc{-__builtin_inf(), __builtin_inf(), 0u};
  const bool series = ser.dh_keys || ser.wet;
  if (series) tile_step<TXW, NTH, TY, READ_FT, THERMAL, SRC, true>(io, K, R, x0, y0, sm, acc);
  else tile_step<TXW, NTH, TY, READ_FT, THERMAL, SRC, false>(io, K, R, x0, y0, sm, acc);
  if (series) series_flush<NW>(acc, ser, red_k, red_c);
}

// AL: p and the (Fx, Fy) pairs share one LDS buffer (never with SRC).
template <int TXW, int NTH, int TY, bool READ_FT, bool THERMAL, bool SRC, bool AL>
__global__ __launch_bounds__(NTH) void k_step_f32(TileIO io, F32K K, TileRun R, int ntx,
                                                  int ntiles, StepSeries ser) {
  __shared__ Smem<TXW, TY, SRC, AL> sm;
  __shared__ unsigned long long red_k[2 * (NTH / 64)];
  __shared__ unsigned int red_c[NTH / 64];
  single_film_step<TXW, NTH, TY, READ_FT, THERMAL, SRC>(io, K, R, ntx, ntiles, ser, sm, red_k,
                                                        red_c);
}

// The same kernel under its own name, launched for the steps whose
// distributions and pressure are not written (f_out = p_out = null): the
// identical body with the same runtime checks, so both give the same bits,
// and a kernel trace tells the two kinds of step apart.
template <int TXW, int NTH, int TY, bool READ_FT, bool THERMAL, bool SRC, bool AL>
__global__ __launch_bounds__(NTH) void k_step_f32_unwritten(TileIO io, F32K K, TileRun R,
                                                            int ntx, int ntiles, StepSeries ser) {
  __shared__ Smem<TXW, TY, SRC, AL> sm;
  __shared__ unsigned long long red_k[2 * (NTH / 64)];
  __shared__ unsigned int red_c[NTH / 64];
  single_film_step<TXW, NTH, TY, READ_FT, THERMAL, SRC>(io, K, R, ntx, ntiles, ser, sm, red_k,
                                                        red_c);
}

template <int TXW, int NTH, int TY, bool SRC>
void launch_single(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                   int Lx, int Ly, int64_t fstride, Enc enc, int nts, hipStream_t s) {
  const int ntx = (Lx + TXW - 1) / TXW, nty = (Ly + TY - 1) / TY;
  TileIO t{io.h_in, io.ux_in, io.uy_in, io.ft, io.f_out, io.h_out, io.ux_out, io.uy_out,
           io.p_out, io.kappa};
  TileRun R;
  R.Lx = Lx;
  R.Ly = Ly;
  R.small = film_is_small(Lx, Ly, TXW, TY);
  R.pairs = (!R.small && (Ly % 2) == 0) ? 1 : 0;
  R.nts = nts;
  R.use_kappa = io.kappa != nullptr;
  R.body = sp.body;
  R.fstride = fstride;
  R.seed = sp.seed;
  R.step = step;
  R.wet_thresh = sp.wet_thresh;
  const F32K K = make_f32k(sp, enc);
  const dim3 grid(ntx * nty), block(NTH);
  // The shared p / (Fx, Fy) buffer raises a 64 x 16 tile's residency from 5
  // to 6 workgroups per CU for one more barrier per tile: it pays only where
  // the film has more tiles than 5 per CU can hold at once.
  const bool al = !SRC && TY >= 16 && (int64_t)ntx * nty > 5 * (int64_t)device_cus();
#define SW_L(KN, RF, TH)                                                                        \
  do {                                                                                          \
    if (al)                                                                                     \
      hipLaunchKernelGGL((KN<TXW, NTH, TY, RF, TH, SRC, !SRC>), grid, block, 0, s, t, K, R, ntx, \
                         ntx * nty, ser);                                                       \
    else                                                                                        \
      hipLaunchKernelGGL((KN<TXW, NTH, TY, RF, TH, SRC, false>), grid, block, 0, s, t, K, R,     \
                         ntx, ntx * nty, ser);                                                  \
  } while (0)
  if (sp.read_ft) {
    if (sp.thermal) SW_L(k_step_f32, true, true);
    else SW_L(k_step_f32, true, false);
  } else if (!io.f_out && !io.p_out) {  // a step whose f and p nobody reads
    if (sp.thermal) SW_L(k_step_f32_unwritten, false, true);
    else SW_L(k_step_f32_unwritten, false, false);
  } else {
    if (sp.thermal) SW_L(k_step_f32, false, true);
    else SW_L(k_step_f32, false, false);
  }
#undef SW_L
}

// ---------------------------------------------------------------- batch of films
// A fixed grid walks the tiles of all resident films.  Block b sits on XCD
// b % 8 and takes tiles of that XCD's contiguous share, so neighbouring
// tiles (shared halos) stay in one L2.  Films are found by binary search of
// the tile prefix table; each film's parity and step number follow from the
// device launch counter, so the launch sequence can be a captured graph.
// ANY_FT = false is the kernel for batches whose films all have τ = 1 (the
// host launches it only then): without the f_t path it fits 5 workgroups per
// CU, the LDS limit; with it, at least 4 waves per SIMD (5 would spill).
template <bool ANY_FT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(ANY_FT ? 4 : 5))) void
k_batch_f32(const BatchFilm* __restrict__ films, const int* __restrict__ tstart,
            const BatchHdr* __restrict__ hdr, const unsigned long long* __restrict__ counter,
            int nts) {
  // LDS for the taller tile; a film of 8-row tiles uses its front part
  // p and the (Fx, Fy) pairs apart: sharing them costs this kernel spills
  using SM16 = Smem<TX, kBatchTileY, false, false>;
  using SM8 = Smem<TX, 8, false, false>;
  static_assert(sizeof(SM8) <= sizeof(SM16), "8-row tiles fit the 16-row tile's LDS");
  __shared__ SM16 sm;
  SM8& sm8 = *reinterpret_cast<SM8*>(&sm);
  const int total = hdr->total_tiles, nf = hdr->nfilms;
  const unsigned long long L = *counter;
  // XCD x = b mod 8 owns list indices 8x .. 8x + 7 of every 64 (neighbours
  // in a film row share its L2), and its G / 8 workgroups take its indices
  // strided, so the workgroups in flight cover one band of the list; odd
  // launches walk backwards.  The film index moves one way.
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b & 7, jj = b >> 3, GX = G >> 3;
  const int nx = ((total + 63) >> 6) << 3;  // indices per XCD, tail included
  const bool back = (L & 1ull) != 0;
  auto index_of = [&](int k) {  // k-th index of XCD x
    return ((k >> 3) << 6) + (x << 3) + (k & 7);
  };
  const int nmine = jj < nx ? (nx - 1 - jj) / GX + 1 : 0;
  SeriesAcc acc{0.0, 0.0, 0u};
  int lo = 0, nxt = 0, fst = 0;  // film of the tile, its successor's and its own first tile
  bool found = false;
  for (int i = 0; i < nmine; ++i) {
    const int k = back ? nx - 1 - jj - i * GX : jj + i * GX;
    const int tile = index_of(k);
    if (tile >= total) continue;  // the last 64 indices' tail
    if (!found) {
      int hi = nf - 1;
      lo = 0;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tstart[mid] <= tile) lo = mid;
        else hi = mid - 1;
      }
      nxt = lo + 1 < nf ? tstart[lo + 1] : total;
      fst = tstart[lo];
      found = true;
    }
    if (back) {
      while (tile < fst) fst = tstart[--lo];  // earlier film
    } else {
      while (tile >= nxt) {  // later film
        ++lo;
        nxt = lo + 1 < nf ? tstart[lo + 1] : total;
      }
    }
    const BatchFilm& F = films[lo];
    if (L < F.admit || L >= F.end) continue;
    if (!ANY_FT && F.read_ft) continue;  // never resident when this kernel runs
    const unsigned long long n = L - F.admit;
    const int cur = (int)(n & 1ull);
    TileIO io;
    io.h_in = F.h[cur];
    io.ux_in = F.ux[cur];
    io.uy_in = F.uy[cur];
    io.h_out = F.h[cur ^ 1];
    io.ux_out = F.ux[cur ^ 1];
    io.uy_out = F.uy[cur ^ 1];
    io.p_out = F.p;
    io.kappa = nullptr;
    if (F.read_ft) {
      io.ft = F.f[cur];
      io.f_out = F.f[cur ^ 1];
    } else {
      io.ft = nullptr;
      io.f_out = F.f[0];
    }
    if (F.last_only && !F.read_ft && L + 1 != F.end) {  // outputs nobody reads before the end
      io.f_out = nullptr;
      io.p_out = nullptr;
    }
    TileRun R;
    R.Lx = F.Lx;
    R.Ly = F.Ly;
    R.small = F.small;
    R.pairs = F.pairs;
    R.nts = nts;
    R.use_kappa = 0;
    R.body = F.body;
    R.fstride = F.fstride;
    R.seed = F.seed;
    R.step = (uint64_t)F.t0 + n + 1;
    R.wet_thresh = 0.0;
    const int local = tile - tstart[lo];
    const int x0 = (local % F.ntx) * TX, y0 = (local / F.ntx) * F.ty;
    const F32K K = F.K;
    if (F.ty == 8) {
      if (ANY_FT && F.read_ft) {
        if (F.thermal) tile_step<TX, NT, 8, true, true, false, false>(io, K, R, x0, y0, sm8, acc);
        else tile_step<TX, NT, 8, true, false, false, false>(io, K, R, x0, y0, sm8, acc);
      } else {
        if (F.thermal) tile_step<TX, NT, 8, false, true, false, false>(io, K, R, x0, y0, sm8, acc);
        else tile_step<TX, NT, 8, false, false, false, false>(io, K, R, x0, y0, sm8, acc);
      }
    } else {
      constexpr int TY = kBatchTileY;
      if (ANY_FT && F.read_ft) {
        if (F.thermal) tile_step<TX, NT, TY, true, true, false, false>(io, K, R, x0, y0, sm, acc);
        else tile_step<TX, NT, TY, true, false, false, false>(io, K, R, x0, y0, sm, acc);
      } else {
        if (F.thermal) tile_step<TX, NT, TY, false, true, false, false>(io, K, R, x0, y0, sm, acc);
        else tile_step<TX, NT, TY, false, false, false, false>(io, K, R, x0, y0, sm, acc);
      }
    }
    __syncthreads();
  }
}

// Batch of films on strips: one wavefront per (film, strip, segment) item.
// The fixed grid strides over each XCD's contiguous share of the items.
__global__ __launch_bounds__(64) void k_batch_strip_f32(const BatchFilm* __restrict__ films,
                                                      const int* __restrict__ tstart,
                                                      const BatchHdr* __restrict__ hdr,
                                                      const unsigned long long* __restrict__ counter,
                                                      int nts) {
  const int total = hdr->total_tiles, nf = hdr->nfilms;
  const unsigned long long L = *counter;
  const int xcd = blockIdx.x & 7, nj = gridDim.x >> 3;
  const int chunk = (total + 7) >> 3;
  const int tend = min(xcd * chunk + chunk, total);
  SeriesAcc acc{0.0, 0.0, 0u};
  for (int item = xcd * chunk + (blockIdx.x >> 3); item < tend; item += nj) {
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tstart[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const BatchFilm& F = films[lo];
    if (L < F.admit || L >= F.end) continue;
    const unsigned long long n = L - F.admit;
    const int cur = (int)(n & 1ull);
    TileIO io;
    io.h_in = F.h[cur];
    io.ux_in = F.ux[cur];
    io.uy_in = F.uy[cur];
    io.h_out = F.h[cur ^ 1];
    io.ux_out = F.ux[cur ^ 1];
    io.uy_out = F.uy[cur ^ 1];
    io.p_out = F.p;
    io.kappa = nullptr;
    if (F.read_ft) {
      io.ft = F.f[cur];
      io.f_out = F.f[cur ^ 1];
    } else {
      io.ft = nullptr;
      io.f_out = F.f[0];
    }
    if (F.last_only && !F.read_ft && L + 1 != F.end) {  // outputs nobody reads before the end
      io.f_out = nullptr;
      io.p_out = nullptr;
    }
    TileRun R;
    R.Lx = F.Lx;
    R.Ly = F.Ly;
    R.small = F.small;
    R.pairs = 0;
    R.nts = nts;
    R.use_kappa = 0;
    R.body = F.body;
    R.fstride = F.fstride;
    R.seed = F.seed;
    R.step = (uint64_t)F.t0 + n + 1;
    R.wet_thresh = 0.0;
    const int local = item - tstart[lo];
    const int sg = local / F.ntx, st = local - sg * F.ntx;
    const int y0 = sg * kStripSeg;
    const int S = min(kStripSeg, F.Ly - y0);
    const F32K K = F.K;
    if (F.read_ft) {
      with_store_mode(nts, [&](auto mode) {
        constexpr int NTS = decltype(mode)::value;
        if (F.thermal) strip_run<true, true, false, NTS>(io, K, R, st * kStripOut, y0, S, acc);
        else strip_run<true, false, false, NTS>(io, K, R, st * kStripOut, y0, S, acc);
      });
    } else {
      with_store_mode(nts, [&](auto mode) {
        constexpr int NTS = decltype(mode)::value;
        if (F.thermal) strip_run<false, true, false, NTS>(io, K, R, st * kStripOut, y0, S, acc);
        else strip_run<false, false, false, NTS>(io, K, R, st * kStripOut, y0, S, acc);
      });
    }
  }
}

__global__ void k_batch_advance(unsigned long long* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter += 1ull;
}

}  // namespace

F32K make_f32k(const StepParams& sp, const Enc& enc) {
  F32K K;
  K.gamma = (float)sp.gamma;
  K.kap = (float)sp.kappa;
  K.hmin = (float)sp.hmin;
  K.hcrit = (float)sp.hcrit;
  K.c6mu = (float)sp.c6mu;
  K.c6d = (float)sp.c6d;
  K.c3dd = (float)sp.c3dd;
  K.kb26 = (float)sp.kb26;
  K.omeg = (float)sp.omeg;
  K.itau = (float)sp.itau;
  K.c56g = (float)((5.0 / 6.0) * sp.g);
  K.g0 = (float)(1.5 * sp.g);
  K.href = enc.href;
  K.bfx = (float)(sp.bax * sp.bfac);
  K.bfy = (float)(sp.bay * sp.bfac);
  K.n = sp.n;
  K.m = sp.m;
  // the (9,3) and (3,2) laws are evaluated in factored form
  K.powmode = (sp.n == 9 && sp.m == 3) ? 1 : ((sp.n == 3 && sp.m == 2) ? 2 : 0);
  K.guard = sp.fp64_guard ? 1 : 0;
  K.hguard = (float)(4.0 * sp.hmin - sp.hcrit);
  K.gamma_d = sp.gamma;
  K.kap_d = sp.kappa;
  K.hmin_d = sp.hmin;
  K.hcrit_d = sp.hcrit;
  K.href_d = (double)enc.href;
  return K;
}

int film_is_small(int Lx, int Ly, int tile_x, int tile_y) {
  return (Lx < tile_x + 6 || Ly < tile_y + 6) ? 1 : 0;
}

// variant: 0 = default (collide at the destination), 1 = collide at the
// source node, 2 = collide at the destination, 3 = LDS-free strips.
// nts: non-temporal stores.
void launch_strip(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                  int Lx, int Ly, int64_t fstride, Enc enc, int seg, int nts, hipStream_t s) {
  TileIO t{io.h_in, io.ux_in, io.uy_in, io.ft, io.f_out, io.h_out, io.ux_out, io.uy_out,
           io.p_out, io.kappa};
  TileRun R;
  R.Lx = Lx;
  R.Ly = Ly;
  R.small = strip_is_small(Lx, Ly);
  R.pairs = 0;
  R.nts = nts;
  R.use_kappa = io.kappa != nullptr;
  R.body = sp.body;
  R.fstride = fstride;
  R.seed = sp.seed;
  R.step = step;
  R.wet_thresh = sp.wet_thresh;
  const F32K K = make_f32k(sp, enc);
  const int nstrips = (Lx + kStripOut - 1) / kStripOut, nseg = (Ly + seg - 1) / seg;
  const int items = nstrips * nseg, chunk = (items + 7) / 8;
  const dim3 grid(8 * chunk), block(64);
#define SW_L(RF, TH)                                                                           \
  hipLaunchKernelGGL((k_strip_f32<RF, TH>), grid, block, 0, s, t, K, R, nstrips, seg, items, chunk, \
                     ser)
  if (sp.read_ft) {
    if (sp.thermal) SW_L(true, true);
    else SW_L(true, false);
  } else {
    if (sp.thermal) SW_L(false, true);
    else SW_L(false, false);
  }
#undef SW_L
}

int strip_is_small(int Lx, int Ly) { return (Lx < 64 || Ly < 8) ? 1 : 0; }

void launch_step_f32(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                     int Lx, int Ly, int64_t fstride, Enc enc, int tile_y, int variant, int nts,
                     hipStream_t s) {
  const bool t8 = tile_y == 8;
  if (variant == 3) {
    launch_strip(io, sp, ser, step, Lx, Ly, fstride, enc, tile_y > 0 ? tile_y : kStripSeg, nts, s);
  } else if (variant == 1) {
    if (t8) launch_single<TX, NT, 8, true>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
    else launch_single<TX, NT, 16, true>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
  } else {
    if (t8) launch_single<TX, NT, 8, false>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
    else launch_single<TX, NT, 16, false>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
  }
}

void launch_batch_f32(const BatchFilm* films, const int* tstart, const BatchHdr* hdr,
                      unsigned long long* counter, int grid, int nts, bool any_ft,
                      hipStream_t s) {
  if (kBatchStrips)
    hipLaunchKernelGGL(k_batch_strip_f32, dim3(grid), dim3(64), 0, s, films, tstart, hdr, counter,
                       nts);
  else if (any_ft)
    hipLaunchKernelGGL(k_batch_f32<true>, dim3(grid), dim3(NT), 0, s, films, tstart, hdr,
                       counter, nts);
  else
    hipLaunchKernelGGL(k_batch_f32<false>, dim3(grid), dim3(NT), 0, s, films, tstart, hdr,
                       counter, nts);
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, s, counter);
}

int batch_blocks_per_cu(bool any_ft) {
  int per_cu = 0;
  const hipError_t e =
      kBatchStrips
          ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_strip_f32, 64, 0)
      : any_ft ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_f32<true>, NT, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_f32<false>, NT, 0);
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  return per_cu;
}

}  // namespace swalbe
