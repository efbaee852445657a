// step_f32.hip -- the fused thin-film step for FP32 films (deviation
// storage), tuned for gfx950.
//
// Same step and tile decomposition as the FP64 kernel of step.hip (64 x TY
// tile, h on a 3-halo, p on 2, F on 1), with the FP32 liberties the
// tolerance-checked path allows:
//   * FMA contraction and v_rcp_f32 reciprocals instead of IEEE division;
//   * Box-Muller with v_log/v_sin/v_cos/v_sqrt_f32, one Philox4x32-10 call
//     per pair of x-adjacent nodes (the pair's four words);
//   * flat thread mappings for every tile phase (no divergent tail columns);
//   * SRC = true: each F node collides all nine of its populations once and
//     parks them in LDS, so the stream is a gather; SRC = false keeps F in
//     LDS and collides at the destination;
//   * PERSIST = true: a grid of resident workgroups walks its XCD's share of
//     the tiles and prefetches the next tile's halo into registers while it
//     computes the current one (software pipelining of HBM and ALU work).
// Height and distributions are deviations from a rest film of height href
// (Enc): dh = h - href, f' = f - feq(href, 0).
#include "swalbe_internal.h"

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

struct F32K {
  float gamma, kap, hmin, hcrit, c6mu, c6d, c3dd, kb26, omeg, itau, c56g, g0, href;
  float bfx, bfy;  // body force α·fac
  int n, m, powmode;
};

struct F32Run {
  int use_kappa, body, Lx, Ly, ntx, ntiles, small, pairs, nts;
  int64_t fstride;  // distance between distribution planes (elements)
  double wet_thresh;
  uint64_t seed, step;
  StepSeries ser;
};

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

__device__ __forceinline__ void store_f(float* p, float v, int nts) {
  if (nts) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <int TY, bool READ_FT, bool THERMAL, bool SRC, bool PERSIST>
__global__ __launch_bounds__(NT) void k_step_f32(StepIO<float> io, F32K K, F32Run R) {
#pragma clang fp contract(fast)
  constexpr int HC = TX + 6, HR = TY + 6, HN = HR * HC;
  constexpr int PC = TX + 4, PR = TY + 4, PN = PR * PC;
  constexpr int UC = TX + 2, UR = TY + 2, UN = UR * UC;
  constexpr int NHI = (HN + NT - 1) / NT, NUI = (UN + NT - 1) / NT;
  __shared__ float sh_h[HN];
  __shared__ float sh_p[PN];
  __shared__ float sh_u[UN], sh_v[UN];
  __shared__ float sh_a[(SRC ? 9 : 2) * UN];
  __shared__ unsigned long long red_k[2 * NW];
  __shared__ unsigned int red_c[NW];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int Lx = R.Lx, Ly = R.Ly;
  const int64_t FS = R.fstride;
  const int small = R.small;
  auto wx = [&](int a) { return small ? wrapm(a, Lx) : wrap1(a, Lx); };
  auto wy = [&](int a) { return small ? wrapm(a, Ly) : wrap1(a, Ly); };

  float hreg[NHI], ureg[NUI], vreg[NUI];
  auto load_regs = [&](int x0, int y0) {
#pragma unroll
    for (int i = 0; i < NHI; ++i) {
      const int e = tid + i * NT;
      if (e < HN) {
        const int r = e / HC, c = e - r * HC;
        hreg[i] = io.h_in[wx(x0 - 3 + c) + (int64_t)Lx * wy(y0 - 3 + r)];
      }
    }
#pragma unroll
    for (int i = 0; i < NUI; ++i) {
      const int e = tid + i * NT;
      if (e < UN) {
        const int r = e / UC, c = e - r * UC;
        const int64_t g = wx(x0 - 1 + c) + (int64_t)Lx * wy(y0 - 1 + r);
        ureg[i] = io.ux_in[g];
        vreg[i] = io.uy_in[g];
      }
    }
  };
  auto regs_to_lds = [&]() {
#pragma unroll
    for (int i = 0; i < NHI; ++i) {
      const int e = tid + i * NT;
      if (e < HN) sh_h[e] = hreg[i];
    }
#pragma unroll
    for (int i = 0; i < NUI; ++i) {
      const int e = tid + i * NT;
      if (e < UN) {
        sh_u[e] = ureg[i];
        sh_v[e] = vreg[i];
      }
    }
  };

  double lmax = -__builtin_inf(), lmin = __builtin_inf();
  unsigned int lcnt = 0;
  const bool series = R.ser.dh_keys || R.ser.wet;
  const float w1 = 1.0f / 9.0f, w5 = 1.0f / 36.0f;

  // phases B, C, D of one tile whose halo is in LDS
  auto process = [&](int x0, int y0) {
    // ---- B: film pressure on the 2-halo
    for (int e = tid; e < PN; e += NT) {
      const int r = e / PC, c = e - r * PC;
      const float* hc = &sh_h[(r + 1) * HC + (c + 1)];
      const float hip = hc[-1], him = hc[1], hjp = hc[-HC], hjm = hc[HC];
      const float hipjp = hc[-HC - 1], himjp = hc[-HC + 1], himjm = hc[HC + 1], hipjm = hc[HC - 1];
      const float L = (2.0f / 3.0f) * (((hjp + hip) + him) + hjm) +
                      (1.0f / 6.0f) * (((hipjp + himjp) + himjm) + hipjm) - (10.0f / 3.0f) * hc[0];
      const float hh = K.href + hc[0];
      const float a = K.hmin * __builtin_amdgcn_rcpf(hh + K.hcrit);
      float kap = K.kap;
      if (R.use_kappa) kap = io.kappa[wx(x0 - 2 + c) + (int64_t)Lx * wy(y0 - 2 + r)];
      sh_p[e] = -K.gamma * (kap * powlaw_f(a, K.n, K.m, K.powmode)) - K.gamma * L;
    }
    __syncthreads();

    // ---- C: forces (and with SRC the collision) on the 1-halo
    auto fnode = [&](int r, int c, float z0, float z1) {
      const float* pc = &sh_p[(r + 1) * PC + (c + 1)];
      const float fip = pc[-1], fim = pc[1], fjp = pc[-PC], fjm = pc[PC];
      const float fipjp = pc[-PC - 1], fimjp = pc[-PC + 1], fimjm = pc[PC + 1], fipjm = pc[PC - 1];
      const float dh = sh_h[(r + 2) * HC + (c + 2)];
      const float hh = K.href + dh;
      const float gx = (-1.0f / 3.0f) * (fip - fim) - (1.0f / 12.0f) * (((fipjp - fimjp) - fimjm) + fipjm);
      const float gy = (-1.0f / 3.0f) * (fjp - fjm) - (1.0f / 12.0f) * (((fipjp + fimjp) - fimjm) - fipjm);
      const int ui = r * UC + c;
      const float u = sh_u[ui], v = sh_v[ui];
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
          float f = K.itau * fe[k] + S[k];
          if constexpr (READ_FT) f += K.omeg * io.ft[k * FS + gsrc];
          sh_a[k * UN + ui] = f;
        }
      } else {
        sh_a[ui] = Fx;
        sh_a[UN + ui] = Fy;
      }
    };
    if (R.pairs) {
      // 34 pair columns cover global x0-2 .. x0+65; tile column c = 2 pc - 1 + b
      constexpr int NPC = (TX + 4) / 2;
      for (int t = tid; t < UR * NPC; t += NT) {
        const int r = t / NPC, pc = t - r * NPC;
        float z[4] = {0.f, 0.f, 0.f, 0.f};
        if constexpr (THERMAL) {
          const int gxp = wx(x0 - 2 + 2 * pc);
          const int64_t node = gxp + (int64_t)Lx * wy(y0 - 1 + r);
          uint32_t w[4];
          philox_ctr(R.seed, R.step, (uint32_t)(node >> 1), kTagThermal, w);
          box_muller_f32(w[0], w[1], &z[0], &z[1]);
          box_muller_f32(w[2], w[3], &z[2], &z[3]);
        }
        const int c0 = 2 * pc - 1;
        if (c0 >= 0) fnode(r, c0, z[0], z[1]);
        if (c0 + 1 < UC) fnode(r, c0 + 1, z[2], z[3]);
      }
    } else {
      for (int t = tid; t < UN; t += NT) {
        const int r = t / UC, c = t - r * UC;
        float z0 = 0.f, z1 = 0.f;
        if constexpr (THERMAL) {
          const int64_t node = wx(x0 - 1 + c) + (int64_t)Lx * wy(y0 - 1 + r);
          uint32_t w[4];
          philox_ctr(R.seed, R.step, (uint32_t)(node >> 1), kTagThermal, w);
          const int b = (int)(node & 1);
          box_muller_f32(w[2 * b], w[2 * b + 1], &z0, &z1);
        }
        fnode(r, c, z0, z1);
      }
    }
    __syncthreads();

    // ---- D: stream (gather) and moments
    const int gxo = x0 + lane;
    for (int row = wv; row < TY; row += NW) {
      const int gyo = y0 + row;
      if (gxo >= Lx || gyo >= Ly) continue;
      const int64_t xo = gxo + (int64_t)Lx * gyo;
      float fo[kQ];
#pragma unroll
      for (int k = 0; k < kQ; ++k) {
        const int su = (row + 1 - cy(k)) * UC + (lane + 1 - cx(k));
        float f;
        if constexpr (SRC) {
          f = sh_a[k * UN + su];
        } else {
          const float dh = sh_h[(row + 3 - cy(k)) * HC + (lane + 3 - cx(k))];
          const float hh = K.href + dh;
          const float u = sh_u[su], v = sh_v[su];
          const float usq = u * u + v * v;
          const float gg = dh * (hh + K.href);
          const float Fx = sh_a[su], Fy = sh_a[UN + su];
          float fe, S;
          if (k == 0) {
            fe = (dh - K.c56g * gg) - (2.0f / 3.0f) * hh * usq;
            S = 0.0f;
          } else {
            const float a = k == 1 ? u : k == 2 ? v : k == 3 ? -u : k == 4 ? -v
                          : k == 5 ? u + v : k == 6 ? v - u : k == 7 ? -(u + v) : u - v;
            const float w = k <= 4 ? w1 : w5;
            fe = w * (K.g0 * gg + hh * ((4.5f * a + 3.0f) * a - 1.5f * usq));
            const float cF = k <= 4 ? (1.0f / 3.0f) : (1.0f / 24.0f);
            const float Fa = k == 1 ? Fx : k == 2 ? Fy : k == 3 ? -Fx : k == 4 ? -Fy
                           : k == 5 ? Fx + Fy : k == 6 ? Fy - Fx : k == 7 ? -(Fx + Fy) : Fx - Fy;
            S = cF * Fa;
          }
          f = K.itau * fe + S;
          if constexpr (READ_FT) {
            const int sx = wrap1(gxo - cx(k), Lx), sy = wrap1(gyo - cy(k), Ly);
            f += K.omeg * io.ft[k * FS + sx + (int64_t)Lx * sy];
          }
        }
        fo[k] = f;
        store_f(&io.f_out[k * FS + xo], f, R.nts);
      }
      const float s = (((((((fo[0] + fo[1]) + fo[2]) + fo[3]) + fo[4]) + fo[5]) + fo[6]) + fo[7]) + fo[8];
      const float ih = __builtin_amdgcn_rcpf(K.href + s);
      store_f(&io.h_out[xo], s, R.nts);
      store_f(&io.ux_out[xo], ((((fo[1] - fo[3]) + fo[5]) - fo[6]) - fo[7] + fo[8]) * ih, R.nts);
      store_f(&io.uy_out[xo], ((((fo[2] - fo[4]) + fo[5]) + fo[6]) - fo[7] - fo[8]) * ih, R.nts);
      store_f(&io.p_out[xo], sh_p[(row + 2) * PC + (lane + 2)], R.nts);
      if (series) {
        const double hv = (double)K.href + (double)sh_h[(row + 3) * HC + (lane + 3)];
        lmax = hv > lmax ? hv : lmax;
        lmin = hv < lmin ? hv : lmin;
        lcnt += hv > R.wet_thresh ? 1u : 0u;
      }
    }
  };

  const int ntx = R.ntx, ntiles = R.ntiles;
  if constexpr (PERSIST) {
    // XCD x = blockIdx % 8 walks tiles [x C, (x+1) C) with stride nj
    const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3, nj = gridDim.x >> 3;
    const int chunk = (ntiles + 7) >> 3;
    const int tend = min(xcd * chunk + chunk, ntiles);
    int tile = xcd * chunk + j;
    if (tile < tend) load_regs((tile % ntx) * TX, (tile / ntx) * TY);
    while (tile < tend) {
      regs_to_lds();
      __syncthreads();
      const int next = tile + nj;
      if (next < tend) load_regs((next % ntx) * TX, (next / ntx) * TY);
      process((tile % ntx) * TX, (tile / ntx) * TY);
      __syncthreads();
      tile = next;
    }
  } else {
    int tile = blockIdx.x;
    if ((ntiles & 7) == 0) tile = (tile & 7) * (ntiles >> 3) + (tile >> 3);
    const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY;
    load_regs(x0, y0);
    regs_to_lds();
    __syncthreads();
    process(x0, y0);
  }

  // ---- per-step series: max/min/count of h_t over the domain (exact)
  if (series) {
    unsigned long long kmax = dkey(lmax), kmin = dkey(lmin);
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a = __shfl_xor(kmax, o), b = __shfl_xor(kmin, o);
      kmax = a > kmax ? a : kmax;
      kmin = b < kmin ? b : kmin;
      lcnt += __shfl_xor(lcnt, o);
    }
    __syncthreads();
    if (lane == 0) {
      red_k[2 * wv] = kmax;
      red_k[2 * wv + 1] = kmin;
      red_c[wv] = lcnt;
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < NW; ++w) {
        kmax = red_k[2 * w] > kmax ? red_k[2 * w] : kmax;
        kmin = red_k[2 * w + 1] < kmin ? red_k[2 * w + 1] : kmin;
        lcnt += red_c[w];
      }
      if (R.ser.dh_keys) {
        atomicMax(&R.ser.dh_keys[0], kmax);
        atomicMin(&R.ser.dh_keys[1], kmin);
      }
      if (R.ser.wet) atomicAdd(&R.ser.wet[0], (unsigned long long)lcnt);
    }
  }
}

int g_cu_count = 0;

template <int TY, bool SRC, bool PERSIST>
void launch_f32_cfg(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                    int Lx, int Ly, int64_t fstride, Enc enc, int nts, hipStream_t s) {
  const int ntx = (Lx + TX - 1) / TX, nty = (Ly + TY - 1) / TY;
  F32Run R;
  R.ntx = ntx;
  R.ntiles = ntx * nty;
  R.small = (Lx < TX + 6 || Ly < TY + 6) ? 1 : 0;
  R.pairs = (!R.small && (Lx % 2) == 0) ? 1 : 0;
  R.use_kappa = io.kappa != nullptr;
  R.body = sp.body;
  R.Lx = Lx;
  R.Ly = Ly;
  R.nts = nts;
  R.fstride = fstride;
  R.wet_thresh = sp.wet_thresh;
  R.seed = sp.seed;
  R.step = step;
  R.ser = ser;
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
#define SW_KER(RF, TH) k_step_f32<TY, RF, TH, SRC, PERSIST>
  auto pick = [&]() -> const void* {
    if (sp.read_ft) return sp.thermal ? (const void*)SW_KER(true, true) : (const void*)SW_KER(true, false);
    return sp.thermal ? (const void*)SW_KER(false, true) : (const void*)SW_KER(false, false);
  };
  int grid = R.ntiles;
  if (PERSIST) {
    if (!g_cu_count) {
      int dev = 0;
      hipGetDevice(&dev);
      hipDeviceGetAttribute(&g_cu_count, hipDeviceAttributeMultiprocessorCount, dev);
      if (g_cu_count <= 0) g_cu_count = 256;
    }
    static int cache[2][2] = {{0, 0}, {0, 0}};
    int& per_cu = cache[sp.read_ft ? 1 : 0][sp.thermal ? 1 : 0];
    if (per_cu < 1) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pick(), NT, 0) != hipSuccess ||
          per_cu < 1)
        per_cu = 1;
    }
    grid = g_cu_count * per_cu;
    grid = (grid + 7) / 8 * 8;
    const int need = (R.ntiles + 7) / 8 * 8;
    if (grid > need) grid = need;
  }
  if (sp.read_ft) {
    if (sp.thermal)
      hipLaunchKernelGGL((SW_KER(true, true)), dim3(grid), dim3(NT), 0, s, io, K, R);
    else
      hipLaunchKernelGGL((SW_KER(true, false)), dim3(grid), dim3(NT), 0, s, io, K, R);
  } else {
    if (sp.thermal)
      hipLaunchKernelGGL((SW_KER(false, true)), dim3(grid), dim3(NT), 0, s, io, K, R);
    else
      hipLaunchKernelGGL((SW_KER(false, false)), dim3(grid), dim3(NT), 0, s, io, K, R);
  }
#undef SW_KER
}

}  // namespace

// variant: 0 = default (collide at the destination, one tile per
// workgroup), 1 = collide at the source node, 2 = collide at the destination,
// 3 = persistent + source, 4 = persistent + destination.  nts: non-temporal
// stores.
void launch_step_f32(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                     int Lx, int Ly, int64_t fstride, Enc enc, int tile_y, int variant, int nts,
                     hipStream_t s) {
  const int v = variant == 0 ? 2 : variant;
  const bool t8 = tile_y == 8;
  switch (v) {
    case 1:
      if (t8) launch_f32_cfg<8, true, false>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      else launch_f32_cfg<16, true, false>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      break;
    case 2:
      if (t8) launch_f32_cfg<8, false, false>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      else launch_f32_cfg<16, false, false>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      break;
    case 3:
      if (t8) launch_f32_cfg<8, true, true>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      else launch_f32_cfg<16, true, true>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      break;
    default:
      if (t8) launch_f32_cfg<8, false, true>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      else launch_f32_cfg<16, false, true>(io, sp, ser, step, Lx, Ly, fstride, enc, nts, s);
      break;
  }
}

}  // namespace swalbe
