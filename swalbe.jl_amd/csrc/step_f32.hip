// step_f32.hip -- the fused thin-film step for FP32 films (deviation
// storage), tuned for instruction count on gfx950.
//
// Same step and tile decomposition as the FP64 kernel of step.hip (64 x TY
// tile, h on a 3-halo, p on 2, F on 1), with the FP32 liberties the
// tolerance-checked path allows:
//   * FMA contraction and v_rcp_f32 reciprocals instead of IEEE division;
//   * Box-Muller with v_log/v_sin/v_cos/v_sqrt_f32, one Philox4x32-10 call
//     per pair of x-adjacent nodes (the pair's four words);
//   * halo rows loaded by whole wavefronts (one row offset per row);
//   * SRC = true: each F node collides all nine of its populations once and
//     parks them in LDS; the stream is then a gather from LDS (no per-source
//     recomputation of feq).  SRC = false keeps F in LDS and collides at the
//     destination.
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

template <int TY, bool READ_FT, bool THERMAL, bool SRC>
__global__ __launch_bounds__(NT) void k_step_f32(StepIO<float> io, F32K K, int use_kappa,
                                                 int body, StepSeries ser, double wet_thresh,
                                                 uint64_t seed, uint64_t step, int Lx, int Ly,
                                                 int ntx, int ntiles, int small, int pairs) {
#pragma clang fp contract(fast)
  constexpr int HC = TX + 6, HR = TY + 6;
  constexpr int PC = TX + 4, PR = TY + 4;
  constexpr int UC = TX + 2, UR = TY + 2;
  constexpr int UN = UR * UC;
  __shared__ float sh_h[HR * HC];
  __shared__ float sh_p[PR * PC];
  __shared__ float sh_u[UN], sh_v[UN];
  __shared__ float sh_a[(SRC ? 9 : 2) * UN];
  __shared__ unsigned long long red_k[2 * NW];
  __shared__ unsigned int red_c[NW];

  int tile = blockIdx.x;
  if ((ntiles & 7) == 0) tile = (tile & 7) * (ntiles >> 3) + (tile >> 3);
  const int x0 = (tile % ntx) * TX;
  const int y0 = (tile / ntx) * TY;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int64_t N = (int64_t)Lx * Ly;
  auto wx = [&](int a) { return small ? wrapm(a, Lx) : wrap1(a, Lx); };
  auto wy = [&](int a) { return small ? wrapm(a, Ly) : wrap1(a, Ly); };

  // ---- phase A: halo rows of h (3) and u (1), a wavefront per row
  {
    const int ca = wx(x0 - 3 + lane);
    const int cb = wx(x0 + 61 + (lane & 7));  // columns 64..69 for lanes 0..5
    for (int r = wv; r < HR; r += NW) {
      const float* row = io.h_in + (int64_t)Lx * wy(y0 - 3 + r);
      sh_h[r * HC + lane] = row[ca];
      if (lane < 6) sh_h[r * HC + 64 + lane] = row[cb];
    }
    const int ua = wx(x0 - 1 + lane);
    const int ub = wx(x0 + 63 + (lane & 1));
    for (int r = wv; r < UR; r += NW) {
      const int64_t ro = (int64_t)Lx * wy(y0 - 1 + r);
      sh_u[r * UC + lane] = io.ux_in[ro + ua];
      sh_v[r * UC + lane] = io.uy_in[ro + ua];
      if (lane < 2) {
        sh_u[r * UC + 64 + lane] = io.ux_in[ro + ub];
        sh_v[r * UC + 64 + lane] = io.uy_in[ro + ub];
      }
    }
  }
  __syncthreads();

  // ---- phase B: film pressure on the 2-halo
  auto pnode = [&](int r, int c) {
    const float* hc = &sh_h[(r + 1) * HC + (c + 1)];
    const float hip = hc[-1], him = hc[1], hjp = hc[-HC], hjm = hc[HC];
    const float hipjp = hc[-HC - 1], himjp = hc[-HC + 1], himjm = hc[HC + 1], hipjm = hc[HC - 1];
    const float L = (2.0f / 3.0f) * (((hjp + hip) + him) + hjm) +
                    (1.0f / 6.0f) * (((hipjp + himjp) + himjm) + hipjm) - (10.0f / 3.0f) * hc[0];
    const float hh = K.href + hc[0];
    const float a = K.hmin * __builtin_amdgcn_rcpf(hh + K.hcrit);
    float kap = K.kap;
    if (use_kappa) kap = io.kappa[wx(x0 - 2 + c) + (int64_t)Lx * wy(y0 - 2 + r)];
    sh_p[r * PC + c] = -K.gamma * (kap * powlaw_f(a, K.n, K.m, K.powmode)) - K.gamma * L;
  };
  for (int r = wv; r < PR; r += NW) {
    pnode(r, lane);
    if (lane < 4) pnode(r, 64 + lane);
  }
  __syncthreads();

  // ---- phase C: forces (and with SRC the collision) on the 1-halo
  const float w1 = 1.0f / 9.0f, w5 = 1.0f / 36.0f;
  auto fnode = [&](int r, int c, float z0, float z1) {
    const float* pc = &sh_p[(r + 1) * PC + (c + 1)];
    const float fip = pc[-1], fim = pc[1], fjp = pc[-PC], fjm = pc[PC];
    const float fipjp = pc[-PC - 1], fimjp = pc[-PC + 1], fimjm = pc[PC + 1], fipjm = pc[PC - 1];
    const float dh = sh_h[(r + 2) * HC + (c + 2)];
    const float hh = K.href + dh;
    const float gx = (-1.0f / 3.0f) * (fip - fim) - (1.0f / 12.0f) * (((fipjp - fimjp) - fimjm) + fipjm);
    const float gy = (-1.0f / 3.0f) * (fjp - fjm) - (1.0f / 12.0f) * (((fipjp + fimjp) - fimjm) - fipjm);
    const int u_i = r * UC + c;
    const float u = sh_u[u_i], v = sh_v[u_i];
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
    if (body) {
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
      // equilibrium deviations feq'_k and force terms S_k of this node
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
        if constexpr (READ_FT) f += K.omeg * io.ft[k * N + gsrc];
        sh_a[k * UN + u_i] = f;
      }
    } else {
      sh_a[u_i] = Fx;
      sh_a[UN + u_i] = Fy;
    }
  };

  if (pairs) {
    // 34 pair columns cover global x0-2 .. x0+65; tile column c = 2 pc - 1 + b
    constexpr int NPC = 34;
    for (int t = tid; t < UR * NPC; t += NT) {
      const int r = t / NPC, pc = t - r * NPC;
      float z[4] = {0.f, 0.f, 0.f, 0.f};
      if constexpr (THERMAL) {
        const int gx = wx(x0 - 2 + 2 * pc);
        const int64_t node = gx + (int64_t)Lx * wy(y0 - 1 + r);
        uint32_t w[4];
        philox_ctr(seed, step, (uint32_t)(node >> 1), kTagThermal, w);
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
        philox_ctr(seed, step, (uint32_t)(node >> 1), kTagThermal, w);
        const int b = (int)(node & 1);
        box_muller_f32(w[2 * b], w[2 * b + 1], &z0, &z1);
      }
      fnode(r, c, z0, z1);
    }
  }
  __syncthreads();

  // ---- phase D: stream (gather) and moments
  double lmax = -__builtin_inf(), lmin = __builtin_inf();
  unsigned int lcnt = 0;
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
          f += K.omeg * io.ft[k * N + sx + (int64_t)Lx * sy];
        }
      }
      fo[k] = f;
      io.f_out[k * N + xo] = f;
    }
    const float s = (((((((fo[0] + fo[1]) + fo[2]) + fo[3]) + fo[4]) + fo[5]) + fo[6]) + fo[7]) + fo[8];
    const float ih = __builtin_amdgcn_rcpf(K.href + s);
    io.h_out[xo] = s;
    io.ux_out[xo] = ((((fo[1] - fo[3]) + fo[5]) - fo[6]) - fo[7] + fo[8]) * ih;
    io.uy_out[xo] = ((((fo[2] - fo[4]) + fo[5]) + fo[6]) - fo[7] - fo[8]) * ih;
    io.p_out[xo] = sh_p[(row + 2) * PC + (lane + 2)];
    if (ser.dh_keys || ser.wet) {
      const double hv = (double)K.href + (double)sh_h[(row + 3) * HC + (lane + 3)];
      lmax = hv > lmax ? hv : lmax;
      lmin = hv < lmin ? hv : lmin;
      lcnt += hv > wet_thresh ? 1u : 0u;
    }
  }

  if (ser.dh_keys || ser.wet) {
    unsigned long long kmax = dkey(lmax), kmin = dkey(lmin);
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long a = __shfl_xor(kmax, o), b = __shfl_xor(kmin, o);
      kmax = a > kmax ? a : kmax;
      kmin = b < kmin ? b : kmin;
      lcnt += __shfl_xor(lcnt, o);
    }
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
      if (ser.dh_keys) {
        atomicMax(&ser.dh_keys[0], kmax);
        atomicMin(&ser.dh_keys[1], kmin);
      }
      if (ser.wet) atomicAdd(&ser.wet[0], (unsigned long long)lcnt);
    }
  }
}

template <int TY, bool SRC>
void launch_f32_ty(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                   int Lx, int Ly, Enc enc, hipStream_t s) {
  const int ntx = (Lx + TX - 1) / TX, nty = (Ly + TY - 1) / TY;
  const int ntiles = ntx * nty;
  const int small = (Lx < TX + 6 || Ly < TY + 6) ? 1 : 0;
  const int pairs = (!small && (Lx % 2) == 0) ? 1 : 0;
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
  // the fused FP32 step evaluates the (9,3) and (3,2) laws in factored form
  K.powmode = (sp.n == 9 && sp.m == 3) ? 1 : ((sp.n == 3 && sp.m == 2) ? 2 : 0);
  const int use_kappa = io.kappa != nullptr;
  dim3 grid(ntiles), block(NT);
#define SW_L(RF, TH)                                                                           \
  hipLaunchKernelGGL((k_step_f32<TY, RF, TH, SRC>), grid, block, 0, s, io, K, use_kappa,       \
                     sp.body, ser, sp.wet_thresh, sp.seed, step, Lx, Ly, ntx, ntiles, small, pairs)
  if (sp.read_ft) {
    if (sp.thermal) SW_L(true, true);
    else SW_L(true, false);
  } else {
    if (sp.thermal) SW_L(false, true);
    else SW_L(false, false);
  }
#undef SW_L
}

}  // namespace

// variant: 0 = default (collide at source), 1 = collide at source, 2 = collide at destination
void launch_step_f32(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                     int Lx, int Ly, Enc enc, int tile_y, int variant, hipStream_t s) {
  const bool src = variant != 2;
  if (src) {
    if (tile_y == 8) launch_f32_ty<8, true>(io, sp, ser, step, Lx, Ly, enc, s);
    else launch_f32_ty<16, true>(io, sp, ser, step, Lx, Ly, enc, s);
  } else {
    if (tile_y == 8) launch_f32_ty<8, false>(io, sp, ser, step, Lx, Ly, enc, s);
    else if (tile_y == 32) launch_f32_ty<32, false>(io, sp, ser, step, Lx, Ly, enc, s);
    else launch_f32_ty<16, false>(io, sp, ser, step, Lx, Ly, enc, s);
  }
}

}  // namespace swalbe
