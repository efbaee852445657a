// ops.hip -- one kernel per reference op (the unfused, drop-in granularity).
//
// Every FP64 expression is written in the evaluation order of the reference
// broadcast it replaces and the library is built with -ffp-contract=off, so
// each op reproduces the reference's Float64 arithmetic bit for bit.  FP32
// arrays may carry the deviation encoding of Film (decode on load, encode on
// store); raw caller arrays use zero_enc().
#include "swalbe_internal.h"

namespace swalbe {

namespace {

constexpr int kBlock1D = 256;
constexpr int kBX = 64, kBY = 4;  // one wavefront per row for 2D stencils

template <typename T>
__device__ __forceinline__ T dec_h(T v, const Enc& e) {
  if constexpr (sizeof(T) == 4) return e.href + v;
  else return v;
}
template <typename T>
__device__ __forceinline__ T enc_h(T v, const Enc& e) {
  if constexpr (sizeof(T) == 4) return v - e.href;
  else return v;
}
template <typename T>
__device__ __forceinline__ T dec_f(T v, int k, const Enc& e) {
  if constexpr (sizeof(T) == 4) return e.R[k] + v;
  else return v;
}
template <typename T>
__device__ __forceinline__ T enc_f(T v, int k, const Enc& e) {
  if constexpr (sizeof(T) == 4) return v - e.R[k];
  else return v;
}

__device__ __forceinline__ int wrapi(int a, int L) {
  a %= L;
  return a < 0 ? a + L : a;
}

inline dim3 grid2d(int Lx, int Ly) { return dim3((Lx + kBX - 1) / kBX, (Ly + kBY - 1) / kBY); }
inline dim3 grid1d(int64_t N) {
  int64_t b = (N + kBlock1D - 1) / kBlock1D;
  if (b > 65535LL * 32) b = 65535LL * 32;
  return dim3((unsigned)(b < 1 ? 1 : b));
}

// ---------------------------------------------------------------- equilibrium
// src/equilibrium.jl:63-116
template <typename T>
__global__ void k_equilibrium(T* __restrict__ feq, const T* __restrict__ h,
                              const T* __restrict__ ux, const T* __restrict__ uy,
                              T* __restrict__ vsq, T g, int64_t N, int64_t fs, Enc e) {
  const T g0 = T(1.5) * g;
  const T w1 = T(1.0 / 9.0), w5 = T(1.0 / 36.0);
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    const T hh = dec_h(h[x], e), u = ux[x], v = uy[x];
    const T usq = u * u + v * v;
    vsq[x] = usq;
    T f[kQ];
    f[0] = hh * ((T(1) - T(5.0 / 6.0) * g * hh) - T(2.0 / 3.0) * usq);
    f[1] = w1 * hh * (((g0 * hh + T(3) * u) + T(4.5) * (u * u)) - T(1.5) * usq);
    f[2] = w1 * hh * (((g0 * hh + T(3) * v) + T(4.5) * (v * v)) - T(1.5) * usq);
    f[3] = w1 * hh * (((g0 * hh - T(3) * u) + T(4.5) * (u * u)) - T(1.5) * usq);
    f[4] = w1 * hh * (((g0 * hh - T(3) * v) + T(4.5) * (v * v)) - T(1.5) * usq);
    T a = u + v;
    f[5] = w5 * hh * (((g0 * hh + T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
    a = v - u;
    f[6] = w5 * hh * (((g0 * hh + T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
    a = u + v;
    f[7] = w5 * hh * (((g0 * hh - T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
    a = u - v;
    f[8] = w5 * hh * (((g0 * hh + T(3) * a) + T(4.5) * (a * a)) - T(1.5) * usq);
#pragma unroll
    for (int k = 0; k < kQ; ++k) feq[k * fs + x] = enc_f(f[k], k, e);
  }
}

// ---------------------------------------------------------------- BGK + stream
// src/collide.jl:70-105: collision with the force term, pull stream
// f_k(x) <- fo_k(x - c_k), written to fdst (the caller makes fout and ftemp
// both refer to it, as `fout .= ftemp` does).
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

template <typename T>
__global__ void k_bgk_stream(T* __restrict__ fdst, const T* __restrict__ feq,
                             const T* __restrict__ ftemp, const T* __restrict__ Fx,
                             const T* __restrict__ Fy, T omeg, T itau, int Lx, int Ly, int64_t fs,
                             Enc e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= Lx || j >= Ly) return;
  const int64_t x = i + (int64_t)Lx * j;
#pragma unroll
  for (int k = 0; k < kQ; ++k) {
    const int is = wrapi(i - cx(k), Lx), js = wrapi(j - cy(k), Ly);
    const int64_t y = is + (int64_t)Lx * js;
    const T ft = dec_f(ftemp[k * fs + y], k, e);
    const T fe = dec_f(feq[k * fs + y], k, e);
    T fo = omeg * ft + itau * fe;
    // (a + b) - c*F is written as (a + b) + (-(c*F)); IEEE negation is exact
    if (k > 0) fo = fo + force_term<T>(k, Fx[y], Fy[y]);
    fdst[k * fs + x] = enc_f(fo, k, e);
  }
}

// ---------------------------------------------------------------- moments
// src/moments.jl:43-52: h = f0 + f1 + ... + f8 (sum!), u = Σ c f / h.
template <typename T>
__global__ void k_moments(T* __restrict__ h, T* __restrict__ ux, T* __restrict__ uy,
                          const T* __restrict__ f, int64_t N, int64_t fs, Enc e) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    T v[kQ];
#pragma unroll
    for (int k = 0; k < kQ; ++k) v[k] = dec_f(f[k * fs + x], k, e);
    T s = v[0];
#pragma unroll
    for (int k = 1; k < kQ; ++k) s = s + v[k];
    h[x] = enc_h(s, e);
    ux[x] = (((((v[1] - v[3]) + v[5]) - v[6]) - v[7]) + v[8]) / s;
    uy[x] = (((((v[2] - v[4]) + v[5]) + v[6]) - v[7]) - v[8]) / s;
  }
}

// ---------------------------------------------------------------- film pressure
// src/pressure.jl:72-155.  Neighbour names follow the reference: hip is
// circshift(h, (1,0)), i.e. h[i-1, j].
template <typename T>
__device__ __forceinline__ T lap9(T hip, T hjp, T him, T hjm, T hipjp, T himjp, T himjm, T hipjm,
                                  T hc) {
  return (T(2.0 / 3.0) * (((hjp + hip) + him) + hjm) +
          T(1.0 / 6.0) * (((hipjp + himjp) + himjm) + hipjm)) -
         T(10.0 / 3.0) * hc;
}

template <typename T>
__device__ __forceinline__ T power_broad(T a, int n) {
  T t = T(1);
  for (int i = 0; i < n; ++i) t *= a;
  return t;
}

template <typename T>
__device__ __forceinline__ T powlaw(T a, int n, int m, int mode) {
  if (mode == 1) {  // fast_93
    const T t = (a * a) * a;
    return ((t * t) * t) - t;
  }
  if (mode == 2) return ((a * a) * a) - (a * a);  // fast_32
  return power_broad(a, n) - power_broad(a, m);
}

template <typename T>
__global__ void k_filmpressure(T* __restrict__ p, const T* __restrict__ h, T gamma, T kap,
                               const T* __restrict__ theta, const T* __restrict__ kapf, int n,
                               int m, T hmin, T hcrit, int mode, int Lx, int Ly, Enc e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= Lx || j >= Ly) return;
  const int im = wrapi(i - 1, Lx), ip = wrapi(i + 1, Lx);
  const int jm = wrapi(j - 1, Ly), jp = wrapi(j + 1, Ly);
  auto H = [&](int a, int b) { return h[a + (int64_t)Lx * b]; };
  const int64_t x = i + (int64_t)Lx * j;
  // Laplacian of the stored values (the FP32 offset href cancels exactly).
  const T L = lap9(H(im, j), H(i, jm), H(ip, j), H(i, jp), H(im, jm), H(ip, jm), H(ip, jp),
                   H(im, jp), H(i, j));
  const T hh = dec_h(H(i, j), e);
  T k = kap;
  if (kapf) k = kapf[x];
  else if (theta) k = (T)kappa((double)theta[x], n, m, (double)hmin);
  const T a = hmin / (hh + hcrit);
  const T p0 = -gamma * (k * powlaw(a, n, m, mode));
  p[x] = p0 - gamma * L;
}

// ---------------------------------------------------------------- gradients
// ∇f! (src/differences.jl:153-206) and h∇p! (src/forcing.jl:168-187):
// out = a * (-1/3 (f[i-1]-f[i+1]) - 1/12 (...)); `a` is a field or absent.
template <typename T>
__global__ void k_grad(T* __restrict__ ox, T* __restrict__ oy, const T* __restrict__ f,
                       const T* __restrict__ a, const T* __restrict__ hh, int Lx, int Ly,
                       Enc e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= Lx || j >= Ly) return;
  const int im = wrapi(i - 1, Lx), ip = wrapi(i + 1, Lx);
  const int jm = wrapi(j - 1, Ly), jp = wrapi(j + 1, Ly);
  auto F = [&](int u, int v) { return f[u + (int64_t)Lx * v]; };
  const T fip = F(im, j), fim = F(ip, j), fjp = F(i, jm), fjm = F(i, jp);
  const T fipjp = F(im, jm), fimjp = F(ip, jm), fimjm = F(ip, jp), fipjm = F(im, jp);
  const T gx = T(-1.0 / 3.0) * (fip - fim) - T(1.0 / 12.0) * (((fipjp - fimjp) - fimjm) + fipjm);
  const T gy = T(-1.0 / 3.0) * (fjp - fjm) - T(1.0 / 12.0) * (((fipjp + fimjp) - fimjm) - fipjm);
  const int64_t x = i + (int64_t)Lx * j;
  if (hh) {  // h∇p!: height * gradient
    const T hv = dec_h(hh[x], e);
    ox[x] = hv * gx;
    oy[x] = hv * gy;
  } else if (a) {
    ox[x] = a[x] * gx;
    oy[x] = a[x] * gy;
  } else {
    ox[x] = gx;
    oy[x] = gy;
  }
}

// ∇²f!(output, f, γ): src/differences.jl:57-75
template <typename T>
__global__ void k_laplacian(T* __restrict__ out, const T* __restrict__ f, T gamma, int Lx,
                            int Ly) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= Lx || j >= Ly) return;
  const int im = wrapi(i - 1, Lx), ip = wrapi(i + 1, Lx);
  const int jm = wrapi(j - 1, Ly), jp = wrapi(j + 1, Ly);
  auto F = [&](int u, int v) { return f[u + (int64_t)Lx * v]; };
  out[i + (int64_t)Lx * j] =
      gamma * lap9(F(im, j), F(i, jm), F(ip, j), F(i, jp), F(im, jm), F(ip, jm), F(ip, jp),
                   F(im, jp), F(i, j));
}

// ---------------------------------------------------------------- slip, noise, body force
// slippage!: src/forcing.jl:42-46
template <typename T>
__global__ void k_slippage(T* __restrict__ sx, T* __restrict__ sy, const T* __restrict__ h,
                           const T* __restrict__ ux, const T* __restrict__ uy, T c6mu, T c6d,
                           T c3dd, int64_t N, Enc e) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    const T hh = dec_h(h[x], e);
    const T den = (T(2) * (hh * hh) + c6d * hh) + c3dd;
    sx[x] = (c6mu * hh * ux[x]) / den;
    sy[x] = (c6mu * hh * uy[x]) / den;
  }
}

// thermal!: src/forcing.jl:297-311 with counter-based normals
template <typename T>
__global__ void k_thermal(T* __restrict__ kx, T* __restrict__ ky, const T* __restrict__ h,
                          T kb26, T delta, T d3dd, uint64_t seed, uint64_t step, int64_t N,
                          int Lx, Enc e) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    double z0, z1;
    normal_pair(seed, step, noise_index((int)(x % Lx), (int)(x / Lx), Lx), kTagThermal, &z0, &z1);
    const T hh = dec_h(h[x], e);
    const T amp = sqrt(kb26 * hh / (((T(2) * hh) * hh + (T(6) * hh) * delta) + d3dd));
    kx[x] = (T)z0 * amp;
    ky[x] = (T)z1 * amp;
  }
}

// inclination!: src/forcing.jl:363-368
template <typename T>
__global__ void k_inclination(T* __restrict__ Fx, T* __restrict__ Fy, const T* __restrict__ h,
                              T ax, T ay, T fac, int64_t N, Enc e) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    const T hh = dec_h(h[x], e);
    Fx[x] = Fx[x] + (hh * ax) * fac;
    Fy[x] = Fy[x] + (hh * ay) * fac;
  }
}

// F = -h∇p - slip [- kbt]: src/simulate.jl:18-19
template <typename T>
__global__ void k_force_sum(T* __restrict__ Fx, T* __restrict__ Fy, const T* __restrict__ hx,
                            const T* __restrict__ hy, const T* __restrict__ sx,
                            const T* __restrict__ sy, const T* __restrict__ kx,
                            const T* __restrict__ ky, int64_t N) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    T a = -hx[x] - sx[x];
    T b = -hy[x] - sy[x];
    if (kx) {
      a = a - kx[x];
      b = b - ky[x];
    }
    Fx[x] = a;
    Fy[x] = b;
  }
}

template <typename T>
__global__ void k_kappa(T* __restrict__ kap, const T* __restrict__ theta, int n, int m,
                        double hmin, int64_t N) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x)
    kap[x] = (T)kappa((double)theta[x], n, m, hmin);
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, int64_t N) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x)
    p[x] = v;
}

template <typename T>
__global__ void k_init_random(T* __restrict__ h, int dist, uint64_t seed, double base,
                              double amp, double href, int64_t N, int Lx) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    double v;
    if (dist == 0) {
      v = base + amp * uniform(seed, 0, (uint32_t)x, kTagInitUniform);
    } else {
      double z0, z1;
      normal_pair(seed, 0, noise_index((int)(x % Lx), (int)(x / Lx), Lx), kTagInitNormal, &z0,
                  &z1);
      v = base + amp * z0;
    }
    h[x] = (T)(v - href);
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
template <typename T>
void launch_equilibrium(T* feq, const T* h, const T* ux, const T* uy, T* vsq, double g, int Lx,
                        int Ly, int64_t fs, Enc enc, hipStream_t s) {
  const int64_t N = (int64_t)Lx * Ly;
  hipLaunchKernelGGL(k_equilibrium<T>, grid1d(N), dim3(kBlock1D), 0, s, feq, h, ux, uy, vsq,
                     (T)g, N, fs, enc);
}
template <typename T>
void launch_bgk_and_stream(T* fdst, const T* feq, const T* ftemp, const T* Fx, const T* Fy,
                           double tau, int Lx, int Ly, int64_t fs, Enc enc, hipStream_t s) {
  const double omeg = 1.0 - 1.0 / tau, itau = 1.0 / tau;
  hipLaunchKernelGGL(k_bgk_stream<T>, grid2d(Lx, Ly), dim3(kBX, kBY), 0, s, fdst, feq, ftemp, Fx,
                     Fy, (T)omeg, (T)itau, Lx, Ly, fs, enc);
}
template <typename T>
void launch_moments(T* h, T* ux, T* uy, const T* f, int Lx, int Ly, int64_t fs, Enc enc,
                    hipStream_t s) {
  const int64_t N = (int64_t)Lx * Ly;
  hipLaunchKernelGGL(k_moments<T>, grid1d(N), dim3(kBlock1D), 0, s, h, ux, uy, f, N, fs, enc);
}
template <typename T>
void launch_filmpressure(T* p, const T* h, double gamma, double kappa_scalar, const T* theta,
                         const T* kapf, int n, int m, double hmin, double hcrit, int powmode,
                         int Lx, int Ly, Enc enc, hipStream_t s) {
  hipLaunchKernelGGL(k_filmpressure<T>, grid2d(Lx, Ly), dim3(kBX, kBY), 0, s, p, h, (T)gamma,
                     (T)kappa_scalar, theta, kapf, n, m, (T)hmin, (T)hcrit, powmode, Lx, Ly, enc);
}
template <typename T>
void launch_kappa_field(T* kap, const T* theta, int n, int m, double hmin, int64_t N,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_kappa<T>, grid1d(N), dim3(kBlock1D), 0, s, kap, theta, n, m, hmin, N);
}
template <typename T>
void launch_grad(T* ox, T* oy, const T* f, const T* a, const T* h, int Lx, int Ly, Enc enc,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_grad<T>, grid2d(Lx, Ly), dim3(kBX, kBY), 0, s, ox, oy, f, a, h, Lx, Ly,
                     enc);
}
template <typename T>
void launch_laplacian(T* out, const T* f, double gamma, int Lx, int Ly, hipStream_t s) {
  hipLaunchKernelGGL(k_laplacian<T>, grid2d(Lx, Ly), dim3(kBX, kBY), 0, s, out, f, (T)gamma, Lx,
                     Ly);
}
template <typename T>
void launch_slippage(T* sx, T* sy, const T* h, const T* ux, const T* uy, double delta, double mu,
                     int Lx, int Ly, Enc enc, hipStream_t s) {
  const int64_t N = (int64_t)Lx * Ly;
  hipLaunchKernelGGL(k_slippage<T>, grid1d(N), dim3(kBlock1D), 0, s, sx, sy, h, ux, uy,
                     (T)(6.0 * mu), (T)(6.0 * delta), (T)(3.0 * (delta * delta)), N, enc);
}
template <typename T>
void launch_thermal(T* kx, T* ky, const T* h, double kbt, double mu, double delta, uint64_t seed,
                    uint64_t step, int Lx, int Ly, Enc enc, hipStream_t s) {
  const int64_t N = (int64_t)Lx * Ly;
  hipLaunchKernelGGL(k_thermal<T>, grid1d(N), dim3(kBlock1D), 0, s, kx, ky, h,
                     (T)(((2.0 * kbt) * mu) * 6.0), (T)delta, (T)((3.0 * delta) * delta), seed,
                     step, N, Lx, enc);
}
template <typename T>
void launch_inclination(T* Fx, T* Fy, const T* h, double ax, double ay, double fac, int Lx,
                        int Ly, Enc enc, hipStream_t s) {
  const int64_t N = (int64_t)Lx * Ly;
  hipLaunchKernelGGL(k_inclination<T>, grid1d(N), dim3(kBlock1D), 0, s, Fx, Fy, h, (T)ax, (T)ay,
                     (T)fac, N, enc);
}
template <typename T>
void launch_force_sum(T* Fx, T* Fy, const T* hgpx, const T* hgpy, const T* sx, const T* sy,
                      const T* kx, const T* ky, int64_t N, hipStream_t s) {
  hipLaunchKernelGGL(k_force_sum<T>, grid1d(N), dim3(kBlock1D), 0, s, Fx, Fy, hgpx, hgpy, sx, sy,
                     kx, ky, N);
}
template <typename T>
void launch_fill(T* p, double v, int64_t N, hipStream_t s) {
  hipLaunchKernelGGL(k_fill<T>, grid1d(N), dim3(kBlock1D), 0, s, p, (T)v, N);
}
template <typename T>
void launch_init_random(T* h, int dist, uint64_t seed, double base, double amp, double href,
                        int Lx, int Ly, hipStream_t s) {
  const int64_t N = (int64_t)Lx * Ly;
  hipLaunchKernelGGL(k_init_random<T>, grid1d(N), dim3(kBlock1D), 0, s, h, dist, seed, base, amp,
                     href, N, Lx);
}

#define SW_INSTANTIATE(T)                                                                        \
  template void launch_equilibrium<T>(T*, const T*, const T*, const T*, T*, double, int, int,   \
                                      int64_t, Enc, hipStream_t);                                \
  template void launch_bgk_and_stream<T>(T*, const T*, const T*, const T*, const T*, double, int, \
                                         int, int64_t, Enc, hipStream_t);                        \
  template void launch_moments<T>(T*, T*, T*, const T*, int, int, int64_t, Enc, hipStream_t);    \
  template void launch_filmpressure<T>(T*, const T*, double, double, const T*, const T*, int, int, \
                                       double, double, int, int, int, Enc, hipStream_t);         \
  template void launch_kappa_field<T>(T*, const T*, int, int, double, int64_t, hipStream_t);      \
  template void launch_grad<T>(T*, T*, const T*, const T*, const T*, int, int, Enc, hipStream_t); \
  template void launch_laplacian<T>(T*, const T*, double, int, int, hipStream_t);                 \
  template void launch_slippage<T>(T*, T*, const T*, const T*, const T*, double, double, int, int, \
                                   Enc, hipStream_t);                                            \
  template void launch_thermal<T>(T*, T*, const T*, double, double, double, uint64_t, uint64_t,   \
                                  int, int, Enc, hipStream_t);                                   \
  template void launch_inclination<T>(T*, T*, const T*, double, double, double, int, int, Enc,    \
                                      hipStream_t);                                              \
  template void launch_force_sum<T>(T*, T*, const T*, const T*, const T*, const T*, const T*,     \
                                    const T*, int64_t, hipStream_t);                             \
  template void launch_fill<T>(T*, double, int64_t, hipStream_t);                                \
  template void launch_init_random<T>(T*, int, uint64_t, double, double, double, int, int,        \
                                      hipStream_t);

SW_INSTANTIATE(double)
SW_INSTANTIATE(float)

}  // namespace swalbe
