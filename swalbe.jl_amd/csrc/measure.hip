// measure.hip -- on-device measurements of a 2D film (src/measures.jl):
// wetted! (count(h > thresh), maximum(h)), fluid_dry! (the mask h > thresh)
// and surfacearea! (Σ sqrt(|∇h|² + 1) over the wet nodes with the ∇f_simple!
// stencil, and the reduced surface energy).  One kernel computes all of
// them; the fused time loop enqueues it before a step on the state's
// stream, so a measurement callback costs no host synchronisation.
//
// Reductions run in a fixed order (a fixed grid, shuffles, then one thread
// over the block partials), so repeated runs give the same bits.
#include <algorithm>
#include <mutex>

#include "swalbe_internal.h"

namespace swalbe {

// The block partials of swalbe_measure_ptr_noalloc: a device array of
// 4 x 512 doubles (16 KiB) that each GPU holds from the library's load until
// the process ends.  Every call on one GPU uses the same array.  In host code
// `scratch` names a host-side stand-in, never the device memory: the device
// address of each GPU's copy comes from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[4 * 512];

namespace {

constexpr int kMT = 256;
// Blocks of one measurement: at most kMaxParts, whatever the film size (the
// kernel strides over the film), so 4 * kMaxParts doubles hold every partial.
constexpr int kMaxParts = 512;
static_assert(sizeof(scratch) == sizeof(double) * 4 * kMaxParts, "scratch holds every partial");

int measure_parts(int64_t N) {
  return (int)std::min<int64_t>(kMaxParts, std::max<int64_t>(1, N / 4096));
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of the current GPU's scratch array, looked up once per
// device (with that device current) and kept.
int scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "swalbe_measure_ptr_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("swalbe_measure_ptr_noalloc: device index beyond the scratch table");
    return SWALBE_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                      "swalbe_measure_ptr_noalloc: hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return SWALBE_OK;
}

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double a = __shfl_xor(v, o);
    v = a > v ? a : v;
  }
  return v;
}

// partials[4b + 0..3] = {count, max, Σ surface, Σ cos πθ} of block b's nodes
template <typename T>
__global__ __launch_bounds__(kMT) void k_measure(const T* __restrict__ h, double href, int Lx,
                                                 int Ly, double thresh, int surface,
                                                 const T* __restrict__ theta,
                                                 uint8_t* __restrict__ mask,
                                                 double* __restrict__ partials) {
  __shared__ double red[4][kMT / 64];
  const int64_t N = (int64_t)Lx * Ly;
  double cnt = 0.0, mx = -__builtin_inf(), surf = 0.0, cs = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kMT + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * kMT) {
    const double hv = href + (double)h[i];
    const bool wet = hv > thresh;
    mx = hv > mx ? hv : mx;
    if (mask) mask[i] = wet ? 1 : 0;
    if (!wet) continue;
    cnt += 1.0;
    if (surface) {
      const int x = (int)(i % Lx), y = (int)(i / Lx);
      const int xm = x == 0 ? Lx - 1 : x - 1, xp = x == Lx - 1 ? 0 : x + 1;
      const int ym = y == 0 ? Ly - 1 : y - 1, yp = y == Ly - 1 ? 0 : y + 1;
      auto at = [&](int xx, int yy) { return href + (double)h[xx + (int64_t)Lx * yy]; };
      // ∇f_simple! (src/measures.jl:175-192): fip = f[i-1, j], fim = f[i+1, j], ...
      const double fip = at(xm, y), fim = at(xp, y), fjp = at(x, ym), fjm = at(x, yp);
      const double fipjp = at(xm, ym), fimjp = at(xp, ym), fimjm = at(xp, yp), fipjm = at(xm, yp);
      const double gx = ((-1.0 / 3.0) * (fip - fim)) -
                        ((1.0 / 12.0) * (((fipjp - fimjp) - fimjm) + fipjm));
      const double gy = ((-1.0 / 3.0) * (fjp - fjm)) -
                        ((1.0 / 12.0) * (((fipjp + fimjp) - fimjm) - fipjm));
      surf += sqrt((gx * gx + gy * gy) + 1.0);
      if (theta) cs += cospi((double)theta[i]);
    }
  }
  cnt = wsum(cnt);
  mx = wmax(mx);
  surf = wsum(surf);
  cs = wsum(cs);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = cnt;
    red[1][w] = mx;
    red[2][w] = surf;
    red[3][w] = cs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double c0 = red[0][0], m0 = red[1][0], s0 = red[2][0], k0 = red[3][0];
    for (int k = 1; k < kMT / 64; ++k) {
      c0 += red[0][k];
      m0 = red[1][k] > m0 ? red[1][k] : m0;
      s0 += red[2][k];
      k0 += red[3][k];
    }
    partials[4 * blockIdx.x + 0] = c0;
    partials[4 * blockIdx.x + 1] = m0;
    partials[4 * blockIdx.x + 2] = s0;
    partials[4 * blockIdx.x + 3] = k0;
  }
}

// out = {count, max, area, reduced energy}; the energy subtracts
// count · cos πθ (scalar θ) or Σ cos πθ over the wet nodes (θ field)
__global__ void k_measure_finish(const double* __restrict__ partials, int nparts, double cos_s,
                                 int theta_field, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double c = 0.0, m = -__builtin_inf(), s = 0.0, k = 0.0;
  for (int b = 0; b < nparts; ++b) {
    c += partials[4 * b + 0];
    m = partials[4 * b + 1] > m ? partials[4 * b + 1] : m;
    s += partials[4 * b + 2];
    k += partials[4 * b + 3];
  }
  out[0] = c;
  out[1] = m;
  out[2] = s;
  out[3] = s - (theta_field ? k : c * cos_s);
}

// ∇f_simple!(outx, outy, f, dgrad) on caller arrays (src/measures.jl:175-192)
template <typename T>
__global__ void k_grad_simple(T* __restrict__ gx, T* __restrict__ gy, const T* __restrict__ f,
                              int Lx, int Ly) {
  const int64_t N = (int64_t)Lx * Ly;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % Lx), y = (int)(i / Lx);
    const int xm = x == 0 ? Lx - 1 : x - 1, xp = x == Lx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? Ly - 1 : y - 1, yp = y == Ly - 1 ? 0 : y + 1;
    auto at = [&](int xx, int yy) { return f[xx + (int64_t)Lx * yy]; };
    const T fip = at(xm, y), fim = at(xp, y), fjp = at(x, ym), fjm = at(x, yp);
    const T fipjp = at(xm, ym), fimjp = at(xp, ym), fimjm = at(xp, yp), fipjm = at(xm, yp);
    gx[i] = ((T(-1.0) / T(3.0)) * (fip - fim)) -
            ((T(1.0) / T(12.0)) * (((fipjp - fimjp) - fimjm) + fipjm));
    gy[i] = ((T(-1.0) / T(3.0)) * (fjp - fjm)) -
            ((T(1.0) / T(12.0)) * (((fipjp + fimjp) - fimjm) - fipjm));
  }
}

}  // namespace

template <typename T>
void launch_measure(const T* h, double href, int Lx, int Ly, double thresh, int surface,
                    double cos_s, const T* theta, uint8_t* mask, double* partials, int nparts,
                    double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_measure<T>, dim3(nparts), dim3(kMT), 0, s, h, href, Lx, Ly, thresh,
                     surface, theta, mask, partials);
  hipLaunchKernelGGL(k_measure_finish, dim3(1), dim3(64), 0, s, partials, nparts, cos_s,
                     theta ? 1 : 0, out);
}
template void launch_measure<double>(const double*, double, int, int, double, int, double,
                                     const double*, uint8_t*, double*, int, double*, hipStream_t);
template void launch_measure<float>(const float*, double, int, int, double, int, double,
                                    const float*, uint8_t*, double*, int, double*, hipStream_t);

}  // namespace swalbe

using namespace swalbe;

extern "C" {

int swalbe_measure_ptr(int32_t dtype, const void* h, const void* theta_field, double theta,
                       int32_t Lx, int32_t Ly, double thresh, void* mask, double* out_dev,
                       void* stream) {
  if (!h || !out_dev || Lx < 1 || Ly < 1 || (dtype != SWALBE_F64 && dtype != SWALBE_F32)) {
    set_error("swalbe_measure_ptr: bad arguments");
    return SWALBE_EINVAL;
  }
  const int64_t N = (int64_t)Lx * Ly;
  const int nparts = measure_parts(N);
  double* partials = nullptr;
  hipError_t e = hipMalloc((void**)&partials, sizeof(double) * 4 * nparts);
  if (e != hipSuccess) return hip_fail(e, "swalbe_measure_ptr");
  const double cs = swalbe::cospi(theta);
  if (dtype == SWALBE_F64)
    launch_measure<double>((const double*)h, 0.0, Lx, Ly, thresh, 1, cs,
                           (const double*)theta_field, (uint8_t*)mask, partials, nparts, out_dev,
                           (hipStream_t)stream);
  else
    launch_measure<float>((const float*)h, 0.0, Lx, Ly, thresh, 1, cs, (const float*)theta_field,
                          (uint8_t*)mask, partials, nparts, out_dev, (hipStream_t)stream);
  e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize((hipStream_t)stream);
  (void)hipFree(partials);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return hip_fail(e, "swalbe_measure_ptr");
  return SWALBE_OK;
}

int swalbe_measure_ptr_noalloc(int32_t dtype, const void* h, const void* theta_field,
                               double theta, int32_t Lx, int32_t Ly, double thresh, void* mask,
                               double* out_dev, void* stream) {
  if (!h || !out_dev || Lx < 1 || Ly < 1 || (dtype != SWALBE_F64 && dtype != SWALBE_F32)) {
    set_error("swalbe_measure_ptr_noalloc: bad arguments");
    return SWALBE_EINVAL;
  }
  const int64_t N = (int64_t)Lx * Ly;
  const int nparts = measure_parts(N);
  if (nparts > kMaxParts) {  // cannot happen: the grid is capped at kMaxParts blocks
    set_error("swalbe_measure_ptr_noalloc: more block partials than the scratch array holds");
    return SWALBE_EINVAL;
  }
  double* partials = nullptr;
  const int rc = scratch_address(&partials);
  if (rc != SWALBE_OK) return rc;
  // every block writes its own 4 partials before the finishing kernel reads
  // them, so nothing left by an earlier call needs clearing
  const double cs = swalbe::cospi(theta);
  if (dtype == SWALBE_F64)
    launch_measure<double>((const double*)h, 0.0, Lx, Ly, thresh, 1, cs,
                           (const double*)theta_field, (uint8_t*)mask, partials, nparts, out_dev,
                           (hipStream_t)stream);
  else
    launch_measure<float>((const float*)h, 0.0, Lx, Ly, thresh, 1, cs, (const float*)theta_field,
                          (uint8_t*)mask, partials, nparts, out_dev, (hipStream_t)stream);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "swalbe_measure_ptr_noalloc");
  return SWALBE_OK;
}

int swalbe_grad_simple_ptr(int32_t dtype, void* outx, void* outy, const void* f, int32_t Lx,
                           int32_t Ly, void* stream) {
  if (!outx || !outy || !f || Lx < 1 || Ly < 1 || (dtype != SWALBE_F64 && dtype != SWALBE_F32)) {
    set_error("swalbe_grad_simple_ptr: bad arguments");
    return SWALBE_EINVAL;
  }
  const int64_t N = (int64_t)Lx * Ly;
  const int blocks = (int)std::min<int64_t>(65535, (N + 255) / 256);
  if (dtype == SWALBE_F64)
    hipLaunchKernelGGL(k_grad_simple<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (double*)outx, (double*)outy, (const double*)f, Lx, Ly);
  else
    hipLaunchKernelGGL(k_grad_simple<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (float*)outx, (float*)outy, (const float*)f, Lx, Ly);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "swalbe_grad_simple_ptr");
  return SWALBE_OK;
}

}  // extern "C"
