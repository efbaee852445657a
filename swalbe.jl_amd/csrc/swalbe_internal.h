// swalbe_internal.h -- device-side state layout and the launchers shared by
// the translation units of libswalbe_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <string>

#include "swalbe_hip.h"
#include "swalbe_math.h"

namespace swalbe {

// The tile a workgroup of a single-film step takes (tiles row-major, ntx
// per row).  The workgroups in flight (the part holds about 1280) then
// cover one narrow band of tile rows: their loads and stores fall on few
// DRAM pages, and vertical neighbours are in flight together.  Consecutive
// blocks go round the 8 XCDs, so where ntx is a multiple of 8 the blocks of
// XCD x take the x-th eighth of every row (ntx / 8 adjacent columns):
// horizontal neighbours share that XCD's L2 (their halo columns) except at
// the eighths' edges.  Odd steps walk backwards, so a step starts on the
// tiles whose inputs the previous step wrote last, still in L2 and the
// Infinity Cache.  Tiles are independent: the order changes no result.
__device__ __forceinline__ int tile_of_block(int ntx, int ntiles, uint64_t step) {
  const int b = (int)blockIdx.x;
  const bool back = (step & 1) != 0;
  if ((ntx & 7) != 0) return back ? ntiles - 1 - b : b;
  const int x = b & 7, J = ntiles >> 3;
  const int j = back ? J - 1 - (b >> 3) : (b >> 3);
  const int cpr = ntx >> 3;  // columns of each XCD per row of tiles
  const int row = j / cpr, k = j - row * cpr;
  return row * ntx + x * cpr + k;
}

// Compute units of the current device, asked once per device (256 if the
// runtime cannot say).  Threads that race on the first call store the same
// value; the atomics keep that race defined.
inline int device_cus() {
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cus[dev].load(std::memory_order_relaxed);
  if (c == 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        c <= 0)
      c = 256;
    cus[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

// The FP64 step's order for a film whose inputs (h, ux, uy) fit the eight
// L2s (4 MB each, so in_bytes <= 32 MB) and whose tile count is a multiple
// of 8: XCD x takes the x-th eighth of the tile list, a band of rows, and
// reads the same band step after step from its own L2.  Odd steps walk each
// band backwards.  Other films take tile_of_block.
__device__ __forceinline__ int tile_of_block_f64(int ntx, int ntiles, uint64_t step,
                                                 int64_t in_bytes) {
  if ((ntiles & 7) != 0 || in_bytes > ((int64_t)32 << 20))
    return tile_of_block(ntx, ntiles, step);
  const int b = (int)blockIdx.x, share = ntiles >> 3, j = b >> 3;
  return (b & 7) * share + ((step & 1) != 0 ? share - 1 - j : j);
}

// FP32 states store height and distributions as deviations from a rest
// state of reference height href: dh = h - href and f'_k = f_k - R_k with
// R = feq(href, u = 0).  FP64 states use href = 0 and R = 0 (plain values).
struct Enc {
  float href;
  float R[kQ];
};

Enc make_enc(double href, double g);
Enc zero_enc();

// Layout of a device-resident film: Julia column-major fields, x fastest.
struct Film {
  int Lx = 0, Ly = 0;
  int dtype = SWALBE_F64;
  size_t esz = 8;
  int64_t N = 0;
  int64_t fstride = 0;                // elements between distribution planes (>= N)
  double href = 0.0, g_enc = 0.0;  // encoding parameters (FP32 only)
  Enc enc{};
  void* f[2] = {nullptr, nullptr};    // distribution ping-pong buffers
  int fout_buf = 0, ftemp_buf = 1;    // which buffer is fout / ftemp
  void* hbuf[2] = {nullptr, nullptr};
  void* uxbuf[2] = {nullptr, nullptr};
  void* uybuf[2] = {nullptr, nullptr};
  int hu_cur = 0;                     // which h/ux/uy buffer holds the current moments
  void* feq = nullptr;
  void* vsq = nullptr;
  void* pressure = nullptr;
  void* Fx = nullptr;
  void* Fy = nullptr;
  void* slipx = nullptr;
  void* slipy = nullptr;
  void* hgpx = nullptr;
  void* hgpy = nullptr;
  void* kbtx = nullptr;
  void* kbty = nullptr;
  void* theta = nullptr;              // θ field, allocated on first upload
  void* kappa = nullptr;              // κ(θ) field derived from theta
  bool kappa_valid = false;
  int kappa_n = 0, kappa_m = 0;
  double kappa_hmin = 0.0;
  void* height() const { return hbuf[hu_cur]; }
  void* velx() const { return uxbuf[hu_cur]; }
  void* vely() const { return uybuf[hu_cur]; }
};

// ---------------------------------------------------------------- per-op launchers
// All arrays are device pointers of element type T (double or float); `enc`
// decodes FP32 deviation storage (zero_enc() for plain arrays).
// Distribution arrays hold plane k at offset k*fs (fs = Lx*Ly for caller arrays).
template <typename T>
void launch_equilibrium(T* feq, const T* h, const T* ux, const T* uy, T* vsq, double g, int Lx,
                        int Ly, int64_t fs, Enc enc, hipStream_t s);
template <typename T>
void launch_bgk_and_stream(T* fdst, const T* feq, const T* ftemp, const T* Fx, const T* Fy,
                           double tau, int Lx, int Ly, int64_t fs, Enc enc, hipStream_t s);
template <typename T>
void launch_moments(T* h, T* ux, T* uy, const T* f, int Lx, int Ly, int64_t fs, Enc enc,
                    hipStream_t s);
template <typename T>
void launch_filmpressure(T* p, const T* h, double gamma, double kappa_scalar, const T* theta,
                         const T* kapf, int n, int m, double hmin, double hcrit, int powmode,
                         int Lx, int Ly, Enc enc, hipStream_t s);
template <typename T>
void launch_grad(T* ox, T* oy, const T* f, const T* a, const T* h, int Lx, int Ly, Enc enc,
                 hipStream_t s);
template <typename T>
void launch_laplacian(T* out, const T* f, double gamma, int Lx, int Ly, hipStream_t s);
template <typename T>
void launch_slippage(T* sx, T* sy, const T* h, const T* ux, const T* uy, double delta, double mu,
                     int Lx, int Ly, Enc enc, hipStream_t s);
template <typename T>
void launch_thermal(T* kx, T* ky, const T* h, double kbt, double mu, double delta, uint64_t seed,
                    uint64_t step, int Lx, int Ly, Enc enc, hipStream_t s);
template <typename T>
void launch_inclination(T* Fx, T* Fy, const T* h, double ax, double ay, double fac, int Lx,
                        int Ly, Enc enc, hipStream_t s);
template <typename T>
void launch_force_sum(T* Fx, T* Fy, const T* hgpx, const T* hgpy, const T* sx, const T* sy,
                      const T* kx, const T* ky, int64_t N, hipStream_t s);
template <typename T>
void launch_fill(T* p, double v, int64_t N, hipStream_t s);
template <typename T>
void launch_init_random(T* h, int dist, uint64_t seed, double base, double amp, double href,
                        int Lx, int Ly, hipStream_t s);
// Initial height of a film from a swalbe_film_init spec (stored as h - href).
template <typename T>
void launch_film_init(T* h, int Lx, int Ly, const swalbe_film_init& in, double href,
                      hipStream_t s);
// κ(θ) per node from a θ field (same arithmetic as the scalar κ).
template <typename T>
void launch_kappa_field(T* kap, const T* theta, int n, int m, double hmin, int64_t N,
                        hipStream_t s);

// ---------------------------------------------------------------- fused step
struct StepParams {
  double gamma, kappa, hmin, hcrit;
  int n, m, powmode;          // powmode 0: power_broad, 1: fast_93, 2: fast_32
  double omeg, itau, g;
  double c6mu, c6d, c3dd;     // slip constants 6μ, 6δ, 3δ²
  int thermal;
  double kb26, delta, d3dd;   // ((2 kbt) μ) 6, δ, (3δ)δ
  uint64_t seed;
  int body;
  double bax, bay, bfac;
  int use_theta;
  int read_ft;
  int materialize;
  double wet_thresh;
  int fp64_guard;  // FP32 kernels: disjoining pressure in FP64 where h + hcrit < 4 hmin
};

StepParams make_step_params(const swalbe_params* p, const swalbe_step_opts* o, int powmode);

struct StepSeries {
  unsigned long long* dh_keys;  // [2] ordered keys of max / min h, or null
  unsigned long long* wet;      // [1] count of h > wet_thresh, or null
};

template <typename T>
struct StepIO {
  const T* h_in;
  const T* ux_in;
  const T* uy_in;
  const T* ft;      // f_t (ftemp); read when read_ft
  T* f_out;
  T* h_out;
  T* ux_out;
  T* uy_out;
  T* p_out;
  const T* kappa;   // κ(θ) field (θ-field runs), or null for the scalar κ
};

template <typename T>
void launch_step(const StepIO<T>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                 int Lx, int Ly, int64_t fstride, Enc enc, int tile_y, int nts, hipStream_t s);
// FP32 films: variant 0/1 collide at the source node, 2 at the destination.
void launch_step_f32(const StepIO<float>& io, const StepParams& sp, StepSeries ser, uint64_t step,
                     int Lx, int Ly, int64_t fstride, Enc enc, int tile_y, int variant, int nts,
                     hipStream_t s);

// Deterministic FP64 reductions of a field: sum, max, min, count(> thresh).
template <typename T>
void launch_reduce(const T* h, int64_t N, double href, double thresh, double* partials,
                   int nparts, hipStream_t s);
void launch_reduce_finish(const double* partials, int nparts, double* out, hipStream_t s);

// On-device measurements of a height field (href added to every value):
// out = {count(h > thresh), max h, Σ sqrt(|∇h|² + 1) over h > thresh,
// reduced energy}; surface = 0 skips the gradient sums; mask (N bytes) gets
// h > thresh when non-null.  partials holds 4 doubles per block (nparts).
template <typename T>
void launch_measure(const T* h, double href, int Lx, int Ly, double thresh, int surface,
                    double cos_s, const T* theta, uint8_t* mask, double* partials, int nparts,
                    double* out, hipStream_t s);

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Ordered 64-bit key of a double (monotone in the value) and back.
SW_HD unsigned long long dkey(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
SW_HD double dkey_inv(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __builtin_bit_cast(double, b);
}

}  // namespace swalbe
