// probe.hip -- libswalbe_probe.so, a measurement tool of bench.py and not
// part of the library's interface: the data-movement ceiling of the fused
// step's access pattern.  A kernel reads `nread` planes and writes `nwrite`
// planes over the same 64 x TY tiles, in the order the step takes them
// (swalbe::tile_of_block), with no arithmetic,
// which calibrates how close the step runs to what its traffic pattern
// allows on the part.  Every store streams past the caches: that is the
// fastest way the part moves this pattern (with cached stores, or with three
// of thirteen planes cached, it moves about two thirds as much).
//
//   int swalbe_pattern_probe(int32_t Lx, int32_t Ly, int32_t nread, int32_t nwrite,
//                            int32_t tile_y, int32_t iters, double* ms_per_launch);
//
// returns 0, or 2 for bad arguments, 3 when the planes cannot be allocated,
// 4 for a HIP error; the mean device time per launch over `iters` launches.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "swalbe_internal.h"

namespace {

// NR / NW > 0: compile-time plane counts (all loads of a row issued at
// once); 0: the runtime counts nread / nwrite.
template <int TY, int NR, int NW>
__global__ __launch_bounds__(256) void k_pattern(const float* __restrict__ in,
                                                 float* __restrict__ out, int Lx, int Ly,
                                                 int ntx, int ntiles, int nread, int nwrite,
                                                 int64_t plane, int back) {
  const int tile = swalbe::tile_of_block(ntx, ntiles, (uint64_t)back);
  const int x0 = (tile % ntx) * 64, y0 = (tile / ntx) * TY;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gx = x0 + lane;
  const int nr = NR > 0 ? NR : nread, nw = NW > 0 ? NW : nwrite;
  float acc[TY / 4];
#pragma unroll
  for (int i = 0; i < TY / 4; ++i) {
    const int gy = y0 + wv + 4 * i;
    const int64_t o = (int64_t)gy * Lx + gx;
    float a = 0.f;
    if (gx < Lx && gy < Ly) {
#pragma unroll
      for (int r = 0; r < (NR > 0 ? NR : 16); ++r)
        if (r < nr) a += in[r * plane + o];
    }
    acc[i] = a;
  }
#pragma unroll
  for (int i = 0; i < TY / 4; ++i) {
    const int gy = y0 + wv + 4 * i;
    const int64_t o = (int64_t)gy * Lx + gx;
    if (gx < Lx && gy < Ly) {
#pragma unroll
      for (int w = 0; w < (NW > 0 ? NW : 16); ++w)
        if (w < nw) __builtin_nontemporal_store(acc[i] + (float)w, &out[w * plane + o]);
    }
  }
}

}  // namespace

extern "C" int swalbe_pattern_probe(int32_t Lx, int32_t Ly, int32_t nread, int32_t nwrite,
                                    int32_t tile_y, int32_t iters, double* ms_per_launch) {
  if (Lx < 1 || Ly < 1 || nread < 0 || nread > 16 || nwrite < 1 || nwrite > 16 || iters < 1 ||
      (tile_y != 8 && tile_y != 16) || !ms_per_launch)
    return 2;
  const int64_t plane = (int64_t)Lx * Ly;
  float *in = nullptr, *out = nullptr;
  if (hipMalloc(&in, sizeof(float) * plane * (nread > 0 ? nread : 1)) != hipSuccess) return 3;
  if (hipMalloc(&out, sizeof(float) * plane * nwrite) != hipSuccess) {
    (void)hipFree(in);
    return 3;
  }
  (void)hipMemset(in, 0, sizeof(float) * plane * (nread > 0 ? nread : 1));
  const int ntx = (Lx + 63) / 64, ntiles = ntx * ((Ly + tile_y - 1) / tile_y);
  const bool step = nread == 3 && nwrite == 13;  // the fused step's own counts
  int parity = 0;
  auto launch = [&]() {
    const int back = parity;
    parity ^= 1;
#define SW_P(TY, R, W)                                                                      \
  hipLaunchKernelGGL((k_pattern<TY, R, W>), dim3(ntiles), dim3(256), 0, 0, in, out, Lx, Ly, \
                     ntx, ntiles, nread, nwrite, plane, back)
    if (tile_y == 8) {
      if (step) SW_P(8, 3, 13);
      else SW_P(8, 0, 0);
    } else {
      if (step) SW_P(16, 3, 13);
      else SW_P(16, 0, 0);
    }
#undef SW_P
  };
  for (int i = 0; i < 3; ++i) launch();
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, 0);
  for (int i = 0; i < iters; ++i) launch();
  (void)hipEventRecord(e1, 0);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(in);
  (void)hipFree(out);
  if (e != hipSuccess) return 4;
  *ms_per_launch = ms / iters;
  return 0;
}
