// swalbe_batch.h -- FP32 step constants and the device descriptors of a
// batch of independent films (shared by step_f32.hip and batch.hip).
#pragma once

#include "swalbe_internal.h"

namespace swalbe {

// FP32 constants of one film's step.
struct F32K {
  float gamma, kap, hmin, hcrit, c6mu, c6d, c3dd, kb26, omeg, itau, c56g, g0, href;
  float bfx, bfy;  // body force α·fac
  int n, m, powmode;
  // FP64 disjoining pressure where the film nears hmin: nodes with
  // h < hguard (= 4 hmin - hcrit) when guard is set
  int guard;
  float hguard;
  double gamma_d, kap_d, hmin_d, hcrit_d, href_d;
};

F32K make_f32k(const StepParams& sp, const Enc& enc);
int film_is_small(int Lx, int Ly, int tile_x, int tile_y);

// Rows per tile of the batched kernel: every film uses 8 or 16 (its own
// choice, by default 16); the kernel's LDS holds the taller tile.
constexpr int kBatchTileY = 16;
// Default rows per segment of the strip kernel; films the strip kernel
// treats with modulo wrapping (narrower than a strip or very short).
constexpr int kStripSeg = 64;
constexpr int kStripOut = 58;  // output columns per strip
int strip_is_small(int Lx, int Ly);
// Batches step their films on strips (true) or 64 x kBatchTileY LDS tiles.
// A film alone gives the same bits with the matching single-film variant
// (3, or 2 with tile_y = kBatchTileY).
constexpr bool kBatchStrips = false;
constexpr int kBatchVariant = kBatchStrips ? 3 : 2;

// One resident film of a batch (device memory, read with scalar loads).
struct BatchFilm {
  float* h[2];
  float* ux[2];
  float* uy[2];
  float* p;
  float* f[2];  // f[1] only when τ ≠ 1
  // ntx, ntiles: strips and (strip, segment) items, or tile columns and tiles
  int Lx, Ly, ntx, ntiles, small, pairs, read_ft, thermal, body;
  int last_only;  // τ = 1: distributions and pressure written on the last step only
  int ty;         // rows per tile of this film (8 or 16)
  int64_t fstride;
  uint64_t seed;
  int64_t t0;                // step number before admission
  unsigned long long admit;  // launch counter value at admission
  unsigned long long end;    // launch counter value at which the film stops
  F32K K;
};

struct BatchHdr {
  int nfilms;
  int total_tiles;
};

// One batched step (all resident films) followed by the counter increment.
// any_ft: some resident film has τ ≠ 1 (reads f_t); the kernel without that
// path needs fewer registers and holds one more workgroup per CU.
void launch_batch_f32(const BatchFilm* films, const int* tstart, const BatchHdr* hdr,
                      unsigned long long* counter, int grid, int nts, bool any_ft,
                      hipStream_t s);
int batch_blocks_per_cu(bool any_ft);

}  // namespace swalbe
