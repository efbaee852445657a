// batch.hip -- batches of independent FP32 films stepped together: one
// launch per time step covers every resident film (any mix of sizes and
// parameters), films are admitted and retired between launches, and a run of
// steps can be replayed as a captured HIP graph.
//
// Memory: one device arena per batch; each film takes h, ux, uy (two
// buffers each), p and one distribution buffer (two when τ ≠ 1), carved
// first-fit.  A film's parity and step number are derived on the device
// from a launch counter, so admissions only rewrite the film table.
#include <string.h>

#include <algorithm>
#include <deque>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "swalbe_batch.h"

using namespace swalbe;

namespace {

constexpr size_t kAlign = 256;
constexpr int kRing = 4;

struct Slot {
  bool used = false;
  BatchFilm d{};
  size_t off = 0, bytes = 0;
  Enc enc{};
  int64_t nsteps = 0;
};

template <typename T>
__global__ void k_film_init(T* __restrict__ h, int Lx, int Ly, swalbe_film_init in, double cth1,
                            double cth2, double href) {
  const int64_t N = (int64_t)Lx * Ly;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < N;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(x % Lx), j = (int)(x / Lx);
    double v = in.base;
    if (in.kind == SWALBE_INIT_NORMAL) {
      double z0, z1;
      normal_pair(in.seed, 0, noise_index(i, j, Lx), kTagInitNormal, &z0, &z1);
      v = in.base + in.amp * z0;
    } else if (in.kind == SWALBE_INIT_CAP || in.kind == SWALBE_INIT_TWO_CAPS) {
      // spherical cap (cos(asin(r/R)) - cos πθ) R inside r <= R, else the
      // precursor (src/initialvalues.jl:203-224); two caps add and remove
      // one precursor (src/initialvalues.jl:277-323); 1-based centres
      auto cap = [&](double R, double cth, double cx_, double cy_) {
        const double dx = (i + 1) - cx_, dy = (j + 1) - cy_;
        const double circ = sqrt(dx * dx + dy * dy);
        double hv = in.precursor;
        if (circ <= R) {
          const double s = circ / R;
          hv = (sqrt(1.0 - s * s) - cth) * R;
        }
        return hv < 0.0 ? in.precursor : hv;
      };
      v = cap(in.r1, cth1, in.c1x, in.c1y);
      if (in.kind == SWALBE_INIT_TWO_CAPS) {
        v = v + cap(in.r2, cth2, in.c2x, in.c2y) - in.precursor;
        if (v < 0.0) v = in.precursor;
      }
      if (in.clamp && v < in.precursor) v = in.precursor;
    }
    h[x] = (T)(v - href);
  }
}

inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

}  // namespace

namespace swalbe {
template <typename T>
void launch_film_init(T* h, int Lx, int Ly, const swalbe_film_init& in, double href,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_film_init<T>, dim3(256), dim3(256), 0, s, h, Lx, Ly, in, cospi(in.theta1),
                     cospi(in.theta2), href);
}
template void launch_film_init<float>(float*, int, int, const swalbe_film_init&, double,
                                      hipStream_t);
template void launch_film_init<double>(double*, int, int, const swalbe_film_init&, double,
                                       hipStream_t);
}  // namespace swalbe

struct swalbe_batch {
  int device = 0;
  hipStream_t stream = nullptr;     // every step, init and table update, in order
  hipStream_t dl_stream = nullptr;  // non-blocking: downloads of finished films
  char* arena = nullptr;
  size_t capacity = 0, used_bytes = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  std::vector<Slot> slots;
  // device film table: [BatchHdr | tstart[max_films] | BatchFilm[max_films]]
  char* d_table = nullptr;
  size_t off_tstart = 0, off_films = 0, table_bytes = 0;
  unsigned long long* d_counter = nullptr;
  // pinned staging copies of the table, reused round-robin once their
  // stream-ordered upload has completed
  char* staging[kRing] = {};
  hipEvent_t staging_ev[kRing] = {};
  bool staging_busy[kRing] = {};
  int staging_next = 0;
  // one event per enqueue call, keyed by the launch counter after it
  std::deque<std::pair<unsigned long long, hipEvent_t>> steps_ev;
  std::vector<hipEvent_t> ev_pool;
  unsigned long long counter = 0;  // launches enqueued so far
  bool dirty = true;
  int nactive = 0;
  int grid_any = 0;        // workgroups of the kernel with the τ ≠ 1 path
  int grid_tau1 = 0;       // workgroups of the kernel for films that all have τ = 1
  bool any_ft = false;     // a film in the uploaded table has τ ≠ 1
  bool graph_any = false;  // the kernel the captured graph launches
  int nts = 2;  // f and p stream past the caches: 3 % faster than cached for batches
  int last_only = 0;  // films admitted with distributions / pressure written at their end only
  int autotune = 0;   // films admitted with the tile height measured fastest for their shape
  int forced_ty = 0;  // films admitted with this tile height (0: default or autotuned)
  int graph_steps = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  const BatchFilm* films() const { return (const BatchFilm*)(d_table + off_films); }
  const int* tstart() const { return (const int*)(d_table + off_tstart); }
  const BatchHdr* hdr() const { return (const BatchHdr*)d_table; }
};

namespace {

#define SW_HIP(call)                                  \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

int bad(const char* msg) {
  set_error(msg);
  return SWALBE_EINVAL;
}

bool arena_alloc(swalbe_batch* b, size_t bytes, size_t* off) {
  for (auto it = b->free_.begin(); it != b->free_.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t noff = it->first + bytes;
      b->free_.erase(it);
      if (rest) b->free_[noff] = rest;
      b->used_bytes += bytes;
      return true;
    }
  }
  return false;
}

void arena_free(swalbe_batch* b, size_t off, size_t bytes) {
  b->used_bytes -= bytes;
  auto it = b->free_.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != b->free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    b->free_.erase(nx);
  }
  if (it != b->free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      b->free_.erase(it);
    }
  }
}

size_t film_bytes(int Lx, int Ly, bool two_f) {
  const size_t N = (size_t)Lx * Ly;
  const size_t plane = align_up(N * 4);
  return plane * (7 + 9 * (two_f ? 2 : 1));
}

// Stream-ordered table update: launches enqueued before it still see the old
// table, launches after it the new one, so admissions never wait for the GPU.
int upload_table(swalbe_batch* b) {
  if (!b->dirty) return SWALBE_OK;
  const int r = b->staging_next;
  b->staging_next = (r + 1) % kRing;
  if (b->staging_busy[r]) SW_HIP(hipEventSynchronize(b->staging_ev[r]));
  char* st = b->staging[r];
  int* tstart = (int*)(st + b->off_tstart);
  BatchFilm* films = (BatchFilm*)(st + b->off_films);
  int nf = 0, total = 0;
  bool any_ft = false;
  for (auto& s : b->slots) {
    if (!s.used) continue;
    tstart[nf] = total;
    films[nf] = s.d;
    total += s.d.ntiles;
    any_ft = any_ft || s.d.read_ft;
    ++nf;
  }
  *(BatchHdr*)st = BatchHdr{nf, total};
  const size_t bytes = b->off_films + (size_t)nf * sizeof(BatchFilm);
  SW_HIP(hipMemcpyAsync(b->d_table, st, bytes, hipMemcpyHostToDevice, b->stream));
  SW_HIP(hipEventRecord(b->staging_ev[r], b->stream));
  b->staging_busy[r] = true;
  b->nactive = nf;
  b->any_ft = any_ft;
  b->dirty = false;
  return SWALBE_OK;
}

int record_steps_event(swalbe_batch* b) {
  hipEvent_t ev;
  if (!b->ev_pool.empty()) {
    ev = b->ev_pool.back();
    b->ev_pool.pop_back();
  } else {
    SW_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  SW_HIP(hipEventRecord(ev, b->stream));
  b->steps_ev.emplace_back(b->counter, ev);
  if (b->steps_ev.size() > 64) {  // the oldest is long done: recycle it
    SW_HIP(hipEventSynchronize(b->steps_ev.front().second));
    b->ev_pool.push_back(b->steps_ev.front().second);
    b->steps_ev.pop_front();
  }
  return SWALBE_OK;
}

// Blocks until the device has run every launch up to launch counter `target`.
int wait_counter(swalbe_batch* b, unsigned long long target) {
  while (!b->steps_ev.empty()) {
    auto [c, ev] = b->steps_ev.front();
    if (c >= target) {
      SW_HIP(hipEventSynchronize(ev));
      return SWALBE_OK;
    }
    // an earlier enqueue call: done once the one covering `target` is
    SW_HIP(hipEventSynchronize(ev));
    b->ev_pool.push_back(ev);
    b->steps_ev.pop_front();
  }
  return SWALBE_OK;
}

int enqueue_steps(swalbe_batch* b, int64_t nsteps) {
  if (int rc = upload_table(b)) return rc;
  if (nsteps == 0) return SWALBE_OK;
  // the table just enqueued decides the kernel: the one without the f_t
  // path while every film has τ = 1
  const bool any = b->any_ft;
  const int grid = any ? b->grid_any : b->grid_tau1;
  int64_t done = 0;
  if (b->graph_steps > 0 && nsteps >= b->graph_steps) {
    if (b->gexec && b->graph_any != any) {  // captured for the other kernel
      SW_HIP(hipGraphExecDestroy(b->gexec));
      SW_HIP(hipGraphDestroy(b->graph));
      b->gexec = nullptr;
      b->graph = nullptr;
    }
    if (!b->gexec) {
      SW_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal));
      for (int i = 0; i < b->graph_steps; ++i)
        launch_batch_f32(b->films(), b->tstart(), b->hdr(), b->d_counter, grid, b->nts, any,
                         b->stream);
      hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      SW_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->graph_any = any;
    }
    while (nsteps - done >= b->graph_steps) {
      SW_HIP(hipGraphLaunch(b->gexec, b->stream));
      done += b->graph_steps;
    }
  }
  for (; done < nsteps; ++done)
    launch_batch_f32(b->films(), b->tstart(), b->hdr(), b->d_counter, grid, b->nts, any,
                     b->stream);
  SW_HIP(hipGetLastError());
  b->counter += (unsigned long long)nsteps;
  return record_steps_event(b);
}

void release(swalbe_batch* b) {
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (auto& pe : b->steps_ev) hipEventDestroy(pe.second);
  for (hipEvent_t ev : b->ev_pool) hipEventDestroy(ev);
  for (int r = 0; r < kRing; ++r) {
    if (b->staging_ev[r]) hipEventDestroy(b->staging_ev[r]);
    if (b->staging[r]) hipHostFree(b->staging[r]);
  }
  void* ptrs[] = {b->arena, b->d_table, b->d_counter};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (b->dl_stream) hipStreamDestroy(b->dl_stream);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

// ---------------------------------------------------------------- autotuner
// The tile height (8 or 16 rows) of a film shape is chosen by measurement:
// a temporary batch of copies of the film (enough to fill the device, as in
// a batch) is stepped with each height and the faster one kept.  Both
// heights give the same bits, so the choice changes the speed only.
// Results are cached per device and per (shape, τ ≠ 1, noise, body force).
struct TuneKey {
  int dev, Lx, Ly, read_ft, thermal, body;
  bool operator<(const TuneKey& o) const {
    return std::tie(dev, Lx, Ly, read_ft, thermal, body) <
           std::tie(o.dev, o.Lx, o.Ly, o.read_ft, o.thermal, o.body);
  }
};
struct TuneEntry {
  int ty;
  double ms8, ms16;  // per batched step of the copies
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tuned;

constexpr double kTuneNodes = 4.0e6;  // nodes per timed batch of copies
constexpr int kTuneWarm = 2, kTuneSteps = 8;

// Milliseconds per batched step of `copies` copies of the film with tile
// height ty: the faster of two timed runs of kTuneSteps steps.
int time_copies(const swalbe_params* p, int Lx, int Ly, const swalbe_film_init* init, int ty,
                int copies, double* ms_per_step) {
  const size_t fb = film_bytes(Lx, Ly, p->tau != 1.0);
  swalbe_batch* t = nullptr;
  if (int rc = swalbe_batch_create(SWALBE_F32, fb * (size_t)copies + kAlign, copies, &t))
    return rc;
  t->forced_ty = ty;
  int rc = SWALBE_OK;
  for (int i = 0; i < copies && rc == SWALBE_OK; ++i) {
    int32_t slot = 0;
    rc = swalbe_batch_add(t, p, Lx, Ly, init, kTuneWarm + 2 * kTuneSteps, &slot);
  }
  double best = 0.0;
  if (rc == SWALBE_OK) rc = swalbe_batch_step(t, kTuneWarm);
  for (int r = 0; r < 2 && rc == SWALBE_OK; ++r) {
    double ms = 0.0;
    rc = swalbe_batch_step_timed(t, kTuneSteps, &ms);
    if (rc == SWALBE_OK && (r == 0 || ms < best)) best = ms;
  }
  swalbe_batch_destroy(t);
  if (rc == SWALBE_OK) *ms_per_step = best / kTuneSteps;
  return rc;
}

int tuned_tile_y(const swalbe_params* p, int Lx, int Ly, const swalbe_film_init* init,
                 TuneEntry* out) {
  int dev = 0;
  hipGetDevice(&dev);
  const TuneKey key{dev, Lx, Ly, p->tau != 1.0 ? 1 : 0, init->thermal ? 1 : 0,
                    init->body ? 1 : 0};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      *out = it->second;
      return SWALBE_OK;
    }
  }
  // as many copies as make ~4 M nodes, within a quarter of the free memory
  const size_t fb = film_bytes(Lx, Ly, p->tau != 1.0);
  int copies = (int)std::min(1024.0, std::max(1.0, std::ceil(kTuneNodes / ((double)Lx * Ly))));
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
    while (copies > 1 && fb * (size_t)copies > free_b / 4) copies /= 2;
  TuneEntry e{kBatchTileY, 0.0, 0.0};
  if (int rc = time_copies(p, Lx, Ly, init, 16, copies, &e.ms16)) return rc;
  if (int rc = time_copies(p, Lx, Ly, init, 8, copies, &e.ms8)) return rc;
  e.ty = e.ms8 < 0.98 * e.ms16 ? 8 : 16;  // 16 on a near tie (less halo work)
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = e;
  }
  *out = e;
  return SWALBE_OK;
}

}  // namespace

extern "C" {

int swalbe_batch_set_autotune(swalbe_batch* b, int32_t on) {
  if (!b) return bad("null batch");
  b->autotune = on ? 1 : 0;
  return SWALBE_OK;
}

int swalbe_batch_set_tile_y(swalbe_batch* b, int32_t rows) {
  if (!b || (rows != 0 && rows != 8 && rows != 16))
    return bad("swalbe_batch_set_tile_y: rows is 0, 8 or 16");
  b->forced_ty = rows;
  return SWALBE_OK;
}

int swalbe_batch_tune_shape(const swalbe_params* p, int32_t Lx, int32_t Ly,
                            const swalbe_film_init* init, int32_t* tile_y, double* ms_rows8,
                            double* ms_rows16) {
  if (!p || !init || !tile_y || Lx < 1 || Ly < 1 || p->n < 0 || p->m < 0)
    return bad("swalbe_batch_tune_shape: bad arguments");
  TuneEntry e{};
  if (int rc = tuned_tile_y(p, Lx, Ly, init, &e)) return rc;
  *tile_y = e.ty;
  if (ms_rows8) *ms_rows8 = e.ms8;
  if (ms_rows16) *ms_rows16 = e.ms16;
  return SWALBE_OK;
}

int swalbe_batch_film_tile_y(swalbe_batch* b, int32_t slot, int32_t* tile_y) {
  if (!b || !tile_y || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used)
    return bad("swalbe_batch_film_tile_y: bad slot");
  *tile_y = b->slots[slot].d.ty;
  return SWALBE_OK;
}

int swalbe_batch_film_bytes(int32_t Lx, int32_t Ly, double tau, size_t* bytes) {
  if (!bytes || Lx < 1 || Ly < 1) return bad("swalbe_batch_film_bytes: bad arguments");
  *bytes = film_bytes(Lx, Ly, tau != 1.0);
  return SWALBE_OK;
}

int swalbe_batch_create(int32_t dtype, size_t capacity_bytes, int32_t max_films,
                        swalbe_batch** out) {
  if (!out || dtype != SWALBE_F32 || capacity_bytes < kAlign || max_films < 1)
    return bad("swalbe_batch_create: FP32 only, capacity and max_films must be positive");
  swalbe_batch* b = new swalbe_batch();
  hipGetDevice(&b->device);
  hipError_t e = hipMalloc(&b->arena, capacity_bytes);
  if (e != hipSuccess) {
    b->arena = nullptr;
    release(b);
    set_error(std::string("swalbe_batch_create: arena hipMalloc failed: ") + hipGetErrorString(e));
    return SWALBE_ENOMEM;
  }
  b->capacity = capacity_bytes / kAlign * kAlign;
  b->free_[0] = b->capacity;
  b->slots.resize(max_films);
  b->off_tstart = 64;
  b->off_films = (b->off_tstart + sizeof(int) * (size_t)max_films + 63) / 64 * 64;
  b->table_bytes = b->off_films + sizeof(BatchFilm) * (size_t)max_films;
  int rc = SWALBE_OK;
  if (hipMalloc(&b->d_table, b->table_bytes) != hipSuccess ||
      hipMalloc(&b->d_counter, sizeof(unsigned long long)) != hipSuccess) {
    set_error("swalbe_batch_create: table allocation failed");
    rc = SWALBE_ENOMEM;
  }
  for (int r = 0; r < kRing && rc == SWALBE_OK; ++r) {
    if (hipHostMalloc((void**)&b->staging[r], b->table_bytes, hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreateWithFlags(&b->staging_ev[r], hipEventDisableTiming) != hipSuccess) {
      b->staging[r] = nullptr;
      set_error("swalbe_batch_create: pinned staging allocation failed");
      rc = SWALBE_ENOMEM;
    }
  }
  if (rc == SWALBE_OK && (hipStreamCreate(&b->stream) != hipSuccess ||
                          hipStreamCreateWithFlags(&b->dl_stream, hipStreamNonBlocking) !=
                              hipSuccess))
    rc = hip_fail(hipErrorUnknown, "hipStreamCreate");
  if (rc == SWALBE_OK) {
    const unsigned long long zero = 0;
    const BatchHdr hdr{0, 0};
    if (hipMemcpy(b->d_counter, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_table, &hdr, sizeof(hdr), hipMemcpyHostToDevice) != hipSuccess)
      rc = hip_fail(hipErrorUnknown, "batch table init");
  }
  if (rc != SWALBE_OK) {
    release(b);
    return rc;
  }
  int cus = 0;
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  if (cus <= 0) cus = 256;
  b->grid_any = (cus * batch_blocks_per_cu(true) + 7) / 8 * 8;
  b->grid_tau1 = (cus * batch_blocks_per_cu(false) + 7) / 8 * 8;
  *out = b;
  return SWALBE_OK;
}

int swalbe_batch_destroy(swalbe_batch* b) {
  if (b) release(b);
  return SWALBE_OK;
}

int swalbe_batch_set_graph(swalbe_batch* b, int32_t steps_per_graph) {
  if (!b || steps_per_graph < 0) return bad("swalbe_batch_set_graph: bad arguments");
  hipSetDevice(b->device);
  if (b->gexec) {
    hipStreamSynchronize(b->stream);
    hipGraphExecDestroy(b->gexec);
    hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
  }
  b->graph_steps = steps_per_graph;
  return SWALBE_OK;
}

int swalbe_batch_set_nontemporal(swalbe_batch* b, int32_t on) {
  if (!b) return bad("null batch");
  if (on < 0 || on > 2) return bad("nontemporal is 0, 1 or 2");
  b->nts = on;
  if (b->gexec) return swalbe_batch_set_graph(b, b->graph_steps);  // re-capture with the new flag
  return SWALBE_OK;
}

int swalbe_batch_set_final_outputs_only(swalbe_batch* b, int32_t on) {
  if (!b) return bad("null batch");
  b->last_only = on ? 1 : 0;
  return SWALBE_OK;
}

int swalbe_batch_add(swalbe_batch* b, const swalbe_params* p, int32_t Lx, int32_t Ly,
                     const swalbe_film_init* init, int64_t nsteps, int32_t* slot_out) {
  if (!b || !p || !init || !slot_out || Lx < 1 || Ly < 1 || nsteps < 0)
    return bad("swalbe_batch_add: bad arguments");
  if ((int64_t)Lx * (Ly + 1) > (int64_t)0xFFFFFFFFLL)  // 32-bit noise stream indices
    return bad("swalbe_batch_add: film too large");
  if (p->n < 0 || p->m < 0) return bad("swalbe_batch_add: negative exponent");
  hipSetDevice(b->device);
  int ty = kBatchTileY;
  if (b->forced_ty) {
    ty = b->forced_ty;
  } else if (b->autotune) {
    TuneEntry e{};
    bool cached = false;
    {
      int dev = 0;
      hipGetDevice(&dev);
      std::lock_guard<std::mutex> lk(g_tune_mu);
      cached = g_tuned.count(TuneKey{dev, Lx, Ly, p->tau != 1.0 ? 1 : 0, init->thermal ? 1 : 0,
                                     init->body ? 1 : 0}) > 0;
    }
    // a shape seen for the first time is timed on an idle device
    if (!cached) SW_HIP(hipStreamSynchronize(b->stream));
    if (int rc = tuned_tile_y(p, Lx, Ly, init, &e)) return rc;
    ty = e.ty;
  }
  int slot = -1;
  for (size_t i = 0; i < b->slots.size(); ++i)
    if (!b->slots[i].used) {
      slot = (int)i;
      break;
    }
  if (slot < 0) return bad("swalbe_batch_add: no free film slot");
  const bool two_f = p->tau != 1.0;
  const size_t bytes = film_bytes(Lx, Ly, two_f);
  size_t off = 0;
  if (!arena_alloc(b, bytes, &off)) {
    set_error("swalbe_batch_add: the batch arena has no room for this film");
    return SWALBE_ENOMEM;
  }
  Slot& s = b->slots[slot];
  s = Slot();
  s.used = true;
  s.off = off;
  s.bytes = bytes;
  s.enc = make_enc(init->href, p->g);
  s.nsteps = nsteps;
  const size_t N = (size_t)Lx * Ly;
  const size_t plane = align_up(N * 4);
  char* base = b->arena + off;
  float* q[7];
  for (int i = 0; i < 7; ++i) q[i] = (float*)(base + i * plane);
  BatchFilm& d = s.d;
  d.h[0] = q[0];
  d.h[1] = q[1];
  d.ux[0] = q[2];
  d.ux[1] = q[3];
  d.uy[0] = q[4];
  d.uy[1] = q[5];
  d.p = q[6];
  d.fstride = (int64_t)(plane / 4);
  d.f[0] = (float*)(base + 7 * plane);
  d.f[1] = two_f ? (float*)(base + 7 * plane + 9 * plane) : nullptr;
  d.Lx = Lx;
  d.Ly = Ly;
  if (kBatchStrips) {
    d.ntx = (Lx + kStripOut - 1) / kStripOut;
    d.ntiles = d.ntx * ((Ly + kStripSeg - 1) / kStripSeg);
    d.small = strip_is_small(Lx, Ly);
    d.pairs = 0;
  } else {
    d.ntx = (Lx + 63) / 64;
    d.ntiles = d.ntx * ((Ly + ty - 1) / ty);
    d.small = film_is_small(Lx, Ly, 64, ty);
    d.pairs = (!d.small && (Ly % 2) == 0) ? 1 : 0;
  }
  d.ty = ty;
  d.read_ft = two_f ? 1 : 0;
  d.last_only = b->last_only;
  d.thermal = init->thermal ? 1 : 0;
  d.body = init->body ? 1 : 0;
  d.seed = init->seed;
  d.t0 = 0;
  d.admit = b->counter;
  d.end = b->counter + (unsigned long long)nsteps;
  swalbe_step_opts o{};
  o.thermal = init->thermal;
  o.body = init->body;
  o.seed = init->seed;
  o.body_ax = init->body_ax;
  o.body_ay = init->body_ay;
  o.body_fac = init->body_fac;
  const StepParams sp = make_step_params(p, &o, 0);
  d.K = make_f32k(sp, s.enc);
  // initial state: h from the spec, u = 0, p = 0, f_t = feq(h, 0) when τ ≠ 1
  // heights are stored relative to the FP32 reference height the kernel adds
  launch_film_init<float>(d.h[0], Lx, Ly, *init, (double)s.enc.href, b->stream);
  for (int i = 2; i < 7; ++i) launch_fill<float>(q[i], 0.0, (int64_t)N, b->stream);
  if (two_f)
    launch_equilibrium<float>(d.f[0], d.h[0], d.ux[0], d.uy[0], d.p, p->g, Lx, Ly, d.fstride,
                              s.enc, b->stream);
  else
    launch_fill<float>(d.f[0], 0.0, (int64_t)9 * d.fstride, b->stream);
  launch_fill<float>(d.p, 0.0, (int64_t)N, b->stream);
  SW_HIP(hipGetLastError());
  b->dirty = true;
  *slot_out = slot;
  return SWALBE_OK;
}

int swalbe_batch_remove(swalbe_batch* b, int32_t slot) {
  if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used)
    return bad("swalbe_batch_remove: bad slot");
  // stream-ordered: a film added later into this memory is initialised
  // after every launch already enqueued
  Slot& s = b->slots[slot];
  arena_free(b, s.off, s.bytes);
  s.used = false;
  b->dirty = true;
  return SWALBE_OK;
}

int swalbe_batch_step(swalbe_batch* b, int64_t nsteps) {
  if (!b || nsteps < 0) return bad("swalbe_batch_step: bad arguments");
  hipSetDevice(b->device);
  return enqueue_steps(b, nsteps);
}

int swalbe_batch_step_timed(swalbe_batch* b, int64_t nsteps, double* ms) {
  if (!b || nsteps < 0 || !ms) return bad("swalbe_batch_step_timed: bad arguments");
  hipSetDevice(b->device);
  if (int rc = upload_table(b)) return rc;  // table upload outside the timed region
  hipEvent_t e0, e1;
  SW_HIP(hipEventCreate(&e0));
  SW_HIP(hipEventCreate(&e1));
  SW_HIP(hipEventRecord(e0, b->stream));
  if (int rc = enqueue_steps(b, nsteps)) return rc;
  SW_HIP(hipEventRecord(e1, b->stream));
  SW_HIP(hipEventSynchronize(e1));
  float f = 0.f;
  SW_HIP(hipEventElapsedTime(&f, e0, e1));
  *ms = f;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return SWALBE_OK;
}

int swalbe_batch_wait(swalbe_batch* b, int64_t steps_behind) {
  if (!b || steps_behind < 0) return bad("swalbe_batch_wait: bad arguments");
  hipSetDevice(b->device);
  if ((unsigned long long)steps_behind >= b->counter) return SWALBE_OK;
  return wait_counter(b, b->counter - (unsigned long long)steps_behind);
}

int swalbe_batch_sync(swalbe_batch* b) {
  if (!b) return bad("null batch");
  hipSetDevice(b->device);
  SW_HIP(hipStreamSynchronize(b->stream));
  return SWALBE_OK;
}

int swalbe_batch_film_status(swalbe_batch* b, int32_t slot, int64_t* steps_done,
                             int32_t* finished) {
  if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used)
    return bad("swalbe_batch_film_status: bad slot");
  const BatchFilm& d = b->slots[slot].d;
  const unsigned long long now = std::min(b->counter, d.end);
  if (steps_done) *steps_done = (int64_t)(now - d.admit);
  if (finished) *finished = b->counter >= d.end ? 1 : 0;
  return SWALBE_OK;
}

int swalbe_batch_download(swalbe_batch* b, int32_t slot, int32_t field, void* host, size_t bytes) {
  if (!b || !host || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used)
    return bad("swalbe_batch_download: bad arguments");
  hipSetDevice(b->device);
  const Slot& s = b->slots[slot];
  const BatchFilm& d = s.d;
  const int64_t N = (int64_t)d.Lx * d.Ly;
  const unsigned long long n = std::min(b->counter, d.end) - d.admit;
  const int cur = (int)(n & 1ull);
  const float* src = nullptr;
  int64_t count = N;
  double off = 0.0;
  switch (field) {
    case SWALBE_HEIGHT: src = d.h[cur]; off = (double)s.enc.href; break;
    case SWALBE_VELX: src = d.ux[cur]; break;
    case SWALBE_VELY: src = d.uy[cur]; break;
    case SWALBE_PRESSURE: src = d.p; break;
    case SWALBE_FOUT: src = d.read_ft ? d.f[cur] : d.f[0]; count = 9 * N; break;
    default: return bad("swalbe_batch_download: field must be HEIGHT, VELX, VELY, PRESSURE or FOUT");
  }
  if (d.last_only && !d.read_ft && (field == SWALBE_FOUT || field == SWALBE_PRESSURE) &&
      b->counter < d.end)
    return bad("swalbe_batch_download: this film writes its distributions and pressure on its "
               "last step only (swalbe_batch_set_final_outputs_only); they are not ready yet");
  size_t hes;
  if (bytes == (size_t)count * 8) hes = 8;
  else if (bytes == (size_t)count * 4) hes = 4;
  else return bad("swalbe_batch_download: byte count does not match the field size");
  std::vector<float> tmp(count);
  // a finished film is not touched by later launches: wait only for the
  // enqueue call that ran its last step, so the copy overlaps newer steps
  if (n > 0 && b->counter >= d.end) {
    if (int rc = wait_counter(b, d.end)) return rc;
  } else {
    SW_HIP(hipStreamSynchronize(b->stream));
  }
  if (field == SWALBE_FOUT)
    SW_HIP(hipMemcpy2DAsync(tmp.data(), N * 4, src, d.fstride * 4, N * 4, 9,
                            hipMemcpyDeviceToHost, b->dl_stream));
  else
    SW_HIP(hipMemcpyAsync(tmp.data(), src, N * 4, hipMemcpyDeviceToHost, b->dl_stream));
  SW_HIP(hipStreamSynchronize(b->dl_stream));
  for (int64_t i = 0; i < count; ++i) {
    const double o = field == SWALBE_FOUT ? (double)s.enc.R[i / N] : off;
    const double v = (double)tmp[i] + o;
    if (hes == 8) ((double*)host)[i] = v;
    else ((float*)host)[i] = (float)v;
  }
  return SWALBE_OK;
}

int swalbe_batch_stats(swalbe_batch* b, int32_t* nfilms, int64_t* nodes, size_t* bytes_used,
                       size_t* bytes_free) {
  if (!b) return bad("null batch");
  int nf = 0;
  int64_t nn = 0;
  for (auto& s : b->slots)
    if (s.used) {
      ++nf;
      nn += (int64_t)s.d.Lx * s.d.Ly;
    }
  if (nfilms) *nfilms = nf;
  if (nodes) *nodes = nn;
  if (bytes_used) *bytes_used = b->used_bytes;
  if (bytes_free) *bytes_free = b->capacity - b->used_bytes;
  return SWALBE_OK;
}

}  // extern "C"
