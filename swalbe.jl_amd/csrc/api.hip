// api.hip -- the C ABI of libswalbe_hip (declared in include/swalbe_hip.h).
//
// A swalbe_state owns every field of one film on one device and one HIP
// stream (default priority).  Calls on one state are not re-entrant;
// distinct states may be driven from distinct threads.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "swalbe_batch.h"

struct swalbe_state {
  swalbe::Film film;
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;  // the stream created with the state
  double* partials = nullptr;  // reduction scratch, 4 doubles per block
  int nparts = 0;
  double* red_out = nullptr;   // 4 doubles
};

namespace swalbe {

namespace {
thread_local std::string g_last_error;

constexpr int kMaxParts = 1024;

inline bool is_dist(int field) {
  return field == SWALBE_FOUT || field == SWALBE_FTEMP || field == SWALBE_FEQ;
}
inline bool is_height(int field) { return field == SWALBE_HEIGHT; }
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return SWALBE_EHIP;
}

Enc zero_enc() {
  Enc e;
  e.href = 0.f;
  for (int k = 0; k < kQ; ++k) e.R[k] = 0.f;
  return e;
}

// R = feq(href, u = 0) (src/equilibrium.jl:74-114 at u = 0).
Enc make_enc(double href, double g) {
  Enc e;
  e.href = (float)href;
  const double r0 = href * (1.0 - (5.0 / 6.0) * g * href);
  const double r1 = (1.0 / 9.0) * href * (1.5 * g * href);
  const double r5 = (1.0 / 36.0) * href * (1.5 * g * href);
  e.R[0] = (float)r0;
  for (int k = 1; k <= 4; ++k) e.R[k] = (float)r1;
  for (int k = 5; k <= 8; ++k) e.R[k] = (float)r5;
  return e;
}

}  // namespace swalbe

using namespace swalbe;

#define SW_CHECK_STATE(st)                   \
  do {                                       \
    if (!(st)) {                             \
      set_error("null swalbe_state handle"); \
      return SWALBE_EINVAL;                  \
    }                                        \
  } while (0)

#define SW_HIP(call)                                  \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

namespace {

// Device pointer and element count of a field; distributions have 9 planes.
void* field_ptr(swalbe_state* st, int field, int64_t* count) {
  Film& F = st->film;
  *count = F.N;
  switch (field) {
    case SWALBE_FOUT: *count = 9 * F.N; return F.f[F.fout_buf];
    case SWALBE_FTEMP: *count = 9 * F.N; return F.f[F.ftemp_buf];
    case SWALBE_FEQ: *count = 9 * F.N; return F.feq;
    case SWALBE_HEIGHT: return F.height();
    case SWALBE_VELX: return F.velx();
    case SWALBE_VELY: return F.vely();
    case SWALBE_VSQ: return F.vsq;
    case SWALBE_PRESSURE: return F.pressure;
    case SWALBE_FX: return F.Fx;
    case SWALBE_FY: return F.Fy;
    case SWALBE_SLIPX: return F.slipx;
    case SWALBE_SLIPY: return F.slipy;
    case SWALBE_HGRADPX: return F.hgpx;
    case SWALBE_HGRADPY: return F.hgpy;
    case SWALBE_KBTX: return F.kbtx;
    case SWALBE_KBTY: return F.kbty;
    case SWALBE_THETA: return F.theta;
    default: return nullptr;
  }
}

// Before a write of fout or ftemp: if both names refer to one buffer, give
// the written one the other buffer so the other name keeps its content.
void unalias_for_write(Film& F, int field) {
  if (F.fout_buf != F.ftemp_buf) return;
  if (field == SWALBE_FOUT) F.fout_buf = 1 - F.ftemp_buf;
  if (field == SWALBE_FTEMP) F.ftemp_buf = 1 - F.fout_buf;
}

// encoding of a plain value of `field` (plane k for distributions)
inline double enc_offset(const Film& F, int field, int k) {
  if (F.dtype != SWALBE_F32) return 0.0;
  if (is_height(field)) return F.enc.href;
  if (is_dist(field)) return F.enc.R[k];
  return 0.0;
}

int alloc_field(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return SWALBE_ENOMEM;
  }
  return SWALBE_OK;
}

template <typename T>
int fill_field(swalbe_state* st, int field, double value) {
  Film& F = st->film;
  int64_t count;
  T* p = (T*)field_ptr(st, field, &count);
  if (!p) {
    set_error("field not allocated");
    return SWALBE_EINVAL;
  }
  if (is_dist(field)) {
    for (int k = 0; k < kQ; ++k)
      launch_fill<T>(p + k * F.fstride, value - enc_offset(F, field, k), F.N, st->stream);
  } else {
    launch_fill<T>(p, value - enc_offset(F, field, 0), F.N, st->stream);
  }
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

// κ(θ) field for θ-field runs, recomputed when θ or (n, m, hmin) change.
int ensure_kappa(swalbe_state* st, int n, int m, double hmin) {
  Film& F = st->film;
  if (!F.theta) {
    set_error("no θ field uploaded (SWALBE_THETA)");
    return SWALBE_EINVAL;
  }
  if (!F.kappa)
    if (int rc = alloc_field(&F.kappa, F.esz * F.N)) return rc;
  if (F.kappa_valid && F.kappa_n == n && F.kappa_m == m && F.kappa_hmin == hmin) return SWALBE_OK;
  if (F.dtype == SWALBE_F64)
    launch_kappa_field<double>((double*)F.kappa, (const double*)F.theta, n, m, hmin, F.N, st->stream);
  else
    launch_kappa_field<float>((float*)F.kappa, (const float*)F.theta, n, m, hmin, F.N, st->stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "kappa field");
  F.kappa_valid = true;
  F.kappa_n = n;
  F.kappa_m = m;
  F.kappa_hmin = hmin;
  return SWALBE_OK;
}

// host <-> device copies of a field in the state's element type; the
// distribution planes sit fstride elements apart on the device.
int copy_h2d(swalbe_state* st, int field, void* dst, const void* src) {
  const Film& F = st->film;
  hipError_t e;
  if (is_dist(field))
    e = hipMemcpy2D(dst, F.fstride * F.esz, src, F.N * F.esz, F.N * F.esz, kQ, hipMemcpyHostToDevice);
  else
    e = hipMemcpy(dst, src, F.N * F.esz, hipMemcpyHostToDevice);
  return e == hipSuccess ? SWALBE_OK : hip_fail(e, "copy to device");
}
int copy_d2h(swalbe_state* st, int field, void* dst, const void* src) {
  const Film& F = st->film;
  hipError_t e;
  if (is_dist(field))
    e = hipMemcpy2D(dst, F.N * F.esz, src, F.fstride * F.esz, F.N * F.esz, kQ, hipMemcpyDeviceToHost);
  else
    e = hipMemcpy(dst, src, F.N * F.esz, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SWALBE_OK : hip_fail(e, "copy to host");
}

int set_device(swalbe_state* st) {
  hipError_t e = hipSetDevice(st->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return SWALBE_OK;
}

}  // namespace

extern "C" {

const char* swalbe_last_error(void) { return g_last_error.c_str(); }

int swalbe_version(void) { return 10000; }

int swalbe_device_count(int32_t* n) {
  if (!n) return SWALBE_EINVAL;
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = (e == hipSuccess) ? c : 0;
  return SWALBE_OK;
}

int swalbe_create(int32_t Lx, int32_t Ly, int32_t dtype, int32_t kind, swalbe_state** out) {
  if (!out || Lx < 1 || Ly < 1 || (dtype != SWALBE_F64 && dtype != SWALBE_F32) ||
      (kind != SWALBE_SIMPLE && kind != SWALBE_THERMAL)) {
    set_error("swalbe_create: invalid argument");
    return SWALBE_EINVAL;
  }
  // noise stream indices 2 (x + Lx floor(y/2)) + y mod 2 are 32-bit
  if ((int64_t)Lx * (Ly + 1) > (int64_t)0xFFFFFFFFLL) {
    set_error("swalbe_create: Lx (Ly + 1) must stay below 2^32 (32-bit noise stream indices)");
    return SWALBE_EINVAL;
  }
  swalbe_state* st = new swalbe_state();
  Film& F = st->film;
  F.Lx = Lx;
  F.Ly = Ly;
  F.dtype = dtype;
  F.esz = dtype == SWALBE_F64 ? 8 : 4;
  F.N = (int64_t)Lx * Ly;
  // distribution planes may be staggered by a pad (elements, env SWALBE_PLANE_PAD)
  {
    int64_t pad = 0;
    if (const char* e = getenv("SWALBE_PLANE_PAD")) pad = atoll(e);
    if (pad < 0) pad = 0;
    F.fstride = F.N + pad;
  }
  F.href = dtype == SWALBE_F32 ? 1.0 : 0.0;
  F.g_enc = 0.0;
  F.enc = dtype == SWALBE_F32 ? make_enc(F.href, 0.0) : zero_enc();
  hipGetDevice(&st->device);
  int rc = SWALBE_OK;
  const size_t fb = F.esz * F.N;
  void** fields[] = {&F.f[0], &F.f[1], &F.feq};
  for (void** p : fields)
    if (rc == SWALBE_OK) rc = alloc_field(p, 9 * F.esz * F.fstride);
  void** scal[] = {&F.hbuf[0], &F.hbuf[1], &F.uxbuf[0], &F.uxbuf[1], &F.uybuf[0], &F.uybuf[1],
                   &F.vsq,     &F.pressure, &F.Fx,     &F.Fy,      &F.slipx,    &F.slipy,
                   &F.hgpx,    &F.hgpy,     &F.kbtx,   &F.kbty};
  for (void** p : scal)
    if (rc == SWALBE_OK) rc = alloc_field(p, fb);
  st->nparts = (int)std::min<int64_t>(kMaxParts, std::max<int64_t>(1, F.N / 4096));
  if (rc == SWALBE_OK) rc = alloc_field((void**)&st->partials, sizeof(double) * 4 * kMaxParts);
  if (rc == SWALBE_OK) rc = alloc_field((void**)&st->red_out, sizeof(double) * 4);
  if (rc == SWALBE_OK) {
    hipError_t e = hipStreamCreate(&st->stream);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamCreate");
    st->own_stream = st->stream;
  }
  if (rc != SWALBE_OK) {
    swalbe_destroy(st);
    return rc;
  }
  // Sys: every array zero, height = 1 (src/initialize.jl:492-508)
  const int ids[] = {SWALBE_FOUT,  SWALBE_FTEMP,    SWALBE_FEQ, SWALBE_VELX, SWALBE_VELY,
                     SWALBE_VSQ,   SWALBE_PRESSURE, SWALBE_FX,  SWALBE_FY,   SWALBE_SLIPX,
                     SWALBE_SLIPY, SWALBE_HGRADPX,  SWALBE_HGRADPY, SWALBE_KBTX, SWALBE_KBTY};
  F.fout_buf = 0;
  F.ftemp_buf = 1;
  for (int id : ids) {
    rc = dtype == SWALBE_F64 ? fill_field<double>(st, id, 0.0) : fill_field<float>(st, id, 0.0);
    if (rc != SWALBE_OK) break;
  }
  if (rc == SWALBE_OK)
    rc = dtype == SWALBE_F64 ? fill_field<double>(st, SWALBE_HEIGHT, 1.0)
                             : fill_field<float>(st, SWALBE_HEIGHT, 1.0);
  if (rc == SWALBE_OK) {
    hipError_t e = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
  }
  if (rc != SWALBE_OK) {
    swalbe_destroy(st);
    return rc;
  }
  (void)kind;  // every state carries kbtx/kbty; kind only documents intent
  *out = st;
  return SWALBE_OK;
}

int swalbe_destroy(swalbe_state* st) {
  if (!st) return SWALBE_OK;
  hipSetDevice(st->device);
  if (st->stream) hipStreamSynchronize(st->stream);
  if (st->own_stream && st->own_stream != st->stream) hipStreamSynchronize(st->own_stream);
  Film& F = st->film;
  void* ptrs[] = {F.f[0],  F.f[1],  F.feq,   F.hbuf[0], F.hbuf[1], F.uxbuf[0], F.uxbuf[1],
                  F.uybuf[0], F.uybuf[1], F.vsq, F.pressure, F.Fx, F.Fy, F.slipx, F.slipy,
                  F.hgpx,  F.hgpy,  F.kbtx,  F.kbty,    F.theta,   F.kappa, st->partials,
                  st->red_out};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (st->own_stream) hipStreamDestroy(st->own_stream);
  delete st;
  return SWALBE_OK;
}

int swalbe_set_stream(swalbe_state* st, void* stream) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  // work already enqueued on the old stream finishes first
  SW_HIP(hipStreamSynchronize(st->stream));
  st->stream = stream ? (hipStream_t)stream : st->own_stream;
  return SWALBE_OK;
}

int swalbe_upload(swalbe_state* st, int32_t field, const void* host, size_t bytes) {
  SW_CHECK_STATE(st);
  if (field < 0 || field >= SWALBE_NFIELDS || !host) {
    set_error("swalbe_upload: bad field id or null host pointer");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  const int64_t count = is_dist(field) ? 9 * F.N : F.N;
  size_t hes;
  if (bytes == (size_t)count * 8) hes = 8;
  else if (bytes == (size_t)count * 4) hes = 4;
  else {
    set_error("swalbe_upload: byte count does not match the field size");
    return SWALBE_EINVAL;
  }
  if (field == SWALBE_THETA && !F.theta)
    if (int rc = alloc_field(&F.theta, F.esz * F.N)) return rc;
  if (field == SWALBE_THETA) F.kappa_valid = false;
  if (field == SWALBE_FOUT || field == SWALBE_FTEMP) unalias_for_write(F, field);
  int64_t c2;
  void* dst = field_ptr(st, field, &c2);
  std::vector<unsigned char> tmp((size_t)count * F.esz);
  for (int64_t i = 0; i < count; ++i) {
    const double v = hes == 8 ? ((const double*)host)[i] : (double)((const float*)host)[i];
    const double off = enc_offset(F, field, (int)(i / F.N));
    if (F.esz == 8) ((double*)tmp.data())[i] = v - off;
    else ((float*)tmp.data())[i] = (float)(v - off);
  }
  SW_HIP(hipStreamSynchronize(st->stream));
  return copy_h2d(st, field, dst, tmp.data());
}

// Stored representation of a field (esz bytes per element; FP32 heights and
// distributions as deviations from the reference film): exact round trips.
int swalbe_download_raw(swalbe_state* st, int32_t field, void* host, size_t bytes) {
  SW_CHECK_STATE(st);
  if (field < 0 || field >= SWALBE_NFIELDS || !host) {
    set_error("swalbe_download_raw: bad field id or null host pointer");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  int64_t count;
  const void* src = field_ptr(st, field, &count);
  if (!src) {
    set_error("swalbe_download_raw: field not allocated");
    return SWALBE_EINVAL;
  }
  if (bytes != (size_t)count * F.esz) {
    set_error("swalbe_download_raw: byte count must be the stored size (elements x 4 or 8)");
    return SWALBE_EINVAL;
  }
  SW_HIP(hipStreamSynchronize(st->stream));
  return copy_d2h(st, field, host, src);
}

int swalbe_upload_raw(swalbe_state* st, int32_t field, const void* host, size_t bytes) {
  SW_CHECK_STATE(st);
  if (field < 0 || field >= SWALBE_NFIELDS || !host) {
    set_error("swalbe_upload_raw: bad field id or null host pointer");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  const int64_t count = is_dist(field) ? 9 * F.N : F.N;
  if (bytes != (size_t)count * F.esz) {
    set_error("swalbe_upload_raw: byte count must be the stored size (elements x 4 or 8)");
    return SWALBE_EINVAL;
  }
  if (field == SWALBE_THETA && !F.theta)
    if (int rc = alloc_field(&F.theta, F.esz * F.N)) return rc;
  if (field == SWALBE_THETA) F.kappa_valid = false;
  if (field == SWALBE_FOUT || field == SWALBE_FTEMP) unalias_for_write(F, field);
  int64_t c2;
  void* dst = field_ptr(st, field, &c2);
  SW_HIP(hipStreamSynchronize(st->stream));
  return copy_h2d(st, field, dst, host);
}

// FP32 reference height and gravity of the deviation storage (href = 0 for FP64).
int swalbe_get_reference_height(swalbe_state* st, double* href, double* g) {
  SW_CHECK_STATE(st);
  if (href) *href = st->film.dtype == SWALBE_F32 ? st->film.href : 0.0;
  if (g) *g = st->film.g_enc;
  return SWALBE_OK;
}

int swalbe_copy(swalbe_state* st, int32_t dst_field, int32_t src_field) {
  SW_CHECK_STATE(st);
  if (dst_field < 0 || dst_field >= SWALBE_NFIELDS || src_field < 0 ||
      src_field >= SWALBE_NFIELDS || dst_field == SWALBE_THETA || src_field == SWALBE_THETA ||
      is_dist(dst_field) != is_dist(src_field) || is_height(dst_field) != is_height(src_field)) {
    set_error("swalbe_copy: fields must be both distributions or both the same kind of plane");
    return SWALBE_EINVAL;
  }
  if (dst_field == src_field) return SWALBE_OK;
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  if (dst_field == SWALBE_FOUT || dst_field == SWALBE_FTEMP) unalias_for_write(F, dst_field);
  int64_t c;
  void* dst = field_ptr(st, dst_field, &c);
  const void* src = field_ptr(st, src_field, &c);
  if (dst == src) return SWALBE_OK;
  const int64_t rows = is_dist(dst_field) ? kQ : 1;
  SW_HIP(hipMemcpy2DAsync(dst, F.fstride * F.esz, src, F.fstride * F.esz, F.N * F.esz, rows,
                          hipMemcpyDeviceToDevice, st->stream));
  return SWALBE_OK;
}

int swalbe_download(swalbe_state* st, int32_t field, void* host, size_t bytes) {
  SW_CHECK_STATE(st);
  if (field < 0 || field >= SWALBE_NFIELDS || !host) {
    set_error("swalbe_download: bad field id or null host pointer");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  int64_t count;
  const void* src = field_ptr(st, field, &count);
  if (!src) {
    set_error("swalbe_download: field not allocated (upload SWALBE_THETA first)");
    return SWALBE_EINVAL;
  }
  size_t hes;
  if (bytes == (size_t)count * 8) hes = 8;
  else if (bytes == (size_t)count * 4) hes = 4;
  else {
    set_error("swalbe_download: byte count does not match the field size");
    return SWALBE_EINVAL;
  }
  std::vector<unsigned char> tmp((size_t)count * F.esz);
  SW_HIP(hipStreamSynchronize(st->stream));
  if (int rc = copy_d2h(st, field, tmp.data(), src)) return rc;
  for (int64_t i = 0; i < count; ++i) {
    const double raw = F.esz == 8 ? ((double*)tmp.data())[i] : (double)((float*)tmp.data())[i];
    const double off = enc_offset(F, field, (int)(i / F.N));
    const double v = off == 0.0 ? raw : raw + off;
    if (hes == 8) ((double*)host)[i] = v;
    else ((float*)host)[i] = (float)v;
  }
  return SWALBE_OK;
}

int swalbe_fill(swalbe_state* st, int32_t field, double value) {
  SW_CHECK_STATE(st);
  if (field < 0 || field >= SWALBE_NFIELDS) {
    set_error("swalbe_fill: bad field id");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  if (field == SWALBE_THETA && !F.theta)
    if (int rc = alloc_field(&F.theta, F.esz * F.N)) return rc;
  if (field == SWALBE_THETA) F.kappa_valid = false;
  if (field == SWALBE_FOUT || field == SWALBE_FTEMP) unalias_for_write(F, field);
  return F.dtype == SWALBE_F64 ? fill_field<double>(st, field, value)
                               : fill_field<float>(st, field, value);
}

int swalbe_synchronize(swalbe_state* st) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  SW_HIP(hipStreamSynchronize(st->stream));
  return SWALBE_OK;
}

int swalbe_set_reference_height(swalbe_state* st, double href, double g) {
  SW_CHECK_STATE(st);
  Film& F = st->film;
  if (F.dtype != SWALBE_F32) return SWALBE_OK;
  if (!std::isfinite(href) || !std::isfinite(g)) {
    set_error("swalbe_set_reference_height: non-finite value");
    return SWALBE_EINVAL;
  }
  // re-encode resident height and distributions through the host (rare call)
  const int ids[] = {SWALBE_HEIGHT, SWALBE_FOUT, SWALBE_FTEMP, SWALBE_FEQ};
  std::vector<std::vector<double>> keep;
  const bool aliased = F.fout_buf == F.ftemp_buf;
  for (int id : ids) {
    const int64_t count = is_dist(id) ? 9 * F.N : F.N;
    keep.emplace_back(count);
    if (id == SWALBE_FTEMP && aliased) continue;
    if (int rc = swalbe_download(st, id, keep.back().data(), count * 8)) return rc;
  }
  F.href = href;
  F.g_enc = g;
  F.enc = make_enc(href, g);
  for (size_t i = 0; i < 4; ++i) {
    const int id = ids[i];
    if (id == SWALBE_FTEMP && aliased) continue;
    // write without un-aliasing: re-encoding keeps buffer identities
    int64_t count;
    void* dst = field_ptr(st, id, &count);
    std::vector<unsigned char> tmp((size_t)count * 4);
    for (int64_t j = 0; j < count; ++j)
      ((float*)tmp.data())[j] = (float)(keep[i][j] - enc_offset(F, id, (int)(j / F.N)));
    if (int rc = copy_h2d(st, id, dst, tmp.data())) return rc;
  }
  return SWALBE_OK;
}

// ---------------------------------------------------------------- per-op entry points
#define SW_DISPATCH(st, CALL_F64, CALL_F32)                 \
  do {                                                      \
    if ((st)->film.dtype == SWALBE_F64) {                   \
      using T = double;                                     \
      CALL_F64;                                             \
    } else {                                                \
      using T = float;                                      \
      CALL_F32;                                             \
    }                                                       \
  } while (0)

#define SW_BOTH(st, CALL) SW_DISPATCH(st, CALL, CALL)

int swalbe_equilibrium(swalbe_state* st, double g) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  SW_BOTH(st, launch_equilibrium<T>((T*)F.feq, (const T*)F.height(), (const T*)F.velx(),
                                    (const T*)F.vely(), (T*)F.vsq, g, F.Lx, F.Ly, F.fstride, F.enc,
                                    st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_bgk_and_stream(swalbe_state* st, double tau) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  const int dst = F.fout_buf != F.ftemp_buf ? F.fout_buf : 1 - F.ftemp_buf;
  SW_BOTH(st, launch_bgk_and_stream<T>((T*)F.f[dst], (const T*)F.feq, (const T*)F.f[F.ftemp_buf],
                                       (const T*)F.Fx, (const T*)F.Fy, tau, F.Lx, F.Ly, F.fstride,
                                       F.enc, st->stream));
  SW_HIP(hipGetLastError());
  F.fout_buf = F.ftemp_buf = dst;
  return SWALBE_OK;
}

int swalbe_moments(swalbe_state* st) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  SW_BOTH(st, launch_moments<T>((T*)F.height(), (T*)F.velx(), (T*)F.vely(),
                                (const T*)F.f[F.fout_buf], F.Lx, F.Ly, F.fstride, F.enc,
                                st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

static int powmode_for(int32_t n, int32_t m, int32_t raw) {
  if (!raw) return 0;
  if (n == 9 && m == 3) return 1;
  if (n == 3 && m == 2) return 2;
  return -1;
}

int swalbe_filmpressure(swalbe_state* st, double gamma, double theta, int32_t n, int32_t m,
                        double hmin, double hcrit, int32_t use_theta_field, int32_t raw) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  const int mode = powmode_for(n, m, raw);
  if (mode < 0) {
    set_error("DomainError: (n, m) must be (9, 3) or (3, 2) for the raw-array filmpressure!");
    return SWALBE_EDOMAIN;
  }
  if (n < 0 || m < 0) {
    set_error("filmpressure: negative exponent");
    return SWALBE_EINVAL;
  }
  if (use_theta_field)
    if (int rc = ensure_kappa(st, n, m, hmin)) return rc;
  const double kap = kappa(theta, n, m, hmin);
  SW_BOTH(st, launch_filmpressure<T>((T*)F.pressure, (const T*)F.height(), gamma, kap, nullptr,
                                     use_theta_field ? (const T*)F.kappa : nullptr, n, m, hmin,
                                     hcrit, mode, F.Lx, F.Ly, F.enc, st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_hgradp(swalbe_state* st) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  SW_BOTH(st, launch_grad<T>((T*)F.hgpx, (T*)F.hgpy, (const T*)F.pressure, nullptr,
                             (const T*)F.height(), F.Lx, F.Ly, F.enc, st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_slippage(swalbe_state* st, double delta, double mu) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  SW_BOTH(st, launch_slippage<T>((T*)F.slipx, (T*)F.slipy, (const T*)F.height(),
                                 (const T*)F.velx(), (const T*)F.vely(), delta, mu, F.Lx, F.Ly,
                                 F.enc, st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_thermal(swalbe_state* st, double kbt, double mu, double delta, uint64_t seed,
                   uint64_t step) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  SW_BOTH(st, launch_thermal<T>((T*)F.kbtx, (T*)F.kbty, (const T*)F.height(), kbt, mu, delta,
                                seed, step, F.Lx, F.Ly, F.enc, st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_inclination(swalbe_state* st, double ax, double ay, double fac) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  SW_BOTH(st, launch_inclination<T>((T*)F.Fx, (T*)F.Fy, (const T*)F.height(), ax, ay, fac, F.Lx,
                                    F.Ly, F.enc, st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_force_sum(swalbe_state* st, int32_t with_thermal) {
  SW_CHECK_STATE(st);
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  SW_BOTH(st, launch_force_sum<T>((T*)F.Fx, (T*)F.Fy, (const T*)F.hgpx, (const T*)F.hgpy,
                                  (const T*)F.slipx, (const T*)F.slipy,
                                  with_thermal ? (const T*)F.kbtx : nullptr,
                                  with_thermal ? (const T*)F.kbty : nullptr, F.N, st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

// ---------------------------------------------------------------- reductions
static int reduce_enqueue(swalbe_state* st, const void* p, double href, double thresh,
                          double* out_dev) {
  Film& F = st->film;
  SW_BOTH(st, launch_reduce<T>((const T*)p, F.N, href, thresh, st->partials, st->nparts,
                               st->stream));
  launch_reduce_finish(st->partials, st->nparts, out_dev, st->stream);
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_reduce(swalbe_state* st, int32_t field, double thresh, double out[4]) {
  SW_CHECK_STATE(st);
  if (!out || field < 0 || field >= SWALBE_NFIELDS || is_dist(field)) {
    set_error("swalbe_reduce: bad field (scalar fields only)");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  int64_t count;
  const void* p = field_ptr(st, field, &count);
  if (!p) {
    set_error("swalbe_reduce: field not allocated");
    return SWALBE_EINVAL;
  }
  if (int rc = reduce_enqueue(st, p, enc_offset(st->film, field, 0), thresh, st->red_out)) return rc;
  SW_HIP(hipMemcpyAsync(out, st->red_out, 4 * sizeof(double), hipMemcpyDeviceToHost, st->stream));
  SW_HIP(hipStreamSynchronize(st->stream));
  return SWALBE_OK;
}

int swalbe_measure(swalbe_state* st, double thresh, int32_t use_theta_field, double theta,
                   int32_t surface, double out[4], uint8_t* host_mask, size_t mask_bytes) {
  SW_CHECK_STATE(st);
  Film& F = st->film;
  if (!out || (host_mask && mask_bytes != (size_t)F.N)) {
    set_error("swalbe_measure: null output or mask size != Lx*Ly");
    return SWALBE_EINVAL;
  }
  if (use_theta_field && !F.theta) {
    set_error("swalbe_measure: no θ field uploaded");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  uint8_t* mask = nullptr;
  if (host_mask) SW_HIP(hipMalloc((void**)&mask, F.N));
  SW_BOTH(st, launch_measure<T>((const T*)F.height(), enc_offset(F, SWALBE_HEIGHT, 0), F.Lx, F.Ly,
                                thresh, surface, cospi(theta),
                                use_theta_field ? (const T*)F.theta : nullptr, mask, st->partials,
                                st->nparts, st->red_out, st->stream));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, st->red_out, 4 * sizeof(double), hipMemcpyDeviceToHost, st->stream);
  if (e == hipSuccess && mask)
    e = hipMemcpyAsync(host_mask, mask, F.N, hipMemcpyDeviceToHost, st->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
  if (mask) (void)hipFree(mask);
  if (e != hipSuccess) return hip_fail(e, "swalbe_measure");
  return SWALBE_OK;
}

int swalbe_init_height(swalbe_state* st, const swalbe_film_init* init) {
  SW_CHECK_STATE(st);
  if (!init || init->kind < SWALBE_INIT_FLAT || init->kind > SWALBE_INIT_TWO_CAPS) {
    set_error("swalbe_init_height: bad init spec");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  // the same offset as uploads: the FP32 reference height the kernels add
  const double off = enc_offset(F, SWALBE_HEIGHT, 0);
  SW_BOTH(st, launch_film_init<T>((T*)F.height(), F.Lx, F.Ly, *init, off, st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_device_memory(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) return hip_fail(e, "hipMemGetInfo");
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return SWALBE_OK;
}

int swalbe_init_random(swalbe_state* st, int32_t dist, uint64_t seed, double base, double amp) {
  SW_CHECK_STATE(st);
  if (dist != 0 && dist != 1) {
    set_error("swalbe_init_random: dist must be 0 (uniform) or 1 (normal)");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  const double off = enc_offset(F, SWALBE_HEIGHT, 0);
  SW_BOTH(st, launch_init_random<T>((T*)F.height(), dist, seed, base, amp, off, F.Lx, F.Ly,
                                    st->stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- fused time loop
namespace swalbe {
StepParams make_step_params(const swalbe_params* p, const swalbe_step_opts* o, int powmode) {
  StepParams sp{};
  sp.gamma = p->gamma;
  sp.kappa = kappa(p->theta, p->n, p->m, p->hmin);
  sp.hmin = p->hmin;
  sp.hcrit = p->hcrit;
  sp.n = p->n;
  sp.m = p->m;
  sp.powmode = powmode;
  sp.omeg = 1.0 - 1.0 / p->tau;
  sp.itau = 1.0 / p->tau;
  sp.g = p->g;
  sp.c6mu = 6.0 * p->mu;
  sp.c6d = 6.0 * p->delta;
  sp.c3dd = 3.0 * (p->delta * p->delta);
  sp.thermal = o && o->thermal;
  sp.kb26 = ((2.0 * p->kbt) * p->mu) * 6.0;
  sp.delta = p->delta;
  sp.d3dd = (3.0 * p->delta) * p->delta;
  sp.seed = o ? o->seed : 0;
  sp.body = o && o->body;
  sp.bax = o ? o->body_ax : 0.0;
  sp.bay = o ? o->body_ay : 0.0;
  sp.bfac = o ? o->body_fac : 0.0;
  sp.use_theta = o && o->use_theta_field;
  sp.read_ft = sp.omeg != 0.0;
  sp.materialize = o && o->materialize;
  sp.wet_thresh = o ? o->wet_thresh : 0.0;
  sp.fp64_guard = o ? !o->no_fp64_guard : 1;
  return sp;
}

}  // namespace swalbe

namespace {

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_init_keys(unsigned long long* p, int64_t nsteps) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nsteps;
       i += (int64_t)gridDim.x * blockDim.x) {
    p[2 * i] = 0ull;
    p[2 * i + 1] = ~0ull;
  }
}

// Snapshots of the height: a device copy per snapshot into a ring slot on
// the state's stream, then a device-to-host copy of that slot on a second
// stream into pinned staging, decoded once the slot comes round again.
struct SnapRing {
  static constexpr int R = 4;
  void* dev[R] = {};
  void* host[R] = {};
  hipEvent_t copied[R] = {};  // D2H of the slot done
  hipEvent_t ready = nullptr;  // snapshot's device copy done (state stream)
  hipStream_t cs = nullptr;
  int64_t pending[R];          // snapshot index held by each slot, -1 = none
  ~SnapRing() {
    if (cs) hipStreamSynchronize(cs);
    for (int r = 0; r < R; ++r) {
      if (dev[r]) hipFree(dev[r]);
      if (host[r]) hipHostFree(host[r]);
      if (copied[r]) hipEventDestroy(copied[r]);
    }
    if (ready) hipEventDestroy(ready);
    if (cs) hipStreamDestroy(cs);
  }
};

struct LoopBuffers {
  unsigned long long* keys = nullptr;  // 2 per step
  unsigned long long* wet = nullptr;   // 1 per step
  double* mass = nullptr;              // 4 per tdump
  double* meas = nullptr;              // 4 per step: count, max, area, energy
  uint8_t* fluid = nullptr;            // Lx*Ly per step
  ~LoopBuffers() {
    if (keys) hipFree(keys);
    if (wet) hipFree(wet);
    if (mass) hipFree(mass);
    if (meas) hipFree(meas);
    if (fluid) hipFree(fluid);
  }
};

template <typename T>
void enqueue_step(swalbe_state* st, const StepParams& sp, StepSeries ser, uint64_t t, int tile_y,
                  int variant, int nts, bool outs) {
  Film& F = st->film;
  const int cur = F.hu_cur;
  const int fdst = F.fout_buf != F.ftemp_buf ? F.fout_buf : 1 - F.ftemp_buf;
  StepIO<T> io{};
  io.h_in = (const T*)F.hbuf[cur];
  io.ux_in = (const T*)F.uxbuf[cur];
  io.uy_in = (const T*)F.uybuf[cur];
  io.ft = (const T*)F.f[F.ftemp_buf];
  io.f_out = (T*)F.f[fdst];
  io.h_out = (T*)F.hbuf[1 - cur];
  io.ux_out = (T*)F.uxbuf[1 - cur];
  io.uy_out = (T*)F.uybuf[1 - cur];
  io.p_out = (T*)F.pressure;
  io.kappa = sp.use_theta ? (const T*)F.kappa : nullptr;
  if (!outs) {  // distributions and pressure of a step nobody reads (τ = 1)
    io.f_out = nullptr;
    io.p_out = nullptr;
  }
  if constexpr (sizeof(T) == 8)
    launch_step<T>(io, sp, ser, t, F.Lx, F.Ly, F.fstride, F.enc, tile_y, nts, st->stream);
  else
    launch_step_f32(io, sp, ser, t, F.Lx, F.Ly, F.fstride, F.enc, tile_y, variant, nts,
                    st->stream);
  F.fout_buf = F.ftemp_buf = fdst;
  F.hu_cur = 1 - cur;
}

// Recompute the intermediates of the last step from the previous moments,
// which the ping-pong buffers still hold (same kernels as the per-op path).
template <typename T>
void materialize_last(swalbe_state* st, const swalbe_params* p, const swalbe_step_opts* o,
                      uint64_t tlast) {
  Film& F = st->film;
  const int prev = 1 - F.hu_cur;
  const T* h = (const T*)F.hbuf[prev];
  const T* ux = (const T*)F.uxbuf[prev];
  const T* uy = (const T*)F.uybuf[prev];
  launch_grad<T>((T*)F.hgpx, (T*)F.hgpy, (const T*)F.pressure, nullptr, h, F.Lx, F.Ly, F.enc,
                 st->stream);
  launch_slippage<T>((T*)F.slipx, (T*)F.slipy, h, ux, uy, p->delta, p->mu, F.Lx, F.Ly, F.enc,
                     st->stream);
  if (o->thermal)
    launch_thermal<T>((T*)F.kbtx, (T*)F.kbty, h, p->kbt, p->mu, p->delta, o->seed, tlast, F.Lx,
                      F.Ly, F.enc, st->stream);
  launch_force_sum<T>((T*)F.Fx, (T*)F.Fy, (const T*)F.hgpx, (const T*)F.hgpy, (const T*)F.slipx,
                      (const T*)F.slipy, o->thermal ? (const T*)F.kbtx : nullptr,
                      o->thermal ? (const T*)F.kbty : nullptr, F.N, st->stream);
  if (o->body)
    launch_inclination<T>((T*)F.Fx, (T*)F.Fy, h, o->body_ax, o->body_ay, o->body_fac, F.Lx, F.Ly,
                          F.enc, st->stream);
  launch_equilibrium<T>((T*)F.feq, h, ux, uy, (T*)F.vsq, p->g, F.Lx, F.Ly, F.fstride, F.enc,
                        st->stream);
}

int run_loop(swalbe_state* st, const swalbe_params* p, int64_t nsteps, const swalbe_step_opts* o,
             double* ms_out) {
  SW_CHECK_STATE(st);
  if (!p || nsteps < 0) {
    set_error("swalbe_time_loop: null params or negative step count");
    return SWALBE_EINVAL;
  }
  if (int rc = set_device(st)) return rc;
  Film& F = st->film;
  swalbe_step_opts zero{};
  if (!o) o = &zero;
  if (o->use_theta_field)
    if (int rc = ensure_kappa(st, p->n, p->m, p->hmin)) return rc;
  if (p->n < 0 || p->m < 0) {
    set_error("swalbe_time_loop: negative exponent");
    return SWALBE_EINVAL;
  }
  if (o->variant < 0 || o->variant > 3) {
    set_error("swalbe_time_loop: variant must be 0, 1, 2 or 3");
    return SWALBE_EINVAL;
  }
  if (o->nontemporal < 0 || o->nontemporal > 2) {
    set_error("swalbe_time_loop: nontemporal must be 0, 1 or 2");
    return SWALBE_EINVAL;
  }
  const bool strips = o->variant == 3;
  // default tile height: 16 rows for FP32.  FP64: 16 rows where the film's
  // inputs (h, ux, uy) exceed the eight 4 MB L2s (the tile then shares one LDS
  // buffer between p and the forces, 3 workgroups per CU: 32.2 against 30.8
  // GLUPS at 4096²), else 8 (38.4 against 33.8 on a 1024² film)
  const bool f64_big = (int64_t)F.Lx * F.Ly * 24 > ((int64_t)32 << 20);
  const int tile_y = o->tile_y > 0 ? o->tile_y
                                   : (strips ? kStripSeg
                                             : (F.dtype == SWALBE_F64 ? (f64_big ? 16 : 8) : 16));
  if (strips ? (tile_y < 4 || tile_y > 4096) : (tile_y != 8 && tile_y != 16)) {
    set_error("swalbe_time_loop: tile_y must be 8 or 16 (4..4096 rows per segment for variant 3)");
    return SWALBE_EINVAL;
  }
  const StepParams sp = make_step_params(p, o, 0);
  // with τ = 1 a step reads no distributions, so those of every step but the
  // last (and its pressure) are dead unless the caller wants them written
  const bool final_only = o->final_outputs_only && !sp.read_ft;
  LoopBuffers lb;
  const bool want_maxh = o->maxh_series != nullptr;
  const bool want_dh = o->dh_series != nullptr || want_maxh, want_wet = o->wet_series != nullptr;
  const bool want_mass = o->mass_series != nullptr && p->tdump > 0;
  const bool want_surf = o->area_series != nullptr || o->energy_series != nullptr;
  const bool want_fluid = o->fluid_series != nullptr;
  const bool want_meas = want_surf || want_fluid;
  if (ms_out && (want_dh || want_wet || want_mass || want_meas)) {
    set_error("swalbe_time_loop_timed: series are not recorded by the timed variant");
    return SWALBE_EINVAL;
  }
  if (want_surf && o->surf_theta_field && !F.theta) {
    set_error("swalbe_time_loop: surface energy with the θ field, but no θ field uploaded");
    return SWALBE_EINVAL;
  }
  if (want_fluid && (double)nsteps * (double)F.N > 4e9) {
    set_error("swalbe_time_loop: the fluid_dry series needs nsteps * Lx * Ly bytes on the device "
              "(more than 4 GB); run fewer steps per call");
    return SWALBE_EINVAL;
  }
  if (want_meas) SW_HIP(hipMalloc((void**)&lb.meas, sizeof(double) * 4 * std::max<int64_t>(1, nsteps)));
  if (want_fluid) SW_HIP(hipMalloc((void**)&lb.fluid, (size_t)F.N * std::max<int64_t>(1, nsteps)));
  const double meas_thresh = o->surf_thresh > 0.0 ? o->surf_thresh : 0.055;
  if (want_dh) SW_HIP(hipMalloc(&lb.keys, sizeof(unsigned long long) * 2 * std::max<int64_t>(1, nsteps)));
  if (want_wet) SW_HIP(hipMalloc(&lb.wet, sizeof(unsigned long long) * std::max<int64_t>(1, nsteps)));
  int64_t nmass = 0;
  for (int64_t it = 1; it <= nsteps && want_mass; ++it)
    if ((o->t0 + it) % p->tdump == 0) ++nmass;
  if (want_mass && nmass) SW_HIP(hipMalloc(&lb.mass, sizeof(double) * 4 * nmass));
  if (want_dh)  // even slots hold max keys (start 0), odd slots min keys (start ~0)
    hipLaunchKernelGGL(k_init_keys, dim3(64), dim3(256), 0, st->stream, lb.keys, nsteps);
  if (want_wet) hipLaunchKernelGGL(k_fill_u64, dim3(64), dim3(256), 0, st->stream, lb.wet, 0ull, nsteps);
  const bool want_snap = o->snap_every > 0 && o->snap_series != nullptr;
  if (ms_out && want_snap) {
    set_error("swalbe_time_loop_timed: snapshots are not taken by the timed variant");
    return SWALBE_EINVAL;
  }
  SnapRing ring;
  const size_t plane_bytes = (size_t)F.N * F.esz;
  const double snap_off = enc_offset(F, SWALBE_HEIGHT, 0);
  auto snap_drain = [&](int r) -> int {  // decode the snapshot held by slot r
    if (ring.pending[r] < 0) return SWALBE_OK;
    SW_HIP(hipEventSynchronize(ring.copied[r]));
    double* dst = o->snap_series + ring.pending[r] * F.N;
    if (F.esz == 8) {
      const double* src = (const double*)ring.host[r];
      for (int64_t i = 0; i < F.N; ++i) dst[i] = src[i] + snap_off;
    } else {
      const float* src = (const float*)ring.host[r];
      for (int64_t i = 0; i < F.N; ++i) dst[i] = (double)src[i] + snap_off;
    }
    ring.pending[r] = -1;
    return SWALBE_OK;
  };
  if (want_snap) {
    for (int r = 0; r < SnapRing::R; ++r) {
      ring.pending[r] = -1;
      SW_HIP(hipMalloc(&ring.dev[r], plane_bytes));
      SW_HIP(hipHostMalloc(&ring.host[r], plane_bytes, hipHostMallocDefault));
      SW_HIP(hipEventCreateWithFlags(&ring.copied[r], hipEventDisableTiming));
    }
    SW_HIP(hipEventCreateWithFlags(&ring.ready, hipEventDisableTiming));
    SW_HIP(hipStreamCreateWithFlags(&ring.cs, hipStreamNonBlocking));
  }
  int64_t ksnap = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms_out) {
    SW_HIP(hipEventCreate(&e0));
    SW_HIP(hipEventCreate(&e1));
    SW_HIP(hipEventRecord(e0, st->stream));
  }
  int64_t km = 0;
  for (int64_t it = 1; it <= nsteps; ++it) {
    const int64_t t = o->t0 + it;
    if (want_mass && t % p->tdump == 0) {
      if (int rc = reduce_enqueue(st, F.height(), enc_offset(F, SWALBE_HEIGHT, 0), 0.0,
                                  lb.mass + 4 * km))
        return rc;
      ++km;
    }
    if (want_meas) {  // measurement callbacks on the step's input height, no host sync
      SW_BOTH(st, launch_measure<T>(
                      (const T*)F.height(), enc_offset(F, SWALBE_HEIGHT, 0), F.Lx, F.Ly,
                      meas_thresh, want_surf ? 1 : 0, cospi(p->theta),
                      want_surf && o->surf_theta_field ? (const T*)F.theta : nullptr,
                      want_fluid ? lb.fluid + (size_t)F.N * (it - 1) : nullptr, st->partials,
                      st->nparts, lb.meas + 4 * (it - 1), st->stream));
    }
    StepSeries ser{want_dh ? lb.keys + 2 * (it - 1) : nullptr, want_wet ? lb.wet + (it - 1) : nullptr};
    const bool outs = !final_only || it == nsteps;
    if (F.dtype == SWALBE_F64)
      enqueue_step<double>(st, sp, ser, (uint64_t)t, tile_y, o->variant, o->nontemporal, outs);
    else
      enqueue_step<float>(st, sp, ser, (uint64_t)t, tile_y, o->variant, o->nontemporal, outs);
    if (want_snap && t % o->snap_every == 0) {
      const int r = (int)(ksnap % SnapRing::R);
      if (int rc = snap_drain(r)) return rc;
      SW_HIP(hipMemcpyAsync(ring.dev[r], F.height(), plane_bytes, hipMemcpyDeviceToDevice,
                            st->stream));
      SW_HIP(hipEventRecord(ring.ready, st->stream));
      SW_HIP(hipStreamWaitEvent(ring.cs, ring.ready, 0));
      SW_HIP(hipMemcpyAsync(ring.host[r], ring.dev[r], plane_bytes, hipMemcpyDeviceToHost, ring.cs));
      SW_HIP(hipEventRecord(ring.copied[r], ring.cs));
      ring.pending[r] = ksnap++;
    }
  }
  for (int r = 0; r < SnapRing::R && want_snap; ++r)
    if (int rc = snap_drain((int)((ksnap + r) % SnapRing::R))) return rc;
  if (ms_out) {
    SW_HIP(hipEventRecord(e1, st->stream));
    SW_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SW_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_out = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  SW_HIP(hipGetLastError());
  if (o->materialize && nsteps > 0) {
    if (F.dtype == SWALBE_F64) materialize_last<double>(st, p, o, (uint64_t)(o->t0 + nsteps));
    else materialize_last<float>(st, p, o, (uint64_t)(o->t0 + nsteps));
    SW_HIP(hipGetLastError());
  }
  if (want_dh) {
    std::vector<unsigned long long> keys(2 * nsteps);
    SW_HIP(hipMemcpyAsync(keys.data(), lb.keys, keys.size() * 8, hipMemcpyDeviceToHost, st->stream));
    SW_HIP(hipStreamSynchronize(st->stream));
    for (int64_t i = 0; i < nsteps; ++i) {
      if (o->dh_series) o->dh_series[i] = dkey_inv(keys[2 * i]) - dkey_inv(keys[2 * i + 1]);
      if (want_maxh) o->maxh_series[i] = dkey_inv(keys[2 * i]);
    }
  }
  if (want_meas && nsteps > 0) {
    std::vector<double> m(4 * nsteps);
    SW_HIP(hipMemcpyAsync(m.data(), lb.meas, m.size() * 8, hipMemcpyDeviceToHost, st->stream));
    if (want_fluid)
      SW_HIP(hipMemcpyAsync(o->fluid_series, lb.fluid, (size_t)F.N * nsteps,
                            hipMemcpyDeviceToHost, st->stream));
    SW_HIP(hipStreamSynchronize(st->stream));
    for (int64_t i = 0; i < nsteps; ++i) {
      if (o->area_series) o->area_series[i] = m[4 * i + 2];
      if (o->energy_series) o->energy_series[i] = m[4 * i + 3];
    }
  }
  if (want_wet) {
    std::vector<unsigned long long> w(nsteps);
    SW_HIP(hipMemcpyAsync(w.data(), lb.wet, w.size() * 8, hipMemcpyDeviceToHost, st->stream));
    SW_HIP(hipStreamSynchronize(st->stream));
    for (int64_t i = 0; i < nsteps; ++i) o->wet_series[i] = (int64_t)w[i];
  }
  if (want_mass && nmass) {
    std::vector<double> m(4 * nmass);
    SW_HIP(hipMemcpyAsync(m.data(), lb.mass, m.size() * 8, hipMemcpyDeviceToHost, st->stream));
    SW_HIP(hipStreamSynchronize(st->stream));
    for (int64_t i = 0; i < nmass; ++i) o->mass_series[i] = m[4 * i];
  }
  return SWALBE_OK;
}

}  // namespace

extern "C" {

int swalbe_time_loop(swalbe_state* st, const swalbe_params* p, int64_t nsteps,
                     const swalbe_step_opts* opts) {
  return run_loop(st, p, nsteps, opts, nullptr);
}

int swalbe_time_loop_timed(swalbe_state* st, const swalbe_params* p, int64_t nsteps,
                           const swalbe_step_opts* opts, double* ms_out) {
  if (!ms_out) {
    set_error("swalbe_time_loop_timed: null ms_out");
    return SWALBE_EINVAL;
  }
  return run_loop(st, p, nsteps, opts, ms_out);
}

// ---------------------------------------------------------------- raw-array twins
static int check_raw(int32_t dtype, int32_t Lx, int32_t Ly, std::initializer_list<const void*> ptrs) {
  if ((dtype != SWALBE_F64 && dtype != SWALBE_F32) || Lx < 1 || Ly < 1) {
    set_error("raw op: bad dtype or size");
    return SWALBE_EINVAL;
  }
  for (const void* p : ptrs)
    if (!p) {
      set_error("raw op: null device pointer");
      return SWALBE_EINVAL;
    }
  return SWALBE_OK;
}

#define SW_RAW_DISPATCH(dtype, CALL) \
  do {                               \
    if ((dtype) == SWALBE_F64) {     \
      using T = double;              \
      CALL;                          \
    } else {                         \
      using T = float;               \
      CALL;                          \
    }                                \
  } while (0)

int swalbe_equilibrium_ptr(int32_t dtype, void* feq, const void* h, const void* ux, const void* uy,
                           void* vsq, double g, int32_t Lx, int32_t Ly, void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {feq, h, ux, uy, vsq})) return rc;
  SW_RAW_DISPATCH(dtype, launch_equilibrium<T>((T*)feq, (const T*)h, (const T*)ux, (const T*)uy,
                                               (T*)vsq, g, Lx, Ly, (int64_t)Lx * Ly, zero_enc(),
                                               (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_bgk_and_stream_ptr(int32_t dtype, void* fout, const void* feq, void* ftemp,
                              const void* Fx, const void* Fy, double tau, int32_t Lx, int32_t Ly,
                              void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {fout, feq, ftemp, Fx, Fy})) return rc;
  if (fout == ftemp) {
    set_error("swalbe_bgk_and_stream_ptr: fout and ftemp must be distinct arrays");
    return SWALBE_EINVAL;
  }
  SW_RAW_DISPATCH(dtype, launch_bgk_and_stream<T>((T*)fout, (const T*)feq, (const T*)ftemp,
                                                  (const T*)Fx, (const T*)Fy, tau, Lx, Ly,
                                                  (int64_t)Lx * Ly, zero_enc(), (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  const size_t bytes = (size_t)9 * Lx * Ly * (dtype == SWALBE_F64 ? 8 : 4);
  SW_HIP(hipMemcpyAsync(ftemp, fout, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return SWALBE_OK;
}

int swalbe_moments_ptr(int32_t dtype, void* h, void* ux, void* uy, const void* fout, int32_t Lx,
                       int32_t Ly, void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {h, ux, uy, fout})) return rc;
  SW_RAW_DISPATCH(dtype, launch_moments<T>((T*)h, (T*)ux, (T*)uy, (const T*)fout, Lx, Ly,
                                           (int64_t)Lx * Ly, zero_enc(), (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_filmpressure_ptr(int32_t dtype, void* p, const void* h, double gamma, double theta,
                            const void* theta_field, int32_t n, int32_t m, double hmin, double hcrit,
                            int32_t raw, int32_t Lx, int32_t Ly, void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {p, h})) return rc;
  const int mode = powmode_for(n, m, raw);
  if (mode < 0) {
    set_error("DomainError: (n, m) must be (9, 3) or (3, 2) for the raw-array filmpressure!");
    return SWALBE_EDOMAIN;
  }
  const double kap = kappa(theta, n, m, hmin);
  SW_RAW_DISPATCH(dtype, launch_filmpressure<T>((T*)p, (const T*)h, gamma, kap,
                                                (const T*)theta_field, nullptr, n, m, hmin, hcrit,
                                                mode, Lx, Ly, zero_enc(), (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_grad_ptr(int32_t dtype, void* outx, void* outy, const void* f, const void* a, int32_t Lx,
                    int32_t Ly, void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {outx, outy, f})) return rc;
  SW_RAW_DISPATCH(dtype, launch_grad<T>((T*)outx, (T*)outy, (const T*)f, (const T*)a, nullptr, Lx,
                                        Ly, zero_enc(), (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_laplacian_ptr(int32_t dtype, void* out, const void* f, double gamma, int32_t Lx,
                         int32_t Ly, void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {out, f})) return rc;
  SW_RAW_DISPATCH(dtype, launch_laplacian<T>((T*)out, (const T*)f, gamma, Lx, Ly,
                                             (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_hgradp_ptr(int32_t dtype, void* hgpx, void* hgpy, const void* p, const void* h,
                      int32_t Lx, int32_t Ly, void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {hgpx, hgpy, p, h})) return rc;
  SW_RAW_DISPATCH(dtype, launch_grad<T>((T*)hgpx, (T*)hgpy, (const T*)p, nullptr, (const T*)h, Lx,
                                        Ly, zero_enc(), (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_slippage_ptr(int32_t dtype, void* slipx, void* slipy, const void* h, const void* ux,
                        const void* uy, double delta, double mu, int32_t Lx, int32_t Ly,
                        void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {slipx, slipy, h, ux, uy})) return rc;
  SW_RAW_DISPATCH(dtype, launch_slippage<T>((T*)slipx, (T*)slipy, (const T*)h, (const T*)ux,
                                            (const T*)uy, delta, mu, Lx, Ly, zero_enc(),
                                            (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_thermal_ptr(int32_t dtype, void* kbtx, void* kbty, const void* h, double kbt, double mu,
                       double delta, uint64_t seed, uint64_t step, int32_t Lx, int32_t Ly,
                       void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {kbtx, kbty, h})) return rc;
  SW_RAW_DISPATCH(dtype, launch_thermal<T>((T*)kbtx, (T*)kbty, (const T*)h, kbt, mu, delta, seed,
                                           step, Lx, Ly, zero_enc(), (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_inclination_ptr(int32_t dtype, void* Fx, void* Fy, const void* h, double ax, double ay,
                           double fac, int32_t Lx, int32_t Ly, void* stream) {
  if (int rc = check_raw(dtype, Lx, Ly, {Fx, Fy, h})) return rc;
  SW_RAW_DISPATCH(dtype, launch_inclination<T>((T*)Fx, (T*)Fy, (const T*)h, ax, ay, fac, Lx, Ly,
                                               zero_enc(), (hipStream_t)stream));
  SW_HIP(hipGetLastError());
  return SWALBE_OK;
}

int swalbe_host_normals(uint64_t seed, uint64_t step, uint32_t tag, int64_t n, double* z0,
                        double* z1) {
  if (n < 0 || (n > 0 && (!z0 || !z1))) {
    set_error("swalbe_host_normals: bad arguments");
    return SWALBE_EINVAL;
  }
  for (int64_t i = 0; i < n; ++i) normal_pair(seed, step, (uint32_t)i, tag, &z0[i], &z1[i]);
  return SWALBE_OK;
}

int swalbe_host_uniforms(uint64_t seed, uint64_t step, uint32_t tag, int64_t n, double* u) {
  if (n < 0 || (n > 0 && !u)) {
    set_error("swalbe_host_uniforms: bad arguments");
    return SWALBE_EINVAL;
  }
  for (int64_t i = 0; i < n; ++i) u[i] = uniform(seed, step, (uint32_t)i, tag);
  return SWALBE_OK;
}

double swalbe_host_cospi(double x) { return cospi(x); }

int swalbe_malloc(void** dptr, size_t bytes) {
  if (!dptr) return SWALBE_EINVAL;
  return alloc_field(dptr, bytes);
}

int swalbe_free(void* dptr) {
  if (dptr) SW_HIP(hipFree(dptr));
  return SWALBE_OK;
}

int swalbe_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  SW_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SWALBE_OK;
}

int swalbe_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  SW_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SWALBE_OK;
}

}  // extern "C"
