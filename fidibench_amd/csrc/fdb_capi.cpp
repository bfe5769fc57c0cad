// fdb_capi.cpp -- host side of the C ABI declared in
// include/fidibench_amd/fdb_upwind.h.
//
// The handle mirrors the reference's Upwind<NDIMS> object
// (pletzer/fidibench upwind/cxx/upwind.cxx:19-135): the field lives in
// HBM as two ping-pong buffers; checksum()/std()/print()/saveVTK() copy it
// back in chunks and reproduce the reference's host arithmetic and output
// byte for byte.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fidibench_amd/fdb_upwind.h"
#include "fdb_internal.h"

namespace {
thread_local std::string g_err;
}  // namespace

namespace fdb {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace fdb

namespace {

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();  // clear the sticky launch error, if any
  return e == hipErrorOutOfMemory ? FDB_ERR_NOMEM : FDB_ERR_HIP;
}

#define FDB_HIP(call)                                  \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return hip_fail(_e, #call);  \
  } while (0)

bool valid_dtype(int d) {
  return d == FDB_F64 || d == FDB_F32 || d == FDB_BF16 || d == FDB_BF16_SR;
}

// Host copies of the device conversions (bf16: round to nearest even on
// the binary32 value; a double is first rounded to float).
inline uint16_t h_f32_to_bf16(float x) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu))
    return (uint16_t)((b >> 16) | 0x0040u);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}
inline float h_bf16_to_f32(uint16_t h) {
  uint32_t b = (uint32_t)h << 16;
  float x;
  std::memcpy(&x, &b, 4);
  return x;
}
inline double h_widen(int dtype, const void* p, uint64_t i) {
  switch (dtype) {
    case FDB_F64: return ((const double*)p)[i];
    case FDB_F32: return (double)((const float*)p)[i];
    default: return (double)h_bf16_to_f32(((const uint16_t*)p)[i]);
  }
}
inline void h_narrow(int dtype, void* p, uint64_t i, double x) {
  switch (dtype) {
    case FDB_F64: ((double*)p)[i] = x; break;
    case FDB_F32: ((float*)p)[i] = (float)x; break;
    default: ((uint16_t*)p)[i] = h_f32_to_bf16((float)x); break;
  }
}
// Same-type copies must not round-trip through double (bf16 NaN payloads,
// signed zeros are kept either way, but a plain copy is exact and faster).
void h_convert(int sd, const void* src, int dd, void* dst, uint64_t n) {
  sd = fdb::storage_dtype(sd);
  dd = fdb::storage_dtype(dd);
  if (sd == dd) {
    std::memcpy(dst, src, n * fdb::dtype_size(sd));
    return;
  }
  if (sd == FDB_BF16 && dd == FDB_F32) {
    for (uint64_t i = 0; i < n; ++i) ((float*)dst)[i] = h_bf16_to_f32(((const uint16_t*)src)[i]);
    return;
  }
  if (sd == FDB_F32 && dd == FDB_BF16) {
    for (uint64_t i = 0; i < n; ++i) ((uint16_t*)dst)[i] = h_f32_to_bf16(((const float*)src)[i]);
    return;
  }
  for (uint64_t i = 0; i < n; ++i) h_narrow(dd, dst, i, h_widen(sd, src, i));
}

constexpr uint64_t kStageBytes = 64ull << 20;  // host staging chunk

// Device scratch for the *_dev reductions, one per device, grow-only.
struct DevScratch {
  std::mutex mu;
  std::vector<std::pair<double*, size_t>> per_dev;
};
DevScratch g_scratch;

int dev_scratch(size_t bytes, double** out) {
  int dev = 0;
  FDB_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch.mu);
  if ((int)g_scratch.per_dev.size() <= dev) g_scratch.per_dev.resize(dev + 1, {nullptr, 0});
  auto& s = g_scratch.per_dev[dev];
  if (s.second < bytes) {
    if (s.first) FDB_HIP(hipFree(s.first));
    s.first = nullptr;
    s.second = 0;
    FDB_HIP(hipMalloc((void**)&s.first, bytes));
    s.second = bytes;
  }
  *out = s.first;
  return FDB_OK;
}

// A malformed tuning variable is an argument error of every launching
// entry point (reported before any device work).
int check_tuning() {
  if (const char* m = fdb::tuning_error()) return fail(FDB_ERR_INVALID, m);
  return FDB_OK;
}

int check_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(FDB_ERR_NODEV, "no HIP device available");
  return FDB_OK;
}

}  // namespace

// ----------------------------------------------------------------- handle

struct fdb_upwind {
  int ndims = 0;
  uint64_t n[3] = {1, 1, 1};
  double v[3] = {0, 0, 0};
  double len[3] = {0, 0, 0};
  int dtype = FDB_F64;
  size_t esz = 8;
  uint64_t ntot = 0;
  void* buf[2] = {nullptr, nullptr};
  int cur = 0;
  hipStream_t stream = nullptr;
  double* partials = nullptr;  // 6 * kReduceBlocks + 6
  int time_block = 1;
  uint64_t seed = 0;        // stochastic rounding (FDB_BF16_SR)
  uint64_t steps_done = 0;  // steps advanced since creation or set_seed
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* stage = nullptr;  // pinned host staging
};

namespace {

int release(fdb_upwind* h) {
  if (!h) return FDB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void*& b : h->buf)
    if (b) (void)hipFree(b);
  if (h->partials) (void)hipFree(h->partials);
  if (h->stage) (void)hipHostFree(h->stage);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FDB_OK;
}

// Streams the current field to the host in chunks, calling fn(values,
// count, first_index) with each chunk widened to double.
template <typename Fn>
int for_each_chunk(fdb_upwind* h, Fn fn) {
  const uint64_t per = kStageBytes / h->esz;
  std::vector<double> wide(std::min<uint64_t>(per, h->ntot));
  const char* src = (const char*)h->buf[h->cur];
  for (uint64_t i0 = 0; i0 < h->ntot; i0 += per) {
    const uint64_t cnt = std::min<uint64_t>(per, h->ntot - i0);
    FDB_HIP(hipMemcpyAsync(h->stage, src + i0 * h->esz, cnt * h->esz, hipMemcpyDeviceToHost,
                           h->stream));
    FDB_HIP(hipStreamSynchronize(h->stream));
    h_convert(h->dtype, h->stage, FDB_F64, wide.data(), cnt);
    fn(wide.data(), cnt, i0);
  }
  return FDB_OK;
}

int run_advect(fdb_upwind* h, int nsteps, double dt, int* launches) {
  fdb::Grid3 g;
  if (!fdb::make_grid(h->ndims, h->n, h->v, h->len, dt, &g))
    return fail(FDB_ERR_INVALID, "advect: invalid geometry");
  int total = 0;
  int left = nsteps;
  while (left > 0) {
    const int k = h->time_block == 0 ? fdb::plan_pass(h->dtype, g, left, fdb::kMaxTimeBlock)
                                     : std::min(left, h->time_block);
    int l = 0;
    FDB_HIP(fdb::launch_advect_pass(h->dtype, g, h->buf, &h->cur, k, h->stream, &l,
                                    h->steps_done, h->seed));
    h->steps_done += (uint64_t)k;
    total += l;
    left -= k;
  }
  if (launches) *launches = total;
  return FDB_OK;
}

}  // namespace

extern "C" {

const char* fdb_last_error(void) { return g_err.c_str(); }

int fdb_version(void) { return 100; }

int fdb_upwind_create(int ndims, const double* velocity, const double* lengths,
                      const uint64_t* num_cells, int dtype, fdb_upwind_t** out) {
  if (!out) return fail(FDB_ERR_INVALID, "create: out is null");
  *out = nullptr;
  if (ndims < 1 || ndims > 3) return fail(FDB_ERR_INVALID, "create: ndims must be 1, 2 or 3");
  if (!velocity || !lengths || !num_cells) return fail(FDB_ERR_INVALID, "create: null array");
  if (!valid_dtype(dtype)) return fail(FDB_ERR_INVALID, "create: unknown dtype");
  uint64_t ntot = 1;
  for (int j = 0; j < ndims; ++j) {
    if (num_cells[j] == 0) return fail(FDB_ERR_INVALID, "create: numCells must be positive");
    if (ntot > (uint64_t)INT64_MAX / num_cells[j])
      return fail(FDB_ERR_INVALID, "create: too many cells");
    ntot *= num_cells[j];
  }
  if (!fdb::fits_bytes(ntot, fdb::dtype_size(dtype)))
    return fail(FDB_ERR_INVALID, "create: field size in bytes overflows");
  if (int rc = check_tuning()) return rc;
  if (int rc = check_device()) return rc;
  std::unique_ptr<fdb_upwind, int (*)(fdb_upwind*)> h(new fdb_upwind, release);
  h->ndims = ndims;
  for (int j = 0; j < ndims; ++j) {
    h->n[j] = num_cells[j];
    h->v[j] = velocity[j];
    h->len[j] = lengths[j];
  }
  h->dtype = dtype;
  h->esz = fdb::dtype_size(dtype);
  h->ntot = ntot;
  FDB_HIP(hipStreamCreate(&h->stream));
  FDB_HIP(hipEventCreate(&h->ev0));
  FDB_HIP(hipEventCreate(&h->ev1));
  for (void*& b : h->buf) FDB_HIP(hipMalloc(&b, ntot * h->esz));
  FDB_HIP(hipMalloc((void**)&h->partials, (6 * fdb::kReduceBlocks + 6) * sizeof(double)));
  FDB_HIP(hipHostMalloc(&h->stage, kStageBytes));
  // The reference ctor: zero field, f[0] = 1 (upwind.cxx:45-48).
  fdb::Grid3 g;
  fdb::make_grid(ndims, h->n, h->v, h->len, 0.0, &g);
  FDB_HIP(fdb::launch_init(dtype, g, h->len, FDB_INIT_DELTA, nullptr, 0, 0, h->buf[0],
                           h->stream));
  FDB_HIP(hipStreamSynchronize(h->stream));
  *out = h.release();
  return FDB_OK;
}

int fdb_upwind_destroy(fdb_upwind_t* h) { return release(h); }

int fdb_upwind_advect(fdb_upwind_t* h, int num_time_steps, double delta_time) {
  if (!h) return fail(FDB_ERR_INVALID, "advect: null handle");
  if (num_time_steps <= 0) return FDB_OK;  // the reference's loop runs no step
  if (int rc = run_advect(h, num_time_steps, delta_time, nullptr)) return rc;
  FDB_HIP(hipStreamSynchronize(h->stream));
  return FDB_OK;
}

int fdb_upwind_advect_timed(fdb_upwind_t* h, int num_time_steps, double delta_time, float* ms,
                            int* launches) {
  if (!h) return fail(FDB_ERR_INVALID, "advect: null handle");
  int l = 0;
  FDB_HIP(hipEventRecord(h->ev0, h->stream));
  if (num_time_steps > 0)
    if (int rc = run_advect(h, num_time_steps, delta_time, &l)) return rc;
  FDB_HIP(hipEventRecord(h->ev1, h->stream));
  FDB_HIP(hipEventSynchronize(h->ev1));
  float t = 0.f;
  FDB_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
  if (ms) *ms = t;
  if (launches) *launches = l;
  return FDB_OK;
}

int fdb_upwind_checksum(fdb_upwind_t* h, double* out) {
  if (!h || !out) return fail(FDB_ERR_INVALID, "checksum: null argument");
  double s = 0.0;  // std::accumulate(f.begin(), f.end(), 0.0), upwind.cxx:92
  int rc = for_each_chunk(h, [&](const double* f, uint64_t n, uint64_t) {
    for (uint64_t i = 0; i < n; ++i) s = s + f[i];
  });
  if (rc) return rc;
  *out = s;
  return FDB_OK;
}

int fdb_upwind_std(fdb_upwind_t* h, double* out) {
  if (!h || !out) return fail(FDB_ERR_INVALID, "std: null argument");
  double cs = 0.0;
  if (int rc = fdb_upwind_checksum(h, &cs)) return rc;
  const double mean = cs / (double)h->ntot;  // upwind.cxx:96
  double res = 0;
  int rc = for_each_chunk(h, [&](const double* f, uint64_t n, uint64_t) {
    for (uint64_t i = 0; i < n; ++i) {
      const double d = f[i] - mean;
      res += d * d;
    }
  });
  if (rc) return rc;
  *out = std::sqrt(res / (double)h->ntot);  // upwind.cxx:102
  return FDB_OK;
}

int fdb_upwind_print(fdb_upwind_t* h) {
  if (!h) return fail(FDB_ERR_INVALID, "print: null handle");
  int rc = for_each_chunk(h, [&](const double* f, uint64_t n, uint64_t i0) {
    for (uint64_t i = 0; i < n; ++i) std::cout << (size_t)(i0 + i) << " " << f[i] << '\n';
  });
  std::cout.flush();
  return rc;
}

int fdb_upwind_save_vtk(fdb_upwind_t* h, const char* filename) {
  if (!h || !filename) return fail(FDB_ERR_INVALID, "saveVTK: null argument");
  std::vector<double> f(h->ntot);
  int rc = for_each_chunk(h, [&](const double* src, uint64_t n, uint64_t i0) {
    std::memcpy(f.data() + i0, src, n * sizeof(double));
  });
  if (rc) return rc;
  return fdb_host_save_vtk(h->ndims, h->n, h->len, f.data(), filename);
}

// Raw field file: this 64-byte header, then the field's ntot values in the
// handle's storage type (FDB_BF16_SR as bf16 bits), little-endian,
// row-major with the last axis contiguous -- numpy.fromfile(name, dtype,
// offset=64) reads it.  steps_done and seed let a stochastic-rounding run
// resume with the same random bits.
struct RawHeader {
  char magic[8];  // "FDBFLD01"
  uint32_t ndims, dtype;
  uint64_t n[3];  // axes 0..ndims-1; the rest 1
  uint64_t steps_done, seed;
  uint64_t reserved;
};
static_assert(sizeof(RawHeader) == 64, "raw header is 64 bytes");
constexpr char kRawMagic[8] = {'F', 'D', 'B', 'F', 'L', 'D', '0', '1'};

int fdb_upwind_save_raw(fdb_upwind_t* h, const char* filename) {
  if (!h || !filename) return fail(FDB_ERR_INVALID, "save_raw: null argument");
  std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(filename, "wb"), &std::fclose);
  if (!fp) return fail(FDB_ERR_IO, std::string("save_raw: cannot open ") + filename);
  RawHeader hd{};
  std::memcpy(hd.magic, kRawMagic, sizeof(kRawMagic));
  hd.ndims = (uint32_t)h->ndims;
  hd.dtype = (uint32_t)h->dtype;
  for (int j = 0; j < 3; ++j) hd.n[j] = j < h->ndims ? h->n[j] : 1;
  hd.steps_done = h->steps_done;
  hd.seed = h->seed;
  if (std::fwrite(&hd, sizeof(hd), 1, fp.get()) != 1)
    return fail(FDB_ERR_IO, "save_raw: write failed");
  const uint64_t per = kStageBytes / h->esz;
  const char* src = (const char*)h->buf[h->cur];
  for (uint64_t i0 = 0; i0 < h->ntot; i0 += per) {
    const uint64_t cnt = std::min<uint64_t>(per, h->ntot - i0);
    FDB_HIP(hipMemcpyAsync(h->stage, src + i0 * h->esz, cnt * h->esz, hipMemcpyDeviceToHost,
                           h->stream));
    FDB_HIP(hipStreamSynchronize(h->stream));
    if (std::fwrite(h->stage, h->esz, cnt, fp.get()) != cnt)
      return fail(FDB_ERR_IO, "save_raw: write failed");
  }
  if (std::fflush(fp.get()) != 0) return fail(FDB_ERR_IO, "save_raw: write failed");
  return FDB_OK;
}

namespace {

// Opens a raw field file and checks its header: the magic, a valid ndims
// and dtype, axes of at least one cell whose product times the item size
// fits 63 bits, and a file size of exactly header + payload.  On success
// the file is positioned at the payload.
int open_raw(const char* filename, const char* who, RawHeader* hd,
             std::unique_ptr<FILE, int (*)(FILE*)>* out) {
  std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(filename, "rb"), &std::fclose);
  if (!fp) return fail(FDB_ERR_IO, std::string(who) + ": cannot open " + filename);
  if (std::fread(hd, sizeof(*hd), 1, fp.get()) != 1 ||
      std::memcmp(hd->magic, kRawMagic, sizeof(kRawMagic)) != 0)
    return fail(FDB_ERR_IO, std::string(who) + ": not a raw field file");
  if (hd->ndims < 1 || hd->ndims > 3 || !valid_dtype((int)hd->dtype))
    return fail(FDB_ERR_IO, std::string(who) + ": bad ndims or dtype in the header");
  uint64_t ntot = 1;
  for (int j = 0; j < 3; ++j) {
    if (hd->n[j] == 0 || (j >= (int)hd->ndims && hd->n[j] != 1) ||
        ntot > (uint64_t)INT64_MAX / hd->n[j])
      return fail(FDB_ERR_IO, std::string(who) + ": bad axes in the header");
    ntot *= hd->n[j];
  }
  const size_t esz = fdb::dtype_size((int)hd->dtype);
  if (ntot > ((uint64_t)INT64_MAX - sizeof(*hd)) / esz)
    return fail(FDB_ERR_IO, std::string(who) + ": bad axes in the header");
  // The whole payload must be there (and nothing after it).
  const uint64_t want = sizeof(*hd) + ntot * esz;
  if (std::fseek(fp.get(), 0, SEEK_END) != 0 || (uint64_t)std::ftell(fp.get()) != want ||
      std::fseek(fp.get(), (long)sizeof(*hd), SEEK_SET) != 0)
    return fail(FDB_ERR_IO, std::string(who) + ": file size does not match its header");
  *out = std::move(fp);
  return FDB_OK;
}

}  // namespace

int fdb_host_raw_info(const char* filename, int* ndims, int* dtype, uint64_t* num_cells,
                      uint64_t* steps_done, uint64_t* seed) {
  if (!filename) return fail(FDB_ERR_INVALID, "raw_info: null filename");
  RawHeader hd{};
  std::unique_ptr<FILE, int (*)(FILE*)> fp(nullptr, &std::fclose);
  if (int rc = open_raw(filename, "raw_info", &hd, &fp)) return rc;
  if (ndims) *ndims = (int)hd.ndims;
  if (dtype) *dtype = (int)hd.dtype;
  if (num_cells)
    for (int j = 0; j < (int)hd.ndims; ++j) num_cells[j] = hd.n[j];
  if (steps_done) *steps_done = hd.steps_done;
  if (seed) *seed = hd.seed;
  return FDB_OK;
}

int fdb_upwind_load_raw(fdb_upwind_t* h, const char* filename) {
  if (!h || !filename) return fail(FDB_ERR_INVALID, "load_raw: null argument");
  RawHeader hd{};
  std::unique_ptr<FILE, int (*)(FILE*)> fp(nullptr, &std::fclose);
  if (int rc = open_raw(filename, "load_raw", &hd, &fp)) return rc;
  bool same = hd.ndims == (uint32_t)h->ndims && hd.dtype == (uint32_t)h->dtype;
  for (int j = 0; j < 3 && same; ++j) same = hd.n[j] == (j < h->ndims ? h->n[j] : 1);
  if (!same)
    return fail(FDB_ERR_INVALID, "load_raw: the file's shape or storage type differs");
  const uint64_t per = kStageBytes / h->esz;
  char* dst = (char*)h->buf[h->cur];
  for (uint64_t i0 = 0; i0 < h->ntot; i0 += per) {
    const uint64_t cnt = std::min<uint64_t>(per, h->ntot - i0);
    FDB_HIP(hipStreamSynchronize(h->stream));  // staging buffer free
    if (std::fread(h->stage, h->esz, cnt, fp.get()) != cnt)
      return fail(FDB_ERR_IO, "load_raw: read failed");
    FDB_HIP(hipMemcpyAsync(dst + i0 * h->esz, h->stage, cnt * h->esz, hipMemcpyHostToDevice,
                           h->stream));
  }
  FDB_HIP(hipStreamSynchronize(h->stream));
  h->steps_done = hd.steps_done;
  h->seed = hd.seed;
  return FDB_OK;
}

int fdb_upwind_ntot(const fdb_upwind_t* h, uint64_t* out) {
  if (!h || !out) return fail(FDB_ERR_INVALID, "ntot: null argument");
  *out = h->ntot;
  return FDB_OK;
}

int fdb_upwind_dtype(const fdb_upwind_t* h, int* out) {
  if (!h || !out) return fail(FDB_ERR_INVALID, "dtype: null argument");
  *out = h->dtype;
  return FDB_OK;
}

int fdb_upwind_set_field(fdb_upwind_t* h, const void* host, int host_dtype) {
  if (!h || !host) return fail(FDB_ERR_INVALID, "set_field: null argument");
  if (!valid_dtype(host_dtype)) return fail(FDB_ERR_INVALID, "set_field: unknown dtype");
  const size_t hs = fdb::dtype_size(host_dtype);
  const uint64_t per = kStageBytes / std::max(h->esz, hs);
  char* dst = (char*)h->buf[h->cur];
  for (uint64_t i0 = 0; i0 < h->ntot; i0 += per) {
    const uint64_t cnt = std::min<uint64_t>(per, h->ntot - i0);
    FDB_HIP(hipStreamSynchronize(h->stream));  // staging buffer free
    h_convert(host_dtype, (const char*)host + i0 * hs, h->dtype, h->stage, cnt);
    FDB_HIP(hipMemcpyAsync(dst + i0 * h->esz, h->stage, cnt * h->esz, hipMemcpyHostToDevice,
                           h->stream));
  }
  FDB_HIP(hipStreamSynchronize(h->stream));
  return FDB_OK;
}

int fdb_upwind_get_field(fdb_upwind_t* h, void* host, int host_dtype) {
  if (!h || !host) return fail(FDB_ERR_INVALID, "get_field: null argument");
  if (!valid_dtype(host_dtype)) return fail(FDB_ERR_INVALID, "get_field: unknown dtype");
  const size_t hs = fdb::dtype_size(host_dtype);
  const uint64_t per = kStageBytes / h->esz;
  const char* src = (const char*)h->buf[h->cur];
  for (uint64_t i0 = 0; i0 < h->ntot; i0 += per) {
    const uint64_t cnt = std::min<uint64_t>(per, h->ntot - i0);
    FDB_HIP(hipMemcpyAsync(h->stage, src + i0 * h->esz, cnt * h->esz, hipMemcpyDeviceToHost,
                           h->stream));
    FDB_HIP(hipStreamSynchronize(h->stream));
    h_convert(h->dtype, h->stage, host_dtype, (char*)host + i0 * hs, cnt);
  }
  return FDB_OK;
}

int fdb_upwind_field_device(fdb_upwind_t* h, void** out) {
  if (!h || !out) return fail(FDB_ERR_INVALID, "field_device: null argument");
  *out = h->buf[h->cur];
  return FDB_OK;
}

int fdb_upwind_init_field(fdb_upwind_t* h, int kind, const double* params, int nparams,
                          uint64_t seed) {
  if (!h) return fail(FDB_ERR_INVALID, "init_field: null handle");
  if (nparams > 0 && !params) return fail(FDB_ERR_INVALID, "init_field: null params");
  fdb::Grid3 g;
  fdb::make_grid(h->ndims, h->n, h->v, h->len, 0.0, &g);
  hipError_t e = fdb::launch_init(h->dtype, g, h->len, kind, params, nparams, seed,
                                  h->buf[h->cur], h->stream);
  if (e == hipErrorInvalidValue) return fail(FDB_ERR_INVALID, "init_field: bad kind/params");
  FDB_HIP(e);
  FDB_HIP(hipStreamSynchronize(h->stream));
  return FDB_OK;
}

int fdb_upwind_mass(fdb_upwind_t* h, double* out) {
  if (!h || !out) return fail(FDB_ERR_INVALID, "mass: null argument");
  double* out2 = h->partials + 6 * fdb::kReduceBlocks;
  FDB_HIP(fdb::launch_sum(h->dtype, h->buf[h->cur], h->ntot, h->partials, out2, h->stream));
  double r[2];
  FDB_HIP(hipMemcpyAsync(r, out2, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  FDB_HIP(hipStreamSynchronize(h->stream));
  *out = r[0] + r[1];
  return FDB_OK;
}

int fdb_upwind_set_seed(fdb_upwind_t* h, uint64_t seed) {
  if (!h) return fail(FDB_ERR_INVALID, "set_seed: null handle");
  h->seed = seed;
  h->steps_done = 0;
  return FDB_OK;
}

int fdb_upwind_set_time_block(fdb_upwind_t* h, int k) {
  if (!h) return fail(FDB_ERR_INVALID, "set_time_block: null handle");
  if (k < 0 || k > fdb::kMaxTimeBlock)
    return fail(FDB_ERR_INVALID, "set_time_block: k out of range (0..8)");
  h->time_block = k;
  return FDB_OK;
}

// ------------------------------------------------------- device pointers

int fdb_advect_dev_seeded(int ndims, const uint64_t* num_cells, const double* velocity,
                          const double* lengths, double delta_time, int dtype, void* a, void* b,
                          int nsteps, int time_block, uint64_t seed, uint64_t step0,
                          void* stream, int* result_in_b) {
  fdb::Grid3 g;
  if (!fdb::make_grid(ndims, num_cells, velocity, lengths, delta_time, &g))
    return fail(FDB_ERR_INVALID, "advect_dev: invalid geometry");
  if (!valid_dtype(dtype)) return fail(FDB_ERR_INVALID, "advect_dev: unknown dtype");
  if (!a || !b || a == b) return fail(FDB_ERR_INVALID, "advect_dev: need two distinct buffers");
  if (time_block < 0 || time_block > fdb::kMaxTimeBlock)
    return fail(FDB_ERR_INVALID, "advect_dev: time_block out of range (0..8)");
  if (int rc = check_tuning()) return rc;
  void* const bufs[2] = {a, b};
  int cur = 0;
  int left = nsteps;
  uint64_t stepno = step0;  // keys the stochastic rounding of each step
  while (left > 0) {
    const int k = time_block == 0 ? fdb::plan_pass(dtype, g, left, fdb::kMaxTimeBlock)
                                  : std::min(left, time_block);
    FDB_HIP(fdb::launch_advect_pass(dtype, g, bufs, &cur, k, (hipStream_t)stream, nullptr,
                                    stepno, seed));
    stepno += (uint64_t)k;
    left -= k;
  }
  if (result_in_b) *result_in_b = cur;
  return FDB_OK;
}

int fdb_advect_dev(int ndims, const uint64_t* num_cells, const double* velocity,
                   const double* lengths, double delta_time, int dtype, void* a, void* b,
                   int nsteps, int time_block, void* stream, int* result_in_b) {
  return fdb_advect_dev_seeded(ndims, num_cells, velocity, lengths, delta_time, dtype, a, b,
                               nsteps, time_block, 0, 0, stream, result_in_b);
}

int fdb_plan_passes(int ndims, const uint64_t* num_cells, const double* velocity,
                    const double* lengths, double delta_time, int dtype, int nsteps,
                    int time_block, int* ks, int max_ks, int* count) {
  fdb::Grid3 g;
  if (!fdb::make_grid(ndims, num_cells, velocity, lengths, delta_time, &g))
    return fail(FDB_ERR_INVALID, "plan_passes: invalid geometry");
  if (!valid_dtype(dtype)) return fail(FDB_ERR_INVALID, "plan_passes: unknown dtype");
  if (time_block < 0 || time_block > fdb::kMaxTimeBlock)
    return fail(FDB_ERR_INVALID, "plan_passes: time_block out of range (0..8)");
  if (nsteps < 0 || !count || (max_ks > 0 && !ks))
    return fail(FDB_ERR_INVALID, "plan_passes: invalid arguments");
  if (int rc = check_tuning()) return rc;
  int n = 0;
  for (int left = nsteps; left > 0;) {
    int k = time_block == 0 ? fdb::plan_pass(dtype, g, left, fdb::kMaxTimeBlock)
                            : std::min(left, time_block);
    left -= k;
    while (k > 0) {  // the launches of one pass: deepest supported depth first
      int K = 1;
      for (int c : {8, 7, 6, 4, 3, 2})
        if (c <= k && fdb::fused_pass_supported(dtype, g, c)) {
          K = c;
          break;
        }
      if (n < max_ks) ks[n] = K;
      ++n;
      k -= K;
    }
  }
  *count = n;
  return FDB_OK;
}

int fdb_adjoint_step_dev(int ndims, const uint64_t* num_cells, const double* velocity,
                         const double* lengths, double delta_time, int lam_dtype,
                         const void* lam_next, void* lam, int f_dtype, const void* f_t,
                         double* vgrad_acc, void* stream) {
  fdb::Grid3 g;
  if (!fdb::make_grid(ndims, num_cells, velocity, lengths, delta_time, &g))
    return fail(FDB_ERR_INVALID, "adjoint_step_dev: invalid geometry");
  if (lam_dtype != FDB_F64 && lam_dtype != FDB_F32 && lam_dtype != FDB_BF16)
    return fail(FDB_ERR_INVALID, "adjoint_step_dev: lam dtype must be F64, F32 or BF16");
  if (f_t && !valid_dtype(f_dtype)) return fail(FDB_ERR_INVALID, "adjoint_step_dev: f dtype");
  if (!lam_next || !lam || lam_next == lam)
    return fail(FDB_ERR_INVALID, "adjoint_step_dev: need two distinct buffers");
  if (f_t && !vgrad_acc) return fail(FDB_ERR_INVALID, "adjoint_step_dev: vgrad_acc is null");
  if (int rc = check_tuning()) return rc;
  double* part = nullptr;
  if (f_t) {
    const int64_t nb = fdb::adjoint_blocks(lam_dtype, f_dtype, g, lam_next, lam, f_t);
    if (int rc = dev_scratch((size_t)nb * 6 * sizeof(double), &part)) return rc;
  }
  hipError_t e = fdb::launch_adjoint_step(lam_dtype, f_t ? f_dtype : lam_dtype, g, lam_next, lam,
                                          f_t, part, vgrad_acc, (hipStream_t)stream);
  FDB_HIP(e);
  if (f_t) FDB_HIP(hipStreamSynchronize((hipStream_t)stream));  // scratch reuse
  return FDB_OK;
}

int fdb_sum_dev(int dtype, const void* x, uint64_t n, double* out, void* stream) {
  if (!valid_dtype(dtype) || !x || !out) return fail(FDB_ERR_INVALID, "sum_dev: bad argument");
  double* part = nullptr;
  if (int rc = dev_scratch((6 * fdb::kReduceBlocks + 6) * sizeof(double), &part)) return rc;
  double* out2 = part + 6 * fdb::kReduceBlocks;
  FDB_HIP(fdb::launch_sum(dtype, x, n, part, out2, (hipStream_t)stream));
  double r[2];
  FDB_HIP(hipMemcpyAsync(r, out2, sizeof(r), hipMemcpyDeviceToHost, (hipStream_t)stream));
  FDB_HIP(hipStreamSynchronize((hipStream_t)stream));
  *out = r[0] + r[1];
  return FDB_OK;
}

int fdb_misfit_dev(int f_dtype, const void* f, const void* target, uint64_t n, int lam_dtype,
                   void* lam, double* loss, void* stream) {
  if (!valid_dtype(f_dtype) ||
      (lam_dtype != FDB_F64 && lam_dtype != FDB_F32 && lam_dtype != FDB_BF16) || !f || !target ||
      !lam || !loss)
    return fail(FDB_ERR_INVALID, "misfit_dev: bad argument");
  double* part = nullptr;
  if (int rc = dev_scratch((6 * fdb::kReduceBlocks + 6) * sizeof(double), &part)) return rc;
  double* out2 = part + 6 * fdb::kReduceBlocks;
  FDB_HIP(fdb::launch_misfit(f_dtype, f, target, n, lam_dtype, lam, part, out2,
                             (hipStream_t)stream));
  double r[2];
  FDB_HIP(hipMemcpyAsync(r, out2, sizeof(r), hipMemcpyDeviceToHost, (hipStream_t)stream));
  FDB_HIP(hipStreamSynchronize((hipStream_t)stream));
  *loss = r[0] + r[1];
  return FDB_OK;
}

int fdb_compare_dev(int a_dtype, const void* a, int b_dtype, const void* b, uint64_t n,
                    double* out4, void* stream) {
  if (!valid_dtype(a_dtype) || !valid_dtype(b_dtype) || !a || !b || !out4)
    return fail(FDB_ERR_INVALID, "compare_dev: bad argument");
  double* part = nullptr;
  if (int rc = dev_scratch((6 * fdb::kReduceBlocks + 6) * sizeof(double), &part)) return rc;
  FDB_HIP(fdb::launch_compare(a_dtype, a, b_dtype, b, n, part, (hipStream_t)stream));
  std::vector<double> h(6 * (size_t)fdb::kReduceBlocks);
  FDB_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                         (hipStream_t)stream));
  FDB_HIP(hipStreamSynchronize((hipStream_t)stream));
  double mx = 0.0, s = 0.0, c = 0.0, mb = 0.0;
  for (int i = 0; i < fdb::kReduceBlocks; ++i) {
    const double* p = &h[6 * (size_t)i];
    mx = std::max(mx, p[0]);
    mb = std::max(mb, p[3]);
    const double sum = s + p[1];
    const double bb = sum - s;
    c += (s - (sum - bb)) + (p[1] - bb) + p[2];
    s = sum;
  }
  out4[0] = mx;
  out4[1] = n ? std::sqrt((s + c) / (double)n) : 0.0;
  out4[2] = mb;
  out4[3] = mb > 0.0 ? mx / mb : 0.0;
  return FDB_OK;
}

int fdb_init_field_dev(int ndims, const uint64_t* num_cells, const double* lengths, int dtype,
                       int kind, const double* params, int nparams, uint64_t seed, void* f,
                       void* stream) {
  double zeros[3] = {0, 0, 0};
  fdb::Grid3 g;
  if (!fdb::make_grid(ndims, num_cells, zeros, lengths, 0.0, &g))
    return fail(FDB_ERR_INVALID, "init_field_dev: invalid geometry");
  if (!valid_dtype(dtype) || !f) return fail(FDB_ERR_INVALID, "init_field_dev: bad argument");
  if (nparams > 0 && !params) return fail(FDB_ERR_INVALID, "init_field_dev: null params");
  hipError_t e =
      fdb::launch_init(dtype, g, lengths, kind, params, nparams, seed, f, (hipStream_t)stream);
  if (e == hipErrorInvalidValue) return fail(FDB_ERR_INVALID, "init_field_dev: bad kind/params");
  FDB_HIP(e);
  return FDB_OK;
}

// ------------------------------------------------------------ host helpers

double fdb_host_checksum(const double* f, uint64_t n) {
  double s = 0.0;
  for (uint64_t i = 0; i < n; ++i) s = s + f[i];
  return s;
}

double fdb_host_std(const double* f, uint64_t n) {
  const double mean = fdb_host_checksum(f, n) / (double)n;
  double res = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const double d = f[i] - mean;
    res += d * d;
  }
  return std::sqrt(res / (double)n);
}

// The legacy-VTK text layout of saveVTK.h:1-60: VTK wants the first
// dimension fastest, so the reference's axes are written in reverse
// (last axis as X); missing axes collapse to one point at 0.0; values go
// through the default ostream formatting (6 significant digits), ten per
// line.
int fdb_host_save_vtk(int ndims, const uint64_t* num_cells, const double* lengths,
                      const double* f, const char* filename) {
  if (ndims < 1 || ndims > 3 || !num_cells || !lengths || !f || !filename)
    return fail(FDB_ERR_INVALID, "save_vtk: bad argument");
  std::ofstream out(filename, std::ios_base::out);
  if (!out) return fail(FDB_ERR_IO, std::string("save_vtk: cannot open ") + filename);
  uint64_t ntot = 1;
  double dx[3];
  for (int j = 0; j < ndims; ++j) {
    ntot *= num_cells[j];
    dx[j] = lengths[j] / (double)num_cells[j];
  }
  out << "# vtk DataFile Version 2.0\n" << "upwind.cxx\n" << "ASCII\n"
      << "DATASET RECTILINEAR_GRID\n" << "DIMENSIONS";
  out << ' ';
  if (ndims > 2) out << (size_t)(num_cells[2] + 1); else out << 1;
  out << ' ';
  if (ndims > 1) out << (size_t)(num_cells[1] + 1); else out << 1;
  out << ' ' << (size_t)(num_cells[0] + 1);
  auto axis = [&](const char* name, int j, bool present) {
    out << '\n' << name << ' ';
    if (present) {
      out << (size_t)(num_cells[j] + 1) << " double\n";
      for (uint64_t i = 0; i < num_cells[j] + 1; ++i) out << ' ' << 0.0 + dx[j] * (double)i;
    } else {
      out << "1 double\n" << "0.0\n";
    }
  };
  axis("X_COORDINATES", 2, ndims > 2);
  axis("Y_COORDINATES", 1, ndims > 1);
  axis("Z_COORDINATES", 0, true);
  out << "\nCELL_DATA " << (size_t)ntot << '\n' << "SCALARS f double 1\n"
      << "LOOKUP_TABLE default\n";
  for (uint64_t i = 0; i < ntot; ++i) {
    out << f[i] << " ";
    if ((i + 1) % 10 == 0) out << '\n';
  }
  out << '\n';
  out.close();
  if (!out) return fail(FDB_ERR_IO, std::string("save_vtk: write failed for ") + filename);
  return FDB_OK;
}

}  // extern "C"
