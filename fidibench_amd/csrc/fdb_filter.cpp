// fdb_filter.cpp -- host side of the C ABI for the generic periodic stencil
// filter (include/fidibench_amd/fdb_filter.h): the one-process form of
// pletzer/fidibench cxx/Filter.cpp.  The kernels are in fdb_filter_kernels.hip.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/fidibench_amd/fdb_filter.h"
#include "fdb_internal.h"

struct fdb_filter {
  int nd = 0;
  uint64_t dims[3] = {1, 1, 1};
  double xmins[3] = {0, 0, 0}, xmaxs[3] = {1, 1, 1};
  fdb::FilterGeom g{};
  fdb::FilterPlan plan{};  // kernel and tiling, fixed at creation
  double* buf[2] = {nullptr, nullptr};
  int in = 0;              // buf[in] holds the input
  bool out_is_in = false;  // after copyOutToIn the output equals the input
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

int ffail(int code, const std::string& msg) {
  fdb::set_last_error(msg);
  return code;
}

int fhip(hipError_t e, const char* what) {
  fdb::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? FDB_ERR_NOMEM : FDB_ERR_HIP;
}

#define FLT_HIP(call)                               \
  do {                                              \
    hipError_t _e = (call);                         \
    if (_e != hipSuccess) return fhip(_e, #call);   \
  } while (0)

// Output buffer as the reference sees it.
const double* out_buf(const fdb_filter* h) { return h->out_is_in ? h->buf[h->in] : h->buf[1 - h->in]; }

// Before the input changes, give the output its own copy if it mirrors
// the input (the reference's output keeps its values).
int detach_output(fdb_filter* h) {
  if (!h->out_is_in) return FDB_OK;
  FLT_HIP(hipMemcpyAsync(h->buf[1 - h->in], h->buf[h->in], h->g.ntot * sizeof(double),
                         hipMemcpyDeviceToDevice, h->s));
  h->out_is_in = false;
  return FDB_OK;
}

void free_filter(fdb_filter* h) {
  if (!h) return;
  if (h->s) (void)hipStreamSynchronize(h->s);
  for (double* b : h->buf)
    if (b) (void)hipFree(b);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

}  // namespace

extern "C" {

int fdb_filter_create(int ndims, const uint64_t* dims, const double* xmins, const double* xmaxs,
                      int nterms, const int* offsets, const double* weights,
                      fdb_filter_t** out) {
  if (!out) return ffail(FDB_ERR_INVALID, "filter_create: out is null");
  *out = nullptr;
  if (ndims < 1 || ndims > 3 || !dims || !xmins || !xmaxs)
    return ffail(FDB_ERR_INVALID, "filter_create: ndims must be 1..3 with dims, xmins, xmaxs");
  if (nterms < 1 || nterms > fdb::kFilterMaxTerms || !offsets || !weights)
    return ffail(FDB_ERR_INVALID, "filter_create: 1..32 stencil terms with offsets and weights");
  fdb::FilterGeom g{};
  g.ntot = 1;
  for (int a = 0; a < 3; ++a) g.n[a] = 1;
  for (int a = 0; a < ndims; ++a) {
    if (dims[a] == 0 || dims[a] > (uint64_t)INT32_MAX)
      return ffail(FDB_ERR_INVALID, "filter_create: dims must be in 1..2^31-1");
    g.n[a] = (int64_t)dims[a];
    if (g.ntot > INT64_MAX / g.n[a]) return ffail(FDB_ERR_INVALID, "filter_create: too many cells");
    g.ntot *= g.n[a];
  }
  if (!fdb::fits_bytes((uint64_t)g.ntot, sizeof(double)))
    return ffail(FDB_ERR_INVALID, "filter_create: field size in bytes overflows");
  // std::map< std::vector<int>, double > order: lexicographic offsets; of
  // duplicate offsets the first inserted is kept (std::map::insert).
  std::vector<std::pair<std::vector<int>, double>> terms;
  for (int t = 0; t < nterms; ++t) {
    std::vector<int> o(offsets + (size_t)t * ndims, offsets + (size_t)(t + 1) * ndims);
    for (int v : o)
      if (v < -fdb::kFilterMaxOffset || v > fdb::kFilterMaxOffset)
        return ffail(FDB_ERR_INVALID, "filter_create: stencil offsets must be within +-8");
    bool dup = false;
    for (const auto& p : terms) dup = dup || p.first == o;
    if (!dup) terms.emplace_back(o, weights[t]);
  }
  std::stable_sort(terms.begin(), terms.end(),
                   [](const auto& x, const auto& y) { return x.first < y.first; });
  g.nt = (int)terms.size();
  for (int t = 0; t < g.nt; ++t) {
    for (int a = 0; a < 3; ++a) g.off[t][a] = a < ndims ? terms[t].first[a] : 0;
    g.w[t] = terms[t].second;
  }
  // (int)i + offset, then %= (size_t)n: a negative index becomes the
  // unsigned remainder of its two's-complement value.
  for (int a = 0; a < 3; ++a)
    for (int k = 1; k <= fdb::kFilterMaxOffset; ++k)
      g.neg[a][k - 1] = (int64_t)((uint64_t)(-(int64_t)k) % (uint64_t)g.n[a]);

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return ffail(FDB_ERR_NODEV, "filter_create: no GPU (this library has no CPU path)");
  fdb_filter* h = new fdb_filter;
  h->nd = ndims;
  for (int a = 0; a < ndims; ++a) {
    h->dims[a] = dims[a];
    h->xmins[a] = xmins[a];
    h->xmaxs[a] = xmaxs[a];
  }
  h->g = g;
  hipError_t e = fdb::plan_filter(g, &h->plan);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    e = hipMalloc((void**)&h->buf[i], g.ntot * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(h->buf[i], 0, g.ntot * sizeof(double), h->s);
  }
  if (e == hipSuccess) e = hipEventCreate(&h->e0);
  if (e == hipSuccess) e = hipEventCreate(&h->e1);
  if (e == hipSuccess) e = hipStreamSynchronize(h->s);
  if (e != hipSuccess) {
    const int rc = fhip(e, "filter_create");
    free_filter(h);
    return rc;
  }
  *out = h;
  return FDB_OK;
}

int fdb_filter_destroy(fdb_filter_t* h) {
  free_filter(h);
  return FDB_OK;
}

int fdb_filter_set_in_data(fdb_filter_t* h, const double* host) {
  if (!h || !host) return ffail(FDB_ERR_INVALID, "filter_set_in_data: null argument");
  if (int rc = detach_output(h)) return rc;
  FLT_HIP(hipMemcpyAsync(h->buf[h->in], host, h->g.ntot * sizeof(double), hipMemcpyHostToDevice,
                         h->s));
  FLT_HIP(hipStreamSynchronize(h->s));
  return FDB_OK;
}

int fdb_filter_set_in_sin(fdb_filter_t* h) {
  if (!h) return ffail(FDB_ERR_INVALID, "filter_set_in_sin: null handle");
  if (int rc = detach_output(h)) return rc;
  FLT_HIP(fdb::launch_filter_init_sin(h->g, h->nd, h->xmins, h->xmaxs, h->buf[h->in], h->s));
  FLT_HIP(hipStreamSynchronize(h->s));
  return FDB_OK;
}

int fdb_filter_apply(fdb_filter_t* h) {
  if (!h) return ffail(FDB_ERR_INVALID, "filter_apply: null handle");
  FLT_HIP(fdb::launch_filter(h->plan, h->buf[h->in], h->buf[1 - h->in], h->s));
  h->out_is_in = false;
  FLT_HIP(hipStreamSynchronize(h->s));
  return FDB_OK;
}

int fdb_filter_copy_out_to_in(fdb_filter_t* h) {
  if (!h) return ffail(FDB_ERR_INVALID, "filter_copy_out_to_in: null handle");
  if (!h->out_is_in) {
    h->in = 1 - h->in;
    h->out_is_in = true;
  }
  return FDB_OK;
}

int fdb_filter_iterate_timed(fdb_filter_t* h, int niters, float* ms) {
  if (!h || niters < 0) return ffail(FDB_ERR_INVALID, "filter_iterate_timed: bad argument");
  FLT_HIP(hipEventRecord(h->e0, h->s));
  for (int i = 0; i < niters; ++i) {
    FLT_HIP(fdb::launch_filter(h->plan, h->buf[h->in], h->buf[1 - h->in], h->s));
    h->in = 1 - h->in;  // copyOutToIn
    h->out_is_in = true;
  }
  FLT_HIP(hipEventRecord(h->e1, h->s));
  FLT_HIP(hipEventSynchronize(h->e1));
  float t = 0.f;
  FLT_HIP(hipEventElapsedTime(&t, h->e0, h->e1));
  if (ms) *ms = t;
  return FDB_OK;
}

int fdb_filter_get_data(fdb_filter_t* h, int which, double* host) {
  if (!h || !host || (which != 0 && which != 1))
    return ffail(FDB_ERR_INVALID, "filter_get_data: bad argument");
  const double* src = which == 0 ? h->buf[h->in] : out_buf(h);
  FLT_HIP(hipMemcpyAsync(host, src, h->g.ntot * sizeof(double), hipMemcpyDeviceToHost, h->s));
  FLT_HIP(hipStreamSynchronize(h->s));
  return FDB_OK;
}

int fdb_filter_kernel(const fdb_filter_t* h, const char** name) {
  if (!h || !name) return ffail(FDB_ERR_INVALID, "filter_kernel: null argument");
  switch (h->plan.kind) {
    case fdb::kFilterStar: *name = "star"; break;
    case fdb::kFilterPlanes: *name = "planes"; break;
    case fdb::kFilterPlanesMid: *name = "planes_mid"; break;
    case fdb::kFilterPlanesWide: *name = "planes_wide"; break;
    case fdb::kFilterRows: *name = "rows"; break;
    default: *name = "gather"; break;
  }
  return FDB_OK;
}

int fdb_filter_checksum(fdb_filter_t* h, int which, double* out) {
  if (!h || !out || (which != 0 && which != 1))
    return ffail(FDB_ERR_INVALID, "filter_checksum: bad argument");
  const double* src = which == 0 ? h->buf[h->in] : out_buf(h);
  // localSum += data[i] in order (Filter.cpp:470-480), chunk by chunk.
  const uint64_t chunk = (uint64_t)1 << 22;
  std::vector<double> tmp((size_t)std::min<uint64_t>(chunk, (uint64_t)h->g.ntot));
  double sum = 0;
  for (uint64_t i0 = 0; i0 < (uint64_t)h->g.ntot; i0 += chunk) {
    const uint64_t cnt = std::min<uint64_t>(chunk, (uint64_t)h->g.ntot - i0);
    FLT_HIP(hipMemcpyAsync(tmp.data(), src + i0, cnt * sizeof(double), hipMemcpyDeviceToHost,
                           h->s));
    FLT_HIP(hipStreamSynchronize(h->s));
    for (uint64_t i = 0; i < cnt; ++i) sum += tmp[i];
  }
  *out = sum;
  return FDB_OK;
}

int fdb_filter_save_vtk(fdb_filter_t* h, const char* filename) {
  if (!h || !filename) return ffail(FDB_ERR_INVALID, "filter_save_vtk: bad argument");
  std::vector<double> f((size_t)h->g.ntot);
  if (int rc = fdb_filter_get_data(h, 1, f.data())) return rc;
  return fdb_host_write_vtk(h->nd, h->dims, h->xmins, h->xmaxs, f.data(), "outData", filename);
}

int fdb_host_write_vtk(int ndims, const uint64_t* dims, const double* xmins,
                       const double* xmaxs, const double* field, const char* name,
                       const char* filename) {
  if (ndims < 1 || ndims > 3 || !dims || !xmins || !xmaxs || !field || !name || !filename)
    return ffail(FDB_ERR_INVALID, "write_vtk: bad argument");
  // writeVTK.cpp:19-95: missing axes get one cell spanning [0, 1].
  size_t cells[3] = {1, 1, 1}, nodes[3] = {2, 2, 2};
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  for (int a = 0; a < ndims; ++a) {
    cells[a] = (size_t)dims[a];
    nodes[a] = cells[a] + 1;
    lo[a] = xmins[a];
    hi[a] = xmaxs[a];
  }
  std::ofstream file(filename);
  if (!file) return ffail(FDB_ERR_IO, std::string("write_vtk: cannot open ") + filename);
  file << "# vtk DataFile Version 2.0\n";
  file << "produced by laplacian\n";
  file << "ASCII\n";
  file << "DATASET STRUCTURED_GRID\n";
  file << "DIMENSIONS " << nodes[0] << ' ' << nodes[1] << ' ' << nodes[2] << '\n';
  file << "POINTS " << nodes[0] * nodes[1] * nodes[2] << " float\n";
  // vertices, column-major (axis 0 fastest)
  for (size_t k = 0; k < nodes[2]; ++k)
    for (size_t j = 0; j < nodes[1]; ++j)
      for (size_t i = 0; i < nodes[0]; ++i) {
        const double x = lo[0] + (hi[0] - lo[0]) * i / double(cells[0]);
        const double y = lo[1] + (hi[1] - lo[1]) * j / double(cells[1]);
        const double z = lo[2] + (hi[2] - lo[2]) * k / double(cells[2]);
        file << x << ' ' << y << ' ' << z << '\n';
      }
  const size_t ncell = cells[0] * cells[1] * cells[2];
  file << "CELL_DATA " << ncell << '\n';
  file << "SCALARS " << name << " float\n";
  file << "LOOKUP_TABLE default\n";
  for (size_t i = 0; i < ncell; ++i) file << field[i] << '\n';
  file.close();
  if (!file) return ffail(FDB_ERR_IO, std::string("write_vtk: write failed on ") + filename);
  return FDB_OK;
}

}  // extern "C"
