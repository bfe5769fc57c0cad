// fdb_noalloc.hip -- fdb_adjoint_step_dev_noalloc: the adjoint step with
// velocity products for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// fdb_adjoint_step_dev keeps the per-block partial sums of the velocity
// products in a grow-only buffer from hipMalloc and synchronises the stream
// before returning (the buffer may be reallocated by the next call).  This
// variant keeps them in `scratch`, a file-scope device array of a fixed
// size that every GPU the library runs on holds until the process ends,
// and returns once the work is enqueued.  Its device address is looked up
// once per device with hipGetSymbolAddress (the array's name in host code
// is not a device address) and kept per device.
//
// All calls on one GPU share that GPU's array: calls ordered on one stream
// are safe (the reduction that reads a call's partials finishes before the
// next call's adjoint kernel writes them), calls that overlap on one GPU
// (different streams without ordering) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "fdb_internal.h"
#include "../../include/fidibench_amd/fdb_upwind.h"

namespace {

// Per-block partial slots (6 doubles each: three {sum, compensation}
// pairs): 3 MiB per device.
constexpr int64_t kScratchSlots = 65536;
__device__ __align__(256) double scratch[6 * kScratchSlots];

std::mutex g_mu;
std::vector<double*> g_scratch_ptr;  // per device id; nullptr until looked up

int nfail(int code, const std::string& msg) {
  fdb::set_last_error(msg);
  return code;
}

int nhip(hipError_t e, const char* what) {
  (void)hipGetLastError();
  return nfail(FDB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device address of the current device's copy of `scratch`.
int scratch_ptr(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nhip(e, "adjoint_step_dev_noalloc: hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return nhip(e, "adjoint_step_dev_noalloc: hipGetSymbolAddress");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return FDB_OK;
}

}  // namespace

extern "C" int fdb_adjoint_step_dev_noalloc(int ndims, const uint64_t* num_cells,
                                            const double* velocity, const double* lengths,
                                            double delta_time, int lam_dtype,
                                            const void* lam_next, void* lam, int f_dtype,
                                            const void* f_t, double* vgrad_acc, void* stream) {
  fdb::Grid3 g;
  if (!fdb::make_grid(ndims, num_cells, velocity, lengths, delta_time, &g))
    return nfail(FDB_ERR_INVALID, "adjoint_step_dev_noalloc: invalid geometry");
  if (lam_dtype != FDB_F64 && lam_dtype != FDB_F32 && lam_dtype != FDB_BF16)
    return nfail(FDB_ERR_INVALID, "adjoint_step_dev_noalloc: lam dtype must be F64, F32 or BF16");
  if (f_t && fdb::dtype_size(f_dtype) == 0)
    return nfail(FDB_ERR_INVALID, "adjoint_step_dev_noalloc: f dtype");
  if (!lam_next || !lam || lam_next == lam)
    return nfail(FDB_ERR_INVALID, "adjoint_step_dev_noalloc: need two distinct buffers");
  if (f_t && !vgrad_acc)
    return nfail(FDB_ERR_INVALID, "adjoint_step_dev_noalloc: vgrad_acc is null");
  if (const char* m = fdb::tuning_error()) return nfail(FDB_ERR_INVALID, m);
  double* part = nullptr;
  if (f_t) {
    const int64_t nb = fdb::adjoint_blocks(lam_dtype, f_dtype, g, lam_next, lam, f_t);
    if (nb <= 0 || nb > kScratchSlots)
      return nfail(FDB_ERR_INVALID,
                   "adjoint_step_dev_noalloc: this grid needs " + std::to_string(nb) +
                       " partial-sum slots; the fixed scratch holds " +
                       std::to_string(kScratchSlots) + " (use fdb_adjoint_step_dev)");
    if (int rc = scratch_ptr(&part)) return rc;
  }
  const hipError_t e = fdb::launch_adjoint_step(lam_dtype, f_t ? f_dtype : lam_dtype, g, lam_next,
                                                lam, f_t, part, vgrad_acc, (hipStream_t)stream);
  if (e != hipSuccess) return nhip(e, "adjoint_step_dev_noalloc");
  return FDB_OK;
}
