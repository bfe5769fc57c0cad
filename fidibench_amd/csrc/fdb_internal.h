// fdb_internal.h -- declarations shared by the HIP kernels (fdb_kernels.hip)
// and the host side of the C ABI (fdb_capi.cpp).  Not installed.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <string>

namespace fdb {

struct Grid3;

// Thread-local message returned by fdb_last_error().
void set_last_error(const std::string& msg);

// Number of blocks the adjoint kernel launches for these buffers (sizes
// its per-block partial sums).
int64_t adjoint_blocks(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                       void* lam, const void* f_t);

// Geometry of one advect call, in the padded 3-D form every kernel uses:
// an NDIMS < 3 problem is placed on the trailing axes and the missing
// leading axes get extent 1 and act = 0 (no term), so the arithmetic of
// Upwind<1> and Upwind<2> is reproduced exactly.
struct Grid3 {
  int64_t n[3];   // extents, axis 0 slowest, axis 2 contiguous
  int up[3];      // upDirection: -1 if v >= 0 else +1 (upwind.cxx:34-35)
  int act[3];     // 1 if the axis is one of the reference's NDIMS axes
  double c[3];    // ((dt*v)*u)/dx, upwind.cxx:72
  double dc[3];   // d c / d v = (dt*u)/dx with u fixed
  int64_t ntot;
};

// Builds Grid3 from the reference's ctor arguments (upwind.cxx:23-49)
// and deltaTime; returns false on invalid sizes.
bool make_grid(int ndims, const uint64_t* n, const double* v, const double* len, double dt,
               Grid3* g);

// Bytes per stored cell for a storage dtype (FDB_F64/F32/BF16), 0 if bad.
size_t dtype_size(int dtype);

// Largest temporal-blocking depth the fused multi-step kernel supports.
constexpr int kMaxTimeBlock = 8;

// Advances `k` steps starting from bufs[*cur] (two distinct buffers of
// the field's size); on return *cur indexes the buffer holding the
// result and the other buffer's contents are undefined.  k == 1 is one
// launch of the streaming kernel; k > 1 runs the temporally blocked
// kernel when the shape fits it, else k streaming launches.  *launches
// (optional) receives the number of kernel launches.
// For FDB_BF16_SR the stochastic rounding of step number step0 + i uses
// a key hashed from (seed, step0 + i).
hipError_t launch_advect_pass(int dtype, const Grid3& g, void* const bufs[2], int* cur, int k,
                              hipStream_t s, int* launches, uint64_t step0 = 0,
                              uint64_t seed = 0);

// One launch of k >= 2 fused steps from src into dst (distinct buffers);
// src is not written.  hipErrorNotSupported when no single launch does k
// steps for this storage type and shape (the caller then steps one by one).
hipError_t launch_fused_pass(int dtype, const Grid3& g, const void* src, void* dst, int k,
                             hipStream_t s);

// The storage layout of a dtype (FDB_BF16_SR is stored as FDB_BF16).
inline int storage_dtype(int dtype) { return dtype == 3 ? 2 : dtype; }

// ---- generic periodic stencil (cxx/Filter.cpp), kernels in fdb_filter.hip

constexpr int kFilterMaxTerms = 32;   // stencil terms
constexpr int kFilterMaxOffset = 8;   // |offset| per axis

// Geometry of a stencil filter.  Axes are the reference's: axis 0 is the
// contiguous one (Filter's MultiArrayIter is column-major); missing axes
// have extent 1 and offset 0.  The terms are in std::map order of their
// offset vectors.  neg[a][k-1] is where index -k lands along axis a under
// the reference's wrap, (size_t)(-k) % n[a] (cxx/Filter.cpp:231-233).
struct FilterGeom {
  int64_t n[3];
  int64_t ntot;
  int nt;
  int off[kFilterMaxTerms][3];
  double w[kFilterMaxTerms];
  int64_t neg[3][kFilterMaxOffset];
};

// out[i] = sum over terms, in order, of w * in[wrapped neighbour].
hipError_t launch_filter(const FilterGeom& g, const double* in, double* out, hipStream_t s);

// in[i] = prod_a sin(2 pi x_a), x_a = xmin_a + (i_a + 0.5) * (xmax_a - xmin_a) / n_a
// (the input function of laplacian/cxx/laplacian.cxx:22-28), with the
// device's sin (not bit-identical to the host's libm).
hipError_t launch_filter_init_sin(const FilterGeom& g, int ndims, const double* xmins,
                                  const double* xmaxs, double* in, hipStream_t s);

// lam = A^T lam_next; if f_t != nullptr also adds the three velocity
// gradient dot products into acc (device, 3 x {sum, comp}).
hipError_t launch_adjoint_step(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                               void* lam, const void* f_t, double* partials, double* acc,
                               hipStream_t s);

// Number of per-block partial slots the reductions need (each slot is
// 6 doubles: up to three {sum, comp} pairs).
constexpr int kReduceBlocks = 2048;

// Compensated FP64 sum of n values of dtype; result {sum, comp} at out2.
hipError_t launch_sum(int dtype, const void* x, uint64_t n, double* partials, double* out2,
                      hipStream_t s);

// lam = f - target (lam_dtype); {sum, comp} of 0.5*(f-target)^2 at out2.
hipError_t launch_misfit(int f_dtype, const void* f, const void* target, uint64_t n,
                         int lam_dtype, void* lam, double* partials, double* out2,
                         hipStream_t s);

// Synthetic initial field (see fdb_init_kind).  params are copied by value.
hipError_t launch_init(int dtype, const Grid3& g, const double* len, int kind,
                       const double* params, int nparams, uint64_t seed, void* f, hipStream_t s);

// Element-wise dtype conversion on the device (n values).
hipError_t launch_convert(int src_dtype, const void* src, int dst_dtype, void* dst, uint64_t n,
                          hipStream_t s);

// Per-block {max|a-b|, sum, comp of sum (a-b)^2, max|b|} at
// partials[block*6 + 0..3] for kReduceBlocks blocks.
hipError_t launch_compare(int a_dtype, const void* a, int b_dtype, const void* b, uint64_t n,
                          double* partials, hipStream_t s);

// Zero the 6-double accumulator with a kernel (no memset nodes, so the
// call stays graph-capturable).
hipError_t launch_zero(double* p, int n, hipStream_t s);

}  // namespace fdb
