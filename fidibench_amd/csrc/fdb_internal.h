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

// True when ntot cells of `bytes` each fit a size_t byte count (so no
// allocation size wraps around).
inline bool fits_bytes(uint64_t ntot, size_t bytes) {
  return bytes != 0 && ntot <= SIZE_MAX / bytes;
}

// Message naming a tuning variable (FDB_ROWS_CFG*, FDB_ADJ_CFG*,
// FDB_TB_CFG_*) that does not parse or names a launch shape with no
// instantiated kernel; nullptr when every such variable is valid.  Read
// once per process.
const char* tuning_error();

// Key of the stochastic rounding of step `step` of a run seeded `seed`.
uint32_t sr_step_key(uint64_t seed, uint64_t step);

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
// src is not written.  With levels != nullptr the states after steps
// 1..k-1 are also written, to levels[0..k-2] (FP64, k = 4 or 8).  For
// FDB_BF16_SR the steps are number step0 .. step0+k-1 of a run seeded
// `seed` (their stochastic-rounding keys).  hipErrorNotSupported when no
// single launch does this for the storage type and shape (the caller then
// steps one by one).
hipError_t launch_fused_pass(int dtype, const Grid3& g, const void* src, void* dst, int k,
                             hipStream_t s, void* const* levels = nullptr, uint64_t step0 = 0,
                             uint64_t seed = 0);

// Whether launch_fused_pass would run k fused steps of this storage type
// (FDB_F64, FDB_F32, FDB_BF16, FDB_BF16_SR) and shape in one launch.
bool fused_pass_supported(int dtype, const Grid3& g, int k);

// Steps of the next launch when `left` (>= 1) steps remain and at most
// `cap` may be fused: the first pass of the least-time split of `left`
// into one-step launches and supported fused passes (1 = a one-step
// launch).  Every split gives the same results.
int plan_pass(int dtype, const Grid3& g, int left, int cap);

// The storage layout of a dtype (FDB_BF16_SR is stored as FDB_BF16).
inline int storage_dtype(int dtype) { return dtype == 3 ? 2 : dtype; }

// K transpose steps in one pass (lam/f_t storage FP64/FP64, FP32/FP32 or
// FP32/bf16; all three axes present): lam = (A^T)^K lam_next, where fs[l]
// is the forward state of the l-th step (f_{t-l} for a pass starting from
// lam_{t+1}); the velocity products of all K steps are added into acc
// (device, 3 x {sum, comp}), as K launch_adjoint_step calls would add
// them (lam bit-identical; the products are the same, summed in another
// order).  partials needs adjoint_fused_blocks(...) slots of 6 doubles;
// hipErrorNotSupported when the shape has no fused form or needs more than
// max_blocks slots.
hipError_t launch_adjoint_fused(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                                void* lam, const void* const* fs, int k, double* partials,
                                int64_t max_blocks, double* acc, hipStream_t s);
int64_t adjoint_fused_blocks(int lam_dtype, int f_dtype, const Grid3& g, int k);

// K transpose steps in one pass whose last step also adds its velocity
// products against f (the forward state that step consumes) into acc:
// K - 1 plain transpose steps and one launch_adjoint_step with f_t = f,
// lam bit-identical, the products summed in another order.  k must be
// adjoint_last_depth(...), which returns 0 when the pairing or shape has
// no such pass and sets *nb to the partial slots it needs.
hipError_t launch_adjoint_last(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                               void* lam, const void* f, int k, double* partials,
                               int64_t max_blocks, double* acc, hipStream_t s);
int adjoint_last_depth(int lam_dtype, int f_dtype, const Grid3& g, int64_t* nb);

// Launch sizing shared by the marching kernels.
int device_cus();
// Blocks of kernel `kern` (host stub) with `threads` threads and `dyn_lds`
// bytes of dynamic LDS that the device holds at once.
int64_t kernel_capacity(const void* kern, int threads, size_t dyn_lds);
// Chunk count for a march over n0 planes (fdb_kernels.hip).
int64_t march_chunks(int64_t tiles, int64_t n0, int extra, int64_t capacity, int64_t lo,
                     int64_t hi, bool fill);

// ---- generic periodic stencil (cxx/Filter.cpp), kernels in fdb_filter_kernels.hip

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

// Tiling of the marching kernels: a TX x TY tile of axes 0/1 with rx/ry
// halos (LX x LY values per plane), a ring of D = 2 rz + 1 planes, tiles
// of `chunk` consecutive axis-2 planes each.
struct FilterMarch {
  int rx, ry, rz, LX, LY, D;
  int64_t tiles_x, tiles_y, chunk;
  int lofs[kFilterMaxTerms];  // oy * LX + ox, per term
  int zo[kFilterMaxTerms];    // oz, per term
};

// How one filter application is launched, fixed once per geometry.
struct FilterPlan {
  int kind;      // kFilterGather, kFilterPlanes(Mid, Wide), kFilterRows or kFilterStar
  FilterGeom g;  // geometry as launched (kFilterRows: viewed as n0 x 1 x n1)
  FilterMarch m;
  size_t lds;    // dynamic LDS bytes per block
  int64_t blocks;
};
constexpr int kFilterGather = 0;  // one thread per output cell, neighbours from HBM/L2
constexpr int kFilterPlanes = 1;  // 3-D: 128 x 8 tiles marched along axis 2
constexpr int kFilterRows = 2;    // 2-D: 512-wide tiles marched along axis 1
constexpr int kFilterPlanesWide = 3;  // 3-D, halo tiles up to radius 8
constexpr int kFilterPlanesMid = 4;   // 3-D, 128 x 4 tiles, radius 2..3
constexpr int kFilterStar = 5;        // 3-D full star stencils of radius 2 (m.rz = 2)

// Picks the kernel and tiling for g (needs the device current).
hipError_t plan_filter(const FilterGeom& g, FilterPlan* p);

// out[i] = sum over terms, in order, of w * in[wrapped neighbour].
hipError_t launch_filter(const FilterPlan& p, const double* in, double* out, hipStream_t s);

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
