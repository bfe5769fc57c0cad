// fdb_filter_kernels.hip -- the generic periodic stencil filter of pletzer/fidibench
// cxx/Filter.cpp on one GPU: Filter::applyFilter (Filter.cpp:192-263) with
// one process, where every neighbour is local.  Per output cell,
//     out = 0;  for each (offset, weight) in std::map order:  out += weight * in[nbr]
// with each product and sum rounded separately (no contraction), so the
// result is bit-identical to the reference's loop order.  The neighbour
// index per axis is the reference's ((int)i + offset) % (size_t)n
// (Filter.cpp:231-233): the periodic neighbour for offsets that stay >= 0,
// and the unsigned remainder (size_t)(i + offset) % n for negative indices,
// which equals the periodic neighbour only when n divides 2^64.
//
// Layout: the reference's column-major order (axis 0 contiguous).  One
// block = 256 consecutive cells of one axis-0 row; the row's axis-1/axis-2
// indices are block-uniform, so only the axis-0 wrap is per lane.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fdb_internal.h"

#pragma clang fp contract(off)

namespace fdb {
namespace {

constexpr int kFilterThreads = 256;

__device__ inline int64_t filter_wrap(int64_t idx, int a, const FilterGeom& g) {
  if (idx < 0) return g.neg[a][-idx - 1];
  return idx < g.n[a] ? idx : idx % g.n[a];
}

// NT > 0: the term count known at compile time (fully unrolled, every
// neighbour load issued before the sums); NT = 0: g.nt terms.
template <int NT>
__global__ __launch_bounds__(kFilterThreads) void filter_kernel(const double* __restrict__ in,
                                                               double* __restrict__ out,
                                                               FilterGeom g, int64_t xblocks) {
  const int64_t b = blockIdx.x;
  const int64_t row = b / xblocks;  // i1 + n1 * i2, block-uniform
  const int64_t i0 = (b - row * xblocks) * kFilterThreads + threadIdx.x;
  if (i0 >= g.n[0]) return;
  const int64_t i1 = row % g.n[1];
  const int64_t i2 = row / g.n[1];
  constexpr int kUnroll = NT > 0 ? NT : 1;
  const int nt = NT > 0 ? NT : g.nt;
  double acc = 0.0;
  if constexpr (NT > 0) {
    double v[kUnroll];
#pragma unroll
    for (int t = 0; t < kUnroll; ++t) {
      const int64_t j0 = filter_wrap(i0 + g.off[t][0], 0, g);
      const int64_t j1 = filter_wrap(i1 + g.off[t][1], 1, g);
      const int64_t j2 = filter_wrap(i2 + g.off[t][2], 2, g);
      v[t] = in[j0 + g.n[0] * (j1 + g.n[1] * j2)];
    }
#pragma unroll
    for (int t = 0; t < kUnroll; ++t) acc = acc + g.w[t] * v[t];
  } else {
    for (int t = 0; t < nt; ++t) {
      const int64_t j0 = filter_wrap(i0 + g.off[t][0], 0, g);
      const int64_t j1 = filter_wrap(i1 + g.off[t][1], 1, g);
      const int64_t j2 = filter_wrap(i2 + g.off[t][2], 2, g);
      acc = acc + g.w[t] * in[j0 + g.n[0] * (j1 + g.n[1] * j2)];
    }
  }
  __builtin_nontemporal_store(acc, out + i0 + g.n[0] * row);
}

struct SinArgs {
  int nd;
  double x0[3], dx[3];
};

__global__ __launch_bounds__(kFilterThreads) void filter_sin_kernel(double* __restrict__ in,
                                                                    FilterGeom g, SinArgs s,
                                                                    int64_t xblocks) {
  const int64_t b = blockIdx.x;
  const int64_t row = b / xblocks;
  const int64_t i0 = (b - row * xblocks) * kFilterThreads + threadIdx.x;
  if (i0 >= g.n[0]) return;
  const int64_t idx[3] = {i0, row % g.n[1], row / g.n[1]};
  double res = 1;
  for (int a = 0; a < s.nd; ++a) {
    const double pos = s.x0[a] + ((double)idx[a] + 0.5) * s.dx[a];
    res *= sin(2.0 * M_PI * pos);
  }
  in[i0 + g.n[0] * row] = res;
}

}  // namespace

hipError_t launch_filter(const FilterGeom& g, const double* in, double* out, hipStream_t s) {
  const int64_t xblocks = (g.n[0] + kFilterThreads - 1) / kFilterThreads;
  const int64_t nb = xblocks * g.n[1] * g.n[2];
  if (nb <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  const dim3 grid((unsigned)nb), block(kFilterThreads);
  switch (g.nt) {
    case 3: hipLaunchKernelGGL(filter_kernel<3>, grid, block, 0, s, in, out, g, xblocks); break;
    case 5: hipLaunchKernelGGL(filter_kernel<5>, grid, block, 0, s, in, out, g, xblocks); break;
    case 7: hipLaunchKernelGGL(filter_kernel<7>, grid, block, 0, s, in, out, g, xblocks); break;
    default: hipLaunchKernelGGL(filter_kernel<0>, grid, block, 0, s, in, out, g, xblocks); break;
  }
  return hipGetLastError();
}

hipError_t launch_filter_init_sin(const FilterGeom& g, int ndims, const double* xmins,
                                  const double* xmaxs, double* in, hipStream_t s) {
  SinArgs a = {};
  a.nd = ndims;
  for (int j = 0; j < ndims; ++j) {
    a.x0[j] = xmins[j];
    a.dx[j] = (xmaxs[j] - xmins[j]) / (double)g.n[j];  // Filter.cpp:104-109
  }
  const int64_t xblocks = (g.n[0] + kFilterThreads - 1) / kFilterThreads;
  const int64_t nb = xblocks * g.n[1] * g.n[2];
  if (nb <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(filter_sin_kernel, dim3((unsigned)nb), dim3(kFilterThreads), 0, s, in, g, a,
                     xblocks);
  return hipGetLastError();
}

}  // namespace fdb
