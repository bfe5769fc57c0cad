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
// Layout: the reference's column-major order (axis 0 contiguous).  Two
// kernels: a plane-marching one with an LDS ring (2-D/3-D grids, below),
// and a gather one for 1-D grids and stencils too wide for the ring, where
// one block = 256 consecutive cells of one axis-0 row (the row's axis-1/2
// indices block-uniform, only the axis-0 wrap per lane).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>

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

// ---- plane-marching kernel: one block = a TX x TY tile of axes 0/1 over a
// chunk of axis-2 planes.  The planes z-rz .. z+rz around the output plane
// sit in an LDS ring, each with an rx/ry halo (loaded through the same
// wrap, so an LDS value is exactly in[wrapped neighbour]); every cell then
// sums its terms from LDS in map order.  The next plane's global loads are
// issued into registers before the current plane is computed.
constexpr int kTX = 128, kTY = 8;        // tile; 256 threads = 64 x 4
constexpr int kLoadX = 3, kLoadY = 6;    // LX <= 192, LY <= 24 (|offset| <= 8)

struct MarchTile {
  int rx, ry, rz, LX, LY, D;
  int64_t tiles_x, tiles_y, chunk;
  int lofs[kFilterMaxTerms];  // oy * LX + ox
  int zo[kFilterMaxTerms];    // oz
};

template <int NT>
__global__ __launch_bounds__(256) void filter_march_kernel(const double* __restrict__ in,
                                                           double* __restrict__ out,
                                                           FilterGeom g, MarchTile m) {
  extern __shared__ double ring[];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int64_t b = blockIdx.x;
  const int64_t bx = b % m.tiles_x;
  const int64_t by = (b / m.tiles_x) % m.tiles_y;
  const int64_t bz = b / (m.tiles_x * m.tiles_y);
  const int64_t x0 = bx * kTX, y0 = by * kTY;
  const int64_t z0 = bz * m.chunk;
  const int64_t z1 = (z0 + m.chunk < g.n[2]) ? z0 + m.chunk : g.n[2];
  const int planeLDS = m.LX * m.LY;

  // Global offsets (within a plane) of this thread's halo-tile elements.
  int64_t goff[kLoadY][kLoadX];
  bool gval[kLoadY][kLoadX];
#pragma unroll
  for (int k = 0; k < kLoadY; ++k) {
    const int ly = ty + 4 * k;
    const int64_t yw = filter_wrap(y0 - m.ry + ly, 1, g);
#pragma unroll
    for (int p = 0; p < kLoadX; ++p) {
      const int lx = tx + 64 * p;
      gval[k][p] = ly < m.LY && lx < m.LX;
      goff[k][p] = filter_wrap(x0 - m.rx + lx, 0, g) + g.n[0] * yw;
    }
  }
  auto slot = [&](int64_t q) { return (int)(((q % m.D) + m.D) % m.D); };
  auto fetch = [&](int64_t q, double (&v)[kLoadY][kLoadX]) {
    const double* pl = in + filter_wrap(q, 2, g) * g.n[0] * g.n[1];
#pragma unroll
    for (int k = 0; k < kLoadY; ++k)
#pragma unroll
      for (int p = 0; p < kLoadX; ++p)
        if (gval[k][p]) v[k][p] = pl[goff[k][p]];
  };
  auto put = [&](int64_t q, const double (&v)[kLoadY][kLoadX]) {
    double* dst = ring + slot(q) * planeLDS;
#pragma unroll
    for (int k = 0; k < kLoadY; ++k)
#pragma unroll
      for (int p = 0; p < kLoadX; ++p)
        if (gval[k][p]) dst[(ty + 4 * k) * m.LX + tx + 64 * p] = v[k][p];
  };

  double v[kLoadY][kLoadX];
  for (int64_t q = z0 - m.rz; q < z0 + m.rz; ++q) {  // warm-up planes
    fetch(q, v);
    put(q, v);
  }
  if (z0 < z1) fetch(z0 + m.rz, v);
  const int nt = NT > 0 ? NT : g.nt;
  for (int64_t z = z0; z < z1; ++z) {
    put(z + m.rz, v);
    __syncthreads();
    if (z + 1 < z1) fetch(z + 1 + m.rz, v);  // in flight during the sums
    double* ob = out + z * g.n[0] * g.n[1];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t y = y0 + ty + 4 * j;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int64_t x = x0 + tx + 64 * i;
        const int base = (ty + 4 * j + m.ry) * m.LX + tx + 64 * i + m.rx;
        double acc = 0.0;
        if constexpr (NT > 0) {
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc = acc + g.w[t] * ring[slot(z + m.zo[t]) * planeLDS + base + m.lofs[t]];
        } else {
          for (int t = 0; t < nt; ++t)
            acc = acc + g.w[t] * ring[slot(z + m.zo[t]) * planeLDS + base + m.lofs[t]];
        }
        if (x < g.n[0] && y < g.n[1]) __builtin_nontemporal_store(acc, ob + x + g.n[0] * y);
      }
    }
    __syncthreads();  // the next put overwrites the oldest plane
  }
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
  // The marching kernel for 2-D/3-D grids whose stencil ring fits in 64 KB
  // of LDS; the gather kernel otherwise (1-D grids, wide stencils).
  MarchTile m = {};
  for (int t = 0; t < g.nt; ++t) {
    m.rx = std::max(m.rx, std::abs(g.off[t][0]));
    m.ry = std::max(m.ry, std::abs(g.off[t][1]));
    m.rz = std::max(m.rz, std::abs(g.off[t][2]));
  }
  m.LX = kTX + 2 * m.rx;
  m.LY = kTY + 2 * m.ry;
  m.D = 2 * m.rz + 1;
  const size_t lds = (size_t)m.D * m.LX * m.LY * sizeof(double);
  if (g.n[1] >= kTY && lds <= 65536) {
    for (int t = 0; t < g.nt; ++t) {
      m.lofs[t] = g.off[t][1] * m.LX + g.off[t][0];
      m.zo[t] = g.off[t][2];
    }
    m.tiles_x = (g.n[0] + kTX - 1) / kTX;
    m.tiles_y = (g.n[1] + kTY - 1) / kTY;
    const int64_t tiles = m.tiles_x * m.tiles_y;
    // about 2048 blocks (8 per CU), each marching a chunk of planes
    const int64_t nch = std::max<int64_t>(1, std::min<int64_t>(g.n[2], (2048 + tiles - 1) / tiles));
    m.chunk = (g.n[2] + nch - 1) / nch;
    const int64_t nb = tiles * ((g.n[2] + m.chunk - 1) / m.chunk);
    if (nb <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)nb), block(64, 4);
    switch (g.nt) {
      case 5: hipLaunchKernelGGL(filter_march_kernel<5>, grid, block, lds, s, in, out, g, m); break;
      case 7: hipLaunchKernelGGL(filter_march_kernel<7>, grid, block, lds, s, in, out, g, m); break;
      default: hipLaunchKernelGGL(filter_march_kernel<0>, grid, block, lds, s, in, out, g, m); break;
    }
    return hipGetLastError();
  }
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
