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
#include <type_traits>
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
// sit in an LDS ring, each the LX x LY halo tile (rx/ry halos loaded
// through the same wrap, so an LDS value is exactly in[wrapped
// neighbour]); every cell then sums its terms from LDS in map order.
// The halo tile is loaded flat: thread t takes elements t, t + 256, ...
// (NL per thread, consecutive threads on consecutive x), held in a
// register ring PF planes deep, so PF planes of loads are in flight while
// a plane is summed.
// Block BX x (256/BX) threads, each summing CX x CY cells: TX = BX * CX,
// TY = (256/BX) * CY.  Shapes: 64 x 4 threads, 2 x 2 cells (128 x 8 tiles)
// for 3-D grids; 256 x 1 threads, 2 x 1 cells (512 x 1 tiles) for 2-D
// grids, which march along their axis 1 viewed as axis 2 of an
// n0 x 1 x n1 grid.
template <int BX, int CX, int CY, int NL, int PF>
struct MarchShape {
  static constexpr int kBX = BX, kBY = 256 / BX, kCX = CX, kCY = CY, kNL = NL, kPF = PF;
  static constexpr int kTX = BX * CX, kTY = (256 / BX) * CY;
};
using March3D = MarchShape<64, 2, 2, 6, 2>;       // halo tile <= 1536 values (radius 1)
using March3DMid = MarchShape<64, 2, 1, 6, 2>;    // 128 x 4 tiles, <= 1536 (radius 2..3)
using March3DWide = MarchShape<64, 2, 2, 14, 2>;  // <= 3584 values (radius <= 8)
using March2D = MarchShape<256, 2, 1, 3, 4>;      // <= 768 values (rx <= 128)

// Hardware hands blocks to the 8 XCDs round-robin by block id; renumber
// them so each XCD takes one contiguous range of tiles (neighbouring tiles,
// whose halos overlap, then share that XCD's L2).
__device__ inline int64_t xcd_contiguous(int64_t i, int64_t nb) {
  const int64_t x = i & 7, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (i >> 3);
}

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it stay in flight (__syncthreads() would also wait for them).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Marks a loaded value as used here, so the compiler waits for it before
// this point rather than at a loop header, where the wait it merges for
// loads issued before the loop would also cover the loop's own stores.
__device__ inline void settle(double x) { asm volatile("" ::"v"(x)); }

// Every memory access of the plane loop is unconditional, so the compiler's
// wait counts let the prefetched planes stay in flight: lanes outside the
// halo tile load a valid (wrapped) address and write a spare LDS area of
// 64 values after each ring plane; output cells past the grid's edge are
// stored through a raw buffer resource whose range check drops them.
template <int NT, class S>
__global__ __launch_bounds__(256) void filter_march_kernel(const double* __restrict__ in,
                                                           double* __restrict__ out,
                                                           FilterGeom g, FilterMarch m) {
  extern __shared__ double ring[];
  constexpr int BX = S::kBX, BY = S::kBY, NL = S::kNL, PF = S::kPF;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * BX + tx;
  const int64_t b = xcd_contiguous(blockIdx.x, gridDim.x);
  const int64_t bx = b % m.tiles_x;
  const int64_t by = (b / m.tiles_x) % m.tiles_y;
  const int64_t bz = b / (m.tiles_x * m.tiles_y);
  const int64_t x0 = bx * S::kTX, y0 = by * S::kTY;
  const int64_t z0 = bz * m.chunk;  // < n[2]: the grid has ceil(n2 / chunk) chunks
  const int64_t z1 = (z0 + m.chunk < g.n[2]) ? z0 + m.chunk : g.n[2];
  const int ne = m.LX * m.LY;  // halo-tile values per ring plane
  const int ps = ne + 64;      // ring plane stride (spare area after each plane)

  // Global offsets (within a plane) and LDS slots of this thread's
  // halo-tile elements.
  int64_t goff[NL];
  int loff[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int e = tid + 256 * k;
    const bool ok = e < ne;
    const int ly = ok ? e / m.LX : 0;
    const int lx = ok ? e - ly * m.LX : 0;
    goff[k] = filter_wrap(x0 - m.rx + lx, 0, g) + g.n[0] * filter_wrap(y0 - m.ry + ly, 1, g);
    loff[k] = ok ? e : ne + (tid & 63);
  }
  // Output offsets (bytes from the tile's first row in a plane); cells
  // outside the grid get an offset past the buffer's range.
  int soff[S::kCY][S::kCX];
#pragma unroll
  for (int j = 0; j < S::kCY; ++j)
#pragma unroll
    for (int i = 0; i < S::kCX; ++i) {
      const int64_t x = x0 + tx + BX * i, y = y0 + ty + BY * j;
      soff[j][i] = (x < g.n[0] && y < g.n[1]) ? (int)((x + g.n[0] * (y - y0)) * 8) : 0x7fffffff;
    }
  const int64_t plane = g.n[0] * g.n[1];
  const int64_t tile_bytes = (g.n[1] - y0) * g.n[0] * 8;  // tile's first row to the plane's end
  const int nrec = (int)(tile_bytes < 0x7fffffffLL ? tile_bytes : 0x7fffffffLL);

  // Ring slots: cur is the slot of the output plane, advanced by one per
  // plane; plane cur + d (|d| <= 2 rz) wraps once at most.
  auto wrapD = [&](int x) { return x < 0 ? x + m.D : (x >= m.D ? x - m.D : x); };
  auto fetch = [&](int64_t q, double (&v)[NL]) {
    const double* pl = in + filter_wrap(q, 2, g) * plane;
#pragma unroll
    for (int k = 0; k < NL; ++k) v[k] = pl[goff[k]];
  };
  auto put = [&](int sl, const double (&v)[NL]) {
    double* dst = ring + sl * ps;
#pragma unroll
    for (int k = 0; k < NL; ++k) dst[loff[k]] = v[k];
  };
  const int nt = NT > 0 ? NT : g.nt;
  auto sums = [&](int64_t z, int cur) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(out + z * plane + y0 * g.n[0], 0, nrec, 0x00020000);
#pragma unroll
    for (int j = 0; j < S::kCY; ++j) {
#pragma unroll
      for (int i = 0; i < S::kCX; ++i) {
        const int base = (ty + BY * j + m.ry) * m.LX + tx + BX * i + m.rx;
        double acc = 0.0;
        if constexpr (NT > 0) {
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc = acc + g.w[t] * ring[wrapD(cur + m.zo[t]) * ps + base + m.lofs[t]];
        } else {
          for (int t = 0; t < nt; ++t)
            acc = acc + g.w[t] * ring[wrapD(cur + m.zo[t]) * ps + base + m.lofs[t]];
        }
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, acc), rs, soff[j][i], 0,
                                              2 /* nt */);
      }
    }
  };

  double v[PF][NL];
  int cur = m.rz;  // slot of plane z0; plane z0 - rz takes slot 0
  for (int d = -m.rz; d < m.rz; ++d) {  // warm-up planes
    fetch(z0 + d, v[0]);
    put(cur + d, v[0]);
  }
#pragma unroll
  for (int u = 0; u < PF; ++u) fetch((z0 + u < z1 ? z0 + u : z1 - 1) + m.rz, v[u]);
  // One plane: store plane zz + rz (prefetched), refill its registers with
  // plane zz + PF + rz (clamped to the chunk: a re-read, not a branch),
  // sum plane zz.
  auto step = [&](int64_t zz, double (&vu)[NL]) {
    put(wrapD(cur + m.rz), vu);
    lds_barrier();
    fetch((zz + PF < z1 ? zz + PF : z1 - 1) + m.rz, vu);
    sums(zz, cur);
    cur = cur + 1 == m.D ? 0 : cur + 1;
    lds_barrier();  // the next put overwrites the oldest plane
  };
  int64_t z = z0;
  for (; z + PF <= z1; z += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) step(z + u, v[u]);
  }
#pragma unroll
  for (int u = 0; u < PF - 1; ++u)
    if (z + u < z1) step(z + u, v[u]);  // block-uniform
}

// ---- star kernel: full 3-D star stencils of radius R, the 6R + 1 terms
// (+-k, 0, 0), (0, +-k, 0), (0, 0, +-k) for k = 1..R and the centre (the
// 4th-order Laplacian at R = 2; at R = 1 the ring kernel below, whose
// three planes take 33 KB, measured faster: 334-339 vs 317-319 G
// cell-updates/s for the 7-point Laplacian at 512^3).  In std::map order their
// terms are the axis-0 offsets -R..-1, the axis-1 offsets -R..-1, the
// axis-2 offsets -R..R (the centre among them), the axis-1 offsets 1..R
// and the axis-0 offsets 1..R; the weights stay runtime values.  The
// axis-2 terms read only the cell's own column, so each thread keeps its
// cells' values of planes z-R .. z+R in registers (loaded R planes ahead)
// and LDS holds only the output plane's halo tile, double-buffered so one
// barrier per plane suffices (26 KB at R = 2 against the ring kernel's
// five planes, 45 KB: twice the blocks per CU).  A plane's tile is put
// together from the threads' own cells (their queue's middle entry) and
// the halo ring around them, loaded flat across the block, so every value
// is read from memory once per plane.  Memory accesses in the plane loop
// are unconditional, as in the ring kernel.
template <int R>
struct StarShape {
  static constexpr int kBX = 64, kBY = 4, kCX = 2, kCY = 2;
  static constexpr int kTX = kBX * kCX, kTY = kBY * kCY;
  static constexpr int kLX = kTX + 2 * R, kLY = kTY + 2 * R, kNE = kLX * kLY;
  static constexpr int kNH = kNE - kTX * kTY;    // halo-ring values per plane
  static constexpr int kNLH = (kNH + 255) / 256;  // per thread
  static constexpr size_t kLds = 2 * (kNE + 64) * sizeof(double);
};

template <int R>
__global__ __launch_bounds__(256) void filter_star_kernel(const double* __restrict__ in,
                                                          double* __restrict__ out,
                                                          FilterGeom g, FilterMarch m) {
  using S = StarShape<R>;
  constexpr int BX = S::kBX, BY = S::kBY, CX = S::kCX, CY = S::kCY;
  constexpr int LX = S::kLX, NE = S::kNE, NH = S::kNH, NLH = S::kNLH, D = 2 * R + 1;
  static_assert(R == 1 || R == 2, "the plane loop is unrolled for 4 or 6 queue slots");
  __shared__ double tile[2][NE + 64];  // + a spare area for lanes without a ring value
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * BX + tx;
  const int64_t b = xcd_contiguous(blockIdx.x, gridDim.x);
  const int64_t bx = b % m.tiles_x;
  const int64_t by = (b / m.tiles_x) % m.tiles_y;
  const int64_t bz = b / (m.tiles_x * m.tiles_y);
  const int64_t x0 = bx * S::kTX, y0 = by * S::kTY;
  const int64_t z0 = bz * m.chunk;
  const int64_t z1 = (z0 + m.chunk < g.n[2]) ? z0 + m.chunk : g.n[2];
  const int64_t plane = g.n[0] * g.n[1];

  // The thread's own cells: in-plane offsets (wrapped: cells past the
  // grid's edge hold the wrapped neighbour, as the halo does), tile slots
  // and store offsets (past the buffer's range outside the grid).
  int64_t gown[CY][CX];
  int lown[CY][CX], soff[CY][CX];
#pragma unroll
  for (int j = 0; j < CY; ++j)
#pragma unroll
    for (int i = 0; i < CX; ++i) {
      const int64_t x = x0 + tx + BX * i, y = y0 + ty + BY * j;
      gown[j][i] = filter_wrap(x, 0, g) + g.n[0] * filter_wrap(y, 1, g);
      lown[j][i] = (ty + BY * j + R) * LX + tx + BX * i + R;
      soff[j][i] = (x < g.n[0] && y < g.n[1]) ? (int)((x + g.n[0] * (y - y0)) * 8) : 0x7fffffff;
    }
  // This thread's share of the halo ring: the R rows above and below the
  // tile (full width), then R columns left and right of each tile row.
  int64_t ghal[NLH];
  int lhal[NLH];
#pragma unroll
  for (int k = 0; k < NLH; ++k) {
    const int e = tid + 256 * k;
    int lx = 0, ly = 0;
    if (e < 2 * R * LX) {
      ly = e / LX;
      lx = e - ly * LX;
      if (ly >= R) ly += S::kTY;
    } else if (e < NH) {
      const int e2 = e - 2 * R * LX;
      const int row = e2 / (2 * R), c = e2 - row * (2 * R);
      ly = R + row;
      lx = c < R ? c : LX - 2 * R + c;
    }
    ghal[k] = e < NH ? filter_wrap(x0 - R + lx, 0, g) + g.n[0] * filter_wrap(y0 - R + ly, 1, g)
                     : 0;
    lhal[k] = e < NH ? ly * LX + lx : NE + (tid & 63);
  }
  const int64_t tile_bytes = (g.n[1] - y0) * g.n[0] * 8;  // tile's first row to the plane's end
  const int nrec = (int)(tile_bytes < 0x7fffffffLL ? tile_bytes : 0x7fffffffLL);

  // Planes inside the loop are indexed 0 <= p < n2 + R <= 2 n2, where the
  // reference's wrap is p or p - n2.
  auto ld_own = [&](int64_t p, double (&v)[CY][CX]) {
    const double* pl = in + (p < g.n[2] ? p : p - g.n[2]) * plane;
#pragma unroll
    for (int j = 0; j < CY; ++j)
#pragma unroll
      for (int i = 0; i < CX; ++i) v[j][i] = pl[gown[j][i]];
  };
  auto ld_hal = [&](int64_t p, double (&v)[NLH]) {
    const double* pl = in + (p < g.n[2] ? p : p - g.n[2]) * plane;
#pragma unroll
    for (int k = 0; k < NLH; ++k) v[k] = pl[ghal[k]];
  };

  // Own cells of planes z-R .. z+R and of the prefetched plane z+R+1 in a
  // circular queue of NQ = D + 1 slots, plane p in slot (p - z0 + R) mod NQ;
  // the plane loop is unrolled NQ times (NQ is even), so every slot, halo
  // buffer and LDS buffer index is a constant and nothing is copied.
  constexpr int NQ = D + 1;
  double q[NQ][CY][CX], h[2][NLH];
#pragma unroll
  for (int d = 0; d < D; ++d) {  // (warm-up planes may be negative: the full wrap)
    const double* pl = in + filter_wrap(z0 - R + d, 2, g) * plane;
#pragma unroll
    for (int j = 0; j < CY; ++j)
#pragma unroll
      for (int i = 0; i < CX; ++i) q[d][j][i] = pl[gown[j][i]];
  }
  ld_hal(z0, h[0]);
#pragma unroll
  for (int d = 0; d < D; ++d)
#pragma unroll
    for (int j = 0; j < CY; ++j)
#pragma unroll
      for (int i = 0; i < CX; ++i) settle(q[d][j][i]);
#pragma unroll
  for (int k = 0; k < NLH; ++k) settle(h[0][k]);

  // Output plane z at phase U = (z - z0) mod NQ.
  auto step = [&](int64_t z, auto phase) __attribute__((always_inline)) {
    constexpr int U = decltype(phase)::value;
    double* t = tile[U & 1];
#pragma unroll
    for (int j = 0; j < CY; ++j)
#pragma unroll
      for (int i = 0; i < CX; ++i) t[lown[j][i]] = q[(U + R) % NQ][j][i];
#pragma unroll
    for (int k = 0; k < NLH; ++k) t[lhal[k]] = h[U & 1][k];
    lds_barrier();
    // The next plane's loads (after the chunk's last plane: this one again).
    const int64_t zn = z + 1 < z1 ? z + 1 : z;
    ld_own(zn + R, q[(U + D) % NQ]);
    ld_hal(zn, h[(U + 1) & 1]);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(out + z * plane + y0 * g.n[0], 0, nrec, 0x00020000);
#pragma unroll
    for (int j = 0; j < CY; ++j) {
#pragma unroll
      for (int i = 0; i < CX; ++i) {
        const int c = lown[j][i];
        double acc = 0.0;
        int u = 0;  // term index in map order
#pragma unroll
        for (int k = -R; k < 0; ++k) acc = acc + g.w[u++] * t[c + k];
#pragma unroll
        for (int k = -R; k < 0; ++k) acc = acc + g.w[u++] * t[c + k * LX];
#pragma unroll
        for (int k = -R; k <= R; ++k) acc = acc + g.w[u++] * q[(U + R + k + NQ) % NQ][j][i];
#pragma unroll
        for (int k = 1; k <= R; ++k) acc = acc + g.w[u++] * t[c + k * LX];
#pragma unroll
        for (int k = 1; k <= R; ++k) acc = acc + g.w[u++] * t[c + k];
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, acc), rs, soff[j][i], 0,
                                              2 /* nt */);
      }
    }
  };
  int64_t z = z0;
  for (; z + NQ <= z1; z += NQ) {
    step(z, std::integral_constant<int, 0>{});
    step(z + 1, std::integral_constant<int, 1>{});
    step(z + 2, std::integral_constant<int, 2>{});
    step(z + 3, std::integral_constant<int, 3>{});
    if constexpr (NQ > 4) {
      step(z + 4, std::integral_constant<int, 4 % NQ>{});
      step(z + 5, std::integral_constant<int, 5 % NQ>{});
    }
  }
  // The last planes (block-uniform branches).
  if (z < z1) step(z, std::integral_constant<int, 0>{});
  if (z + 1 < z1) step(z + 1, std::integral_constant<int, 1>{});
  if (z + 2 < z1) step(z + 2, std::integral_constant<int, 2>{});
  if constexpr (NQ > 4) {
    if (z + 3 < z1) step(z + 3, std::integral_constant<int, 3>{});
    if (z + 4 < z1) step(z + 4, std::integral_constant<int, 4 % NQ>{});
  }
}

// Radius R if g is a full 3-D star stencil of radius R = 2 with its terms
// in map order (see filter_star_kernel), else 0.
int star_radius(const FilterGeom& g) {
  for (int R = 2; R <= 2; ++R) {
    if (g.nt != 6 * R + 1) continue;
    int expect[kFilterMaxTerms][3] = {};
    int u = 0;
    for (int k = -R; k < 0; ++k) expect[u++][0] = k;
    for (int k = -R; k < 0; ++k) expect[u++][1] = k;
    for (int k = -R; k <= R; ++k) expect[u++][2] = k;
    for (int k = 1; k <= R; ++k) expect[u++][1] = k;
    for (int k = 1; k <= R; ++k) expect[u++][0] = k;
    bool same = true;
    for (int t = 0; t < g.nt && same; ++t)
      for (int a = 0; a < 3; ++a) same = same && g.off[t][a] == expect[t][a];
    if (same) return R;
  }
  return 0;
}

template <int R>
bool star_plan(const FilterGeom& g, FilterMarch& m, int64_t& blocks) {
  using S = StarShape<R>;
  m = FilterMarch{};
  m.rx = m.ry = m.rz = R;
  m.LX = S::kLX;
  m.LY = S::kLY;
  m.D = 2 * R + 1;
  // (n2 > 2R: the loop's planes p < n2 + R wrap to p or p - n2)
  if (g.n[2] <= 2 * R || g.n[1] < S::kTY) return false;
  if ((int64_t)S::kTY * g.n[0] * 8 >= 0x7fffffffLL) return false;  // store offsets
  m.tiles_x = (g.n[0] + S::kTX - 1) / S::kTX;
  m.tiles_y = (g.n[1] + S::kTY - 1) / S::kTY;
  const int64_t tiles = m.tiles_x * m.tiles_y;
  const int64_t cap =
      kernel_capacity(reinterpret_cast<const void*>(&filter_star_kernel<R>), 256, 0);
  const int64_t per_round = std::max<int64_t>(1, cap / tiles);
  const int64_t nch = march_chunks(tiles, g.n[2], 2 * R, cap,
                                   std::max<int64_t>(1, per_round / 2), 4 * per_round, true);
  m.chunk = (g.n[2] + nch - 1) / nch;
  blocks = tiles * ((g.n[2] + m.chunk - 1) / m.chunk);
  return blocks > 0 && blocks <= 0x7fffffffLL;
}

template <int NT, class S>
const void* march_kernel_ptr() {
  return reinterpret_cast<const void*>(&filter_march_kernel<NT, S>);
}

// Term counts with an unrolled kernel: 2nd-order Laplacians (5, 7),
// 4th-order ones (9, 13), the 27-point box; others loop over g.nt.
template <class S>
const void* march_kernel_for(int nt) {
  switch (nt) {
    case 5: return march_kernel_ptr<5, S>();
    case 7: return march_kernel_ptr<7, S>();
    case 9: return march_kernel_ptr<9, S>();
    case 13: return march_kernel_ptr<13, S>();
    case 27: return march_kernel_ptr<27, S>();
    default: return march_kernel_ptr<0, S>();
  }
}

// LDS of one workgroup on gfx950; rings above the default 64 KB of dynamic
// LDS need the kernel's limit raised (hipFuncSetAttribute).
constexpr size_t kMaxRingLds = 160 * 1024;

// Radii, ring size and tiling of the marching kernel for geometry g and
// shape S; false if the ring does not fit in the workgroup's LDS (or the
// runtime refuses to raise the kernel's limit) or the halo tile exceeds
// the shape's loads (256 x NL values).  The axis-2 chunk count fills the
// device's resident blocks in whole rounds, counting the 2 rz warm-up
// planes a chunk reads before its first output plane.
template <class S>
bool march_plan(const FilterGeom& g, FilterMarch& m, size_t& lds, int64_t& blocks) {
  m = FilterMarch{};
  for (int t = 0; t < g.nt; ++t) {
    m.rx = std::max(m.rx, std::abs(g.off[t][0]));
    m.ry = std::max(m.ry, std::abs(g.off[t][1]));
    m.rz = std::max(m.rz, std::abs(g.off[t][2]));
  }
  m.LX = S::kTX + 2 * m.rx;
  m.LY = S::kTY + 2 * m.ry;
  m.D = 2 * m.rz + 1;
  lds = (size_t)m.D * (m.LX * m.LY + 64) * sizeof(double);
  if (m.LX * m.LY > 256 * S::kNL || lds > kMaxRingLds) return false;
  // store offsets within one tile (TY rows of a plane) fit in 31 bits
  if ((int64_t)S::kTY * g.n[0] * 8 >= 0x7fffffffLL) return false;
  for (int t = 0; t < g.nt; ++t) {
    m.lofs[t] = g.off[t][1] * m.LX + g.off[t][0];
    m.zo[t] = g.off[t][2];
  }
  m.tiles_x = (g.n[0] + S::kTX - 1) / S::kTX;
  m.tiles_y = (g.n[1] + S::kTY - 1) / S::kTY;
  const int64_t tiles = m.tiles_x * m.tiles_y;
  const void* kern = march_kernel_for<S>(g.nt);
  if (lds > 65536 &&
      hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
          hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const int64_t cap = kernel_capacity(kern, 256, lds);
  const int64_t per_round = std::max<int64_t>(1, cap / tiles);
  const int64_t nch =
      march_chunks(tiles, g.n[2], 2 * m.rz, cap, std::max<int64_t>(1, per_round / 2),
                   4 * per_round, true);
  m.chunk = (g.n[2] + nch - 1) / nch;
  blocks = tiles * ((g.n[2] + m.chunk - 1) / m.chunk);
  return blocks > 0 && blocks <= 0x7fffffffLL;
}

template <class S>
hipError_t launch_march(const FilterPlan& p, const double* in, double* out, hipStream_t s) {
  const dim3 grid((unsigned)p.blocks), block(S::kBX, S::kBY);
  switch (p.g.nt) {
#define FDB_MARCH_NT(N)                                                                      \
  case N:                                                                                    \
    hipLaunchKernelGGL((filter_march_kernel<N, S>), grid, block, p.lds, s, in, out, p.g, p.m); \
    break;
    FDB_MARCH_NT(5)
    FDB_MARCH_NT(7)
    FDB_MARCH_NT(9)
    FDB_MARCH_NT(13)
    FDB_MARCH_NT(27)
#undef FDB_MARCH_NT
    default:
      hipLaunchKernelGGL((filter_march_kernel<0, S>), grid, block, p.lds, s, in, out, p.g, p.m);
      break;
  }
  return hipGetLastError();
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

hipError_t plan_filter(const FilterGeom& g, FilterPlan* p) {
  // 2-D grids (no axis-2 extent or offsets): rows marched along axis 1,
  // the grid viewed as n0 x 1 x n1 (same memory order).  3-D grids with
  // n1 >= 8: full star stencils of radius 2 by the star kernel, others by
  // planes marched along axis 2 (rings up to 160 KB of LDS).
  // Otherwise (1-D grids, stencils wider than the ring allows): the
  // gather kernel.
  *p = FilterPlan{};
  bool flat = g.n[2] == 1 && g.n[1] > 1;
  for (int t = 0; t < g.nt && flat; ++t) flat = g.off[t][2] == 0;
  if (flat) {
    FilterGeom h = g;
    h.n[1] = 1;
    h.n[2] = g.n[1];
    for (int t = 0; t < g.nt; ++t) {
      h.off[t][1] = 0;
      h.off[t][2] = g.off[t][1];
    }
    for (int k = 0; k < kFilterMaxOffset; ++k) {
      h.neg[1][k] = 0;
      h.neg[2][k] = g.neg[1][k];
    }
    if (march_plan<March2D>(h, p->m, p->lds, p->blocks)) {
      p->kind = kFilterRows;
      p->g = h;
      return hipSuccess;
    }
  }
  const int star = g.n[2] > 1 ? star_radius(g) : 0;
  if (star == 2 && star_plan<2>(g, p->m, p->blocks)) {
    p->kind = kFilterStar;
    p->g = g;
    p->lds = 0;  // static
    return hipSuccess;
  }
  if (g.n[1] >= March3D::kTY && march_plan<March3D>(g, p->m, p->lds, p->blocks)) {
    p->kind = kFilterPlanes;
    p->g = g;
    return hipSuccess;
  }
  if (g.n[1] >= March3DMid::kTY && march_plan<March3DMid>(g, p->m, p->lds, p->blocks)) {
    p->kind = kFilterPlanesMid;
    p->g = g;
    return hipSuccess;
  }
  if (g.n[1] >= March3DWide::kTY && march_plan<March3DWide>(g, p->m, p->lds, p->blocks)) {
    p->kind = kFilterPlanesWide;
    p->g = g;
    return hipSuccess;
  }
  p->kind = kFilterGather;
  p->g = g;
  p->lds = 0;
  const int64_t xblocks = (g.n[0] + kFilterThreads - 1) / kFilterThreads;
  p->blocks = xblocks * g.n[1] * g.n[2];
  if (p->blocks <= 0 || p->blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  return hipSuccess;
}

hipError_t launch_filter(const FilterPlan& p, const double* in, double* out, hipStream_t s) {
  if (p.kind == kFilterRows) return launch_march<March2D>(p, in, out, s);
  if (p.kind == kFilterPlanes) return launch_march<March3D>(p, in, out, s);
  if (p.kind == kFilterPlanesMid) return launch_march<March3DMid>(p, in, out, s);
  if (p.kind == kFilterPlanesWide) return launch_march<March3DWide>(p, in, out, s);
  if (p.kind == kFilterStar) {
    const dim3 grid((unsigned)p.blocks), block(64, 4);
    hipLaunchKernelGGL(filter_star_kernel<2>, grid, block, 0, s, in, out, p.g, p.m);
    return hipGetLastError();
  }
  const FilterGeom& g = p.g;
  const int64_t xblocks = (g.n[0] + kFilterThreads - 1) / kFilterThreads;
  const dim3 grid((unsigned)p.blocks), block(kFilterThreads);
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
