// fdb_kernels.hip -- CDNA4 (gfx950) kernels of the upwind advection path.
//
// Arithmetic contract (pletzer/fidibench upwind/cxx/upwind.cxx:64-84): per
// cell i and axis j = 0, 1, 2 in that order,
//     g = g - c_j * (old[up_j(i)] - old[i]),   g starting at old[i],
// with c_j = ((dt*v_j)*u_j)/dx_j computed once on the host in FP64.  The
// whole file is compiled without floating-point contraction (the pragma
// below and -ffp-contract=off), so no multiply-add is fused and the FP64
// results equal the reference's bit for bit.
//
// Layout: row-major, axis 2 contiguous (upwind.cxx:39-44).  Two buffers
// ping-pong between steps instead of the reference's per-step oldF copy.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <mutex>

#include "fdb_internal.h"
#include "../../include/fidibench_amd/fdb_upwind.h"

#pragma clang fp contract(off)

namespace fdb {
namespace {

// ------------------------------------------------------------ storage types

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Storage type TS and the type arithmetic is done in (M).
template <typename TS> struct St;
// rnd(x) == ld(st(x)): x rounded to the storage type, kept in M.
template <> struct St<double> {
  using M = double;
  __device__ static M ld(double x) { return x; }
  __device__ static double st(M x) { return x; }
  __device__ static M rnd(M x) { return x; }
};
template <> struct St<float> {
  using M = float;
  __device__ static M ld(float x) { return x; }
  __device__ static float st(M x) { return x; }
  __device__ static M rnd(M x) { return x; }
};
// bf16 rounding on the device is gfx950's v_cvt_pk_bf16_f32: round to
// nearest, ties to even, the rule the host side (fdb_capi.cpp) and the
// oracle implement in software.  rnd converts the pair (+0, x): the bf16
// of x lands in the high half over sixteen zero bits, which is the binary32
// value of that bf16 -- one instruction instead of a convert and a shift.
template <> struct St<uint16_t> {
  using M = float;
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef __bf16 b2 __attribute__((ext_vector_type(2)));
  __device__ static M ld(uint16_t x) { return bf16_to_f32(x); }
  __device__ static uint16_t st(M x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
  __device__ static M rnd(M x) {
    const f2 v = {0.0f, x};
    return __builtin_bit_cast(float, __builtin_convertvector(v, b2));
  }
};

// Stochastic rounding of step results for bf16 storage (FDB_BF16_SR): 16
// random bits are added below the kept bits of the binary32 value, then
// truncated, so x rounds up with probability equal to the discarded
// fraction (unbiased).  Cells pair along axis 1: the pair of (i0, i1, i2)
// is the cell (i0, i1 & ~1, i2), named by its flat index.  The bits are
// one half of a pair word: lowbias32 (C. Wellons' 32-bit integer hash) of
// the pair index, xor a per-step key made on the host from the seed and
// the step's index, through one multiply-xorshift round (sr_round): the
// low half for the even row, the high half for the odd one (one word per
// two cells; a thread or lane holds both rows of a pair, so no exchange is
// needed).  The pair's lowbias32 does not depend on the step, so a fused
// pass hashes each pair once per plane and pays one multiply per pair and
// step.  Inf stays Inf, NaN stays quiet.
__host__ __device__ inline uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// x ^= x >> 16, x *= 0x7feb352d, x ^= x >> 16: lowbias32's first round
// and its closing shift, which carries the product's high half into the
// low half (the even cell's bits).
__device__ inline uint32_t sr_round(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  return x ^ (x >> 16);
}

// sr_round(w ^ k) with its first, xor-linear step done on each operand
// beforehand: sr_round_pre(w ^ (w >> 16) ^ k ^ (k >> 16)) == sr_round(w ^ k).
__host__ __device__ inline uint32_t sr_fold(uint32_t x) { return x ^ (x >> 16); }
__device__ inline uint32_t sr_round_pre(uint32_t x) {
  x *= 0x7feb352du;
  return x ^ (x >> 16);
}

// lowbias32 of a pair index folded to 32 bits: the step-independent part.
__device__ inline uint32_t sr_pair_word(uint64_t pair) {
  return lowbias32((uint32_t)(pair ^ (pair >> 32)));
}

__device__ inline uint32_t sr_pair_hash(uint64_t pair, uint32_t key) {
  return sr_round(sr_pair_word(pair) ^ key);
}

// x rounded with 16 random bits r16: (bits(x) + r16) >> 16 for finite x
// (a carry into the exponent rounds up to the next binade, or to Inf past
// the largest finite value), Inf kept, NaN kept quiet.  gfx950's
// v_cvt_sr_bf16_f32 does exactly this with the random bits taken from bits
// 31:16 of its second operand (checked against this rule over 10^6 random
// values and every special class: NaN payloads, +-Inf, +-0, denormals,
// the largest finite value), so it is called with r16 << 16.
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint16_t bf16_sr_bits(float x, uint32_t r16) {
  const s16x2 z = {0, 0};
  return (uint16_t)__builtin_bit_cast(uint32_t,
                                      __builtin_amdgcn_cvt_sr_bf16_f32(z, x, r16 << 16, false));
}

// Two values of one row rounded with their pairs' words h0, h1 (the even
// row takes the low 16 bits, the odd row the high 16); packed low = first.
__device__ inline uint32_t bf16x2_sr(float a, float b, uint32_t h0, uint32_t h1, bool odd) {
  s16x2 v = {0, 0};
  v = __builtin_amdgcn_cvt_sr_bf16_f32(v, a, odd ? h0 : h0 << 16, false);
  v = __builtin_amdgcn_cvt_sr_bf16_f32(v, b, odd ? h1 : h1 << 16, true);
  return __builtin_bit_cast(uint32_t, v);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two floats to two bf16 in one 32-bit word (low half = first), one
// v_cvt_pk_bf16_f32.
__device__ inline uint32_t pack_bf16x2(float lo, float hi) {
  const f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

template <typename TS, int VEC>
struct alignas(sizeof(TS) * VEC) Pack {
  TS v[VEC];
};

template <typename TS, int VEC>
__device__ inline void load_pack(const TS* p, typename St<TS>::M (&out)[VEC]) {
  const Pack<TS, VEC> pk = *reinterpret_cast<const Pack<TS, VEC>*>(p);
#pragma unroll
  for (int v = 0; v < VEC; ++v) out[v] = St<TS>::ld(pk.v[v]);
}

template <typename TS, int VEC>
__device__ inline void store_pack(TS* p, const typename St<TS>::M (&in)[VEC]) {
  Pack<TS, VEC> pk;
#pragma unroll
  for (int v = 0; v < VEC; ++v) pk.v[v] = St<TS>::st(in[v]);
  *reinterpret_cast<Pack<TS, VEC>*>(p) = pk;
}

// Value of the neighbouring lane through a DPP wavefront shift (a VALU
// move, no LDS round trip): lane_below gives lane L the value of lane
// L-1, lane_above that of lane L+1; lane 0 (resp. 63) keeps its own.
template <int CTRL, typename T>
__device__ inline T dpp_shift(T x) {
  static_assert(sizeof(T) % 4 == 0, "32-bit lanes");
  constexpr int W = sizeof(T) / 4;
  struct Words { int w[W]; };
  Words a = __builtin_bit_cast(Words, x), r;
#pragma unroll
  for (int i = 0; i < W; ++i) r.w[i] = __builtin_amdgcn_update_dpp(a.w[i], a.w[i], CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(T, r);
}
template <typename T> __device__ inline T lane_below(T x) { return dpp_shift<0x138>(x); }  // wave_shr:1
template <typename T> __device__ inline T lane_above(T x) { return dpp_shift<0x130>(x); }  // wave_shl:1

// The same shifts where the edge lane's result is never used: one DPP move
// per 32 bits, no copy of the old value (the edge lane reads 0).
template <int CTRL, typename T>
__device__ inline T dpp_shift_any(T x) {
  constexpr int W = sizeof(T) / 4;
  struct Words { int w[W]; };
  Words a = __builtin_bit_cast(Words, x), r;
#pragma unroll
  for (int i = 0; i < W; ++i) r.w[i] = __builtin_amdgcn_mov_dpp(a.w[i], CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(T, r);
}

// Marks a loaded value as used here, so the compiler waits for its load
// at this point.  Called on the loads issued before a march loop: without
// it, the wait for them is placed inside the loop where the loop's first
// use is, and the merged wait count there also stalls on the next plane's
// loads issued at the top of each iteration (no prefetch overlap).
template <typename T>
__device__ inline void settle(const T& x) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "one or two VGPRs");
  asm volatile("" ::"v"(x));
}

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it stay in flight (a full __syncthreads() fence would wait for
// them).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline int64_t wrap(int64_t x, int64_t n) {
  return x < 0 ? x + n : (x >= n ? x - n : x);
}

// Coefficients in the arithmetic type: FP64 as computed, or rounded once
// to float for the reduced-precision storage types.
template <typename M> struct Coef { M c[3]; };

// Everything one march kernel needs, passed by value.
struct MarchArgs {
  int64_t n0, n1, n2;
  int dir0, dir1, dir2;     // neighbour offsets: +u_j forward, -u_j adjoint
  int act0, act1, act2;
  int64_t chunk;            // planes per block along axis 0
  int64_t tiles2, tiles1;   // tiles along axis 2 and axis 1
  uint32_t sr_key;          // stochastic-rounding key of this step (BF16_SR)
};

// Adds npart per-block {sum, comp} x3 partials into acc (device, 6
// doubles) in a fixed order: one block, deterministic.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ part,
                                                              int64_t npart, int npairs,
                                                              double* __restrict__ acc,
                                                              int accumulate) {
  __shared__ double red[256][6];
  const int t = threadIdx.x;
  double s[3] = {0, 0, 0}, e[3] = {0, 0, 0};
  for (int64_t i = t; i < npart; i += 256) {
    for (int j = 0; j < npairs; ++j) {
      const double x = part[i * 6 + 2 * j];
      const double sum = s[j] + x;
      const double bb = sum - s[j];
      e[j] += (s[j] - (sum - bb)) + (x - bb) + part[i * 6 + 2 * j + 1];
      s[j] = sum;
    }
  }
  for (int j = 0; j < 3; ++j) {
    red[t][2 * j] = s[j];
    red[t][2 * j + 1] = e[j];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      for (int j = 0; j < npairs; ++j) {
        const double s1 = red[t][2 * j], s2 = red[t + w][2 * j];
        const double sum = s1 + s2;
        const double bb = sum - s1;
        const double err = (s1 - (sum - bb)) + (s2 - bb);
        red[t][2 * j] = sum;
        red[t][2 * j + 1] = red[t][2 * j + 1] + red[t + w][2 * j + 1] + err;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int j = 0; j < npairs; ++j) {
      double s1 = red[0][2 * j], c1 = red[0][2 * j + 1];
      if (accumulate) {
        const double s2 = acc[2 * j];
        const double sum = s1 + s2;
        const double bb = sum - s1;
        const double err = (s1 - (sum - bb)) + (s2 - bb);
        c1 = c1 + acc[2 * j + 1] + err;
        s1 = sum;
      }
      acc[2 * j] = s1;
      acc[2 * j + 1] = c1;
    }
  }
}

// ----------------------------------------------------------- reductions --

template <typename TS>
__device__ inline double widen(const TS* x, uint64_t i) {
  return (double)St<TS>::ld(x[i]);
}

// Neumaier sum per thread, TwoSum tree per block -> partials[block*6+{0,1}].
template <typename TS>
__global__ __launch_bounds__(256) void sum_kernel(const TS* __restrict__ x, uint64_t n,
                                                  double* __restrict__ partials) {
  __shared__ double red[256][2];
  const int t = threadIdx.x;
  double s = 0.0, e = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < n; i += (uint64_t)gridDim.x * 256) {
    const double v = widen(x, i);
    const double sum = s + v;
    e += (__builtin_fabs(s) >= __builtin_fabs(v)) ? (s - sum) + v : (v - sum) + s;
    s = sum;
  }
  red[t][0] = s;
  red[t][1] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      const double s1 = red[t][0], s2 = red[t + w][0];
      const double sum = s1 + s2;
      const double bb = sum - s1;
      red[t][1] = red[t][1] + red[t + w][1] + ((s1 - (sum - bb)) + (s2 - bb));
      red[t][0] = sum;
    }
    __syncthreads();
  }
  if (t < 2) partials[blockIdx.x * 6 + t] = red[0][t];
}

// lam = f - target; partial sums of 0.5*(f-target)^2 (product error kept).
template <typename TF, typename TL>
__global__ __launch_bounds__(256) void misfit_kernel(const TF* __restrict__ f,
                                                     const TF* __restrict__ tg, uint64_t n,
                                                     TL* __restrict__ lam,
                                                     double* __restrict__ partials) {
  __shared__ double red[256][2];
  const int t = threadIdx.x;
  double s = 0.0, e = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < n; i += (uint64_t)gridDim.x * 256) {
    const double d = widen(f, i) - widen(tg, i);
    lam[i] = St<TL>::st((typename St<TL>::M)d);
    const double h = 0.5 * d;  // exact scaling
    const double pr = h * d;
    const double pe = __builtin_fma(h, d, -pr);
    const double sum = s + pr;
    e += (__builtin_fabs(s) >= __builtin_fabs(pr)) ? (s - sum) + pr : (pr - sum) + s;
    e += pe;
    s = sum;
  }
  red[t][0] = s;
  red[t][1] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      const double s1 = red[t][0], s2 = red[t + w][0];
      const double sum = s1 + s2;
      const double bb = sum - s1;
      red[t][1] = red[t][1] + red[t + w][1] + ((s1 - (sum - bb)) + (s2 - bb));
      red[t][0] = sum;
    }
    __syncthreads();
  }
  if (t < 2) partials[blockIdx.x * 6 + t] = red[0][t];
}

// Per-block {max|a-b|, sum (a-b)^2 (Neumaier pair), max|b|} of two fields
// widened to double -> partials[block*6 + 0..3].
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void compare_kernel(const TA* __restrict__ a,
                                                      const TB* __restrict__ b, uint64_t n,
                                                      double* __restrict__ partials) {
  __shared__ double red[256][4];
  const int t = threadIdx.x;
  double mx = 0.0, s = 0.0, e = 0.0, mb = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + t; i < n; i += (uint64_t)gridDim.x * 256) {
    const double x = widen(a, i), y = widen(b, i);
    const double d = x - y;
    mx = fmax(mx, __builtin_fabs(d));
    mb = fmax(mb, __builtin_fabs(y));
    const double q = d * d;
    const double sum = s + q;
    e += (s >= q) ? (s - sum) + q : (q - sum) + s;
    s = sum;
  }
  red[t][0] = mx;
  red[t][1] = s;
  red[t][2] = e;
  red[t][3] = mb;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      red[t][0] = fmax(red[t][0], red[t + w][0]);
      const double s1 = red[t][1], s2 = red[t + w][1];
      const double sum = s1 + s2;
      const double bb = sum - s1;
      red[t][2] = red[t][2] + red[t + w][2] + ((s1 - (sum - bb)) + (s2 - bb));
      red[t][1] = sum;
      red[t][3] = fmax(red[t][3], red[t + w][3]);
    }
    __syncthreads();
  }
  if (t < 4) partials[blockIdx.x * 6 + t] = red[0][t];
}

__global__ void zero_kernel(double* p, int n) {
  const int t = threadIdx.x;
  if (t < n) p[t] = 0.0;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void convert_kernel(const TA* __restrict__ a, TB* __restrict__ b,
                                                      uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    const double x = (double)St<TA>::ld(a[i]);
    b[i] = St<TB>::st((typename St<TB>::M)x);
  }
}

// --------------------------------------------------------- initial fields

// SplitMix64 (Steele, Lea & Flood 2014), counter form: draw i of `seed`.
__host__ __device__ inline uint64_t splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct InitArgs {
  int64_t n[3];
  int act[3];
  double dx[3];
  int kind;
  int nparams;
  uint64_t seed;
  double p[5 * 8 + 1];
};

template <typename TS>
__global__ __launch_bounds__(256) void init_kernel(TS* __restrict__ f, InitArgs a) {
  const uint64_t ntot = (uint64_t)(a.n[0] * a.n[1] * a.n[2]);
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < ntot;
       i += (uint64_t)gridDim.x * 256) {
    const int64_t i2 = (int64_t)(i % (uint64_t)a.n[2]);
    const int64_t r = (int64_t)(i / (uint64_t)a.n[2]);
    const int64_t i1 = r % a.n[1];
    const int64_t i0 = r / a.n[1];
    const double x[3] = {((double)i0 + 0.5) * a.dx[0], ((double)i1 + 0.5) * a.dx[1],
                         ((double)i2 + 0.5) * a.dx[2]};
    double val = 0.0;
    if (a.kind == FDB_INIT_DELTA) {
      val = (i == 0) ? 1.0 : 0.0;
    } else if (a.kind == FDB_INIT_UNIFORM) {
      val = a.p[0] + (a.p[1] - a.p[0]) * ((double)(splitmix64(a.seed, i) >> 11) * 0x1.0p-53);
    } else if (a.kind == FDB_INIT_SPHERE || a.kind == FDB_INIT_GAUSS) {
      double s = 0.0;
      bool first = true;
      for (int j = 0; j < 3; ++j) {
        if (!a.act[j]) continue;
        const double d = x[j] - a.p[j];
        s = first ? d * d : s + d * d;
        first = false;
      }
      if (a.kind == FDB_INIT_SPHERE) {
        val = (s < a.p[3] * a.p[3]) ? 1.0 : 0.0;
        if (a.p[4] != 0.0)
          val = val + a.p[4] * ((double)(splitmix64(a.seed, i) >> 11) * 0x1.0p-53);
      } else {
        val = a.p[4] * exp(-s / a.p[3]);
      }
    } else if (a.kind == FDB_INIT_GAUSS_SUM) {
      const int m = (int)a.p[0];
      for (int k = 0; k < m; ++k) {
        const double* q = a.p + 1 + 5 * k;
        double s = 0.0;
        bool first = true;
        for (int j = 0; j < 3; ++j) {
          if (!a.act[j]) continue;
          const double d = x[j] - q[j];
          s = first ? d * d : s + d * d;
          first = false;
        }
        val = val + q[4] * exp(-s / (q[3] * q[3]));
      }
    }
    f[i] = St<TS>::st((typename St<TS>::M)val);
  }
}

// ---------------------------------------------- one-step forward, rows --
//
// The streaming forward kernel.  Block = 64 lanes (axis 2, VEC cells per
// lane, one 16-byte access) x WY waves (axis 1); every thread holds R
// consecutive rows of axis 1 in registers, so the axis-1 neighbour is a
// register of the same thread except for one halo row per plane.  The
// block marches along axis 0 through `chunk` planes in the direction that
// makes the axis-0 neighbour the plane it loaded one iteration earlier,
// and loads plane p+1 while it computes plane p.  HBM traffic is one read
// and one write of each cell plus the halo rows and one extra plane per
// chunk, which mostly hit L2 (tiles sharing them run on the same XCD).
template <int BYTES> struct RawT;
template <> struct RawT<2> { using T = uint16_t; };
template <> struct RawT<4> { using T = uint32_t; };
template <> struct RawT<8> { using T = uint32_t __attribute__((ext_vector_type(2))); };
template <> struct RawT<16> { using T = uint32_t __attribute__((ext_vector_type(4))); };

template <typename TS, int VEC, bool NT>
__device__ inline void store_pack_nt(TS* p, const typename St<TS>::M (&in)[VEC]) {
  Pack<TS, VEC> pk;
  if constexpr (sizeof(TS) == 2 && VEC % 2 == 0) {
    uint32_t w[VEC / 2];
#pragma unroll
    for (int v = 0; v < VEC / 2; ++v) w[v] = pack_bf16x2(in[2 * v], in[2 * v + 1]);
    pk = __builtin_bit_cast(Pack<TS, VEC>, w);
  } else {
#pragma unroll
    for (int v = 0; v < VEC; ++v) pk.v[v] = St<TS>::st(in[v]);
  }
  using R = typename RawT<sizeof(Pack<TS, VEC>)>::T;
  const R raw = __builtin_bit_cast(R, pk);
  if constexpr (NT)
    __builtin_nontemporal_store(raw, reinterpret_cast<R*>(p));
  else
    *reinterpret_cast<R*>(p) = raw;
}

// bf16 store with stochastic rounding (cell = flat index of the first).
// VEC consecutive cells of one row rounded with their pairs' words h[v]
// (odd: the row is the odd one of its pairs).
template <int VEC, bool NT>
__device__ inline void store_pack_sr(uint16_t* p, const float (&in)[VEC],
                                     const uint32_t (&h)[VEC], bool odd) {
  Pack<uint16_t, VEC> pk;
  if constexpr (VEC % 2 == 0) {
    uint32_t w[VEC / 2];
#pragma unroll
    for (int v = 0; v < VEC; v += 2) w[v / 2] = bf16x2_sr(in[v], in[v + 1], h[v], h[v + 1], odd);
    pk = __builtin_bit_cast(Pack<uint16_t, VEC>, w);
  } else {
#pragma unroll
    for (int v = 0; v < VEC; ++v) pk.v[v] = bf16_sr_bits(in[v], odd ? h[v] >> 16 : h[v] & 0xffffu);
  }
  using R = typename RawT<sizeof(Pack<uint16_t, VEC>)>::T;
  const R raw = __builtin_bit_cast(R, pk);
  if constexpr (NT)
    __builtin_nontemporal_store(raw, reinterpret_cast<R*>(p));
  else
    *reinterpret_cast<R*>(p) = raw;
}

// Per-thread compensated accumulators of the three velocity-gradient dot
// products (adjoint kernel only).
struct VAcc {
  double s[3] = {0.0, 0.0, 0.0}, e[3] = {0.0, 0.0, 0.0};
  __device__ void add(int j, double x) {  // Neumaier
    const double t = s[j] + x;
    e[j] += (__builtin_fabs(s[j]) >= __builtin_fabs(x)) ? (s[j] - t) + x : (x - t) + s[j];
    s[j] = t;
  }
};

// ADJ = true: the transpose step (a.dir* = -u) plus, when ft != nullptr,
// the products f_t[i] * (lam_next[i + dir_j e_j] - lam_next[i]) summed per
// plane in FP64 and added to acc (sum_i lam_next[i] (f_t[up_j(i)] - f_t[i])
// by the periodic shift i -> i + dir_j e_j).
template <typename TS, typename TF, int VEC, int R, int WY, bool NT, bool ADJ, bool SR>
__device__ inline void rows_tile(const TS* __restrict__ in, TS* __restrict__ out,
                                 const MarchArgs& a, const Coef<typename St<TS>::M>& cf,
                                 int64_t b, const TF* __restrict__ ft, VAcc& acc) {
  using M = typename St<TS>::M;
  const bool with_f = ADJ && ft != nullptr;
  const int lane = threadIdx.x;
  const int wave = threadIdx.y;

  const int64_t ti2 = b % a.tiles2;
  b /= a.tiles2;
  const int64_t ti1 = b % a.tiles1;
  const int64_t ch = b / a.tiles1;

  const int64_t i2 = (ti2 * 64 + lane) * VEC;
  const bool col_ok = i2 < a.n2;
  const int64_t i2c = col_ok ? i2 : 0;
  const int64_t r0 = (ti1 * WY + wave) * R;
  const int64_t nvalid = r0 < a.n1 ? (a.n1 - r0 < R ? a.n1 - r0 : R) : 0;  // rows to store

  // Row offsets within a plane (rows past the end re-read the last row).
  int64_t roff[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t row = (r0 + r < a.n1) ? r0 + r : a.n1 - 1;
    roff[r] = row * a.n2;
  }
  // The one axis-1 neighbour row outside the thread's rows, and the row
  // that needs it.
  const bool up1 = a.dir1 < 0;
  const int rl = nvalid > 0 ? (int)nvalid - 1 : 0;
  // (r0 < n1 whenever nvalid > 0, so both wraps stay in [0, n1); a wave
  // past the last row reads row 0 and stores nothing.)
  const int64_t hrow =
      nvalid == 0 ? 0 : (up1 ? wrap(r0 - 1, a.n1) : wrap(r0 + rl + 1, a.n1));
  const int64_t hoff = hrow * a.n2 + i2c;
  // Axis-2 neighbour from outside this lane's cells.
  const bool lo2 = a.dir2 < 0;
  const bool edge = lo2 ? (lane == 0 || i2c == 0) : (lane == 63 || i2c + VEC >= a.n2);
  const int64_t i2e = lo2 ? wrap(i2c - 1, a.n2) : wrap(i2c + VEC, a.n2);

  const int64_t plane = a.n1 * a.n2;
  const int64_t p_lo = ch * a.chunk;
  const int64_t p_hi = (p_lo + a.chunk < a.n0) ? p_lo + a.chunk : a.n0;
  const int64_t step = a.dir0 < 0 ? 1 : -1;
  int64_t p = a.dir0 < 0 ? p_lo : p_hi - 1;
  const int64_t cnt = p_hi - p_lo;

  M prev[R][VEC];
  {
    const TS* base = in + wrap(p + a.dir0, a.n0) * plane + i2c;
#pragma unroll
    for (int r = 0; r < R; ++r) load_pack<TS, VEC>(base + roff[r], prev[r]);
  }
  using MF = typename St<TF>::M;
  M cur[R][VEC], hal[VEC], ed[R];
  MF fc[ADJ ? R : 1][VEC];
  auto load_plane = [&](int64_t pp, M (&c)[R][VEC], M (&h)[VEC], M (&e)[R]) {
    const TS* base = in + pp * plane;
#pragma unroll
    for (int r = 0; r < R; ++r) load_pack<TS, VEC>(base + roff[r] + i2c, c[r]);
    load_pack<TS, VEC>(base + hoff, h);
#pragma unroll
    for (int r = 0; r < R; ++r) e[r] = edge ? St<TS>::ld(base[roff[r] + i2e]) : M(0);
  };
  auto load_f = [&](int64_t pp, MF (&f)[ADJ ? R : 1][VEC]) {
    if constexpr (ADJ) {
      if (with_f) {
        const TF* fb = ft + pp * plane + i2c;
#pragma unroll
        for (int r = 0; r < R; ++r) load_pack<TF, VEC>(fb + roff[r], f[r]);
      }
    }
  };
  load_plane(p, cur, hal, ed);
  load_f(p, fc);
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      settle(prev[r][v]);
      settle(cur[r][v]);
      if constexpr (ADJ) settle(fc[r][v]);
    }
    settle(ed[r]);
  }
#pragma unroll
  for (int v = 0; v < VEC; ++v) settle(hal[v]);

  for (int64_t k = 0; k < cnt; ++k) {
    M nxt[R][VEC], nhal[VEC], ned[R];
    MF fn[ADJ ? R : 1][VEC];
    if (k + 1 < cnt) {
      load_plane(p + step, nxt, nhal, ned);
      load_f(p + step, fn);
    }
    double ps[3] = {0.0, 0.0, 0.0};
    // axis-2 neighbours: lane shifts are done by all lanes together
    M n2[R][VEC];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (lo2) {
        const M sh = lane_below(cur[r][VEC - 1]);
        n2[r][0] = edge ? ed[r] : sh;
#pragma unroll
        for (int v = 1; v < VEC; ++v) n2[r][v] = cur[r][v - 1];
      } else {
        const M sh = lane_above(cur[r][0]);
        n2[r][VEC - 1] = edge ? ed[r] : sh;
#pragma unroll
        for (int v = 0; v < VEC - 1; ++v) n2[r][v] = cur[r][v + 1];
      }
    }
    TS* obase = out + p * plane + i2c;
    uint32_t hw[SR ? VEC : 1];  // SR: the pair words of the current row pair
#pragma unroll
    for (int r = 0; r < R; ++r) {
      M o[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const M x = cur[r][v];
        M n1;
        if (up1)
          n1 = (r == 0) ? hal[v] : cur[r > 0 ? r - 1 : 0][v];
        else
          n1 = (r == rl) ? hal[v] : cur[r + 1 < R ? r + 1 : R - 1][v];
        M g = x;
        if (a.act0) g = g - cf.c[0] * (prev[r][v] - x);
        if (a.act1) g = g - cf.c[1] * (n1 - x);
        if (a.act2) g = g - cf.c[2] * (n2[r][v] - x);
        o[v] = g;
        if constexpr (ADJ) {
          if (with_f && col_ok && r < nvalid) {
            const double xd = (double)x, fv = (double)fc[r][v];
            ps[0] += fv * ((double)prev[r][v] - xd);
            ps[1] += fv * ((double)n1 - xd);
            ps[2] += fv * ((double)n2[r][v] - xd);
          }
        }
        prev[r][v] = x;
      }
      if constexpr (SR) {
        // Pair words of this row's cells, shared with the row's partner
        // when the thread holds both (computed at the pair's first row).
        const int64_t grow = r0 + r;
        const bool odd = (grow & 1) != 0;
        if (!odd || r == 0) {
          const uint64_t base = (uint64_t)(p * plane + (grow - (odd ? 1 : 0)) * a.n2 + i2c);
#pragma unroll
          for (int v = 0; v < VEC; ++v) hw[v] = sr_pair_hash(base + (uint64_t)v, a.sr_key);
        }
        if (col_ok && r < nvalid) store_pack_sr<VEC, NT>(obase + roff[r], o, hw, odd);
      } else if (col_ok && r < nvalid) {
        store_pack_nt<TS, VEC, NT>(obase + roff[r], o);
      }
    }
    if constexpr (ADJ) {
      if (with_f) {
#pragma unroll
        for (int j = 0; j < 3; ++j) acc.add(j, ps[j]);
      }
    }
    if (k + 1 < cnt) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) cur[r][v] = nxt[r][v];
        ed[r] = ned[r];
        if constexpr (ADJ) {
#pragma unroll
          for (int v = 0; v < VEC; ++v) fc[r][v] = fn[r][v];
        }
      }
#pragma unroll
      for (int v = 0; v < VEC; ++v) hal[v] = nhal[v];
    }
    p += step;
  }
}

// Block reduction of the per-thread accumulators into partials[block*6].
template <int NT_>
__device__ inline void block_reduce_vacc(const VAcc& acc, double* __restrict__ partials) {
  __shared__ double red[NT_][6];
  const int t = threadIdx.y * 64 + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    red[t][2 * j] = acc.s[j];
    red[t][2 * j + 1] = acc.e[j];
  }
  __syncthreads();
  for (int w = NT_ / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double s1 = red[t][2 * j], s2 = red[t + w][2 * j];
        const double sum = s1 + s2;
        const double bb = sum - s1;
        red[t][2 * j + 1] =
            red[t][2 * j + 1] + red[t + w][2 * j + 1] + ((s1 - (sum - bb)) + (s2 - bb));
        red[t][2 * j] = sum;
      }
    }
    __syncthreads();
  }
  if (t < 6) partials[blockIdx.x * 6 + t] = red[0][t];
}

// Tiles are numbered axis-2 fastest, then axis 1, then chunk of axis 0.
// Hardware hands blocks to the 8 XCDs round-robin by block id; each XCD
// gets a contiguous range of tile numbers so tiles that share halo rows
// and chunk-boundary planes share that XCD's L2.  With gridDim.x < ntiles
// the grid is persistent: block (xcd, l) takes tiles start+l, start+l+n8..
template <typename TS, typename TF, int VEC, int R, int WY, bool NT, bool ADJ, bool SR>
__global__ __launch_bounds__(64 * WY) void rows_kernel(const TS* __restrict__ in,
                                                       TS* __restrict__ out, MarchArgs a,
                                                       Coef<typename St<TS>::M> cf,
                                                       int64_t ntiles, const TF* __restrict__ ft,
                                                       double* __restrict__ partials) {
  const int64_t nb = gridDim.x;
  VAcc acc;
  if ((nb & 7) == 0) {
    const int64_t xcd = blockIdx.x & 7, l = blockIdx.x >> 3, n8 = nb >> 3;
    const int64_t per = (ntiles + 7) >> 3;
    const int64_t lo = xcd * per;
    const int64_t hi = lo + per < ntiles ? lo + per : ntiles;
    for (int64_t t = lo + l; t < hi; t += n8)
      rows_tile<TS, TF, VEC, R, WY, NT, ADJ, SR>(in, out, a, cf, t, ft, acc);
  } else {
    for (int64_t t = blockIdx.x; t < ntiles; t += nb)
      rows_tile<TS, TF, VEC, R, WY, NT, ADJ, SR>(in, out, a, cf, t, ft, acc);
  }
  if constexpr (ADJ) {
    if (ft != nullptr) block_reduce_vacc<64 * WY>(acc, partials);
  }
}

// ------------------------------------------ temporally blocked forward --
//
// K steps per pass over HBM.  A block of WY waves owns an extended tile:
// 64 lanes along axis 2 (one cell each) x WY*R rows of axis 1 (R rows in
// registers per wave), reaching K cells past its output region on the
// upwind side of both axes; it outputs (64-K) x (WY*R-K) cells per
// plane.  The block marches along axis 0 from upwind to downwind and, at
// every plane, computes levels 1..K of the step in registers.  Level l of
// plane p needs level l-1 of plane p -- rows of the same wave, the
// neighbouring lane (DPP), and one edge row of the neighbouring wave,
// handed over through LDS at every level -- and of plane p-upwind (kept
// per level in prev[l-1]).  It starts K planes before its chunk so that
// every level is valid when output begins.  Cells within l of the upwind
// tile edges hold wrong values at level l, but they never reach an output
// cell (an output cell's level-K value depends only on cells at most K
// away upwind).  Each value is produced by the same operations on the same
// inputs as K separate steps, so the result is bit-identical.
//
// The upwind signs of axes 1 and 2 are template parameters (LO1: the
// axis-1 neighbour is row r-1; LO2: the axis-2 neighbour is column c-1),
// so the plane loop has no direction branches.  Row offsets are uniform
// per wave (scalar registers); each lane adds one 32-bit column offset.
struct TbArgs {
  int64_t n0, n1, n2;
  int u0;
  int64_t chunk;               // output planes per tile along axis 0
  int64_t tiles2, tiles1, ntiles;
  uint32_t sr_key[kMaxTimeBlock];  // kTbSR: sr_fold(key) of level l's step
};

// Modes of the temporally blocked kernel.
//   kTbPlain     K forward steps, level K stored to out.
//   kTbStoreAll  also levels 1..K-1 stored, to aux.lv[0..K-2] (a forward
//                that keeps every state).
//   kTbAdjoint   the fused transpose (adjoint) pass.  The kernel is called
//                with the adjoint's directions (-u), so level l+1 =
//                A^T (level l), and aux.fs[l] is the forward state of that
//                adjoint step (level l of the pass consumes f_{t-l} when the
//                pass starts from lam_{t+1}).  At every output cell it also
//                sums f[i] * (lam[i + dir_j e_j] - lam[i]) for j = 0, 1, 2 --
//                the one-step adjoint's velocity products -- in FP64
//                (differences taken in FP64 as there), Neumaier-compensated
//                per thread and reduced per block into aux.partials[block*6].
//   kTbAdjLast   the fused transpose pass with the products of its last
//                level only, against aux.fs[0] (the forward state that
//                level's step consumes): K - 1 plain transpose steps, then
//                the one-step adjoint with products.
//   kTbSR        K forward steps in bf16 storage, each level rounded
//                stochastically with the bits a one-step launch of that step
//                draws (sr_pair_hash of the cell's pair and the step's key,
//                low half for the even row, high half for the odd one),
//                so the result is the one-step launches' bit for bit.  A
//                lane holds both rows of each of its pairs (even rows per
//                wave, even K, even n1), so each lane hashes one word per
//                pair of rows and level, with no exchange between lanes.
constexpr int kTbPlain = 0, kTbStoreAll = 1, kTbAdjoint = 2, kTbSR = 3, kTbAdjLast = 4;
struct TbAux {
  const void* fs[kMaxTimeBlock];
  void* lv[kMaxTimeBlock];
  double* partials;
};

// Raw buffer access to one row: the row's start (uniform) goes into the
// resource, the lane's byte offset into a VGPR, so no lane does 64-bit
// address arithmetic.  Offsets stay below the row length (no access is
// dropped); stores are nontemporal (aux 2).
template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t row_rsrc(const T* row, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(row), 0, bytes, 0x00020000);
}
using u32x2_t = uint32_t __attribute__((ext_vector_type(2)));
__device__ inline double buf_ld(__amdgpu_buffer_rsrc_t rs, uint32_t off, const double*) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
}
__device__ inline float buf_ld(__amdgpu_buffer_rsrc_t rs, uint32_t off, const float*) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}
__device__ inline uint16_t buf_ld(__amdgpu_buffer_rsrc_t rs, uint32_t off, const uint16_t*) {
  return __builtin_amdgcn_raw_buffer_load_b16(rs, off, 0, 0);
}
__device__ inline void buf_st_nt(double v, __amdgpu_buffer_rsrc_t rs, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), rs, off, 0, 2);
}
__device__ inline void buf_st_nt(float v, __amdgpu_buffer_rsrc_t rs, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, off, 0, 2);
}
__device__ inline void buf_st_nt(uint16_t v, __amdgpu_buffer_rsrc_t rs, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b16(v, rs, off, 0, 2);
}

// x mod n in [0, n) for any x (n > 0).
__device__ inline int64_t pmod(int64_t x, int64_t n) {
  x %= n;
  return x < 0 ? x + n : x;
}

template <typename TS, int R, int K, int WY, bool LO1, bool LO2, int MODE = kTbPlain,
          typename TF = TS>
__global__ __launch_bounds__(64 * WY) void tb_kernel(const TS* __restrict__ in,
                                                     TS* __restrict__ out, TbArgs a,
                                                     Coef<typename St<TS>::M> cf, TbAux adj) {
  using M = typename St<TS>::M;
  constexpr bool ALL = MODE == kTbStoreAll, SR = MODE == kTbSR, LASTP = MODE == kTbAdjLast;
  // PROD: velocity products (at every level, or at the last one only);
  // NF states held, one per level with products.
  constexpr bool PROD = MODE == kTbAdjoint || LASTP;
  constexpr int NF = MODE == kTbAdjoint ? K : 1;
  static_assert(!SR || sizeof(TS) == 2, "stochastic rounding is for bf16 storage");
  // FP64 forward passes exchange the edge rows one level ahead (PIPE; see
  // the level loop): their 200-240 VGPRs leave one 8-wave block per CU, so
  // the exchange's latency and the per-level barrier are not hidden by a
  // second block.  FP32 and bf16 blocks fit two per CU, which hides them,
  // and there the ordering below (exchange within the level) measured
  // faster (FP32 K=8 2480 vs 2160 G cell-updates/s at 512^3).
  // The last-level products (LASTP) fit this ordering too: row er's axis-1
  // neighbour is read at the start of the level, so every row's three
  // differences are at hand when it is computed.
  constexpr bool PIPE = MODE != kTbAdjoint && sizeof(M) == 8;
  // Edge row handed to the neighbouring wave, double-buffered by the
  // parity of the running level count (one barrier per level suffices).
  __shared__ M edge[2][WY][64];
  int64_t t = blockIdx.x;
  const int64_t nb = gridDim.x;
  // Hardware hands blocks to the 8 XCDs round-robin by block id: renumber
  // them so each XCD takes one contiguous range of tiles (neighbouring
  // tiles, whose halos overlap, then share that XCD's L2), for any grid.
  {
    const int64_t x = t & 7, q = nb >> 3, rm = nb & 7;
    t = x * q + (x < rm ? x : rm) + (t >> 3);
  }
  if (t >= a.ntiles) return;                               // whole block
  const int lane = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int64_t t2 = t % a.tiles2;
  const int64_t t1 = (t / a.tiles2) % a.tiles1;
  const int64_t ch = t / (a.tiles2 * a.tiles1);

  // Columns: output columns o2 .. o2+63-K; the halo is on the upwind side.
  const int64_t o2 = t2 * (64 - K);
  const int64_t cl = LO2 ? o2 - K + lane : o2 + lane;  // unwrapped column of this lane
  const uint32_t col = (uint32_t)pmod(cl, a.n2);
  const bool col_out = (LO2 ? lane >= K : lane < 64 - K) && cl >= 0 && cl < a.n2;
  const uint32_t cb = col * (uint32_t)sizeof(TS);         // byte offset in the row
  const int rbytes = (int)(a.n2 * (int64_t)sizeof(TS));  // row length (< 2^31)
  // Rows of this wave: block rows w*R .. w*R+R-1 of the extended tile,
  // whose output rows are o1 .. o1+WY*R-1-K.
  const int64_t o1 = t1 * (WY * R - K);
  const int64_t rb = (LO1 ? o1 - K : o1) + (int64_t)w * R;
  int64_t roff[R];  // uniform in-plane row offsets
  bool row_out[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t rl = rb + r;
    const int br = w * R + r;
    roff[r] = pmod(rl, a.n1) * a.n2;
    row_out[r] = (LO1 ? br >= K : br < WY * R - K) && rl >= 0 && rl < a.n1;
  }
  // SR: rows 2j and 2j+1 of the wave are the two rows of a pair (the
  // wave's first row is even: R, K and the tile's row count are even, and
  // an even n1 keeps the parity of wrapped rows); sword[j] is the pair
  // word (sr_fold) of this lane's column for the current plane.
  constexpr int J = (R + 1) / 2;
  static_assert(!SR || (R % 2 == 0 && K % 2 == 0), "SR passes pair whole rows");
  uint32_t sword[SR ? J : 1];
  // Planes: output p_lo .. p_hi-1; march K planes earlier (upwind side).
  const int64_t plane = a.n1 * a.n2;
  const int64_t p_lo = ch * a.chunk;
  const int64_t p_hi = (p_lo + a.chunk < a.n0) ? p_lo + a.chunk : a.n0;
  const int64_t step = a.u0 < 0 ? 1 : -1;
  const int64_t first = a.u0 < 0 ? p_lo - K : p_hi - 1 + K;
  const int64_t cnt = p_hi - p_lo + K;

  M prev[K][R];
#pragma unroll
  for (int l = 0; l < K; ++l)
#pragma unroll
    for (int r = 0; r < R; ++r) prev[l][r] = M(0);
  M x[R], nx[R];
  int64_t pw = pmod(first, a.n0);  // wrapped index of the current plane
  {
    const TS* base = in + pw * plane;
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = St<TS>::ld(buf_ld(row_rsrc(base + roff[r], rbytes), cb, in));
#pragma unroll
    for (int r = 0; r < R; ++r) settle(x[r]);
#pragma unroll
    for (int r = 0; r < R; ++r) nx[r] = x[r];  // defined where a prefetch is skipped
  }
  // Two planes per trip of the plane loop (the registers of prev[l] and of
  // level l alternate between planes instead of being copied at the end of
  // each plane) where that fits the registers: K <= 6, and the FP64
  // shapes sized for it, 56 row-levels per wave (8 rows at K = 7, 7 at
  // K = 8).  (FP64 with 9 rows at K = 7, 8 would spill; the 5-row K = 8
  // shape keeps one plane per trip within 128 VGPRs.)
  constexpr bool kTwoPlanes = !PROD && (K <= 6 || (sizeof(M) == 8 && R * K == 56));
  constexpr bool kCondPrefetch = PIPE && !kTwoPlanes;
  constexpr bool kWaitBeforeStores = (MODE == kTbPlain || MODE == kTbSR) && !kCondPrefetch;
  // PROD: fv[l] holds the state of level l at the current plane (LASTP:
  // fv[0], the last level's); after its level it is reloaded with the next
  // plane's (always a valid plane, so every load is unconditional).  The
  // states are needed at output cells only: halo lanes and rows load the
  // nearest output cell of the tile instead (a line the tile loads
  // anyway), so the states cost no halo traffic.
  M fv[NF][R];
  VAcc acc;
  uint32_t colf = col;
  int64_t rofff[PROD ? R : 1];
  if constexpr (PROD) {
    const int64_t c_hi = (o2 + 63 - K < a.n2 - 1) ? o2 + 63 - K : a.n2 - 1;
    colf = (uint32_t)(cl < o2 ? o2 : (cl > c_hi ? c_hi : cl));
    const int64_t r_hi = (o1 + WY * R - K - 1 < a.n1 - 1) ? o1 + WY * R - K - 1 : a.n1 - 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t rl = rb + r;
      rofff[r] = (rl < o1 ? o1 : (rl > r_hi ? r_hi : rl)) * a.n2;
    }
#pragma unroll
    for (int l = 0; l < NF; ++l) {
      const TF* fb = (const TF*)adj.fs[l] + pw * plane;
#pragma unroll
      for (int r = 0; r < R; ++r) fv[l][r] = (M)St<TF>::ld((fb + rofff[r])[colf]);
    }
  }
  // One plane, all K levels; each level in place: row r from level l to
  // l+1 with axis-1 neighbour n1 (two planes per trip with kTwoPlanes).
  int lvl = 0;
  auto plane_step = [&](const int64_t m) __attribute__((always_inline)) {
    int64_t pn = pw + step;
    pn = pn < 0 ? a.n0 - 1 : (pn == a.n0 ? 0 : pn);
    if constexpr (SR) {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const uint64_t pair = (uint64_t)(pw * plane + roff[2 * j]) + col;
        sword[j] = sr_fold(sr_pair_word(pair));  // once per plane, for all K levels
      }
    }
    if (!kCondPrefetch || m + 1 < cnt) {
      // The next plane.  Loaded unconditionally (after the last plane the
      // current one again, from L2, unused) where registers allow: a
      // prefetch under a branch makes the compiler merge nx across the
      // branch right after the loads, which waits for them there.  FP64 at
      // K > 6 keeps the branch (unconditional loads there spill).
      const TS* base = in + (m + 1 < cnt ? pn : pw) * plane;
#pragma unroll
      for (int r = 0; r < R; ++r) nx[r] = St<TS>::ld(buf_ld(row_rsrc(base + roff[r], rbytes), cb, in));
    }
    double ps[3] = {0.0, 0.0, 0.0};  // PROD: this plane's products (LASTP: last level)
#pragma unroll
    for (int l = 0; l < K; ++l, ++lvl) {
      // Level l+1 from level l in three phases over all rows -- the three
      // differences, the three products, then g = x - p0 - p1 - p2 -- so
      // independent rows interleave (each cell still sees the same
      // operations in the same order).  The row that needs the
      // neighbouring wave's edge row (row 0 with LO1, else row R-1) gets
      // its axis-1 difference after the barrier.
      const int buf = lvl & 1;
      constexpr int er = LO1 ? 0 : R - 1;  // the block-edge row of this wave
      // The neighbouring wave whose edge row this wave's row er reads.  The
      // block's own upwind-edge wave reads its own slot instead: its row er
      // is a halo row, where any finite or non-finite value is harmless (no
      // output cell depends on it), so the read needs no branch.
      const int wn = LO1 ? (w > 0 ? w - 1 : w) : (w + 1 < WY ? w + 1 : w);
      if constexpr (WY > 1 && !PIPE) edge[buf][w][lane] = x[LO1 ? R - 1 : 0];
      if constexpr (PIPE) {
        // Edge exchange off the critical path.  The row this wave hands to
        // its neighbour (wr) is never the row that needs the neighbour's
        // edge (er), so its level-l+1 value is computed first and written
        // to the other parity's slot at once; the neighbour's level-l edge,
        // written during the previous level (or with the plane's loads),
        // was fenced by the previous barrier and is read at the start of
        // the level; row er is finished last and the level ends with the
        // barrier.  A slot is rewritten two levels after it was written,
        // after a barrier that every reader of it had passed.  The block's
        // upwind-edge wave reads its own slot: its row er is a halo row,
        // where any value is harmless.  Per cell the operations and their
        // order are those of a one-step launch.
        constexpr int wr = LO1 ? R - 1 : 0;
        M n1 = x[er];  // the block's upwind edge: halo row, any value
        if constexpr (WY > 1) n1 = edge[buf][wn][lane];
        M g[R], p2e = M(0);
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = LO1 ? R - 1 - q : q;  // wr first, er last
          // The edge lane (63, or 0 with LO2) is a halo column: any value.
          const M n2 = LO2 ? dpp_shift_any<0x138>(x[r]) : dpp_shift_any<0x130>(x[r]);
          const M d0 = prev[l][r] - x[r];
          const M d2 = n2 - x[r];
          if constexpr (LASTP) {
            if (l + 1 == K) {  // the products of the pass's last step, at output cells
              const M v1 = r == er ? n1 : x[LO1 ? (r > 0 ? r - 1 : 0) : (r + 1 < R ? r + 1 : R - 1)];
              const double fd = (double)fv[0][r];
              const bool ok = m >= K && col_out && row_out[r];
              ps[0] += ok ? fd * d0 : 0.0;
              ps[1] += ok ? fd * (v1 - x[r]) : 0.0;
              ps[2] += ok ? fd * d2 : 0.0;
            }
          }
          const M p0 = cf.c[0] * d0;
          const M p2 = cf.c[2] * d2;
          g[r] = x[r] - p0;
          if (r != er) {
            const M d1 = x[LO1 ? r - 1 : r + 1] - x[r];
            const M p1 = cf.c[1] * d1;
            g[r] = g[r] - p1;
            g[r] = g[r] - p2;
          } else {
            p2e = p2;
          }
          if (WY > 1 && l + 1 < K && r == wr) {
            edge[buf ^ 1][w][lane] = St<TS>::rnd(g[wr]);  // level l+1 of row wr
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        const M p1e = cf.c[1] * (n1 - x[er]);
        g[er] = g[er] - p1e;
        g[er] = g[er] - p2e;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          prev[l][r] = x[r];
          x[r] = St<TS>::rnd(g[r]);  // each level as stored
        }
        if constexpr (LASTP) {
          if (l + 1 == K) {  // f_0 at the next plane's output cells
            const TF* fb = (const TF*)adj.fs[0] + pn * plane;
#pragma unroll
            for (int r = 0; r < R; ++r) fv[0][r] = (M)St<TF>::ld((fb + rofff[r])[colf]);
          }
        }
        if constexpr (WY > 1) {
          // The next plane's row wr (its level 0) as late as possible, so
          // the plane's prefetch has the whole plane to arrive.
          if (l + 1 == K) edge[buf ^ 1][w][lane] = nx[wr];
          lds_barrier();
        }
      } else {
      M d0[R], d1[R], d2[R];
      M nb2[PROD && sizeof(M) < 8 ? R : 1];  // axis-2 neighbours, for FP32 products
#pragma unroll
      for (int r = 0; r < R; ++r) {
        // The edge lane (63, or 0 with LO2) is a halo column: any value.
        const M n2 = LO2 ? dpp_shift_any<0x138>(x[r]) : dpp_shift_any<0x130>(x[r]);
        if constexpr (PROD && sizeof(M) < 8) nb2[r] = n2;
        d0[r] = prev[l][r] - x[r];
        d2[r] = n2 - x[r];
        if (r != er) d1[r] = x[LO1 ? r - 1 : r + 1] - x[r];
      }
      M n1 = x[er];  // the block's upwind edge: halo row, any value
      if constexpr (WY > 1) {
        lds_barrier();
        // (A conditional read here measured faster than the branch-free
        // one of the PIPE path: FP32 K=8 2480 vs 2040 G cell-updates/s.)
        const M e = edge[buf][LO1 ? (w > 0 ? w - 1 : 0) : (w + 1 < WY ? w + 1 : WY - 1)][lane];
        if (LO1 ? w > 0 : w + 1 < WY) n1 = e;
      }
      d1[er] = n1 - x[er];
      if constexpr (PROD) {
        if (!LASTP || l + 1 == K) {
          const int fl = LASTP ? 0 : l;
          const bool cells = m >= K && col_out;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const double fd = (double)fv[fl][r];
            double q0, q1, q2;
            if constexpr (sizeof(M) == 8) {  // the FP64 differences are d0..d2
              q0 = fd * d0[r];
              q1 = fd * d1[r];
              q2 = fd * d2[r];
            } else {
              const double xd = (double)x[r];
              const M v1 =
                  r == er ? n1 : x[LO1 ? (r > 0 ? r - 1 : 0) : (r + 1 < R ? r + 1 : R - 1)];
              q0 = fd * ((double)prev[l][r] - xd);
              q1 = fd * ((double)v1 - xd);
              q2 = fd * ((double)nb2[r] - xd);
            }
            const bool ok = cells && row_out[r];
            ps[0] += ok ? q0 : 0.0;
            ps[1] += ok ? q1 : 0.0;
            ps[2] += ok ? q2 : 0.0;
          }
          const TF* fb = (const TF*)adj.fs[fl] + pn * plane;
#pragma unroll
          for (int r = 0; r < R; ++r) fv[fl][r] = (M)St<TF>::ld((fb + rofff[r])[colf]);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        d0[r] = cf.c[0] * d0[r];
        d1[r] = cf.c[1] * d1[r];
        d2[r] = cf.c[2] * d2[r];
      }
      uint32_t word = 0;  // SR: the pair word of rows 2j, 2j+1
#pragma unroll
      for (int r = 0; r < R; ++r) {
        M g = x[r];
        g = g - d0[r];
        g = g - d1[r];
        g = g - d2[r];
        const M gone = prev[l][r];  // SR: a bf16 value, so its low 16 bits are zero
        prev[l][r] = x[r];
        if constexpr (SR) {
          // The pair's word (at the pair's first row, for both): the even
          // row takes its low half, the odd row its high half, which
          // v_cvt_sr_bf16_f32 reads from bits 31:16 of its operand.  The
          // conversion writes the high half of a register whose low half is
          // kept, so it is given a dead bf16 value (low half zero).
          if (r % 2 == 0) word = sr_round_pre(sword[r / 2] ^ a.sr_key[l]);
          const uint32_t rbits = (r % 2 == 0) ? word << 16 : word;
          x[r] = __builtin_bit_cast(
              float, __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(s16x2, gone), g,
                                                      rbits, true));
        } else {
          x[r] = St<TS>::rnd(g);  // each level as stored
        }
      }
      }  // PIPE
      if constexpr (ALL) {
        if (l + 1 < K && m >= K && col_out) {  // level l+1 (level K below)
          TS* ob = (TS*)adj.lv[l] + pw * plane;
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (row_out[r]) buf_st_nt(St<TS>::st(x[r]), row_rsrc(ob + roff[r], rbytes), cb);
        }
      }
    }
    if constexpr (PROD) {
#pragma unroll
      for (int j = 0; j < 3; ++j) acc.add(j, ps[j]);
    }
    // The next plane's loads, issued at the top of this plane, are waited
    // for here, before this plane's stores: the memory counter retires in
    // issue order, so a wait placed after the (row-conditional) stores would
    // also wait for their write acknowledgements, in every wave at once.
    // (The store-all and adjoint modes issue memory operations within the
    // level loop, which a full wait here would catch.)
    if constexpr (kWaitBeforeStores) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    if (m >= K && col_out) {
      TS* ob = out + pw * plane;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (row_out[r]) buf_st_nt(St<TS>::st(x[r]), row_rsrc(ob + roff[r], rbytes), cb);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = nx[r];
    pw = pn;
  };
  if constexpr (WY > 1 && PIPE) {
    // Level 0 of the first plane's edge row, for the first level's read.
    edge[0][w][lane] = x[LO1 ? R - 1 : 0];
    lds_barrier();
  }
  int64_t m = 0;
  if constexpr (kTwoPlanes) {
    for (; m + 1 < cnt; m += 2) {
      plane_step(m);
      plane_step(m + 1);
    }
    if (m < cnt) plane_step(m);
  } else {  // two planes per trip would spill (FP64 9 rows at K = 7, 8;
            // the adjoint) or lose a wave per SIMD (FP32, BF16 at K = 7, 8;
            // FP64 5 rows at K = 8)
    for (; m < cnt; ++m) plane_step(m);
  }
  if constexpr (PROD) block_reduce_vacc<64 * WY>(acc, adj.partials);
}

// ------------------------------------------------------------ dispatch ---

template <typename TS>
constexpr int max_vec() { return (int)(16 / sizeof(TS)); }

// Launch shape of the rows kernel.  Defaults come from sweeps at 512^3 on
// MI355X (see RowsCfg); FDB_ROWS_CFG="R,WY,CHUNK,NT[,GRID]" overrides
// every storage type, FDB_ROWS_CFG_<bytes> one of them (_8 FP64, _4 FP32,
// _2 BF16) and FDB_ADJ_CFG_<bytes> the adjoint.  CHUNK 0 = automatic,
// GRID 0 = one block per tile.
struct RowsCfg {
  // FP64 forward: R=8 rows x 2 waves, ~1024 equal tiles (64-plane chunks
  // at 512^3, one tile per block) and streaming stores run at the
  // plain-copy bandwidth of the device; fewer rows per thread or more
  // waves per block lose 2-8 %, R=16 loses 25 % (one wave per SIMD), and
  // grids that leave some blocks a second tile lose up to a third.
  int R = 8, WY = 2, chunk = 0, nt = 1, grid = 0;
};

// Launch shape of the temporally blocked kernel (tb_cfg).
struct TbCfg {
  int R, WY, blocks;
};

constexpr int64_t kTargetBlocks = 1024;

}  // namespace

// The device's compute units (queried once).
int device_cus() {
  static int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      n = 256;
    return n > 0 ? n : 256;
  }();
  return cus;
}

int64_t kernel_capacity(const void* kern, int threads, size_t dyn_lds) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, dyn_lds) !=
      hipSuccess) {
    (void)hipGetLastError();  // a failed query only loses the estimate
    per_cu = 1;
  }
  return (int64_t)std::max(per_cu, 1) * device_cus();
}

// Number of chunks of the marched axis, between lo and hi, for `tiles`
// tiles per chunk: every block is one tile, the device holds `capacity`
// blocks at once, and a chunk of c planes costs c + extra plane visits
// (planes read before the first output plane).  Picks the count that
// wastes least on the last, partly filled round of blocks and on the extra
// planes.  With fill = false a single, partly filled round counts as full
// (a kernel that saturates HBM with fewer resident blocks); with
// fill = true it counts by its fill (a kernel whose rate scales with
// resident waves).
int64_t march_chunks(int64_t tiles, int64_t n0, int extra, int64_t capacity, int64_t lo,
                     int64_t hi, bool fill) {
  int64_t best = 1;
  double best_eff = -1.0;
  for (int64_t nch = std::max<int64_t>(1, lo); nch <= hi && nch <= n0; ++nch) {
    const int64_t chunk = (n0 + nch - 1) / nch;
    const int64_t real = (n0 + chunk - 1) / chunk;
    const int64_t blocks = tiles * real;
    const int64_t rounds = (blocks + capacity - 1) / capacity;
    const double used =
        (!fill && rounds == 1) ? 1.0 : (double)blocks / (double)(rounds * capacity);
    const double eff = used * (double)n0 / (double)(real * (chunk + extra));
    if (eff > best_eff + 1e-9) {
      best_eff = eff;
      best = real;
    }
  }
  return best;
}

namespace {

// Blocks of `kern` (with `threads` threads, static LDS only) the device
// holds at once.
template <typename Kern>
int64_t resident_blocks(Kern kern, int threads) {
  return kernel_capacity(reinterpret_cast<const void*>(kern), threads, 0);
}

// (R, WY) pairs the rows kernel is instantiated for (FDB_ROWS_V below).
constexpr int kRowsShapes[][2] = {{8, 2}, {4, 2}, {4, 4}, {8, 1}, {2, 4}, {2, 8}, {2, 2}, {1, 8}};
// (R, K, WY) of the temporally blocked forward (FDB_TB below).
constexpr int kTbShapes[][3] = {{9, 2, 4}, {9, 3, 4}, {9, 4, 4}, {9, 6, 8},
                                {9, 7, 8}, {9, 8, 8}, {5, 8, 8}};
// FP64-only shapes: two planes per trip fit in 256 VGPRs.
constexpr int kTbShapes64[][3] = {{8, 6, 8}, {8, 7, 8}, {7, 8, 8}};

// Default launch shape of the rows kernel.  bytes: storage of the stepped
// field (lam for the adjoint); fbytes: the adjoint's f_t storage (0: none).
RowsCfg rows_default(int bytes, bool adj, int fbytes) {
  RowsCfg c;
  // 512^3 sweeps: FP32 R=4 x 2 waves (715 G cells/s, 2.0x FP64), BF16
  // R=2 x 4 waves.  The adjoint carries f_t as well (1024x1024x768
  // sweeps): R=8 x 1 wave (FP32 lam and f_t, 5.3 TB/s), R=8 x 2 (FP64,
  // 5.3 TB/s), R=4 x 4 for an FP32 lam with a bf16 f_t (4.9 TB/s; R=8
  // spills there, and every spilled f_t load waits for its data at once).
  if (adj) {
    c.R = 8;
    c.WY = 1;
    if (bytes == 8 && fbytes == 8) c.WY = 2;
    if (bytes == 4 && fbytes == 2) {
      c.R = 4;
      c.WY = 4;
    }
    if (bytes == 2) {  // bf16 lam: 8 cells per lane, 1 row x 8 waves (512^3 sweep)
      c.R = 1;
      c.WY = 8;
    }
  } else if (bytes == 4) {
    c.R = 4;
  } else if (bytes == 2) {
    c.R = 2;
    c.WY = 4;
  }
  return c;
}

// FP64 at K = 7: 8 rows x 8 waves, two planes per plane-loop trip (512^3:
// 1516 vs 1431 G cell-updates/s for 9 rows, one plane per trip; at K = 8
// the 7-row two-plane shape loses, 1395 vs 1480 G).
TbCfg tb_default(int bytes, int K) {
  if (bytes == 8 && K == 7) return TbCfg{8, 8, 0};
  return K >= 6 ? TbCfg{9, 8, 0} : TbCfg{9, 4, 0};
}

// Launch shapes, read once per process from the environment (a
// function-local static, so concurrent first calls are safe).  Tuning
// variables:
//   FDB_ROWS_CFG / FDB_ROWS_CFG_<bytes>  = "R,WY,CHUNK,NT[,GRID]"  (forward)
//   FDB_ADJ_CFG_<bytes> / _<bytes>_<fbytes> = same                   (adjoint)
//   FDB_TB_CFG_<bytes>_<K>               = "R,WY,BLOCKS"           (fused)
// A variable that does not parse or names a shape without an instantiated
// kernel makes `error` non-empty; the entry points then fail with
// FDB_ERR_INVALID instead of launching (tuning_error()).
struct Tuning {
  RowsCfg rows[2][9][9];  // [adj][bytes][fbytes]
  TbCfg tb[9][kMaxTimeBlock + 1];  // [bytes][K]
  bool tb_env[9][kMaxTimeBlock + 1] = {};  // set by FDB_TB_CFG_<bytes>_<K>
  std::string error;
};

bool rows_shape_ok(const RowsCfg& c) {
  if (c.chunk < 0 || c.grid < 0 || c.nt != 1) return false;  // non-temporal stores only
  for (const auto& s : kRowsShapes)
    if (s[0] == c.R && s[1] == c.WY) return true;
  return false;
}

bool tb_shape_ok(const TbCfg& c, int K, int bytes) {
  if (c.blocks < 0) return false;
  for (const auto& s : kTbShapes)
    if (s[0] == c.R && s[1] == K && s[2] == c.WY) return true;
  if (bytes == 8)
    for (const auto& s : kTbShapes64)
      if (s[0] == c.R && s[1] == K && s[2] == c.WY) return true;
  return false;
}

// Parses exactly a comma-separated list of 3 to 5 decimal integers;
// returns how many, 0 if the text is anything else.
int parse_ints(const char* s, int* v, int lo, int hi) {
  int n = 0;
  while (n < hi) {
    char* end = nullptr;
    const long x = strtol(s, &end, 10);
    if (end == s || x < INT32_MIN || x > INT32_MAX) return 0;
    v[n++] = (int)x;
    if (*end == '\0') return n >= lo ? n : 0;
    if (*end != ',') return 0;
    s = end + 1;
  }
  return 0;
}

Tuning load_tuning() {
  Tuning t;
  auto bad = [&](const char* name, const char* val) {
    if (t.error.empty())
      t.error = std::string("tuning variable ") + name + "=\"" + val +
                "\" does not name an instantiated launch shape";
  };
  auto rows_env = [&](const char* name, RowsCfg& c) {
    const char* s = getenv(name);
    if (!s) return false;
    int v[5] = {0, 0, 0, 0, 0};
    const int got = parse_ints(s, v, 4, 5);
    RowsCfg n = c;
    n.R = v[0];
    n.WY = v[1];
    n.chunk = v[2];
    n.nt = v[3];
    n.grid = v[4];
    if (got == 0 || !rows_shape_ok(n)) {
      bad(name, s);
      return false;
    }
    c = n;
    return true;
  };
  const int sizes[3] = {8, 4, 2};
  char name[48];
  for (int b : sizes) {
    RowsCfg f = rows_default(b, false, 0);
    snprintf(name, sizeof(name), "FDB_ROWS_CFG_%d", b);
    if (!rows_env(name, f)) rows_env("FDB_ROWS_CFG", f);
    for (int fb = 0; fb < 9; ++fb) t.rows[0][b][fb] = f;
    for (int fb : {0, 8, 4, 2}) {
      RowsCfg a = rows_default(b, true, fb);
      snprintf(name, sizeof(name), "FDB_ADJ_CFG_%d_%d", b, fb);
      if (!rows_env(name, a)) {
        snprintf(name, sizeof(name), "FDB_ADJ_CFG_%d", b);
        rows_env(name, a);
      }
      t.rows[1][b][fb] = a;
    }
    for (int K = 2; K <= kMaxTimeBlock; ++K) {
      TbCfg c = tb_default(b, K);
      snprintf(name, sizeof(name), "FDB_TB_CFG_%d_%d", b, K);
      if (const char* s = getenv(name)) {
        int v[3] = {0, 0, 0};
        if (parse_ints(s, v, 3, 3) == 3 && tb_shape_ok(TbCfg{v[0], v[1], v[2]}, K, b))
          c = TbCfg{v[0], v[1], v[2]};
        else
          bad(name, s);
        t.tb_env[b][K] = true;
      }
      t.tb[b][K] = c;
    }
  }
  return t;
}

const Tuning& tuning() {
  static const Tuning t = load_tuning();
  return t;
}

RowsCfg rows_cfg(int bytes, bool adj, int fbytes = 0) {
  return tuning().rows[adj ? 1 : 0][bytes][adj ? fbytes : 0];
}

// Launches the rows kernel for one step (forward: dirs = u, adjoint:
// dirs = -u).  With dry = true only reports the block count.
template <typename TS, typename TF, bool ADJ, bool SR = false>
hipError_t rows_dispatch(const Grid3& g, const TS* in, TS* out, const int dirs[3],
                         const TF* ft, double* partials, int64_t* nblocks_out, hipStream_t s,
                         bool dry, uint32_t sr_key = 0) {
  using M = typename St<TS>::M;
  const RowsCfg cfg = rows_cfg((int)sizeof(TS), ADJ, ft ? (int)sizeof(TF) : 0);
  MarchArgs a;
  a.sr_key = sr_key;
  a.n0 = g.n[0];
  a.n1 = g.n[1];
  a.n2 = g.n[2];
  a.dir0 = dirs[0];
  a.dir1 = dirs[1];
  a.dir2 = dirs[2];
  a.act0 = g.act[0];
  a.act1 = g.act[1];
  a.act2 = g.act[2];
  Coef<M> cf;
  for (int j = 0; j < 3; ++j) cf.c[j] = (M)g.c[j];
  // Widest access that keeps every row start aligned (f_t included, and
  // no wider than 16 bytes of f_t) ...
  int vec = std::min(max_vec<TS>(), max_vec<TF>());
  uintptr_t align_bits = (uintptr_t)in | (uintptr_t)out;
  while (vec > 1 && ((a.n2 % vec) != 0 || (align_bits % (vec * sizeof(TS))) != 0 ||
                     ((uintptr_t)ft % (vec * sizeof(TF))) != 0))
    vec >>= 1;
  // ... narrowed while that leaves fewer idle lanes in a row's last tile
  // (a 768-cell bf16 row takes two 512-cell tiles at 8 cells per lane, a
  // quarter of the lanes idle, but three exact tiles at 4).
  auto idle = [&](int v) {
    const int64_t w = 64 * (int64_t)v;
    return ((a.n2 + w - 1) / w) * w - a.n2;
  };
  while (vec > 1 && idle(vec / 2) < idle(vec)) vec >>= 1;
  const int R = cfg.R, WY = cfg.WY;
  a.tiles2 = (a.n2 + 64 * vec - 1) / (64 * vec);
  a.tiles1 = (a.n1 + (int64_t)WY * R - 1) / ((int64_t)WY * R);
  const int64_t tiles = a.tiles2 * a.tiles1;
  // Axis-0 chunk: an explicit one, or equal chunks for about kTargetBlocks
  // tiles (4 per CU) and up to 4x more, whichever count fills the last
  // round of resident blocks best (occupancy of the kernel launched); each
  // block takes one tile.
  hipError_t result = hipErrorInvalidValue;
  auto go = [&](auto kern, int W) {
    int64_t chunk = cfg.chunk;
    if (chunk <= 0) {
      const int64_t lo = std::max<int64_t>(1, (kTargetBlocks + tiles / 2) / tiles);
      const int64_t nch =
          march_chunks(tiles, a.n0, 1, resident_blocks(kern, 64 * W), lo, 4 * lo, false);
      chunk = (a.n0 + nch - 1) / nch;
    }
    a.chunk = chunk;
    const int64_t ntiles = tiles * ((a.n0 + chunk - 1) / chunk);
    const int64_t nb = (cfg.grid > 0 && cfg.grid < ntiles) ? cfg.grid : ntiles;
    if (nblocks_out) *nblocks_out = nb;
    if (dry) {
      result = hipSuccess;
      return;
    }
    if (nb <= 0 || nb > 0x7fffffffLL) {
      result = hipErrorInvalidConfiguration;
      return;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64, W), 0, s, in, out, a, cf, ntiles, ft,
                       partials);
    result = hipGetLastError();
  };
  bool launched = false;
#define FDB_ROWS(V, RR, W, NTF)                                                 \
  if constexpr (V <= max_vec<TS>() && V <= max_vec<TF>()) {                     \
    if (!launched && vec == V && R == RR && WY == W && (cfg.nt != 0) == NTF) { \
      go(rows_kernel<TS, TF, V, RR, W, NTF, ADJ, SR>, W);                       \
      launched = true;                                                          \
    }                                                                           \
  }
#define FDB_ROWS_V(RR, W, NTF) \
  FDB_ROWS(8, RR, W, NTF) FDB_ROWS(4, RR, W, NTF) FDB_ROWS(2, RR, W, NTF) FDB_ROWS(1, RR, W, NTF)
  FDB_ROWS_V(8, 2, true)
  FDB_ROWS_V(4, 2, true)
  FDB_ROWS_V(4, 4, true)
  FDB_ROWS_V(8, 1, true)
  FDB_ROWS_V(2, 4, true)
  FDB_ROWS_V(2, 8, true)
  FDB_ROWS_V(2, 2, true)
  FDB_ROWS_V(1, 8, true)
#undef FDB_ROWS_V
#undef FDB_ROWS
  return launched ? result : hipErrorInvalidValue;
}

// Temporally blocked pass: K steps in one launch.  Returns
// hipErrorNotSupported when the shape does not fit (the caller then runs
// one-step launches): all three axes present (an absent axis has no term
// at all, which a fused pass would not reproduce for Inf/NaN values).
// Launch shape (rows per wave, waves per block, target tiles per launch;
// 0 = chunk count from the kernel's occupancy, below): sweeps at 512^3
// and 1024^3 on MI355X, all bit-exact, gave 9 rows x 8 waves at K >= 6
// (72 rows per block, 64 output rows at K=8) for FP64 and FP32 and 9 rows
// x 4 waves at K <= 4 (tb_default; FDB_TB_CFG_<bytes>_<K> overrides).
TbCfg tb_cfg(int bytes, int K) { return tuning().tb[bytes][K]; }

// levels != nullptr: also store levels 1..K-1 to levels[0..K-2] (K = 4
// or 8, FP64 only: with one 4-byte cell per lane the FP32 kernel writes its
// 8 levels at ~2 TB/s, slower than one-step launches storing 16 B per lane).
// Launch shape of the stochastic-rounding bf16 pass (K = 2, 4 or 8): an
// even number of rows per wave and an even K, so the wave's rows start at
// an even row and hold whole pairs; within 128 VGPRs, i.e. four waves per
// SIMD like the round-to-nearest pass.
TbCfg tb_sr_cfg(int K) { return K == 8 ? TbCfg{8, 8, 0} : TbCfg{8, 4, 0}; }

// The fused SR pass needs an even n1 (wrapped rows keep their parity).
bool tb_sr_ok(const Grid3& g, int K) {
  return (K == 2 || K == 4 || K == 8) && g.n[1] % 2 == 0;
}

// FP64 at K = 6: 9 rows x 8 waves (at 512^3 72 tiles, two rounds of 504
// blocks) or 8 rows (81 tiles, one round of 243), whichever marches fewer
// row-planes per resident block slot on this grid: the march's rounds x
// (chunk + K) planes, times the rows per wave.  Measured: 512^3 picks 8
// rows, 549 against 570 us per launch; 1024^3 9 rows, 4751 against 4844.
int fp64_k6_rows(const Grid3& g) {
  constexpr int K = 6, WY = 8;
  auto cost = [&](int R, int64_t cap) {
    const int64_t tiles =
        ((g.n[2] + (64 - K) - 1) / (64 - K)) * ((g.n[1] + (WY * R - K) - 1) / (WY * R - K));
    const int64_t nch = march_chunks(tiles, g.n[0], K, cap, 1, 64, true);
    const int64_t chunk = (g.n[0] + nch - 1) / nch;
    const int64_t blocks = tiles * ((g.n[0] + chunk - 1) / chunk);
    return (blocks + cap - 1) / cap * (chunk + K) * R;
  };
  const int64_t c9 = cost(9, resident_blocks(tb_kernel<double, 9, K, WY, false, false>, 64 * WY));
  const int64_t c8 = cost(8, resident_blocks(tb_kernel<double, 8, K, WY, false, false>, 64 * WY));
  return c8 < c9 ? 8 : 9;
}

// visits != nullptr: launch nothing; store the pass's length in row-plane
// visits of one resident block slot instead (rounds of resident blocks x
// planes per block, warm-up planes included, x rows per wave), from the same
// shape, chunking and occupancy the launch would use.
template <typename TS>
hipError_t tb_dispatch(const Grid3& g, const TS* in, TS* out, int K, hipStream_t s,
                       void* const* levels = nullptr, const uint32_t* sr_keys = nullptr,
                       int64_t* visits = nullptr) {
  using M = typename St<TS>::M;
  const bool sr = sr_keys != nullptr;
  if (sr && (sizeof(TS) != 2 || !tb_sr_ok(g, K))) return hipErrorNotSupported;
  TbCfg cfg = sr ? tb_sr_cfg(K) : tb_cfg((int)sizeof(TS), K);
  if constexpr (sizeof(TS) == 8) {
    if (K == 6 && !sr && levels == nullptr && !tuning().tb_env[8][6]) cfg.R = fp64_k6_rows(g);
  }
  const int R = cfg.R, WY = cfg.WY;
  if (!(g.act[0] && g.act[1] && g.act[2]) || WY * R <= K) return hipErrorNotSupported;
  TbArgs a;
  a.n0 = g.n[0];
  a.n1 = g.n[1];
  a.n2 = g.n[2];
  a.u0 = g.up[0];
  // Rows are addressed by 32-bit byte offsets; longer rows take one-step
  // launches.
  if (a.n2 * (int64_t)sizeof(TS) > INT32_MAX) return hipErrorNotSupported;
  a.tiles2 = (a.n2 + (64 - K) - 1) / (64 - K);
  a.tiles1 = (a.n1 + (WY * R - K) - 1) / (WY * R - K);
  const int64_t tiles = a.tiles2 * a.tiles1;
  const bool lo1 = g.up[1] < 0, lo2 = g.up[2] < 0;
  Coef<M> cf;
  for (int j = 0; j < 3; ++j) cf.c[j] = (M)g.c[j];
  TbAux aux{};
  const bool all = levels != nullptr;
  if (all && sizeof(TS) != 8) return hipErrorNotSupported;
  for (int l = 0; all && l + 1 < K; ++l) aux.lv[l] = levels[l];
  for (int l = 0; sr && l < K; ++l) a.sr_key[l] = sr_fold(sr_keys[l]);
  const int mode = all ? kTbStoreAll : (sr ? kTbSR : kTbPlain);
  bool launched = false;
#define FDB_TB_DIR(RR, KK, W, L1, L2, MD)                                                   \
  if (!launched && R == RR && K == KK && WY == W && lo1 == L1 && lo2 == L2 &&               \
      mode == MD) {                                                                         \
    auto kern = tb_kernel<TS, RR, KK, W, L1, L2, MD>;                                       \
    const int64_t cap = resident_blocks(kern, 64 * W);                                      \
    int64_t nch;                                                                            \
    if (cfg.blocks > 0) {                                                                   \
      nch = std::max<int64_t>(1, (cfg.blocks + tiles / 2) / tiles);                         \
    } else {                                                                                \
      nch = march_chunks(tiles, a.n0, KK, cap, 1, 64, true);                                \
    }                                                                                       \
    a.chunk = (a.n0 + nch - 1) / nch;                                                       \
    a.ntiles = tiles * ((a.n0 + a.chunk - 1) / a.chunk);                                    \
    if (visits) {                                                                           \
      *visits = (a.ntiles + cap - 1) / cap * (a.chunk + KK) * RR;                           \
    } else {                                                                                \
      if (a.ntiles > 0x7fffffffLL) return hipErrorInvalidConfiguration;                     \
      hipLaunchKernelGGL(kern, dim3((unsigned)a.ntiles), dim3(64, W), 0, s, in, out, a, cf, \
                         aux);                                                              \
    }                                                                                       \
    launched = true;                                                                        \
  }
#define FDB_TB(RR, KK, W, MD)                                                          \
  FDB_TB_DIR(RR, KK, W, false, false, MD) FDB_TB_DIR(RR, KK, W, false, true, MD)       \
  FDB_TB_DIR(RR, KK, W, true, false, MD) FDB_TB_DIR(RR, KK, W, true, true, MD)
  FDB_TB(9, 2, 4, kTbPlain)
  FDB_TB(9, 3, 4, kTbPlain)
  FDB_TB(9, 4, 4, kTbPlain)
  FDB_TB(9, 6, 8, kTbPlain)
  FDB_TB(9, 7, 8, kTbPlain)
  FDB_TB(9, 8, 8, kTbPlain)
  FDB_TB(5, 8, 8, kTbPlain)
  if constexpr (sizeof(TS) == 8) {
    FDB_TB(8, 6, 8, kTbPlain)
    FDB_TB(8, 7, 8, kTbPlain)
    FDB_TB(7, 8, 8, kTbPlain)
    FDB_TB(9, 4, 4, kTbStoreAll)
    FDB_TB(9, 8, 8, kTbStoreAll)
  }
  if constexpr (sizeof(TS) == 2) {
    FDB_TB(8, 2, 4, kTbSR)
    FDB_TB(8, 4, 4, kTbSR)
    FDB_TB(8, 8, 8, kTbSR)
  }
#undef FDB_TB
#undef FDB_TB_DIR
  if (!launched) return hipErrorNotSupported;
  return visits ? hipSuccess : hipGetLastError();
}

// Fused transpose pass (tb_kernel with PROD): K adjoint steps with the
// velocity products, K = 4, 8 waves of 4 rows (FP64) or 8 rows (FP32):
// prev, the K states and the level's differences then stay within 256
// VGPRs without spills (FP64 spills from 5 rows on).  LAST: the products
// of the last step only, against fs[0] (kTbAdjLast): K = 7 with 8 rows
// (FP64, 234 VGPRs, one block per CU like the plain 7-step pass) or K = 8
// with 6 rows (FP32, at most 115 VGPRs, so two blocks per CU).  *nb_out
// gets the block count (= partial slots written); with dry = true nothing
// is launched.  hipErrorNotSupported when the shape has no fused form or
// needs more than max_blocks partial slots.
template <typename TS, typename TF = TS, bool LAST = false>
hipError_t tb_adj_dispatch(const Grid3& g, const TS* in, TS* out, const void* const* fs, int K,
                           double* partials, int64_t max_blocks, int64_t* nb_out, bool dry,
                           hipStream_t s) {
  using M = typename St<TS>::M;
  if (sizeof(TS) < 4 || K > kMaxTimeBlock) return hipErrorNotSupported;
  constexpr int R = LAST ? (sizeof(TS) == 8 ? 8 : 6) : (sizeof(TS) == 8 ? 4 : 8), WY = 8;
  constexpr int KL = LAST ? (sizeof(TS) == 8 ? 7 : 8) : 4;  // the instantiated depth
  constexpr int MD = LAST ? kTbAdjLast : kTbAdjoint;
  if (!(g.act[0] && g.act[1] && g.act[2]) || WY * R <= K) return hipErrorNotSupported;
  TbArgs a;
  a.n0 = g.n[0];
  a.n1 = g.n[1];
  a.n2 = g.n[2];
  a.u0 = -g.up[0];  // the transpose gathers from the downwind side
  if (a.n2 * (int64_t)sizeof(TS) > INT32_MAX) return hipErrorNotSupported;  // row offsets
  a.tiles2 = (a.n2 + (64 - K) - 1) / (64 - K);
  a.tiles1 = (a.n1 + (WY * R - K) - 1) / (WY * R - K);
  const int64_t tiles = a.tiles2 * a.tiles1;
  const bool lo1 = -g.up[1] < 0, lo2 = -g.up[2] < 0;
  Coef<M> cf;
  for (int j = 0; j < 3; ++j) cf.c[j] = (M)g.c[j];
  TbAux adj{};
  for (int l = 0; l < (LAST ? 1 : K); ++l) adj.fs[l] = fs ? fs[l] : nullptr;
  adj.partials = partials;
  bool launched = false;
#define FDB_TBA_DIR(KK, L1, L2)                                                              \
  if (!launched && K == KK && lo1 == L1 && lo2 == L2) {                                      \
    auto kern = tb_kernel<TS, R, KK, WY, L1, L2, MD, TF>;                                    \
    const int64_t nch =                                                                      \
        march_chunks(tiles, a.n0, KK, resident_blocks(kern, 64 * WY), 1, 64, true);          \
    a.chunk = (a.n0 + nch - 1) / nch;                                                        \
    a.ntiles = tiles * ((a.n0 + a.chunk - 1) / a.chunk);                                     \
    if (nb_out) *nb_out = a.ntiles;                                                          \
    if (dry) return hipSuccess;                                                              \
    if (a.ntiles > max_blocks || a.ntiles > 0x7fffffffLL) return hipErrorNotSupported;       \
    hipLaunchKernelGGL(kern, dim3((unsigned)a.ntiles), dim3(64, WY), 0, s, in, out, a, cf,   \
                       adj);                                                                 \
    launched = true;                                                                         \
  }
#define FDB_TBA(KK)                                                  \
  FDB_TBA_DIR(KK, false, false) FDB_TBA_DIR(KK, false, true)         \
  FDB_TBA_DIR(KK, true, false) FDB_TBA_DIR(KK, true, true)
  FDB_TBA(KL)
#undef FDB_TBA
#undef FDB_TBA_DIR
  if (!launched) return hipErrorNotSupported;
  return hipGetLastError();
}

template <typename TS, bool SR = false>
hipError_t step_typed(const Grid3& g, const void* in, void* out, hipStream_t s,
                      uint32_t sr_key = 0) {
  return rows_dispatch<TS, TS, false, SR>(g, (const TS*)in, (TS*)out, g.up, (const TS*)nullptr,
                                          nullptr, nullptr, s, false, sr_key);
}

// The transpose step: directions -u; with f_t the velocity-gradient
// products too (lam in FP64, FP32 or bf16 storage with FP32 arithmetic;
// f_t in any storage type).
hipError_t adjoint_dispatch(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                            void* lam, const void* f_t, double* partials, int64_t* nb,
                            hipStream_t s, bool dry) {
  const int dirs[3] = {-g.up[0], -g.up[1], -g.up[2]};
  f_dtype = f_t == nullptr ? lam_dtype : storage_dtype(f_dtype);
#define FDB_ADJ(TL, TF)                                                                    \
  return rows_dispatch<TL, TF, true>(g, (const TL*)lam_next, (TL*)lam, dirs, (const TF*)f_t, \
                                     partials, nb, s, dry)
  if (lam_dtype == FDB_F64 && f_dtype == FDB_F64) FDB_ADJ(double, double);
  if (lam_dtype == FDB_F64 && f_dtype == FDB_F32) FDB_ADJ(double, float);
  if (lam_dtype == FDB_F64 && f_dtype == FDB_BF16) FDB_ADJ(double, uint16_t);
  if (lam_dtype == FDB_F32 && f_dtype == FDB_F64) FDB_ADJ(float, double);
  if (lam_dtype == FDB_F32 && f_dtype == FDB_F32) FDB_ADJ(float, float);
  if (lam_dtype == FDB_F32 && f_dtype == FDB_BF16) FDB_ADJ(float, uint16_t);
  if (lam_dtype == FDB_BF16 && f_dtype == FDB_F64) FDB_ADJ(uint16_t, double);
  if (lam_dtype == FDB_BF16 && f_dtype == FDB_F32) FDB_ADJ(uint16_t, float);
  if (lam_dtype == FDB_BF16 && f_dtype == FDB_BF16) FDB_ADJ(uint16_t, uint16_t);
#undef FDB_ADJ
  return hipErrorInvalidValue;
}

}  // namespace

// --------------------------------------------------------------- public --

size_t dtype_size(int dtype) {
  switch (dtype) {
    case FDB_F64: return 8;
    case FDB_F32: return 4;
    case FDB_BF16: return 2;
    case FDB_BF16_SR: return 2;
    default: return 0;
  }
}

bool make_grid(int ndims, const uint64_t* n, const double* v, const double* len, double dt,
               Grid3* g) {
  if (ndims < 1 || ndims > 3 || !n || !v || !len) return false;
  const int off = 3 - ndims;
  g->ntot = 1;
  for (int j = 0; j < 3; ++j) {
    g->n[j] = 1;
    g->up[j] = -1;
    g->act[j] = 0;
    g->c[j] = 0.0;
    g->dc[j] = 0.0;
  }
  for (int r = 0; r < ndims; ++r) {
    const int j = r + off;
    if (n[r] == 0 || n[r] > (uint64_t)INT64_MAX) return false;
    g->n[j] = (int64_t)n[r];
    g->act[j] = 1;
    g->up[j] = (v[r] < 0.) ? +1 : -1;               // upwind.cxx:34-35
    const double dx = len[r] / (double)n[r];         // upwind.cxx:36
    g->c[j] = dt * v[r] * (double)g->up[j] / dx;     // upwind.cxx:72
    g->dc[j] = dt * (double)g->up[j] / dx;
    if (g->ntot > INT64_MAX / g->n[j]) return false;
    g->ntot *= g->n[j];
  }
  return true;
}

bool fused_pass_supported(int dtype, const Grid3& g, int k) {
  if (k < 2 || k > kMaxTimeBlock) return false;
  if (dtype != FDB_F64 && dtype != FDB_F32 && dtype != FDB_BF16 && dtype != FDB_BF16_SR)
    return false;
  const int bytes = (int)dtype_size(dtype);
  const bool sr = dtype == FDB_BF16_SR;
  if (sr && !tb_sr_ok(g, k)) return false;
  const TbCfg c = sr ? tb_sr_cfg(k) : tb_cfg(bytes, k);
  return (sr || tb_shape_ok(c, k, bytes)) && g.act[0] && g.act[1] && g.act[2] && c.WY * c.R > k &&
         g.n[2] * (int64_t)bytes <= INT32_MAX;
}

namespace {
// Device time (us) of one launch on MI355X, on cubic grids of kPassSide[i]^3
// cells: index 1 the one-step kernel, 2..8 the fused pass of that many
// steps (0: no such pass).  Rows: FP64, FP32, bf16 round to nearest, bf16
// stochastic rounding.  Below 512^3 a launch's time is bound by the
// latency of its few marched planes rather than by throughput (at 16^3 a
// one-step FP64 launch, 16 planes in sequence, takes longer than a 4-step
// pass), so the times are measured points, not a model.
constexpr int kPassSizes = 4;
constexpr double kPassSide[kPassSizes] = {16, 128, 256, 512};
constexpr float kPassUsAt[kPassSizes][4][kMaxTimeBlock + 1] = {
    {{0, 23.4f, 7.9f, 8.8f, 9.9f, 0, 20.5f, 23.4f, 32.3f},
     {0, 12.1f, 6.8f, 8.8f, 10.2f, 0, 17.9f, 23.8f, 26.5f},
     {0, 11.0f, 9.3f, 8.5f, 9.9f, 0, 20.9f, 25.6f, 28.3f},
     {0, 10.8f, 6.7f, 0, 10.5f, 0, 0, 0, 37.7f}},
    {{0, 11.2f, 11.3f, 14.7f, 20.5f, 0, 30.2f, 35.2f, 44.4f},
     {0, 6.0f, 10.5f, 13.0f, 16.2f, 0, 29.0f, 37.5f, 44.8f},
     {0, 6.1f, 10.8f, 13.2f, 17.6f, 0, 28.0f, 41.1f, 49.5f},
     {0, 5.9f, 13.5f, 0, 24.8f, 0, 0, 0, 76.4f}},
    {{0, 62.8f, 65.0f, 66.9f, 67.9f, 0, 88.3f, 100.4f, 114.5f},
     {0, 30.4f, 39.2f, 41.5f, 43.4f, 0, 57.0f, 70.8f, 80.2f},
     {0, 17.7f, 28.9f, 35.8f, 44.6f, 0, 72.5f, 78.7f, 91.7f},
     {0, 20.7f, 33.9f, 0, 67.7f, 0, 0, 0, 144.0f}},
    {{0, 373, 573, 576, 537, 0, 549, 577, 705},
     {0, 190, 379, 407, 381, 0, 400, 450, 402},
     {0, 99, 228, 233, 241, 0, 394, 356, 453},
     {0, 117, 220, 0, 348, 0, 0, 0, 769}}};
constexpr const float (&kPassUs)[4][kMaxTimeBlock + 1] = kPassUsAt[kPassSizes - 1];
constexpr int kPlanTail = 48;  // steps split exactly; longer runs first take the best depth
int pass_row(int dtype) {
  return dtype == FDB_F64 ? 0 : (dtype == FDB_F32 ? 1 : (dtype == FDB_BF16 ? 2 : 3));
}

// Length of a k-step fused pass over g in plane visits of one resident block
// slot (tb_dispatch's query mode); 0 when there is no such pass.
int64_t fused_pass_visits(int dtype, const Grid3& g, int k) {
  if (!fused_pass_supported(dtype, g, k)) return 0;
  int64_t v = 0;
  hipError_t e = hipErrorNotSupported;
  const uint32_t keys[kMaxTimeBlock] = {};
  switch (dtype) {
    case FDB_F64:
      e = tb_dispatch<double>(g, nullptr, nullptr, k, nullptr, nullptr, nullptr, &v);
      break;
    case FDB_F32:
      e = tb_dispatch<float>(g, nullptr, nullptr, k, nullptr, nullptr, nullptr, &v);
      break;
    case FDB_BF16:
      e = tb_dispatch<uint16_t>(g, nullptr, nullptr, k, nullptr, nullptr, nullptr, &v);
      break;
    case FDB_BF16_SR:
      e = tb_dispatch<uint16_t>(g, nullptr, nullptr, k, nullptr, nullptr, keys, &v);
      break;
    default: break;
  }
  return e == hipSuccess ? v : 0;
}

// Launch times (us) of every depth over g (0: no such pass).  On grids of at
// least 512^3 cells the 512^3 times are scaled to g: the one-step kernel by
// the cell count, a fused pass by its length in plane visits (its blocks take
// equal time per plane there, so the length captures rounds of resident
// blocks, warm-up planes and tiles cut by the grid's extents; at 1024^3 FP64
// this picks 8-step passes, 679 us per step, over 7-step ones, 710).  On
// smaller grids launch times stop scaling that way (a march of few planes is
// bound by the latency of each plane, not by throughput), and each time is
// interpolated between the measured sizes, linearly in log(time) against
// log(cells) (grids below 16^3 take the 16^3 times).
struct PassCosts {
  float us[kMaxTimeBlock + 1];
};

// Plane visits of every pass at 512^3, per row of kPassUs and depth (read
// once per process, on the first grid that needs them).
const int64_t* ref_visits() {
  static const int64_t* ref = [] {
    static int64_t v[4][kMaxTimeBlock + 1] = {};
    const uint64_t n[3] = {512, 512, 512};
    const double vel[3] = {1.0, 1.0, 1.0}, len[3] = {1.0, 1.0, 1.0};
    Grid3 r;
    if (make_grid(3, n, vel, len, 1e-4, &r)) {
      const int dts[4] = {FDB_F64, FDB_F32, FDB_BF16, FDB_BF16_SR};
      for (int d = 0; d < 4; ++d)
        for (int k = 2; k <= kMaxTimeBlock; ++k) v[d][k] = fused_pass_visits(dts[d], r, k);
    }
    return &v[0][0];
  }();
  return ref;
}

PassCosts pass_costs(int dtype, const Grid3& g) {
  const int row = pass_row(dtype);
  const float* tab = kPassUs[row];
  const double cells = (double)g.ntot / (512.0 * 512.0 * 512.0);
  PassCosts c{};
  if (cells < 1.0) {
    // Between the measured sizes i and i + 1 (side^3 cells each).
    const double lc = std::log((double)g.ntot);
    int i = 0;
    while (i + 2 < kPassSizes && lc > 3.0 * std::log(kPassSide[i + 1])) ++i;
    const double l0 = 3.0 * std::log(kPassSide[i]), l1 = 3.0 * std::log(kPassSide[i + 1]);
    const double a = std::min(1.0, std::max(0.0, (lc - l0) / (l1 - l0)));
    for (int k = 1; k <= kMaxTimeBlock; ++k) {
      const float t0 = kPassUsAt[i][row][k], t1 = kPassUsAt[i + 1][row][k];
      if (t0 <= 0.0f || t1 <= 0.0f || (k > 1 && !fused_pass_supported(dtype, g, k))) continue;
      c.us[k] = (float)std::exp((1.0 - a) * std::log((double)t0) + a * std::log((double)t1));
    }
    return c;
  }
  const int64_t* ref = ref_visits() + row * (kMaxTimeBlock + 1);
  c.us[1] = tab[1] * (float)cells;
  for (int k = 2; k <= kMaxTimeBlock; ++k) {
    if (tab[k] <= 0.0f || ref[k] <= 0 || !fused_pass_supported(dtype, g, k)) continue;
    const int64_t v = fused_pass_visits(dtype, g, k);
    if (v > 0) c.us[k] = tab[k] * (float)v / (float)ref[k];
  }
  return c;
}

// pass_costs for the last few (dtype, grid) pairs: a plan asks once per pass.
PassCosts cached_pass_costs(int dtype, const Grid3& g) {
  struct Entry {
    int dtype;
    int64_t n[3];
    int act[3];
    PassCosts c;
  };
  static std::mutex mu;
  static Entry cache[8];
  static int used = 0, next = 0;
  std::lock_guard<std::mutex> lock(mu);
  for (int i = 0; i < used; ++i) {
    const Entry& e = cache[i];
    if (e.dtype == dtype && e.n[0] == g.n[0] && e.n[1] == g.n[1] && e.n[2] == g.n[2] &&
        e.act[0] == g.act[0] && e.act[1] == g.act[1] && e.act[2] == g.act[2])
      return e.c;
  }
  Entry& e = cache[next];
  e = Entry{dtype, {g.n[0], g.n[1], g.n[2]}, {g.act[0], g.act[1], g.act[2]},
            pass_costs(dtype, g)};
  next = (next + 1) % 8;
  used = std::min(used + 1, 8);
  return e.c;
}
}  // namespace

int plan_pass(int dtype, const Grid3& g, int left, int cap) {
  cap = std::min(std::min(cap, left), kMaxTimeBlock);
  if (cap <= 1) return 1;
  const PassCosts pc = cached_pass_costs(dtype, g);
  const float* us = pc.us;
  bool ok[kMaxTimeBlock + 1] = {};
  ok[1] = true;
  for (int k = 2; k <= cap; ++k) ok[k] = us[k] > 0;
  if (left > kPlanTail) {  // the least time per step
    int best = 1;
    for (int k = 2; k <= cap; ++k)
      if (ok[k] && us[k] * best < us[best] * k) best = k;
    return best;
  }
  // Least total time over the splits of `left` (deeper first on ties).
  float t[kPlanTail + 1];
  int first[kPlanTail + 1];
  t[0] = 0.0f;
  first[0] = 0;
  for (int n = 1; n <= left; ++n) {
    t[n] = -1.0f;
    first[n] = 1;
    for (int k = std::min(n, cap); k >= 1; --k) {
      if (!ok[k]) continue;
      const float c = t[n - k] + us[k];
      if (t[n] < 0.0f || c < t[n]) {
        t[n] = c;
        first[n] = k;
      }
    }
  }
  return first[left];
}

const char* tuning_error() {
  const Tuning& t = tuning();
  return t.error.empty() ? nullptr : t.error.c_str();
}

uint32_t sr_step_key(uint64_t seed, uint64_t step) {
  const uint32_t k1 = lowbias32((uint32_t)step + 0x9E3779B9u * (uint32_t)(step >> 32));
  return lowbias32(k1 ^ (uint32_t)seed ^ ((uint32_t)(seed >> 32) * 0x85EBCA6Bu));
}

hipError_t launch_fused_pass(int dtype, const Grid3& g, const void* src, void* dst, int k,
                             hipStream_t s, void* const* levels, uint64_t step0, uint64_t seed) {
  if (k < 2 || src == dst) return hipErrorNotSupported;
  switch (dtype) {
    case FDB_F64: return tb_dispatch<double>(g, (const double*)src, (double*)dst, k, s, levels);
    case FDB_F32: return tb_dispatch<float>(g, (const float*)src, (float*)dst, k, s, levels);
    case FDB_BF16:
      return tb_dispatch<uint16_t>(g, (const uint16_t*)src, (uint16_t*)dst, k, s, levels);
    case FDB_BF16_SR: {
      if (levels || k > kMaxTimeBlock) return hipErrorNotSupported;
      uint32_t keys[kMaxTimeBlock];
      for (int l = 0; l < k; ++l) keys[l] = sr_step_key(seed, step0 + (uint64_t)l);
      return tb_dispatch<uint16_t>(g, (const uint16_t*)src, (uint16_t*)dst, k, s, nullptr,
                                   keys);
    }
    default: return hipErrorNotSupported;
  }
}

hipError_t launch_advect_pass(int dtype, const Grid3& g, void* const bufs[2], int* cur, int k,
                              hipStream_t s, int* launches, uint64_t step0, uint64_t seed) {
  if (bufs[0] == bufs[1] || (*cur != 0 && *cur != 1)) return hipErrorInvalidValue;
  hipError_t err = hipSuccess;
  int left = k, n = 0;
  uint64_t stepno = step0;
  while (left > 0) {
    const void* src = bufs[*cur];
    void* dst = bufs[1 - *cur];
    int done = 0;
    // Temporally blocked pass with the deepest supported K <= left.
    for (int K : {8, 7, 6, 4, 3, 2}) {
      if (K > left) continue;
      switch (dtype) {
        case FDB_F64: err = tb_dispatch<double>(g, (const double*)src, (double*)dst, K, s); break;
        case FDB_F32: err = tb_dispatch<float>(g, (const float*)src, (float*)dst, K, s); break;
        case FDB_BF16:
          err = tb_dispatch<uint16_t>(g, (const uint16_t*)src, (uint16_t*)dst, K, s);
          break;
        // Stochastic rounding: the bits of step stepno + l at level l.
        case FDB_BF16_SR:
          err = launch_fused_pass(dtype, g, src, dst, K, s, nullptr, stepno, seed);
          break;
        default: return hipErrorInvalidValue;
      }
      if (err == hipSuccess) {
        done = K;
        break;
      }
      if (err != hipErrorNotSupported) return err;
    }
    if (done == 0) {
      switch (dtype) {
        case FDB_F64: err = step_typed<double>(g, src, dst, s); break;
        case FDB_F32: err = step_typed<float>(g, src, dst, s); break;
        case FDB_BF16: err = step_typed<uint16_t>(g, src, dst, s); break;
        case FDB_BF16_SR:
          err = step_typed<uint16_t, true>(g, src, dst, s, sr_step_key(seed, stepno));
          break;
        default: return hipErrorInvalidValue;
      }
      if (err != hipSuccess) return err;
      done = 1;
    }
    *cur = 1 - *cur;
    left -= done;
    stepno += (uint64_t)done;
    ++n;
  }
  if (launches) *launches = n;
  return hipSuccess;
}

hipError_t launch_adjoint_step(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                               void* lam, const void* f_t, double* partials, double* acc,
                               hipStream_t s) {
  int64_t nb = 0;
  hipError_t err = adjoint_dispatch(lam_dtype, f_dtype, g, lam_next, lam, f_t, partials, &nb, s,
                                    false);
  if (err != hipSuccess || f_t == nullptr) return err;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partials, nb, 3, acc, 1);
  return hipGetLastError();
}

namespace {

// The (lam, f_t) storage pairings with a fused transpose pass: FP64/FP64,
// FP32/FP32 and FP32 lam over bf16 states (the mixed-precision gradient).
template <bool LAST>
hipError_t adj_fused_typed(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                           void* lam, const void* const* fs, int k, double* partials,
                           int64_t max_blocks, int64_t* nb, bool dry, hipStream_t s) {
  f_dtype = storage_dtype(f_dtype);
  if (lam_dtype == FDB_F64 && f_dtype == FDB_F64)
    return tb_adj_dispatch<double, double, LAST>(g, (const double*)lam_next, (double*)lam, fs,
                                                 k, partials, max_blocks, nb, dry, s);
  if (lam_dtype == FDB_F32 && f_dtype == FDB_F32)
    return tb_adj_dispatch<float, float, LAST>(g, (const float*)lam_next, (float*)lam, fs, k,
                                               partials, max_blocks, nb, dry, s);
  if (lam_dtype == FDB_F32 && f_dtype == FDB_BF16)
    return tb_adj_dispatch<float, uint16_t, LAST>(g, (const float*)lam_next, (float*)lam, fs,
                                                  k, partials, max_blocks, nb, dry, s);
  return hipErrorNotSupported;
}

}  // namespace

hipError_t launch_adjoint_last(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                               void* lam, const void* f, int k, double* partials,
                               int64_t max_blocks, double* acc, hipStream_t s) {
  if (lam_next == lam || f == nullptr || k < 2) return hipErrorNotSupported;
  int64_t nb = 0;
  const void* fs[1] = {f};
  const hipError_t err = adj_fused_typed<true>(lam_dtype, f_dtype, g, lam_next, lam, fs, k,
                                               partials, max_blocks, &nb, false, s);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partials, nb, 3, acc, 1);
  return hipGetLastError();
}

int adjoint_last_depth(int lam_dtype, int f_dtype, const Grid3& g, int64_t* nb) {
  for (int k : {7, 8}) {
    int64_t n = 0;
    if (adj_fused_typed<true>(lam_dtype, f_dtype, g, nullptr, nullptr, nullptr, k, nullptr, 0,
                              &n, true, nullptr) == hipSuccess) {
      if (nb) *nb = n;
      return k;
    }
  }
  return 0;
}

hipError_t launch_adjoint_fused(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                                void* lam, const void* const* fs, int k, double* partials,
                                int64_t max_blocks, double* acc, hipStream_t s) {
  if (lam_next == lam || k < 2) return hipErrorNotSupported;
  int64_t nb = 0;
  const hipError_t err = adj_fused_typed<false>(lam_dtype, f_dtype, g, lam_next, lam, fs, k,
                                                partials, max_blocks, &nb, false, s);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partials, nb, 3, acc, 1);
  return hipGetLastError();
}

int64_t adjoint_fused_blocks(int lam_dtype, int f_dtype, const Grid3& g, int k) {
  int64_t nb = 0;
  (void)adj_fused_typed<false>(lam_dtype, f_dtype, g, nullptr, nullptr, nullptr, k, nullptr, 0,
                               &nb, true, nullptr);
  return nb;
}

int64_t adjoint_blocks(int lam_dtype, int f_dtype, const Grid3& g, const void* lam_next,
                       void* lam, const void* f_t) {
  int64_t nb = 0;
  (void)adjoint_dispatch(lam_dtype, f_dtype, g, lam_next, lam, f_t, nullptr, &nb, nullptr, true);
  return nb;
}

hipError_t launch_sum(int dtype, const void* x, uint64_t n, double* partials, double* out2,
                      hipStream_t s) {
  dtype = storage_dtype(dtype);
  const dim3 grid(kReduceBlocks), block(256);
  switch (dtype) {
    case FDB_F64: hipLaunchKernelGGL(sum_kernel<double>, grid, block, 0, s, (const double*)x, n, partials); break;
    case FDB_F32: hipLaunchKernelGGL(sum_kernel<float>, grid, block, 0, s, (const float*)x, n, partials); break;
    case FDB_BF16: hipLaunchKernelGGL(sum_kernel<uint16_t>, grid, block, 0, s, (const uint16_t*)x, n, partials); break;
    default: return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partials,
                     (int64_t)kReduceBlocks, 1, out2, 0);
  return hipGetLastError();
}

hipError_t launch_misfit(int f_dtype, const void* f, const void* target, uint64_t n,
                         int lam_dtype, void* lam, double* partials, double* out2,
                         hipStream_t s) {
  f_dtype = storage_dtype(f_dtype);
  const dim3 grid(kReduceBlocks), block(256);
#define FDB_MF(TF, TL)                                                                      \
  hipLaunchKernelGGL((misfit_kernel<TF, TL>), grid, block, 0, s, (const TF*)f,              \
                     (const TF*)target, n, (TL*)lam, partials)
  if (f_dtype == FDB_F64 && lam_dtype == FDB_F64) FDB_MF(double, double);
  else if (f_dtype == FDB_F32 && lam_dtype == FDB_F64) FDB_MF(float, double);
  else if (f_dtype == FDB_BF16 && lam_dtype == FDB_F64) FDB_MF(uint16_t, double);
  else if (f_dtype == FDB_F64 && lam_dtype == FDB_F32) FDB_MF(double, float);
  else if (f_dtype == FDB_F32 && lam_dtype == FDB_F32) FDB_MF(float, float);
  else if (f_dtype == FDB_BF16 && lam_dtype == FDB_F32) FDB_MF(uint16_t, float);
  else return hipErrorInvalidValue;
#undef FDB_MF
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partials,
                     (int64_t)kReduceBlocks, 1, out2, 0);
  return hipGetLastError();
}

hipError_t launch_init(int dtype, const Grid3& g, const double* len, int kind,
                       const double* params, int nparams, uint64_t seed, void* f,
                       hipStream_t s) {
  dtype = storage_dtype(dtype);
  InitArgs a = {};
  int r = 0;
  for (int j = 0; j < 3; ++j) {
    a.n[j] = g.n[j];
    a.act[j] = g.act[j];
    a.dx[j] = g.act[j] ? len[r++] / (double)g.n[j] : 1.0;
  }
  a.kind = kind;
  a.seed = seed;
  const int cap = (int)(sizeof(a.p) / sizeof(a.p[0]));
  if (nparams < 0 || nparams > cap) return hipErrorInvalidValue;
  a.nparams = nparams;
  for (int i = 0; i < nparams; ++i) a.p[i] = params[i];
  // Centres are given per user axis (first ndims entries of each triple);
  // move them onto the padded trailing axes the kernel indexes.
  int nd = g.act[0] + g.act[1] + g.act[2];
  auto remap = [&](double* c3) {
    double u[3] = {c3[0], c3[1], c3[2]};
    for (int j = 0; j < 3; ++j) c3[j] = (j >= 3 - nd) ? u[j - (3 - nd)] : 0.0;
  };
  if (kind == FDB_INIT_SPHERE || kind == FDB_INIT_GAUSS) {
    if (nparams < 5) return hipErrorInvalidValue;
    remap(a.p);
  } else if (kind == FDB_INIT_GAUSS_SUM) {
    if (nparams < 1 || (int)a.p[0] < 0 || 1 + 5 * (int)a.p[0] > nparams) return hipErrorInvalidValue;
    for (int k = 0; k < (int)a.p[0]; ++k) remap(a.p + 1 + 5 * k);
  } else if (kind == FDB_INIT_UNIFORM) {
    if (nparams < 2) return hipErrorInvalidValue;
  } else if (kind != FDB_INIT_DELTA) {
    return hipErrorInvalidValue;
  }
  const dim3 grid(8192), block(256);
  switch (dtype) {
    case FDB_F64: hipLaunchKernelGGL(init_kernel<double>, grid, block, 0, s, (double*)f, a); break;
    case FDB_F32: hipLaunchKernelGGL(init_kernel<float>, grid, block, 0, s, (float*)f, a); break;
    case FDB_BF16: hipLaunchKernelGGL(init_kernel<uint16_t>, grid, block, 0, s, (uint16_t*)f, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_convert(int sd, const void* src, int dd, void* dst, uint64_t n, hipStream_t s) {
  sd = storage_dtype(sd);
  dd = storage_dtype(dd);
  const dim3 grid(4096), block(256);
#define FDB_CV(TA, TB) \
  hipLaunchKernelGGL((convert_kernel<TA, TB>), grid, block, 0, s, (const TA*)src, (TB*)dst, n)
  if (sd == FDB_F64 && dd == FDB_F32) FDB_CV(double, float);
  else if (sd == FDB_F64 && dd == FDB_BF16) FDB_CV(double, uint16_t);
  else if (sd == FDB_F32 && dd == FDB_F64) FDB_CV(float, double);
  else if (sd == FDB_F32 && dd == FDB_BF16) FDB_CV(float, uint16_t);
  else if (sd == FDB_BF16 && dd == FDB_F64) FDB_CV(uint16_t, double);
  else if (sd == FDB_BF16 && dd == FDB_F32) FDB_CV(uint16_t, float);
  else return hipErrorInvalidValue;
#undef FDB_CV
  return hipGetLastError();
}

hipError_t launch_compare(int da, const void* a, int db, const void* b, uint64_t n,
                          double* partials, hipStream_t s) {
  da = storage_dtype(da);
  db = storage_dtype(db);
  const dim3 grid(kReduceBlocks), block(256);
#define FDB_CMP(TA, TB) \
  hipLaunchKernelGGL((compare_kernel<TA, TB>), grid, block, 0, s, (const TA*)a, (const TB*)b, n, partials)
#define FDB_CMP_B(TA)                                 \
  if (db == FDB_F64) FDB_CMP(TA, double);            \
  else if (db == FDB_F32) FDB_CMP(TA, float);        \
  else if (db == FDB_BF16) FDB_CMP(TA, uint16_t);    \
  else return hipErrorInvalidValue;
  if (da == FDB_F64) { FDB_CMP_B(double) }
  else if (da == FDB_F32) { FDB_CMP_B(float) }
  else if (da == FDB_BF16) { FDB_CMP_B(uint16_t) }
  else return hipErrorInvalidValue;
#undef FDB_CMP_B
#undef FDB_CMP
  return hipGetLastError();
}

hipError_t launch_zero(double* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace fdb
