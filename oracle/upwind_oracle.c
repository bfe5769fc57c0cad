/*
 * upwind_oracle.c -- CPU restatement of FiDiBench's upwind advection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (fidibench_amd/csrc) never links or calls it.
 *
 * Reference followed (pletzer/fidibench, upwind/cxx/upwind.cxx):
 *   ctor     :23-49   u_j = (v_j < 0.) ? +1 : -1          (:34-35)
 *                     dx_j = L_j / N_j                     (:36)
 *                     dimProd, last axis contiguous        (:39-44)
 *                     f zero-filled, f[0] = 1              (:45-48)
 *   advect   :51-86   per step: oldF = f (:59-62); then for every cell i
 *                     and axis j = 0..NDIMS-1 in order:
 *                       coeff = dt * v_j * u_j / dx_j      (:72, evaluated
 *                               left to right: ((dt*v)*u)/dx)
 *                       up    = i with index j moved to (i_j+u_j+N_j)%N_j
 *                                                          (:75-78)
 *                       f[i]  = f[i] - coeff*(oldF[up] - oldF[i])   (:80)
 *   checksum :91-93   std::accumulate from 0.0: strict left fold
 *   std      :95-103  mean = checksum/ntot; left fold of (f-mean)^2;
 *                     sqrt(res/ntot)
 * The FP64 functions must be compiled with -ffp-contract=off: a fused
 * multiply-add changes the bits (the reference is built without FMA).
 *
 * Extensions with no counterpart in the reference, pinned in tests by
 * mathematics (dot-product test, exact dyadic arithmetic, closed forms,
 * finite differences):
 *   orc_advect_f32 / orc_advect_bf16  reduced-precision storage, float math
 *   orc_advect_bf16_sr                bf16 storage, stochastic rounding
 *   orc_adjoint_step_f64              exact transpose of one forward step
 *   orc_adjoint_step_lowp             the same in float math, f32/bf16 storage
 *   orc_vgrad_terms                   velocity-gradient dot products
 *   orc_misfit_grad_f64               loss, dL/df0 and dL/dv by adjoint
 *   orc_tlm_f64                       tangent-linear model (directional
 *                                     derivative of f_T)
 *   orc_init_*                        synthetic initial fields
 *
 * Also the generic stencil filter of cxx/Filter.cpp (one process),
 * orc_filter_apply, used to check the GPU Filter.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXD 3

/* Geometry shared by every routine: the reference's ctor state (:23-49). */
typedef struct {
  int nd;
  uint64_t n[ORC_MAXD];
  uint64_t stride[ORC_MAXD]; /* dimProd */
  int up[ORC_MAXD];          /* upDirection */
  double coeff[ORC_MAXD];    /* ((dt*v)*u)/dx, upwind.cxx:72 */
  uint64_t ntot;
} orc_geom;

static int orc_geom_init(orc_geom* g, int nd, const uint64_t* n, const double* v,
                         const double* len, double dt) {
  if (nd < 1 || nd > ORC_MAXD) return -1;
  g->nd = nd;
  g->ntot = 1;
  for (int j = 0; j < nd; ++j) {
    if (n[j] == 0) return -1;
    g->n[j] = n[j];
    g->up[j] = (v[j] < 0.) ? +1 : -1;
    double dx = len[j] / (double)n[j];
    g->coeff[j] = dt * v[j] * (double)g->up[j] / dx;
    g->ntot *= n[j];
  }
  g->stride[nd - 1] = 1;
  for (int j = nd - 2; j >= 0; --j) g->stride[j] = g->stride[j + 1] * g->n[j + 1];
  return 0;
}

/* Flat index of the upwind neighbour of cell i along axis j, with the
 * periodic wrap of upwind.cxx:75-76. `dir` is +u_j for the forward
 * operator and -u_j for its transpose. */
static inline uint64_t orc_shift(const orc_geom* g, uint64_t i, int j, int dir) {
  uint64_t ij = (i / g->stride[j]) % g->n[j];
  uint64_t nj = (ij + (uint64_t)(int64_t)dir + g->n[j]) % g->n[j];
  return i - ij * g->stride[j] + nj * g->stride[j];
}

/* ---------------------------------------------------------------- FP64 --- */

/* One forward step old -> f, upwind.cxx:64-84. */
static void orc_step_f64(const orc_geom* g, const double* old, double* f) {
  const int64_t ntot = (int64_t)g->ntot;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ntot; ++i) {
    double fi = old[i];
    for (int j = 0; j < g->nd; ++j) {
      uint64_t up = orc_shift(g, (uint64_t)i, j, g->up[j]);
      fi = fi - g->coeff[j] * (old[up] - old[i]);
    }
    f[i] = fi;
  }
}

/* advect(numTimeSteps, deltaTime) on f in place, upwind.cxx:51-86. */
int orc_advect_f64(int nd, const uint64_t* n, const double* v, const double* len,
                   double dt, int nsteps, double* f) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  double* old = (double*)malloc(g.ntot * sizeof(double));
  if (!old) return -2;
  for (int s = 0; s < nsteps; ++s) {
    memcpy(old, f, g.ntot * sizeof(double)); /* :59-62 */
    orc_step_f64(&g, old, f);
  }
  free(old);
  return 0;
}

/* checksum(), upwind.cxx:91-93: left fold from 0.0. */
double orc_checksum(const double* f, uint64_t n) {
  double s = 0.0;
  for (uint64_t i = 0; i < n; ++i) s = s + f[i];
  return s;
}

/* std(), upwind.cxx:95-103. */
double orc_std(const double* f, uint64_t n) {
  double mean = orc_checksum(f, n) / (double)n;
  double res = 0;
  for (uint64_t i = 0; i < n; ++i) {
    double d = f[i] - mean;
    res += d * d;
  }
  return sqrt(res / (double)n);
}

/* ------------------------------------------------------ reduced precision */

/* bfloat16 <-> float.  Round to nearest, ties to even, on the top 16 bits
 * of the binary32 pattern; NaN stays a quiet NaN. */
uint16_t orc_f32_to_bf16(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu))
    return (uint16_t)((b >> 16) | 0x0040u);
  uint32_t lsb = (b >> 16) & 1u;
  b += 0x7fffu + lsb;
  return (uint16_t)(b >> 16);
}

float orc_bf16_to_f32(uint16_t h) {
  uint32_t b = (uint32_t)h << 16;
  float x;
  memcpy(&x, &b, 4);
  return x;
}

/* FP32 storage and arithmetic: the FP64 coefficient is rounded once to
 * float, then the same operation order as the FP64 step. */
int orc_advect_f32(int nd, const uint64_t* n, const double* v, const double* len,
                   double dt, int nsteps, float* f) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  float cf[ORC_MAXD];
  for (int j = 0; j < nd; ++j) cf[j] = (float)g.coeff[j];
  float* old = (float*)malloc(g.ntot * sizeof(float));
  if (!old) return -2;
  const int64_t ntot = (int64_t)g.ntot;
  for (int s = 0; s < nsteps; ++s) {
    memcpy(old, f, g.ntot * sizeof(float));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ntot; ++i) {
      float fi = old[i];
      for (int j = 0; j < nd; ++j) {
        uint64_t up = orc_shift(&g, (uint64_t)i, j, g.up[j]);
        fi = fi - cf[j] * (old[up] - old[i]);
      }
      f[i] = fi;
    }
  }
  free(old);
  return 0;
}

/* bfloat16 storage, float arithmetic: widen, step in float exactly as
 * orc_advect_f32, round the result once to bfloat16. */
int orc_advect_bf16(int nd, const uint64_t* n, const double* v, const double* len,
                    double dt, int nsteps, uint16_t* f) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  float cf[ORC_MAXD];
  for (int j = 0; j < nd; ++j) cf[j] = (float)g.coeff[j];
  float* old = (float*)malloc(g.ntot * sizeof(float));
  if (!old) return -2;
  const int64_t ntot = (int64_t)g.ntot;
  for (int s = 0; s < nsteps; ++s) {
    for (int64_t i = 0; i < ntot; ++i) old[i] = orc_bf16_to_f32(f[i]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ntot; ++i) {
      float fi = old[i];
      for (int j = 0; j < nd; ++j) {
        uint64_t up = orc_shift(&g, (uint64_t)i, j, g.up[j]);
        fi = fi - cf[j] * (old[up] - old[i]);
      }
      f[i] = orc_f32_to_bf16(fi);
    }
  }
  free(old);
  return 0;
}

/* bfloat16 storage with stochastic rounding of every step's result: 16
 * random bits added below the kept bits of the binary32 value, then
 * truncated.  Cells pair along the second-to-last axis (axis 1 of the grid
 * padded to three axes): a cell's pair is the flat index of the cell with
 * that index rounded down to even (1-D grids: the cell itself).  The bits
 * come from lowbias32 (C. Wellons' integer hash) of the folded pair index,
 * xor a per-step key hashed from (seed, step), through one multiply-
 * xorshift round (lowbias32's first round and closing shift): the low half
 * for the even row, the high half for the odd one. */
static uint32_t orc_lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

uint32_t orc_sr_step_key(uint64_t seed, uint64_t step) {
  uint32_t k1 = orc_lowbias32((uint32_t)step + 0x9E3779B9u * (uint32_t)(step >> 32));
  return orc_lowbias32(k1 ^ (uint32_t)seed ^ ((uint32_t)(seed >> 32) * 0x85EBCA6Bu));
}

static uint32_t orc_sr_round(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  return x ^ (x >> 16);
}

uint32_t orc_sr_bits(uint64_t pair, int odd, uint32_t key) {
  const uint32_t h = orc_sr_round(orc_lowbias32((uint32_t)(pair ^ (pair >> 32))) ^ key);
  return odd ? (h >> 16) : (h & 0xffffu);
}

uint16_t orc_bf16_sr(float x, uint64_t pair, int odd, uint32_t key) {
  uint32_t b;
  memcpy(&b, &x, 4);
  if ((b & 0x7f800000u) == 0x7f800000u)
    return (uint16_t)((b >> 16) | ((b & 0x007fffffu) ? 0x0040u : 0u));
  return (uint16_t)((b + orc_sr_bits(pair, odd, key)) >> 16);
}

/* As orc_advect_bf16, with step number step0 + s rounded stochastically. */
int orc_advect_bf16_sr(int nd, const uint64_t* n, const double* v, const double* len,
                       double dt, int nsteps, uint64_t seed, uint64_t step0, uint16_t* f) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  float cf[ORC_MAXD];
  for (int j = 0; j < nd; ++j) cf[j] = (float)g.coeff[j];
  float* old = (float*)malloc(g.ntot * sizeof(float));
  if (!old) return -2;
  const int64_t ntot = (int64_t)g.ntot;
  for (int s = 0; s < nsteps; ++s) {
    const uint32_t key = orc_sr_step_key(seed, step0 + (uint64_t)s);
    for (int64_t i = 0; i < ntot; ++i) old[i] = orc_bf16_to_f32(f[i]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ntot; ++i) {
      float fi = old[i];
      for (int j = 0; j < nd; ++j) {
        uint64_t up = orc_shift(&g, (uint64_t)i, j, g.up[j]);
        fi = fi - cf[j] * (old[up] - old[i]);
      }
      // the pair along the second-to-last axis
      const uint64_t st = nd >= 2 ? g.stride[nd - 2] : 1;
      const int odd = nd >= 2 ? (int)(((uint64_t)i / st) % g.n[nd - 2] & 1u) : 0;
      f[i] = orc_bf16_sr(fi, (uint64_t)i - (uint64_t)odd * st, odd, key);
    }
  }
  free(old);
  return 0;
}

/* ------------------------------------------------------------- adjoint --- */

/* One adjoint step lam_next -> lam: the transpose of the forward step
 * f' = f - sum_j c_j (S_j f - f), (S_j f)[i] = f[i + u_j e_j].  Its
 * transpose is lam = lam' - sum_j c_j (S_j^T lam' - lam'), with
 * (S_j^T lam')[i] = lam'[i - u_j e_j]: the same coefficients, gathered
 * from the downwind side, in the same axis order. */
void orc_adjoint_step_f64(int nd, const uint64_t* n, const double* v, const double* len,
                          double dt, const double* lam_next, double* lam) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return;
  const int64_t ntot = (int64_t)g.ntot;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ntot; ++i) {
    double li = lam_next[i];
    for (int j = 0; j < g.nd; ++j) {
      uint64_t dn = orc_shift(&g, (uint64_t)i, j, -g.up[j]);
      li = li - g.coeff[j] * (lam_next[dn] - lam_next[i]);
    }
    lam[i] = li;
  }
}

/* Storage kinds of the mixed-precision helpers: 0 FP64, 1 FP32, 2 bf16. */
static double orc_widen(int kind, const void* p, uint64_t i) {
  if (kind == 0) return ((const double*)p)[i];
  if (kind == 1) return (double)((const float*)p)[i];
  return (double)orc_bf16_to_f32(((const uint16_t*)p)[i]);
}

/* One adjoint step with float arithmetic (the FP64 coefficients rounded
 * once to float) on float (kind 1) or bfloat16 (kind 2: widened, result
 * rounded to nearest even) storage, in the operation order of
 * orc_adjoint_step_f64 -- the reduced-precision variants of the forward
 * (orc_advect_f32/bf16) applied to the transpose. */
int orc_adjoint_step_lowp(int nd, const uint64_t* n, const double* v, const double* len,
                          double dt, int kind, const void* lam_next, void* lam) {
  orc_geom g;
  if ((kind != 1 && kind != 2) || orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  float cf[ORC_MAXD];
  for (int j = 0; j < nd; ++j) cf[j] = (float)g.coeff[j];
  const int64_t ntot = (int64_t)g.ntot;
  float* w = (float*)malloc(g.ntot * sizeof(float));
  if (!w) return -2;
  for (int64_t i = 0; i < ntot; ++i) w[i] = (float)orc_widen(kind, lam_next, (uint64_t)i);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ntot; ++i) {
    float li = w[i];
    for (int j = 0; j < nd; ++j) {
      uint64_t dn = orc_shift(&g, (uint64_t)i, j, -g.up[j]);
      li = li - cf[j] * (w[dn] - w[i]);
    }
    if (kind == 1)
      ((float*)lam)[i] = li;
    else
      ((uint16_t*)lam)[i] = orc_f32_to_bf16(li);
  }
  free(w);
  return 0;
}

/* The velocity-gradient dot products of one adjoint step,
 *   D_j = sum_i f_t[i] * (lam_next[i - u_j e_j] - lam_next[i]),
 * with each value widened to double, each difference and product in
 * double, and the sum a long-double left fold (dL/dv_j gets
 * -(dt*u_j/dx_j) * D_j per step).  Kinds as orc_widen. */
int orc_vgrad_terms(int nd, const uint64_t* n, const double* v, const double* len, double dt,
                    int lam_kind, const void* lam_next, int f_kind, const void* f_t,
                    double* out) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  for (int j = 0; j < nd; ++j) {
    long double s = 0.0L;
    for (uint64_t i = 0; i < g.ntot; ++i) {
      uint64_t dn = orc_shift(&g, i, j, -g.up[j]);
      const double d = orc_widen(lam_kind, lam_next, dn) - orc_widen(lam_kind, lam_next, i);
      s += (long double)(orc_widen(f_kind, f_t, i) * d);
    }
    out[j] = (double)s;
  }
  return 0;
}

int orc_adjoint_f64(int nd, const uint64_t* n, const double* v, const double* len,
                    double dt, int nsteps, double* lam) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  double* nxt = (double*)malloc(g.ntot * sizeof(double));
  if (!nxt) return -2;
  for (int s = 0; s < nsteps; ++s) {
    memcpy(nxt, lam, g.ntot * sizeof(double));
    orc_adjoint_step_f64(nd, n, v, len, dt, nxt, lam);
  }
  free(nxt);
  return 0;
}

/* d c_j / d v_j for c_j = ((dt*v_j)*u_j)/dx_j with u_j held fixed. */
static double orc_dcoeff(const orc_geom* g, const double* len, double dt, int j) {
  double dx = len[j] / (double)g->n[j];
  return dt * (double)g->up[j] / dx;
}

/* Misfit loss L = 0.5 * sum (f_T - target)^2 after nsteps forward steps
 * from f0 (FP64), and its gradient by the discrete adjoint:
 *   lam_T = f_T - target;  lam_t = A^T lam_{t+1};
 *   dL/dv_j = sum_t sum_i lam_{t+1}[i] * (-(dt*u_j/dx_j)) * (f_t[up_j(i)] - f_t[i])
 * dt is a fixed input (not a function of v).  All states are stored
 * (small grids only).  Sums over cells are plain left folds in long
 * double. */
int orc_misfit_grad_f64(int nd, const uint64_t* n, const double* v, const double* len,
                        double dt, int nsteps, const double* f0, const double* target,
                        double* loss, double* grad_f0, double* grad_v) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  const uint64_t nt = g.ntot;
  double* states = (double*)malloc((size_t)(nsteps + 1) * nt * sizeof(double));
  double* lam2 = (double*)malloc(nt * sizeof(double));
  if (!states || !lam2) { free(states); free(lam2); return -2; }
  memcpy(states, f0, nt * sizeof(double));
  for (int s = 0; s < nsteps; ++s) orc_step_f64(&g, states + (size_t)s * nt, states + (size_t)(s + 1) * nt);
  const double* fT = states + (size_t)nsteps * nt;
  long double acc = 0.0L;
  for (uint64_t i = 0; i < nt; ++i) {
    double d = fT[i] - target[i];
    grad_f0[i] = d;
    acc += (long double)d * (long double)d;
  }
  *loss = (double)(0.5L * acc);
  long double gv[ORC_MAXD] = {0.0L, 0.0L, 0.0L};
  for (int s = nsteps - 1; s >= 0; --s) {
    const double* ft = states + (size_t)s * nt;
    for (int j = 0; j < nd; ++j) {
      long double dj = 0.0L;
      for (uint64_t i = 0; i < nt; ++i) {
        uint64_t up = orc_shift(&g, i, j, g.up[j]);
        dj += (long double)grad_f0[i] * (long double)(ft[up] - ft[i]);
      }
      gv[j] += -(long double)orc_dcoeff(&g, len, dt, j) * dj;
    }
    memcpy(lam2, grad_f0, nt * sizeof(double));
    orc_adjoint_step_f64(nd, n, v, len, dt, lam2, grad_f0);
  }
  for (int j = 0; j < nd; ++j) grad_v[j] = (double)gv[j];
  free(states);
  free(lam2);
  return 0;
}

/* Tangent-linear model: the directional derivative of f_T along
 * (df0, dv), propagated step by step beside the forward state:
 *   df_{t+1} = A df_t - sum_j (dc_j/dv_j dv_j) (S_j f_t - f_t). */
int orc_tlm_f64(int nd, const uint64_t* n, const double* v, const double* len, double dt,
                int nsteps, const double* f0, const double* df0, const double* dv,
                double* fT, double* dfT) {
  orc_geom g;
  if (orc_geom_init(&g, nd, n, v, len, dt)) return -1;
  const uint64_t nt = g.ntot;
  double* f = (double*)malloc(nt * sizeof(double));
  double* df = (double*)malloc(nt * sizeof(double));
  double* fo = (double*)malloc(nt * sizeof(double));
  double* dfo = (double*)malloc(nt * sizeof(double));
  if (!f || !df || !fo || !dfo) { free(f); free(df); free(fo); free(dfo); return -2; }
  double dc[ORC_MAXD];
  for (int j = 0; j < nd; ++j) dc[j] = orc_dcoeff(&g, len, dt, j) * dv[j];
  memcpy(f, f0, nt * sizeof(double));
  memcpy(df, df0, nt * sizeof(double));
  for (int s = 0; s < nsteps; ++s) {
    memcpy(fo, f, nt * sizeof(double));
    memcpy(dfo, df, nt * sizeof(double));
    for (uint64_t i = 0; i < nt; ++i) {
      double fi = fo[i], di = dfo[i];
      for (int j = 0; j < nd; ++j) {
        uint64_t up = orc_shift(&g, i, j, g.up[j]);
        fi = fi - g.coeff[j] * (fo[up] - fo[i]);
        di = di - g.coeff[j] * (dfo[up] - dfo[i]) - dc[j] * (fo[up] - fo[i]);
      }
      f[i] = fi;
      df[i] = di;
    }
  }
  memcpy(fT, f, nt * sizeof(double));
  memcpy(dfT, df, nt * sizeof(double));
  free(f); free(df); free(fo); free(dfo);
  return 0;
}

/* ------------------------------------------------------ initial fields --- */

/* SplitMix64 (Steele, Lea & Flood 2014), used counter-style: the draw for
 * cell i under `seed` is mix(seed + (i+1)*golden), uniform in [0,1) as
 * (z >> 11) * 2^-53. */
uint64_t orc_splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double orc_uniform(uint64_t seed, uint64_t i) {
  return (double)(orc_splitmix64(seed, i) >> 11) * 0x1.0p-53;
}

/* Sphere indicator plus optional uniform noise:
 *   f = (|x - c|^2 < r^2 ? 1 : 0) + amp * U(seed, i),
 * cell centres x_j = (i_j + 0.5) * (L_j / N_j); |x-c|^2 summed left to
 * right over the axes.  nd must be 3. */
void orc_init_sphere_noise(const uint64_t* n, const double* len, const double* centre,
                           double radius, double amp, uint64_t seed, double* f) {
  const double r2 = radius * radius;
  const double dx0 = len[0] / (double)n[0], dx1 = len[1] / (double)n[1],
               dx2 = len[2] / (double)n[2];
  const int64_t n0 = (int64_t)n[0];
#pragma omp parallel for schedule(static)
  for (int64_t i0 = 0; i0 < n0; ++i0)
    for (uint64_t i1 = 0; i1 < n[1]; ++i1)
      for (uint64_t i2 = 0; i2 < n[2]; ++i2) {
        uint64_t i = ((uint64_t)i0 * n[1] + i1) * n[2] + i2;
        double d0 = ((double)i0 + 0.5) * dx0 - centre[0];
        double d1 = ((double)i1 + 0.5) * dx1 - centre[1];
        double d2 = ((double)i2 + 0.5) * dx2 - centre[2];
        double s = d0 * d0 + d1 * d1 + d2 * d2;
        double val = (s < r2) ? 1.0 : 0.0;
        if (amp != 0.0) val = val + amp * orc_uniform(seed, i);
        f[i] = val;
      }
}

/* Uniform noise in [lo, hi): lo + (hi - lo) * U(seed, i). */
void orc_init_uniform(uint64_t ntot, double lo, double hi, uint64_t seed, double* f) {
  for (uint64_t i = 0; i < ntot; ++i) f[i] = lo + (hi - lo) * orc_uniform(seed, i);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

double orc_wtime(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

/* ------------------------------------------------------------- filter --- */

/* One applyFilter() of pletzer/fidibench cxx/Filter.cpp:192-263 with one
 * process (every neighbour local).  Column-major cells (axis 0 fastest,
 * Filter.cpp:50); the terms are taken in std::map order of their offset
 * vectors (sorted here, lexicographically; of equal offsets the first is
 * kept, as std::map::insert does); per cell out = 0, then out += w * in[nbr]
 * term by term.  The neighbour index per axis is ((int)i + offset) % n with
 * an unsigned (size_t) n, so a negative index wraps as its two's-complement
 * unsigned value (Filter.cpp:231-233). */
typedef struct {
  int off[ORC_MAXD];
  double w;
  int order;
} orc_term;

static int orc_ndims_cmp;
static int orc_term_cmp(const void* a, const void* b) {
  const orc_term* x = (const orc_term*)a;
  const orc_term* y = (const orc_term*)b;
  for (int j = 0; j < orc_ndims_cmp; ++j)
    if (x->off[j] != y->off[j]) return x->off[j] < y->off[j] ? -1 : 1;
  return x->order - y->order;
}

int orc_filter_apply(int nd, const uint64_t* n, int nterms, const int* offsets,
                     const double* weights, const double* in, double* out) {
  if (nd < 1 || nd > ORC_MAXD || nterms < 1) return -1;
  orc_term* t = (orc_term*)malloc(sizeof(orc_term) * (size_t)nterms);
  if (!t) return -2;
  for (int k = 0; k < nterms; ++k) {
    for (int j = 0; j < nd; ++j) t[k].off[j] = offsets[(size_t)k * nd + j];
    t[k].w = weights[k];
    t[k].order = k;
  }
  orc_ndims_cmp = nd;
  qsort(t, (size_t)nterms, sizeof(orc_term), orc_term_cmp);
  int m = 0;  /* drop repeated offsets, keeping the first inserted */
  for (int k = 0; k < nterms; ++k) {
    int same = m > 0;
    for (int j = 0; j < nd && same; ++j) same = t[k].off[j] == t[m - 1].off[j];
    if (!same) t[m++] = t[k];
  }
  uint64_t ntot = 1, stride[ORC_MAXD];
  for (int j = 0; j < nd; ++j) {
    stride[j] = ntot;
    ntot *= n[j];
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)ntot; ++i) {
    double acc = 0;
    for (int k = 0; k < m; ++k) {
      uint64_t nb = 0;
      for (int j = 0; j < nd; ++j) {
        const uint64_t ij = ((uint64_t)i / stride[j]) % n[j];
        const int64_t shifted = (int64_t)ij + t[k].off[j];
        const uint64_t wrapped = (uint64_t)shifted % n[j];
        nb += wrapped * stride[j];
      }
      acc += t[k].w * in[nb];
    }
    out[i] = acc;
  }
  free(t);
  return 0;
}
