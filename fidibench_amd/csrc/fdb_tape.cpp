// fdb_tape.cpp -- reverse-mode gradient of a loss of f_T with respect to
// the initial field and the constant velocity, with binomial
// checkpointing.  No counterpart in the reference (an extension of the
// Upwind path); the adjoint kernel itself is in fdb_kernels.hip.
//
// Schedule: reversing n steps with c free snapshot slots (the initial
// state is held separately) is the classical binomial problem with
// s = c + 1 snapshots (Griewank 1992; Griewank & Walther, ACM TOMS 26,
// 2000).  With the repetition number r = min{r : C(s+r, s) >= n}, the
// first checkpoint is placed at distance
//     m = max(1, C(s+r-2, s), n - C(s-1+r, s-1))
// and the two parts are reversed recursively; the total number of forward
// steps is then r*n - C(s+r, s+1), the optimum.
//
// Memory traffic: the working state is a pointer.  A step reads it and
// writes a scratch buffer -- or, when the schedule stores the state it
// produces, the snapshot slot itself -- so storing and restoring cost no
// copies; with enough memory for every state the forward pass moves the
// same bytes as a plain advect.
//
// Shift-invariant method (FDB_GRAD_SHIFT).  The step is A = I - sum_j c_j
// (S_j - I) with S_j the periodic unit shift along axis j and constant
// c_j, so A, A^T and every S_j are circulant and commute.  With
// lam_{t+1} = (A^T)^(T-1-t) lam_T and f_t = A^t f_0,
//     <lam_{t+1}, (S_j - I) f_t> = <lam_T, A^(T-1-t) (S_j - I) A^t f_0>
//                                = <lam_T, (S_j - I) A^(T-1) f_0>,
// the same for every t: the velocity products of the T adjoint steps all
// equal those of the last one (t = 0), D_j = T * <lam_1, (S_j - I) f_0>.
// The backward is then T-1 plain transpose steps (fused K-step passes of
// the forward kernel with the directions negated; lam_0 bit-identical to
// the checkpointed sweep, whose lam recursion does not involve f either)
// and one transpose step with the products against f_0 (the last level of
// the last fused pass where the storage pairing allows): no state but f_0
// is needed, so nothing is stored or recomputed.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fidibench_amd/fdb_upwind.h"
#include "fdb_internal.h"

namespace {

// C(s + r, s), saturating well below INT64_MAX.
int64_t beta(int64_t s, int64_t r) {
  if (r < 0) return 0;
  const int64_t cap = (int64_t)1 << 60;
  int64_t b = 1;
  for (int64_t i = 1; i <= std::min(s, r); ++i) {
    // b = C(s+r, i) computed incrementally; exact while below the cap.
    const __int128 nb = (__int128)b * (s + r - i + 1) / i;
    if (nb > cap) return cap;
    b = (int64_t)nb;
  }
  return b;
}

int64_t repetitions(int64_t n, int64_t s) {
  int64_t r = 0;
  while (beta(s, r) < n) ++r;
  return r;
}

struct Plan {
  std::vector<std::pair<int, int>> acts;
  int64_t forward = 0;
};

// Precondition: the working state is state `s` and the top stored state
// is state `s`.  Emits the actions that run the adjoint steps e-1 .. s.
void treeverse(Plan& p, int s, int e, int c) {
  const int n = e - s;
  if (n == 1) {
    p.acts.push_back({FDB_CK_ADJOINT, s});
    return;
  }
  if (c == 0) {
    for (int t = e - 1; t >= s; --t) {
      p.acts.push_back({FDB_CK_RESTORE, 0});
      if (t > s) {
        p.acts.push_back({FDB_CK_ADVANCE, t - s});
        p.forward += t - s;
      }
      p.acts.push_back({FDB_CK_ADJOINT, t});
    }
    return;
  }
  const int64_t S = c + 1;
  const int64_t r = repetitions(n, S);
  int64_t m = std::max<int64_t>({1, beta(S, r - 2), (int64_t)n - beta(S - 1, r)});
  m = std::min<int64_t>(m, n - 1);
  p.acts.push_back({FDB_CK_ADVANCE, (int)m});
  p.forward += m;
  p.acts.push_back({FDB_CK_STORE, s + (int)m});
  treeverse(p, s + (int)m, e, c - 1);
  p.acts.push_back({FDB_CK_POP, 0});
  p.acts.push_back({FDB_CK_RESTORE, 0});
  treeverse(p, s, s + (int)m, c);
}

Plan make_plan(int nsteps, int snaps) {
  Plan p;
  if (nsteps > 0) treeverse(p, 0, nsteps, std::max(0, std::min(snaps, nsteps - 1)));
  return p;
}

int tfail(int code, const std::string& m) {
  fdb::set_last_error(m);
  return code;
}

}  // namespace

struct fdb_tape {
  int ndims = 0;
  int nsteps = 0;
  int f_dtype = FDB_F64, lam_dtype = FDB_F64;
  fdb::Grid3 g{};
  size_t fbytes = 0, lbytes = 0;
  int snaps = 0;
  Plan plan;
  size_t resume = 0;  // first plan action left for backward
  // device memory
  void* base = nullptr;          // state 0 (a copy of f0)
  std::vector<void*> slots;      // snapshot storage, used as a stack
  std::vector<int> stack_steps;  // steps of the stored states (top = back)
  void* scratch[2] = {nullptr, nullptr};
  const void* w = nullptr;       // where the working state lives
  int wstep = -1;                // which state it is
  int time_block = fdb::kMaxTimeBlock;
  void* lam_tmp = nullptr;
  double* partials = nullptr;
  size_t partials_bytes = 0;
  double* acc = nullptr;
  bool forward_done = false;
  bool fuse_adj = false;  // this backward may use fused transpose passes
  int64_t fwd_steps = 0, adj_steps = 0;
  int method = FDB_GRAD_CHECKPOINT;
  bool slots_ready = false;    // snapshot slots allocated (released by FDB_GRAD_SHIFT)
  const void* f0_ref = nullptr;  // FDB_GRAD_SHIFT: the caller's f0
};

namespace {

#define TAPE_HIP(call)                                                          \
  do {                                                                          \
    hipError_t _e = (call);                                                     \
    if (_e != hipSuccess) {                                                     \
      (void)hipGetLastError();                                                  \
      return tfail(_e == hipErrorOutOfMemory ? FDB_ERR_NOMEM : FDB_ERR_HIP,     \
                   std::string(#call) + ": " + hipGetErrorString(_e));          \
    }                                                                           \
  } while (0)

void tape_free(fdb_tape* t) {
  if (!t) return;
  (void)hipDeviceSynchronize();
  if (t->base) (void)hipFree(t->base);
  for (void* s : t->slots) (void)hipFree(s);
  for (void* w : t->scratch)
    if (w) (void)hipFree(w);
  if (t->lam_tmp) (void)hipFree(t->lam_tmp);
  if (t->partials) (void)hipFree(t->partials);
  if (t->acc) (void)hipFree(t->acc);
  delete t;
}

// One step from the working state into dst (dst != working state).
int step_into(fdb_tape* t, void* dst, hipStream_t s) {
  void* const pair[2] = {const_cast<void*>(t->w), dst};
  int cur = 0;
  // The step's absolute index keys stochastic rounding (FDB_BF16_SR), so a
  // recomputed state has the same bits as the one the forward produced.
  TAPE_HIP(fdb::launch_advect_pass(t->f_dtype, t->g, pair, &cur, 1, s, nullptr,
                                   (uint64_t)t->wstep, 0));
  t->w = dst;
  t->wstep += 1;
  t->fwd_steps += 1;
  return FDB_OK;
}

// A scratch buffer that is not the working state.
void* free_scratch(fdb_tape* t) { return t->w == t->scratch[0] ? t->scratch[1] : t->scratch[0]; }

// Steps of the next launch over `left` remaining steps (1 = a one-step
// launch): the least-time split of the library's pass planner, at most
// `cap` steps per pass.
int pass_steps(int dtype, const fdb::Grid3& g, int left, int cap) {
  return fdb::plan_pass(dtype, g, left, cap);
}

// Advance the working state by a steps; the last one lands in `last_dst`
// when given (the slot of a snapshot about to be stored).  The states in
// between are not kept, so runs of steps go through fused multi-step
// launches where the storage type and shape allow (same bits as single
// steps); the working state's buffer is only read.
int advance(fdb_tape* t, int a, void* last_dst, hipStream_t s) {
  int left = a;
  while (left > 0) {
    int done = 0;
    const int k = pass_steps(t->f_dtype, t->g, left, t->time_block);
    if (k > 1) {
      void* dst = (k == left && last_dst) ? last_dst : free_scratch(t);
      // Stochastic rounding is keyed by the absolute step (recomputed
      // states get the bits the forward drew).
      const hipError_t e = fdb::launch_fused_pass(t->f_dtype, t->g, t->w, dst, k, s, nullptr,
                                                  (uint64_t)t->wstep, 0);
      if (e != hipErrorNotSupported) {
        TAPE_HIP(e);
        t->w = dst;
        t->wstep += k;
        t->fwd_steps += k;
        done = k;
      }
    }
    if (done == 0) {
      void* dst = (left == 1 && last_dst) ? last_dst : free_scratch(t);
      if (int rc = step_into(t, dst, s)) return rc;
      done = 1;
    }
    left -= done;
  }
  return FDB_OK;
}

const void* top_buffer(const fdb_tape* t) {
  return t->stack_steps.size() == 1 ? t->base : t->slots[t->stack_steps.size() - 2];
}

// The slot the next STORE fills.
void* next_slot(fdb_tape* t) { return t->slots[t->stack_steps.size() - 1]; }

// Buffer of stack entry idx (entry 0 is the initial state).
const void* stack_buffer(const fdb_tape* t, size_t idx) {
  return idx == 0 ? t->base : t->slots[idx - 1];
}

// Steps in the fused transpose pass.
constexpr int kFusedAdjointSteps = 4;

// Length of the fused adjoint run starting at acts[i] = ADJOINT t: the
// schedule continues with (POP, RESTORE, ADJOINT t-1), (POP, RESTORE,
// ADJOINT t-2), ... and states t-1, t-2, ... are the stored states just
// below the stack top, so one pass can read them all.  Returns
// kFusedAdjointSteps when that many steps qualify (and the storage types
// have a fused form), else 1.
int fused_adjoint_run(const fdb_tape* t, size_t i) {
  constexpr int k = kFusedAdjointSteps;
  if (!t->fuse_adj) return 1;
  const auto& acts = t->plan.acts;
  const int step = acts[i].second;
  const size_t n = t->stack_steps.size();
  if (n < (size_t)k || t->stack_steps[n - 1] != step) return 1;
  for (int q = 1; q < k; ++q) {
    const size_t j = i + 3 * (size_t)q;
    if (j >= acts.size() || acts[j - 2].first != FDB_CK_POP ||
        acts[j - 1].first != FDB_CK_RESTORE || acts[j].first != FDB_CK_ADJOINT ||
        acts[j].second != step - q || t->stack_steps[n - 1 - q] != step - q)
      return 1;
  }
  return k;
}

// Executes plan actions from index i; in forward mode stops at the first
// RESTORE or ADJOINT and returns its index through *stop.
int run_plan(fdb_tape* t, size_t i, bool forward_only, hipStream_t s, const void** lam_next,
             void* lam0, int* done, size_t* stop) {
  const auto& acts = t->plan.acts;
  for (; i < acts.size(); ++i) {
    const auto& a = acts[i];
    switch (a.first) {
      case FDB_CK_ADVANCE: {
        const bool store_next = i + 1 < acts.size() && acts[i + 1].first == FDB_CK_STORE;
        // A run of (ADVANCE 1, STORE) pairs -- the schedule keeping every
        // state -- goes through fused passes that write each intermediate
        // state into its snapshot slot.
        if (store_next && a.second == 1) {
          size_t run = 0;
          while (i + 2 * run + 1 < acts.size() && acts[i + 2 * run].first == FDB_CK_ADVANCE &&
                 acts[i + 2 * run].second == 1 && acts[i + 2 * run + 1].first == FDB_CK_STORE)
            ++run;
          const int k = (run >= 8 && t->time_block >= 8) ? 8
                        : (run >= 4 && t->time_block >= 4) ? 4 : 0;
          const size_t n = t->stack_steps.size();
          if (k > 0 && n - 1 + (size_t)k <= t->slots.size()) {
            void* lv[fdb::kMaxTimeBlock];
            for (int q = 0; q < k; ++q) lv[q] = t->slots[n - 1 + q];
            const hipError_t e =
                fdb::launch_fused_pass(t->f_dtype, t->g, t->w, lv[k - 1], k, s, lv);
            if (e == hipSuccess) {
              for (int q = 0; q < k; ++q) t->stack_steps.push_back(t->wstep + q + 1);
              t->w = lv[k - 1];
              t->wstep += k;
              t->fwd_steps += k;
              i += 2 * (size_t)k - 1;  // the last STORE of the run
              break;
            }
            if (e != hipErrorNotSupported) TAPE_HIP(e);
          }
        }
        if (int rc = advance(t, a.second, store_next ? next_slot(t) : nullptr, s)) return rc;
        break;
      }
      case FDB_CK_STORE:
        if (t->w != next_slot(t)) {  // the state was not produced in place
          TAPE_HIP(hipMemcpyAsync(next_slot(t), t->w, t->fbytes, hipMemcpyDeviceToDevice, s));
          t->w = next_slot(t);
        }
        t->stack_steps.push_back(t->wstep);
        break;
      case FDB_CK_POP:
        t->stack_steps.pop_back();
        break;
      case FDB_CK_RESTORE: {
        if (forward_only) {
          *stop = i;
          return FDB_OK;
        }
        // Skip the restore and the advance after it when the working
        // state is already the state they would produce.
        const int top = t->stack_steps.back();
        const bool adv = i + 1 < acts.size() && acts[i + 1].first == FDB_CK_ADVANCE;
        const int target = top + (adv ? acts[i + 1].second : 0);
        if (t->wstep == target) {
          if (adv) ++i;
          break;
        }
        t->w = top_buffer(t);
        t->wstep = top;
        break;
      }
      case FDB_CK_ADJOINT: {
        if (forward_only) {
          *stop = i;
          return FDB_OK;
        }
        if (t->wstep != a.second)
          return tfail(FDB_ERR_INVALID, "tape_backward: schedule out of step (internal)");
        // lam ping-pongs between lam0 and lam_tmp, aiming for the last step
        // to land in lam0 (fdb_tape_backward copies it there otherwise).
        const int k = fused_adjoint_run(t, i);
        *done += k;
        void* dst = ((t->nsteps - *done) % 2 == 0) ? lam0 : t->lam_tmp;
        if (dst == *lam_next) dst = (dst == lam0) ? t->lam_tmp : lam0;
        if (k > 1) {
          // One pass over steps t, t-1, ..., t-k+1 (states from the stack),
          // standing for the k ADJOINTs and the POP/RESTOREs between them.
          const void* fs[kFusedAdjointSteps];
          const size_t n = t->stack_steps.size();
          fs[0] = t->w;
          for (int q = 1; q < k; ++q) fs[q] = stack_buffer(t, n - 1 - q);
          const hipError_t e = fdb::launch_adjoint_fused(
              t->lam_dtype, t->f_dtype, t->g, *lam_next, dst, fs, k, t->partials,
              (int64_t)(t->partials_bytes / (6 * sizeof(double))), t->acc, s);
          if (e == hipSuccess) {
            for (int q = 1; q < k; ++q) t->stack_steps.pop_back();
            t->w = fs[k - 1];
            t->wstep = a.second - (k - 1);
            i += 3 * (size_t)(k - 1);  // the last ADJOINT of the run
            *lam_next = dst;
            t->adj_steps += k;
            break;
          }
          if (e != hipErrorNotSupported) TAPE_HIP(e);
          *done -= k - 1;  // not fused after all: this step alone
          dst = ((t->nsteps - *done) % 2 == 0) ? lam0 : t->lam_tmp;
          if (dst == *lam_next) dst = (dst == lam0) ? t->lam_tmp : lam0;
        }
        TAPE_HIP(fdb::launch_adjoint_step(t->lam_dtype, t->f_dtype, t->g, *lam_next, dst, t->w,
                                          t->partials, t->acc, s));
        *lam_next = dst;
        ++t->adj_steps;
        break;
      }
      default:
        return tfail(FDB_ERR_INVALID, "tape: bad action (internal)");
    }
  }
  if (stop) *stop = acts.size();
  return FDB_OK;
}

// Grows the adjoint kernel's partial-sum scratch to nb blocks.
int ensure_partials(fdb_tape* t, int64_t nb) {
  const size_t pb = (size_t)nb * 6 * sizeof(double);
  if (pb > t->partials_bytes) {
    if (t->partials) TAPE_HIP(hipFree(t->partials));
    t->partials = nullptr;
    t->partials_bytes = 0;
    TAPE_HIP(hipMalloc((void**)&t->partials, pb));
    t->partials_bytes = pb;
  }
  return FDB_OK;
}

// The backward of FDB_GRAD_SHIFT: lam_T -> lam_1 by T-1 transpose steps
// (fused passes of the forward kernel run with the directions negated,
// same coefficients), then lam_1 -> lam_0 with the velocity products
// against f_0, scaled by T (the products of every step are equal; see the
// file comment).  Where the storage pairing has one, the last L steps
// (lam_L -> lam_0) run as one fused pass that forms the products at its
// last level (launch_adjoint_last: FP64 L = 7, FP32 L = 8), which saves
// the one-step launch's pass over lam_1, f_0 and lam_0.  lam_T is only
// read; the passes alternate between lam_tmp and lam0, arranged so that
// the last one writes lam0.
int backward_shift(fdb_tape* t, const void* lam_T, void* lam0, double* grad_v, hipStream_t s) {
  fdb::Grid3 ga = t->g;  // the transpose step: gather from the downwind side
  for (int j = 0; j < 3; ++j) ga.up[j] = -t->g.up[j];
  const int ld = t->lam_dtype;
  int64_t nbl = 0;
  const int kl = fdb::adjoint_last_depth(ld, t->f_dtype, t->g, &nbl);
  const int last = (kl > 1 && kl <= t->time_block && kl <= t->nsteps) ? kl : 1;
  // Pass lengths for the T-last plain steps: the planner's split.
  std::vector<int> ks;
  for (int left = t->nsteps - last; left > 0;) {
    const int k = pass_steps(ld, ga, left, t->time_block);
    ks.push_back(k);
    left -= k;
  }
  const void* src = lam_T;
  for (size_t i = 0; i < ks.size(); ++i) {
    // Counted from the end: the last pass writes lam_tmp, the one before lam0, ...
    void* dst = ((ks.size() - 1 - i) % 2 == 0) ? t->lam_tmp : lam0;
    if (ks[i] > 1) {
      TAPE_HIP(fdb::launch_fused_pass(ld, ga, src, dst, ks[i], s));
    } else {
      TAPE_HIP(fdb::launch_adjoint_step(ld, ld, t->g, src, dst, nullptr, nullptr, nullptr, s));
    }
    src = dst;
    t->adj_steps += ks[i];
  }
  if (last > 1) {
    if (int rc = ensure_partials(t, nbl)) return rc;
    TAPE_HIP(fdb::launch_zero(t->acc, 6, s));
    TAPE_HIP(fdb::launch_adjoint_last(ld, t->f_dtype, t->g, src, lam0, t->f0_ref, last,
                                      t->partials, nbl, t->acc, s));
  } else {
    const int64_t nb = fdb::adjoint_blocks(ld, t->f_dtype, t->g, src, lam0, t->f0_ref);
    if (int rc = ensure_partials(t, nb)) return rc;
    TAPE_HIP(fdb::launch_zero(t->acc, 6, s));
    TAPE_HIP(fdb::launch_adjoint_step(ld, t->f_dtype, t->g, src, lam0, t->f0_ref, t->partials,
                                      t->acc, s));
  }
  t->adj_steps += last;
  double acc[6];
  TAPE_HIP(hipMemcpyAsync(acc, t->acc, sizeof(acc), hipMemcpyDeviceToHost, s));
  TAPE_HIP(hipStreamSynchronize(s));
  const int off = 3 - t->ndims;
  for (int r = 0; r < t->ndims; ++r) {
    const int j = r + off;
    grad_v[r] = -t->g.dc[j] * ((double)t->nsteps * (acc[2 * j] + acc[2 * j + 1]));
  }
  return FDB_OK;
}

}  // namespace

extern "C" {

int fdb_ckpt_plan(int nsteps, int snaps, int* actions, int* args, int cap,
                  int64_t* forward_steps) {
  if (nsteps < 0 || snaps < 0 || cap < 0) return tfail(FDB_ERR_INVALID, "ckpt_plan: negative");
  Plan p = make_plan(nsteps, snaps);
  for (size_t i = 0; i < p.acts.size() && (int)i < cap; ++i) {
    if (actions) actions[i] = p.acts[i].first;
    if (args) args[i] = p.acts[i].second;
  }
  if (forward_steps) *forward_steps = p.forward;
  return (int)std::min<size_t>(p.acts.size(), (size_t)INT32_MAX);
}

int fdb_tape_create(int ndims, const uint64_t* num_cells, const double* velocity,
                    const double* lengths, double delta_time, int nsteps, int f_dtype,
                    int lam_dtype, uint64_t max_bytes, fdb_tape_t** out) {
  if (!out) return tfail(FDB_ERR_INVALID, "tape_create: out is null");
  *out = nullptr;
  fdb::Grid3 g;
  if (!fdb::make_grid(ndims, num_cells, velocity, lengths, delta_time, &g))
    return tfail(FDB_ERR_INVALID, "tape_create: invalid geometry");
  if (nsteps < 1) return tfail(FDB_ERR_INVALID, "tape_create: nsteps must be >= 1");
  if (fdb::dtype_size(f_dtype) == 0) return tfail(FDB_ERR_INVALID, "tape_create: f dtype");
  if (lam_dtype != FDB_F64 && lam_dtype != FDB_F32 && lam_dtype != FDB_BF16)
    return tfail(FDB_ERR_INVALID, "tape_create: lam dtype must be F64, F32 or BF16");
  // Every byte count below (fields, the fixed working set) must not wrap.
  if (!fdb::fits_bytes((uint64_t)g.ntot, 8 * 4))
    return tfail(FDB_ERR_INVALID, "tape_create: field size in bytes overflows");
  if (const char* m = fdb::tuning_error()) return tfail(FDB_ERR_INVALID, m);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return tfail(FDB_ERR_NODEV, "no HIP device available");
  fdb_tape* t = new fdb_tape;
  t->ndims = ndims;
  t->nsteps = nsteps;
  t->f_dtype = f_dtype;
  t->lam_dtype = lam_dtype;
  t->g = g;
  t->fbytes = (size_t)g.ntot * fdb::dtype_size(f_dtype);
  t->lbytes = (size_t)g.ntot * fdb::dtype_size(lam_dtype);
  // Fixed: the initial state, two scratch buffers, one adjoint buffer.
  const uint64_t fixed = 3 * (uint64_t)t->fbytes + t->lbytes;
  if (max_bytes < fixed) {
    tape_free(t);
    return tfail(FDB_ERR_NOMEM, "tape_create: max_bytes below the fixed working set (" +
                                    std::to_string(fixed) + " bytes)");
  }
  t->snaps = (int)std::min<uint64_t>((max_bytes - fixed) / t->fbytes, (uint64_t)(nsteps - 1));
  t->plan = make_plan(nsteps, t->snaps);
  auto fail_free = [&](hipError_t e, const char* what) {
    (void)hipGetLastError();
    tape_free(t);
    return tfail(e == hipErrorOutOfMemory ? FDB_ERR_NOMEM : FDB_ERR_HIP,
                 std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&t->base, t->fbytes)) != hipSuccess) return fail_free(e, "hipMalloc base");
  for (void*& w : t->scratch)
    if ((e = hipMalloc(&w, t->fbytes)) != hipSuccess) return fail_free(e, "hipMalloc scratch");
  if ((e = hipMalloc(&t->lam_tmp, t->lbytes)) != hipSuccess) return fail_free(e, "hipMalloc lam");
  t->slots.assign(t->snaps, nullptr);
  for (void*& sl : t->slots)
    if ((e = hipMalloc(&sl, t->fbytes)) != hipSuccess) return fail_free(e, "hipMalloc snapshot");
  t->slots_ready = true;
  if ((e = hipMalloc((void**)&t->acc, 6 * sizeof(double))) != hipSuccess)
    return fail_free(e, "hipMalloc acc");
  *out = t;
  return FDB_OK;
}

int fdb_tape_forward(fdb_tape_t* t, const void* f0, void* fT, void* stream) {
  if (!t || !f0 || !fT) return tfail(FDB_ERR_INVALID, "tape_forward: null argument");
  if (f0 == fT) return tfail(FDB_ERR_INVALID, "tape_forward: f0 and fT must differ");
  hipStream_t s = (hipStream_t)stream;
  t->fwd_steps = 0;
  t->adj_steps = 0;
  if (t->method == FDB_GRAD_SHIFT) {
    // Plain forward from the caller's f0 (only read), kept for the products.
    t->w = f0;
    t->wstep = 0;
    if (int rc = advance(t, t->nsteps, fT, s)) return rc;
    t->f0_ref = f0;
    TAPE_HIP(hipStreamSynchronize(s));
    t->forward_done = true;
    return FDB_OK;
  }
  if (!t->slots_ready) {  // released by an earlier switch to FDB_GRAD_SHIFT
    for (void*& sl : t->slots) TAPE_HIP(hipMalloc(&sl, t->fbytes));
    t->slots_ready = true;
  }
  TAPE_HIP(hipMemcpyAsync(t->base, f0, t->fbytes, hipMemcpyDeviceToDevice, s));
  t->w = t->base;
  t->wstep = 0;
  t->stack_steps.assign(1, 0);
  // The first descent of the plan (advances and stores), then on to state
  // nsteps-1 (the backward's first adjoint needs it) and one step into fT.
  size_t stop = 0;
  if (int rc = run_plan(t, 0, true, s, nullptr, nullptr, nullptr, &stop)) return rc;
  t->resume = stop;
  if (int rc = advance(t, t->nsteps - 1 - t->wstep, nullptr, s)) return rc;
  const void* keep = t->w;
  const int keep_step = t->wstep;
  if (int rc = step_into(t, fT, s)) return rc;
  t->w = keep;  // fT is the caller's; the working state stays at nsteps-1
  t->wstep = keep_step;
  TAPE_HIP(hipStreamSynchronize(s));
  t->forward_done = true;
  return FDB_OK;
}

int fdb_tape_backward(fdb_tape_t* t, const void* lam_T, void* lam0, double* grad_v,
                      void* stream) {
  if (!t || !lam_T || !lam0 || !grad_v)
    return tfail(FDB_ERR_INVALID, "tape_backward: null argument");
  if (lam_T == lam0) return tfail(FDB_ERR_INVALID, "tape_backward: lam_T and lam0 must differ");
  if (!t->forward_done) return tfail(FDB_ERR_INVALID, "tape_backward: run tape_forward first");
  hipStream_t s = (hipStream_t)stream;
  if (t->method == FDB_GRAD_SHIFT) return backward_shift(t, lam_T, lam0, grad_v, s);
  // Partial-sum scratch of the adjoint kernel, sized for every buffer
  // pairing the sweep uses (f_t is always a tape buffer).
  int64_t nb = 0;
  const void* pairs[4][2] = {{lam_T, lam0}, {lam_T, t->lam_tmp}, {t->lam_tmp, lam0},
                             {lam0, t->lam_tmp}};
  for (auto& p : pairs)
    nb = std::max(nb, fdb::adjoint_blocks(t->lam_dtype, t->f_dtype, t->g, p[0],
                                          const_cast<void*>(p[1]), t->base));
  t->fuse_adj = false;
  if (t->time_block >= kFusedAdjointSteps) {
    const int64_t fb =
        fdb::adjoint_fused_blocks(t->lam_dtype, t->f_dtype, t->g, kFusedAdjointSteps);
    t->fuse_adj = fb > 0;
    nb = std::max(nb, fb);
  }
  if (int rc = ensure_partials(t, nb)) return rc;
  TAPE_HIP(fdb::launch_zero(t->acc, 6, s));
  const void* lam_next = lam_T;
  int done = 0;
  if (int rc = run_plan(t, t->resume, false, s, &lam_next, lam0, &done, nullptr)) return rc;
  if (done != t->nsteps) return tfail(FDB_ERR_INVALID, "tape_backward: incomplete (internal)");
  if (lam_next != lam0)
    TAPE_HIP(hipMemcpyAsync(lam0, lam_next, t->lbytes, hipMemcpyDeviceToDevice, s));
  double acc[6];
  TAPE_HIP(hipMemcpyAsync(acc, t->acc, sizeof(acc), hipMemcpyDeviceToHost, s));
  TAPE_HIP(hipStreamSynchronize(s));
  const int off = 3 - t->ndims;
  for (int r = 0; r < t->ndims; ++r) {
    const int j = r + off;
    grad_v[r] = -t->g.dc[j] * (acc[2 * j] + acc[2 * j + 1]);
  }
  // The sweep consumed the stored states: another backward needs another
  // forward.
  t->forward_done = false;
  return FDB_OK;
}

int fdb_tape_set_time_block(fdb_tape_t* t, int k) {
  if (!t || k < 1 || k > fdb::kMaxTimeBlock)
    return tfail(FDB_ERR_INVALID, "tape_set_time_block: k must be 1..8");
  t->time_block = k;
  return FDB_OK;
}

int fdb_tape_set_method(fdb_tape_t* t, int method) {
  if (!t || (method != FDB_GRAD_CHECKPOINT && method != FDB_GRAD_SHIFT))
    return tfail(FDB_ERR_INVALID, "tape_set_method: unknown method");
  if (method != t->method) t->forward_done = false;  // a backward needs this method's forward
  t->method = method;
  if (method == FDB_GRAD_SHIFT && t->slots_ready) {
    // No snapshots in this method: give their memory back (a later
    // checkpointed forward allocates them again).
    TAPE_HIP(hipDeviceSynchronize());
    for (void*& sl : t->slots) {
      if (sl) TAPE_HIP(hipFree(sl));
      sl = nullptr;
    }
    t->slots_ready = false;
  }
  return FDB_OK;
}

int fdb_tape_stats(const fdb_tape_t* t, int* snaps, int64_t* forward_steps,
                   int64_t* adjoint_steps) {
  if (!t) return tfail(FDB_ERR_INVALID, "tape_stats: null tape");
  if (snaps) *snaps = t->method == FDB_GRAD_SHIFT ? 0 : t->snaps;
  if (forward_steps) *forward_steps = t->fwd_steps;
  if (adjoint_steps) *adjoint_steps = t->adj_steps;
  return FDB_OK;
}

int fdb_tape_destroy(fdb_tape_t* t) {
  tape_free(t);
  return FDB_OK;
}

}  // extern "C"
