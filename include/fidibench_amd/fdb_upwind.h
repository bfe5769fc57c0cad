/*
 * fdb_upwind.h -- C ABI of the MI355X upwind advection library
 * (libfdb_upwind.so).
 *
 * Drop-in boundary for pletzer/fidibench upwind/cxx/upwind.cxx, whose
 * Upwind<NDIMS> class (upwind.cxx:19-135) is the whole hot path.  The
 * reference has no library boundary: its public members are the
 * interface.  Each handle entry point below names the member it replaces.
 *
 * Conventions
 *   - Every function returns an int status: FDB_OK (0) or a negative
 *     FDB_ERR_* code; fdb_last_error() then holds a message for the
 *     calling thread.
 *   - Layout is row-major with the last axis contiguous, exactly as the
 *     reference's dimProd (upwind.cxx:39-44).
 *   - A handle owns two device buffers (ping-pong) and one HIP stream.
 *     Entry points on one handle are not re-entrant; distinct handles are
 *     independent.  Every handle call returns with the stream drained.
 *   - The *_dev entry points operate on caller-owned device pointers on a
 *     caller-supplied stream (a hipStream_t passed as void*, NULL = the
 *     null stream); they return once the work is enqueued unless stated.
 */
#ifndef FDB_UPWIND_H
#define FDB_UPWIND_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdb_upwind fdb_upwind_t;

/* Field storage types.  F64 is the reference's type and is bit-exact with
 * it; F32 and BF16 store the field in 4 and 2 bytes and compute in float.
 * BF16_SR is bf16 storage whose step results are rounded stochastically
 * (unbiased; the random bits are a hash of the seed, the step's index and
 * the index of the cell's pair -- cells pair along the second-to-last axis
 * -- so runs are reproducible).  Host conversions into
 * BF16 and BF16_SR fields round to nearest even. */
enum fdb_dtype { FDB_F64 = 0, FDB_F32 = 1, FDB_BF16 = 2, FDB_BF16_SR = 3 };

enum fdb_status {
  FDB_OK = 0,
  FDB_ERR_INVALID = -1, /* bad argument (sizes, dtype, null pointer) */
  FDB_ERR_HIP = -2,     /* a HIP runtime call failed */
  FDB_ERR_NOMEM = -3,   /* device or host allocation failed */
  FDB_ERR_IO = -4,      /* file could not be written */
  FDB_ERR_NODEV = -5    /* no usable GPU */
};

/* Synthetic initial fields for fdb_upwind_init_field / fdb_init_field_dev
 * (cell centres x_j = (i_j + 0.5) * L_j / N_j). */
enum fdb_init_kind {
  FDB_INIT_DELTA = 0,   /* f[0] = 1, else 0 (the reference ctor, upwind.cxx:45-48) */
  FDB_INIT_SPHERE = 1,  /* params {cx,cy,cz,r,amp}: (|x-c|^2 < r^2) + amp*U(seed,i) */
  FDB_INIT_GAUSS = 2,   /* params {cx,cy,cz,w,amp}: amp*exp(-|x-c|^2 / w) */
  FDB_INIT_GAUSS_SUM = 3,/* params {m, (cx,cy,cz,w,a) x m}: sum_k a*exp(-|x-c_k|^2 / w_k^2) */
  FDB_INIT_UNIFORM = 4  /* params {lo,hi}: lo + (hi-lo)*U(seed,i) */
};

/* Message for the last failing call on this thread ("" if none). */
const char* fdb_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int fdb_version(void);

/* ------------------------------------------------------------ handles --- */

/* Upwind(velocity, lengths, numCells), upwind.cxx:23-49.  ndims is NDIMS
 * (1..3); the field starts as the reference's: zero with f[0] = 1.
 * dtype selects the storage type (FDB_F64 reproduces the reference). */
int fdb_upwind_create(int ndims, const double* velocity, const double* lengths,
                      const uint64_t* num_cells, int dtype, fdb_upwind_t** out);
/* Releases the handle (the reference object's destructor). */
int fdb_upwind_destroy(fdb_upwind_t* h);

/* advect(int numTimeSteps, double deltaTime), upwind.cxx:51-86.  Also
 * the one-step advect(double) of upwindAcc.cxx:49 with num_time_steps=1. */
int fdb_upwind_advect(fdb_upwind_t* h, int num_time_steps, double delta_time);

/* checksum() const, upwind.cxx:91-93: the sequential left fold from 0.0
 * of the field (widened to double), computed on the host after a copy
 * back, so FP64 results match the reference bit for bit. */
int fdb_upwind_checksum(fdb_upwind_t* h, double* out);

/* std() const, upwind.cxx:95-103. */
int fdb_upwind_std(fdb_upwind_t* h, double* out);

/* print() const, upwind.cxx:105-109: "i f[i]" lines on stdout. */
int fdb_upwind_print(fdb_upwind_t* h);

/* saveVTK(filename), upwind/cxx/saveVTK.h:1-60: ASCII RECTILINEAR_GRID
 * with the same bytes as the reference writer. */
int fdb_upwind_save_vtk(fdb_upwind_t* h, const char* filename);

/* Fast binary dump and resume (extension; the reference writes VTK only):
 * a 64-byte header ("FDBFLD01", ndims, dtype, axes, steps done, seed) and
 * the field in its storage type, little-endian, last axis contiguous
 * (numpy.fromfile(name, dtype, offset=64)).  load_raw restores the field
 * and the step count/seed of the stochastic rounding into a handle of the
 * same shape and storage type, so a run can stop and resume bit-exactly;
 * FDB_ERR_INVALID for another shape or type, FDB_ERR_IO for a missing,
 * foreign or truncated file.  These checks come before any data is read
 * and leave the handle unchanged; a read error after them (a file that
 * shrinks while it is read) leaves the field partly replaced. */
int fdb_upwind_save_raw(fdb_upwind_t* h, const char* filename);
int fdb_upwind_load_raw(fdb_upwind_t* h, const char* filename);
/* The header of a raw field file, checked as load_raw checks it (magic,
 * ndims 1..3, dtype, axes, exact file size), without a device: ndims,
 * dtype, num_cells[0..ndims-1], steps done and seed (any may be NULL).
 * FDB_ERR_IO for a missing, foreign, malformed or truncated file. */
int fdb_host_raw_info(const char* filename, int* ndims, int* dtype, uint64_t* num_cells,
                      uint64_t* steps_done, uint64_t* seed);

/* -- extensions of the handle (no member in the reference) -- */

/* Number of cells (the reference's private ntot, upwind.cxx:120). */
int fdb_upwind_ntot(const fdb_upwind_t* h, uint64_t* out);
/* Storage dtype of the handle. */
int fdb_upwind_dtype(const fdb_upwind_t* h, int* out);
/* Replace the field from a host buffer of ntot values of host_dtype
 * (converted to the storage type; bf16 rounds to nearest even). */
int fdb_upwind_set_field(fdb_upwind_t* h, const void* host, int host_dtype);
/* Copy the field to a host buffer of ntot values of host_dtype. */
int fdb_upwind_get_field(fdb_upwind_t* h, void* host, int host_dtype);
/* Device pointer of the current field (valid until the next advect). */
int fdb_upwind_field_device(fdb_upwind_t* h, void** out);
/* Fill the field on the device with a synthetic initial condition. */
int fdb_upwind_init_field(fdb_upwind_t* h, int kind, const double* params, int nparams,
                          uint64_t seed);
/* Sum of the field by an FP64 compensated (Neumaier + TwoSum) reduction
 * on the device: the mass used for drift reports. */
int fdb_upwind_mass(fdb_upwind_t* h, double* out);
/* Steps fused per HBM pass in advect (temporal blocking): 1..8 = passes of
 * at most k steps (1 = one kernel launch per step, the default);
 * FDB_TIME_BLOCK_AUTO (0) = the library's least-time split of each advect
 * into one-step launches and fused passes (fdb_plan_passes lists it).
 * Results are identical for every k.  Extension (no counterpart in the
 * reference). */
#define FDB_TIME_BLOCK_AUTO 0
int fdb_upwind_set_time_block(fdb_upwind_t* h, int k);
/* Seed of the stochastic rounding of an FDB_BF16_SR handle (default 0);
 * also restarts the handle's step count, which indexes the random bits. */
int fdb_upwind_set_seed(fdb_upwind_t* h, uint64_t seed);
/* advect(), returning the device time of its kernels (ms, HIP events on
 * the handle's stream) and the number of kernel launches. */
int fdb_upwind_advect_timed(fdb_upwind_t* h, int num_time_steps, double delta_time,
                            float* ms, int* launches);

/* ------------------------------------------------ device-pointer entry -- */

/* nsteps of advect on a caller-owned field: steps alternate between the
 * buffers a and b (same size, both device memory), starting from a.
 * *result_in_b is set to 1 if the final state is in b, else 0.  The
 * arguments mean what they mean in fdb_upwind_create/advect;
 * time_block as in fdb_upwind_set_time_block (0..8). */
int fdb_advect_dev(int ndims, const uint64_t* num_cells, const double* velocity,
                   const double* lengths, double delta_time, int dtype, void* a, void* b,
                   int nsteps, int time_block, void* stream, int* result_in_b);

/* The pass depths (steps per launch, in launch order) that an advect of
 * nsteps steps with this time_block (0..8, 0 = FDB_TIME_BLOCK_AUTO) makes
 * for this geometry and dtype, as fdb_upwind_advect and fdb_advect_dev
 * run them: *count gets the number of launches, ks[0 .. min(count,
 * max_ks)-1] their depths.  Launches and allocates nothing; on grids of at
 * least 512^3 cells the split reads the kernels' occupancy on the current
 * device (with no device: one block per CU on 256 CUs).  Extension (no
 * counterpart in the reference). */
int fdb_plan_passes(int ndims, const uint64_t* num_cells, const double* velocity,
                    const double* lengths, double delta_time, int dtype, int nsteps,
                    int time_block, int* ks, int max_ks, int* count);

/* fdb_advect_dev with the stochastic-rounding stream made explicit: for
 * FDB_BF16_SR the random bits of the i-th step of this call are keyed by
 * (seed, step0 + i), as a handle seeded `seed` keys its step step0 + i, so
 * calls with step0 = 0, n, 2n, ... continue one run (fdb_advect_dev is
 * seed 0, step0 0).  Other dtypes ignore both.  Extension (no counterpart
 * in the reference). */
int fdb_advect_dev_seeded(int ndims, const uint64_t* num_cells, const double* velocity,
                          const double* lengths, double delta_time, int dtype, void* a, void* b,
                          int nsteps, int time_block, uint64_t seed, uint64_t step0,
                          void* stream, int* result_in_b);

/* One adjoint step (the exact transpose of one forward step, gathering
 * from the downwind side with the same coefficients): lam = A^T lam_next.
 * If f_t is non-null, also accumulates into vgrad_acc (device, 6 doubles:
 * three (sum, compensation) pairs) the velocity-gradient terms
 *   sum_i lam_next[i] * (f_t[up_j(i)] - f_t[i]) * (-(dt*u_j/dx_j)).
 * lam dtype is FDB_F64, FDB_F32 or FDB_BF16 (FP32 arithmetic, each step
 * rounded to nearest even); f_t has dtype f_dtype. */
int fdb_adjoint_step_dev(int ndims, const uint64_t* num_cells, const double* velocity,
                         const double* lengths, double delta_time, int lam_dtype,
                         const void* lam_next, void* lam, int f_dtype, const void* f_t,
                         double* vgrad_acc, void* stream);

/* fdb_adjoint_step_dev without any device allocation, for callers that
 * cannot allocate: code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc.  No counterpart in the reference.  Same arguments
 * and results (lam and the vgrad_acc sums bit-identical); the velocity
 * products' per-block partial sums live in a fixed device array of 65536
 * slots (3 MiB, held by each GPU the library runs on until the process
 * ends) instead of a hipMalloc'd buffer, so a grid needing more slots is
 * FDB_ERR_INVALID (512^3 needs at most a few thousand).  Returns once the
 * work is enqueued.  All calls on one GPU share the array: calls ordered on
 * one stream are safe, calls that overlap on one GPU race. */
int fdb_adjoint_step_dev_noalloc(int ndims, const uint64_t* num_cells, const double* velocity,
                                 const double* lengths, double delta_time, int lam_dtype,
                                 const void* lam_next, void* lam, int f_dtype, const void* f_t,
                                 double* vgrad_acc, void* stream);

/* FP64 compensated sum of n values (dtype) at x; blocks until done. */
int fdb_sum_dev(int dtype, const void* x, uint64_t n, double* out, void* stream);

/* Misfit: lam = f - target (lam_dtype), loss = 0.5*sum((f-target)^2) in
 * compensated FP64.  f and target have dtype f_dtype.  Blocks. */
int fdb_misfit_dev(int f_dtype, const void* f, const void* target, uint64_t n, int lam_dtype,
                   void* lam, double* loss, void* stream);

/* Reverse-mode tape for the gradient of a scalar loss of f_T, where f_T
 * is nsteps forward steps from f0 with a fixed delta_time.
 * fdb_tape_forward computes f_T and keeps checkpoints; fdb_tape_backward
 * takes lam_T = dL/df_T and returns lam_0 = dL/df0 and dL/dv (host,
 * ndims doubles), recomputing the states it did not keep.  Checkpoints
 * follow the optimal binomial schedule (Griewank's revolve) for the
 * number of field copies that fit in max_bytes of device memory.
 * f_dtype is the forward storage type, lam_dtype (FDB_F64, FDB_F32 or
 * FDB_BF16) the adjoint's.  All device buffers are caller-owned; both
 * calls block. */
typedef struct fdb_tape fdb_tape_t;
int fdb_tape_create(int ndims, const uint64_t* num_cells, const double* velocity,
                    const double* lengths, double delta_time, int nsteps, int f_dtype,
                    int lam_dtype, uint64_t max_bytes, fdb_tape_t** out);
int fdb_tape_forward(fdb_tape_t* t, const void* f0, void* fT, void* stream);
int fdb_tape_backward(fdb_tape_t* t, const void* lam_T, void* lam0, double* grad_v,
                      void* stream);
/* Steps per launch, 1..8 (default 8): runs of forward steps between stored
 * states go through fused multi-step passes, and runs of adjoint steps
 * whose states are all stored (lam/f storage FP64/FP64, FP32/FP32 or FP32/bf16,
 * 3-D) through a fused 4-step transpose pass.  lam_0 is bit-identical for
 * every k; dL/dv sums the same products in another order (equal to
 * rounding).  Extension (no counterpart in the reference). */
int fdb_tape_set_time_block(fdb_tape_t* t, int k);
/* How fdb_tape_backward forms the gradient (extension; the reference has
 * no gradient):
 *   FDB_GRAD_CHECKPOINT  (default) every adjoint step with its velocity
 *       products against the forward state of that step, the states
 *       recomputed from binomial checkpoints within max_bytes;
 *   FDB_GRAD_SHIFT  the step and the periodic axis shifts are circulant
 *       and commute, so the products of every step are equal: lam_0 by
 *       T transpose steps (fused passes, bit-identical to the checkpointed
 *       lam_0) and dL/dv = -(dt*u_j/dx_j) * T * <lam_1, S_j f0 - f0>
 *       (equal to the per-step sum up to rounding).  No state is stored:
 *       the tape keeps a reference to the caller's f0, which must stay
 *       unchanged until the last fdb_tape_backward; backward may be called
 *       again (another lam_T) after one forward.
 * Set before fdb_tape_forward; changing it requires a new forward. */
enum fdb_grad_method { FDB_GRAD_CHECKPOINT = 0, FDB_GRAD_SHIFT = 1 };
int fdb_tape_set_method(fdb_tape_t* t, int method);
/* Checkpoint slots used, forward steps and adjoint steps executed so far. */
int fdb_tape_stats(const fdb_tape_t* t, int* snaps, int64_t* forward_steps,
                   int64_t* adjoint_steps);
int fdb_tape_destroy(fdb_tape_t* t);

/* Error of field a against reference field b (n values each, any dtypes,
 * widened to double): out4 = {max|a-b|, sqrt(mean((a-b)^2)), max|b|,
 * max|a-b| / max|b|}.  Blocks. */
int fdb_compare_dev(int a_dtype, const void* a, int b_dtype, const void* b, uint64_t n,
                    double* out4, void* stream);

/* Fill a device field (dtype) with a synthetic initial condition. */
int fdb_init_field_dev(int ndims, const uint64_t* num_cells, const double* lengths, int dtype,
                       int kind, const double* params, int nparams, uint64_t seed, void* f,
                       void* stream);

/* ------------------------------------------------------- host helpers --- */

/* The reference's left-fold checksum and std over a host FP64 array. */
double fdb_host_checksum(const double* f, uint64_t n);
double fdb_host_std(const double* f, uint64_t n);

/* saveVTK over a host FP64 field (the bytes of saveVTK.h:1-60). */
int fdb_host_save_vtk(int ndims, const uint64_t* num_cells, const double* lengths,
                      const double* f, const char* filename);

/* The checkpoint schedule the tape runs, for tests: reversing nsteps
 * steps with `snaps` stored states besides the initial one.  Writes up
 * to cap (action, argument) pairs and returns their count (or a negative
 * status); *forward_steps receives the total forward steps, which equals
 * the binomial optimum r*n - C(s+r, s+1) with s = snaps+1 and r the
 * repetition number.  Actions:
 *   FDB_CK_ADVANCE a  advance the working state by a steps
 *   FDB_CK_STORE      push a copy of the working state
 *   FDB_CK_RESTORE    copy the top stored state into the working state
 *   FDB_CK_POP        drop the top stored state
 *   FDB_CK_ADJOINT t  adjoint of step t (working state must be state t) */
enum { FDB_CK_ADVANCE = 1, FDB_CK_STORE = 2, FDB_CK_RESTORE = 3, FDB_CK_POP = 4,
       FDB_CK_ADJOINT = 5 };
int fdb_ckpt_plan(int nsteps, int snaps, int* actions, int* args, int cap,
                  int64_t* forward_steps);

#ifdef __cplusplus
}
#endif

#endif /* FDB_UPWIND_H */
