/* fdb_filter.h -- C ABI of the generic periodic stencil filter on one
 * MI355X: the single-process form of pletzer/fidibench's Filter class
 * (cxx/Filter.h:17-228, cxx/Filter.cpp) used by laplacian/cxx/laplacian.cxx.
 *
 * Conventions are those of fdb_upwind.h (int status, fdb_last_error()).
 * Data arrays are in the reference's order: column-major, axis 0
 * contiguous (Filter's MultiArrayIter, Filter.cpp:50).  With one process
 * every stencil neighbour is local, so the MPI windows and halo fetches
 * of the reference (Filter.cpp:115-129, 321-356) have no counterpart. */
#ifndef FIDIBENCH_AMD_FDB_FILTER_H
#define FIDIBENCH_AMD_FDB_FILTER_H

#include <stdint.h>

#include "fdb_upwind.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdb_filter fdb_filter_t;

/* Filter(globalDims, xmins, xmaxs, stencil), Filter.cpp:11-78.
 * ndims in 1..3; dims, xmins, xmaxs have ndims entries; the stencil is
 * nterms (<= 32) offset vectors of ndims ints each (|offset| <= 8) and
 * their weights, in any order: they are applied in std::map order
 * (lexicographic offsets) and, as with std::map::insert, the first of
 * duplicate offsets is kept.  Input and output start at zero. */
int fdb_filter_create(int ndims, const uint64_t* dims, const double* xmins, const double* xmaxs,
                      int nterms, const int* offsets, const double* weights,
                      fdb_filter_t** out);

/* ~Filter(), Filter.cpp:80-87. */
int fdb_filter_destroy(fdb_filter_t* h);

/* setInData / setInDataByIndices (Filter.cpp:131-190) from a host array of
 * all cells (FP64, column-major). */
int fdb_filter_set_in_data(fdb_filter_t* h, const double* host);

/* setInData(func) with laplacian.cxx's func, prod_i sin(2 pi x_i) at cell
 * centres (laplacian.cxx:22-28, Filter.cpp:104-109), computed on the GPU
 * (device sin: not bit-identical to the host's). */
int fdb_filter_set_in_sin(fdb_filter_t* h);

/* applyFilter(), Filter.cpp:192-263: output = stencil applied to input. */
int fdb_filter_apply(fdb_filter_t* h);

/* copyOutToIn(), Filter.cpp:441-463 (a buffer swap; output then reads as
 * input, as in the reference after the copy). */
int fdb_filter_copy_out_to_in(fdb_filter_t* h);

/* niters x (applyFilter(); copyOutToIn();) -- the timed loop of
 * laplacian.cxx:80-86 -- with the kernels' time (HIP events) in *ms. */
int fdb_filter_iterate_timed(fdb_filter_t* h, int niters, float* ms);

/* The kernel this filter's applyFilter launches, fixed at creation:
 * "star" (full 3-D star stencils of radius 2: the 4th-order
 * Laplacian), "planes", "planes_mid", "planes_wide" (3-D LDS rings),
 * "rows" (2-D) or "gather".  Extension (no counterpart in the reference). */
int fdb_filter_kernel(const fdb_filter_t* h, const char** name);

/* computeCheckSum("input" | "output"), Filter.cpp:466-485: which = 0 for
 * input, 1 for output; a sequential left fold on the host. */
int fdb_filter_checksum(fdb_filter_t* h, int which, double* out);

/* Copies input (which = 0) or output (1) to a host array. */
int fdb_filter_get_data(fdb_filter_t* h, int which, double* host);

/* saveVTK(filename), Filter.cpp:488-538: writeVTK's STRUCTURED_GRID text
 * (cxx/writeVTK.cpp:12-95) of the output. */
int fdb_filter_save_vtk(fdb_filter_t* h, const char* filename);

/* writeVTK(filename, globalDims, xmins, xmaxs, field, name) on a host
 * field (cxx/writeVTK.cpp:12-95); no GPU needed. */
int fdb_host_write_vtk(int ndims, const uint64_t* dims, const double* xmins,
                       const double* xmaxs, const double* field, const char* name,
                       const char* filename);

#ifdef __cplusplus
}
#endif

#endif
