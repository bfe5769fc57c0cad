// ref_harness.cpp -- drives the reference's own Upwind<NDIMS> class.
//
// TEST INFRASTRUCTURE ONLY (built into oracle/_ref/, never shipped).
// This file contains no reference code: it #includes upwind/cxx/upwind.cxx
// from the reference checkout named on the compiler line (-I), with the
// class's private members opened and its main() renamed, so that any
// initial field can be set and the reference's advect(), checksum(), std()
// and saveVTK() run unmodified (upwind.cxx:19-135, saveVTK.h:1-60).
// The reference main (upwind.cxx:139-217) stays callable for CLI goldens.
#include <vector>
#include <array>
#include <functional>
#include <algorithm>
#include <numeric>
#include <string>
#include <fstream>
#include <limits>
#include <iostream>
#include <sstream>
#include <cstring>
#include <cmath>
#include <cstdint>
#include <omp.h>
#include "CmdLineArgParser.h"

#define private public
#define main fidibench_reference_main
#include "upwind.cxx"
#undef main
#undef private

namespace {

template <size_t ND>
int run(const uint64_t* n, const double* v, const double* len, double dt, int nsteps,
        const double* f_in, double* f_out, double* checksum, double* stddev,
        const char* vtk_path) {
  std::vector<double> vel(v, v + ND), lengths(len, len + ND);
  std::vector<size_t> cells(n, n + ND);
  Upwind<ND> up(vel, lengths, cells);
  if (f_in) std::copy(f_in, f_in + up.ntot, up.f.begin());
  up.advect(nsteps, dt);
  if (f_out) std::copy(up.f.begin(), up.f.end(), f_out);
  if (checksum) *checksum = up.checksum();
  if (stddev) *stddev = up.std();
  if (vtk_path) up.saveVTK(vtk_path);
  return 0;
}

template <size_t ND>
int run_print(const uint64_t* n, const double* v, const double* len, double dt, int nsteps,
              const double* f_in) {
  std::vector<double> vel(v, v + ND), lengths(len, len + ND);
  std::vector<size_t> cells(n, n + ND);
  Upwind<ND> up(vel, lengths, cells);
  if (f_in) std::copy(f_in, f_in + up.ntot, up.f.begin());
  up.advect(nsteps, dt);
  up.print();
  std::cout.flush();
  return 0;
}

}  // namespace

extern "C" {

// Construct, optionally replace the field, advect, then the reference's
// own print() (upwind.cxx:105-109) on stdout.
int ref_upwind_print(int nd, const uint64_t* n, const double* v, const double* len, double dt,
                     int nsteps, const double* f_in) {
  switch (nd) {
    case 1: return run_print<1>(n, v, len, dt, nsteps, f_in);
    case 2: return run_print<2>(n, v, len, dt, nsteps, f_in);
    case 3: return run_print<3>(n, v, len, dt, nsteps, f_in);
    default: return -1;
  }
}

// Construct Upwind<nd>(v, len, n); if f_in is non-null, replace the
// ctor's f[0]=1 field with it; advect(nsteps, dt); copy the field out and
// report checksum() and std(); optionally saveVTK(vtk_path).
int ref_upwind_run(int nd, const uint64_t* n, const double* v, const double* len, double dt,
                   int nsteps, const double* f_in, double* f_out, double* checksum,
                   double* stddev, const char* vtk_path) {
  switch (nd) {
    case 1: return run<1>(n, v, len, dt, nsteps, f_in, f_out, checksum, stddev, vtk_path);
    case 2: return run<2>(n, v, len, dt, nsteps, f_in, f_out, checksum, stddev, vtk_path);
    case 3: return run<3>(n, v, len, dt, nsteps, f_in, f_out, checksum, stddev, vtk_path);
    default: return -1;
  }
}

// Time advect() alone on a 3-D field (the CPU baseline), returning seconds.
double ref_upwind_time_advect(const uint64_t* n, const double* v, const double* len, double dt,
                              int nsteps, const double* f_in, double* checksum) {
  std::vector<double> vel(v, v + 3), lengths(len, len + 3);
  std::vector<size_t> cells(n, n + 3);
  Upwind<3> up(vel, lengths, cells);
  if (f_in) std::copy(f_in, f_in + up.ntot, up.f.begin());
  double t0 = omp_get_wtime();
  up.advect(nsteps, dt);
  double t1 = omp_get_wtime();
  if (checksum) *checksum = up.checksum();
  return t1 - t0;
}

// The reference executable's main (upwindCxx), for stdout goldens.
int ref_upwind_main(int argc, char** argv) {
  int rc = fidibench_reference_main(argc, argv);
  std::cout.flush();
  return rc;
}

int ref_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
