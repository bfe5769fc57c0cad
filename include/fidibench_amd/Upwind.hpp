// Upwind.hpp -- C++ drop-in for pletzer/fidibench upwind/cxx/upwind.cxx's
// Upwind<NDIMS> class (upwind.cxx:19-135), backed by libfdb_upwind.so.
//
// Same constructor and members as the reference:
//   Upwind(velocity, lengths, numCells)        upwind.cxx:23-49
//   void advect(int numTimeSteps, double dt)   upwind.cxx:51-86
//   void advect(double dt)  (one step)         upwindAcc.cxx:49
//   void saveVTK(const std::string&)           saveVTK.h:1-60
//   double checksum() const                    upwind.cxx:91-93
//   double std() const                         upwind.cxx:95-103
//   void print() const                         upwind.cxx:105-109
// The field lives on the GPU; FP64 results are bit-identical to the
// reference.  Failures throw std::runtime_error with the library's
// message (the reference reports none).  Extensions: a storage dtype,
// setField/getField, mass(), setTimeBlock().
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fdb_upwind.h"

namespace fidibench {

namespace detail {
inline void check(int rc, const char* what) {
  if (rc != FDB_OK) throw std::runtime_error(std::string(what) + ": " + fdb_last_error());
}
}  // namespace detail

template <size_t NDIMS>
class Upwind {
  static_assert(NDIMS >= 1 && NDIMS <= 3, "Upwind supports 1, 2 or 3 dimensions");

 public:
  Upwind(const std::vector<double>& velocity, const std::vector<double>& lengths,
         const std::vector<size_t>& numCells, int dtype = FDB_F64) {
    init(velocity, lengths, std::vector<uint64_t>(numCells.begin(), numCells.end()), dtype);
  }
  // The int-sized numCells of the OpenACC variants (upwindAcc.cxx:17-19).
  Upwind(const std::vector<double>& velocity, const std::vector<double>& lengths,
         const std::vector<int>& numCells, int dtype = FDB_F64) {
    std::vector<uint64_t> n;
    for (int c : numCells) {
      if (c <= 0) throw std::runtime_error("Upwind: numCells must be positive");
      n.push_back((uint64_t)c);
    }
    init(velocity, lengths, n, dtype);
  }
  ~Upwind() {
    if (h_) fdb_upwind_destroy(h_);
  }
  Upwind(const Upwind&) = delete;
  Upwind& operator=(const Upwind&) = delete;
  Upwind(Upwind&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  Upwind& operator=(Upwind&& o) noexcept {
    if (this != &o) {
      if (h_) fdb_upwind_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
  }

  void advect(int numTimeSteps, double deltaTime) {
    detail::check(fdb_upwind_advect(h_, numTimeSteps, deltaTime), "advect");
  }
  void advect(double deltaTime) { advect(1, deltaTime); }

  void saveVTK(const std::string& filename) {
    detail::check(fdb_upwind_save_vtk(h_, filename.c_str()), "saveVTK");
  }
  // Extensions: binary dump and resume (fdb_upwind_save_raw / load_raw).
  void saveRaw(const std::string& filename) {
    detail::check(fdb_upwind_save_raw(h_, filename.c_str()), "saveRaw");
  }
  void loadRaw(const std::string& filename) {
    detail::check(fdb_upwind_load_raw(h_, filename.c_str()), "loadRaw");
  }

  double checksum() const {
    double r = 0.0;
    detail::check(fdb_upwind_checksum(h_, &r), "checksum");
    return r;
  }

  double std() const {
    double r = 0.0;
    detail::check(fdb_upwind_std(h_, &r), "std");
    return r;
  }

  void print() const { detail::check(fdb_upwind_print(h_), "print"); }

  // ---- extensions
  size_t size() const {
    uint64_t n = 0;
    detail::check(fdb_upwind_ntot(h_, &n), "ntot");
    return (size_t)n;
  }
  void setField(const std::vector<double>& f) {
    if (f.size() != size()) throw std::runtime_error("setField: size mismatch");
    detail::check(fdb_upwind_set_field(h_, f.data(), FDB_F64), "setField");
  }
  std::vector<double> getField() const {
    std::vector<double> f(size());
    detail::check(fdb_upwind_get_field(h_, f.data(), FDB_F64), "getField");
    return f;
  }
  double mass() const {
    double r = 0.0;
    detail::check(fdb_upwind_mass(h_, &r), "mass");
    return r;
  }
  // Steps fused per pass: 1..8, or FDB_TIME_BLOCK_AUTO (0) for the
  // library's least-time split of each advect; same results for every k.
  void setTimeBlock(int k) { detail::check(fdb_upwind_set_time_block(h_, k), "setTimeBlock"); }
  fdb_upwind_t* handle() const { return h_; }

 private:
  void init(const std::vector<double>& v, const std::vector<double>& L,
            const std::vector<uint64_t>& n, int dtype) {
    if (v.size() < NDIMS || L.size() < NDIMS || n.size() < NDIMS)
      throw std::runtime_error("Upwind: velocity, lengths and numCells need NDIMS entries");
    detail::check(fdb_upwind_create((int)NDIMS, v.data(), L.data(), n.data(), dtype, &h_),
                  "Upwind");
  }
  fdb_upwind_t* h_ = nullptr;
};

}  // namespace fidibench
