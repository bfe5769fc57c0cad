// Filter.hpp -- C++ drop-in for pletzer/fidibench cxx/Filter.h's Filter
// class in its one-process form, backed by libfdb_upwind.so (fdb_filter.h).
//
// Same constructor and members as the reference:
//   Filter(globalDims, xmins, xmaxs, stencil)          Filter.cpp:11-78
//   void setInData(double (*)(const std::vector<double>&))      :131-161
//   void setInDataByIndices(double (*)(const std::vector<size_t>&)) :163-190
//   void applyFilter()                                  :192-263
//   void copyOutToIn()                                  :441-463
//   double computeCheckSum(const std::string&)          :466-485
//   void saveVTK(const std::string&)                    :488-538
//   std::vector<double> getPosition(const std::vector<size_t>&) :104-112
//   int getRank() const, int getNumProcs() const, bool isDecompValid() const
// The data live on the GPU; applyFilter is bit-identical to the
// reference's loop.  With one process the rank is 0, there is one
// process and the decomposition is always valid.  Failures throw
// std::runtime_error.  Extensions: iterate(), getData(), setInSin().
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fdb_filter.h"

namespace fidibench {

class Filter {
 public:
  Filter(const std::vector<size_t>& globalDims, const std::vector<double>& xmins,
         const std::vector<double>& xmaxs, const std::map<std::vector<int>, double>& stencil)
      : dims_(globalDims), xmins_(xmins), xmaxs_(xmaxs) {
    if (xmins.size() != globalDims.size() || xmaxs.size() != globalDims.size())
      throw std::runtime_error("Filter: globalDims, xmins and xmaxs must have one entry per axis");
    std::vector<uint64_t> d(globalDims.begin(), globalDims.end());
    std::vector<int> off;
    std::vector<double> w;
    for (const auto& kv : stencil) {
      if (kv.first.size() != globalDims.size())
        throw std::runtime_error("Filter: stencil offsets must have one entry per axis");
      off.insert(off.end(), kv.first.begin(), kv.first.end());
      w.push_back(kv.second);
    }
    check(fdb_filter_create((int)d.size(), d.data(), xmins.data(), xmaxs.data(), (int)w.size(),
                            off.data(), w.data(), &h_),
          "Filter");
  }
  ~Filter() {
    if (h_) fdb_filter_destroy(h_);
  }
  Filter(const Filter&) = delete;
  Filter& operator=(const Filter&) = delete;
  Filter(Filter&& o) noexcept
      : dims_(std::move(o.dims_)), xmins_(std::move(o.xmins_)), xmaxs_(std::move(o.xmaxs_)),
        h_(std::exchange(o.h_, nullptr)) {}

  std::vector<double> getPosition(const std::vector<size_t>& globalInds) const {
    std::vector<double> pos(dims_.size());
    for (size_t i = 0; i < dims_.size(); ++i) {
      const double delta = (xmaxs_[i] - xmins_[i]) / double(dims_[i]);
      pos[i] = xmins_[i] + (globalInds[i] + 0.5) * delta;  // cell centred
    }
    return pos;
  }

  void setInData(double (*f)(const std::vector<double>&)) {
    std::vector<double> data(ntot());
    std::vector<size_t> inds(dims_.size(), 0);
    for (size_t i = 0; i < data.size(); ++i) {
      data[i] = f(getPosition(inds));
      next(inds);
    }
    check(fdb_filter_set_in_data(h_, data.data()), "setInData");
  }

  void setInDataByIndices(double (*f)(const std::vector<size_t>&)) {
    std::vector<double> data(ntot());
    std::vector<size_t> inds(dims_.size(), 0);
    for (size_t i = 0; i < data.size(); ++i) {
      data[i] = f(inds);
      next(inds);
    }
    check(fdb_filter_set_in_data(h_, data.data()), "setInDataByIndices");
  }

  void applyFilter() { check(fdb_filter_apply(h_), "applyFilter"); }
  void copyOutToIn() { check(fdb_filter_copy_out_to_in(h_), "copyOutToIn"); }

  double computeCheckSum(const std::string& inOrOut) {
    double s = 0.0;
    check(fdb_filter_checksum(h_, inOrOut == "input" ? 0 : 1, &s), "computeCheckSum");
    return s;
  }

  void saveVTK(const std::string& filename) {
    check(fdb_filter_save_vtk(h_, filename.c_str()), "saveVTK");
  }

  int getRank() const { return 0; }
  int getNumProcs() const { return 1; }
  bool isDecompValid() const { return true; }

  // niters x (applyFilter(); copyOutToIn();) on the GPU; kernel time in ms.
  float iterate(int niters) {
    float ms = 0.f;
    check(fdb_filter_iterate_timed(h_, niters, &ms), "iterate");
    return ms;
  }
  std::vector<double> getData(const std::string& inOrOut) {
    std::vector<double> d(ntot());
    check(fdb_filter_get_data(h_, inOrOut == "input" ? 0 : 1, d.data()), "getData");
    return d;
  }
  // The input function of laplacian.cxx computed on the GPU.
  void setInSin() { check(fdb_filter_set_in_sin(h_), "setInSin"); }

 private:
  static void check(int rc, const char* what) {
    if (rc != FDB_OK) throw std::runtime_error(std::string(what) + ": " + fdb_last_error());
  }
  size_t ntot() const {
    size_t n = 1;
    for (size_t d : dims_) n *= d;
    return n;
  }
  // Column-major index order (axis 0 fastest), as the reference iterates.
  void next(std::vector<size_t>& inds) const {
    for (size_t a = 0; a < inds.size(); ++a) {
      if (++inds[a] < dims_[a]) return;
      inds[a] = 0;
    }
  }

  std::vector<size_t> dims_;
  std::vector<double> xmins_, xmaxs_;
  fdb_filter_t* h_ = nullptr;
};

}  // namespace fidibench
