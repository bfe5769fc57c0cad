// san_host.cpp -- CPU test program of the library's host code under the address
// and undefined-behaviour sanitizers (test infrastructure: built by
// `make -C fidibench_amd/csrc sanitize` with the host side of every source
// instrumented, run by tests/test_host_sanitize.py; never shipped).
//
// It calls every entry point whose work is host logic -- the checkpoint
// schedule, the reference's checksum/std folds, the two VTK writers, the
// raw-file header checks on good, truncated, foreign and malformed files,
// the CLI parser, argument validation of the handle/tape/filter
// constructors (which stop before any device work) and the tuning-variable
// parser -- on small and edge-case inputs.  A sanitizer report aborts the
// run; a wrong result exits 1.
//
//   san_host            the checks above
//   san_host tuning     the same process with malformed tuning variables set
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fidibench_amd/CmdLineArgs.hpp"
#include "../../include/fidibench_amd/fdb_filter.h"
#include "../../include/fidibench_amd/fdb_upwind.h"

namespace {

int g_fail = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s (%s)\n", what, fdb_last_error());
    ++g_fail;
  }
}

std::string tmp(const char* name) {
  const char* d = std::getenv("TMPDIR");
  return std::string(d && *d ? d : "/tmp") + "/fdb_san_" + name;
}

void plans() {
  // Every (steps, snapshots) pair on a small grid, caps below the plan
  // length (the output is cut, the count still returned) and T = 400.
  for (int T = 0; T <= 40; T += (T < 10 ? 1 : 7)) {
    for (int s = 0; s <= 6; ++s) {
      int64_t fw = -1;
      const int n = fdb_ckpt_plan(T, s, nullptr, nullptr, 0, &fw);
      check(n >= 0 && fw >= 0, "ckpt_plan length");
      std::vector<int> a(n + 1), b(n + 1);
      check(fdb_ckpt_plan(T, s, a.data(), b.data(), n, &fw) == n, "ckpt_plan fill");
      check(fdb_ckpt_plan(T, s, a.data(), b.data(), n / 2, &fw) == n, "ckpt_plan cut");
    }
  }
  for (int s : {1, 2, 3, 5, 165}) {
    int64_t fw = 0;
    const int n = fdb_ckpt_plan(400, s, nullptr, nullptr, 0, &fw);
    std::vector<int> a(n), b(n);
    check(fdb_ckpt_plan(400, s, a.data(), b.data(), n, &fw) == n && fw > 0, "ckpt_plan 400");
  }
  check(fdb_ckpt_plan(-1, 0, nullptr, nullptr, 0, nullptr) == FDB_ERR_INVALID, "plan T < 0");
  check(fdb_ckpt_plan(5, -1, nullptr, nullptr, 0, nullptr) == FDB_ERR_INVALID, "plan s < 0");
}

void folds_and_vtk() {
  const double f[7] = {1.0, -0.0, 1e-310, 3.5, -2.25, 1e300, -1e300};
  check(fdb_host_checksum(f, 7) == ((((((1.0 + -0.0) + 1e-310) + 3.5) + -2.25) + 1e300) + -1e300),
        "left fold");
  check(fdb_host_checksum(f, 0) == 0.0, "empty fold");
  (void)fdb_host_std(f, 4);
  const uint64_t n3[3] = {3, 4, 5}, n2[2] = {4, 6}, n1[1] = {7};
  const double len[3] = {1.0, 1.5, 2.0}, xmin[3] = {0, 0, 0};
  std::vector<double> g(60);
  for (size_t i = 0; i < g.size(); ++i) g[i] = 0.1 * (double)i - 2.0;
  check(fdb_host_save_vtk(3, n3, len, g.data(), tmp("a.vtk").c_str()) == FDB_OK, "save_vtk 3d");
  check(fdb_host_save_vtk(2, n2, len, g.data(), tmp("b.vtk").c_str()) == FDB_OK, "save_vtk 2d");
  check(fdb_host_save_vtk(1, n1, len, g.data(), tmp("c.vtk").c_str()) == FDB_OK, "save_vtk 1d");
  check(fdb_host_save_vtk(4, n3, len, g.data(), tmp("d.vtk").c_str()) == FDB_ERR_INVALID,
        "save_vtk ndims");
  check(fdb_host_save_vtk(1, n1, len, g.data(), "/nonexistent_dir/x.vtk") == FDB_ERR_IO,
        "save_vtk io");
  check(fdb_host_write_vtk(3, n3, xmin, len, g.data(), "f", tmp("e.vtk").c_str()) == FDB_OK,
        "write_vtk 3d");
  check(fdb_host_write_vtk(1, n1, xmin, len, g.data(), "f", tmp("f.vtk").c_str()) == FDB_OK,
        "write_vtk 1d");
}

void write_file(const std::string& path, const std::vector<unsigned char>& bytes) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) {
    check(false, "write temp file");
    return;
  }
  if (!bytes.empty()) std::fwrite(bytes.data(), 1, bytes.size(), fp);
  std::fclose(fp);
}

std::vector<unsigned char> raw(uint32_t nd, uint32_t dt, uint64_t a, uint64_t b, uint64_t c,
                               size_t payload) {
  std::vector<unsigned char> v(64 + payload, 0);
  std::memcpy(v.data(), "FDBFLD01", 8);
  std::memcpy(v.data() + 8, &nd, 4);
  std::memcpy(v.data() + 12, &dt, 4);
  const uint64_t n[3] = {a, b, c};
  std::memcpy(v.data() + 16, n, 24);
  return v;
}

void raw_headers() {
  const std::string p = tmp("f.raw");
  int nd = 0, dt = 0;
  uint64_t n[3] = {0, 0, 0}, st = 0, sd = 0;
  write_file(p, raw(3, 0, 4, 5, 6, 4 * 5 * 6 * 8));
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_OK && nd == 3 && n[2] == 6,
        "raw good");
  check(fdb_host_raw_info(p.c_str(), nullptr, nullptr, nullptr, nullptr, nullptr) == FDB_OK,
        "raw good, no outputs");
  write_file(p, raw(3, 0, 4, 5, 6, 4 * 5 * 6 * 8 - 1));
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO, "raw truncated");
  write_file(p, raw(2, 1, 1u << 31, 1u << 31, 1, 0));
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO, "raw huge");
  write_file(p, raw(3, 0, ~0ull, ~0ull, ~0ull, 0));
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO, "raw overflow");
  write_file(p, raw(9, 0, 1, 1, 1, 8));
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO, "raw ndims");
  write_file(p, raw(1, 77, 1, 1, 1, 8));
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO, "raw dtype");
  write_file(p, std::vector<unsigned char>(10, 'x'));
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO, "raw foreign");
  write_file(p, {});
  check(fdb_host_raw_info(p.c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO, "raw empty");
  check(fdb_host_raw_info(tmp("none.raw").c_str(), &nd, &dt, n, &st, &sd) == FDB_ERR_IO,
        "raw missing");
  check(fdb_host_raw_info(nullptr, &nd, &dt, n, &st, &sd) == FDB_ERR_INVALID, "raw null");
}

void cli() {
  const std::vector<std::vector<std::string>> cases = {
      {}, {"-numCells", "16", "-numSteps", "3", "-std"}, {"32", "10"},
      {"-numCells"}, {"-numSteps", "5", "-numCells", "24", "-numCells", "20"},
      {"-std", "-std", "-vtk"}, {"-bogus"}, {"-h"}, {"-precision", "bf16", "-timeBlock", "8"},
      {"-numCells", "-3"}, {"-numCells", "99999999999999999999"}};
  for (const auto& c : cases) {
    std::vector<std::string> keep = {"upwindCxx"};
    keep.insert(keep.end(), c.begin(), c.end());
    std::vector<char*> argv;
    for (auto& s : keep) argv.push_back(&s[0]);
    argv.push_back(nullptr);
    fidibench::CmdLineArgs a;
    a.set("-numCells", 128, "Number of cells along each axis.");
    a.set("-numSteps", 10, "Number of time steps.");
    a.set("-vtk", false, "Write VTK files.");
    a.set("-std", false, "Print the standard deviation.");
    a.set("-precision", "f64", "Storage type.");
    a.set("-timeBlock", 1, "Steps per pass.");
    a.set("-dt", 0.5, "Time step.");
    if (a.parse((int)keep.size(), argv.data())) {
      (void)a.get<int>("-numCells");
      (void)a.get<int>("-nope");
      (void)a.get<double>("-dt");
      (void)a.get<std::string>("-precision");
      (void)a.get<bool>("-std");
    }
  }
}

void constructors() {
  const double v[3] = {0.3, -0.2, 0.1}, len[3] = {1, 1, 1};
  const uint64_t ok[3] = {8, 8, 8}, zero[3] = {8, 0, 8};
  const uint64_t wraps[3] = {1ull << 21, 1ull << 21, 1ull << 20};  // 2^62 cells
  fdb_upwind_t* h = nullptr;
  check(fdb_upwind_create(0, v, len, ok, FDB_F64, &h) == FDB_ERR_INVALID, "create ndims");
  check(fdb_upwind_create(3, v, len, zero, FDB_F64, &h) == FDB_ERR_INVALID, "create zero");
  check(fdb_upwind_create(3, v, len, ok, 9, &h) == FDB_ERR_INVALID, "create dtype");
  check(fdb_upwind_create(3, nullptr, len, ok, FDB_F64, &h) == FDB_ERR_INVALID, "create null");
  check(fdb_upwind_create(3, v, len, wraps, FDB_F64, &h) == FDB_ERR_INVALID, "create bytes");
  check(fdb_upwind_create(3, v, len, ok, FDB_F64, nullptr) == FDB_ERR_INVALID, "create out");
  check(fdb_upwind_advect(nullptr, 1, 0.1) == FDB_ERR_INVALID, "advect null");
  check(fdb_upwind_destroy(nullptr) == FDB_OK, "destroy null");
  fdb_tape_t* t = nullptr;
  check(fdb_tape_create(3, ok, v, len, 0.01, 0, FDB_F64, FDB_F64, 1 << 20, &t) ==
            FDB_ERR_INVALID, "tape steps");
  check(fdb_tape_create(3, wraps, v, len, 0.01, 4, FDB_F64, FDB_F64, 1 << 20, &t) ==
            FDB_ERR_INVALID, "tape bytes");
  check(fdb_tape_create(3, ok, v, len, 0.01, 4, FDB_F64, FDB_BF16_SR, 1 << 20, &t) ==
            FDB_ERR_INVALID, "tape lam dtype");
  check(fdb_tape_set_method(nullptr, FDB_GRAD_SHIFT) == FDB_ERR_INVALID, "tape method null");
  fdb_filter_t* f = nullptr;
  const int offs[6] = {0, 0, 1, 0, 9, 0};
  const double w[3] = {1.0, 0.5, 0.25};
  const uint64_t d2[2] = {16, 16};
  check(fdb_filter_create(2, d2, len, len, 3, offs, w, &f) == FDB_ERR_INVALID, "filter offset");
  check(fdb_filter_create(2, d2, len, len, 0, offs, w, &f) == FDB_ERR_INVALID, "filter terms");
  const uint64_t big[3] = {(1ull << 31) - 1, (1ull << 31) - 1, 1};
  check(fdb_filter_create(3, big, len, len, 1, offs, w, &f) == FDB_ERR_INVALID, "filter bytes");
  // Valid arguments get as far as the device: none here.
  const int rc = fdb_upwind_create(3, v, len, ok, FDB_F64, &h);
  check(rc == FDB_ERR_NODEV || (rc == FDB_OK && fdb_upwind_destroy(h) == FDB_OK), "create nodev");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "tuning") == 0) {
    // Malformed tuning variables: every launching constructor refuses
    // with FDB_ERR_INVALID, naming the variable.
    setenv("FDB_TB_CFG_8_6", "9,8,0,1,2,3,4,5,6,7,8,9", 1);
    setenv("FDB_ROWS_CFG", "8,2", 1);
    const double v[3] = {0.3, -0.2, 0.1}, len[3] = {1, 1, 1};
    const uint64_t ok[3] = {8, 8, 8};
    fdb_upwind_t* h = nullptr;
    check(fdb_upwind_create(3, v, len, ok, FDB_F64, &h) == FDB_ERR_INVALID &&
              std::strstr(fdb_last_error(), "FDB_") != nullptr, "tuning refused");
    fdb_tape_t* t = nullptr;
    check(fdb_tape_create(3, ok, v, len, 0.01, 4, FDB_F64, FDB_F64, 1 << 30, &t) ==
              FDB_ERR_INVALID, "tuning refused (tape)");
  } else {
    plans();
    folds_and_vtk();
    raw_headers();
    cli();
    constructors();
  }
  if (g_fail) {
    std::fprintf(stderr, "%d check(s) failed\n", g_fail);
    return 1;
  }
  std::printf("san_host ok%s\n", argc > 1 ? " (tuning)" : "");
  return 0;
}
