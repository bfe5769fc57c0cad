// CmdLineArgs.hpp -- command-line options with the semantics of
// pletzer/fidibench cxx/CmdLineArgParser.{h,cpp}, which the upwindCxx
// executable uses (upwind.cxx:143-157):
//   * options are typed (double, int, string, bool) and kept in name order;
//   * parse() first rejects any token that looks like an option name
//     ('-' followed by a non-digit) and is not registered, printing
//     "<token> is not a valid option." (CmdLineArgParser.h:80-96);
//   * a typed option takes the token after it (atof/atoi/verbatim), the
//     last occurrence winning; a trailing option without a value keeps its
//     default; other tokens, including positionals, are ignored (:98-154);
//   * a bool option flips its value at every occurrence (:155-166);
//   * get<T>() of an unknown name returns -max for int/double, "" for
//     string and false for bool (CmdLineArgParser.cpp:18-59);
//   * "-h" is always registered (CmdLineArgParser.h:29-33).
#pragma once

#include <cctype>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <map>
#include <string>
#include <utility>

namespace fidibench {

class CmdLineArgs {
 public:
  CmdLineArgs() { set("-h", false, "Print help."); }

  void set(const std::string& name, double def, const std::string& help) { dbl_[name] = {def, help}; }
  void set(const std::string& name, int def, const std::string& help) { int_[name] = {def, help}; }
  void set(const std::string& name, const std::string& def, const std::string& help) {
    str_[name] = {def, help};
  }
  void set(const std::string& name, const char* def, const std::string& help) {
    set(name, std::string(def), help);
  }
  void set(const std::string& name, bool def, const std::string& help) { bool_[name] = {def, help}; }

  void setPurpose(const std::string& p) { purpose_ = p; }
  void setFooter(const std::string& f) { footer_ = f; }

  bool parse(int argc, char** argv) {
    exec_ = argc > 0 ? argv[0] : "";
    for (int a = 1; a < argc; ++a) {
      const std::string tok(argv[a]);
      const bool looks_like_option =
          tok.size() >= 2 && tok[0] == '-' && !std::isdigit((unsigned char)tok[1]);
      if (looks_like_option && !known(tok)) {
        std::cout << tok << " is not a valid option.\n";
        return false;
      }
    }
    take_values(dbl_, argc, argv, [](const char* s) { return std::atof(s); });
    take_values(int_, argc, argv, [](const char* s) { return std::atoi(s); });
    take_values(str_, argc, argv, [](const char* s) { return std::string(s); });
    for (auto& kv : bool_)
      for (int a = 1; a < argc; ++a)
        if (kv.first == argv[a]) kv.second.first = !kv.second.first;
    return true;
  }

  template <class T>
  T get(const std::string& name) const;

  void help() const {
    std::cout << purpose_ << std::endl;
    std::cout << exec_ << " [options]\n";
    std::cout << "Usage:\n";
    for (const auto& kv : dbl_)
      std::cout << "\t" << kv.first << " <double#> " << kv.second.second << " ("
                << kv.second.first << ")\n";
    for (const auto& kv : int_)
      std::cout << "\t" << kv.first << " <int#> " << kv.second.second << " (" << kv.second.first
                << ")\n";
    for (const auto& kv : str_)
      std::cout << "\t" << kv.first << " <string> " << kv.second.second << " ("
                << kv.second.first << ")\n";
    for (const auto& kv : bool_)
      std::cout << "\t" << kv.first << ' ' << kv.second.second << " (" << kv.second.first
                << ")\n";
    std::cout << footer_ << std::endl;
  }

 private:
  template <class T>
  using Table = std::map<std::string, std::pair<T, std::string>>;

  bool known(const std::string& t) const {
    return dbl_.count(t) || int_.count(t) || str_.count(t) || bool_.count(t);
  }

  template <class T, class Conv>
  static void take_values(Table<T>& tab, int argc, char** argv, Conv conv) {
    for (auto& kv : tab) {
      bool next_is_value = false;
      for (int a = 1; a < argc; ++a) {
        if (next_is_value) {
          kv.second.first = conv(argv[a]);
          next_is_value = false;
        }
        if (kv.first == argv[a]) next_is_value = true;
      }
    }
  }

  Table<double> dbl_;
  Table<int> int_;
  Table<std::string> str_;
  Table<bool> bool_;
  std::string exec_, purpose_, footer_ = "\n";
};

template <>
inline double CmdLineArgs::get<double>(const std::string& n) const {
  auto it = dbl_.find(n);
  return it == dbl_.end() ? -std::numeric_limits<double>::max() : it->second.first;
}
template <>
inline int CmdLineArgs::get<int>(const std::string& n) const {
  auto it = int_.find(n);
  return it == int_.end() ? -std::numeric_limits<int>::max() : it->second.first;
}
template <>
inline std::string CmdLineArgs::get<std::string>(const std::string& n) const {
  auto it = str_.find(n);
  return it == str_.end() ? std::string() : it->second.first;
}
template <>
inline bool CmdLineArgs::get<bool>(const std::string& n) const {
  auto it = bool_.find(n);
  return it == bool_.end() ? false : it->second.first;
}

}  // namespace fidibench
