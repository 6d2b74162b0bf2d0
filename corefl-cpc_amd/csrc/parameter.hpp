// Typed key/value store for COREFL's input/setup.txt (grammar of Parameter.cpp:38-184).
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

namespace corefl {

using StructVal = std::variant<std::string, int, double>;
using Struct = std::map<std::string, StructVal>;

class Parameter {
public:
  Parameter();
  // Parse a setup text (the content of setup.txt).  Later keys override earlier ones.
  void parse(const std::string &text);
  // Derived values (ngg, n_var, schemes) as Parameter::deduce_known_info does (Parameter.cpp:186-340).
  void deduce();

  int get_int(const std::string &k) const { return at(ints, k); }
  double get_real(const std::string &k) const { return at(reals, k); }
  bool get_bool(const std::string &k) const { return at(bools, k); }
  const std::string &get_string(const std::string &k) const { return at(strings, k); }
  const std::vector<std::string> &get_string_array(const std::string &k) const { return at(string_arrays, k); }
  const Struct &get_struct(const std::string &k) const { return at(structs, k); }
  bool has_struct(const std::string &k) const { return structs.count(k) != 0; }
  bool has_int(const std::string &k) const { return ints.count(k) != 0; }
  bool has_real(const std::string &k) const { return reals.count(k) != 0; }

  void set_int(const std::string &k, int v) { ints[k] = v; }
  void set_real(const std::string &k, double v) { reals[k] = v; }
  void set_string(const std::string &k, const std::string &v) { strings[k] = v; }

private:
  template<class M> static const typename M::mapped_type &at(const M &m, const std::string &k) {
    auto it = m.find(k);
    if (it == m.end()) throw std::runtime_error("setup.txt: missing parameter '" + k + "'");
    return it->second;
  }
  std::map<std::string, int> ints;
  std::map<std::string, double> reals;
  std::map<std::string, bool> bools;
  std::map<std::string, std::string> strings;
  std::map<std::string, std::vector<int>> int_arrays;
  std::map<std::string, std::vector<double>> real_arrays;
  std::map<std::string, std::vector<std::string>> string_arrays;
  std::map<std::string, Struct> structs;
};

// Struct lookup helpers with defaults.
double struct_real(const Struct &s, const std::string &k, double dflt);
int struct_int(const Struct &s, const std::string &k, int dflt);
std::string struct_string(const Struct &s, const std::string &k, const std::string &dflt);

} // namespace corefl
