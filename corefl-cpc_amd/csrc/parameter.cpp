// setup.txt parser and defaults.  Grammar: Parameter.cpp:38-184; defaults: Parameter.cpp:342-523;
// derived values: Parameter.cpp:186-340.
#include "parameter.hpp"

#include <sstream>

namespace corefl {

Parameter::Parameter() {
  // Defaults the RHS/time-advance path reads (Parameter.cpp:342-523).
  ints["problem_type"] = 0;
  strings["canonical_problem"] = "default";
  ints["gridIsBinary"] = 0;
  reals["gridScale"] = 1.0;
  bools["steady"] = true;
  ints["temporal_scheme"] = 3;
  bools["fixed_time_step"] = false;
  reals["cfl"] = 0.8;
  reals["dt"] = 1e-8;
  ints["total_step"] = 1;
  reals["total_simulation_time"] = 1e+10;
  ints["inviscid_scheme"] = 3;
  ints["reconstruction"] = 2;
  ints["limiter"] = 0;
  reals["entropy_fix_factor"] = 0.125;
  ints["viscous_order"] = 2;
  // this build's: thickness (planes) of the z-slabs an air stage is cut into so that flux kernels and updates
  // overlap; -1 chooses 4 march lengths, 0 runs each stage unsliced
  ints["stage_slab_planes"] = -1;
  reals["shockSensor_threshold"] = 0;
  ints["shock_sensor"] = 0;           // 0 Ducros, 1 Jameson, 2 density/pressure jump (Parameter.cpp:384)
  reals["shockSensor_epsilon"] = -1;  // the Ducros sensor's epsilon (Parameter.cpp:386)
  bools["positive_preserving"] = false;
  ints["species"] = 0;
  ints["reaction"] = 0;
  ints["chemSrcMethod"] = 0;
  ints["chem_advance_method"] = 0; // Wu's splitting reaction sub-steps (Parameter.cpp:431)
  ints["hardCodedMech"] = 0;        // hard-coded hydrogen mechanisms (Parameter.cpp:399)
  ints["symmetric_conservative_metric"] = 0; // 2, 4, 6: symmetric-conservative metrics (Parameter.cpp:351)
  strings["mechanism_file"] = "chemistry/mechanism.inp";
  strings["therm_file"] = "chemistry/therm.dat";
  strings["transport_file"] = "chemistry/tran.dat"; // Parameter.cpp:401
  bools["gradPInDiffusionFlux"] = false;            // Parameter.cpp:382
  reals["chemistry_rtol"] = 1e-6; // ROK4E tolerances of the split chemistry sub-steps (new)
  reals["chemistry_atol"] = 1e-8;
  ints["chemistry_max_attempts"] = 100000; // per-cell step-attempt limit of a split chemistry sub-step (new)
  bools["turbulence"] = false;
  strings["default_init"] = "freestream";
  strings["reference_state"] = "freestream";
  ints["groups_init"] = 1;
  reals["prandtl_number"] = 0.72;
  reals["schmidt_number"] = 0.5;
  bools["parallel"] = false;
  // residual monitor: the reference requires output_screen in setup.txt; 0 (here the default) turns it off
  ints["output_screen"] = 0;
  ints["initial"] = 0; // 1: continue from output/flowfield.plt (set in every example setup.txt, read in Initialize.cu:12)
  bools["limit_flow"] = false; // Parameter.cpp:490
  // mixing layer (canonical_problem mixingLayer; MixingLayer.cu:9-193, Parameter.cpp:486-487)
  reals["velocity_ratio"] = -1;
  reals["ma_c"] = 0;
  bools["upper_faster"] = true;
  reals["delta_omega"] = 1e-3;
  ints["characteristic_velocity_ml"] = 1;
  ints["fluctuation_seed"] = 0; // seed of the inflow faces' white noise (new: the reference seeds curand from the clock)
  // sponge layer (Parameter.cpp:467-484; SpongeLayer.cu)
  bools["sponge_layer"] = false;
  ints["sponge_function"] = 0;
  ints["spongeX"] = 0;
  ints["spongeY"] = 0;
  ints["spongeZ"] = 0;
  reals["spongeCoefficient"] = 1.0;
  for (const char *k : {"spongeXMinusStart", "spongeXMinusEnd", "spongeXPlusStart", "spongeXPlusEnd", "spongeYMinusStart",
                        "spongeYMinusEnd", "spongeYPlusStart", "spongeYPlusEnd", "spongeZMinusStart", "spongeZMinusEnd",
                        "spongeZPlusStart", "spongeZPlusEnd"})
    reals[k] = 0;
}

static bool starts_with(const std::string &s, const char *p) { return s.rfind(p, 0) == 0; }

// Read tokens of an array body until "}" (Parameter.cpp:135-160).  A "//" token ends the line.
template<class T, class Conv> static bool read_tokens(std::istringstream &line, std::vector<T> &arr, Conv conv) {
  std::string t;
  while (line >> t) {
    if (t == "}") return false;
    if (t == "//") break;
    arr.push_back(conv(t));
  }
  return true;
}

void Parameter::parse(const std::string &text) {
  std::istringstream file(text);
  std::string input;
  while (std::getline(file, input)) {
    if (input.empty() || starts_with(input, "//") || starts_with(input, "!")) continue;
    std::istringstream line(input);
    std::string type, key, temp;
    if (!(line >> type)) continue;
    line >> key >> temp;
    if (type == "int") {
      int v{};
      line >> v;
      ints[key] = v;
    } else if (type == "real") {
      double v{};
      line >> v;
      reals[key] = v;
    } else if (type == "bool") {
      int v{};
      line >> v;
      bools[key] = v != 0;
    } else if (type == "string") {
      std::string v;
      line >> v;
      strings[key] = v;
    } else if (type == "array" || type == "range") {
      // "array <type> <name> {" ; values may continue on following lines until "}"
      const std::string name = temp;
      std::string brace;
      line >> brace; // "{" (or "=" followed by "{")
      if (brace == "=") line >> brace;
      auto body = [&](auto &arr, auto conv) {
        bool more = read_tokens(line, arr, conv);
        while (more && std::getline(file, input)) {
          std::istringstream l2(input);
          more = read_tokens(l2, arr, conv);
        }
      };
      if (key == "int") {
        std::vector<int> a;
        body(a, [](const std::string &s) { return std::stoi(s); });
        int_arrays[name] = a;
      } else if (key == "real") {
        std::vector<double> a;
        body(a, [](const std::string &s) { return std::stod(s); });
        real_arrays[name] = a;
      } else if (key == "string") {
        std::vector<std::string> a;
        body(a, [](const std::string &s) { return s; });
        string_arrays[name] = a;
      }
    } else if (type == "struct") {
      // Parameter.cpp:162-184: members "type key = value" until a line whose first token is "}"
      Struct st;
      while (std::getline(file, input)) {
        std::istringstream l2(input);
        std::string t, k, eq;
        if (!(l2 >> t)) continue;
        if (t == "}") break;
        if (starts_with(t, "//") || starts_with(t, "!")) continue;
        l2 >> k >> eq;
        if (t == "string") {
          std::string v;
          l2 >> v;
          st[k] = v;
        } else if (t == "int") {
          int v{};
          l2 >> v;
          st[k] = v;
        } else if (t == "real") {
          double v{};
          l2 >> v;
          st[k] = v;
        }
      }
      structs[key] = st;
    }
  }
}

void Parameter::deduce() {
  // ngg and the scheme tags (Parameter.cpp:197-305).
  int ngg = 2;
  const int recon = get_int("reconstruction");
  const int inv = get_int("inviscid_scheme");
  if (recon == 4) ngg = 3;             // WENO5 reconstruction
  if (recon == 5) ngg = 4;             // WENO7 reconstruction
  if (inv == 51 || inv == 52) ngg = 3; // WENO5 LLF split
  if (inv == 71 || inv == 72) ngg = 4; // WENO7 LLF split
  const int vo = get_int("viscous_order");
  if (vo == 6 && ngg < 3) ngg = 3;
  if (vo == 8 && ngg < 4) ngg = 4;
  ints["ngg"] = ngg;
  // n_var = 5 + n_spec (Parameter.cpp:548-612); this build runs the single-species (air) model.
  ints["n_var"] = 5;
  ints["n_spec"] = 0;
  ints["n_scalar"] = 0;
  // periodic label from the boundary-condition structs (Mesh.cpp:639-646)
  int periodic_label = -100;
  if (string_arrays.count("boundary_conditions")) {
    for (const auto &n : string_arrays.at("boundary_conditions")) {
      if (!structs.count(n)) continue;
      const auto &s = structs.at(n);
      if (struct_string(s, "type", "") == "periodic") periodic_label = struct_int(s, "label", 8);
    }
  }
  ints["periodic_label"] = periodic_label;
}

double struct_real(const Struct &s, const std::string &k, double dflt) {
  auto it = s.find(k);
  if (it == s.end()) return dflt;
  if (auto p = std::get_if<double>(&it->second)) return *p;
  if (auto p = std::get_if<int>(&it->second)) return *p;
  return dflt;
}

int struct_int(const Struct &s, const std::string &k, int dflt) {
  auto it = s.find(k);
  if (it == s.end()) return dflt;
  if (auto p = std::get_if<int>(&it->second)) return *p;
  if (auto p = std::get_if<double>(&it->second)) return static_cast<int>(*p);
  return dflt;
}

std::string struct_string(const Struct &s, const std::string &k, const std::string &dflt) {
  auto it = s.find(k);
  if (it == s.end()) return dflt;
  if (auto p = std::get_if<std::string>(&it->second)) return *p;
  return dflt;
}

} // namespace corefl
