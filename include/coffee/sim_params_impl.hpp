// parse_config: reads the TOML subset the reference's config files use
// (key = value with integers, floats, booleans, strings and flat arrays of
// them, '#' comments).  Key names and types follow src/sim_params.cpp:7-146:
// integer keys ignore float values (keeping the default), float keys accept
// integers, `nodes` is read as an array of floats.
#ifndef COFFEE_SIM_PARAMS_IMPL_HPP
#define COFFEE_SIM_PARAMS_IMPL_HPP

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <vector>

namespace Coffee {
namespace detail {

inline std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

struct toml_value {
  std::vector<std::string> items;  // one item for scalars
  bool is_array = false;
};

inline bool is_int_token(const std::string& t) {
  if (t.empty()) return false;
  size_t i = (t[0] == '-' || t[0] == '+') ? 1 : 0;
  if (i >= t.size()) return false;
  for (; i < t.size(); i++)
    if (!std::isdigit((unsigned char)t[i]) && t[i] != '_') return false;
  return true;
}

inline bool is_float_token(const std::string& t) {
  if (is_int_token(t)) return true;
  char* end = nullptr;
  std::strtod(t.c_str(), &end);
  return end && *end == 0 && !t.empty();
}

inline std::map<std::string, toml_value> read_toml(const std::string& filename) {
  std::ifstream in(filename);
  if (!in) throw std::runtime_error("cannot open config file " + filename);
  std::map<std::string, toml_value> out;
  std::string line, pending;
  while (std::getline(in, line)) {
    bool in_str = false;
    for (size_t i = 0; i < line.size(); i++) {
      if (line[i] == '"') in_str = !in_str;
      if (line[i] == '#' && !in_str) {
        line = line.substr(0, i);
        break;
      }
    }
    pending += line;
    // arrays may span lines
    int depth = 0;
    for (char ch : pending) depth += (ch == '[') - (ch == ']');
    if (depth > 0) {
      pending += " ";
      continue;
    }
    std::string l = trim(pending);
    pending.clear();
    if (l.empty() || l[0] == '[') continue;  // tables are not used by the solver configs
    size_t eq = l.find('=');
    if (eq == std::string::npos) continue;
    std::string key = trim(l.substr(0, eq));
    std::string val = trim(l.substr(eq + 1));
    toml_value v;
    if (!val.empty() && val[0] == '[') {
      v.is_array = true;
      std::string body = val.substr(1, val.rfind(']') - 1);
      std::stringstream ss(body);
      std::string item;
      while (std::getline(ss, item, ',')) {
        item = trim(item);
        if (!item.empty()) v.items.push_back(item);
      }
    } else {
      v.items.push_back(val);
    }
    out[key] = v;
  }
  return out;
}

template <typename T>
inline void get_float(const std::map<std::string, toml_value>& m, const char* k, T& dst) {
  auto it = m.find(k);
  if (it == m.end() || it->second.is_array || !is_float_token(it->second.items[0])) return;
  dst = (T)std::strtod(it->second.items[0].c_str(), nullptr);
}

template <typename T>
inline void get_int(const std::map<std::string, toml_value>& m, const char* k, T& dst) {
  auto it = m.find(k);
  if (it == m.end() || it->second.is_array || !is_int_token(it->second.items[0])) return;
  dst = (T)std::strtoll(it->second.items[0].c_str(), nullptr, 10);
}

inline void get_bool(const std::map<std::string, toml_value>& m, const char* k, bool& dst) {
  auto it = m.find(k);
  if (it == m.end() || it->second.is_array) return;
  if (it->second.items[0] == "true") dst = true;
  else if (it->second.items[0] == "false") dst = false;
}

template <typename T>
inline void get_array(const std::map<std::string, toml_value>& m, const char* k, T* dst,
                      bool as_float) {
  auto it = m.find(k);
  if (it == m.end() || !it->second.is_array) return;
  for (size_t i = 0; i < 3 && i < it->second.items.size(); i++) {
    const std::string& t = it->second.items[i];
    if (t == "true" || t == "false") dst[i] = (T)(t == "true");
    else if (as_float) dst[i] = (T)std::strtod(t.c_str(), nullptr);
    else dst[i] = (T)std::strtoll(t.c_str(), nullptr, 10);
  }
}

}  // namespace detail

inline sim_params parse_config(const std::string& filename) {
  using namespace detail;
  auto m = read_toml(filename);
  sim_params r;
  get_float(m, "dt", r.dt);
  get_int(m, "max_steps", r.max_steps);
  get_int(m, "data_interval", r.data_interval);
  get_int(m, "snapshot_interval", r.snapshot_interval);
  get_int(m, "downsample", r.downsample);
  get_bool(m, "slice_x", r.slice_x);
  get_bool(m, "slice_y", r.slice_y);
  get_bool(m, "slice_z", r.slice_z);
  get_bool(m, "slice_xy", r.slice_xy);
  get_float(m, "slice_x_pos", r.slice_x_pos);
  get_float(m, "slice_y_pos", r.slice_y_pos);
  get_float(m, "slice_z_pos", r.slice_z_pos);
  get_int(m, "slice_interval", r.slice_interval);
  get_int(m, "downsample_2d", r.downsample_2d);
  get_array(m, "periodic_boundary", r.periodic_boundary, false);
  get_array(m, "guard", r.guard, false);
  get_array(m, "N", r.N, false);
  get_array(m, "lower", r.lower, true);
  get_array(m, "size", r.size, true);
  get_array(m, "nodes", r.nodes, true);
  get_int(m, "shift_ghost", r.shift_ghost);
  get_float(m, "a", r.a);
  get_bool(m, "calc_current", r.calc_current);
  get_bool(m, "clean_ep", r.clean_ep);
  get_bool(m, "check_egb", r.check_egb);
  get_array(m, "pml", r.pml, false);
  get_int(m, "pmllen", r.pmllen);
  get_float(m, "sigpml", r.sigpml);
  get_float(m, "damp_gamma", r.damp_gamma);
  get_bool(m, "use_edotb_damping", r.use_edotb_damping);
  get_bool(m, "divB_clean", r.divB_clean);
  get_float(m, "ch2", r.ch2);
  get_float(m, "tau", r.tau);
  get_float(m, "KOeps", r.KOeps);
  get_bool(m, "KO_geometry", r.KO_geometry);
  get_int(m, "problem", r.problem);
  get_float(m, "radius", r.radius);
  get_float(m, "omega", r.omega);
  get_float(m, "b0", r.b0);
  get_float(m, "p1", r.p1);
  get_float(m, "p2", r.p2);
  get_float(m, "p3", r.p3);
  get_float(m, "q11", r.q11);
  get_float(m, "q12", r.q12);
  get_float(m, "q13", r.q13);
  get_float(m, "q22", r.q22);
  get_float(m, "q23", r.q23);
  get_float(m, "q_offset_x", r.q_offset_x);
  get_float(m, "q_offset_y", r.q_offset_y);
  get_float(m, "q_offset_z", r.q_offset_z);
  get_float(m, "tp_start", r.tp_start);
  get_float(m, "tp_end", r.tp_end);
  get_float(m, "rpert1", r.rpert1);
  get_float(m, "rpert2", r.rpert2);
  get_float(m, "thp1", r.thp1);
  get_float(m, "thp2", r.thp2);
  get_float(m, "dw0", r.dw0);
  get_float(m, "nT", r.nT);
  get_float(m, "shear", r.shear);
  get_float(m, "ph1", r.ph1);
  get_float(m, "ph2", r.ph2);
  get_float(m, "dph", r.dph);
  get_float(m, "theta0", r.theta0);
  get_float(m, "drpert", r.drpert);
  get_float(m, "tp_start1", r.tp_start1);
  get_float(m, "tp_end1", r.tp_end1);
  get_float(m, "rpert11", r.rpert11);
  get_float(m, "rpert21", r.rpert21);
  get_float(m, "thp11", r.thp11);
  get_float(m, "thp21", r.thp21);
  get_float(m, "dw1", r.dw1);
  get_float(m, "nT1", r.nT1);
  get_float(m, "shear1", r.shear1);
  get_float(m, "ph11", r.ph11);
  get_float(m, "ph21", r.ph21);
  get_float(m, "dph1", r.dph1);
  get_int(m, "pert_type", r.pert_type);
  return r;
}

}  // namespace Coffee

#endif
