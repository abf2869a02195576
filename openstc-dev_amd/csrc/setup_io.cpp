// Setup-file reader with the reference's grammar and file order.
//   grammar: Parameter.cpp:53-163 (read_one_file, read_line_to_array, read_struct)
//   order:   Parameter.h:69-82 (default_settings, 0_global_control ... 10_statistics)
//   derived: Parameter.cpp:165-219 (ngg from the reconstruction, inviscid_type)
#include <fstream>
#include <sstream>

#include "ostc_internal.h"

namespace ostc {

namespace {

void trim_left(std::string &s) {
  const auto p = s.find_first_not_of(" \n\t\r\f\v");
  s = p == std::string::npos ? std::string() : s.substr(p);
}

// one array line: returns false once "}" was read (the array is complete)
template <typename T>
bool array_line(std::istringstream &line, std::vector<T> &arr) {
  std::string tok;
  while (line >> tok) {
    if (tok == "}") return false;
    if (tok == "//") break;
    if constexpr (std::is_same_v<T, double>)
      arr.push_back(std::stod(tok));
    else if constexpr (std::is_same_v<T, int>)
      arr.push_back(std::stoi(tok));
    else
      arr.push_back(tok);
  }
  return true;
}

bool next_line(std::ifstream &f, std::string &input, std::istringstream &line) {
  if (!std::getline(f, input)) return false;
  trim_left(input);
  line.clear();
  line.str(input);
  return true;
}

}  // namespace

void Params::read_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) return;  // a missing setup file contributes nothing, as in the reference
  std::string input, type, key, temp;
  std::istringstream line;
  while (std::getline(f, input)) {
    if (input.rfind("//", 0) == 0 || input.rfind("!", 0) == 0 || input.empty()) continue;
    line.clear();
    line.str(input);
    type.clear();
    key.clear();
    temp.clear();
    line >> type >> key >> temp;
    if (type == "int") {
      int v{};
      line >> v;
      i[key] = v;
    } else if (type == "real") {
      double v{};
      line >> v;
      r[key] = v;
    } else if (type == "bool") {
      bool v{};
      line >> v;
      b[key] = v;
    } else if (type == "string") {
      std::string v;
      line >> v;
      s[key] = v;
    } else if (type == "array") {
      const std::string name = temp;
      if (key == "int") {
        std::vector<int> arr;
        line >> temp;
        while (array_line(line, arr))
          if (!next_line(f, input, line)) break;
        ia[name] = arr;
      } else if (key == "real") {
        std::vector<double> arr;
        line >> temp;
        while (array_line(line, arr))
          if (!next_line(f, input, line)) break;
        ra[name] = arr;
      } else if (key == "string") {
        std::vector<std::string> arr;
        line >> temp >> temp;
        while (array_line(line, arr))
          if (!next_line(f, input, line)) break;
        sa[name] = arr;
      }
    } else if (type == "struct") {
      Struct stc;
      std::string k2, t2;
      while (next_line(f, input, line)) {
        k2.clear();
        line >> k2;
        if (k2 == "}") break;
        if (k2 == "string") {
          std::string v;
          line >> k2 >> t2 >> v;
          stc.emplace(k2, v);
        } else if (k2 == "int") {
          int v{};
          line >> k2 >> t2 >> v;
          stc.emplace(k2, v);
        } else if (k2 == "real") {
          double v{};
          line >> k2 >> t2 >> v;
          stc.emplace(k2, v);
        }
      }
      st[key] = stc;
    }
  }
}

int Params::get_int(const std::string &k, int d) const {
  auto it = i.find(k);
  return it == i.end() ? d : it->second;
}
double Params::get_real(const std::string &k, double d) const {
  auto it = r.find(k);
  return it == r.end() ? d : it->second;
}
bool Params::get_bool(const std::string &k, bool d) const {
  auto it = b.find(k);
  return it == b.end() ? d : it->second;
}
std::string Params::get_string(const std::string &k, const std::string &d) const {
  auto it = s.find(k);
  return it == s.end() ? d : it->second;
}

void read_setup(const std::string &case_dir, Params &p) {
  static const char *files[] = {"default_settings.txt", "0_global_control.txt", "1_grid_information.txt",
                                "2_scheme.txt", "3_species_reactions.txt", "4_laminar_turbulent.txt",
                                "5_boundary_condition.txt", "6_post_process.txt", "7_output_control.txt",
                                "8_initialization.txt", "9_transport_property.txt", "10_statistics.txt"};
  for (const char *f : files) p.read_file(case_dir + "/input_files/setup/" + f);
  p.i.emplace("step", 0);
  // Parameter.cpp:165-219 deduce_known_info
  int ngg = 2;
  const int rec = p.get_int("reconstruction", 2);
  if (rec == 4) ngg = 3;
  if (rec == 5) ngg = 4;
  p.i["ngg"] = ngg;
  int inviscid_type = 0;
  if (p.get_int("inviscid_scheme", 3) == 2) inviscid_type = 2;
  if (rec == 4 || rec == 5) inviscid_type = 1;
  p.i["inviscid_type"] = inviscid_type;
  // Parameter.cpp:30-52: turbulence variables
  int n_turb = 0;
  if (p.get_bool("turbulence", false) && p.get_int("turbulence_method", 1) == 1)
    n_turb = p.get_int("RANS_model", 2) == 1 ? 1 : 2;
  if (!p.get_bool("turbulence", false)) p.i["RANS_model"] = 0;
  p.i["n_turb"] = n_turb;
}

}  // namespace ostc
