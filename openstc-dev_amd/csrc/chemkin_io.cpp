// CHEMKIN mechanism, NASA-7 therm.dat and tran.dat readers with the
// reference's semantics (ChemData.cpp:11-685, Element.cpp:13-160):
//  * species and element indices follow file order;
//  * molecular weight sums element weights in alphabetical element order;
//  * T_low/T_high/T_mid per species from therm columns 45/55/65, defaults
//    300/1000/5000 (the THERMO header line does not change them);
//  * reaction Ea is divided by R_c (cal/mol -> K); the second Arrhenius set
//    of LOW and REV lines is kept as written (no conversion), while a
//    DUPLICATE partner's Ea arrives already converted;
//  * DUPLICATE pairs with identical stoichiometry merge into one reaction
//    (label 3) whose second set is A2/b2/Ea2;
//  * third-body efficiencies default to 1.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

#include "ostc_internal.h"

namespace ostc {

// Standard atomic weights as tabulated by the reference (Element.cpp:13-132);
// only the symbols a combustion mechanism may name are listed.
double atomic_weight(const std::string &e) {
  static const std::map<std::string, double> w = {
      {"H", 1.008},    {"HE", 4.002602}, {"C", 12.011}, {"N", 14.007},          {"O", 15.999},
      {"F", 18.998403163}, {"NE", 20.1797}, {"S", 32.06}, {"CL", 35.45}, {"AR", 39.95},
      {"D", 2.0141017781}, {"TR", 3.0160492820}};
  auto it = w.find(e);
  return it == w.end() ? -1.0 : it->second;
}

namespace {

void upper(std::string &s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
}

void trim_left(std::string &s) {
  const auto p = s.find_first_not_of(" \n\t\r\f\v");
  s = p == std::string::npos ? std::string() : s.substr(p);
}

bool getline_up(std::ifstream &f, std::string &in) {
  if (!std::getline(f, in)) return false;
  trim_left(in);
  upper(in);
  return true;
}

struct Work {
  Mechanism &m;
  std::map<std::string, int> spec;   // name -> index
  std::map<std::string, int> elem;   // name -> index (file order)
  int n_elem = 0;
};

std::string read_species(std::ifstream &f, Work &w) {
  std::string in;
  while (f >> in) {
    if (in[0] == '!') {
      std::getline(f, in);
      continue;
    }
    if (in == "ELEMENTS" || in == "ELEM") break;
  }
  while (f >> in) {
    if (in[0] == '!') {
      std::getline(f, in);
      continue;
    }
    upper(in);
    if (in == "END") continue;
    if (in == "SPECIES" || in == "SPEC") break;
    w.elem.emplace(in, w.n_elem++);
  }
  int ns = 0;
  while (f >> in) {
    if (in[0] == '!') {
      std::getline(f, in);
      continue;
    }
    upper(in);
    if (in == "END") continue;
    if (in == "REACTIONS" || in == "REAC" || in == "THERMO") break;
    if (w.spec.emplace(in, ns).second) {
      w.m.names.push_back(in);
      ++ns;
    }
  }
  if (ns > kMaxSpec) return "more species than the supported maximum of 20";
  Mechanism &m = w.m;
  m.n_spec = ns;
  m.elements.assign(w.n_elem, "");
  for (auto &[n, i] : w.elem) m.elements[i] = n;
  m.elem_comp.assign((size_t)ns * w.n_elem, 0);
  m.mw.assign(ns, 0);
  m.t_low.assign(ns, 300);
  m.t_mid.assign(ns, 1000);
  m.t_high.assign(ns, 5000);
  m.hi.assign((size_t)ns * 7, 0);
  m.lo.assign((size_t)ns * 7, 0);
  m.LJ_potent_inv.assign(ns, 0);
  m.vis_coeff.assign(ns, 0);
  m.WjDivWi_to_One4th.assign((size_t)ns * ns, 0);
  m.sqrt_WiDivWjPl1Mul8.assign((size_t)ns * ns, 0);
  m.binary_diffusivity_coeff.assign((size_t)ns * ns, 0);
  m.kb_over_eps_jk.assign((size_t)ns * ns, 0);
  return "";
}

double field(const std::string &s, size_t pos, size_t len) {
  if (pos >= s.size()) return 0.0;
  return std::stod(s.substr(pos, len));
}

std::string read_therm(std::ifstream &f, Work &w) {
  Mechanism &m = w.m;
  const int ns = m.n_spec;
  std::string in;
  // skip to THERMO, then the global temperature line
  while (getline_up(f, in))
    if (in.rfind("THERMO", 0) == 0) break;
  while (std::getline(f, in)) {
    if (in.empty() || in[0] == '!') continue;
    break;
  }
  std::vector<char> have(ns, 0);
  int n_read = 0;
  std::string key;
  while (getline_up(f, in)) {
    if (in.empty() || in[0] == '!') continue;
    std::istringstream line(in);
    line >> key;
    if (key == "END") continue;
    if (key == "REACTIONS" || key == "TRANSPORT" || key == "TRAN") break;
    if (n_read >= ns) break;
    {
      std::istringstream nm(in.substr(0, 18));
      nm >> key;
    }
    auto it = w.spec.find(key);
    if (it == w.spec.end() || have[it->second]) {
      std::string skip;
      std::getline(f, skip);
      std::getline(f, skip);
      std::getline(f, skip);
      continue;
    }
    const int sp = it->second;
    m.t_low[sp] = field(in, 45, 10);
    m.t_high[sp] = field(in, 55, 10);
    if (in.size() > 65) {
      std::istringstream tm(in.substr(65, 10));
      std::string tok;
      if (tm >> tok) m.t_mid[sp] = std::stod(tok);
    }
    for (int e = 0; e < 4; ++e) {
      if ((size_t)(24 + e * 5) >= in.size()) break;
      std::string comp = in.substr(24 + e * 5, 5);
      trim_left(comp);
      if (comp.empty() || comp[0] == '0') break;
      std::istringstream cs(comp);
      std::string el;
      int cnt = 0;
      cs >> el;
      auto ei = w.elem.find(el);
      if (ei == w.elem.end()) continue;
      cs >> cnt;
      m.elem_comp[(size_t)sp * w.n_elem + ei->second] = cnt;
    }
    double mwt = 0;
    for (auto &[el, idx] : w.elem) mwt += atomic_weight(el) * m.elem_comp[(size_t)sp * w.n_elem + idx];
    m.mw[sp] = mwt;
    std::string l2, l3, l4;
    std::getline(f, l2);
    std::getline(f, l3);
    std::getline(f, l4);
    double *hi = &m.hi[(size_t)sp * 7], *lo = &m.lo[(size_t)sp * 7];
    for (int c = 0; c < 5; ++c) hi[c] = field(l2, 15 * c, 15);
    hi[5] = field(l3, 0, 15);
    hi[6] = field(l3, 15, 15);
    lo[0] = field(l3, 30, 15);
    lo[1] = field(l3, 45, 15);
    lo[2] = field(l3, 60, 15);
    for (int c = 0; c < 4; ++c) lo[3 + c] = field(l4, 15 * c, 15);
    have[sp] = 1;
    ++n_read;
  }
  if (n_read < ns) return "thermodynamic data missing for some species";
  return "";
}

bool polar(double mu) { return mu != 0; }

double reduced_dipole(int i, const std::vector<double> &mu, const std::vector<double> &eps,
                      const std::vector<double> &sig) {
  if (!polar(mu[i])) return 0;
  const double debye = 1e-18 * std::sqrt(1e-7);
  const double ang = std::pow(1e+8, 1.5);
  const double conv = debye * ang / std::sqrt(1.380649e-23);
  return conv * mu[i] / std::sqrt(eps[i] * sig[i] * sig[i] * sig[i]);
}

std::string read_tran(std::ifstream &f, Work &w) {
  Mechanism &m = w.m;
  const int ns = m.n_spec;
  std::vector<double> eps(ns, 0), sig(ns, 0), mu(ns, 0), alpha(ns, 0);
  std::vector<char> have(ns, 0);
  int n_read = 0;
  std::string in, key;
  while (getline_up(f, in)) {
    if (in.empty() || in[0] == '!') continue;
    std::istringstream line(in);
    line >> key;
    if (key.rfind("END", 0) == 0 || key.rfind("REAC", 0) == 0) break;
    auto it = w.spec.find(key);
    if (it == w.spec.end()) continue;
    if (n_read >= ns) break;
    const int sp = it->second;
    if (have[sp]) continue;
    double pass;
    line >> pass >> eps[sp] >> sig[sp] >> mu[sp] >> alpha[sp];
    m.LJ_potent_inv[sp] = 1.0 / eps[sp];
    m.vis_coeff[sp] = 2.6693e-6 * std::sqrt(m.mw[sp]) / (sig[sp] * sig[sp]);
    have[sp] = 1;
    ++n_read;
  }
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      m.WjDivWi_to_One4th[i * ns + j] = std::pow(m.mw[j] / m.mw[i], 0.25);
      m.sqrt_WiDivWjPl1Mul8[i * ns + j] = 1.0 / std::sqrt(8 * (1 + m.mw[i] / m.mw[j]));
    }
  for (int j = 0; j < ns; ++j)
    for (int k = j + 1; k < ns; ++k) {
      const double Wjk = m.mw[j] * m.mw[k] / (m.mw[j] + m.mw[k]);
      double xi = 1;
      if (polar(mu[j]) != polar(mu[k])) {
        const int p = polar(mu[j]) ? j : k, n = polar(mu[j]) ? k : j;
        const double an = alpha[n] / std::pow(sig[n], 3);
        const double mp = reduced_dipole(p, mu, eps, sig);
        xi = 1.0 + 0.25 * an * mp * mp * std::sqrt(eps[p] / eps[n]);
      }
      const double sjk = 0.5 * (sig[j] + sig[k]) * std::pow(xi, -1.0 / 6.0);
      const double bdc = 0.0188324 / (std::sqrt(Wjk) * sjk * sjk);
      m.binary_diffusivity_coeff[j * ns + k] = bdc;
      m.binary_diffusivity_coeff[k * ns + j] = bdc;
      const double ejk = xi * xi * std::sqrt(eps[j] * eps[k]);
      m.kb_over_eps_jk[j * ns + k] = 1.0 / ejk;
      m.kb_over_eps_jk[k * ns + j] = 1.0 / ejk;
    }
  if (n_read < ns) return "transport data missing for some species";
  return "";
}

struct RxnRead {
  Work &w;
  int ns;
  void resize(int nr) {
    Mechanism &m = w.m;
    m.n_reac = nr;
    m.label.resize(nr, 1);
    m.stoi_f.resize((size_t)nr * ns, 0);
    m.stoi_b.resize((size_t)nr * ns, 0);
    m.order.resize(nr, 0);
    m.A.resize(nr, 0);
    m.b.resize(nr, 0);
    m.Ea.resize(nr, 0);
    m.A2.resize(nr, 0);
    m.b2.resize(nr, 0);
    m.Ea2.resize(nr, 0);
    m.third_body_coeff.resize((size_t)nr * ns, 1);
    m.troe_alpha.resize(nr, 0);
    m.troe_t3.resize(nr, 0);
    m.troe_t1.resize(nr, 0);
    m.troe_t2.resize(nr, 0);
  }

  int sp(const std::string &n, std::string &err) {
    auto it = w.spec.find(n);
    if (it == w.spec.end()) {
      err = "unknown species '" + n + "' in reaction";
      return 0;
    }
    return it->second;
  }

  // ChemData.cpp:525-632 read_reaction_line
  void equation(std::string in, int idx, std::string &err) {
    Mechanism &m = w.m;
    auto lp = in.find("(+");
    auto cat = in.find("+M");
    auto rp = in.find(')', lp);
    if (cat != std::string::npos) {
      m.label[idx] = 4;
      if (lp != std::string::npos) {
        m.label[idx] = 5;
        in.erase(in.begin() + lp, in.begin() + rp + 1);
        lp = in.find("(+");
        rp = in.find_last_of(')');
        in.erase(in.begin() + lp, in.begin() + rp + 1);
      } else {
        in.erase(cat, 2);
        cat = in.find("+M");
        in.erase(cat, 2);
      }
    }
    auto ls = in.find('<'), eq = in.find('=');
    if (ls == std::string::npos) ls = eq;
    std::replace(in.begin(), in.begin() + ls + 1, ' ', '<');
    std::string rxn;
    std::istringstream line(in);
    line >> rxn >> m.A[idx] >> m.b[idx] >> m.Ea[idx];
    m.Ea[idx] /= kRc;
    ls = rxn.find('<');
    eq = rxn.find('=');
    auto gt = rxn.find('>');
    if (ls == std::string::npos) {
      if (gt != std::string::npos)
        m.label[idx] = 0;
      else
        gt = eq;
      ls = eq;
    }
    std::string reac = rxn.substr(0, ls), prod = rxn.substr(gt + 1);
    auto parse_side = [&](std::string side, std::vector<int> &stoi, int sign) {
      for (auto plus = side.find('+'); plus != std::string::npos; plus = side.find('+')) {
        if (plus + 1 < side.size() && side[plus + 1] == '+') ++plus;
        int nu = 1;
        std::string name;
        if (std::isdigit((unsigned char)side[0])) {
          nu = std::stoi(side.substr(0, 1));
          side.erase(0, 1);
          name = side.substr(0, plus - 1);
        } else {
          name = side.substr(0, plus);
        }
        stoi[(size_t)idx * ns + sp(name, err)] += nu;
        m.order[idx] += sign * nu;
        side.erase(0, name.size() + 1);
      }
      int nu = 1;
      if (!side.empty() && std::isdigit((unsigned char)side[0])) {
        nu = std::stoi(side.substr(0, 1));
        side.erase(0, 1);
      }
      stoi[(size_t)idx * ns + sp(side, err)] += nu;
      m.order[idx] += sign * nu;
    };
    parse_side(reac, m.stoi_f, -1);
    parse_side(prod, m.stoi_b, +1);
  }

  // ChemData.cpp:634-685 get_auxi_info: returns the next reaction line
  std::string auxiliary(std::ifstream &f, int idx, bool &is_dup, std::string &err) {
    Mechanism &m = w.m;
    std::string in, key;
    while (getline_up(f, in)) {
      std::replace(in.begin(), in.end(), '/', ' ');
      std::istringstream line(in);
      line >> key;
      if (in.empty() || in[0] == '!') continue;
      if (key == "END") break;
      if (in.find('=') != std::string::npos) break;
      if (key == "PLOG") continue;
      if (key == "DUPLICATE" || key == "DUP") {
        is_dup = true;
        m.label[idx] = 3;
        continue;
      }
      if (key == "REV") {
        line >> m.A2[idx] >> m.b2[idx] >> m.Ea2[idx];
        m.label[idx] = 2;
      } else if (key == "LOW") {
        line >> m.A2[idx] >> m.b2[idx] >> m.Ea2[idx];
        continue;
      } else if (key == "TROE") {
        m.label[idx] = 6;
        line >> m.troe_alpha[idx] >> m.troe_t3[idx] >> m.troe_t1[idx];
        std::string t2;
        line >> t2;
        if (!t2.empty()) {
          m.troe_t2[idx] = std::stod(t2);
          m.label[idx] = 7;
        }
      } else {
        std::istringstream tb(in);
        std::string nm;
        while (tb >> nm) {
          double coe{};
          tb >> coe;
          m.third_body_coeff[(size_t)idx * ns + sp(nm, err)] = coe;
        }
      }
    }
    return in;
  }
};

std::string read_reactions(std::ifstream &f, Work &w) {
  Mechanism &m = w.m;
  std::string in;
  bool found = false;
  while (std::getline(f, in)) {
    trim_left(in);
    if (in.rfind("REACTIONS", 0) == 0 || in.rfind("REAC", 0) == 0) {
      found = true;
      break;
    }
  }
  if (!found) return "REACTIONS section not found";
  RxnRead rr{w, m.n_spec};
  rr.resize(100);
  int has_read = 0;
  std::vector<int> dups;
  std::string err;
  getline_up(f, in);
  while (in != "END") {
    if (in.empty() || in[0] == '!') {
      if (!getline_up(f, in)) break;
      continue;
    }
    if (has_read >= m.n_reac) rr.resize(m.n_reac + 100);
    rr.equation(in, has_read, err);
    if (!err.empty()) return err;
    bool is_dup = false;
    in = rr.auxiliary(f, has_read, is_dup, err);
    if (!err.empty()) return err;
    if (is_dup) {
      const int ns = m.n_spec;
      int pos = -1;
      for (size_t r = 0; r < dups.size(); ++r) {
        const int d = dups[r];
        bool same = true;
        for (int s = 0; s < ns; ++s)
          if (m.stoi_f[(size_t)d * ns + s] != m.stoi_f[(size_t)has_read * ns + s] ||
              m.stoi_b[(size_t)d * ns + s] != m.stoi_b[(size_t)has_read * ns + s]) {
            same = false;
            break;
          }
        if (same) {
          pos = (int)r;
          break;
        }
      }
      if (pos >= 0) {
        const int d = dups[pos];
        m.A2[d] = m.A[has_read];
        m.b2[d] = m.b[has_read];
        m.Ea2[d] = m.Ea[has_read];
        dups.erase(dups.begin() + pos);
        for (int s = 0; s < ns; ++s) {
          m.stoi_f[(size_t)has_read * ns + s] = 0;
          m.stoi_b[(size_t)has_read * ns + s] = 0;
        }
        m.label[has_read] = 1;
        m.order[has_read] = 0;
      } else {
        dups.push_back(has_read);
        ++has_read;
      }
      continue;
    }
    ++has_read;
  }
  rr.resize(has_read);
  if (has_read > kMaxReac) return "more reactions than the supported maximum of 50";
  return "";
}

}  // namespace

std::string read_mechanism(const std::string &input_dir, const std::string &mech_file, const std::string &therm_file,
                           const std::string &tran_file, bool with_reactions, Mechanism &m) {
  m = Mechanism();
  Work w{m};
  {
    std::ifstream f(input_dir + "/" + mech_file);
    if (!f) return "cannot open mechanism file " + mech_file;
    auto e = read_species(f, w);
    if (!e.empty()) return e;
  }
  {
    std::ifstream f(input_dir + "/" + therm_file);
    if (!f) return "cannot open therm file " + therm_file;
    auto e = read_therm(f, w);
    if (!e.empty()) return e;
  }
  {
    std::ifstream f(input_dir + "/" + tran_file);
    if (!f) return "cannot open transport file " + tran_file;
    auto e = read_tran(f, w);
    if (!e.empty()) return e;
  }
  if (with_reactions) {
    std::ifstream f(input_dir + "/" + mech_file);
    auto e = read_reactions(f, w);
    if (!e.empty()) return e;
  } else {
    m.n_reac = 0;
  }
  return "";
}

}  // namespace ostc
