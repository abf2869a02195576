// Dumps the species / reaction tables produced by the reference's own CHEMKIN,
// therm.dat and tran.dat readers (ChemData.cpp, Element.cpp, Parameter.cpp and
// gxl_lib/MyString.cpp, compiled unmodified from /root/reference by
// oracle/build_ref.sh).  The dump is the golden record the oracle's chemistry
// tables and the product's parser are both checked against.
//
// usage: ref_chem_dump <case_root> <setup_file>
//   case_root  directory holding input_files/chemistry/...
//   setup_file reference-format setup file naming species/reaction/mechanism keys
//
// Output (stdout): one "key n v0 v1 ..." line per table, doubles printed %.17g.
#include "ChemData.h"
#include "Parameter.h"
#include <cstdio>
#include <string>
#include <vector>
#include <unistd.h>

static void put(const char *key, const std::vector<double> &v) {
  std::printf("%s %zu", key, v.size());
  for (double x : v) std::printf(" %.17g", x);
  std::printf("\n");
}

static void put_i(const char *key, const std::vector<int> &v) {
  std::printf("%s %zu", key, v.size());
  for (int x : v) std::printf(" %d", x);
  std::printf("\n");
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <case_root> <setup_file>\n", argv[0]);
    return 2;
  }
  const std::string setup = argv[2];
  if (chdir(argv[1]) != 0) {
    std::perror("chdir");
    return 2;
  }
  cfd::Parameter parameter(setup);
  cfd::Species spec(parameter);
  cfd::Reaction reac(parameter, spec);
  const int ns = spec.n_spec, nr = reac.n_reac;
  std::printf("n_spec 1 %d\nn_reac 1 %d\n", ns, nr);
  // species names in index order
  std::vector<std::string> names(ns);
  for (auto &[name, idx] : spec.spec_list) names[idx] = name;
  std::printf("names %d", ns);
  for (auto &n : names) std::printf(" %s", n.c_str());
  std::printf("\n");
  put("mw", spec.mw);
  put("t_low", spec.t_low);
  put("t_mid", spec.t_mid);
  put("t_high", spec.t_high);
  std::vector<double> hi(ns * 7), lo(ns * 7);
  for (int i = 0; i < ns; ++i)
    for (int c = 0; c < 7; ++c) {
      hi[i * 7 + c] = spec.high_temp_coeff(i, c);
      lo[i * 7 + c] = spec.low_temp_coeff(i, c);
    }
  put("high_coeff", hi);
  put("low_coeff", lo);
  put("LJ_potent_inv", spec.LJ_potent_inv);
  put("vis_coeff", spec.vis_coeff);
  std::vector<double> a(ns * ns), b(ns * ns), c(ns * ns), d(ns * ns);
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      a[i * ns + j] = spec.WjDivWi_to_One4th(i, j);
      b[i * ns + j] = spec.sqrt_WiDivWjPl1Mul8(i, j);
      c[i * ns + j] = spec.binary_diffusivity_coeff(i, j);
      d[i * ns + j] = spec.kb_over_eps_jk(i, j);
    }
  put("WjDivWi_to_One4th", a);
  put("sqrt_WiDivWjPl1Mul8", b);
  put("binary_diffusivity_coeff", c);
  put("kb_over_eps_jk", d);
  std::vector<int> elem(ns * (int)spec.elem_list.size());
  const int ne = (int)spec.elem_list.size();
  for (int i = 0; i < ns; ++i)
    for (int e = 0; e < ne; ++e) elem[i * ne + e] = spec.elem_comp(i, e);
  std::printf("elements %d", ne);
  std::vector<std::string> enames(ne);
  for (auto &[name, idx] : spec.elem_list) enames[idx] = name;
  for (auto &n : enames) std::printf(" %s", n.c_str());
  std::printf("\n");
  put_i("elem_comp", elem);
  put_i("label", reac.label);
  std::vector<int> sf(nr * ns), sb(nr * ns);
  std::vector<double> tb(nr * ns);
  for (int r = 0; r < nr; ++r)
    for (int s = 0; s < ns; ++s) {
      sf[r * ns + s] = reac.stoi_f(r, s);
      sb[r * ns + s] = reac.stoi_b(r, s);
      tb[r * ns + s] = reac.third_body_coeff(r, s);
    }
  put_i("stoi_f", sf);
  put_i("stoi_b", sb);
  put_i("order", reac.order);
  put("A", reac.A);
  put("b", reac.b);
  put("Ea", reac.Ea);
  put("A2", reac.A2);
  put("b2", reac.b2);
  put("Ea2", reac.Ea2);
  put("third_body_coeff", tb);
  put("troe_alpha", reac.troe_alpha);
  put("troe_t3", reac.troe_t3);
  put("troe_t1", reac.troe_t1);
  put("troe_t2", reac.troe_t2);
  return 0;
}
