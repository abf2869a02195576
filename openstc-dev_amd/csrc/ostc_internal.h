// Internal declarations shared by the host side (setup files, CHEMKIN
// tables, grid metrics, boundary registration, drivers) and the gfx950
// kernels.  Nothing here crosses the C ABI.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <variant>
#include <vector>

namespace ostc {

constexpr double kRu = 8314.462618;       // J/(kmol K)      Constants.h:17
constexpr double kRc = 1.98720425864083;  // cal/(mol K)     Constants.h:19
constexpr double kGammaAir = 1.4;         //                 Constants.h:20
constexpr double kPatm = 101325.0;        //                 Constants.h:22
constexpr double kMwAir = 28.9647;        //                 Constants.h:25
constexpr int kMaxSpec = 20;              // MAX_SPEC_NUMBER (CMakeLists.txt:14)
constexpr int kMaxReac = 50;              // MAX_REAC_NUMBER (CMakeLists.txt:15)

// ------------------------------------------------------------ setup files
// Typed key/value store with the reference's file grammar
// (Parameter.cpp:53-163): "int|real|bool|string key = value",
// "array <type> name { ... }" and "struct name { ... }".
using StructVal = std::variant<std::string, int, double>;
using Struct = std::map<std::string, StructVal>;

struct Params {
  std::map<std::string, int> i;
  std::map<std::string, double> r;
  std::map<std::string, bool> b;
  std::map<std::string, std::string> s;
  std::map<std::string, std::vector<int>> ia;
  std::map<std::string, std::vector<double>> ra;
  std::map<std::string, std::vector<std::string>> sa;
  std::map<std::string, Struct> st;

  void read_file(const std::string &path);
  int get_int(const std::string &k, int dflt) const;
  double get_real(const std::string &k, double dflt) const;
  bool get_bool(const std::string &k, bool dflt) const;
  std::string get_string(const std::string &k, const std::string &dflt) const;
};

// Reads input_files/setup/*.txt in the reference order (Parameter.h:69-82)
// and derives ngg / inviscid_type (Parameter.cpp:165-219).
void read_setup(const std::string &case_dir, Params &p);

// ------------------------------------------------------------ mechanism
// Species and reaction tables with the reference's DParameter content
// (DParameter.cuh:13-92, DParameter.cu:7-159).  Row-major, packed sizes.
struct Mechanism {
  int n_spec = 0, n_reac = 0;
  std::vector<std::string> names;
  std::vector<std::string> elements;
  std::vector<int> elem_comp;  // n_spec x n_elem
  std::vector<double> mw, t_low, t_mid, t_high, hi, lo;  // hi/lo: n_spec x 7
  std::vector<double> LJ_potent_inv, vis_coeff;
  std::vector<double> WjDivWi_to_One4th, sqrt_WiDivWjPl1Mul8, binary_diffusivity_coeff, kb_over_eps_jk;
  std::vector<int> label, stoi_f, stoi_b, order;
  std::vector<double> A, b, Ea, A2, b2, Ea2, third_body_coeff;
  std::vector<double> troe_alpha, troe_t3, troe_t1, troe_t2;
};

// Atomic weight of an element symbol (upper case), -1 if unknown.
double atomic_weight(const std::string &e);

// ChemData.cpp:11-685 semantics.  Paths are relative to <case>/input_files/.
// Returns an empty string on success or an error message.
std::string read_mechanism(const std::string &input_dir, const std::string &mech_file, const std::string &therm_file,
                           const std::string &tran_file, bool with_reactions, Mechanism &m);

// ------------------------------------------------------------ grid
// Reference ghosted layout (gxl_lib/Array.hpp:141-155):
//   p(i,j,k) = (k+g)*d1 + (j+g)*d2 + (i+g), component stride sz.
struct Layout {
  int g = 0;
  int64_t d2 = 0, d1 = 0, sz = 0;
  Layout() = default;
  Layout(int mx, int my, int mz, int gg)
      : g(gg), d2(mx + 2 * gg), d1((int64_t)(my + 2 * gg) * (mx + 2 * gg)),
        sz((int64_t)(my + 2 * gg) * (mx + 2 * gg) * (mz + 2 * gg)) {}
  int64_t at(int i, int j, int k) const { return (int64_t)(k + g) * d1 + (int64_t)(j + g) * d2 + (i + g); }
};

// Ghost nodes (Mesh.cpp:790-944) and Jacobian/metrics (Mesh.cpp:259-521).
// x,y,z interior node coordinates with i fastest.  Outputs: jac (Lg) and
// metric as 9 component-major arrays (metric[(r*3+c)*Lg.sz + p] =
// d xi_r / d x_c), both on the ngg-ghosted layout; xyz (optional) gets the
// ghosted node coordinates.  conservative: the rows in the conservative
// symmetric form that keeps a uniform flow exact, with the differences of
// the AWENO-5 scheme when aweno is set.  Returns false on a non-positive
// Jacobian in the interior.
bool compute_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                     std::vector<double> &jac, std::vector<double> &metric, std::string &err,
                     std::vector<double> *xyz = nullptr, bool conservative = false, bool aweno = false);

// Ghost node steps of Mesh.cpp:790-944 on an (ngg + 1)-ghosted node layout:
// reflected faces first, then (after the inner faces' ghost nodes were taken
// from the neighbours) the edges and corners.
void ghost_node_faces(int mx, int my, int mz, int ngg, const Layout &L, double *A, bool is_z);
void ghost_node_edges(int mx, int my, int mz, int ngg, const Layout &L, double *A);

// An inner face after Mesh.cpp:34-187 InnerFace::register_boundary, 0-based:
// this block's range (the face layer and its ngg ghost layers, full
// tangential extent), the target block's start indices, which of this
// block's directions each target direction follows (src_tar) and the index
// steps on both sides.
struct InnerFace {
  int rs[3] = {0, 0, 0}, re[3] = {0, 0, 0};
  int face = 0, dir = 1;
  int ts[3] = {0, 0, 0}, te[3] = {0, 0, 0};
  int tblk = 0, tface = 0, tdir = 1;
  int src_tar[3] = {0, 0, 0};
  int loop[3] = {1, 1, 1}, tloop[3] = {1, 1, 1};
  int n[3] = {0, 0, 0};
};
// s7 / t7: the two lines of an inner-face pair as in the reference's file
// (1-based; t7[6] the target block, 1-based).
bool register_inner_face(const int *s7, const int *t7, int ngg, int dim, InnerFace &f, std::string &err);
// Mesh.cpp:946-994: ghost nodes beyond face f of block B from neighbour N
// (B[3], N[3]: x, y, z arrays on the layouts LB, LN).
void inner_ghost_nodes(const InnerFace &f, const Layout &LB, double *const *B, const Layout &LN,
                       const double *const *N);

// ------------------------------------------------------------ field files
// Tecplot binary (V112) with ordered zones and nodal double data, the
// reference's output and restart format (FieldIO.h, Initialize.cuh).
struct PltZone {
  int mx = 0, my = 0, mz = 0;
  std::vector<const double *> vars;  // one array of mx*my*mz values (i fastest) per variable
};
struct PltFile {
  std::vector<std::string> names;
  std::vector<std::array<int, 3>> dims;
  std::vector<std::vector<std::vector<double>>> data;  // zone -> variable -> values
  double time = 0;                                     // solution time (of the last zone)
};
bool plt_write(const std::string &path, const std::vector<std::string> &names, const std::vector<PltZone> &zones,
               double time, std::string &err);
bool plt_read(const std::string &path, PltFile &out, std::string &err);

// ------------------------------------------------------------ boundaries
enum BcType { BC_WALL = 2, BC_SYMMETRY = 3, BC_FARFIELD = 4, BC_INFLOW = 5, BC_OUTFLOW = 6, BC_SUBSONIC = 7,
              BC_BACKPRESSURE = 9 };

struct BoundaryFace {
  int raw[6] = {0, 0, 0, 0, 0, 0};  // i1 i2 j1 j2 k1 k2 from the boundary file
  int label = 0;
  int type = 0;                      // resolved BcType
  int face = 0, direction = 1;
  int rs[3] = {0, 0, 0}, re[3] = {0, 0, 0};  // registered (and wall-trimmed) ranges
  int state = -1;                    // index into the solver's inflow / wall tables
};

}  // namespace ostc
