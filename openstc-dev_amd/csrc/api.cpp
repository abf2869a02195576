// C ABI of the solver (include/ostc.h): set-up from the reference's input
// files, device zones in the DZone layout, the stage operations and the
// fused RK3 driver (RK.cuh:19-195 without IO), all stream-ordered on one
// HIP stream per solver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ostc.h"
#include "device_common.h"
#include "flux_kernels.h"
#include "ostc_internal.h"
#include "surrogate.h"

using namespace ostc;

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
  g_err = msg;
  return 1;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

struct Inflow {
  std::string name;
  int label = 5;
  double rho = 0, u = 0, v = 0, w = 0, p = 0, t = 0, vel = 0, mach = 0, mw = kMwAir;
  double mut = 0;  // SST: turbulent viscosity of the state
  std::vector<double> sv;
};

struct WallBC {
  int label = 2;
  bool isothermal = true;
  double t = 300;
};

struct Block {
  int mx = 0, my = 0, mz = 0;
  std::vector<double> jac, metric;  // host copies (ghosted, component-major metric) of a lone block
  std::vector<double> xyz;          // ghosted node coordinates (SST wall distance only)
  std::vector<double> x0, y0, z0;   // interior node coordinates, i fastest
  std::vector<InnerFace> inner;     // inner faces to other blocks, in file order
  std::vector<BoundaryFace> faces;
  std::vector<BCDesc> bcs;          // in application order
  ZoneView z{};
  std::vector<void *> mem;
};

struct Prof {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace

struct ostc_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string case_dir;
  Params prm;
  Mechanism mech;
  DMech *d_mech = nullptr;
  DParams P{};
  int ngg = 2, n_var = 5, n_spec = 0, n_scalar = 0, n_turb = 0;
  std::vector<Block> blocks;
  std::vector<Inflow> inflows;
  std::vector<WallBC> walls;
  bool finalized = false;
  bool fp32 = false;
  StepScratch w{};
  double t_phys = 0;
  int step = 0;
  bool prof = false;
  std::vector<Prof> pending;
  std::map<std::string, std::pair<long, double>> stats;
  std::vector<void *> user_mem;
  SurrogateState *sur = nullptr;
  bool chem_split = false;          // finite-rate source left out of the RHS (operator-split chemistry)
  double steady_chem_dt = 0;        // steady path: inline chemistry step per pseudo-iteration (0: off)
  int steady_chem_mode = -1;        // 0 FP64 point-implicit, 1 surrogate + fallback
  int64_t steady_cells = 0, steady_fb = 0;  // inline chemistry cells and fallbacks since it was set
  bool res_scale_set = false;       // a step-1 residual has set the residual scale
  bool restart = false;             // fields, step and time came from ostc_read_flowfield
  bool conservative_metric = false; // conservative symmetric metrics (uniform flow kept exactly)
  int restart_step = 0;
  double restart_time = 0;
  double *chem_buf = nullptr;       // gather / result buffers of ostc_chem_step (21 doubles per cell)
  int64_t chem_cap = 0;

  ~ostc_solver() {
    for (auto &b : blocks)
      for (void *p : b.mem) (void)hipFree(p);
    for (void *p : user_mem) (void)hipFree(p);
    if (chem_buf) (void)hipFree(chem_buf);
    if (d_mech) (void)hipFree(d_mech);
    (void)hipFree(w.partial);
    (void)hipFree(w.counter);
    (void)hipFree(w.dt);
    (void)hipFree(w.dt_hist);
    (void)hipFree(w.res_hist);
    (void)hipFree(w.res_scale);
    (void)hipFree(w.res_one);
    (void)hipFree(w.blk_dt);
    for (auto &p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    surrogate_destroy(sur);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // RAII timing of one launch group
  struct Timer {
    ostc_solver *s;
    Prof p{};
    bool on;
    Timer(ostc_solver *sv, const char *name) : s(sv), on(sv->prof) {
      if (!on) return;
      p.name = name;
      (void)hipEventCreate(&p.a);
      (void)hipEventCreate(&p.b);
      (void)hipEventRecord(p.a, s->stream);
    }
    ~Timer() {
      if (!on) return;
      (void)hipEventRecord(p.b, s->stream);
      s->pending.push_back(p);
    }
  };

  void drain_profile() {
    if (pending.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (auto &p : pending) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, p.a, p.b);
      auto &e = stats[p.name];
      e.first += 1;
      e.second += ms;
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending.clear();
  }
};

namespace {

// Transport.cu:6-38 Sutherland / compute_viscosity (Wilke's rule) on the host
double host_viscosity(const ostc_solver *s, double t, double mw, const double *Y) {
  if (s->n_spec == 0) return 1.716e-5 * std::pow(t / 273, 1.5) * (273 + 111) / (t + 111);
  const Mechanism &m = s->mech;
  const int ns = m.n_spec;
  std::vector<double> x(ns), vis(ns);
  for (int i = 0; i < ns; ++i) {
    x[i] = Y[i] * mw / m.mw[i];
    const double tdl = t * m.LJ_potent_inv[i];
    const double ci = 1.147 * std::pow(tdl, -0.145) + std::pow(tdl + 0.5, -2);
    vis[i] = m.vis_coeff[i] * std::sqrt(t) / ci;
  }
  double mu = 0;
  for (int i = 0; i < ns; ++i) {
    double vt = 0;
    for (int j = 0; j < ns; ++j) {
      double phi = 1.0;
      if (i != j) {
        const double num = 1 + std::sqrt(vis[i] / vis[j]) * m.WjDivWi_to_One4th[i * ns + j];
        phi = num * num * m.sqrt_WiDivWjPl1Mul8[i * ns + j];
      }
      vt += phi * x[j];
    }
    mu += vis[i] * x[i] / vt;
  }
  return mu;
}

// BoundCond.cpp:83-95 / 269-280: SST state of an inflow or farfield struct:
// mu_t = ratio mu; with a turbulence intensity I, k = 1.5 (|V| I)^2 and
// omega = rho k / mu_t (k, omega stay 0 without one)
int turbulence_state(const ostc_solver *s, const Struct &info, const std::string &name, double viscosity,
                     double velocity, double density, double &mut, double *sv) {
  if (s->n_turb == 0) return 0;
  auto get = [&](const char *k, double &v) {
    auto f = info.find(k);
    if (f == info.end()) return false;
    if (std::holds_alternative<double>(f->second))
      v = std::get<double>(f->second);
    else if (std::holds_alternative<int>(f->second))
      v = std::get<int>(f->second);
    return true;
  };
  double ratio = 0, inten = 0;
  if (!get("turb_viscosity_ratio", ratio)) return fail("'" + name + "' needs turb_viscosity_ratio for SST");
  mut = ratio * viscosity;
  if (get("turbulence_intensity", inten)) {
    sv[s->n_spec] = 1.5 * velocity * velocity * inten * inten;
    sv[s->n_spec + 1] = density * sv[s->n_spec] / mut;
  }
  return 0;
}

// BoundCond.cpp:7-122 Inflow::Inflow for the keys a struct may carry.
int make_inflow(ostc_solver *s, const std::string &name, Inflow &in) {
  auto it = s->prm.st.find(name);
  if (it == s->prm.st.end()) return fail("setup struct '" + name + "' not found");
  const Struct &info = it->second;
  auto get_r = [&](const char *k, double &v) {
    auto f = info.find(k);
    if (f == info.end()) return false;
    if (std::holds_alternative<double>(f->second))
      v = std::get<double>(f->second);
    else if (std::holds_alternative<int>(f->second))
      v = std::get<int>(f->second);
    return true;
  };
  in.name = name;
  auto lf = info.find("label");
  if (lf != info.end() && std::holds_alternative<int>(lf->second)) in.label = std::get<int>(lf->second);
  double mach = -1, pressure = 101325, temperature = -1, velocity = 0, density = -1, u = 1, v = 0, w = 0;
  get_r("mach", mach);
  get_r("pressure", pressure);
  get_r("temperature", temperature);
  get_r("velocity", velocity);
  get_r("density", density);
  get_r("u", u);
  get_r("v", v);
  get_r("w", w);
  const int ns = s->n_spec;
  in.sv.assign(std::max(s->n_scalar, 0), 0.0);
  double mw = kMwAir;
  if (ns > 0) {
    for (int l = 0; l < ns; ++l) get_r(s->mech.names[l].c_str(), in.sv[l]);
    double acc = 0;
    for (int l = 0; l < ns; ++l) acc += in.sv[l] / s->mech.mw[l];
    mw = 1 / acc;
  }
  if (temperature < 0) temperature = pressure * mw / (density * kRu);
  double gamma = kGammaAir;
  if (ns > 0) {
    double cp = 0, cv = 0;
    for (int l = 0; l < ns; ++l) {
      // Species::compute_cp (ChemData.cpp:110-128)
      const double tt = temperature < s->mech.t_low[l] ? s->mech.t_low[l] : temperature;
      const double *c = (temperature < s->mech.t_low[l] || tt < s->mech.t_mid[l]) ? &s->mech.lo[l * 7]
                                                                                  : &s->mech.hi[l * 7];
      double cpi;
      if (temperature < s->mech.t_low[l]) {
        const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
        cpi = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      } else {
        const double t2 = temperature * temperature, t3 = t2 * temperature, t4 = t3 * temperature;
        cpi = c[0] + c[1] * tt + c[2] * t2 + c[3] * t3 + c[4] * t4;
      }
      cpi *= kRu / s->mech.mw[l];
      cp += in.sv[l] * cpi;
      cv += in.sv[l] * (cpi - kRu / s->mech.mw[l]);
    }
    gamma = cp / cv;
  }
  const double c = std::sqrt(gamma * kRu / mw * temperature);
  if (mach < 0)
    mach = velocity / c;
  else
    velocity = mach * c;
  u *= velocity;
  v *= velocity;
  w *= velocity;
  if (density < 0) density = pressure * mw / (kRu * temperature);
  in.rho = density;
  in.u = u;
  in.v = v;
  in.w = w;
  in.p = pressure;
  in.t = temperature;
  in.vel = velocity;
  in.mach = mach;
  in.mw = mw;
  if (turbulence_state(s, info, name, host_viscosity(s, temperature, mw, in.sv.data()), velocity, density, in.mut,
                       in.sv.data()))
    return 1;
  return 0;
}

// BoundCond.cpp:184-297 FarField::FarField: (mach, pressure, temperature),
// (density, velocity, pressure) or (mach, temperature, Reynolds number);
// also a, gamma and the entropy p / rho^gamma of the far state.
int make_farfield(ostc_solver *s, const std::string &name, BCDesc &d) {
  auto it = s->prm.st.find(name);
  if (it == s->prm.st.end()) return fail("setup struct '" + name + "' not found");
  const Struct &info = it->second;
  auto get_r = [&](const char *k, double &v) {
    auto f = info.find(k);
    if (f == info.end()) return false;
    if (std::holds_alternative<double>(f->second))
      v = std::get<double>(f->second);
    else if (std::holds_alternative<int>(f->second))
      v = std::get<int>(f->second);
    return true;
  };
  double mach = -1, pressure = -1, temperature = -1, velocity = 0, density = -1, u = 1, v = 0, w = 0, re = -1;
  get_r("mach", mach);
  get_r("pressure", pressure);
  get_r("temperature", temperature);
  get_r("velocity", velocity);
  get_r("density", density);
  get_r("u", u);
  get_r("v", v);
  get_r("w", w);
  get_r("reynolds_number", re);
  const int ns = s->n_spec;
  std::vector<double> sv(std::max(s->n_scalar, 0), 0.0);
  double mw = kMwAir;
  if (ns > 0) {
    for (int l = 0; l < ns; ++l) get_r(s->mech.names[l].c_str(), sv[l]);
    double acc = 0;
    for (int l = 0; l < ns; ++l) acc += sv[l] / s->mech.mw[l];
    mw = 1 / acc;
  }
  if (temperature < 0) temperature = pressure * mw / (density * kRu);
  double gamma = kGammaAir;
  if (ns > 0) {
    double cp = 0, cv = 0;
    for (int l = 0; l < ns; ++l) {
      const double tt = temperature < s->mech.t_low[l] ? s->mech.t_low[l] : temperature;
      const double *c = (temperature < s->mech.t_low[l] || tt < s->mech.t_mid[l]) ? &s->mech.lo[l * 7]
                                                                                  : &s->mech.hi[l * 7];
      const double cpi = (c[0] + c[1] * tt + c[2] * tt * tt + c[3] * tt * tt * tt + c[4] * tt * tt * tt * tt) * kRu /
                         s->mech.mw[l];
      cp += sv[l] * cpi;
      cv += sv[l] * (cpi - kRu / s->mech.mw[l]);
    }
    gamma = cp / cv;
  }
  const double a = std::sqrt(gamma * kRu / mw * temperature);
  if (mach < 0)
    mach = velocity / a;
  else
    velocity = mach * a;
  u *= velocity;
  v *= velocity;
  w *= velocity;
  if (pressure < 0) {
    if (ns > 0) return fail("farfield '" + name + "': the Reynolds-number form is supported for air only");
    const double mu = 1.716e-5 * std::pow(temperature / 273, 1.5) * (273 + 111) / (temperature + 111);
    density = mu * re / velocity;
    pressure = density * temperature * kRu / mw;
  }
  if (density < 0) density = pressure * mw / (kRu * temperature);
  if (turbulence_state(s, info, name, host_viscosity(s, temperature, mw, sv.data()), velocity, density, d.mut,
                       sv.data()))
    return 1;
  d.rho = density;
  d.u = u;
  d.v = v;
  d.w = w;
  d.p = pressure;
  d.t = temperature;
  d.vel = velocity;
  for (size_t l = 0; l < sv.size() && l < (size_t)kMS + 4; ++l) d.sv[l] = sv[l];
  d.a_inf = a;
  d.gamma_inf = gamma;
  d.entropy = pressure / std::pow(density, gamma);
  d.mw_inf = mw;
  return 0;
}

// p_inf, T_inf, rho_inf and v_inf of the reference state (the inflow or
// farfield struct it names)
int reference_pt(ostc_solver *s, double &p_inf, double &t_inf, double *rho_inf = nullptr, double *v_inf = nullptr) {
  const std::string ref = s->prm.get_string("reference_state", "");
  auto it = s->prm.st.find(ref);
  if (it == s->prm.st.end()) return fail("reference_state struct '" + ref + "' not found");
  auto tf = it->second.find("type");
  const std::string type = tf == it->second.end() ? "inflow" : std::get<std::string>(tf->second);
  if (type == "farfield") {
    BCDesc d{};
    if (make_farfield(s, ref, d)) return 1;
    p_inf = d.p;
    t_inf = d.t;
    if (rho_inf) *rho_inf = d.rho;
    if (v_inf) *v_inf = d.vel;
  } else {
    Inflow in;
    if (make_inflow(s, ref, in)) return 1;
    p_inf = in.p;
    t_inf = in.t;
    if (rho_inf) *rho_inf = in.rho;
    if (v_inf) *v_inf = in.vel;
  }
  return 0;
}

int upload_mech(ostc_solver *s) {
  DMech h{};
  const Mechanism &m = s->mech;
  h.n_spec = m.n_spec;
  h.n_reac = m.n_reac;
  for (int i = 0; i < m.n_spec; ++i) {
    h.mw[i] = m.mw[i];
    h.mw_inv[i] = 1.0 / m.mw[i];
    h.r_mw[i] = kRu / m.mw[i];
    h.t_low[i] = m.t_low[i];
    h.t_mid[i] = m.t_mid[i];
    for (int c = 0; c < 7; ++c) {
      h.hi[i * 7 + c] = m.hi[i * 7 + c];
      h.lo[i * 7 + c] = m.lo[i * 7 + c];
    }
    h.LJ_potent_inv[i] = m.LJ_potent_inv[i];
    h.log_LJ_potent_inv[i] = m.LJ_potent_inv[i] > 0 ? std::log(m.LJ_potent_inv[i]) : 0.0;
    h.vis_coeff[i] = m.vis_coeff[i];
  }
  const int ns = m.n_spec, nr = m.n_reac;
  for (int i = 0; i < ns * ns; ++i) {
    h.WjDivWi_to_One4th[i] = m.WjDivWi_to_One4th[i];
    h.sqrt_WiDivWjPl1Mul8[i] = m.sqrt_WiDivWjPl1Mul8[i];
    h.binary_diffusivity_coeff[i] = m.binary_diffusivity_coeff[i];
    h.kb_over_eps_jk[i] = m.kb_over_eps_jk[i];
    h.log_kb_over_eps_jk[i] = m.kb_over_eps_jk[i] > 0 ? std::log(m.kb_over_eps_jk[i]) : 0.0;
  }
  for (int r = 0; r < nr; ++r) {
    h.label[r] = m.label[r];
    h.order[r] = m.order[r];
    h.A[r] = m.A[r];
    h.b[r] = m.b[r];
    h.Ea[r] = m.Ea[r];
    h.A2[r] = m.A2[r];
    h.b2[r] = m.b2[r];
    h.Ea2[r] = m.Ea2[r];
    h.troe_alpha[r] = m.troe_alpha[r];
    h.troe_t3[r] = m.troe_t3[r];
    h.troe_t1[r] = m.troe_t1[r];
    h.troe_t2[r] = m.troe_t2[r];
  }
  for (int i = 0; i < nr * ns; ++i) {
    h.stoi_f[i] = m.stoi_f[i];
    h.stoi_b[i] = m.stoi_b[i];
    h.third_body_coeff[i] = m.third_body_coeff[i];
  }
  HIPCHK(hipMalloc(&s->d_mech, sizeof(DMech)));
  HIPCHK(hipMemcpy(s->d_mech, &h, sizeof(DMech), hipMemcpyHostToDevice));
  return 0;
}

struct FieldDesc {
  double *ptr;
  int64_t count;
  int ghosts, ncomp;
};

FieldDesc field_desc(ostc_solver *s, Block &b, int f) {
  ZoneView &z = b.z;
  const int64_t sg = z.sz, s0 = z.sz0;
  switch (f) {
    case OSTC_BV: return {z.bv, sg * 6, z.g, 6};
    case OSTC_SV: return {z.sv, sg * s->n_scalar, z.g, s->n_scalar};
    case OSTC_CV: return {z.cv, sg * s->n_var, z.g, s->n_var};
    case OSTC_DQ: return {z.dq, s0 * s->n_var, 0, s->n_var};
    case OSTC_VEL: return {z.vel, sg, z.g, 1};
    case OSTC_ACOUSTIC: return {z.acoustic, sg, z.g, 1};
    case OSTC_MACH: return {z.mach, sg, z.g, 1};
    case OSTC_MUL: return {z.mul, sg, z.g, 1};
    case OSTC_COND: return {z.cond, sg, z.g, 1};
    case OSTC_RHOD: return {z.rhoD, sg * s->n_spec, z.g, s->n_spec};
    case OSTC_GAMMA: return {z.gam, s->n_spec ? sg : 0, z.g, 1};
    case OSTC_CP: return {z.cp, s->n_spec ? sg : 0, z.g, 1};
    case OSTC_DT_LOCAL: return {z.dt_local, s0, 0, 1};
    case OSTC_QN: return {z.qn, sg * s->n_var, z.g, s->n_var};
    case OSTC_BV_LAST: return {z.bv_last, s0 * 4, 0, 4};
    case OSTC_JAC: return {const_cast<double *>(z.jac), sg, z.g, 1};
    case OSTC_METRIC: return {const_cast<double *>(z.metric), sg * 9, z.g, 9};
    case OSTC_CHEM_JAC: return {z.chem_jac, s->P.chemSrcMethod == 1 ? s0 * s->n_spec * s->n_spec
                                            : (s->P.chemSrcMethod == 2 ? s0 * s->n_spec : 0), 0,
                                s->P.chemSrcMethod == 2 ? s->n_spec : s->n_spec * s->n_spec};
    case OSTC_MUT: return {z.mut, z.mut ? sg : 0, z.g, 1};
    case OSTC_TURB_COND: return {z.tcond, z.tcond ? sg : 0, z.g, 1};
    case OSTC_WALL_DIST: return {z.wdist, z.wdist ? sg : 0, z.g, 1};
    case OSTC_TURB_JAC: return {z.turb_jac, z.turb_jac ? s0 * 2 : 0, 0, 2};
  }
  return {nullptr, -1, 0, 0};
}

// WallDistance.cuh:9-81: the nodes of every wall range (registered and
// trimmed) of every block are the wall points; each block's ghosted points
// get their distance on the device.
int compute_wall_distance(ostc_solver *s) {
  std::vector<double> wx, wy, wz;
  for (auto &b : s->blocks) {
    const Layout L(b.mx, b.my, b.mz, s->ngg);
    for (auto &f : b.faces) {
      if (f.type != BC_WALL) continue;
      for (int k = f.rs[2]; k <= f.re[2]; ++k)
        for (int j = f.rs[1]; j <= f.re[1]; ++j)
          for (int i = f.rs[0]; i <= f.re[0]; ++i) {
            const int64_t p = L.at(i, j, k);
            wx.push_back(b.xyz[p]);
            wy.push_back(b.xyz[L.sz + p]);
            wz.push_back(b.xyz[2 * L.sz + p]);
          }
    }
  }
  const int nw = (int)wx.size();
  std::vector<double> wall(3 * (size_t)std::max(nw, 1));
  std::copy(wx.begin(), wx.end(), wall.begin());
  std::copy(wy.begin(), wy.end(), wall.begin() + nw);
  std::copy(wz.begin(), wz.end(), wall.begin() + 2 * nw);
  double *d_wall = nullptr, *d_xyz = nullptr;
  HIPCHK(hipMalloc(&d_wall, wall.size() * sizeof(double)));
  HIPCHK(hipMemcpyAsync(d_wall, wall.data(), wall.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  int rc = 0;
  for (auto &b : s->blocks) {
    if (b.xyz.size() != (size_t)b.z.sz * 3) {
      rc = fail("node coordinates of a block are missing");
      break;
    }
    hipError_t e = hipMalloc(&d_xyz, b.xyz.size() * sizeof(double));
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_xyz, b.xyz.data(), b.xyz.size() * sizeof(double), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = launch_wall_distance(b.z, d_xyz, d_wall, nw, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_xyz);
    d_xyz = nullptr;
    if (e != hipSuccess) {
      rc = fail(std::string("wall distance: ") + hipGetErrorString(e));
      break;
    }
  }
  (void)hipFree(d_wall);
  return rc;
}

int check_blk(ostc_solver *s, int blk) {
  if (!s) return fail("null solver");
  if (blk < 0 || blk >= (int)s->blocks.size()) return fail("block index out of range");
  if (!s->finalized) return fail("ostc_finalize has not been called");
  return 0;
}

// Boundary::register_boundary (Mesh.cpp:14-31)
void register_face(BoundaryFace &b, int ngg, int dim) {
  int s[3] = {b.raw[0], b.raw[2], b.raw[4]}, e[3] = {b.raw[1], b.raw[3], b.raw[5]};
  b.face = 0;
  b.direction = 1;
  for (int d = 0; d < dim; ++d)
    if (s[d] == e[d]) {
      b.face = d;
      if (s[d] == 0) b.direction = -1;
      break;
    }
  for (int d = 0; d < dim; ++d) {
    if (d == b.face) continue;
    if (e[d] < s[d]) std::swap(s[d], e[d]);
    s[d] -= ngg;
    e[d] += ngg;
  }
  if (dim == 2) {
    s[2] = -ngg;
    e[2] = ngg;
  }
  for (int d = 0; d < 3; ++d) {
    b.rs[d] = s[d];
    b.re[d] = e[d];
  }
}

// Ghosted node coordinates ((ngg + 1) layers) of every block, with the
// ghosts beyond inner faces taken from the neighbour (Mesh.cpp:790-944), and
// a check that every inner face's ranges lie inside both blocks.
int build_nodes(ostc_solver *s, std::vector<std::array<std::vector<double>, 3>> &nodes) {
  const int g = s->ngg;
  const size_t nb = s->blocks.size();
  nodes.resize(nb);
  std::vector<Layout> Lc(nb);
  for (size_t bi = 0; bi < nb; ++bi) {
    Block &b = s->blocks[bi];
    Lc[bi] = Layout(b.mx, b.my, b.mz, g + 1);
    const double *src[3] = {b.x0.data(), b.y0.data(), b.z0.data()};
    for (int c = 0; c < 3; ++c) {
      auto &A = nodes[bi][c];
      A.assign(Lc[bi].sz, 0.0);
      for (int k = 0; k < b.mz; ++k)
        for (int j = 0; j < b.my; ++j)
          for (int i = 0; i < b.mx; ++i) A[Lc[bi].at(i, j, k)] = src[c][((size_t)k * b.my + j) * b.mx + i];
      ghost_node_faces(b.mx, b.my, b.mz, g, Lc[bi], A.data(), c == 2);
    }
  }
  auto inside = [&](const Block &b, const int *ijk, int lo_pad, int hi_pad) {
    const int m[3] = {b.mx, b.my, b.mz};
    for (int d = 0; d < 3; ++d)
      if (ijk[d] < -lo_pad || ijk[d] >= m[d] + hi_pad) return false;
    return true;
  };
  for (size_t bi = 0; bi < nb; ++bi) {
    const Block &b = s->blocks[bi];
    for (size_t q = 0; q < b.inner.size(); ++q) {
      const InnerFace &f = b.inner[q];
      const Block &t = s->blocks[f.tblk];
      // the exchange walks f.n points per direction from rs (this block) and
      // ts (target); the node copy reaches one layer further out
      for (int corner = 0; corner < 8; ++corner) {
        int d[3], a[3], at[3];
        for (int x = 0; x < 3; ++x) d[x] = ((corner >> x) & 1) ? f.loop[x] * (f.n[x] - 1) : 0;
        for (int x = 0; x < 3; ++x) {
          a[x] = f.rs[x] + d[x];
          at[x] = f.ts[x] + f.tloop[x] * d[f.src_tar[x]];
        }
        if (!inside(b, a, g, g) || !inside(t, at, g, g))
          return fail("inner face " + std::to_string(q) + " of block " + std::to_string(bi) +
                      " does not fit the blocks it joins");
        if (f.n[f.src_tar[f.tface]] != f.n[f.face] || b.mz != t.mz)
          return fail("inner face " + std::to_string(q) + " of block " + std::to_string(bi) + " is inconsistent");
      }
      double *B[3] = {nodes[bi][0].data(), nodes[bi][1].data(), nodes[bi][2].data()};
      const double *N[3] = {nodes[f.tblk][0].data(), nodes[f.tblk][1].data(), nodes[f.tblk][2].data()};
      inner_ghost_nodes(f, Lc[bi], B, Lc[f.tblk], N);
    }
  }
  for (size_t bi = 0; bi < nb; ++bi) {
    Block &b = s->blocks[bi];
    for (int c = 0; c < 3; ++c) ghost_node_edges(b.mx, b.my, b.mz, g, Lc[bi], nodes[bi][c].data());
  }
  return 0;
}

// inner-face exchange of every block in order (DataCommunication.cuh:27-56)
int exchange_all(ostc_solver *s, bool with_cv) {
  bool any = false;
  for (auto &b : s->blocks) any = any || !b.inner.empty();
  if (!any) return 0;
  ostc_solver::Timer t(s, "inner_exchange");
  for (auto &b : s->blocks)
    for (const InnerFace &f : b.inner) {
      InnerFaceDesc d;
      for (int x = 0; x < 3; ++x) {
        d.rs[x] = f.rs[x];
        d.ts[x] = f.ts[x];
        d.src_tar[x] = f.src_tar[x];
        d.loop[x] = f.loop[x];
        d.tloop[x] = f.tloop[x];
        d.n[x] = f.n[x];
      }
      d.face = f.face;
      d.face_at = f.dir > 0 ? f.rs[f.face] : f.re[f.face];
      HIPCHK(launch_inner_exchange(b.z, s->blocks[f.tblk].z, s->P, s->d_mech, d, with_cv, s->stream));
    }
  return 0;
}

int run_bcs(ostc_solver *s, Block &b, bool with_cv) {
  ostc_solver::Timer t(s, "boundary");
  for (auto &bc : b.bcs) HIPCHK(launch_bc(b.z, s->P, s->d_mech, bc, with_cv, s->stream));
  return 0;
}

// SourceTerm.cuh:11-37 then the fluxes (SteadySim.cuh:73-78, RK.cuh:109-116):
// the SST kernel (mu_t, lambda_t, k / omega sources) initialises dq, the
// chemical source adds to it, the sweeps accumulate.
int rhs_block(ostc_solver *s, Block &b, bool *dt_done = nullptr) {
  bool init = true;
  if (s->n_turb) {
    ostc_solver::Timer t(s, "sst");
    HIPCHK(launch_sst(b.z, s->P, s->stream));
    init = false;
  }
  if (s->P.mix && s->P.reaction == 1 && !s->chem_split) {
    ostc_solver::Timer t(s, "source");
    HIPCHK(launch_source(b.z, s->P, s->d_mech, !init, s->stream));
    init = false;
  }
  static const char *names[3] = {"sweep_i", "sweep_j", "sweep_k"};
  for (int d = 0; d < s->P.dim; ++d) {
    ostc_solver::Timer t(s, names[d]);
    HIPCHK(launch_flux_sweep(b.z, s->P, s->d_mech, d, init && d == 0, s->fp32, s->stream,
                             dt_done && d == 2 ? &s->w : nullptr, dt_done));
  }
  return 0;
}

// k_planes_copied: in 2-D, k_eliminate_k ran right before (the k-ghost
// planes hold copies of k = 0)
int properties_block(ostc_solver *s, Block &b, bool k_planes_copied) {
  ostc_solver::Timer t(s, "properties");
  HIPCHK(launch_properties(b.z, s->P, s->d_mech, k_planes_copied, s->stream));
  return 0;
}

// CFL dt of the stage into w.dt: the minimum over every block
// (TimeAdvanceFunc.cu:155-201), on the device
int global_dt_device(ostc_solver *s) {
  ostc_solver::Timer t(s, "local_dt");
  const int nb = (int)s->blocks.size();
  for (int bi = 0; bi < nb; ++bi)
    HIPCHK(launch_local_dt(s->blocks[bi].z, s->P, s->w, nb == 1 ? s->w.dt : s->w.blk_dt + bi, s->stream));
  if (nb > 1) HIPCHK(launch_min_of(s->w.blk_dt, nb, s->w.dt, s->stream));
  return 0;
}

// residual over every block (Residual.cuh:57-139) into res_hist[slot] or out
int residual_all(ostc_solver *s, int slot, int first, double *out) {
  std::vector<const ZoneView *> zs;
  for (auto &b : s->blocks) zs.push_back(&b.z);
  HIPCHK(launch_residual_blocks(zs, s->w, slot, first, s->stream, out));
  return 0;
}

// Device dt / residual histories for a run of nsteps steps: kept in the
// solver's scratch and only reallocated when a longer run asks for more.
int ensure_hist(ostc_solver *s, int nsteps) {
  if (nsteps < 0) return fail("negative step count");
  if (nsteps <= s->w.hist_cap) return 0;
  HIPCHK(hipStreamSynchronize(s->stream));
  (void)hipFree(s->w.dt_hist);
  (void)hipFree(s->w.res_hist);
  s->w.dt_hist = s->w.res_hist = nullptr;
  s->w.hist_cap = 0;
  HIPCHK(hipMalloc(&s->w.dt_hist, sizeof(double) * 3 * nsteps));
  HIPCHK(hipMalloc(&s->w.res_hist, sizeof(double) * 5 * nsteps));
  s->w.hist_cap = nsteps;
  return 0;
}

}  // namespace

extern "C" {

int ostc_abi_version(void) { return 1; }

const char *ostc_last_error(void) { return g_err.c_str(); }

int ostc_init(const char *case_dir, int device, ostc_solver **out) {
  if (!case_dir || !out) return fail("null argument");
  auto s = std::make_unique<ostc_solver>();
  s->case_dir = case_dir;
  s->device = device;
  read_setup(s->case_dir, s->prm);
  const Params &p = s->prm;
  const int species = p.get_int("species", 0), reaction = p.get_int("reaction", 0);
  if (species != 0 && species != 1) return fail("species = " + std::to_string(species) + " is not supported");
  if (reaction == 2) return fail("flamelet model (reaction = 2) is not supported");
  int n_turb = 0;
  if (p.get_bool("turbulence", false)) {
    // main.cu:17-90: RANS (turbulence_method 1) with k-omega SST (RANS_model 2)
    if (p.get_int("turbulence_method", 1) != 1 || p.get_int("RANS_model", 2) != 2)
      return fail("only RANS with the k-omega SST model (turbulence_method 1, RANS_model 2) is supported");
    n_turb = 2;
  }
  if (species == 1) {
    const std::string in = s->case_dir + "/input_files";
    auto e = read_mechanism(in, p.get_string("mechanism_file", "chemistry/combustion_mechanism.inp"),
                            p.get_string("therm_file", "chemistry/therm.dat"),
                            p.get_string("transport_file", "chemistry/tran.dat"), reaction == 1, s->mech);
    if (!e.empty()) return fail(e);
    s->n_spec = s->mech.n_spec;
    // the point-implicit kernels park the rates of progress of at most 24
    // reactions per cell in LDS
    if (s->mech.n_reac > 24) return fail("mechanisms with more than 24 reactions are not supported by this build");
  }
  // Field.cu:4-60: variables per cell (species, then k and omega)
  s->n_turb = n_turb;
  s->n_scalar = s->n_spec + n_turb;
  s->n_var = 5 + s->n_scalar;
  s->ngg = p.get_int("ngg", 2);
  s->prm.i["n_spec"] = s->n_spec;
  s->prm.i["n_reac"] = s->mech.n_reac;
  s->prm.i["n_var"] = s->n_var;
  s->prm.i["n_scalar"] = s->n_scalar;
  s->prm.i["n_turb"] = n_turb;
  s->prm.i["i_turb_cv"] = 5 + s->n_spec;
  DParams &P = s->P;
  P.g = s->ngg;
  P.nv = s->n_var;
  P.ns = s->n_spec;
  P.nsc = s->n_scalar;
  P.nt = n_turb;
  P.turb_implicit = p.get_int("turb_implicit", 1);
  P.compress = p.get_int("compressibility_correction", 0);
  P.mix = species ? 1 : 0;
  P.reaction = reaction;
  P.chemSrcMethod = p.get_int("chemSrcMethod", 0);
  P.nr = s->mech.n_reac;
  P.inviscid_scheme = p.get_int("inviscid_scheme", 3);
  P.reconstruction = p.get_int("reconstruction", 2);
  P.limiter = p.get_int("limiter", 0);
  P.viscous = p.get_int("viscous_order", 2) != 0;
  P.inviscid_type = p.get_int("inviscid_type", 0);
  P.gradP = p.get_bool("gradPInDiffusionFlux", false) ? 1 : 0;
  P.cfl = p.get_real("cfl", 1.0);
  P.efix = p.get_real("entropy_fix_factor", 0.125);
  P.Pr = p.get_real("prandtl_number", 0.72);
  P.Sc = p.get_real("schmidt_number", 0.9);
  P.Prt = p.get_real("turbulent_prandtl_number", 0.9);
  P.Sct = p.get_real("turbulent_schmidt_number", 0.9);
  // not a reference key: metrics in the conservative symmetric form, which
  // keep a uniform flow exact on curved grids (the reference's are not)
  s->conservative_metric = p.get_bool("conservative_metric", false);
  P.jface = s->conservative_metric ? 1 : 0;
  // AWENO.cu:86: the characteristic interpolation reconstructs only for
  // reconstruction 4 (AWENO-5); reconstruction 5 (WENO7) selects the same
  // kernel in the reference but is not that scheme
  if (P.inviscid_type == 1 && P.reconstruction != 4)
    return fail("reconstruction " + std::to_string(P.reconstruction) +
                " with the characteristic (AWENO) path is not supported: only reconstruction 4 (AWENO-5)");
  // mixtures and k / omega take the multi-species characteristic
  // interpolation (the reference's mixture AWENO is an empty stub,
  // AWENO.cu:10-15; this build defines it in flux_device.h)
  if (n_turb && !P.viscous) return fail("SST needs the viscous flux (viscous_order 2)");
  if (P.mix && !species_supported(P.ns))
    return fail("a " + std::to_string(P.ns) + "-species mechanism: the kernels are compiled for 4 and 9 species");
  if (device < 0) {  // host-only: setup files, tables and grid metrics, no device work
    *out = s.release();
    return 0;
  }
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail("HIP device " + std::to_string(device) + " not available");
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  {
    // device address of the fixed scratch of ostc_chem_batch_noalloc, once per device
    std::string err;
    if (noalloc_scratch_init(device, err)) return fail(err);
  }
  if (upload_mech(s.get())) return 1;
  *out = s.release();
  return 0;
}

int ostc_destroy(ostc_solver *s) {
  if (s && s->device < 0) {
    delete s;
  } else if (s) {
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    delete s;
  }
  return 0;
}

int ostc_get_int(ostc_solver *s, const char *key, int *value) {
  if (!s || !key || !value) return fail("null argument");
  auto it = s->prm.i.find(key);
  if (it != s->prm.i.end()) {
    *value = it->second;
    return 0;
  }
  auto ib = s->prm.b.find(key);
  if (ib != s->prm.b.end()) {
    *value = ib->second ? 1 : 0;
    return 0;
  }
  return fail(std::string("integer key '") + key + "' not set");
}

int ostc_get_real(ostc_solver *s, const char *key, double *value) {
  if (!s || !key || !value) return fail("null argument");
  auto it = s->prm.r.find(key);
  if (it == s->prm.r.end()) return fail(std::string("real key '") + key + "' not set");
  *value = it->second;
  return 0;
}

int ostc_get_table(ostc_solver *s, const char *name, double *out, int64_t count) {
  if (!s || !name || !out) return fail("null argument");
  const Mechanism &m = s->mech;
  const std::map<std::string, const std::vector<double> *> dt = {
      {"mw", &m.mw}, {"t_low", &m.t_low}, {"t_mid", &m.t_mid}, {"t_high", &m.t_high}, {"high_coeff", &m.hi},
      {"low_coeff", &m.lo}, {"LJ_potent_inv", &m.LJ_potent_inv}, {"vis_coeff", &m.vis_coeff},
      {"WjDivWi_to_One4th", &m.WjDivWi_to_One4th}, {"sqrt_WiDivWjPl1Mul8", &m.sqrt_WiDivWjPl1Mul8},
      {"binary_diffusivity_coeff", &m.binary_diffusivity_coeff}, {"kb_over_eps_jk", &m.kb_over_eps_jk},
      {"A", &m.A}, {"b", &m.b}, {"Ea", &m.Ea}, {"A2", &m.A2}, {"b2", &m.b2}, {"Ea2", &m.Ea2},
      {"third_body_coeff", &m.third_body_coeff}, {"troe_alpha", &m.troe_alpha}, {"troe_t3", &m.troe_t3},
      {"troe_t1", &m.troe_t1}, {"troe_t2", &m.troe_t2}};
  const std::map<std::string, const std::vector<int> *> it = {
      {"label", &m.label}, {"stoi_f", &m.stoi_f}, {"stoi_b", &m.stoi_b}, {"order", &m.order},
      {"elem_comp", &m.elem_comp}};
  if (auto f = dt.find(name); f != dt.end()) {
    if ((int64_t)f->second->size() != count) return fail("table size mismatch for " + std::string(name));
    std::copy(f->second->begin(), f->second->end(), out);
    return 0;
  }
  if (auto f = it.find(name); f != it.end()) {
    if ((int64_t)f->second->size() != count) return fail("table size mismatch for " + std::string(name));
    for (int64_t i = 0; i < count; ++i) out[i] = (*f->second)[i];
    return 0;
  }
  return fail(std::string("unknown table ") + name);
}

int ostc_add_block(ostc_solver *s, int mx, int my, int mz, const double *x, const double *y, const double *z,
                   int *blk) {
  if (!s || !x || !y || !z) return fail("null argument");
  if (s->finalized) return fail("blocks must be added before ostc_finalize");
  if (mx < 2 * s->ngg + 2 || my < 2 * s->ngg + 2 || mz < 1) return fail("block too small for the ghost depth");
  if (!s->blocks.empty() && ((mz == 1) != (s->blocks[0].mz == 1))) return fail("mixed 2-D / 3-D blocks");
  Block b;
  b.mx = mx;
  b.my = my;
  b.mz = mz;
  std::string err;
  if (!compute_metrics(mx, my, mz, s->ngg, x, y, z, b.jac, b.metric, err, s->n_turb ? &b.xyz : nullptr,
                       s->conservative_metric, s->P.inviscid_type == 1))
    return fail(err);
  const size_t n = (size_t)mx * my * mz;
  b.x0.assign(x, x + n);
  b.y0.assign(y, y + n);
  b.z0.assign(z, z + n);
  s->blocks.push_back(std::move(b));
  if (blk) *blk = (int)s->blocks.size() - 1;
  return 0;
}

int ostc_read_grid(ostc_solver *s) {
  if (!s) return fail("null solver");
  if (s->finalized) return fail("the grid must be read before ostc_finalize");
  if (!s->blocks.empty()) return fail("blocks were already added");
  const std::string path = s->case_dir + "/input_files/grid/grid   0.grd";
  std::ifstream f(path);
  if (!f) return fail("grid file not found: " + path);
  // Mesh.cpp:599-686: block count, the three extents of every block, then
  // per block all x (i fastest, then j, then k), all y, all z
  std::string line;
  if (!std::getline(f, line)) return fail("grid file is empty");
  int nb = 0;
  try {
    nb = std::stoi(line);
  } catch (...) {
    return fail("grid file: bad block count");
  }
  if (nb <= 0) return fail("grid file: no blocks");
  std::vector<int> dims(3 * (size_t)nb);
  for (int b = 0; b < nb; ++b) {
    if (!std::getline(f, line)) return fail("grid file ends in the block extents");
    std::istringstream ls(line);
    if (!(ls >> dims[3 * b] >> dims[3 * b + 1] >> dims[3 * b + 2])) return fail("grid file: bad block extents");
  }
  // Mesh.cpp:571-578: coordinates scaled to metres unless gridScale is 1
  const double scale = s->prm.get_real("gridScale", 1.0);
  const bool do_scale = std::abs(scale - 1.0) > 1e-15;
  for (int b = 0; b < nb; ++b) {
    const size_t n = (size_t)dims[3 * b] * dims[3 * b + 1] * dims[3 * b + 2];
    std::vector<double> c[3];
    for (int a = 0; a < 3; ++a) {
      c[a].resize(n);
      for (size_t q = 0; q < n; ++q)
        if (!(f >> c[a][q])) return fail("grid file ends in the coordinates of block " + std::to_string(b));
      if (do_scale)
        for (double &v : c[a]) v *= scale;
    }
    if (ostc_add_block(s, dims[3 * b], dims[3 * b + 1], dims[3 * b + 2], c[0].data(), c[1].data(), c[2].data(),
                       nullptr))
      return 1;
  }
  return 0;
}

int ostc_add_inner_face(ostc_solver *s, int blk, const int *src7, const int *tar7) {
  if (!s || !src7 || !tar7) return fail("null argument");
  if (s->finalized) return fail("inner faces must be added before ostc_finalize");
  if (blk < 0 || blk >= (int)s->blocks.size()) return fail("block index out of range");
  InnerFace f;
  std::string err;
  const int dim = s->blocks[0].mz == 1 ? 2 : 3;
  if (!register_inner_face(src7, tar7, s->ngg, dim, f, err)) return fail(err);
  if (f.tblk < 0 || f.tblk >= (int)s->blocks.size()) return fail("inner face targets a block that does not exist");
  s->blocks[blk].inner.push_back(f);
  return 0;
}

int ostc_read_inner_file(ostc_solver *s) {
  if (!s) return fail("null solver");
  const std::string path = s->case_dir + "/input_files/boundary_condition/inner   0.txt";
  std::ifstream f(path);
  if (!f) return fail("inner-face file not found: " + path);
  // Mesh.cpp:710-742: the face count of every block, a header line, then
  // per face a source line and a target line of seven integers
  std::string line;
  if (!std::getline(f, line)) return fail("inner-face file is empty");
  const int nb = (int)s->blocks.size();
  std::vector<int> nface(nb);
  for (int b = 0; b < nb; ++b)
    if (!(f >> nface[b])) return fail("inner-face file: missing face count of block " + std::to_string(b));
  std::getline(f, line);
  std::getline(f, line);
  for (int b = 0; b < nb; ++b)
    for (int q = 0; q < nface[b]; ++q) {
      int a[7] = {0, 0, 0, 0, 0, 0, 0}, t[7] = {0, 0, 0, 0, 0, 0, 0};
      if (!std::getline(f, line)) return fail("inner-face file ends early");
      std::istringstream l1(line);
      for (int &v : a) l1 >> v;
      if (!std::getline(f, line)) return fail("inner-face file ends early");
      std::istringstream l2(line);
      for (int &v : t) l2 >> v;
      if (ostc_add_inner_face(s, b, a, t)) return 1;
    }
  return 0;
}

int ostc_read_mesh(ostc_solver *s) {
  // Mesh.cpp:550-568: grid, boundary conditions, inner faces
  if (ostc_read_grid(s)) return 1;
  if (ostc_read_boundary_file(s)) return 1;
  std::ifstream probe(s->case_dir + "/input_files/boundary_condition/inner   0.txt");
  if (probe && ostc_read_inner_file(s)) return 1;
  return 0;
}

int ostc_block_dims(ostc_solver *s, int blk, int *dims) {
  if (!s || !dims) return fail("null argument");
  if (blk < 0 || blk >= (int)s->blocks.size()) return fail("block index out of range");
  dims[0] = s->blocks[blk].mx;
  dims[1] = s->blocks[blk].my;
  dims[2] = s->blocks[blk].mz;
  return 0;
}

int ostc_block_count(ostc_solver *s, int *n) {
  if (!s || !n) return fail("null argument");
  *n = (int)s->blocks.size();
  return 0;
}

int ostc_inner_faces(ostc_solver *s, int blk, int *count, int *out, int cap) {
  if (!s) return fail("null solver");
  if (blk < 0 || blk >= (int)s->blocks.size()) return fail("block index out of range");
  const auto &v = s->blocks[blk].inner;
  if (count) *count = (int)v.size();
  if (out)
    for (int q = 0; q < (int)v.size() && q < cap; ++q) {
      const InnerFace &f = v[q];
      int *o = out + 31 * q;
      const int vals[31] = {f.rs[0], f.rs[1], f.rs[2], f.re[0], f.re[1], f.re[2], f.face, f.dir,
                            f.ts[0], f.ts[1], f.ts[2], f.te[0], f.te[1], f.te[2], f.tblk, f.tface,
                            f.tdir, f.src_tar[0], f.src_tar[1], f.src_tar[2], f.loop[0], f.loop[1], f.loop[2],
                            f.tloop[0], f.tloop[1], f.tloop[2], f.n[0], f.n[1], f.n[2], 0, 0};
      std::copy(vals, vals + 31, o);
    }
  return 0;
}

int ostc_add_boundary(ostc_solver *s, int blk, int i1, int i2, int j1, int j2, int k1, int k2, int label) {
  if (!s) return fail("null solver");
  if (blk < 0 || blk >= (int)s->blocks.size()) return fail("block index out of range");
  BoundaryFace f;
  const int r[6] = {i1, i2, j1, j2, k1, k2};
  std::copy(r, r + 6, f.raw);
  f.label = label;
  s->blocks[blk].faces.push_back(f);
  return 0;
}

int ostc_read_boundary_file(ostc_solver *s) {
  if (!s) return fail("null solver");
  std::ifstream f(s->case_dir + "/input_files/boundary_condition/boundary   0.txt");
  if (!f) return fail("boundary file not found");
  std::string line;
  for (size_t b = 0; b < s->blocks.size(); ++b) {
    if (!std::getline(f, line)) return fail("boundary file ends early");
    const int n = std::stoi(line);
    for (int q = 0; q < n; ++q) {
      if (!std::getline(f, line)) return fail("boundary file ends early");
      std::istringstream ls(line);
      int v[7] = {0, 0, 0, 0, 0, 0, 0};
      for (int &x : v) ls >> x;
      if (ostc_add_boundary(s, (int)b, v[0], v[1], v[2], v[3], v[4], v[5], v[6])) return 1;
    }
  }
  return 0;
}

int ostc_finalize(ostc_solver *s) {
  if (!s) return fail("null solver");
  if (s->finalized) return fail("already finalized");
  if (s->device < 0) return fail("solver was opened host-only (device < 0)");
  if (s->blocks.empty()) return fail("no block added");
  HIPCHK(hipSetDevice(s->device));
  const int dim = s->blocks[0].mz == 1 ? 2 : 3;
  s->P.dim = dim;
  // BoundCond.cu:167-280: label -> type, named structs first, then defaults
  std::vector<int> labels;
  for (auto &b : s->blocks)
    for (auto &f : b.faces)
      if (std::find(labels.begin(), labels.end(), f.label) == labels.end()) labels.push_back(f.label);
  std::map<int, int> type_of;
  std::vector<std::pair<int, int>> reg;  // (type, label) in registration order
  auto bca = s->prm.sa.find("boundary_conditions");
  if (bca != s->prm.sa.end()) {
    for (auto &name : bca->second) {
      auto st = s->prm.st.find(name);
      if (st == s->prm.st.end()) return fail("boundary condition struct '" + name + "' missing");
      const Struct &info = st->second;
      auto lf = info.find("label");
      auto tf = info.find("type");
      if (lf == info.end() || tf == info.end()) return fail("struct '" + name + "' needs label and type");
      const int label = std::get<int>(lf->second);
      auto li = std::find(labels.begin(), labels.end(), label);
      if (li == labels.end()) continue;
      labels.erase(li);
      const std::string type = std::get<std::string>(tf->second);
      int t = 0;
      if (type == "wall") t = BC_WALL;
      else if (type == "inflow") t = BC_INFLOW;
      else if (type == "outflow") t = BC_OUTFLOW;
      else if (type == "symmetry") t = BC_SYMMETRY;
      else if (type == "farfield") t = BC_FARFIELD;
      else if (type == "subsonic_inflow") t = BC_SUBSONIC;
      else if (type == "back_pressure") t = BC_BACKPRESSURE;
      else return fail("unknown boundary type '" + type + "'");
      type_of[label] = t;
      reg.push_back({t, label});
      if (t == BC_INFLOW) {
        Inflow in;
        if (make_inflow(s, name, in)) return 1;
        in.label = label;
        s->inflows.push_back(in);
      } else if (t == BC_WALL) {
        WallBC wb;
        wb.label = label;
        auto th = info.find("thermal_type");
        if (th != info.end()) wb.isothermal = std::get<std::string>(th->second) == "isothermal";
        auto tw = info.find("temperature");
        if (wb.isothermal && tw != info.end()) wb.t = std::get<double>(tw->second);
        s->walls.push_back(wb);
      }
    }
  }
  for (int lab : labels) {
    if (lab == 2 || lab == 3 || lab == 4 || lab == 5 || lab == 6 || lab == 7 || lab == 9) {
      type_of[lab] = lab;
      reg.push_back({lab, lab});
      if (lab == 2) {
        WallBC wb;
        wb.label = 2;
        s->walls.push_back(wb);
      }
      if (lab == 5) return fail("inflow label 5 has no setup struct");
    } else {
      return fail("boundary label " + std::to_string(lab) + " has no type");
    }
  }
  for (auto &[t, l] : reg)
    if ((t == BC_SUBSONIC || t == BC_BACKPRESSURE) && s->n_spec > 0)
      return fail("subsonic inflow and back pressure boundaries are defined for air only (BoundCond.cpp:300-364)");
  // allocate zones
  const int g = s->ngg;
  int64_t max_cells = 0, red_blocks = 0;
  std::vector<std::array<std::vector<double>, 3>> nodes;
  if (build_nodes(s, nodes)) return 1;
  int *d_bad_interior = nullptr;
  unsigned char *d_bad = nullptr;
  double *d_nodes = nullptr, *d_fsp = nullptr;
  auto free_geo = [&]() {
    (void)hipFree(d_bad_interior);
    (void)hipFree(d_bad);
    (void)hipFree(d_nodes);
    (void)hipFree(d_fsp);
  };
  for (auto &b : s->blocks) {
    ZoneView &z = b.z;
    z.mx = b.mx;
    z.my = b.my;
    z.mz = b.mz;
    z.g = g;
    const Layout L(b.mx, b.my, b.mz, g), L0(b.mx, b.my, b.mz, 0);
    // kernels decompose point indices in 32-bit arithmetic
    if (L.sz >= ((int64_t)1 << 32))
      return fail("block " + std::to_string(&b - &s->blocks[0]) +
                  " has 2^32 or more ghosted points; split it into smaller blocks");
    z.d2 = L.d2;
    z.d1 = L.d1;
    z.sz = L.sz;
    z.e2 = L0.d2;
    z.e1 = L0.d1;
    z.sz0 = L0.sz;
    max_cells = std::max<int64_t>(max_cells, L0.sz);
    red_blocks += (L0.sz + 255) / 256 + 1;
    auto alloc = [&](double **p, int64_t n) -> int {
      n = std::max<int64_t>(n, 1);
      HIPCHK(hipMalloc((void **)p, n * sizeof(double)));
      HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(double), s->stream));
      b.mem.push_back(*p);
      return 0;
    };
    const int nv = s->n_var, ns = s->n_spec;
    double *jac = nullptr, *met = nullptr;
    if (alloc(&z.bv, L.sz * 6) || alloc(&z.sv, L.sz * std::max(s->n_scalar, 1)) || alloc(&z.cv, L.sz * nv) ||
        alloc(&z.qn, L.sz * nv) || alloc(&z.vel, L.sz) || alloc(&z.acoustic, L.sz) || alloc(&z.mach, L.sz) ||
        alloc(&z.mul, L.sz) || alloc(&z.cond, L.sz) || alloc(&z.rhoD, ns ? L.sz * ns : 1) ||
        alloc(&z.gam, ns ? L.sz : 1) || alloc(&z.cp, ns ? L.sz : 1) || alloc(&jac, L.sz) ||
        alloc(&met, L.sz * 9) || alloc(&z.dq, L0.sz * nv) || alloc(&z.dt_local, L0.sz) ||
        alloc(&z.bv_last, L0.sz * 4) ||
        alloc(&z.chem_jac, s->P.chemSrcMethod == 1 ? L0.sz * ns * ns : (s->P.chemSrcMethod == 2 ? L0.sz * ns : 1)) ||
        (s->prm.get_int("implicit_method", 0) == 1 &&
         (alloc(&z.inv_sr, L0.sz * 3) || alloc(&z.dq0, L0.sz * nv) || alloc(&z.dqk, L0.sz * nv) || (ns && alloc(&z.dplur_h, L0.sz * (ns + 1))))))
      return 1;
    // SST: mu_t and lambda_t (ghost lambda_t stays zero: only the source
    // kernel writes it, on the interior), wall distance, source diagonal
    if (s->n_turb && (alloc(&z.mut, L.sz) || alloc(&z.tcond, L.sz) || alloc(&z.wdist, L.sz) ||
                      (s->P.turb_implicit == 1 && alloc(&z.turb_jac, L0.sz * 2))))
      return 1;
    z.jac = jac;
    z.metric = met;
    {
      // Jacobian and metrics on the device from the ghosted nodes
      // (Mesh.cpp:259-521), bit-identical to the host computation
      const size_t bi = &b - &s->blocks[0];
      const Layout Lc(b.mx, b.my, b.mz, g + 1);
      hipError_t e = hipMalloc(&d_nodes, sizeof(double) * 3 * Lc.sz);
      if (e == hipSuccess) e = hipMalloc(&d_bad, L.sz);
      if (e == hipSuccess) e = hipMalloc(&d_bad_interior, sizeof(int));
      if (e == hipSuccess && s->conservative_metric) e = hipMalloc(&d_fsp, sizeof(double) * 9 * Lc.sz);
      if (e == hipSuccess) e = hipMemsetAsync(d_bad_interior, 0, sizeof(int), s->stream);
      for (int c = 0; c < 3 && e == hipSuccess; ++c)
        e = hipMemcpyAsync(d_nodes + c * Lc.sz, nodes[bi][c].data(), sizeof(double) * Lc.sz, hipMemcpyHostToDevice,
                           s->stream);
      if (e == hipSuccess)
        e = launch_metrics(b.mx, b.my, b.mz, g, d_nodes, d_nodes + Lc.sz, d_nodes + 2 * Lc.sz, jac, met, d_bad,
                           d_bad_interior, s->stream, s->conservative_metric ? 1 : 0,
                           s->P.inviscid_type == 1 ? 1 : 0, d_fsp);
      int bad_interior = 0;
      if (e == hipSuccess)
        e = hipMemcpyAsync(&bad_interior, d_bad_interior, sizeof(int), hipMemcpyDeviceToHost, s->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
      free_geo();
      d_nodes = nullptr;
      d_fsp = nullptr;
      d_bad = nullptr;
      d_bad_interior = nullptr;
      if (e != hipSuccess) return fail(std::string("device metrics: ") + hipGetErrorString(e));
      if (bad_interior) return fail("negative Jacobian in the interior of block " + std::to_string(bi));
      if (s->n_turb) {
        // ghosted node coordinates for the wall distance (ngg layers)
        b.xyz.assign((size_t)L.sz * 3, 0);
        for (int k = -g; k < b.mz + g; ++k)
          for (int j = -g; j < b.my + g; ++j)
            for (int i = -g; i < b.mx + g; ++i)
              for (int c = 0; c < 3; ++c) b.xyz[(size_t)c * L.sz + L.at(i, j, k)] = nodes[bi][c][Lc.at(i, j, k)];
      }
    }
    // boundaries: register, trim walls (BoundCond.cu:369-384), order by priority
    for (auto &f : b.faces) {
      register_face(f, g, dim);
      f.type = type_of.at(f.label);
      if (f.type == BC_WALL)
        for (int q = 0; q < 3; ++q) {
          if (q == f.face) continue;
          f.rs[q] += g;
          f.re[q] -= g;
        }
    }
    static const int prio[] = {BC_FARFIELD, BC_OUTFLOW, BC_BACKPRESSURE, BC_INFLOW, BC_SUBSONIC, BC_SYMMETRY, BC_WALL};
    for (int t : prio)
      for (auto &[rt, rl] : reg) {
        if (rt != t) continue;
        for (auto &f : b.faces) {
          if (f.label != rl) continue;
          BCDesc d{};
          d.type = t;
          d.face = f.face;
          d.dir = f.direction;
          for (int q = 0; q < 3; ++q) {
            d.rs[q] = f.rs[q];
            d.re[q] = f.re[q];
          }
          if (t == BC_INFLOW) {
            const Inflow *in = nullptr;
            for (auto &x : s->inflows)
              if (x.label == rl) in = &x;
            if (!in) return fail("no inflow state for label " + std::to_string(rl));
            d.rho = in->rho;
            d.u = in->u;
            d.v = in->v;
            d.w = in->w;
            d.p = in->p;
            d.t = in->t;
            d.vel = in->vel;
            d.mut = in->mut;
            for (size_t l = 0; l < in->sv.size() && l < (size_t)kMS + 4; ++l) d.sv[l] = in->sv[l];
          } else if (t == BC_FARFIELD || t == BC_SUBSONIC || t == BC_BACKPRESSURE) {
            std::string nm;
            for (auto &[n2, v2] : s->prm.st) {
              auto lf = v2.find("label");
              if (lf != v2.end() && std::holds_alternative<int>(lf->second) && std::get<int>(lf->second) == rl) nm = n2;
            }
            if (nm.empty()) return fail("no setup struct for boundary label " + std::to_string(rl));
            const Struct &info = s->prm.st.at(nm);
            auto get_r = [&](const char *k, double &v) {
              auto f = info.find(k);
              if (f == info.end()) return false;
              if (std::holds_alternative<double>(f->second))
                v = std::get<double>(f->second);
              else if (std::holds_alternative<int>(f->second))
                v = std::get<int>(f->second);
              return true;
            };
            if (t == BC_FARFIELD) {
              if (make_farfield(s, nm, d)) return 1;
            } else {
              double p_inf = 0, t_inf = 0, rho_inf = 0, v_inf = 0;
              if (reference_pt(s, p_inf, t_inf, &rho_inf, &v_inf)) return 1;
              if (t == BC_SUBSONIC) {
                // BoundCond.cpp:300-343
                double ptr = 1, ttr = 1;
                if (!get_r("pt_pRef_ratio", ptr) || !get_r("Tt_TRef_ratio", ttr))
                  return fail("subsonic inflow '" + nm + "' needs pt_pRef_ratio and Tt_TRef_ratio");
                d.u = 1;
                d.v = 0;
                d.w = 0;
                get_r("u", d.u);
                get_r("v", d.v);
                get_r("w", d.w);
                d.p_total = ptr * p_inf;
                d.t_total = ttr * t_inf;
                // BoundCond.cpp:327-341: SST state from Sutherland(T_ref), v_inf, rho_inf
                if (s->n_turb) {
                  const double mu = 1.716e-5 * std::pow(t_inf / 273, 1.5) * (273 + 111) / (t_inf + 111);
                  if (turbulence_state(s, info, nm, mu, v_inf, rho_inf, d.mut, d.sv)) return 1;
                }
              } else {
                // BoundCond.cpp:345-364
                double pb = -1, ratio = 1;
                get_r("pressure", pb);
                if (pb < 0) {
                  get_r("p_pRef_ratio", ratio);
                  pb = ratio * p_inf;
                }
                d.p_back = pb;
              }
            }
          } else if (t == BC_WALL) {
            d.t_wall = 300;
            d.isothermal = 1;
            for (auto &x : s->walls)
              if (x.label == rl) {
                d.t_wall = x.t;
                d.isothermal = x.isothermal ? 1 : 0;
              }
          }
          b.bcs.push_back(d);
        }
      }
  }
  if (s->n_turb && compute_wall_distance(s)) return 1;
  // per-run scratch: partials, ticket counter, dt slot, residual scale
  s->w.max_blocks = (int)std::max<int64_t>((max_cells + 255) / 256 + 1, red_blocks);
  HIPCHK(hipMalloc(&s->w.partial, sizeof(double) * 4 * s->w.max_blocks));
  HIPCHK(hipMalloc(&s->w.blk_dt, sizeof(double) * s->blocks.size()));
  HIPCHK(hipMalloc(&s->w.counter, sizeof(unsigned)));
  HIPCHK(hipMemsetAsync(s->w.counter, 0, sizeof(unsigned), s->stream));
  HIPCHK(hipMalloc(&s->w.dt, sizeof(double)));
  HIPCHK(hipMalloc(&s->w.res_scale, sizeof(double) * 4));
  HIPCHK(hipMalloc(&s->w.res_one, sizeof(double) * 5));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->finalized = true;
  return 0;
}

int ostc_block_geometry(ostc_solver *s, int blk, double *jac, double *metric) {
  if (!s) return fail("null solver");
  if (blk < 0 || blk >= (int)s->blocks.size()) return fail("block index out of range");
  const Block &b = s->blocks[blk];
  if (s->finalized) {
    // the metrics in use: computed on the device with the inner faces' ghosts
    HIPCHK(hipSetDevice(s->device));
    if (jac) HIPCHK(hipMemcpyAsync(jac, b.z.jac, sizeof(double) * b.z.sz, hipMemcpyDeviceToHost, s->stream));
    if (metric)
      HIPCHK(hipMemcpyAsync(metric, b.z.metric, sizeof(double) * 9 * b.z.sz, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
  }
  if (jac) std::copy(b.jac.begin(), b.jac.end(), jac);
  if (metric) std::copy(b.metric.begin(), b.metric.end(), metric);
  return 0;
}

int ostc_field_info(ostc_solver *s, int blk, int field, int64_t *count, int *ghosts, int *ncomp) {
  if (check_blk(s, blk)) return 1;
  FieldDesc d = field_desc(s, s->blocks[blk], field);
  if (d.count < 0) return fail("unknown field id");
  if (count) *count = d.count;
  if (ghosts) *ghosts = d.ghosts;
  if (ncomp) *ncomp = d.ncomp;
  return 0;
}

int ostc_upload(ostc_solver *s, int blk, int field, const double *host) {
  if (check_blk(s, blk)) return 1;
  FieldDesc d = field_desc(s, s->blocks[blk], field);
  if (d.count < 0) return fail("unknown field id");
  if (field == OSTC_JAC || field == OSTC_METRIC) return fail("geometry is derived from the grid");
  if (d.count == 0) return 0;
  HIPCHK(hipMemcpyAsync(d.ptr, host, d.count * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_download(ostc_solver *s, int blk, int field, double *host) {
  if (check_blk(s, blk)) return 1;
  FieldDesc d = field_desc(s, s->blocks[blk], field);
  if (d.count < 0) return fail("unknown field id");
  if (d.count == 0) return 0;
  HIPCHK(hipMemcpyAsync(host, d.ptr, d.count * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_bind_zone(ostc_solver *s, int blk, const ostc_zone_view *v) {
  if (check_blk(s, blk)) return 1;
  if (!v) return fail("bind_zone: null view");
  Block &b = s->blocks[blk];
  if (v->mx != b.mx || v->my != b.my || v->mz != b.mz)
    return fail("bind_zone: view extent " + std::to_string(v->mx) + "x" + std::to_string(v->my) + "x" +
                std::to_string(v->mz) + " does not match block " + std::to_string(blk));
  if (v->ngg != s->ngg || v->n_var != s->n_var || v->n_scalar != s->n_scalar)
    return fail("bind_zone: ngg / n_var / n_scalar differ from the setup (" + std::to_string(s->ngg) + ", " +
                std::to_string(s->n_var) + ", " + std::to_string(s->n_scalar) + ")");
  if (!v->bv || !v->cv || (s->n_scalar > 0 && !v->sv)) return fail("bind_zone: bv, cv (and sv) are required");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  // the library's own array of a bound field is freed (never a caller's)
  auto rebind = [&](double *&slot, double *caller) -> int {
    if (!caller || caller == slot) return 0;
    auto it = std::find(b.mem.begin(), b.mem.end(), (void *)slot);
    if (it != b.mem.end()) {
      HIPCHK(hipFree(*it));
      b.mem.erase(it);
    }
    slot = caller;
    return 0;
  };
  if (rebind(b.z.bv, v->bv) || rebind(b.z.cv, v->cv) || (s->n_scalar > 0 && rebind(b.z.sv, v->sv)) ||
      rebind(b.z.qn, v->qn) || rebind(b.z.dq, v->dq))
    return 1;
  return 0;
}

int ostc_local_dt(ostc_solver *s, int blk) {
  if (check_blk(s, blk)) return 1;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(launch_local_dt(s->blocks[blk].z, s->P, s->w, s->w.dt, s->stream));
  return 0;
}

int ostc_store_last(ostc_solver *s, int blk) {
  if (check_blk(s, blk)) return 1;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(launch_store_last(s->blocks[blk].z, s->stream));
  return 0;
}

int ostc_residual(ostc_solver *s, int step, double res[4], double *err_max) {
  if (check_blk(s, 0)) return 1;
  HIPCHK(hipSetDevice(s->device));
  if (residual_all(s, 0, step == 1 ? 1 : 0, s->w.res_one)) return 1;
  if (step == 1) s->res_scale_set = true;
  double h[5];
  HIPCHK(hipMemcpyAsync(h, s->w.res_one, sizeof(h), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (std::isnan(h[4])) return fail("NaN residual at step " + std::to_string(step));
  if (res)
    for (int l = 0; l < 4; ++l) res[l] = h[l];
  if (err_max) *err_max = h[4];
  return 0;
}

int ostc_field_device_ptr(ostc_solver *s, int blk, int field, double **dptr) {
  if (check_blk(s, blk)) return 1;
  FieldDesc d = field_desc(s, s->blocks[blk], field);
  if (d.count < 0) return fail("unknown field id");
  *dptr = d.ptr;
  return 0;
}

int ostc_inflow_state(ostc_solver *s, const char *name, double *out, int n) {
  if (!s || !name || !out) return fail("null argument");
  Inflow in;
  if (make_inflow(s, name, in)) return 1;
  const double v[7] = {in.rho, in.u, in.v, in.w, in.p, in.t, in.vel};
  for (int i = 0; i < n && i < 7; ++i) out[i] = v[i];
  for (int i = 7; i < n && i - 7 < (int)in.sv.size(); ++i) out[i] = in.sv[i - 7];
  if (n > 7 + (int)in.sv.size()) out[7 + in.sv.size()] = in.mut;
  return 0;
}

int ostc_initialize_uniform(ostc_solver *s, int blk) {
  if (check_blk(s, blk)) return 1;
  Inflow in;
  if (make_inflow(s, s->prm.get_string("default_init", s->prm.get_string("reference_state", "")), in)) return 1;
  Block &b = s->blocks[blk];
  const int64_t sz = b.z.sz;
  std::vector<double> bv(sz * 6), sv(sz * std::max(s->n_scalar, 0));
  const double v[6] = {in.rho, in.u, in.v, in.w, in.p, in.t};
  for (int l = 0; l < 6; ++l) std::fill(bv.begin() + l * sz, bv.begin() + (l + 1) * sz, v[l]);
  for (int l = 0; l < s->n_scalar; ++l) std::fill(sv.begin() + l * sz, sv.begin() + (l + 1) * sz, in.sv[l]);
  if (ostc_upload(s, blk, OSTC_BV, bv.data())) return 1;
  if (s->n_scalar > 0 && ostc_upload(s, blk, OSTC_SV, sv.data())) return 1;
  return 0;
}

int ostc_initialize(ostc_solver *s) {
  if (check_blk(s, 0)) return 1;
  HIPCHK(hipSetDevice(s->device));
  // Driver.cu:40-98 phase by phase over all blocks, then RK.cuh:38-48
  if (s->P.dim == 2)
    for (auto &b : s->blocks) HIPCHK(launch_eliminate_k(b.z, s->P, false, s->stream));
  for (auto &b : s->blocks) {
    // the mixture farfield reads gamma of its boundary point; the reference
    // applies the first BCs before any properties exist (Driver.cu:53-92,
    // reading unset memory), here they are computed from the initial field
    if (s->P.mix && properties_block(s, b, true)) return 1;
    if (run_bcs(s, b, false)) return 1;
  }
  for (auto &b : s->blocks) {
    HIPCHK(launch_velocity(b.z, s->stream));
    if (s->n_turb) HIPCHK(launch_init_mut(b.z, s->P, s->d_mech, s->stream));
  }
  if (exchange_all(s, false)) return 1;
  for (auto &b : s->blocks) {
    if (properties_block(s, b, false)) return 1;
    HIPCHK(launch_store_last(b.z, s->stream));
    HIPCHK(launch_cv_from_bv(b.z, s->P, s->d_mech, s->stream));
  }
  // a restart continues the step count and time of the file it was read from
  s->t_phys = s->restart ? s->restart_time : 0;
  s->step = s->restart ? s->restart_step : 0;
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_rhs(ostc_solver *s, int blk) {
  if (check_blk(s, blk)) return 1;
  if (rhs_block(s, s->blocks[blk])) return 1;
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_global_dt(ostc_solver *s, double *dt) {
  if (check_blk(s, 0)) return 1;
  double best = 1e+6;
  for (auto &b : s->blocks) {
    HIPCHK(launch_local_dt(b.z, s->P, s->w, s->w.dt, s->stream));
    double v = 0;
    HIPCHK(hipMemcpyAsync(&v, s->w.dt, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    best = std::min(best, v);
  }
  if (dt) *dt = best;
  return 0;
}

int ostc_rk3_stage(ostc_solver *s, int blk, double dt, int rk) {
  if (check_blk(s, blk)) return 1;
  if (rk < 0 || rk > 2) return fail("rk stage must be 0, 1 or 2");
  Block &b = s->blocks[blk];
  HIPCHK(hipMemcpyAsync(s->w.dt, &dt, sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIPCHK(launch_rk_update(b.z, s->P, s->d_mech, s->w.dt, rk, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_apply_bc(ostc_solver *s, int blk, int with_cv) {
  if (check_blk(s, blk)) return 1;
  if (run_bcs(s, s->blocks[blk], with_cv != 0)) return 1;
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_exchange(ostc_solver *s, int with_cv) {
  if (check_blk(s, 0)) return 1;
  HIPCHK(hipSetDevice(s->device));
  if (exchange_all(s, with_cv != 0)) return 1;
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_update_properties(ostc_solver *s, int blk) {
  if (check_blk(s, blk)) return 1;
  Block &b = s->blocks[blk];
  if (s->P.dim == 2) HIPCHK(launch_eliminate_k(b.z, s->P, true, s->stream));
  if (properties_block(s, b, true)) return 1;
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_step_rk3(ostc_solver *s, int nsteps, double *dt_hist, double *res_hist) {
  if (check_blk(s, 0)) return 1;
  if (nsteps < 0) return fail("negative step count");
  if (nsteps == 0) return 0;
  HIPCHK(hipSetDevice(s->device));
  if (ensure_hist(s, nsteps)) return 1;
  double *d_dt = s->w.dt_hist, *d_res = s->w.res_hist;
  const bool fixed = s->prm.get_bool("fixed_time_step", false);
  const double fixed_dt = s->prm.get_real("dt", 1e-5);
  const int os = std::max(1, s->prm.get_int("output_screen", 1));
  int rc = 0;
  auto body = [&]() -> int {
    if (fixed) HIPCHK(hipMemcpyAsync(s->w.dt, &fixed_dt, sizeof(double), hipMemcpyHostToDevice, s->stream));
    for (int it = 0; it < nsteps; ++it) {
      const int step = s->step + 1 + it;
      for (auto &b : s->blocks) {
        if (step % os == 0) HIPCHK(launch_store_last(b.z, s->stream));
        // qn: stage 0 of k_rk_update stores the step's initial state
      }
      for (int rk = 0; rk < 3; ++rk) {
        // one block: the k-sweep computes the CFL dt where it can
        bool dt_done = false;
        const bool fold_dt = !fixed && s->blocks.size() == 1;
        for (auto &b : s->blocks)
          if (rhs_block(s, b, fold_dt ? &dt_done : nullptr)) return 1;
        if (!fixed && !dt_done && global_dt_device(s)) return 1;
        HIPCHK(launch_copy_dt(s->w.dt, d_dt + it * 3 + rk, s->stream));
        for (auto &b : s->blocks) {
          {
            ostc_solver::Timer t(s, "rk_update");
            HIPCHK(launch_rk_update(b.z, s->P, s->d_mech, s->w.dt, rk, s->stream));
          }
          if (run_bcs(s, b, true)) return 1;
        }
        if (exchange_all(s, true)) return 1;
        if (s->P.dim == 2)
          for (auto &b : s->blocks) HIPCHK(launch_eliminate_k(b.z, s->P, true, s->stream));
        for (auto &b : s->blocks)
          if (properties_block(s, b, true)) return 1;
      }
      if (step % os == 0 || step == 1) {
        ostc_solver::Timer t(s, "residual");
        if (residual_all(s, it, step == 1 ? 1 : 0, nullptr)) return 1;
        if (step == 1) s->res_scale_set = true;
      }
    }
    return 0;
  };
  rc = body();
  std::vector<double> hdt(3 * nsteps), hres(5 * nsteps, NAN);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(hdt.data(), d_dt, sizeof(double) * 3 * nsteps, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hres.data(), d_res, sizeof(double) * 5 * nsteps, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(std::string("step_rk3: ") + hipGetErrorString(e));
  }
  if (rc) {
    // the block state is partly advanced; timing events of the launches
    // that were queued are released with the rest
    s->drain_profile();
    return rc;
  }
  s->drain_profile();
  for (int it = 0; it < nsteps; ++it) {
    const int step = s->step + 1 + it;
    s->t_phys += hdt[it * 3 + 2];
    if (!(step % os == 0 || step == 1))
      for (int l = 0; l < 5; ++l) hres[it * 5 + l] = NAN;
    else if (std::isnan(hres[it * 5 + 4]))
      return fail("NaN residual at step " + std::to_string(step));
  }
  s->step += nsteps;
  if (dt_hist) std::copy(hdt.begin(), hdt.end(), dt_hist);
  if (res_hist) std::copy(hres.begin(), hres.end(), res_hist);
  return 0;
}

int chem_interior(ostc_solver *s, Block &b, double dt, int mode, int64_t *nf_out);

static int implicit_block(ostc_solver *s, Block &b) {
  const int method = s->prm.get_int("implicit_method", 0);
  ostc_solver::Timer t(s, "implicit");
  // split chemistry: the right-hand side has no chemical source, so the
  // implicit treatment has no chemistry Jacobian either
  DParams P = s->P;
  if (s->chem_split) P.reaction = 0;
  if (method == 1) {
    std::vector<BCDesc> walls;
    for (auto &bc : b.bcs)
      if (bc.type == BC_WALL) walls.push_back(bc);
    HIPCHK(launch_dplur(b.z, P, s->d_mech, s->prm.get_int("DPLUR_inner_step", 3), walls, s->stream));
  } else {
    HIPCHK(launch_chem_implicit(b.z, P, s->stream));
    if (s->n_turb && s->P.turb_implicit == 1) HIPCHK(launch_turb_implicit(b.z, s->P, s->stream));
  }
  return 0;
}

int ostc_implicit(ostc_solver *s, int blk) {
  if (check_blk(s, blk)) return 1;
  HIPCHK(hipSetDevice(s->device));
  if (implicit_block(s, s->blocks[blk])) return 1;
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_step_steady(ostc_solver *s, int nsteps, double *res_hist) {
  if (check_blk(s, 0)) return 1;
  if (nsteps < 0) return fail("negative step count");
  if (nsteps == 0) return 0;
  HIPCHK(hipSetDevice(s->device));
  if (ensure_hist(s, nsteps)) return 1;
  double *d_res = s->w.res_hist;
  const int os = std::max(1, s->prm.get_int("output_screen", 1));
  auto body = [&]() -> int {
    // ghost cv follows for the AWENO path, which reads it
    const bool with_cv = s->P.inviscid_type == 1;
    for (int it = 0; it < nsteps; ++it) {
      const int step = s->step + 1 + it;
      for (auto &b : s->blocks) {
        if (step % os == 0) HIPCHK(launch_store_last(b.z, s->stream));
        if (rhs_block(s, b)) return 1;
        {
          ostc_solver::Timer t(s, "local_dt");
          HIPCHK(launch_local_dt(b.z, s->P, s->w, s->w.dt, s->stream));
        }
        if (implicit_block(s, b)) return 1;
        {
          ostc_solver::Timer t(s, "update_bv");
          HIPCHK(launch_update_bv(b.z, s->P, s->d_mech, s->stream));
        }
        if (s->steady_chem_dt > 0) {
          // inline operator-split chemistry after the flow update
          int64_t nf = 0;
          if (chem_interior(s, b, s->steady_chem_dt, s->steady_chem_mode, &nf)) return 1;
          s->steady_cells += (int64_t)b.mx * b.my * b.mz;
          s->steady_fb += nf;
        }
        if (run_bcs(s, b, with_cv)) return 1;
      }
      if (exchange_all(s, with_cv)) return 1;
      if (s->P.dim == 2)
        for (auto &b : s->blocks) HIPCHK(launch_eliminate_k(b.z, s->P, with_cv, s->stream));
      for (auto &b : s->blocks)
        if (properties_block(s, b, true)) return 1;
      if (step % os == 0 || step == 1) {
        ostc_solver::Timer t(s, "residual");
        if (residual_all(s, it, step == 1 ? 1 : 0, nullptr)) return 1;
        if (step == 1) s->res_scale_set = true;
      }
    }
    return 0;
  };
  int rc = body();
  std::vector<double> hres(5 * nsteps, NAN);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(hres.data(), d_res, sizeof(double) * 5 * nsteps, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(std::string("step_steady: ") + hipGetErrorString(e));
  }
  if (rc) {
    s->drain_profile();
    return rc;
  }
  s->drain_profile();
  for (int it = 0; it < nsteps; ++it) {
    const int step = s->step + 1 + it;
    if (!(step % os == 0 || step == 1))
      for (int l = 0; l < 5; ++l) hres[it * 5 + l] = NAN;
    else if (std::isnan(hres[it * 5 + 4]))
      return fail("NaN residual at step " + std::to_string(step));
  }
  s->step += nsteps;
  if (res_hist) std::copy(hres.begin(), hres.end(), res_hist);
  return 0;
}

int ostc_time(ostc_solver *s, double *t, int *step) {
  if (!s) return fail("null solver");
  if (t) *t = s->t_phys;
  if (step) *step = s->step;
  return 0;
}

int ostc_set_flux_precision(ostc_solver *s, int fp32) {
  if (!s) return fail("null solver");
  s->fp32 = fp32 != 0;
  return 0;
}

// the chemistry entry points need a reacting mechanism of a compiled species
// count; the surrogate (mode 1) is built for the 9-species mechanism
static int need_chem(const ostc_solver *s, int mode) {
  if (!(s->n_spec == 4 || s->n_spec == 9) || s->mech.n_reac <= 0)
    return fail("chemistry needs a reacting mechanism of 4 or 9 species");
  if (mode == 1 && s->n_spec != 9) return fail("the surrogate is built for the 9-species mechanism");
  return 0;
}

int ostc_chem_batch(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y, double dt,
                    double *dY, double *dT, int mode, int64_t *n_fallback) {
  if (!s) return fail("null solver");
  if (need_chem(s, mode)) return 1;
  if (n < 0) return fail("negative cell count");
  HIPCHK(hipSetDevice(s->device));
  if (mode == 0) {
    ostc_solver::Timer t(s, "chem_batch_fp64");
    HIPCHK(launch_chem_batch(s->d_mech, s->mech.n_reac, n, rho, T, Y, dt, dY, dT, s->stream, s->n_spec));
    if (n_fallback) *n_fallback = 0;
  } else if (mode == 1) {
    if (!s->sur || !surrogate_ready(s->sur)) return fail("no trained surrogate: train or load one first");
    ostc_solver::Timer t(s, "chem_batch_surrogate");
    int64_t nf = 0;
    std::string err;
    if (surrogate_infer(s->sur, s->d_mech, n, rho, T, Y, dt, dY, dT, &nf, s->stream, err)) return fail(err);
    if (n_fallback) *n_fallback = nf;
  } else {
    return fail("mode must be 0 (point-implicit) or 1 (surrogate)");
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  s->drain_profile();
  return 0;
}

int ostc_chem_batch_noalloc(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y,
                            double dt, double *dY, double *dT, int mode, int64_t *n_fallback) {
  if (!s) return fail("null solver");
  if (s->device < 0) return fail("solver has no device");
  if (need_chem(s, mode)) return 1;
  if (n < 0) return fail("negative cell count");
  if (n > kNoAllocCells)
    return fail("ostc_chem_batch_noalloc takes at most " + std::to_string(kNoAllocCells) + " cells, got " +
                std::to_string(n));
  HIPCHK(hipSetDevice(s->device));
  if (mode == 0) {
    HIPCHK(launch_chem_batch(s->d_mech, s->mech.n_reac, n, rho, T, Y, dt, dY, dT, s->stream, s->n_spec));
    if (n_fallback) *n_fallback = 0;
  } else if (mode == 1) {
    if (!s->sur || !surrogate_ready(s->sur)) return fail("no trained surrogate: train or load one first");
    int64_t nf = 0;
    std::string err;
    if (surrogate_infer_noalloc(s->sur, s->d_mech, n, rho, T, Y, dt, dY, dT, &nf, s->stream, err)) return fail(err);
    if (n_fallback) *n_fallback = nf;
  } else {
    return fail("mode must be 0 (point-implicit) or 1 (surrogate)");
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_chem_states(ostc_solver *s, int blk, double *rho, double *T, double *Y) {
  if (check_blk(s, blk)) return 1;
  if (!rho || !T || !Y) return fail("null argument");
  if (!s->P.mix) return fail("chemistry states need a multi-species mechanism");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(launch_chem_gather(s->blocks[blk].z, s->P, rho, T, Y, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_set_steady_chem(ostc_solver *s, double dt, int mode) {
  if (!s) return fail("null solver");
  if (dt > 0) {
    if (!(s->P.mix && s->P.reaction == 1)) return fail("inline chemistry needs a reacting mechanism");
    if (mode != 0 && mode != 1) return fail("mode must be 0 (point-implicit) or 1 (surrogate)");
    if (need_chem(s, mode)) return 1;
    if (mode == 1 && (!s->sur || !surrogate_ready(s->sur))) return fail("no trained surrogate: train or load one first");
    s->chem_split = true;
    s->steady_chem_dt = dt;
    s->steady_chem_mode = mode;
  } else {
    s->chem_split = false;
    s->steady_chem_dt = 0;
    s->steady_chem_mode = -1;
  }
  s->steady_cells = s->steady_fb = 0;
  return 0;
}

int ostc_steady_chem_stats(ostc_solver *s, int64_t *cells, int64_t *fallback) {
  if (!s) return fail("null solver");
  if (cells) *cells = s->steady_cells;
  if (fallback) *fallback = s->steady_fb;
  return 0;
}

// ---------------------------------------------------------------- field files
// Variable names of the output file (FieldIO.h:422-455): coordinates, the
// basic variables and the Mach number, the species, k and omega, mu_t.
static std::vector<std::string> output_names(const ostc_solver *s) {
  std::vector<std::string> v{"x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"};
  for (int l = 0; l < s->n_spec; ++l) v.push_back(s->mech.names[l]);
  if (s->n_turb) {
    v.push_back("tke");
    v.push_back("omega");
    v.push_back("mut");
  }
  return v;
}

int ostc_write_flowfield(ostc_solver *s, const char *root) {
  if (check_blk(s, 0)) return 1;
  if (!root) return fail("null output directory");
  HIPCHK(hipSetDevice(s->device));
  const std::string base = std::string(root) + "/output";
  std::error_code ec;
  std::filesystem::create_directories(base + "/field", ec);
  std::filesystem::create_directories(base + "/message", ec);
  if (ec) return fail("cannot create " + base + ": " + ec.message());
  const auto names = output_names(s);
  // interior points of every block (the reference writes ngg_output = 0)
  std::vector<std::vector<double>> store;
  std::vector<PltZone> zones;
  for (auto &b : s->blocks) {
    const int64_t n = (int64_t)b.mx * b.my * b.mz;
    auto take = [&](const double *dev, int comps) -> int {
      std::vector<double> h((size_t)b.z.sz * comps);
      HIPCHK(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
      for (int c = 0; c < comps; ++c) {
        std::vector<double> v((size_t)n);
        size_t q = 0;
        for (int k = 0; k < b.mz; ++k)
          for (int j = 0; j < b.my; ++j)
            for (int i = 0; i < b.mx; ++i) v[q++] = h[(size_t)c * b.z.sz + b.z.at(i, j, k)];
        store.push_back(std::move(v));
      }
      return 0;
    };
    const size_t first = store.size();
    store.push_back(b.x0);
    store.push_back(b.y0);
    store.push_back(b.z0);
    if (take(b.z.bv, 6) || take(b.z.mach, 1)) return 1;
    if (s->n_scalar > 0 && take(b.z.sv, s->n_scalar)) return 1;
    if (s->n_turb && take(b.z.mut, 1)) return 1;
    PltZone zn;
    zn.mx = b.mx;
    zn.my = b.my;
    zn.mz = b.mz;
    for (size_t q = first; q < store.size(); ++q) zn.vars.push_back(nullptr);
    zones.push_back(zn);
  }
  // pointers after all arrays are in place
  size_t q = 0;
  for (auto &zn : zones)
    for (auto &p : zn.vars) p = store[q++].data();
  std::string err;
  if (!plt_write(base + "/field/flowfield.plt", names, zones, s->t_phys, err)) return fail(err);
  // FieldIO.h:457-462 and Residual.cuh:123-139: the step and the residual scale
  std::ofstream st(base + "/message/step.txt");
  st << s->step;
  if (!st) return fail("cannot write " + base + "/message/step.txt");
  if (s->res_scale_set) {
    double sc[4];
    HIPCHK(hipMemcpyAsync(sc, s->w.res_scale, sizeof(sc), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    std::ofstream rs(base + "/message/residual_scale.txt");
    rs.precision(17);
    rs << sc[0] << '\n' << sc[1] << '\n' << sc[2] << '\n' << sc[3] << '\n';
  }
  return 0;
}

int ostc_read_flowfield(ostc_solver *s, const char *root) {
  if (check_blk(s, 0)) return 1;
  if (!root) return fail("null output directory");
  HIPCHK(hipSetDevice(s->device));
  const std::string base = std::string(root) + "/output";
  PltFile f;
  std::string err;
  if (!plt_read(base + "/field/flowfield.plt", f, err)) return fail(err);
  if (f.dims.size() != s->blocks.size())
    return fail("the flow-field file has " + std::to_string(f.dims.size()) + " zones, the case " +
                std::to_string(s->blocks.size()) + " blocks");
  // Initialize.cuh:238-313 identify_variable_labels: 0-5 basic variables,
  // 6 + l species l, 6 + n_spec + t turbulence variable t; -1 unused
  const int ns = s->n_spec, nt = s->n_turb;
  bool have_spec = false, have_turb = false;
  std::vector<int> label;
  for (auto &name : f.names) {
    std::string n = name;
    for (char &c : n) c = (char)std::toupper((unsigned char)c);
    int l = -1;
    if (n == "DENSITY" || n == "ROE" || n == "RHO") l = 0;
    else if (n == "U") l = 1;
    else if (n == "V") l = 2;
    else if (n == "W") l = 3;
    else if (n == "P" || n == "PRESSURE") l = 4;
    else if (n == "T" || n == "TEMPERATURE") l = 5;
    else {
      for (int q = 0; q < ns; ++q) {
        std::string sp = s->mech.names[q];
        for (char &c : sp) c = (char)std::toupper((unsigned char)c);
        if (n == sp) {
          l = 6 + q;
          have_spec = true;
          break;
        }
      }
      if (nt == 2 && (n == "K" || n == "TKE")) {
        l = 6 + ns;
        have_turb = true;
      } else if (nt == 2 && n == "OMEGA") {
        l = 6 + ns + 1;
        have_turb = true;
      }
    }
    label.push_back(l);
  }
  Inflow in;
  if ((ns && !have_spec) || (nt && !have_turb)) {
    if (make_inflow(s, s->prm.get_string("default_init", s->prm.get_string("reference_state", "")), in)) return 1;
  }
  for (size_t bi = 0; bi < s->blocks.size(); ++bi) {
    Block &b = s->blocks[bi];
    const auto &d = f.dims[bi];
    // a file written with ngg_file ghost layers (FieldIO.h ngg_output)
    const int gf = (d[0] - b.mx) / 2;
    if (gf < 0 || gf > s->ngg || d[0] != b.mx + 2 * gf || d[1] != b.my + 2 * gf ||
        (b.mz > 1 ? d[2] != b.mz + 2 * gf : (d[2] != 1 && d[2] != 1 + 2 * gf)))
      return fail("zone " + std::to_string(bi) + " of the flow-field file does not fit block " + std::to_string(bi));
    const int gk = d[2] == b.mz + 2 * gf ? gf : 0;
    std::vector<double> bv((size_t)b.z.sz * 6), sv((size_t)b.z.sz * std::max(s->n_scalar, 1));
    HIPCHK(hipMemcpyAsync(bv.data(), b.z.bv, bv.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    if (s->n_scalar > 0)
      HIPCHK(hipMemcpyAsync(sv.data(), b.z.sv, sv.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t v = 0; v < label.size(); ++v) {
      const int l = label[v];
      if (l < 0) continue;
      double *dst = l < 6 ? bv.data() + (size_t)l * b.z.sz : sv.data() + (size_t)(l - 6) * b.z.sz;
      const std::vector<double> &src = f.data[bi][v];
      size_t q = 0;
      for (int k = -gk; k < b.mz + gk; ++k)
        for (int j = -gf; j < b.my + gf; ++j)
          for (int i = -gf; i < b.mx + gf; ++i) dst[b.z.at(i, j, k)] = src[q++];
    }
    // Initialize.cu:49-121: groups the file lacks take the default inflow state
    for (int k = 0; k < b.mz; ++k)
      for (int j = 0; j < b.my; ++j)
        for (int i = 0; i < b.mx; ++i) {
          const int64_t p = b.z.at(i, j, k);
          if (ns && !have_spec)
            for (int l = 0; l < ns; ++l) sv[(size_t)l * b.z.sz + p] = in.sv[l];
          if (nt && !have_turb)
            for (int l = 0; l < nt; ++l) sv[(size_t)(ns + l) * b.z.sz + p] = in.sv[ns + l];
        }
    HIPCHK(hipMemcpyAsync(b.z.bv, bv.data(), bv.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    if (s->n_scalar > 0)
      HIPCHK(hipMemcpyAsync(b.z.sv, sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  // Initialize.cuh:234-237 step, Driver.cu:21-27 residual scale
  int step = 0;
  std::ifstream st(base + "/message/step.txt");
  if (st) st >> step;
  std::ifstream rs(base + "/message/residual_scale.txt");
  if (rs) {
    double sc[4] = {1e-20, 1e-20, 1e-20, 1e-20};
    rs >> sc[0] >> sc[1] >> sc[2] >> sc[3];
    HIPCHK(hipMemcpyAsync(s->w.res_scale, sc, sizeof(sc), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->res_scale_set = true;
  }
  s->restart = true;
  s->restart_step = step;
  s->restart_time = f.time;
  return 0;
}

int ostc_set_chem_split(ostc_solver *s, int on) {
  if (!s) return fail("null solver");
  if (on && !(s->P.mix && s->P.reaction == 1)) return fail("operator-split chemistry needs a reacting mixture");
  s->chem_split = on != 0;
  return 0;
}

// one chemistry step dt of every interior cell of a block (gather, FP64
// integrator or surrogate + fallback, scatter); *nf: fallback cells
int chem_interior(ostc_solver *s, Block &b, double dt, int mode, int64_t *nf_out) {
  const int64_t n = (int64_t)b.mx * b.my * b.mz;
  if (n > s->chem_cap) {
    if (s->chem_buf) HIPCHK(hipFree(s->chem_buf));
    s->chem_buf = nullptr;
    s->chem_cap = 0;
    HIPCHK(hipMalloc(&s->chem_buf, sizeof(double) * 21 * n));
    s->chem_cap = n;
  }
  double *rho = s->chem_buf, *T = rho + n, *Y = T + n, *dY = Y + 9 * n, *dT = dY + 9 * n;
  int64_t nf = 0;
  {
    ostc_solver::Timer t(s, "chem_step");
    HIPCHK(launch_chem_gather(b.z, s->P, rho, T, Y, s->stream));
    if (mode == 0) {
      HIPCHK(launch_chem_batch(s->d_mech, s->mech.n_reac, n, rho, T, Y, dt, dY, dT, s->stream, s->n_spec));
    } else {
      std::string err;
      if (surrogate_infer(s->sur, s->d_mech, n, rho, T, Y, dt, dY, dT, &nf, s->stream, err)) return fail(err);
    }
    HIPCHK(launch_chem_scatter(b.z, s->P, s->d_mech, dY, dT, s->stream));
  }
  *nf_out = nf;
  return 0;
}

int ostc_chem_step(ostc_solver *s, int blk, double dt, int mode, int64_t *n_fallback) {
  if (check_blk(s, blk)) return 1;
  if (!(dt > 0)) return fail("dt must be positive");
  if (mode != 0 && mode != 1) return fail("mode must be 0 (point-implicit) or 1 (surrogate)");
  if (need_chem(s, mode)) return 1;
  if (mode == 1 && (!s->sur || !surrogate_ready(s->sur))) return fail("no trained surrogate: train or load one first");
  HIPCHK(hipSetDevice(s->device));
  Block &b = s->blocks[blk];
  int64_t nf = 0;
  if (chem_interior(s, b, dt, mode, &nf)) return 1;
  if (run_bcs(s, b, true)) return 1;
  if (s->P.dim == 2) HIPCHK(launch_eliminate_k(b.z, s->P, true, s->stream));
  if (properties_block(s, b, true)) return 1;
  HIPCHK(hipStreamSynchronize(s->stream));
  s->drain_profile();
  if (n_fallback) *n_fallback = nf;
  return 0;
}

static int need_sur(ostc_solver *s) {
  if (!s) return fail("null solver");
  if (s->n_spec != 9 || s->mech.n_reac <= 0) return fail("the surrogate needs the 9-species reacting mechanism");
  if (s->device < 0) return fail("the surrogate needs a device");
  if (!s->sur) {
    std::string err;
    if (surrogate_create(s->mech, s->device, &s->sur, err)) return fail(err);
  }
  return 0;
}

#define SURCALL(expr)                  \
  do {                                 \
    std::string err;                   \
    if (expr) return fail(err);        \
  } while (0)

int ostc_surrogate_init(ostc_solver *s, uint64_t seed) {
  if (need_sur(s)) return 1;
  SURCALL(surrogate_init_params(s->sur, seed, s->stream, err));
  return 0;
}

int ostc_surrogate_param_count(ostc_solver *s, int64_t *n) {
  if (need_sur(s)) return 1;
  if (!n) return fail("null argument");
  *n = surrogate_param_count(s->sur);
  return 0;
}

int ostc_surrogate_get_params(ostc_solver *s, float *host) {
  if (need_sur(s)) return 1;
  if (!host) return fail("null argument");
  SURCALL(surrogate_get_params(s->sur, host, s->stream, err));
  return 0;
}

int ostc_surrogate_set_params(ostc_solver *s, const float *host) {
  if (need_sur(s)) return 1;
  if (!host) return fail("null argument");
  SURCALL(surrogate_set_params(s->sur, host, s->stream, err));
  return 0;
}

int ostc_surrogate_forward(ostc_solver *s, int64_t B, const float *x, float *y) {
  if (need_sur(s)) return 1;
  if (!x || !y) return fail("null argument");
  SURCALL(surrogate_forward(s->sur, B, x, y, s->stream, err));
  return 0;
}

int ostc_surrogate_gradient(ostc_solver *s, int64_t B, const float *x, const float *t, float *grad, double *loss) {
  if (need_sur(s)) return 1;
  if (!x || !t || !grad) return fail("null argument");
  SURCALL(surrogate_gradient(s->sur, B, x, t, grad, loss, s->stream, err));
  return 0;
}

int ostc_surrogate_train(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y, double dt,
                         const ostc_surrogate_train_cfg *cfg, double *loss_hist) {
  if (need_sur(s)) return 1;
  SurrogateTrainCfg c;
  if (cfg) {
    c.epochs = cfg->epochs;
    c.batch = cfg->batch;
    c.lr = cfg->lr;
    c.lr_min = cfg->lr_min;
    c.beta1 = cfg->beta1;
    c.beta2 = cfg->beta2;
    c.eps = cfg->eps;
    c.seed = cfg->seed;
  }
  if (c.epochs < 1) return fail("epochs must be >= 1");
  ostc_solver::Timer t(s, "surrogate_train");
  SURCALL(surrogate_train(s->sur, s->d_mech, n, rho, T, Y, dt, c, loss_hist, s->stream, err));
  return 0;
}

int ostc_surrogate_set_gate(ostc_solver *s, const ostc_surrogate_gate *g) {
  if (need_sur(s)) return 1;
  if (!g) return fail("null argument");
  if (!(g->cons_tol >= 0) || !(g->dy_max >= 0) || !(g->dT_max >= 0) || !(g->margin >= 0) || !(g->sig_min >= 0))
    return fail("gate thresholds must be non-negative");
  SurrogateGate q;
  q.cons_tol = g->cons_tol;
  q.dy_max = g->dy_max;
  q.dT_max = g->dT_max;
  q.margin = g->margin;
  q.sig_min = g->sig_min;
  surrogate_set_gate(s->sur, q);
  return 0;
}

int ostc_surrogate_get_gate(ostc_solver *s, ostc_surrogate_gate *g) {
  if (need_sur(s)) return 1;
  if (!g) return fail("null argument");
  const SurrogateGate q = surrogate_get_gate(s->sur);
  g->cons_tol = q.cons_tol;
  g->dy_max = q.dy_max;
  g->dT_max = q.dT_max;
  g->margin = q.margin;
  g->sig_min = q.sig_min;
  return 0;
}

int ostc_surrogate_norm(ostc_solver *s, double *out) {
  if (need_sur(s)) return 1;
  if (!out) return fail("null argument");
  return surrogate_get_norm(s->sur, out);
}

int ostc_surrogate_save(ostc_solver *s, const char *path) {
  if (need_sur(s)) return 1;
  if (!path) return fail("null path");
  SURCALL(surrogate_save(s->sur, path, s->stream, err));
  return 0;
}

int ostc_surrogate_load(ostc_solver *s, const char *path) {
  if (need_sur(s)) return 1;
  if (!path) return fail("null path");
  SURCALL(surrogate_load(s->sur, path, s->stream, err));
  return 0;
}

int ostc_surrogate_set_kernel(ostc_solver *s, int kernel) {
  if (need_sur(s)) return 1;
  if (kernel < 0 || kernel > 3)
    return fail("kernel must be 0 (layer GEMMs), 1 (fused, 128 cells), 2 (fused, 64) or 3 (fused, 128, pipelined)");
  surrogate_set_kernel(s->sur, kernel);
  return 0;
}

int ostc_surrogate_network_outputs(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y,
                                   float *host_out) {
  if (need_sur(s)) return 1;
  if (!host_out) return fail("null argument");
  SURCALL(surrogate_network_outputs(s->sur, s->d_mech, n, rho, T, Y, host_out, s->stream, err));
  return 0;
}

int ostc_surrogate_last_stats(ostc_solver *s, double *out) {
  if (need_sur(s)) return 1;
  if (!out) return fail("null argument");
  return surrogate_last_stats(s->sur, out);
}

int ostc_surrogate_gate_reasons(ostc_solver *s, int64_t *out) {
  if (need_sur(s)) return 1;
  if (!out) return fail("null argument");
  std::string err;
  if (surrogate_gate_reasons(s->sur, out, s->stream, err)) return fail(err);
  return 0;
}

int ostc_dev_alloc(ostc_solver *s, int64_t bytes, void **dptr) {
  if (!s || !dptr) return fail("null argument");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipMalloc(dptr, std::max<int64_t>(bytes, 8)));
  s->user_mem.push_back(*dptr);
  return 0;
}

int ostc_dev_free(ostc_solver *s, void *dptr) {
  if (!s) return fail("null solver");
  auto it = std::find(s->user_mem.begin(), s->user_mem.end(), dptr);
  if (it == s->user_mem.end()) return fail("pointer not allocated by ostc_dev_alloc");
  s->user_mem.erase(it);
  HIPCHK(hipFree(dptr));
  return 0;
}

int ostc_memcpy_h2d(ostc_solver *s, void *dst, const void *src, int64_t bytes) {
  if (!s) return fail("null solver");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_memcpy_d2h(ostc_solver *s, void *dst, const void *src, int64_t bytes) {
  if (!s) return fail("null solver");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_memcpy_d2d(ostc_solver *s, void *dst, const void *src, int64_t bytes) {
  if (!s) return fail("null solver");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_sync(ostc_solver *s) {
  if (!s) return fail("null solver");
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int ostc_profile(ostc_solver *s, int enable) {
  if (!s) return fail("null solver");
  s->drain_profile();
  s->prof = enable != 0;
  if (enable) s->stats.clear();
  return 0;
}

int ostc_profile_report(ostc_solver *s, char *buf, int len) {
  if (!s || !buf || len <= 0) return fail("null argument");
  s->drain_profile();
  std::ostringstream o;
  for (auto &[name, v] : s->stats) o << name << " " << v.first << " " << v.second << "\n";
  const std::string r = o.str();
  std::snprintf(buf, (size_t)len, "%s", r.c_str());
  return 0;
}

}  // extern "C"
