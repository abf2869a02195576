// Tecplot binary (V112) files of ordered zones with nodal double data: the
// restart and output format the reference writes (FieldIO.h:58-186 header,
// 206-420 data section, 456-600 per-output refresh) and reads back
// (Initialize.cuh:64-230).  Strings are stored as one 32-bit integer per
// character plus a terminating 0; numbers are little-endian.
#include <cstdint>
#include <cstring>
#include <fstream>

#include "ostc_internal.h"

namespace ostc {

namespace {

struct Out {
  std::ofstream f;
  void i32(int32_t v) { f.write((const char *)&v, 4); }
  void f32(float v) { f.write((const char *)&v, 4); }
  void f64(double v) { f.write((const char *)&v, 8); }
  void str(const std::string &s) {
    for (unsigned char c : s) i32((int32_t)c);
    i32(0);
  }
};

struct In {
  const std::vector<char> &b;
  size_t o = 0;
  bool ok = true;
  explicit In(const std::vector<char> &buf) : b(buf) {}
  template <typename T>
  T get() {
    T v{};
    if (o + sizeof(T) > b.size()) {
      ok = false;
      return v;
    }
    std::memcpy(&v, b.data() + o, sizeof(T));
    o += sizeof(T);
    return v;
  }
  std::string str() {
    std::string s;
    while (ok) {
      const int32_t c = get<int32_t>();
      if (!ok || c == 0) break;
      s += (char)c;
    }
    return s;
  }
  void skip(size_t n) {
    if (o + n > b.size()) ok = false;
    o += n;
  }
};

}  // namespace

bool plt_write(const std::string &path, const std::vector<std::string> &names, const std::vector<PltZone> &zones,
               double time, std::string &err) {
  Out o;
  o.f.open(path, std::ios::binary | std::ios::trunc);
  if (!o.f) {
    err = "cannot open " + path + " for writing";
    return false;
  }
  const int32_t nv = (int32_t)names.size();
  // header: magic, byte order, file type (full), title, variables
  o.f.write("#!TDV112", 8);
  o.i32(1);
  o.i32(0);
  o.str("Solution file");
  o.i32(nv);
  for (auto &n : names) o.str(n);
  for (size_t z = 0; z < zones.size(); ++z) {
    o.f32(299.0f);                    // zone marker
    o.str("zone " + std::to_string(z));
    o.i32(-1);                        // parent zone
    o.i32(-2);                        // strand id
    o.f64(time);                      // solution time
    o.i32(-1);                        // zone colour
    o.i32(0);                         // ordered
    o.i32(0);                         // nodal data
    o.i32(0);                         // no raw face neighbours
    o.i32(0);                         // no user face connections
    o.i32(zones[z].mx);
    o.i32(zones[z].my);
    o.i32(zones[z].mz);
    o.i32(0);                         // no auxiliary data
  }
  o.f32(357.0f);                      // end of header
  for (const PltZone &zn : zones) {
    const size_t n = (size_t)zn.mx * zn.my * zn.mz;
    if (zn.vars.size() != (size_t)nv) {
      err = "zone variable count differs from the variable names";
      return false;
    }
    o.f32(299.0f);
    for (int l = 0; l < nv; ++l) o.i32(2);  // double
    o.i32(0);                               // no passive variables
    o.i32(0);                               // no variable sharing
    o.i32(-1);                              // no connectivity sharing
    for (int l = 0; l < nv; ++l) {
      const double *v = zn.vars[l];
      double lo = n ? v[0] : 0, hi = n ? v[0] : 0;
      for (size_t q = 1; q < n; ++q) {
        lo = std::min(lo, v[q]);
        hi = std::max(hi, v[q]);
      }
      o.f64(lo);
      o.f64(hi);
    }
    for (int l = 0; l < nv; ++l) o.f.write((const char *)zn.vars[l], (std::streamsize)(n * sizeof(double)));
  }
  if (!o.f) {
    err = "write failed: " + path;
    return false;
  }
  return true;
}

bool plt_read(const std::string &path, PltFile &out, std::string &err) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  const std::streamsize size = f.tellg();
  f.seekg(0);
  std::vector<char> buf((size_t)size);
  if (!f.read(buf.data(), size)) {
    err = "cannot read " + path;
    return false;
  }
  In in(buf);
  if (size < 8 || std::memcmp(buf.data(), "#!TDV", 5) != 0) {
    err = path + ": not a Tecplot binary file";
    return false;
  }
  in.skip(8);
  (void)in.get<int32_t>();  // byte order
  (void)in.get<int32_t>();  // file type
  (void)in.str();           // title
  const int32_t nv = in.get<int32_t>();
  if (!in.ok || nv <= 0 || nv > 4096) {
    err = path + ": bad variable count";
    return false;
  }
  out = PltFile{};
  for (int l = 0; l < nv; ++l) out.names.push_back(in.str());
  while (in.ok) {
    const float marker = in.get<float>();
    if (marker == 357.0f) break;
    if (marker != 299.0f) {
      err = path + ": unexpected zone marker";
      return false;
    }
    (void)in.str();  // zone name
    in.skip(8);      // parent zone, strand id
    out.time = in.get<double>();
    in.skip(20);     // colour, zone type, data location, face neighbours, user connections
    std::array<int, 3> d{};
    for (int &v : d) v = in.get<int32_t>();
    (void)in.get<int32_t>();  // auxiliary data flag (none is written)
    out.dims.push_back(d);
  }
  for (size_t z = 0; z < out.dims.size() && in.ok; ++z) {
    const size_t n = (size_t)out.dims[z][0] * out.dims[z][1] * out.dims[z][2];
    in.skip(4 + 4 * (size_t)nv + 12 + 16 * (size_t)nv);
    std::vector<std::vector<double>> vars((size_t)nv);
    for (int l = 0; l < nv && in.ok; ++l) {
      if (in.o + n * 8 > buf.size()) {
        in.ok = false;
        break;
      }
      vars[l].resize(n);
      std::memcpy(vars[l].data(), buf.data() + in.o, n * 8);
      in.o += n * 8;
    }
    out.data.push_back(std::move(vars));
  }
  if (!in.ok) {
    err = path + ": file ends early";
    return false;
  }
  return true;
}

}  // namespace ostc
