// sht_restart.cpp -- the restart state of the dynamical core as a netCDF
// classic file (64-bit offset form, "CDF\2"), written and read without a
// netCDF library.
//
// What it holds is what DCPAM hands RestartFileOutput and takes back from
// RestartFileGet (atm/dcpam_main_mod.f90:3063-3066, :4134-4137): the grid
// fields U, V, Temp, the tracer mixing ratios QMix and the surface pressure
// Ps at the two time levels B (t - dt) and N (t) of the leapfrog scheme,
// arrays xyz(0:im-1, 1:jm, 1:km), xyzf(0:im-1, 1:jm, 1:km, 1:nq) and
// xy(0:im-1, 1:jm) (:4120-4130), plus the coordinates and the restart time.
// DCPAM writes it through gtool5 (outside the tree), whose variable names and
// attributes are not in the reference; the names here follow the Fortran
// arguments (UB, VB, TempB, QMixB, PsB, UN, ...), each with long_name and
// units attributes.  File layout (all big-endian): magic, numrecs = 0, the
// dimensions lon, lat, sig, nq, the global attributes, the variables (the
// coordinates, then the ten fields); data follow the header in that order,
// lon fastest, as Fortran stores the arrays.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sht_host.h"

namespace sht {
namespace host {
namespace {

constexpr uint32_t kDimTag = 0x0A, kVarTag = 0x0B, kAttTag = 0x0C;
constexpr uint32_t kChar = 2, kDouble = 6;

struct Field {
  const char *name, *long_name, *units;
  int rank;  // 2: (lat, lon); 3: (sig, lat, lon); 4: (nq, sig, lat, lon)
};
// the ten fields in the order of RestartFileOutput's arguments
const Field kFields[10] = {
    {"UB", "eastward wind (time level B)", "m s-1", 3},
    {"VB", "northward wind (time level B)", "m s-1", 3},
    {"TempB", "temperature (time level B)", "K", 3},
    {"QMixB", "mass mixing ratio of tracers (time level B)", "1", 4},
    {"PsB", "surface pressure (time level B)", "Pa", 2},
    {"UN", "eastward wind (time level N)", "m s-1", 3},
    {"VN", "northward wind (time level N)", "m s-1", 3},
    {"TempN", "temperature (time level N)", "K", 3},
    {"QMixN", "mass mixing ratio of tracers (time level N)", "1", 4},
    {"PsN", "surface pressure (time level N)", "Pa", 2},
};

void put32(std::string &b, uint32_t v) {
  for (int s = 24; s >= 0; s -= 8) b.push_back((char)((v >> s) & 0xff));
}
void put64(std::string &b, uint64_t v) {
  for (int s = 56; s >= 0; s -= 8) b.push_back((char)((v >> s) & 0xff));
}
void pad4(std::string &b) {
  while (b.size() % 4) b.push_back('\0');
}
void put_name(std::string &b, const std::string &s) {
  put32(b, (uint32_t)s.size());
  b += s;
  pad4(b);
}
void put_text_att(std::string &b, const char *name, const std::string &v) {
  put_name(b, name);
  put32(b, kChar);
  put32(b, (uint32_t)v.size());
  b += v;
  pad4(b);
}
void put_double_att(std::string &b, const char *name, double v) {
  put_name(b, name);
  put32(b, kDouble);
  put32(b, 1);
  uint64_t u;
  std::memcpy(&u, &v, 8);
  put64(b, u);
}

struct Var {
  std::string name, long_name, units;
  std::vector<int> dims;  // dimension ids, slowest first
  uint64_t count = 0;     // doubles
  const double *src = nullptr;
};

// header of the whole file; begins[] receives each variable's data offset
std::string header(const std::vector<std::pair<std::string, uint32_t>> &dims, double time,
                   const std::vector<Var> &vars, std::vector<uint64_t> &begins) {
  // Two passes: offsets are fixed-width fields, so the header length does not
  // depend on their values.
  std::string b;
  for (int pass = 0; pass < 2; ++pass) {
    const uint64_t hlen = b.size();
    b.clear();
    b += "CDF";
    b.push_back('\2');
    put32(b, 0);  // numrecs: no record dimension
    put32(b, kDimTag);
    put32(b, (uint32_t)dims.size());
    for (const auto &d : dims) {
      put_name(b, d.first);
      put32(b, d.second);
    }
    put32(b, kAttTag);
    put32(b, 3);
    put_text_att(b, "title", "dynamical core restart state (time levels B and N)");
    put_text_att(b, "source", "libsht_hip.so sht_restart_write");
    put_double_att(b, "time", time);
    put32(b, kVarTag);
    put32(b, (uint32_t)vars.size());
    uint64_t off = hlen;
    begins.assign(vars.size(), 0);
    for (size_t i = 0; i < vars.size(); ++i) {
      const Var &v = vars[i];
      put_name(b, v.name);
      put32(b, (uint32_t)v.dims.size());
      for (int d : v.dims) put32(b, (uint32_t)d);
      put32(b, kAttTag);
      put32(b, 2);
      put_text_att(b, "long_name", v.long_name);
      put_text_att(b, "units", v.units);
      put32(b, kDouble);
      const uint64_t bytes = 8 * v.count;
      put32(b, bytes > 0xfffffffcull ? 0xffffffffu : (uint32_t)bytes);  // vsize (4-byte padded)
      put64(b, off);
      begins[i] = off;
      off += bytes;
    }
  }
  return b;
}

std::vector<Var> layout(const RestartDims &d, const double *lon, const double *lat, const double *sig,
                        const double *const *fields) {
  std::vector<Var> v;
  const uint64_t nxy = (uint64_t)d.im * d.jm, nxyz = nxy * d.km;
  v.push_back({"lon", "longitude", "radian", {0}, (uint64_t)d.im, lon});
  v.push_back({"lat", "latitude (Gaussian, south to north)", "radian", {1}, (uint64_t)d.jm, lat});
  v.push_back({"sig", "sigma level", "1", {2}, (uint64_t)d.km, sig});
  for (int i = 0; i < 10; ++i) {
    const Field &f = kFields[i];
    Var x{f.name, f.long_name, f.units, {}, 0, fields ? fields[i] : nullptr};
    if (f.rank == 2) {
      x.dims = {1, 0};
      x.count = nxy;
    } else if (f.rank == 3) {
      x.dims = {2, 1, 0};
      x.count = nxyz;
    } else {
      x.dims = {3, 2, 1, 0};
      x.count = nxyz * d.nq;
    }
    v.push_back(x);
  }
  return v;
}

bool write_doubles(FILE *f, const double *p, uint64_t n) {
  std::vector<unsigned char> buf;
  const uint64_t chunk = 1 << 16;
  buf.resize(8 * chunk);
  for (uint64_t i0 = 0; i0 < n; i0 += chunk) {
    const uint64_t m = n - i0 < chunk ? n - i0 : chunk;
    for (uint64_t i = 0; i < m; ++i) {
      uint64_t u;
      std::memcpy(&u, p + i0 + i, 8);
      for (int s = 0; s < 8; ++s) buf[8 * i + s] = (unsigned char)(u >> (56 - 8 * s));
    }
    if (std::fwrite(buf.data(), 1, 8 * m, f) != 8 * m) return false;
  }
  return true;
}

// ---- reading
struct Reader {
  const std::vector<unsigned char> &b;
  size_t pos = 0;
  bool ok = true;
  uint32_t u32() {
    if (pos + 4 > b.size()) {
      ok = false;
      return 0;
    }
    const uint32_t v = (uint32_t)b[pos] << 24 | (uint32_t)b[pos + 1] << 16 | (uint32_t)b[pos + 2] << 8 | b[pos + 3];
    pos += 4;
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return hi << 32 | u32();
  }
  std::string name() {
    const uint32_t n = u32();
    if (!ok || n > b.size() - pos) {
      ok = false;
      return {};
    }
    std::string s(reinterpret_cast<const char *>(&b[pos]), n);
    pos += (n + 3) & ~3u;
    if (pos > b.size()) ok = false;
    return s;
  }
  void skip_atts() {
    const uint32_t tag = u32(), n = u32();
    if (!ok || (tag != kAttTag && !(tag == 0 && n == 0))) {
      ok = false;
      return;
    }
    for (uint32_t i = 0; i < n && ok; ++i) {
      (void)name();
      const uint32_t type = u32(), cnt = u32();
      static const uint32_t size[7] = {0, 1, 1, 2, 4, 4, 8};
      if (!ok || type < 1 || type > 6 || cnt > b.size()) {
        ok = false;
        return;
      }
      pos += ((uint64_t)cnt * size[type] + 3) & ~3ull;
      if (pos > b.size()) ok = false;
    }
  }
};

}  // namespace

int restart_write(const char *path, const RestartDims &d, double time, const double *lon, const double *lat,
                  const double *sig, const double *const *fields, std::string &msg) {
  auto bad = [&](int code, const std::string &why) {
    msg = "sht_restart_write: " + why;
    return code;
  };
  if (!path || !lon || !lat || !sig || !fields) return bad(SHT_ERR_INVALID, "NULL argument");
  if (d.im < 1 || d.jm < 1 || d.km < 1 || d.nq < 1) return bad(SHT_ERR_INVALID, "dimensions must be >= 1");
  for (int i = 0; i < 10; ++i)
    if (!fields[i]) return bad(SHT_ERR_INVALID, std::string("field ") + kFields[i].name + " is NULL");
  const std::vector<std::pair<std::string, uint32_t>> dims = {
      {"lon", (uint32_t)d.im}, {"lat", (uint32_t)d.jm}, {"sig", (uint32_t)d.km}, {"nq", (uint32_t)d.nq}};
  const std::vector<Var> vars = layout(d, lon, lat, sig, fields);
  std::vector<uint64_t> begins;
  const std::string h = header(dims, time, vars, begins);
  const std::string tmp = std::string(path) + ".tmp";
  FILE *f = std::fopen(tmp.c_str(), "wb");
  if (!f) return bad(SHT_ERR_INVALID, "cannot open " + tmp);
  bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
  for (const Var &v : vars) ok = ok && write_doubles(f, v.src, v.count);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return bad(SHT_ERR_INVALID, std::string("cannot write ") + path);
  }
  return SHT_OK;
}

int restart_read(const char *path, const RestartDims &d, double *time, double *const *fields, std::string &msg) {
  auto bad = [&](int code, const std::string &why) {
    msg = "sht_restart_read: " + why;
    return code;
  };
  if (!path || !fields) return bad(SHT_ERR_INVALID, "NULL argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return bad(SHT_ERR_INVALID, std::string("cannot open ") + path);
  std::vector<unsigned char> b;
  unsigned char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof chunk, f)) > 0) b.insert(b.end(), chunk, chunk + n);
  std::fclose(f);
  Reader r{b};
  if (b.size() < 8 || std::memcmp(b.data(), "CDF", 3) != 0 || (b[3] != 1 && b[3] != 2))
    return bad(SHT_ERR_INVALID, "not a netCDF classic file");
  const bool wide = b[3] == 2;
  r.pos = 4;
  if (r.u32() != 0) return bad(SHT_ERR_UNSUPPORTED, "record variables are not read");
  // dimensions
  std::vector<std::pair<std::string, uint32_t>> dims;
  uint32_t tag = r.u32(), nd = r.u32();
  if (!r.ok || (tag != kDimTag && !(tag == 0 && nd == 0)) || nd > 1024) return bad(SHT_ERR_INVALID, "bad dimension list");
  for (uint32_t i = 0; i < nd && r.ok; ++i) {
    std::string nm = r.name();
    dims.push_back({nm, r.u32()});
  }
  // global attributes: the restart time
  const size_t gatt = r.pos;
  r.skip_atts();
  if (!r.ok) return bad(SHT_ERR_INVALID, "bad attribute list");
  if (time) {
    *time = 0.0;
    Reader a{b};
    a.pos = gatt + 4;
    const uint32_t na = a.u32();
    for (uint32_t i = 0; i < na && a.ok; ++i) {
      const std::string nm = a.name();
      const uint32_t type = a.u32(), cnt = a.u32();
      static const uint32_t size[7] = {0, 1, 1, 2, 4, 4, 8};
      if (!a.ok || type < 1 || type > 6) break;
      if (nm == "time" && type == kDouble && cnt == 1) {
        const uint64_t u = a.u64();
        std::memcpy(time, &u, 8);
        break;
      }
      a.pos += ((uint64_t)cnt * size[type] + 3) & ~3ull;
    }
  }
  // variables
  tag = r.u32();
  const uint32_t nv = r.u32();
  if (!r.ok || (tag != kVarTag && !(tag == 0 && nv == 0)) || nv > 4096) return bad(SHT_ERR_INVALID, "bad variable list");
  const uint64_t nxy = (uint64_t)d.im * d.jm, nxyz = nxy * d.km;
  int found = 0;
  for (uint32_t i = 0; i < nv && r.ok; ++i) {
    const std::string nm = r.name();
    const uint32_t rank = r.u32();
    if (!r.ok || rank > 8) return bad(SHT_ERR_INVALID, "bad variable " + nm);
    uint64_t count = 1;
    std::vector<uint32_t> lens;
    for (uint32_t k = 0; k < rank; ++k) {
      const uint32_t id = r.u32();
      if (!r.ok || id >= dims.size()) return bad(SHT_ERR_INVALID, "bad dimension id in " + nm);
      lens.push_back(dims[id].second);
      count *= dims[id].second;
    }
    r.skip_atts();
    const uint32_t type = r.u32();
    (void)r.u32();  // vsize
    const uint64_t begin = wide ? r.u64() : r.u32();
    if (!r.ok) return bad(SHT_ERR_INVALID, "truncated header");
    for (int fi = 0; fi < 10; ++fi) {
      if (nm != kFields[fi].name) continue;
      const uint64_t want = kFields[fi].rank == 2 ? nxy : kFields[fi].rank == 3 ? nxyz : nxyz * d.nq;
      if (type != kDouble || count != want || lens.back() != (uint32_t)d.im)
        return bad(SHT_ERR_INVALID, std::string("variable ") + nm + " does not match the requested grid");
      if (begin > b.size() || 8 * count > b.size() - begin)
        return bad(SHT_ERR_INVALID, std::string("variable ") + nm + " runs past the end of the file");
      if (fields[fi]) {
        const unsigned char *p = &b[begin];
        for (uint64_t e = 0; e < count; ++e) {
          uint64_t u = 0;
          for (int s = 0; s < 8; ++s) u = u << 8 | p[8 * e + s];
          std::memcpy(fields[fi] + e, &u, 8);
        }
      }
      ++found;
    }
  }
  if (!r.ok) return bad(SHT_ERR_INVALID, "truncated header");
  if (found != 10) return bad(SHT_ERR_INVALID, "the file lacks some of the ten restart fields");
  return SHT_OK;
}

}  // namespace host
}  // namespace sht

using namespace sht;

extern "C" {

int sht_restart_write(const char *path, const sht_restart_desc *d, const double *const fields[10]) {
  if (!d) {
    host::set_error("sht_restart_write: desc is NULL");
    return SHT_ERR_INVALID;
  }
  std::string msg;
  host::RestartDims dims{d->im, d->jm, d->km, d->nq};
  const int rc = host::restart_write(path, dims, d->time, d->lon, d->lat, d->sig, fields, msg);
  if (rc) host::set_error(msg);
  return rc;
}

int sht_restart_read(const char *path, const sht_restart_desc *d, double *time, double *const fields[10]) {
  if (!d) {
    host::set_error("sht_restart_read: desc is NULL");
    return SHT_ERR_INVALID;
  }
  std::string msg;
  host::RestartDims dims{d->im, d->jm, d->km, d->nq};
  const int rc = host::restart_read(path, dims, time, fields, msg);
  if (rc) host::set_error(msg);
  return rc;
}

}  // extern "C"
