// Point-set loaders (include/hd_io.h): WKT, OFF, PLY and the .bin cache,
// and the reference's normalise / translate preprocessing. Host code only.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hd.h"
#include "hd_internal.h"
#include "hd_io.h"

namespace hd {
namespace {

// --- .bin: uint64 count, then count points (loader.h:150-174) ------------

int ReadBin(const std::string& path, int nd, uint64_t limit,
            std::vector<float>* out) {
  std::ifstream ifs(path, std::ios::binary);
  if (!ifs) return Fail(HD_EINVAL, "cannot open %s", path.c_str());
  uint64_t n = 0;
  ifs.read(reinterpret_cast<char*>(&n), sizeof(n));
  if (!ifs) return Fail(HD_EINVAL, "%s: no point count", path.c_str());
  n = std::min(n, limit);
  out->resize(n * nd);
  ifs.read(reinterpret_cast<char*>(out->data()),
           (std::streamsize)(sizeof(float) * n * nd));
  if (!ifs) return Fail(HD_EINVAL, "%s: fewer points than its count", path.c_str());
  return HD_OK;
}

int WriteBin(const std::string& path, int nd, const std::vector<float>& pts) {
  std::ofstream ofs(path, std::ios::binary);
  if (!ofs) return Fail(HD_EINVAL, "cannot write %s", path.c_str());
  const uint64_t n = pts.size() / nd;
  ofs.write(reinterpret_cast<const char*>(&n), sizeof(n));
  ofs.write(reinterpret_cast<const char*>(pts.data()),
            (std::streamsize)(sizeof(float) * pts.size()));
  if (!ofs) return Fail(HD_EINVAL, "cannot write %s", path.c_str());
  return HD_OK;
}

// --- OFF: "OFF", counts (same or next line), n_dims numbers per vertex
// (loader.h:103-148; like the reference, exactly n_dims numbers are read
// per vertex) ---------------------------------------------------------------

int ReadOff(const std::string& path, int nd, uint64_t limit,
            std::vector<float>* out) {
  std::ifstream ifs(path);
  if (!ifs) return Fail(HD_EINVAL, "cannot open %s", path.c_str());
  std::string line;
  std::getline(ifs, line);
  if (line.compare(0, 3, "OFF") != 0)
    return Fail(HD_EINVAL, "%s: not an OFF file", path.c_str());
  line = line.substr(3);
  if (line.find_first_not_of(" \t\r") == std::string::npos) std::getline(ifs, line);
  std::istringstream hs(line);
  uint64_t nv = 0, nf = 0, ne = 0;
  hs >> nv >> nf >> ne;
  if (nv == 0) return Fail(HD_EINVAL, "%s: no vertex count", path.c_str());
  nv = std::min(nv, limit);
  out->resize(nv * nd);
  for (uint64_t i = 0; i < nv * nd; ++i)
    if (!(ifs >> (*out)[i]))
      return Fail(HD_EINVAL, "%s: fewer vertices than its count", path.c_str());
  return HD_OK;
}

// --- WKT (loader.h:26-101) --------------------------------------------------
// A geometry is a keyword and nested parenthesised lists of coordinate
// tuples. The points kept: every tuple of a POINT, LINESTRING or
// MULTILINESTRING; the first (outer) ring of a POLYGON and of each polygon
// of a MULTIPOLYGON.

struct WktCursor {
  const std::string& s;
  size_t i = 0;
  void Skip() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  bool Eat(char c) {
    Skip();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  bool Number(double* v) {
    Skip();
    const char* b = s.c_str() + i;
    char* e = nullptr;
    *v = std::strtod(b, &e);
    if (e == b) return false;
    i += (size_t)(e - b);
    return true;
  }
};

// "( x y [z], x y [z], ... )": appends the first nd coordinates of each
// tuple when keep; returns false on malformed input.
bool WktTuples(WktCursor& c, int nd, bool keep, std::vector<double>* pts) {
  if (!c.Eat('(')) return false;
  while (true) {
    double v[4] = {0, 0, 0, 0};
    int k = 0;
    double x;
    while (k < 4 && c.Number(&x)) v[k++] = x;
    if (k < 2) return false;
    if (keep)
      for (int d = 0; d < nd; ++d) pts->push_back(d < k ? v[d] : 0.0);
    if (c.Eat(',')) continue;
    return c.Eat(')');
  }
}

// "( ring, ring, ... )": keeps the first ring only when outer_only.
bool WktRings(WktCursor& c, int nd, bool keep, bool outer_only,
              std::vector<double>* pts) {
  if (!c.Eat('(')) return false;
  for (int r = 0;; ++r) {
    if (!WktTuples(c, nd, keep && (!outer_only || r == 0), pts)) return false;
    if (c.Eat(',')) continue;
    return c.Eat(')');
  }
}

bool StartsWith(const std::string& s, const char* p) {
  return s.compare(0, std::strlen(p), p) == 0;
}

int ReadWkt(const std::string& path, int nd, uint64_t limit,
            std::vector<float>* out) {
  std::ifstream ifs(path);
  if (!ifs) return Fail(HD_EINVAL, "cannot open %s", path.c_str());
  std::string line;
  std::vector<double> pts;
  uint64_t line_no = 0;
  while (std::getline(ifs, line) && pts.size() / nd < limit) {
    ++line_no;
    const size_t first = line.find_first_not_of(" \t\r");
    if (first == std::string::npos) continue;
    line = line.substr(first);
    // Longer keywords first (MULTIPOLYGON before POLYGON etc.).
    const char* kw = nullptr;
    for (const char* k : {"MULTIPOLYGON", "MULTILINESTRING", "POLYGON",
                          "LINESTRING", "POINT", "GEOMETRYCOLLECTION"})
      if (StartsWith(line, k)) {
        kw = k;
        break;
      }
    if (kw == nullptr)
      return Fail(HD_EINVAL, "%s:%llu: bad geometry", path.c_str(),
                  (unsigned long long)line_no);
    if (std::strcmp(kw, "GEOMETRYCOLLECTION") == 0) continue;  // loader.h:92
    WktCursor c{line};
    c.i = std::strlen(kw);
    // An optional dimension tag ("POINT Z (...)").
    c.Skip();
    while (c.i < line.size() && std::isalpha((unsigned char)line[c.i])) ++c.i;
    bool ok;
    if (std::strcmp(kw, "POINT") == 0 || std::strcmp(kw, "LINESTRING") == 0) {
      ok = WktTuples(c, nd, true, &pts);
    } else if (std::strcmp(kw, "POLYGON") == 0) {
      ok = WktRings(c, nd, true, true, &pts);
    } else if (std::strcmp(kw, "MULTILINESTRING") == 0) {
      ok = WktRings(c, nd, true, false, &pts);
    } else {  // MULTIPOLYGON: ( polygon, polygon, ... )
      ok = c.Eat('(');
      while (ok) {
        ok = WktRings(c, nd, true, true, &pts);
        if (!ok || !c.Eat(',')) break;
      }
      ok = ok && c.Eat(')');
    }
    if (!ok)
      return Fail(HD_EINVAL, "%s:%llu: malformed %s", path.c_str(),
                  (unsigned long long)line_no, kw);
  }
  const uint64_t n = std::min<uint64_t>(pts.size() / nd, limit);
  out->resize(n * nd);
  for (uint64_t i = 0; i < n * nd; ++i) (*out)[i] = (float)pts[i];
  return HD_OK;
}

// --- PLY (ply_loader.h:37-101) ---------------------------------------------
// Header: "ply", "format <ascii|binary_little_endian|binary_big_endian> 1.0",
// element and property lines (scalar "property <type> <name>" or
// "property list <count type> <item type> <name>"), "end_header". The
// vertex element's x, y (, z) are read in FP64; like the reference, each
// axis is then shifted so that its minimum is 0 (an axis of zero range is
// left as is) and rounded to FP32.

struct PlyProp {
  std::string name;
  int type = 0;        // scalar type, or the list's item type
  int count_type = 0;  // list count type (0: scalar property)
};
struct PlyElement {
  std::string name;
  uint64_t count = 0;
  std::vector<PlyProp> props;
};

// 1..8: int8 uint8 int16 uint16 int32 uint32 float32 float64.
int PlyType(const std::string& t) {
  static const char* names[][2] = {{"char", "int8"},    {"uchar", "uint8"},
                                   {"short", "int16"},  {"ushort", "uint16"},
                                   {"int", "int32"},    {"uint", "uint32"},
                                   {"float", "float32"}, {"double", "float64"}};
  for (int k = 0; k < 8; ++k)
    if (t == names[k][0] || t == names[k][1]) return k + 1;
  return 0;
}
int PlySize(int type) {
  static const int sizes[] = {0, 1, 1, 2, 2, 4, 4, 4, 8};
  return sizes[type];
}

// One binary value as FP64.
bool PlyBinary(std::istream& in, int type, bool big_endian, double* v) {
  unsigned char b[8];
  const int n = PlySize(type);
  in.read(reinterpret_cast<char*>(b), n);
  if (!in) return false;
  if (big_endian) std::reverse(b, b + n);
  switch (type) {
    case 1: { int8_t x; std::memcpy(&x, b, 1); *v = x; break; }
    case 2: { uint8_t x; std::memcpy(&x, b, 1); *v = x; break; }
    case 3: { int16_t x; std::memcpy(&x, b, 2); *v = x; break; }
    case 4: { uint16_t x; std::memcpy(&x, b, 2); *v = x; break; }
    case 5: { int32_t x; std::memcpy(&x, b, 4); *v = x; break; }
    case 6: { uint32_t x; std::memcpy(&x, b, 4); *v = x; break; }
    case 7: { float x; std::memcpy(&x, b, 4); *v = x; break; }
    default: { double x; std::memcpy(&x, b, 8); *v = x; break; }
  }
  return true;
}

int ReadPly(const std::string& path, int nd, uint64_t limit,
            std::vector<float>* out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return Fail(HD_EINVAL, "cannot open %s", path.c_str());
  std::string line;
  std::getline(in, line);
  if (line.compare(0, 3, "ply") != 0)
    return Fail(HD_EINVAL, "%s: not a PLY file", path.c_str());
  int format = -1;  // 0 ascii, 1 little endian, 2 big endian
  std::vector<PlyElement> elems;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string w;
    ls >> w;
    if (w == "format") {
      std::string f;
      ls >> f;
      format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1
             : f == "binary_big_endian" ? 2 : -1;
      if (format < 0) return Fail(HD_EINVAL, "%s: unknown format %s", path.c_str(), f.c_str());
    } else if (w == "element") {
      PlyElement e;
      ls >> e.name >> e.count;
      elems.push_back(e);
    } else if (w == "property") {
      if (elems.empty()) return Fail(HD_EINVAL, "%s: property before element", path.c_str());
      PlyProp p;
      std::string t;
      ls >> t;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> p.name;
        p.count_type = PlyType(ct);
        p.type = PlyType(it);
        if (!p.count_type || !p.type)
          return Fail(HD_EINVAL, "%s: bad list property", path.c_str());
      } else {
        p.type = PlyType(t);
        ls >> p.name;
        if (!p.type) return Fail(HD_EINVAL, "%s: bad property type %s", path.c_str(), t.c_str());
      }
      elems.back().props.push_back(p);
    } else if (w == "end_header") {
      break;
    }
  }
  if (format < 0) return Fail(HD_EINVAL, "%s: no format line", path.c_str());
  std::vector<double> v;  // x, y, z of the read vertices
  bool found = false;
  for (const PlyElement& e : elems) {
    const bool vertex = e.name == "vertex";
    int axis_of[64];
    if (e.props.size() > 64) return Fail(HD_EINVAL, "%s: too many properties", path.c_str());
    int have = 0;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const std::string& n = e.props[k].name;
      axis_of[k] = !vertex || e.props[k].count_type ? -1
                   : n == "x" ? 0 : n == "y" ? 1 : n == "z" ? 2 : -1;
      if (axis_of[k] >= nd) axis_of[k] = -1;
      if (axis_of[k] >= 0) ++have;
    }
    if (vertex && have < nd)
      return Fail(HD_EINVAL, "%s: vertex lacks x/y%s", path.c_str(), nd == 3 ? "/z" : "");
    const uint64_t rows = vertex ? std::min(e.count, limit) : e.count;
    if (vertex) v.assign(rows * 3, 0.0);
    for (uint64_t r = 0; r < rows; ++r) {
      for (size_t k = 0; k < e.props.size(); ++k) {
        const PlyProp& p = e.props[k];
        uint64_t items = 1;
        double x = 0;
        if (p.count_type) {
          if (format == 0) {
            if (!(in >> x)) return Fail(HD_EINVAL, "%s: truncated", path.c_str());
          } else if (!PlyBinary(in, p.count_type, format == 2, &x)) {
            return Fail(HD_EINVAL, "%s: truncated", path.c_str());
          }
          items = (uint64_t)x;
        }
        for (uint64_t t = 0; t < items; ++t) {
          if (format == 0) {
            if (!(in >> x)) return Fail(HD_EINVAL, "%s: truncated", path.c_str());
          } else if (!PlyBinary(in, p.type, format == 2, &x)) {
            return Fail(HD_EINVAL, "%s: truncated", path.c_str());
          }
          if (axis_of[k] >= 0) v[r * 3 + axis_of[k]] = x;
        }
      }
    }
    if (vertex) {
      found = true;
      break;  // elements after the vertices are not needed
    }
  }
  if (!found) return Fail(HD_EINVAL, "%s: no vertex element", path.c_str());
  const uint64_t n = v.size() / 3;
  double lo[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) {
    double mn = n ? v[d] : 0.0, mx = mn;
    for (uint64_t i = 0; i < n; ++i) {
      mn = std::min(mn, v[i * 3 + d]);
      mx = std::max(mx, v[i * 3 + d]);
    }
    lo[d] = mn;
  }
  out->resize(n * nd);
  for (uint64_t i = 0; i < n; ++i)
    for (int d = 0; d < nd; ++d) (*out)[i * nd + d] = (float)(v[i * 3 + d] - lo[d]);
  return HD_OK;
}

std::string CachePath(const std::string& folder, const std::string& path,
                      uint64_t limit) {
  std::string esc = path;
  std::replace(esc.begin(), esc.end(), '/', '_');
  return folder + "/points_" + esc + "_limit_" + std::to_string(limit) + ".bin";
}

}  // namespace
}  // namespace hd

using namespace hd;

extern "C" {

int hd_load_points(const char* path, int input_type, int n_dims,
                   uint64_t limit, const char* serialize_folder, float** out,
                   uint64_t* n_out) {
  ClearError();
  if (!path || !out || !n_out) return Fail(HD_EINVAL, "null argument");
  *out = nullptr;
  *n_out = 0;
  if (n_dims != 2 && n_dims != 3) return Fail(HD_EINVAL, "n_dims must be 2 or 3");
  if (limit == 0) return Fail(HD_EINVAL, "limit must be positive");
  std::vector<float> pts;
  const std::string p(path);
  const bool cacheable = serialize_folder && serialize_folder[0] &&
                         (input_type == HD_INPUT_WKT || input_type == HD_INPUT_OFF);
  std::string cache;
  int rc = HD_OK;
  if (cacheable) {
    const std::string folder(serialize_folder);
    if (mkdir(folder.c_str(), 0755) != 0 && errno != EEXIST)
      return Fail(HD_EINVAL, "cannot create %s", folder.c_str());
    cache = CachePath(folder, p, limit);
  }
  if (cacheable && access(cache.c_str(), R_OK) == 0) {
    rc = ReadBin(cache, n_dims, limit, &pts);
  } else {
    switch (input_type) {
      case HD_INPUT_WKT: rc = ReadWkt(p, n_dims, limit, &pts); break;
      case HD_INPUT_OFF: rc = ReadOff(p, n_dims, limit, &pts); break;
      case HD_INPUT_PLY: rc = ReadPly(p, n_dims, limit, &pts); break;
      case HD_INPUT_BIN: rc = ReadBin(p, n_dims, limit, &pts); break;
      default: return Fail(HD_EINVAL, "unknown input type %d", input_type);
    }
    if (rc == HD_OK && cacheable && !pts.empty())
      rc = WriteBin(cache, n_dims, pts);
  }
  if (rc != HD_OK) return rc;
  if (pts.empty()) return Fail(HD_EINVAL, "%s: no points", path);
  float* buf = static_cast<float*>(std::malloc(sizeof(float) * pts.size()));
  if (!buf) return Fail(HD_ENOMEM, "out of host memory");
  std::memcpy(buf, pts.data(), sizeof(float) * pts.size());
  *out = buf;
  *n_out = pts.size() / n_dims;
  return HD_OK;
}

void hd_free_points(float* points) { std::free(points); }

int hd_normalize_points(float* p, uint64_t n, int nd) {
  ClearError();
  if (!p && n) return Fail(HD_EINVAL, "null argument");
  if (nd != 2 && nd != 3) return Fail(HD_EINVAL, "n_dims must be 2 or 3");
  if (n == 0) return HD_OK;
  float lo[3], hi[3];
  for (int d = 0; d < nd; ++d) lo[d] = hi[d] = p[d];
  for (uint64_t i = 0; i < n; ++i)
    for (int d = 0; d < nd; ++d) {
      lo[d] = std::min(lo[d], p[i * nd + d]);
      hi[d] = std::max(hi[d], p[i * nd + d]);
    }
  float ext = 0;
  for (int d = 0; d < nd; ++d) ext = std::max(ext, hi[d] - lo[d]);
  if (ext == 0) ext = 1;
  for (uint64_t i = 0; i < n; ++i)
    for (int d = 0; d < nd; ++d) p[i * nd + d] = (p[i * nd + d] - lo[d]) / ext;
  return HD_OK;
}

int hd_translate_points(float* p, uint64_t n, int nd, int dim, double ratio) {
  ClearError();
  if (!p && n) return Fail(HD_EINVAL, "null argument");
  if (nd != 2 && nd != 3) return Fail(HD_EINVAL, "n_dims must be 2 or 3");
  if (dim < 0 || dim >= nd) return Fail(HD_EINVAL, "dim out of range");
  if (n == 0) return HD_OK;
  float lo = p[dim], hi = p[dim];
  for (uint64_t i = 0; i < n; ++i) {
    lo = std::min(lo, p[i * nd + dim]);
    hi = std::max(hi, p[i * nd + dim]);
  }
  // FP32 extent times an FP64 ratio, added in FP64 and rounded once (the
  // reference's `float += double`).
  const double shift = (double)(hi - lo) * ratio;
  for (uint64_t i = 0; i < n; ++i)
    p[i * nd + dim] = (float)((double)p[i * nd + dim] + shift);
  return HD_OK;
}

}  // extern "C"
