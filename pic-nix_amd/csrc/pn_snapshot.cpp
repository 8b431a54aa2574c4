// pn_snapshot.cpp -- the reference's snapshot and profile files.
//
// FieldDiag (pic/diag/field.hpp:66-185) and ParticleDiag (pic/diag/particle.hpp:
// 56-130) as one process writes them in the default "mpiio" I/O mode: a raw
// data file of native-endian f8 values, chunk after chunk in chunk-id order,
// and a JSON file with the nixio metadata of each dataset {datatype,
// description, offset, size, ndim, shape} (nix/nixio.cpp:351-364) and the meta
// record {endian, rawfile, layout, decimate, time, step, chunk_id_range}.  The
// JSON layout is nlohmann::json's pretty print with indent 2 (keys sorted).
// The application profile (nix/application.cpp:11-26, saved as msgpack by
// nix/statehandler.hpp:44-63) holds the configuration and the chunk map that
// the reference's post-processing uses to reassemble chunked arrays
// (script/picnix/utils.py:153-181).
//
// Host-side code: the arrays come through the library's own getters.
#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/picnix.h"
#include "pn_sim.hpp"

namespace {

using picnix::Geom;
using picnix::P32;
using picnix::PArr;
using picnix::Sim;

// ---------------------------------------------------------------------------
// a JSON value (the subset the files use) with nlohmann's pretty print and
// msgpack encodings
// ---------------------------------------------------------------------------

struct Json {
  enum Kind { Null, Bool, Int, Float, Str, Arr, Obj };
  Kind                        kind = Null;
  bool                        b    = false;
  int64_t                     i    = 0;
  double                      f    = 0;
  std::string                 s;
  std::vector<Json>           a;
  std::map<std::string, Json> o;

  static Json boolean(bool v)
  {
    Json j;
    j.kind = Bool;
    j.b    = v;
    return j;
  }
  static Json integer(int64_t v)
  {
    Json j;
    j.kind = Int;
    j.i    = v;
    return j;
  }
  static Json real(double v)
  {
    Json j;
    j.kind = Float;
    j.f    = v;
    return j;
  }
  static Json string(std::string v)
  {
    Json j;
    j.kind = Str;
    j.s    = std::move(v);
    return j;
  }
  static Json array()
  {
    Json j;
    j.kind = Arr;
    return j;
  }
  template <typename T>
  static Json ints(const std::vector<T>& v)
  {
    Json j = array();
    for (T x : v)
      j.a.push_back(integer((int64_t)x));
    return j;
  }
  Json& operator[](const std::string& k)
  {
    kind = Obj;
    return o[k];
  }
  void push(Json v)
  {
    kind = Arr;
    a.push_back(std::move(v));
  }
};

void dump_string(const std::string& v, std::string& out)
{
  out += '"';
  for (unsigned char c : v) {
    switch (c) {
    case '"': out += "\\\""; break;
    case '\\': out += "\\\\"; break;
    case '\n': out += "\\n"; break;
    case '\t': out += "\\t"; break;
    case '\r': out += "\\r"; break;
    case '\b': out += "\\b"; break;
    case '\f': out += "\\f"; break;
    default:
      if (c < 0x20) {
        char buf[8];
        std::snprintf(buf, sizeof(buf), "\\u%04x", c);
        out += buf;
      } else {
        out += (char)c;
      }
    }
  }
  out += '"';
}

// shortest round-trip digits; integral-looking values get ".0" (as nlohmann)
void dump_real(double v, std::string& out)
{
  if (!std::isfinite(v)) {
    out += "null";
    return;
  }
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof(buf), v);
  std::string t(buf, r.ptr);
  if (t.find_first_of(".eE") == std::string::npos)
    t += ".0";
  out += t;
}

void dump(const Json& j, std::string& out, int indent, int level)
{
  const std::string pad((size_t)indent * (level + 1), ' ');
  const std::string end((size_t)indent * level, ' ');
  switch (j.kind) {
  case Json::Null: out += "null"; break;
  case Json::Bool: out += j.b ? "true" : "false"; break;
  case Json::Int: out += std::to_string(j.i); break;
  case Json::Float: dump_real(j.f, out); break;
  case Json::Str: dump_string(j.s, out); break;
  case Json::Arr:
    if (j.a.empty()) {
      out += "[]";
      break;
    }
    out += "[\n";
    for (size_t k = 0; k < j.a.size(); k++) {
      out += pad;
      dump(j.a[k], out, indent, level + 1);
      out += k + 1 < j.a.size() ? ",\n" : "\n";
    }
    out += end + "]";
    break;
  case Json::Obj:
    if (j.o.empty()) {
      out += "{}";
      break;
    }
    out += "{\n";
    for (auto it = j.o.begin(); it != j.o.end(); ++it) {
      out += pad;
      dump_string(it->first, out);
      out += ": ";
      dump(it->second, out, indent, level + 1);
      out += std::next(it) != j.o.end() ? ",\n" : "\n";
    }
    out += end + "}";
    break;
  }
}

void put_be(std::vector<uint8_t>& out, uint64_t v, int nbytes)
{
  for (int k = nbytes - 1; k >= 0; k--)
    out.push_back((uint8_t)(v >> (8 * k)));
}

// msgpack with the smallest integer, string, array and map forms
void msgpack(const Json& j, std::vector<uint8_t>& out)
{
  switch (j.kind) {
  case Json::Null: out.push_back(0xc0); break;
  case Json::Bool: out.push_back(j.b ? 0xc3 : 0xc2); break;
  case Json::Int: {
    const int64_t v = j.i;
    if (v >= 0) {
      if (v < 128)
        out.push_back((uint8_t)v);
      else if (v <= 0xff)
        out.push_back(0xcc), put_be(out, (uint64_t)v, 1);
      else if (v <= 0xffff)
        out.push_back(0xcd), put_be(out, (uint64_t)v, 2);
      else if (v <= 0xffffffffll)
        out.push_back(0xce), put_be(out, (uint64_t)v, 4);
      else
        out.push_back(0xcf), put_be(out, (uint64_t)v, 8);
    } else if (v >= -32) {
      out.push_back((uint8_t)(int8_t)v);
    } else if (v >= -128) {
      out.push_back(0xd0), put_be(out, (uint64_t)v, 1);
    } else if (v >= -32768) {
      out.push_back(0xd1), put_be(out, (uint64_t)v, 2);
    } else if (v >= -2147483648ll) {
      out.push_back(0xd2), put_be(out, (uint64_t)v, 4);
    } else {
      out.push_back(0xd3), put_be(out, (uint64_t)v, 8);
    }
    break;
  }
  case Json::Float: {
    uint64_t bits;
    std::memcpy(&bits, &j.f, 8);
    out.push_back(0xcb);
    put_be(out, bits, 8);
    break;
  }
  case Json::Str: {
    const size_t n = j.s.size();
    if (n < 32)
      out.push_back((uint8_t)(0xa0 | n));
    else if (n <= 0xff)
      out.push_back(0xd9), put_be(out, n, 1);
    else if (n <= 0xffff)
      out.push_back(0xda), put_be(out, n, 2);
    else
      out.push_back(0xdb), put_be(out, n, 4);
    out.insert(out.end(), j.s.begin(), j.s.end());
    break;
  }
  case Json::Arr: {
    const size_t n = j.a.size();
    if (n < 16)
      out.push_back((uint8_t)(0x90 | n));
    else if (n <= 0xffff)
      out.push_back(0xdc), put_be(out, n, 2);
    else
      out.push_back(0xdd), put_be(out, n, 4);
    for (const Json& e : j.a)
      msgpack(e, out);
    break;
  }
  case Json::Obj: {
    const size_t n = j.o.size();
    if (n < 16)
      out.push_back((uint8_t)(0x80 | n));
    else if (n <= 0xffff)
      out.push_back(0xde), put_be(out, n, 2);
    else
      out.push_back(0xdf), put_be(out, n, 4);
    for (const auto& kv : j.o) {
      msgpack(Json::string(kv.first), out);
      msgpack(kv.second, out);
    }
    break;
  }
  }
}

// ---------------------------------------------------------------------------
// config.toml -> JSON with types (the profile's "configuration"): tables,
// arrays of tables, dotted keys, strings, booleans, integers, floats and
// one-line arrays
// ---------------------------------------------------------------------------

std::string trim(const std::string& t)
{
  size_t a = t.find_first_not_of(" \t\r");
  if (a == std::string::npos)
    return "";
  size_t b = t.find_last_not_of(" \t\r");
  return t.substr(a, b - a + 1);
}

std::string strip_comment(const std::string& line)
{
  char quote = 0;
  for (size_t k = 0; k < line.size(); k++) {
    const char ch = line[k];
    if (quote) {
      if (ch == quote)
        quote = 0;
    } else if (ch == '"' || ch == '\'') {
      quote = ch;
    } else if (ch == '#') {
      return trim(line.substr(0, k));
    }
  }
  return trim(line);
}

std::vector<std::string> split_top(const std::string& t, char sep)
{
  std::vector<std::string> parts;
  int                      depth = 0;
  char                     quote = 0;
  std::string              cur;
  for (char ch : t) {
    if (quote) {
      if (ch == quote)
        quote = 0;
      cur += ch;
      continue;
    }
    if (ch == '"' || ch == '\'')
      quote = ch;
    else if (ch == '[' || ch == '{')
      depth++;
    else if (ch == ']' || ch == '}')
      depth--;
    if (ch == sep && depth == 0) {
      parts.push_back(trim(cur));
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!trim(cur).empty())
    parts.push_back(trim(cur));
  return parts;
}

Json toml_value(const std::string& v)
{
  if (v.size() >= 2 && (v[0] == '"' || v[0] == '\''))
    return Json::string(v.substr(1, v.size() - 2));
  if (v == "true" || v == "false")
    return Json::boolean(v == "true");
  if (!v.empty() && v[0] == '[') {
    Json a = Json::array();
    for (const std::string& e : split_top(v.substr(1, v.size() - 2), ','))
      a.push(toml_value(e));
    return a;
  }
  std::string n;
  for (char ch : v)
    if (ch != '_')
      n += ch;
  if (n.find_first_of(".eE") == std::string::npos && n.find("inf") == std::string::npos &&
      n.find("nan") == std::string::npos) {
    char*         endp = nullptr;
    const int64_t iv   = std::strtoll(n.c_str(), &endp, 0);
    if (endp && *endp == '\0')
      return Json::integer(iv);
  }
  return Json::real(std::strtod(n.c_str(), nullptr));
}

Json* walk(Json& root, const std::vector<std::string>& path)
{
  Json* cur = &root;
  for (const std::string& k : path) {
    Json& nx = (*cur)[k];
    cur      = nx.kind == Json::Arr && !nx.a.empty() ? &nx.a.back() : &nx;
  }
  return cur;
}

std::vector<std::string> dotted(const std::string& key)
{
  std::vector<std::string> out;
  for (const std::string& p : split_top(key, '.'))
    out.push_back(p.size() >= 2 && (p[0] == '"' || p[0] == '\'') ? p.substr(1, p.size() - 2) : p);
  return out;
}

Json toml_to_json(const std::string& text)
{
  Json               root;
  root.kind = Json::Obj;
  Json*              table = &root;
  std::istringstream in(text);
  std::string        line;
  while (std::getline(in, line)) {
    line = strip_comment(line);
    if (line.empty())
      continue;
    if (line.rfind("[[", 0) == 0) {
      auto  path = dotted(trim(line.substr(2, line.find("]]") - 2)));
      auto  last = path.back();
      path.pop_back();
      Json* parent = walk(root, path);
      Json& arr    = (*parent)[last];
      arr.kind     = Json::Arr;
      Json t;
      t.kind = Json::Obj;
      arr.a.push_back(t);
      table = &arr.a.back();
    } else if (line[0] == '[') {
      table       = walk(root, dotted(trim(line.substr(1, line.find(']') - 1))));
      table->kind = Json::Obj;
    } else {
      const size_t eq = line.find('=');
      if (eq == std::string::npos)
        continue;
      auto  path = dotted(trim(line.substr(0, eq)));
      auto  last = path.back();
      path.pop_back();
      Json* t    = walk(*table, path);
      (*t)[last] = toml_value(trim(line.substr(eq + 1)));
    }
  }
  return root;
}

// ---------------------------------------------------------------------------
// files
// ---------------------------------------------------------------------------

std::string step_name(int step)
{
  char buf[32];
  std::snprintf(buf, sizeof(buf), "%08d", step); // nix::format_step (nix/nix.hpp:217-220)
  return buf;
}

int32_t endian_flag()
{
  const uint8_t b[4] = {1, 0, 0, 0};
  int32_t       f;
  std::memcpy(&f, b, 4);
  return f;
}

void put_metadata(Json& dataset, const std::string& name, const std::string& desc, size_t disp,
                  size_t size, const std::vector<int64_t>& shape)
{
  Json& d          = dataset[name];
  d["datatype"]    = Json::string("f8");
  d["description"] = Json::string(desc);
  d["offset"]      = Json::integer((int64_t)disp);
  d["size"]        = Json::integer((int64_t)size);
  d["ndim"]        = Json::integer((int64_t)shape.size());
  d["shape"]       = Json::ints(shape);
}

bool write_json(const std::string& path, const Json& root)
{
  std::string out;
  dump(root, out, 2, 0);
  std::ofstream ofs(path);
  ofs << out;
  return (bool)ofs;
}

int fail(Sim& s, int code, const std::string& msg)
{
  s.err = msg;
  return code;
}

// the interior extent of a chunk and the ghost-inclusive one
struct ChunkShape {
  int n[3]; // interior (z, y, x)
  int e[3]; // ghost-inclusive (z, y, x)
  int nb;
};

ChunkShape chunk_shape(const Sim& s)
{
  ChunkShape c{};
  for (int d = 0; d < 3; d++) {
    c.n[d] = s.g.dims[d];
    c.e[d] = s.g.dims[d] + 2 * s.g.Nb;
  }
  c.nb = s.g.Nb;
  return c;
}

// XtensorPacker3D::decimate_size and FieldDiag::calc_decimated_size
// (nix/xtensor_packer3d.hpp:194-203, pic/diag/field.hpp:124-136)
int packed_size(int size, int decimate) { return size <= decimate ? 1 : size / decimate; }
int meta_size(int size, int decimate)
{
  if (size <= decimate)
    return 1;
  if (decimate <= 0 || size % decimate != 0)
    return size;
  return size / decimate;
}

// Colocated E and B at cell centres (XtensorPacker3D::colocate_field_{1,2,3}d,
// nix/xtensor_packer3d.hpp:206-302); x is the ghost-inclusive [z][y][x][6]
// chunk array, y receives the interior values.
void colocate(const ChunkShape& c, const double* x, std::vector<double>& y)
{
  const int  ez = c.e[0], ey = c.e[1], ex = c.e[2], nb = c.nb;
  const bool d3 = c.n[0] > 1, d2 = c.n[1] > 1;
  auto       X  = [&](int iz, int iy, int ix, int k) {
    return x[(((size_t)iz * ey + iy) * ex + ix) * 6 + k];
  };
  y.assign((size_t)ez * ey * ex * 6, 0.0);
  auto Y = [&](int iz, int iy, int ix, int k) -> double& {
    return y[(((size_t)iz * ey + iy) * ex + ix) * 6 + k];
  };
  for (int iz = nb; iz < nb + c.n[0]; iz++)
    for (int iy = nb; iy < nb + c.n[1]; iy++)
      for (int ix = nb; ix < nb + c.n[2]; ix++) {
        if (d3) {
          Y(iz, iy, ix, 0) = 0.50 * (X(iz, iy, ix, 0) + X(iz, iy, ix + 1, 0));
          Y(iz, iy, ix, 1) = 0.50 * (X(iz, iy, ix, 1) + X(iz, iy + 1, ix, 1));
          Y(iz, iy, ix, 2) = 0.50 * (X(iz, iy, ix, 2) + X(iz + 1, iy, ix, 2));
          Y(iz, iy, ix, 3) = 0.25 * (X(iz, iy, ix, 3) + X(iz + 1, iy + 1, ix, 3) +
                                     X(iz, iy + 1, ix, 3) + X(iz + 1, iy, ix, 3));
          Y(iz, iy, ix, 4) = 0.25 * (X(iz, iy, ix, 4) + X(iz + 1, iy, ix + 1, 4) +
                                     X(iz + 1, iy, ix, 4) + X(iz, iy, ix + 1, 4));
          Y(iz, iy, ix, 5) = 0.25 * (X(iz, iy, ix, 5) + X(iz, iy + 1, ix + 1, 5) +
                                     X(iz, iy, ix + 1, 5) + X(iz, iy + 1, ix, 5));
        } else if (d2) {
          Y(iz, iy, ix, 0) = 0.50 * (X(iz, iy, ix, 0) + X(iz, iy, ix + 1, 0));
          Y(iz, iy, ix, 1) = 0.50 * (X(iz, iy, ix, 1) + X(iz, iy + 1, ix, 1));
          Y(iz, iy, ix, 2) = X(iz, iy, ix, 2);
          Y(iz, iy, ix, 3) = 0.50 * (X(iz, iy, ix, 3) + X(iz, iy + 1, ix, 3));
          Y(iz, iy, ix, 4) = 0.50 * (X(iz, iy, ix, 4) + X(iz, iy, ix + 1, 4));
          Y(iz, iy, ix, 5) = 0.25 * (X(iz, iy, ix, 5) + X(iz, iy + 1, ix + 1, 5) +
                                     X(iz, iy, ix + 1, 5) + X(iz, iy + 1, ix, 5));
        } else {
          Y(iz, iy, ix, 0) = 0.50 * (X(iz, iy, ix, 0) + X(iz, iy, ix + 1, 0));
          Y(iz, iy, ix, 1) = X(iz, iy, ix, 1);
          Y(iz, iy, ix, 2) = X(iz, iy, ix, 2);
          Y(iz, iy, ix, 3) = X(iz, iy, ix, 3);
          Y(iz, iy, ix, 4) = 0.50 * (X(iz, iy, ix, 4) + X(iz, iy, ix + 1, 4));
          Y(iz, iy, ix, 5) = 0.50 * (X(iz, iy, ix, 5) + X(iz, iy, ix + 1, 5));
        }
      }
}

// Block averages of the interior (XtensorPacker3D::decimate_field,
// nix/xtensor_packer3d.hpp:151-192: each output value sums factor * input over
// the block in (z, y, x) order); decimate < 0: the raw interior
// (pack_array_raw, :37-61).  x is ghost-inclusive with nc components.
void pack_interior(const ChunkShape& c, const double* x, int nc, int decimate,
                   std::vector<double>& out)
{
  const int ey = c.e[1], ex = c.e[2], nb = c.nb;
  auto      at = [&](int iz, int iy, int ix, int k) {
    return x[(((size_t)iz * ey + iy) * ex + ix) * nc + k];
  };
  if (decimate < 0) {
    for (int iz = nb; iz < nb + c.n[0]; iz++)
      for (int iy = nb; iy < nb + c.n[1]; iy++)
        for (int ix = nb; ix < nb + c.n[2]; ix++)
          for (int k = 0; k < nc; k++)
            out.push_back(at(iz, iy, ix, k));
    return;
  }
  int sz[3], bs[3];
  for (int d = 0; d < 3; d++) {
    sz[d] = packed_size(c.n[d], decimate);
    bs[d] = c.n[d] / sz[d];
  }
  const double        factor = 1.0 / (bs[0] * bs[1] * bs[2]);
  std::vector<double> z((size_t)sz[0] * sz[1] * sz[2] * nc, 0.0);
  for (int bz = 0; bz < bs[0]; bz++)
    for (int by = 0; by < bs[1]; by++)
      for (int bx = 0; bx < bs[2]; bx++)
        for (int jz = 0; jz < sz[0]; jz++)
          for (int jy = 0; jy < sz[1]; jy++)
            for (int jx = 0; jx < sz[2]; jx++)
              for (int k = 0; k < nc; k++)
                z[(((size_t)jz * sz[1] + jy) * sz[2] + jx) * nc + k] +=
                    factor * at(nb + jz * bs[0] + bz, nb + jy * bs[1] + by, nb + jx * bs[2] + bx, k);
  out.insert(out.end(), z.begin(), z.end());
}

// ---------------------------------------------------------------------------
// checkpoint: the device state as raw blocks (our own layout; the content of
// PicChunk::pack, pic/pic_chunk.cpp:59-79: fields, current, Friedman history,
// particles per species) plus the application record of StateHandler
// (nix/statehandler.hpp:36-63, nix/application.cpp:11-48)
// ---------------------------------------------------------------------------

struct Block {
  std::string name;
  const void* dev; // device pointer (save) or destination (load)
  size_t      bytes;
};

std::vector<Block> state_blocks(Sim& s, const std::vector<int32_t>& np)
{
  const Geom&        g   = s.g;
  const size_t       vol = (size_t)g.vol;
  std::vector<Block> b;
  if (s.fp32) {
    b.push_back({"E", s.f_E, 3 * vol * sizeof(float)});
    b.push_back({"B", s.f_B, 3 * vol * sizeof(float)});
    b.push_back({"J", s.f_J, 4 * vol * sizeof(float)});
    if (s.f_ff)
      b.push_back({"ff", s.f_ff, 9 * vol * sizeof(float)});
  } else {
    b.push_back({"E", s.d_E, 3 * vol * sizeof(double)});
    b.push_back({"B", s.d_B, 3 * vol * sizeof(double)});
    b.push_back({"J", s.d_J, 4 * vol * sizeof(double)});
    if (s.d_ff)
      b.push_back({"ff", s.d_ff, 9 * vol * sizeof(double)});
  }
  b.push_back({"cellstart", s.ss.cellstart, (size_t)g.nseg * (g.nbin + 1) * sizeof(int32_t)});
  for (int seg = 0; seg < g.nseg; seg++) {
    const size_t off = (size_t)seg * g.cap, n = (size_t)np[seg];
    const std::string t = "p" + std::to_string(seg) + ".";
    if (s.fp32) {
      const P32& q = s.qa;
      b.push_back({t + "rx", q.rx + off, n * sizeof(float)});
      b.push_back({t + "ry", q.ry + off, n * sizeof(float)});
      b.push_back({t + "rz", q.rz + off, n * sizeof(float)});
      b.push_back({t + "ux", q.ux + off, n * sizeof(float)});
      b.push_back({t + "uy", q.uy + off, n * sizeof(float)});
      b.push_back({t + "uz", q.uz + off, n * sizeof(float)});
      b.push_back({t + "id", q.id + off, n * sizeof(int64_t)});
      b.push_back({t + "cell", q.cell + off, n * sizeof(int32_t)});
    } else {
      const PArr& q = s.pa;
      b.push_back({t + "x", q.x + off, n * sizeof(double)});
      b.push_back({t + "y", q.y + off, n * sizeof(double)});
      b.push_back({t + "z", q.z + off, n * sizeof(double)});
      b.push_back({t + "ux", q.ux + off, n * sizeof(double)});
      b.push_back({t + "uy", q.uy + off, n * sizeof(double)});
      b.push_back({t + "uz", q.uz + off, n * sizeof(double)});
      b.push_back({t + "id", q.id + off, n * sizeof(int64_t)});
    }
  }
  return b;
}

// the application record (Application::to_json) with the library's state flags
Json app_state(Sim& s, int step, double time)
{
  const auto&  cfg = s.cfg;
  const double now =
      std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch()).count();
  Json state;
  state["timestamp"]     = Json::real(now);
  state["wclock"]        = Json::real(now);
  state["ndims"]         = Json::ints(std::vector<int>{cfg.N[0], cfg.N[1], cfg.N[2]});
  state["cdims"]         = Json::ints(std::vector<int>{cfg.C[0], cfg.C[1], cfg.C[2]});
  state["curstep"]       = Json::integer(step);
  state["curtime"]       = Json::real(time);
  state["thisrank"]      = Json::integer(0);
  state["nprocess"]      = Json::integer(1);
  state["nthread"]       = Json::integer(1);
  state["configuration"] = toml_to_json(s.toml);
  // ChunkMap::to_json (nix/chunkmap.cpp:45-61): chunkid[z][y][x], coord[id] = (x, y, z)
  Json& cm          = state["chunkmap"];
  cm["size"]        = Json::integer(s.g.nchunk);
  cm["ndim"]        = Json::integer(3);
  cm["shape"]       = Json::ints(std::vector<int>{cfg.C[0], cfg.C[1], cfg.C[2]});
  cm["periodicity"] = Json::ints(std::vector<int>{1, 1, s.g.xwall ? 0 : 1});
  Json ids          = Json::array();
  for (int cz = 0; cz < cfg.C[0]; cz++) {
    Json plane = Json::array();
    for (int cy = 0; cy < cfg.C[1]; cy++) {
      std::vector<int> row(cfg.C[2]);
      for (int cx = 0; cx < cfg.C[2]; cx++)
        row[cx] = s.id_of_coord[((size_t)cz * cfg.C[1] + cy) * cfg.C[2] + cx];
      plane.push(Json::ints(row));
    }
    ids.push(plane);
  }
  cm["chunkid"] = ids;
  Json coord    = Json::array();
  for (int id = 0; id < s.g.nchunk; id++)
    coord.push(Json::ints(
        std::vector<int>{s.coord[id * 3 + 2], s.coord[id * 3 + 1], s.coord[id * 3 + 0]}));
  cm["coord"]    = coord;
  cm["boundary"] = Json::ints(std::vector<int>{0, s.g.nchunk});
  return state;
}

bool write_msgpack(const std::string& path, const Json& j)
{
  std::vector<uint8_t> bytes;
  msgpack(j, bytes);
  std::error_code             ec;
  const std::filesystem::path p(path);
  if (p.has_parent_path())
    std::filesystem::create_directories(p.parent_path(), ec);
  std::ofstream ofs(p, std::ios::binary);
  ofs.write(reinterpret_cast<const char*>(bytes.data()), (std::streamsize)bytes.size());
  return (bool)ofs;
}

// a msgpack reader for what write_msgpack writes
struct Reader {
  const std::vector<uint8_t>& b;
  size_t                      p = 0;
  uint64_t                    be(int n)
  {
    uint64_t v = 0;
    for (int k = 0; k < n; k++)
      v = (v << 8) | b.at(p++);
    return v;
  }
  Json read()
  {
    const uint8_t t = b.at(p++);
    if (t <= 0x7f)
      return Json::integer(t);
    if (t >= 0xe0)
      return Json::integer((int8_t)t);
    if ((t & 0xe0) == 0xa0)
      return str(t & 0x1f);
    if ((t & 0xf0) == 0x90)
      return arr(t & 0x0f);
    if ((t & 0xf0) == 0x80)
      return obj(t & 0x0f);
    switch (t) {
    case 0xc0: return Json();
    case 0xc2: return Json::boolean(false);
    case 0xc3: return Json::boolean(true);
    case 0xcc: return Json::integer((int64_t)be(1));
    case 0xcd: return Json::integer((int64_t)be(2));
    case 0xce: return Json::integer((int64_t)be(4));
    case 0xcf: return Json::integer((int64_t)be(8));
    case 0xd0: return Json::integer((int8_t)be(1));
    case 0xd1: return Json::integer((int16_t)be(2));
    case 0xd2: return Json::integer((int32_t)be(4));
    case 0xd3: return Json::integer((int64_t)be(8));
    case 0xcb: {
      const uint64_t bits = be(8);
      double         f;
      std::memcpy(&f, &bits, 8);
      return Json::real(f);
    }
    case 0xd9: return str((size_t)be(1));
    case 0xda: return str((size_t)be(2));
    case 0xdb: return str((size_t)be(4));
    case 0xdc: return arr((size_t)be(2));
    case 0xdd: return arr((size_t)be(4));
    case 0xde: return obj((size_t)be(2));
    case 0xdf: return obj((size_t)be(4));
    default: throw std::runtime_error("unsupported msgpack type in state file");
    }
  }
  Json str(size_t n)
  {
    if (p + n > b.size())
      throw std::runtime_error("truncated state file");
    Json j = Json::string(std::string(b.begin() + p, b.begin() + p + n));
    p += n;
    return j;
  }
  Json arr(size_t n)
  {
    Json j = Json::array();
    for (size_t k = 0; k < n; k++)
      j.push(read());
    return j;
  }
  Json obj(size_t n)
  {
    Json j;
    j.kind = Json::Obj;
    for (size_t k = 0; k < n; k++) {
      Json key = read();
      j[key.s] = read();
    }
    return j;
  }
};

bool same_json(const Json& a, const Json& b)
{
  if (a.kind != b.kind)
    return false;
  switch (a.kind) {
  case Json::Null: return true;
  case Json::Bool: return a.b == b.b;
  case Json::Int: return a.i == b.i;
  case Json::Float: return a.f == b.f;
  case Json::Str: return a.s == b.s;
  case Json::Arr:
    if (a.a.size() != b.a.size())
      return false;
    for (size_t k = 0; k < a.a.size(); k++)
      if (!same_json(a.a[k], b.a[k]))
        return false;
    return true;
  case Json::Obj:
    if (a.o.size() != b.o.size())
      return false;
    for (const auto& kv : a.o) {
      auto it = b.o.find(kv.first);
      if (it == b.o.end() || !same_json(kv.second, it->second))
        return false;
    }
    return true;
  }
  return false;
}

} // namespace

extern "C" {

int pn_write_field_diag(pn_sim* h, const char* dir, int step, double time, int decimate)
{
  Sim* sp = picnix::sim_state(h);
  if (!sp || !dir)
    return PN_ERR_ARG;
  Sim& s = *sp;
  if (decimate == 0)
    return fail(s, PN_ERR_ARG, "decimate must be >= 1 (colocated) or < 0 (raw)");
  if (s.fp32)
    return fail(s, PN_ERR_UNSUPPORTED, "the field diagnostic writes moments, FP64 only");
  const ChunkShape c  = chunk_shape(s);
  const int        nk = s.g.nchunk, ns = s.g.Ns;
  const size_t     ng = (size_t)c.e[0] * c.e[1] * c.e[2];
  std::vector<double> data, buf, col;
  Json                dataset;
  dataset.kind = Json::Obj;
  int nz = meta_size(c.n[0], decimate), ny = meta_size(c.n[1], decimate),
      nx = meta_size(c.n[2], decimate);
  // electromagnetic field (pic/diag/field.hpp:144-160)
  buf.resize(ng * 6);
  for (int k = 0; k < nk; k++) {
    int rc = pn_get_fields(h, k, PN_FIELD_UF, buf.data());
    if (rc)
      return rc;
    if (decimate < 0) {
      pack_interior(c, buf.data(), 6, decimate, data);
    } else {
      colocate(c, buf.data(), col);
      pack_interior(c, col.data(), 6, decimate, data);
    }
  }
  const size_t uf_bytes = data.size() * sizeof(double);
  {
    const size_t size = (size_t)nz * ny * nx * 6 * sizeof(double);
    put_metadata(dataset, "uf", "electromagnetic field", 0, uf_bytes,
                 {(int64_t)(uf_bytes / size), nz, ny, nx, 6});
  }
  // moment (pic/diag/field.hpp:162-180): deposit + Mom exchange first
  int rc = pn_deposit_moment(h, -1);
  if (!rc)
    rc = pn_exchange(h, PN_BOUNDARY_MOM);
  if (rc)
    return rc;
  buf.resize(ng * 14 * ns);
  for (int k = 0; k < nk; k++) {
    if ((rc = pn_get_fields(h, k, PN_FIELD_UM, buf.data())))
      return rc;
    pack_interior(c, buf.data(), 14 * ns, decimate, data);
  }
  {
    const size_t um_bytes = data.size() * sizeof(double) - uf_bytes;
    const size_t size     = (size_t)nz * ny * nx * ns * 14 * sizeof(double);
    put_metadata(dataset, "um", "moment", uf_bytes, um_bytes,
                 {(int64_t)(um_bytes / size), nz, ny, nx, ns, 14});
  }
  namespace fs = std::filesystem;
  std::error_code ec;
  fs::create_directories(dir, ec);
  const std::string name = step_name(step);
  {
    std::ofstream ofs(fs::path(dir) / (name + ".data"), std::ios::binary);
    ofs.write(reinterpret_cast<const char*>(data.data()), data.size() * sizeof(double));
    if (!ofs)
      return fail(s, PN_ERR_ARG, "cannot write " + (fs::path(dir) / (name + ".data")).string());
  }
  Json root;
  Json& meta            = root["meta"];
  meta["endian"]        = Json::integer(endian_flag());
  meta["rawfile"]       = Json::string(name + ".data");
  meta["layout"]        = Json::integer(1); // nix::ARRAY_LAYOUT, row-major
  meta["decimate"]      = Json::integer(decimate);
  meta["time"]          = Json::real(time);
  meta["step"]          = Json::integer(step);
  meta["chunk_id_range"] = Json::ints(std::vector<int>{0, nk - 1});
  root["dataset"]       = dataset;
  if (!write_json((fs::path(dir) / (name + ".json")).string(), root))
    return fail(s, PN_ERR_ARG, "cannot write the field metadata file");
  return PN_OK;
}

int pn_write_particle_diag(pn_sim* h, const char* dir, int step, double time, double fraction,
                           int seed)
{
  Sim* sp = picnix::sim_state(h);
  if (!sp || !dir)
    return PN_ERR_ARG;
  Sim&                s  = *sp;
  const int           nk = s.g.nchunk, ns = s.g.Ns;
  std::vector<double> data, xu;
  Json                dataset;
  dataset.kind = Json::Obj;
  for (int is = 0; is < ns; is++) {
    const size_t disp0 = data.size() * sizeof(double);
    for (int k = 0; k < nk; k++) {
      int64_t n = 0;
      int     rc = pn_get_np(h, k, is, &n);
      if (rc)
        return rc;
      xu.resize((size_t)std::max<int64_t>(n, 1) * 7);
      if ((rc = pn_get_particles(h, k, is, xu.data(), n, &n)))
        return rc;
      // ParticlePacker::generate_random_index (pic/diag/particle.hpp:30-46)
      const int            N = (int)n;
      const int            M = std::min(static_cast<int>(N * fraction), N);
      std::mt19937_64      engine(seed);
      std::vector<int64_t> index(N);
      std::iota(index.begin(), index.end(), 0);
      if (M < N) {
        std::shuffle(index.begin(), index.end(), engine);
        index.resize(M);
        std::sort(index.begin(), index.end());
      }
      for (int64_t ip : index)
        data.insert(data.end(), xu.begin() + ip * 7, xu.begin() + ip * 7 + 7);
    }
    const size_t nbyte = data.size() * sizeof(double) - disp0;
    char         name[16], desc[40];
    std::snprintf(name, sizeof(name), "up%02d", is);
    std::snprintf(desc, sizeof(desc), "particle species %02d", is);
    put_metadata(dataset, name, desc, disp0, nbyte, {(int64_t)(nbyte / (7 * sizeof(double))), 7});
  }
  namespace fs = std::filesystem;
  std::error_code ec;
  fs::create_directories(dir, ec);
  const std::string fname = step_name(step);
  {
    std::ofstream ofs(fs::path(dir) / (fname + ".data"), std::ios::binary);
    ofs.write(reinterpret_cast<const char*>(data.data()), data.size() * sizeof(double));
    if (!ofs)
      return fail(s, PN_ERR_ARG, "cannot write the particle data file");
  }
  Json  root;
  Json& meta      = root["meta"];
  meta["endian"]  = Json::integer(endian_flag());
  meta["rawfile"] = Json::string(fname + ".data");
  meta["layout"]  = Json::integer(1);
  meta["time"]    = Json::real(time);
  meta["step"]    = Json::integer(step);
  root["dataset"] = dataset;
  if (!write_json((fs::path(dir) / (fname + ".json")).string(), root))
    return fail(s, PN_ERR_ARG, "cannot write the particle metadata file");
  return PN_OK;
}

int pn_write_profile(pn_sim* h, const char* path, int step, double time)
{
  Sim* sp = picnix::sim_state(h);
  if (!sp || !path)
    return PN_ERR_ARG;
  if (!write_msgpack(path, app_state(*sp, step, time)))
    return fail(*sp, PN_ERR_ARG, std::string("cannot write ") + path);
  return PN_OK;
}

int pn_save_state(pn_sim* h, const char* prefix, int step, double time)
{
  Sim* sp = picnix::sim_state(h);
  if (!sp || !prefix)
    return PN_ERR_ARG;
  Sim& s = *sp;
  try {
    if (hipStreamSynchronize(s.stream) != hipSuccess)
      return fail(s, PN_ERR_HIP, "stream synchronisation failed before the checkpoint");
    std::vector<int32_t> np(s.g.nseg);
    if (hipMemcpy(np.data(), s.d_np, np.size() * sizeof(int32_t), hipMemcpyDeviceToHost) !=
        hipSuccess)
      return fail(s, PN_ERR_HIP, "particle counts copy failed");
    Json state          = app_state(s, step, time);
    Json& lib           = state["picnix_amd"];
    lib["precision"]    = Json::string(s.fp32 ? "fp32" : "fp64");
    lib["sorted"]       = Json::boolean(s.sorted);
    lib["rho_valid"]    = Json::boolean(s.rho_valid);
    lib["np"]           = Json::ints(np);
    if (s.g.xwall) {
      // MainChunk's option["random_seed"] and option["boundary"] state per chunk
      Json& sh = lib["shock"];
      sh.kind  = Json::Obj;
      Json ids = Json::array();
      for (int64_t v : s.shk.lastid)
        ids.push(Json::integer(v));
      sh["random_seed"] = Json::ints(s.shk.random_seed);
      sh["lastid"]      = ids;
      sh["influx"]      = Json::ints(s.shk.influx);
      sh["efflux"]      = Json::ints(s.shk.efflux);
    }
    lib["rawfile"]      = Json::string(std::filesystem::path(std::string(prefix) + ".data").filename().string());
    Json&        blocks = lib["blocks"];
    blocks.kind         = Json::Obj;
    std::ofstream ofs(std::string(prefix) + ".data", std::ios::binary);
    if (!ofs)
      return fail(s, PN_ERR_ARG, std::string("cannot write ") + prefix + ".data");
    std::vector<char> host;
    size_t            disp = 0;
    for (const Block& bl : state_blocks(s, np)) {
      host.resize(bl.bytes);
      if (bl.bytes &&
          hipMemcpy(host.data(), bl.dev, bl.bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(s, PN_ERR_HIP, "device copy failed for block " + bl.name);
      ofs.write(host.data(), (std::streamsize)bl.bytes);
      Json& e     = blocks[bl.name];
      e["offset"] = Json::integer((int64_t)disp);
      e["size"]   = Json::integer((int64_t)bl.bytes);
      disp += bl.bytes;
    }
    if (!ofs)
      return fail(s, PN_ERR_ARG, std::string("write failed: ") + prefix + ".data");
    if (!write_msgpack(std::string(prefix) + ".msgpack", state))
      return fail(s, PN_ERR_ARG, std::string("cannot write ") + prefix + ".msgpack");
  } catch (const std::exception& e) {
    return fail(s, PN_ERR_ARG, e.what());
  }
  return PN_OK;
}

int pn_load_state(pn_sim* h, const char* prefix, int* step, double* time)
{
  Sim* sp = picnix::sim_state(h);
  if (!sp || !prefix)
    return PN_ERR_ARG;
  Sim& s = *sp;
  try {
    std::ifstream in(std::string(prefix) + ".msgpack", std::ios::binary);
    if (!in)
      return fail(s, PN_ERR_ARG, std::string("cannot open ") + prefix + ".msgpack");
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(in)),
                               std::istreambuf_iterator<char>());
    Reader rd{bytes};
    Json   state = rd.read();
    // Application::from_json (nix/application.cpp:28-48): the same grid, chunks,
    // process count and [parameter] table
    Json cur = app_state(s, 0, 0.0);
    bool ok  = same_json(cur["ndims"], state["ndims"]) && same_json(cur["cdims"], state["cdims"]) &&
              same_json(cur["nprocess"], state["nprocess"]) &&
              same_json(cur["configuration"]["parameter"], state["configuration"]["parameter"]);
    Json& lib = state["picnix_amd"];
    ok        = ok && lib["precision"].s == (s.fp32 ? "fp32" : "fp64") &&
         same_json(cur["configuration"]["application"]["option"],
                   state["configuration"]["application"]["option"]);
    if (!ok)
      return fail(s, PN_ERR_CONFIG, "trying to load an inconsistent state");
    std::vector<int32_t> np;
    for (const Json& v : lib["np"].a)
      np.push_back((int32_t)v.i);
    if ((int)np.size() != s.g.nseg)
      return fail(s, PN_ERR_CONFIG, "state file has another segment count");
    for (int32_t n : np)
      if (n < 0 || n > s.g.cap)
        return fail(s, PN_ERR_CAPACITY, "state file holds more particles than a segment's capacity");
    std::ifstream dat(std::string(prefix) + ".data", std::ios::binary);
    if (!dat)
      return fail(s, PN_ERR_ARG, std::string("cannot open ") + prefix + ".data");
    if (hipStreamSynchronize(s.stream) != hipSuccess)
      return fail(s, PN_ERR_HIP, "stream synchronisation failed before the restart");
    std::vector<char> host;
    for (const Block& bl : state_blocks(s, np)) {
      auto it = lib["blocks"].o.find(bl.name);
      if (it == lib["blocks"].o.end() || (size_t)it->second.o.at("size").i != bl.bytes)
        return fail(s, PN_ERR_CONFIG, "state file lacks or mis-sizes block " + bl.name);
      host.resize(bl.bytes);
      dat.seekg(it->second.o.at("offset").i);
      dat.read(host.data(), (std::streamsize)bl.bytes);
      if (!dat)
        return fail(s, PN_ERR_ARG, "truncated state data for block " + bl.name);
      if (bl.bytes && hipMemcpy(const_cast<void*>(bl.dev), host.data(), bl.bytes,
                                hipMemcpyHostToDevice) != hipSuccess)
        return fail(s, PN_ERR_HIP, "device copy failed for block " + bl.name);
    }
    if (hipMemcpy(s.d_np, np.data(), np.size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
        hipSuccess)
      return fail(s, PN_ERR_HIP, "particle counts copy failed");
    if (s.g.xwall) {
      auto it = lib.o.find("shock");
      if (it == lib.o.end())
        return fail(s, PN_ERR_CONFIG, "state file lacks the shock boundary state");
      Json& sh = it->second;
      for (int c = 0; c < s.g.nchunk; c++) {
        s.shk.random_seed[c] = (int)sh["random_seed"].a.at(c).i;
        s.shk.lastid[c]      = sh["lastid"].a.at(c).i;
      }
      for (size_t k = 0; k < s.shk.influx.size(); k++) {
        s.shk.influx[k] = (int)sh["influx"].a.at(k).i;
        s.shk.efflux[k] = (int)sh["efflux"].a.at(k).i;
      }
    }
    s.sorted        = lib["sorted"].b;
    s.rho_valid     = lib["rho_valid"].b;
    s.counted       = false;
    s.np_host_valid = false;
    // recorded step graphs refer to the buffer roles of before
    for (auto& ge : s.gexec)
      if (ge) {
        (void)hipGraphExecDestroy(ge);
        ge = nullptr;
      }
    s.buf0_x = nullptr;
    if (step)
      *step = (int)state["curstep"].i;
    if (time)
      *time = state["curtime"].f;
  } catch (const std::exception& e) {
    return fail(s, PN_ERR_CONFIG, std::string("bad state file: ") + e.what());
  }
  return PN_OK;
}

} // extern "C"
