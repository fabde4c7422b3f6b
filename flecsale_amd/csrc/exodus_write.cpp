// Exodus-II output: what the apps' `output` task writes (apps/hydro/
// tasks.h:703-734 writes density and pressure, apps/maire_hydro/tasks.h:
// 661-692 density, through flecsi-sp's io_exodus on the SEACAS libraries,
// neither vendored).  One file per call, holding the mesh as it is at that
// time (a Lagrangian mesh moves), one element block, one time step and the
// named cell variables.
//
// The file is netCDF "classic" with 64-bit offsets (CDF-2), written here
// directly per the netCDF file-format specification (header lists, then the
// fixed-size variables, then the one record of the record variables; every
// value big-endian), with the dimension, variable and attribute names the
// Exodus-II data model defines (coordx/coordy/coordz, connect1 with its
// elem_type, time_whole, name_elem_var, vals_elem_var<i>eb1, ...), so Exodus
// readers and generic netCDF readers both read it.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

namespace fsale {
namespace {

enum : uint32_t { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum : uint32_t { TAG_DIMENSION = 0x0A, TAG_VARIABLE = 0x0B, TAG_ATTRIBUTE = 0x0C };

constexpr int kLenName = 33;   // Exodus names: 32 characters and a terminator

uint32_t type_size(uint32_t t) { return t == NC_CHAR ? 1 : t == NC_DOUBLE ? 8 : 4; }
uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

class Buf {
 public:
  std::vector<uint8_t> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void f64(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    u64(u);
  }
  void f32(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u32(u);
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
};

struct Att {
  std::string name;
  uint32_t type;
  std::string text;   // NC_CHAR
  double num;         // one NC_INT / NC_FLOAT / NC_DOUBLE value
};

struct Var {
  std::string name;
  std::vector<int> dims;   // dimension ids; a leading record dimension makes a record variable
  uint32_t type;
  std::vector<Att> atts;
  std::vector<uint8_t> data;   // big-endian values (one record's for a record variable)
  bool record = false;
  uint64_t vsize = 0, begin = 0;
};

struct Dim {
  std::string name;
  uint64_t len;   // 0: the record (unlimited) dimension
};

void put_atts(Buf &h, const std::vector<Att> &atts) {
  if (atts.empty()) {
    h.u32(0);
    h.u32(0);
    return;
  }
  h.u32(TAG_ATTRIBUTE);
  h.u32((uint32_t)atts.size());
  for (const Att &a : atts) {
    h.name(a.name);
    h.u32(a.type);
    if (a.type == NC_CHAR) {
      h.u32((uint32_t)a.text.size());
      h.b.insert(h.b.end(), a.text.begin(), a.text.end());
    } else {
      h.u32(1);
      if (a.type == NC_INT) h.u32((uint32_t)(int32_t)a.num);
      if (a.type == NC_FLOAT) h.f32((float)a.num);
      if (a.type == NC_DOUBLE) h.f64(a.num);
    }
    h.pad();
  }
}

// the header with every variable's begin offset set, 64-bit offsets
std::vector<uint8_t> header(const std::vector<Dim> &dims, const std::vector<Att> &gatts,
                            const std::vector<Var> &vars, uint32_t numrecs) {
  Buf h;
  h.b = {'C', 'D', 'F', 2};
  h.u32(numrecs);
  h.u32(TAG_DIMENSION);
  h.u32((uint32_t)dims.size());
  for (const Dim &d : dims) {
    h.name(d.name);
    h.u32((uint32_t)d.len);
  }
  put_atts(h, gatts);
  h.u32(TAG_VARIABLE);
  h.u32((uint32_t)vars.size());
  for (const Var &v : vars) {
    h.name(v.name);
    h.u32((uint32_t)v.dims.size());
    for (int id : v.dims) h.u32((uint32_t)id);
    put_atts(h, v.atts);
    h.u32(v.type);
    h.u32(v.vsize > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)v.vsize);
    h.u64(v.begin);
  }
  return std::move(h.b);
}

int fail_write(const std::string &msg) {
  set_message(msg);
  return FSALE_ERR_ARG;
}

// names padded to kLenName bytes each (a char [n][len_name] variable)
std::vector<uint8_t> name_table(const std::vector<std::string> &names) {
  std::vector<uint8_t> t(names.size() * kLenName, 0);
  for (size_t i = 0; i < names.size(); ++i)
    std::memcpy(&t[i * kLenName], names[i].data(), std::min<size_t>(names[i].size(), kLenName - 1));
  return t;
}

int write_exodus(const char *path, const char *title, int dim, int nv, const double *coords, int nc, int vpc,
                 const int32_t *cells, double time, int nvars, const char *const *names,
                 const double *const *values) {
  if (!path || !coords || !cells || (nvars > 0 && (!names || !values)) || nvars < 0) return FSALE_ERR_ARG;
  if (!((dim == 2 && vpc == 4) || (dim == 3 && vpc == 8)))
    return fail_write("Exodus output takes quad4 (2D) and hex8 (3D) meshes");
  if (nv <= 0 || nc <= 0) return fail_write("Exodus output of an empty mesh");
  for (int i = 0; i < nvars; ++i) {
    if (!names[i] || !values[i]) return FSALE_ERR_ARG;
    if (std::strlen(names[i]) >= (size_t)kLenName) return fail_write("Exodus variable name over 32 characters");
  }
  for (size_t i = 0; i < (size_t)nc * vpc; ++i)
    if (cells[i] < 0 || cells[i] >= nv) return fail_write("Exodus output: a cell names a vertex out of range");

  // dimensions
  std::vector<Dim> dims = {{"len_string", kLenName}, {"len_line", 81}, {"four", 4}, {"len_name", kLenName},
                           {"time_step", 0}, {"num_dim", (uint64_t)dim}, {"num_nodes", (uint64_t)nv},
                           {"num_elem", (uint64_t)nc}, {"num_el_blk", 1}, {"num_el_in_blk1", (uint64_t)nc},
                           {"num_nod_per_el1", (uint64_t)vpc}};
  const int D_LEN_NAME = 3, D_TIME = 4, D_DIM = 5, D_NODES = 6, D_BLK = 8, D_IN_BLK = 9, D_PER_EL = 10;
  int D_VAR = -1;
  if (nvars > 0) {
    D_VAR = (int)dims.size();
    dims.push_back({"num_elem_var", (uint64_t)nvars});
  }

  const std::vector<Att> gatts = {{"api_version", NC_FLOAT, "", 5.22}, {"version", NC_FLOAT, "", 5.22},
                                  {"floating_point_word_size", NC_INT, "", 8}, {"file_size", NC_INT, "", 1},
                                  {"maximum_name_length", NC_INT, "", 32}, {"int64_status", NC_INT, "", 0},
                                  {"title", NC_CHAR, title ? std::string(title) : std::string(), 0}};

  std::vector<Var> vars;
  auto add = [&](const std::string &name, std::vector<int> d, uint32_t type, std::vector<Att> atts = {}) -> Var & {
    vars.push_back(Var{name, std::move(d), type, std::move(atts), {}, false, 0, 0});
    return vars.back();
  };
  {
    Var &v = add("time_whole", {D_TIME}, NC_DOUBLE);
    Buf b;
    b.f64(time);
    v.data = std::move(b.b);
  }
  {
    Var &v = add("eb_status", {D_BLK}, NC_INT);
    Buf b;
    b.u32(1);
    v.data = std::move(b.b);
  }
  {
    Var &v = add("eb_prop1", {D_BLK}, NC_INT, {{"name", NC_CHAR, "ID", 0}});
    Buf b;
    b.u32(1);
    v.data = std::move(b.b);
  }
  add("eb_names", {D_BLK, D_LEN_NAME}, NC_CHAR).data = name_table({""});
  add("coor_names", {D_DIM, D_LEN_NAME}, NC_CHAR).data =
      name_table(dim == 2 ? std::vector<std::string>{"x", "y"} : std::vector<std::string>{"x", "y", "z"});
  static const char *const kCoord[3] = {"coordx", "coordy", "coordz"};
  for (int d = 0; d < dim; ++d) {
    Var &v = add(kCoord[d], {D_NODES}, NC_DOUBLE);
    Buf b;
    b.b.reserve((size_t)nv * 8);
    for (int i = 0; i < nv; ++i) b.f64(coords[(size_t)i * dim + d]);
    v.data = std::move(b.b);
  }
  {
    Var &v = add("connect1", {D_IN_BLK, D_PER_EL}, NC_INT, {{"elem_type", NC_CHAR, dim == 2 ? "QUAD4" : "HEX8", 0}});
    Buf b;
    b.b.reserve((size_t)nc * vpc * 4);
    for (size_t i = 0; i < (size_t)nc * vpc; ++i) b.u32((uint32_t)(cells[i] + 1));   // 1-based
    v.data = std::move(b.b);
  }
  if (nvars > 0) {
    std::vector<std::string> vn(names, names + nvars);
    add("name_elem_var", {D_VAR, D_LEN_NAME}, NC_CHAR).data = name_table(vn);
    Var &tab = add("elem_var_tab", {D_BLK, D_VAR}, NC_INT);
    Buf b;
    for (int i = 0; i < nvars; ++i) b.u32(1);
    tab.data = std::move(b.b);
    for (int i = 0; i < nvars; ++i) {
      Var &v = add("vals_elem_var" + std::to_string(i + 1) + "eb1", {D_TIME, D_IN_BLK}, NC_DOUBLE);
      Buf c;
      c.b.reserve((size_t)nc * 8);
      for (int k = 0; k < nc; ++k) c.f64(values[i][k]);
      v.data = std::move(c.b);
    }
  }

  // sizes and offsets: fixed-size variables in order, then the record
  for (Var &v : vars) {
    v.record = !v.dims.empty() && v.dims[0] == D_TIME;
    uint64_t n = type_size(v.type);
    for (size_t k = v.record ? 1 : 0; k < v.dims.size(); ++k) n *= dims[v.dims[k]].len;
    if (v.data.size() != n) return fail_write("Exodus output: internal size mismatch in " + v.name);
    v.vsize = pad4(n);
  }
  uint64_t off = header(dims, gatts, vars, 1).size();
  for (Var &v : vars)
    if (!v.record) {
      v.begin = off;
      off += v.vsize;
    }
  for (Var &v : vars)
    if (v.record) {
      v.begin = off;
      off += v.vsize;
    }
  for (const Var &v : vars)
    if (!v.record && v.vsize > 0xFFFFFFFCull) return fail_write("Exodus output: a variable over 4 GiB");

  FILE *f = std::fopen(path, "wb");
  if (!f) return fail_write(std::string("cannot write ") + path);
  const std::vector<uint8_t> h = header(dims, gatts, vars, 1);
  bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
  static const uint8_t zeros[4] = {0, 0, 0, 0};
  for (int pass = 0; pass < 2 && ok; ++pass)
    for (const Var &v : vars) {
      if (v.record != (pass == 1)) continue;
      ok = ok && std::fwrite(v.data.data(), 1, v.data.size(), f) == v.data.size();
      ok = ok && std::fwrite(zeros, 1, v.vsize - v.data.size(), f) == v.vsize - v.data.size();
    }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail_write(std::string("cannot write ") + path);
  return FSALE_OK;
}

}  // namespace

int exodus_write(const char *path, const char *title, int dim, int nv, const double *coords, int nc, int vpc,
                 const int32_t *cells, double time, int nvars, const char *const *names,
                 const double *const *values) {
  try {
    return write_exodus(path, title, dim, nv, coords, nc, vpc, cells, time, nvars, names, values);
  } catch (const std::bad_alloc &) {
    return fail_write(std::string("out of host memory writing ") + (path ? path : ""));
  }
}

}  // namespace fsale

extern "C" int fsale_exodus_write(const char *path, const char *title, int32_t dim, int32_t num_vertices,
                                  const double *coords, int32_t num_cells, int32_t verts_per_cell,
                                  const int32_t *cells, double time, int32_t num_vars, const char *const *names,
                                  const double *const *values) {
  return fsale::exodus_write(path, title, dim, num_vertices, coords, num_cells, verts_per_cell, cells, time,
                             num_vars, names, values);
}
