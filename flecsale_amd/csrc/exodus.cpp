// Exodus-II mesh files, the meshes the reference's apps read with `-m`
// (data/meshes/*.g; apps/hydro/2d/CMakeLists.txt:38-48 runs hydro_2d -m
// square_32x32.g).  The reference reads them through flecsi-sp's Exodus
// definition on the SEACAS / netCDF libraries (not vendored); this reader
// takes what the hydro paths use: vertex coordinates (coordx/coordy/coordz,
// or the older 2-D `coord` array) and the element blocks' connectivity
// (connect1..connectN, 1-based, concatenated in block order).  Vertex and
// cell numbering are the file's (the burton mesh keeps Exodus order).
//
// Encodings: the netCDF "classic" formats (CDF-1, CDF-2 64-bit offset, CDF-5)
// are parsed here directly (header lists and big-endian variables, per the
// netCDF file-format specification); the netCDF-4 encoding is an HDF5 file,
// read through the HDF5 C library when one can be loaded at run time
// (FSALE_HDF5_LIB, else libhdf5.so on the library path).
#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

namespace fsale {
namespace {

struct ExoMesh {
  int dim = 0, nv = 0, nc = 0, vpc = 0;
  std::vector<double> coords;   // [nv][dim]
  std::vector<int32_t> cells;   // [nc][vpc], 0-based
};

int fail_mesh(const std::string &msg) {
  set_message(msg);
  return FSALE_ERR_MESH;
}

// checks shared by both encodings: cell type and index range
int finish_mesh(ExoMesh &m) {
  if (!((m.dim == 2 && m.vpc == 4) || (m.dim == 3 && m.vpc == 8)))
    return fail_mesh("Exodus element type with " + std::to_string(m.vpc) + " nodes in " + std::to_string(m.dim) +
                     "D: the hydro paths take quad4 (2D) and hex8 (3D) elements");
  for (int32_t &v : m.cells) {
    v -= 1;
    if (v < 0 || v >= m.nv) return fail_mesh("Exodus connectivity refers to a node out of range");
  }
  return FSALE_OK;
}

// ------------------------------------------------------- netCDF classic
class NcClassic {
 public:
  struct Var {
    std::string name;
    std::vector<uint64_t> dims;
    int type = 0;
    uint64_t begin = 0;
  };

  explicit NcClassic(std::vector<uint8_t> bytes) : b_(std::move(bytes)) {}

  int parse(std::string &err) {
    if (b_.size() < 8 || b_[0] != 'C' || b_[1] != 'D' || b_[2] != 'F') return bad(err, "not a netCDF file");
    ver_ = b_[3];
    if (ver_ != 1 && ver_ != 2 && ver_ != 5) return bad(err, "unknown netCDF version");
    pos_ = 4;
    uint64_t numrecs;
    if (!count(numrecs)) return bad(err, "truncated header");
    // dimensions
    uint64_t tag, n;
    if (!u32(tag) || !count(n)) return bad(err, "truncated header");
    if (tag == 0x0A) {
      for (uint64_t i = 0; i < n; ++i) {
        std::string name;
        uint64_t len;
        if (!text(name) || !count(len)) return bad(err, "bad dimension list");
        dims_.push_back({name, len});
      }
    } else if (tag != 0 || n != 0) {
      return bad(err, "bad dimension list");
    }
    if (!attributes()) return bad(err, "bad attribute list");
    // variables
    if (!u32(tag) || !count(n)) return bad(err, "truncated header");
    if (tag == 0x0B) {
      for (uint64_t i = 0; i < n; ++i) {
        Var v;
        uint64_t rank;
        if (!text(v.name) || !count(rank)) return bad(err, "bad variable list");
        for (uint64_t r = 0; r < rank; ++r) {
          uint64_t id;
          if (!count(id) || id >= dims_.size()) return bad(err, "bad variable dimensions");
          v.dims.push_back(id);
        }
        uint64_t type, vsize;
        if (!attributes() || !u32(type) || !count(vsize) || !offset(v.begin)) return bad(err, "bad variable");
        v.type = (int)type;
        vars_.push_back(v);
      }
    } else if (tag != 0 || n != 0) {
      return bad(err, "bad variable list");
    }
    return FSALE_OK;
  }

  bool has_dim(const char *name, uint64_t *len) const {
    for (const auto &d : dims_)
      if (d.first == name) {
        if (len) *len = d.second;
        return true;
      }
    return false;
  }
  const Var *var(const std::string &name) const {
    for (const auto &v : vars_)
      if (v.name == name) return &v;
    return nullptr;
  }
  // a non-record variable's values as doubles or int32
  template <class T>
  bool read(const std::string &name, std::vector<T> &out, std::vector<uint64_t> *shape = nullptr) const {
    const Var *v = var(name);
    if (!v) return false;
    uint64_t n = 1;
    std::vector<uint64_t> sh;
    for (uint64_t id : v->dims) {
      const uint64_t len = dims_[id].second;
      if (len == 0 || len > b_.size() || n > b_.size() / len) return false;   // record variable, or larger than the file
      n *= len;
      sh.push_back(len);
    }
    static const int size[] = {0, 1, 1, 2, 4, 4, 8, 1, 2, 4, 8, 8};
    if (v->type < 1 || v->type > 11) return false;
    const int w = size[v->type];
    if (v->begin > b_.size() || n > (b_.size() - v->begin) / w) return false;
    out.resize(n);
    const uint8_t *p = b_.data() + v->begin;
    for (uint64_t i = 0; i < n; ++i, p += w) {
      uint64_t raw = 0;
      for (int k = 0; k < w; ++k) raw = (raw << 8) | p[k];
      double val;
      switch (v->type) {
        case 1: val = (double)(int8_t)raw; break;
        case 3: val = (double)(int16_t)raw; break;
        case 4: val = (double)(int32_t)(uint32_t)raw; break;
        case 5: {
          uint32_t u = (uint32_t)raw;
          float f;
          std::memcpy(&f, &u, 4);
          val = f;
          break;
        }
        case 6: std::memcpy(&val, &raw, 8); break;
        case 9: val = (double)(uint32_t)raw; break;
        case 10: val = (double)(int64_t)raw; break;
        default: val = (double)raw; break;
      }
      out[i] = (T)val;
    }
    if (shape) *shape = sh;
    return true;
  }

 private:
  int bad(std::string &err, const char *what) {
    err = what;
    return FSALE_ERR_MESH;
  }
  bool u32(uint64_t &v) {
    if (pos_ + 4 > b_.size()) return false;
    v = ((uint64_t)b_[pos_] << 24) | ((uint64_t)b_[pos_ + 1] << 16) | ((uint64_t)b_[pos_ + 2] << 8) | b_[pos_ + 3];
    pos_ += 4;
    return true;
  }
  bool u64(uint64_t &v) {
    uint64_t hi, lo;
    if (!u32(hi) || !u32(lo)) return false;
    v = (hi << 32) | lo;
    return true;
  }
  // counts and lengths: 4 bytes, 8 in CDF-5; offsets: 4 bytes in CDF-1
  bool count(uint64_t &v) { return ver_ == 5 ? u64(v) : u32(v); }
  bool offset(uint64_t &v) { return ver_ == 1 ? u32(v) : u64(v); }
  bool text(std::string &s) {
    uint64_t n;
    if (!count(n) || pos_ + n > b_.size()) return false;
    s.assign((const char *)b_.data() + pos_, n);
    pos_ += (n + 3) / 4 * 4;
    return pos_ <= b_.size();
  }
  bool attributes() {
    uint64_t tag, n;
    if (!u32(tag) || !count(n)) return false;
    if (tag == 0 && n == 0) return true;
    if (tag != 0x0C) return false;
    static const int size[] = {0, 1, 1, 2, 4, 4, 8, 1, 2, 4, 8, 8};
    for (uint64_t i = 0; i < n; ++i) {
      std::string name;
      uint64_t type, len;
      if (!text(name) || !u32(type) || !count(len) || type < 1 || type > 11) return false;
      pos_ += (len * size[type] + 3) / 4 * 4;
      if (pos_ > b_.size()) return false;
    }
    return true;
  }

  std::vector<uint8_t> b_;
  size_t pos_ = 0;
  int ver_ = 0;
  std::vector<std::pair<std::string, uint64_t>> dims_;
  std::vector<Var> vars_;
};

int read_classic(std::vector<uint8_t> bytes, ExoMesh &m) {
  NcClassic nc(std::move(bytes));
  std::string err;
  int rc = nc.parse(err);
  if (rc) return fail_mesh("Exodus (netCDF classic): " + err);
  uint64_t dim = 0, nv = 0, nblk = 0;
  if (!nc.has_dim("num_dim", &dim) || !nc.has_dim("num_nodes", &nv) || !nc.has_dim("num_el_blk", &nblk))
    return fail_mesh("not an Exodus file (num_dim / num_nodes / num_el_blk missing)");
  if (nv > (uint64_t)INT32_MAX || nblk > 1000000) return fail_mesh("Exodus mesh too large");
  m.dim = (int)dim;
  m.nv = (int)nv;
  if (m.dim != 2 && m.dim != 3)
    return fail_mesh("Exodus mesh of dimension " + std::to_string(m.dim) + ": the hydro paths are 2D and 3D");
  // coordinates are read (and so checked against the file's size) before
  // anything is sized from the header's node count
  static const char *names[3] = {"coordx", "coordy", "coordz"};
  std::vector<double> comp[3], c;
  if (nc.var("coordx")) {
    for (int d = 0; d < m.dim; ++d)
      if (!nc.read(names[d], comp[d]) || comp[d].size() != nv) return fail_mesh(std::string("Exodus: bad ") + names[d]);
  } else {
    if (!nc.read("coord", c) || c.size() != nv * dim) return fail_mesh("Exodus: no coordinates");
    for (int d = 0; d < m.dim; ++d) comp[d].assign(c.begin() + (size_t)d * m.nv, c.begin() + (size_t)(d + 1) * m.nv);
  }
  m.coords.assign((size_t)m.nv * m.dim, 0.0);
  for (int d = 0; d < m.dim; ++d)
    for (int i = 0; i < m.nv; ++i) m.coords[(size_t)i * m.dim + d] = comp[d][i];
  for (uint64_t b = 1; b <= nblk; ++b) {
    std::vector<int32_t> conn;
    std::vector<uint64_t> shape;
    if (!nc.read("connect" + std::to_string(b), conn, &shape) || shape.size() != 2)
      return fail_mesh("Exodus: bad connect" + std::to_string(b));
    if (m.vpc && (int)shape[1] != m.vpc) return fail_mesh("Exodus: element blocks of different node counts");
    m.vpc = (int)shape[1];
    m.cells.insert(m.cells.end(), conn.begin(), conn.end());
    if (m.cells.size() / m.vpc > (size_t)INT32_MAX) return fail_mesh("Exodus mesh too large");
  }
  m.nc = m.vpc ? (int)(m.cells.size() / m.vpc) : 0;
  return finish_mesh(m);
}

// --------------------------------------------------------- netCDF-4 / HDF5
// The few HDF5 entry points used, loaded at run time (HDF5 1.10+: hid_t is
// 64-bit).
struct Hdf5 {
  typedef int64_t hid_t;
  void *h = nullptr;
  int (*open)() = nullptr;
  hid_t (*fopen)(const char *, unsigned, hid_t) = nullptr;
  int (*fclose)(hid_t) = nullptr;
  hid_t (*dopen)(hid_t, const char *, hid_t) = nullptr;
  int (*dclose)(hid_t) = nullptr;
  hid_t (*dspace)(hid_t) = nullptr;
  int (*sndims)(hid_t) = nullptr;
  int (*sdims)(hid_t, uint64_t *, uint64_t *) = nullptr;
  int (*sclose)(hid_t) = nullptr;
  int (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *) = nullptr;
  int (*lexists)(hid_t, const char *, hid_t) = nullptr;
  int (*eset)(hid_t, void *, void *) = nullptr;
  hid_t *t_double = nullptr, *t_int = nullptr;

  bool load() {
    if (h) return true;
    // FSALE_HDF5_LIB, the library path, then a conda installation's
    const char *env = std::getenv("FSALE_HDF5_LIB");
    const char *tries[] = {env, "libhdf5.so", "libhdf5.so.310", "libhdf5.so.200", "libhdf5.so.103",
                           "/opt/conda/lib/libhdf5.so"};
    for (const char *name : tries)
      if (name && (h = dlopen(name, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) return false;
    auto sym = [&](const char *n) { return dlsym(h, n); };
    open = (int (*)())sym("H5open");
    fopen = (hid_t(*)(const char *, unsigned, hid_t))sym("H5Fopen");
    fclose = (int (*)(hid_t))sym("H5Fclose");
    dopen = (hid_t(*)(hid_t, const char *, hid_t))sym("H5Dopen2");
    dclose = (int (*)(hid_t))sym("H5Dclose");
    dspace = (hid_t(*)(hid_t))sym("H5Dget_space");
    sndims = (int (*)(hid_t))sym("H5Sget_simple_extent_ndims");
    sdims = (int (*)(hid_t, uint64_t *, uint64_t *))sym("H5Sget_simple_extent_dims");
    sclose = (int (*)(hid_t))sym("H5Sclose");
    dread = (int (*)(hid_t, hid_t, hid_t, hid_t, hid_t, void *))sym("H5Dread");
    lexists = (int (*)(hid_t, const char *, hid_t))sym("H5Lexists");
    eset = (int (*)(hid_t, void *, void *))sym("H5Eset_auto2");
    t_double = (hid_t *)sym("H5T_NATIVE_DOUBLE_g");
    t_int = (hid_t *)sym("H5T_NATIVE_INT_g");
    const bool ok = open && fopen && fclose && dopen && dclose && dspace && sndims && sdims && sclose && dread &&
                    lexists && t_double && t_int;
    if (!ok || open() < 0) {
      dlclose(h);
      h = nullptr;
      return false;
    }
    if (eset) eset(0, nullptr, nullptr);   // H5E_DEFAULT: no error stack printing
    return true;
  }
};

Hdf5 &hdf5() {
  static Hdf5 lib;
  return lib;
}

// dataset `name` of `file` into out (shape in *shape); false if absent
template <class T>
bool h5_read(Hdf5 &L, int64_t file, const char *name, int64_t type, std::vector<T> &out,
             std::vector<uint64_t> *shape) {
  if (L.lexists(file, name, 0) <= 0) return false;
  const int64_t ds = L.dopen(file, name, 0);
  if (ds < 0) return false;
  const int64_t sp = L.dspace(ds);
  const int nd = sp >= 0 ? L.sndims(sp) : -1;
  bool ok = nd >= 1 && nd <= 4;
  uint64_t dims[4] = {0, 0, 0, 0};
  if (ok) ok = L.sdims(sp, dims, nullptr) >= 0;
  if (sp >= 0) L.sclose(sp);
  uint64_t n = 1;
  for (int i = 0; ok && i < nd; ++i) {
    if (dims[i] != 0 && n > ((uint64_t)1 << 31) / dims[i]) ok = false;   // beyond what the library's arrays index
    n *= dims[i];
  }
  if (ok) {
    out.resize(n);
    ok = L.dread(ds, type, 0, 0, 0, out.data()) >= 0;   // H5S_ALL, H5S_ALL, H5P_DEFAULT
  }
  L.dclose(ds);
  if (ok && shape) shape->assign(dims, dims + nd);
  return ok;
}

int read_exodus_file(const char *path, ExoMesh &m);

int read_hdf5(const char *path, ExoMesh &m) {
  Hdf5 &L = hdf5();
  if (!L.load())
    return fail_mesh("netCDF-4 (HDF5) Exodus file: no HDF5 library could be loaded (set FSALE_HDF5_LIB)");
  const int64_t f = L.fopen(path, 0u, 0);   // H5F_ACC_RDONLY, H5P_DEFAULT
  if (f < 0) return fail_mesh(std::string("cannot open ") + path + " as HDF5");
  int rc = FSALE_OK;
  std::vector<double> cx, cy, cz;
  if (!h5_read(L, f, "coordx", *L.t_double, cx, nullptr) || !h5_read(L, f, "coordy", *L.t_double, cy, nullptr)) {
    rc = fail_mesh("Exodus (HDF5): no coordx / coordy");
  } else {
    m.dim = h5_read(L, f, "coordz", *L.t_double, cz, nullptr) ? 3 : 2;
    m.nv = (int)cx.size();
    if (cy.size() != cx.size() || (m.dim == 3 && cz.size() != cx.size())) rc = fail_mesh("Exodus (HDF5): coordinates");
    m.coords.resize((size_t)m.nv * m.dim);
    for (int i = 0; !rc && i < m.nv; ++i) {
      m.coords[(size_t)i * m.dim] = cx[i];
      m.coords[(size_t)i * m.dim + 1] = cy[i];
      if (m.dim == 3) m.coords[(size_t)i * m.dim + 2] = cz[i];
    }
    for (int b = 1; !rc; ++b) {
      std::vector<int32_t> conn;
      std::vector<uint64_t> shape;
      const std::string name = "connect" + std::to_string(b);
      if (!h5_read(L, f, name.c_str(), *L.t_int, conn, &shape)) {
        if (b == 1) rc = fail_mesh("Exodus (HDF5): no connect1");
        break;
      }
      if (shape.size() != 2 || (m.vpc && (int)shape[1] != m.vpc)) {
        rc = fail_mesh("Exodus (HDF5): bad " + name);
        break;
      }
      m.vpc = (int)shape[1];
      m.cells.insert(m.cells.end(), conn.begin(), conn.end());
    }
  }
  L.fclose(f);
  if (rc) return rc;
  m.nc = (int)(m.cells.size() / m.vpc);
  return finish_mesh(m);
}

int read_exodus(const char *path, ExoMesh &m) {
  if (!path) return FSALE_ERR_ARG;
  try {
    return read_exodus_file(path, m);
  } catch (const std::bad_alloc &) {
    return fail_mesh(std::string("out of host memory reading ") + path);
  }
}

int read_exodus_file(const char *path, ExoMesh &m) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return fail_mesh(std::string("cannot open ") + path);
  std::vector<uint8_t> bytes;
  uint8_t buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0) bytes.insert(bytes.end(), buf, buf + got);
  std::fclose(fp);
  static const uint8_t h5sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  if (bytes.size() >= 8 && std::memcmp(bytes.data(), h5sig, 8) == 0) return read_hdf5(path, m);
  return read_classic(std::move(bytes), m);
}

}  // namespace
}  // namespace fsale

using namespace fsale;

extern "C" {

int fsale_exodus_info(const char *path, int32_t *dim, int32_t *num_cells, int32_t *num_vertices,
                      int32_t *verts_per_cell) {
  ExoMesh m;
  int rc = read_exodus(path, m);
  if (rc) return rc;
  if (dim) *dim = m.dim;
  if (num_cells) *num_cells = m.nc;
  if (num_vertices) *num_vertices = m.nv;
  if (verts_per_cell) *verts_per_cell = m.vpc;
  return FSALE_OK;
}

int fsale_exodus_read(const char *path, double *coords, int32_t *cells) {
  if (!coords || !cells) return FSALE_ERR_ARG;
  ExoMesh m;
  int rc = read_exodus(path, m);
  if (rc) return rc;
  std::memcpy(coords, m.coords.data(), m.coords.size() * sizeof(double));
  std::memcpy(cells, m.cells.data(), m.cells.size() * sizeof(int32_t));
  return FSALE_OK;
}

}  // extern "C"
