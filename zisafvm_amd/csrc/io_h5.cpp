// libzfv_io.so: the reference's mesh and snapshot files with HDF5 (declared
// in include/zfv_io.h).  Host code only; the rate-of-change engine
// (libzfv_hip.so) does not depend on it.
//
//   mesh files      load_grid_gmsh_h5, src/zisa/grid/grid.cpp:889-901
//   renumbering     renumber_grid, src/renumber_grid.cpp:46-135
//   snapshots       save_state / load_state, src/zisa/model/models.cpp:7-27;
//                   AllVariables::save / load, src/zisa/model/all_variables.cpp:74-113
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/zfv_io.h"
#include "common.hpp"
#include "setup.hpp"

namespace {

thread_local std::string g_err;

using zfv::ZfvError;

// the conserved variables' dataset names (euler_var_t::labels)
const char *const kLabels[5] = {"rho", "mv1", "mv2", "mv3", "E"};

// RAII holders of HDF5 ids
struct H5Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Id() {
    if (id >= 0 && close) close(id);
  }
  H5Id(const H5Id &) = delete;
  H5Id &operator=(const H5Id &) = delete;
};

void quiet_hdf5() {
  // errors are reported through return codes, not printed by the library
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}

hid_t open_read(const std::string &path) {
  quiet_hdf5();
  hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) throw ZfvError(ZFV_ERR_INVALID, "cannot open " + path + " for reading");
  return f;
}

hid_t create_file(const std::string &path) {
  quiet_hdf5();
  hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) throw ZfvError(ZFV_ERR_INVALID, "cannot create " + path);
  return f;
}

bool has_dataset(hid_t f, const char *name) { return H5Lexists(f, name, H5P_DEFAULT) > 0; }

// dimensions of a dataset (rank 0 for a scalar)
std::vector<hsize_t> dims_of(hid_t f, const char *name) {
  if (!has_dataset(f, name)) throw ZfvError(ZFV_ERR_INVALID, std::string("no dataset '") + name + "'");
  H5Id d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
  H5Id s(H5Dget_space(d.id), H5Sclose);
  const int rank = H5Sget_simple_extent_ndims(s.id);
  std::vector<hsize_t> dims(size_t(std::max(rank, 0)));
  if (rank > 0) H5Sget_simple_extent_dims(s.id, dims.data(), nullptr);
  return dims;
}

// reads a whole dataset converted to the memory type (HDF5 converts between
// integer widths and signedness, and between float widths, on read)
template <class T>
std::vector<T> read_all(hid_t f, const char *name, hid_t mem_type, std::vector<hsize_t> *dims_out = nullptr) {
  std::vector<hsize_t> dims = dims_of(f, name);
  size_t count = 1;
  for (hsize_t d : dims) count *= size_t(d);
  std::vector<T> out(count);
  H5Id d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
  if (count > 0 && H5Dread(d.id, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
    throw ZfvError(ZFV_ERR_INVALID, std::string("cannot read dataset '") + name + "'");
  if (dims_out) *dims_out = dims;
  return out;
}

template <class T>
T read_scalar(hid_t f, const char *name, hid_t mem_type) {
  std::vector<T> v = read_all<T>(f, name, mem_type);
  if (v.size() != 1) throw ZfvError(ZFV_ERR_INVALID, std::string("dataset '") + name + "' is not a scalar");
  return v[0];
}

void write_array(hid_t f, const char *name, hid_t file_type, hid_t mem_type, const void *data,
                 const std::vector<hsize_t> &dims) {
  H5Id s(dims.empty() ? H5Screate(H5S_SCALAR) : H5Screate_simple(int(dims.size()), dims.data(), nullptr), H5Sclose);
  H5Id d(H5Dcreate2(f, name, file_type, s.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (d.id < 0 || H5Dwrite(d.id, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    throw ZfvError(ZFV_ERR_INVALID, std::string("cannot write dataset '") + name + "'");
}

struct MshH5 {
  int n_dims = 0;
  std::vector<double> vertices;      // n_vertices x 3
  std::vector<uint64_t> vidx;        // n_cells x (n_dims + 1)
  int64_t n_vertices = 0, n_cells = 0;
};

MshH5 read_msh(const std::string &path) {
  H5Id f(open_read(path), H5Fclose);
  MshH5 m;
  m.n_dims = int(read_scalar<int64_t>(f.id, "n_dims", H5T_NATIVE_INT64));
  if (m.n_dims != 2 && m.n_dims != 3) throw ZfvError(ZFV_ERR_INVALID, "n_dims must be 2 or 3");
  std::vector<hsize_t> vd, cd;
  m.vertices = read_all<double>(f.id, "vertices", H5T_NATIVE_DOUBLE, &vd);
  m.vidx = read_all<uint64_t>(f.id, "vertex_indices", H5T_NATIVE_UINT64, &cd);
  if (vd.size() != 2 || vd[1] != 3) throw ZfvError(ZFV_ERR_INVALID, "vertices must be n_vertices x 3");
  if (cd.size() != 2 || int(cd[1]) != m.n_dims + 1)
    throw ZfvError(ZFV_ERR_INVALID, "vertex_indices must be n_cells x (n_dims + 1)");
  m.n_vertices = int64_t(vd[0]);
  m.n_cells = int64_t(cd[0]);
  for (uint64_t v : m.vidx)
    if (v >= uint64_t(m.n_vertices)) throw ZfvError(ZFV_ERR_INVALID, "vertex index out of range");
  return m;
}

void write_msh(const std::string &path, const MshH5 &m) {
  H5Id f(create_file(path), H5Fclose);
  write_array(f.id, "vertices", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, m.vertices.data(),
              {hsize_t(m.n_vertices), 3});
  write_array(f.id, "vertex_indices", H5T_STD_U64LE, H5T_NATIVE_UINT64, m.vidx.data(),
              {hsize_t(m.n_cells), hsize_t(m.n_dims + 1)});
  const uint64_t nd = uint64_t(m.n_dims);
  write_array(f.id, "n_dims", H5T_STD_U64LE, H5T_NATIVE_UINT64, &nd, {});
}

double tet_volume(const double *a, const double *b, const double *c, const double *d) {
  double u[3], v[3], w[3];
  for (int k = 0; k < 3; ++k) {
    u[k] = b[k] - a[k];
    v[k] = c[k] - a[k];
    w[k] = d[k] - a[k];
  }
  return std::abs(u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                  u[2] * (v[0] * w[1] - v[1] * w[0])) /
         6.0;
}

template <class F>
int guarded(F &&f) {
  try {
    f();
    return ZFV_OK;
  } catch (const ZfvError &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_err = e.what();
    return ZFV_ERR_INVALID;
  }
}

}  // namespace

extern "C" {

const char *zfv_io_last_error(void) { return g_err.c_str(); }

int zfv_io_read_msh_h5(const char *path, zfv_mesh_buf **out) {
  return guarded([&] {
    MshH5 m = read_msh(path);
    if (m.n_cells > int64_t(std::numeric_limits<int32_t>::max()) ||
        m.n_vertices > int64_t(std::numeric_limits<int32_t>::max()))
      throw ZfvError(ZFV_ERR_INVALID, "mesh too large for 32-bit cell and vertex ids");
    auto *b = new zfv::MeshBuf;
    const int nv = m.n_dims + 1;
    b->n_dims = m.n_dims;
    b->n_cells = int(m.n_cells);
    b->vertex_ids.resize(size_t(m.n_cells) * nv);
    b->corners.resize(size_t(m.n_cells) * nv * 3);
    b->ghost.assign(size_t(m.n_cells), 0);
    for (int64_t i = 0; i < m.n_cells; ++i)
      for (int k = 0; k < nv; ++k) {
        const uint64_t v = m.vidx[size_t(i) * nv + k];
        b->vertex_ids[size_t(i) * nv + k] = int32_t(v);
        for (int d = 0; d < 3; ++d) b->corners[(size_t(i) * nv + k) * 3 + d] = m.vertices[3 * v + d];
      }
    *out = reinterpret_cast<zfv_mesh_buf *>(b);
  });
}

int zfv_io_mesh_view(const zfv_mesh_buf *buf, zfv_mesh *view) {
  const auto *m = reinterpret_cast<const zfv::MeshBuf *>(buf);
  view->n_dims = m->n_dims;
  view->n_cells = m->n_cells;
  view->vertex_ids = m->vertex_ids.data();
  view->corners = m->corners.data();
  view->ghost_mask = m->ghost.data();
  return ZFV_OK;
}

int zfv_io_mesh_free(zfv_mesh_buf *buf) {
  delete reinterpret_cast<zfv::MeshBuf *>(buf);
  return ZFV_OK;
}

int zfv_io_write_msh_h5(const char *path, int32_t n_dims, int64_t n_vertices, const double *vertices,
                        int64_t n_cells, const int64_t *vertex_indices) {
  return guarded([&] {
    if (n_dims != 2 && n_dims != 3) throw ZfvError(ZFV_ERR_INVALID, "n_dims must be 2 or 3");
    if (n_vertices < 0 || n_cells < 0) throw ZfvError(ZFV_ERR_INVALID, "negative sizes");
    MshH5 m;
    m.n_dims = n_dims;
    m.n_vertices = n_vertices;
    m.n_cells = n_cells;
    m.vertices.assign(vertices, vertices + size_t(n_vertices) * 3);
    m.vidx.resize(size_t(n_cells) * (n_dims + 1));
    for (size_t q = 0; q < m.vidx.size(); ++q) {
      if (vertex_indices[q] < 0 || vertex_indices[q] >= n_vertices)
        throw ZfvError(ZFV_ERR_INVALID, "vertex index out of range");
      m.vidx[q] = uint64_t(vertex_indices[q]);
    }
    write_msh(path, m);
  });
}

int zfv_io_renumber_msh_h5(const char *path, int64_t *order) {
  return guarded([&] {
    MshH5 m = read_msh(path);
    const int nv = m.n_dims + 1;
    const int64_t n = m.n_cells;
    auto vx = [&](int64_t i, int k) { return &m.vertices[3 * m.vidx[size_t(i) * nv + k]]; };
    // sanity_check: tetrahedron volumes within a factor 100 (the tool's
    // check reads four vertices per cell; it applies to tetrahedra only)
    if (m.n_dims == 3 && n > 0) {
      double vmin = std::numeric_limits<double>::max(), vmax = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        const double v = tet_volume(vx(i, 0), vx(i, 1), vx(i, 2), vx(i, 3));
        vmin = std::min(vmin, v);
        vmax = std::max(vmax, v);
      }
      if (vmax / vmin > 100.0) throw ZfvError(ZFV_ERR_INVALID, "Faulty cell found.");
    }
    // the tool's cell centres: first vertex, then every vertex added again,
    // divided by the vertex count (src/renumber_grid.cpp:76-81), and its
    // bounding box (maxima start at the smallest positive double, :65-73)
    std::vector<double> ctr(static_cast<size_t>(n) * 3);
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::numeric_limits<double>::max();
      hi[d] = std::numeric_limits<double>::min();
    }
    for (int64_t i = 0; i < n; ++i) {
      double c[3];
      for (int d = 0; d < 3; ++d) c[d] = vx(i, 0)[d];
      for (int k = 0; k < nv; ++k)
        for (int d = 0; d < 3; ++d) c[d] += vx(i, k)[d];
      for (int d = 0; d < 3; ++d) {
        c[d] /= double(nv);
        ctr[3 * size_t(i) + d] = c[d];
        lo[d] = std::min(c[d], lo[d]);
        hi[d] = std::max(c[d], hi[d]);
      }
    }
    // normalised to [0, 1) as the tool does, then a Hilbert index with the
    // tool's 64 / n_dims bits per axis (32 in 2D, 21 in 3D)
    const int bits = m.n_dims == 2 ? 32 : 21;
    std::vector<uint64_t> key(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
      uint32_t q[3] = {0, 0, 0};
      for (int d = 0; d < m.n_dims; ++d) {
        const double t = (ctr[3 * size_t(i) + d] - lo[d]) / (hi[d] - lo[d] + 1e-10 * hi[d]);
        const double s = std::ldexp(std::min(std::max(t, 0.0), 1.0), bits);
        q[d] = s >= std::ldexp(1.0, bits) ? uint32_t((uint64_t(1) << bits) - 1) : uint32_t(s);
      }
      key[size_t(i)] = m.n_dims == 2 ? zfv::hilbert2(q[0], q[1], bits) : zfv::hilbert3(q[0], q[1], q[2], bits);
    }
    // sigma sorted by index; ties keep the file's order (a stable sort, so
    // the result does not depend on the sort implementation)
    std::vector<int64_t> sigma(static_cast<size_t>(n));
    std::iota(sigma.begin(), sigma.end(), int64_t(0));
    std::stable_sort(sigma.begin(), sigma.end(), [&](int64_t a, int64_t b) { return key[size_t(a)] < key[size_t(b)]; });
    std::vector<uint64_t> vidx(m.vidx.size());
    for (int64_t k = 0; k < n; ++k)
      for (int j = 0; j < nv; ++j) vidx[size_t(k) * nv + j] = m.vidx[size_t(sigma[size_t(k)]) * nv + j];
    m.vidx.swap(vidx);
    const std::string tmp = std::string(path) + "_";
    write_msh(tmp, m);
    if (std::rename(tmp.c_str(), path) != 0) throw ZfvError(ZFV_ERR_INVALID, "cannot rename " + tmp);
    if (order)
      for (int64_t k = 0; k < n; ++k) order[k] = sigma[size_t(k)];
  });
}

static void write_cvars(hid_t f, int64_t n_cells, const double *cvars) {
  std::vector<double> comp(static_cast<size_t>(n_cells));
  for (int v = 0; v < 5; ++v) {
    for (int64_t i = 0; i < n_cells; ++i) comp[size_t(i)] = cvars[5 * size_t(i) + v];
    write_array(f, kLabels[v], H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, comp.data(), {hsize_t(n_cells)});
  }
  const uint64_t n_avars = 0;
  write_array(f, "n_avars", H5T_STD_U64LE, H5T_NATIVE_UINT64, &n_avars, {});
}

int zfv_io_write_snapshot(const char *path, int64_t n_cells, const double *cvars, double t, int64_t n_steps) {
  return guarded([&] {
    if (n_cells < 0 || n_steps < 0) throw ZfvError(ZFV_ERR_INVALID, "negative sizes");
    H5Id f(create_file(path), H5Fclose);
    write_array(f.id, "time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &t, {});
    const uint64_t ns = uint64_t(n_steps);
    write_array(f.id, "n_steps", H5T_STD_U64LE, H5T_NATIVE_UINT64, &ns, {});
    write_cvars(f.id, n_cells, cvars);
  });
}

int zfv_io_write_steady_state(const char *path, int64_t n_cells, const double *cvars) {
  return guarded([&] {
    if (n_cells < 0) throw ZfvError(ZFV_ERR_INVALID, "negative sizes");
    H5Id f(create_file(path), H5Fclose);
    write_cvars(f.id, n_cells, cvars);
  });
}

int zfv_io_snapshot_cells(const char *path, int64_t *n_cells) {
  return guarded([&] {
    H5Id f(open_read(path), H5Fclose);
    std::vector<hsize_t> d = dims_of(f.id, kLabels[0]);
    if (d.size() != 1) throw ZfvError(ZFV_ERR_INVALID, "'rho' is not one-dimensional");
    *n_cells = int64_t(d[0]);
  });
}

int zfv_io_read_snapshot(const char *path, int64_t n_cells, double *cvars, double *t, int64_t *n_steps) {
  return guarded([&] {
    H5Id f(open_read(path), H5Fclose);
    if (t) *t = read_scalar<double>(f.id, "time", H5T_NATIVE_DOUBLE);
    if (n_steps) *n_steps = read_scalar<int64_t>(f.id, "n_steps", H5T_NATIVE_INT64);
    for (int v = 0; v < 5; ++v) {
      std::vector<hsize_t> d;
      std::vector<double> comp = read_all<double>(f.id, kLabels[v], H5T_NATIVE_DOUBLE, &d);
      if (d.size() != 1 || int64_t(d[0]) != n_cells)
        throw ZfvError(ZFV_ERR_INVALID, std::string("'") + kLabels[v] + "' does not hold n_cells values");
      for (int64_t i = 0; i < n_cells; ++i) cvars[5 * size_t(i) + v] = comp[size_t(i)];
    }
    const int64_t n_avars = read_scalar<int64_t>(f.id, "n_avars", H5T_NATIVE_INT64);
    if (n_avars != 0) throw ZfvError(ZFV_ERR_INVALID, "tracers (n_avars > 0) are not carried");
  });
}

}  // extern "C"
