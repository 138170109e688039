/*
 * zfv_io.h — C-ABI of libzfv_io.so, the host-side file formats around the
 * rate-of-change engine: the reference's mesh files (.msh.h5) and its
 * snapshot / restart files, read and written with HDF5.  A separate library
 * from libzfv_hip.so so that the engine itself needs no HDF5.
 *
 * Formats (all datasets at the file's root):
 *   mesh (.msh.h5; load_grid_gmsh_h5, src/zisa/grid/grid.cpp:889-901):
 *     n_dims          scalar integer, 2 or 3
 *     vertices        n_vertices x 3 doubles
 *     vertex_indices  n_cells x (n_dims + 1) integers (any integer type on
 *                     read; written as 64-bit unsigned, the reference's int_t)
 *   snapshot (save_state, src/zisa/model/models.cpp:7-16, and
 *   AllVariables::save, src/zisa/model/all_variables.cpp:74-90):
 *     time            scalar double
 *     n_steps         scalar integer
 *     rho mv1 mv2 mv3 E   n_cells doubles each (euler_var_t::labels,
 *                     src/zisa/model/euler_variables.cpp:9-27)
 *     n_avars         scalar integer (0: no tracers), then mq0, mq1, ...
 *   steady state (DumpSnapshot::do_steady_state,
 *   include/zisa/io/dump_snapshot_impl.hpp:35-39): the snapshot without
 *   time and n_steps.
 * Errors: negative return codes (ZFV_ERR_* of zfv.h) and zfv_io_last_error().
 */
#ifndef ZFV_IO_H
#define ZFV_IO_H

#include <stdint.h>

#include "zfv.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *zfv_io_last_error(void);

/* Reads a .msh.h5 mesh (load_grid_gmsh_h5) into a mesh buffer with per-cell
 * corners and no ghost cells (the reference marks ghosts per experiment,
 * mask_ghost_cells, src/zisa/grid/grid.cpp:1122-1136); view it with
 * zfv_io_mesh_view, free it with zfv_io_mesh_free. */
int zfv_io_read_msh_h5(const char *path, zfv_mesh_buf **out);
int zfv_io_mesh_view(const zfv_mesh_buf *buf, zfv_mesh *view);
int zfv_io_mesh_free(zfv_mesh_buf *buf);

/* Writes a .msh.h5 mesh from shared vertices and per-cell vertex indices. */
int zfv_io_write_msh_h5(const char *path, int32_t n_dims, int64_t n_vertices, const double *vertices,
                        int64_t n_cells, const int64_t *vertex_indices);

/* The reference's grid renumbering tool (renumber_grid, src/renumber_grid.cpp:46-135):
 * reads path, orders the cells along a Hilbert curve of their (reference-
 * defined) centres normalised to the bounding box, writes the permuted
 * vertex_indices with the same vertices and n_dims to path + "_", and
 * renames it over path.  The 3D volume check of the tool (sanity_check)
 * rejects meshes whose largest tetrahedron is more than 100 times the
 * smallest.  If order is not NULL it receives the permutation (new cell k
 * was old cell order[k]). */
int zfv_io_renumber_msh_h5(const char *path, int64_t *order);

/* Snapshot of the conserved variables (n_cells x 5 row-major, the caller's
 * AllVariables::cvars) at time t after n_steps steps, no tracers. */
int zfv_io_write_snapshot(const char *path, int64_t n_cells, const double *cvars, double t, int64_t n_steps);

/* Steady state file (cvars only, no time). */
int zfv_io_write_steady_state(const char *path, int64_t n_cells, const double *cvars);

/* Restart (load_full_state, src/zisa/model/load_full_state.cpp:8-13):
 * reads the conserved variables of a snapshot (n_cells must match the file)
 * and its time and step count (either may be NULL; a steady state file has
 * neither, and asking for them is an error). */
int zfv_io_read_snapshot(const char *path, int64_t n_cells, double *cvars, double *t, int64_t *n_steps);

/* Number of cells of a snapshot or steady state file (length of "rho"). */
int zfv_io_snapshot_cells(const char *path, int64_t *n_cells);

#ifdef __cplusplus
}
#endif

#endif /* ZFV_IO_H */
