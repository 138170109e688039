// Shared host-side types of libzfv_hip.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/zfv.h"

namespace zfv {

// zfv_create in two halves (kernels.hip), for the serving runtime's setup
// workers: create_host builds the setup image with no device call,
// create_device allocates, uploads and runs the setup kernels; on failure
// the caller deletes the handle with zfv_destroy
int create_host(const zfv_mesh *mesh, const zfv_params *params, zfv_handle **out);
int create_device(zfv_handle *h);

struct ZfvError : std::runtime_error {
  int code;
  ZfvError(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
};

struct MeshBuf {
  int n_dims = 2;
  int n_cells = 0;
  std::vector<int32_t> vertex_ids;
  std::vector<double> corners;
  std::vector<uint8_t> ghost;
};

MeshBuf *mesh_lattice2d(int n, double jitter, uint64_t seed, bool px, bool py, int ghost_layers);
MeshBuf *mesh_ball3d(int n, double radius, double ghost_radius, double jitter, uint64_t seed);

}  // namespace zfv
