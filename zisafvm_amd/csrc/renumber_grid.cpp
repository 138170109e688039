// zfv_renumber_grid [--grid] <mesh.msh.h5>...: the reference's grid
// renumbering tool (src/renumber_grid.cpp:145-185 runs renumber_grid on the
// file given by --grid): each file's cells are put in Hilbert order in place.
#include <cstdio>
#include <cstring>

#include "../../include/zfv_io.h"

int main(int argc, char **argv) {
  int done = 0;
  for (int k = 1; k < argc; ++k) {
    if (std::strcmp(argv[k], "--grid") == 0) continue;
    if (std::strcmp(argv[k], "-h") == 0 || std::strcmp(argv[k], "--help") == 0) break;
    if (zfv_io_renumber_msh_h5(argv[k], nullptr) != ZFV_OK) {
      std::fprintf(stderr, "%s: %s\n", argv[k], zfv_io_last_error());
      return 1;
    }
    std::printf("renumbered %s\n", argv[k]);
    ++done;
  }
  if (done == 0) {
    std::fprintf(stderr, "usage: %s [--grid] <grid.msh.h5>...\n", argv[0]);
    return 2;
  }
  return 0;
}
