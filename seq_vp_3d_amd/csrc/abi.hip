// Library-level entry points: version and workspace sizes of every operation.
#include "abi_util.h"

extern "C" const char* seqvp_version(void) { return "seqvp-hip gfx950 1"; }

extern "C" size_t seqvp_workspace_bytes(int op, int64_t nx, int64_t ny, int64_t nz) {
  if (nx < 2 || ny < 2 || nz < 2) return 0;
  switch (op) {
    case 6:
      if (nx < 3 || ny < 3 || nz < 3) return 0;
      return seqvp_dst_workspace_bytes(nx - 2, ny - 2, nz - 2, 3);
    case 7:
      return seqvp_vec_workspace_bytes(3 * (nx - 2) * (ny - 2) * (nz - 2));
    default:
      return seqvp_q1_workspace_bytes(op, nx, ny, nz);
  }
}
