// Tile kernels for D = 2 (see kernels.hpp).
#include "launch.hpp"
#define TILE_D 2
#include "tile_launch.inc"

namespace nfftm {
hipError_t launch_tile_d2(bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds,
                          hipStream_t s, const void* in, void* out) {
  return dbl ? tile_w<double>(spread, W, NT, a, lds, s, in, out)
             : tile_w<float>(spread, W, NT, a, lds, s, in, out);
}
}  // namespace nfftm
