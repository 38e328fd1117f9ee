// Tile kernels for D = 3 (see kernels.hpp).
#include "launch.hpp"
#define TILE_D 3
#include "tile_launch.inc"

namespace nfftm {
hipError_t launch_tile_d3(bool spread, bool dbl, int W, int NT, const KArgs& a, size_t lds,
                          hipStream_t s, const void* in, void* out) {
  return dbl ? tile_w<double>(spread, W, NT, a, lds, s, in, out)
             : tile_w<float>(spread, W, NT, a, lds, s, in, out);
}
}  // namespace nfftm
