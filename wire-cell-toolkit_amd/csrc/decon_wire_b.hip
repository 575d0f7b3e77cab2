// Wire-pass instantiations of the fused deconvolution: 8- and 16-bin tiles.
#include "decon_kernels.hpp"

namespace wct {
WCT_INSTANTIATE_WIRE(8)
WCT_INSTANTIATE_WIRE(16)
}  // namespace wct
