// Wire-pass instantiations of the fused deconvolution: 2- and 4-bin tiles.
#include "decon_kernels.hpp"

namespace wct {
WCT_INSTANTIATE_WIRE(2)
WCT_INSTANTIATE_WIRE(4)
}  // namespace wct
