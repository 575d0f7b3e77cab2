// Row-pass instantiations of the fused deconvolution for the less common
// row lengths (up to 2048 and above 12288 complex samples).
#include "decon_kernels.hpp"

namespace wct {
WCT_INSTANTIATE_ROW(4)
WCT_INSTANTIATE_ROW(8)
WCT_INSTANTIATE_ROW(32)
}  // namespace wct
