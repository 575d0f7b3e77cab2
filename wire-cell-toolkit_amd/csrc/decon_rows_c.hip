// Row-pass instantiations of the fused deconvolution: rows of up to 4096
// and 5120 complex samples (8000- and 10000-tick frames).
#include "decon_kernels.hpp"

namespace wct {
WCT_INSTANTIATE_ROW(16)
WCT_INSTANTIATE_ROW(20)
}  // namespace wct
