// Row-pass instantiations of the fused deconvolution: rows of up to 6144
// complex samples.
#include "decon_kernels.hpp"

namespace wct {
WCT_INSTANTIATE_ROW(24)
}  // namespace wct
