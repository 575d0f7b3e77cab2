// Row-pass instantiations of the fused deconvolution: rows of up to 3072
// complex samples (6000-tick frames).
#include "decon_kernels.hpp"

namespace wct {
WCT_INSTANTIATE_ROW(12)
}  // namespace wct
