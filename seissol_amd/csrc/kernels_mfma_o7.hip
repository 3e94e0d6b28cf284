// Matrix-core element-update kernels of ORDER 7.
#include "mfma_kernels.h"

template struct sdg::mfma::MfmaOrder<7>;
