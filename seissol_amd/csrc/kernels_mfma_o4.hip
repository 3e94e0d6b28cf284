// Matrix-core element-update kernels of ORDER 4.
#include "mfma_kernels.h"

template struct sdg::mfma::MfmaOrder<4>;
