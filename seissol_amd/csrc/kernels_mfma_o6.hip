// Matrix-core element-update kernels of ORDER 6.
#include "mfma_kernels.h"

template struct sdg::mfma::MfmaOrder<6>;
