// Matrix-core element-update kernels of ORDER 5.
#include "mfma_kernels.h"

template struct sdg::mfma::MfmaOrder<5>;
