// Matrix-core element-update kernels of ORDER 3.
#include "mfma_kernels.h"

template struct sdg::mfma::MfmaOrder<3>;
