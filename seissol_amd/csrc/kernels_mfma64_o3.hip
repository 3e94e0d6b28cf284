// FP64 matrix-core local pass of ORDER 3.
#include "mfma64_kernels.h"

template struct sdg::mfma64::MfmaOrder64<3>;
