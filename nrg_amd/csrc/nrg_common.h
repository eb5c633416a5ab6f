// nrg_common.h -- error reporting shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/nrg_chem.h"

namespace nrg {

// Records a formatted message for nrg_chem_last_error() and returns status.
__attribute__((visibility("hidden"))) int set_error(int status, const char *fmt, ...);

}  // namespace nrg

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return ::nrg::set_error(NRG_CHEM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
