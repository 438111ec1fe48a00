// Shared helpers of the gfx950 kernels: error state of the C ABI and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/tenet_amd.h"

namespace tn {

void set_error(const std::string& msg);

inline int fail(const std::string& msg) {
  set_error(msg);
  return 1;
}

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

}  // namespace tn
