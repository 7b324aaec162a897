// Host-side helpers for the C-ABI: error codes, last-error string, launch checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/sgmc.h"

namespace sgmc {

void set_error(const char* fmt, ...);

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return SGMC_E_HIP;
  }
  return SGMC_OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace sgmc

#define SGMC_REQUIRE(cond, code, ...)   \
  do {                                  \
    if (!(cond)) {                      \
      ::sgmc::set_error(__VA_ARGS__);   \
      return (code);                    \
    }                                   \
  } while (0)
