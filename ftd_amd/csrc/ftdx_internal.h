// ftdx_internal.h -- error helpers shared by the libftdx.so translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace ftdx {

// Record `msg` as this thread's last error and return `code`.
int fail(int code, const std::string& msg);
// Record a HIP error (name and text) with context `what`; returns FTDX_ERR_HIP.
int hip_fail(hipError_t e, const char* what);

}  // namespace ftdx
