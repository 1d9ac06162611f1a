// Shared helpers of the libxentica_hip.so translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

namespace xen {
// Record ``msg`` as the calling thread's last error; returns ``code`` (or -1).
int fail(const std::string &msg, int code);
// Map a HIP status to 0 or a recorded error.
int check(hipError_t err, const char *what);
}  // namespace xen
