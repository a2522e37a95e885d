// Error state of the C ABI (thread-local last-error string).
#include <hip/hip_runtime.h>

#include <string>

#include "common.h"
#include "xfe_hip.h"

namespace xfe {

namespace {
thread_local std::string tls_error;
}

void set_error(const std::string& msg) { tls_error = msg; }
void clear_error() { tls_error.clear(); }

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
    std::string m = std::string("HIP error '") + hipGetErrorString(e) + "' in " + what + " at " +
                    file + ":" + std::to_string(line);
    throw Error(m);
}

}  // namespace xfe

extern "C" const char* xfe_last_error(void) { return xfe::tls_error.c_str(); }
