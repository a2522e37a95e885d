// Shared host helpers of the HIP library: error state, checked calls and
// grow-only device buffers.
#ifndef XFE_COMMON_H
#define XFE_COMMON_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>

namespace xfe {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

void set_error(const std::string& msg);
void clear_error();

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define XFE_CHECK(expr)                                                   \
    do {                                                                  \
        const hipError_t xfe_e_ = (expr);                                 \
        if (xfe_e_ != hipSuccess) ::xfe::throw_hip(xfe_e_, #expr, __FILE__, __LINE__); \
    } while (0)

// Grow-only device allocation.  Buffers of the per-thread workspaces live
// until xfe_release_workspace() or process exit (the destructor does not call
// into the HIP runtime, which may already be torn down at exit).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    void ensure(size_t count) {
        if (count == 0) count = 1;
        if (count <= cap) return;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        XFE_CHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        cap = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace xfe

#endif
