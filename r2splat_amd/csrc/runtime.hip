// Error reporting and small utility kernels shared by every entry point.
#include "common.hpp"

#include <cstdarg>

namespace r2s {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

int check_launch(const char* where) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: kernel launch failed: %s", where, hipGetErrorString(e));
        return R2S_ERR_HIP;
    }
    return R2S_OK;
}

}  // namespace r2s

extern "C" int r2s_abi_version(void) { return 4; }

extern "C" const char* r2s_last_error(void) { return r2s::g_error; }
