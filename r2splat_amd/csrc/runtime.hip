// Error reporting and small utility kernels shared by every entry point.
#include "common.hpp"

#include <cstdarg>
#include <limits>

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

extern "C" int r2s_abi_version(void) { return 5; }

extern "C" const char* r2s_last_error(void) { return r2s::g_error; }

// ---------------------------------------------------------------------------
// Value checks: one workgroup row (blockIdx.y) per check, a grid-stride scan
// of its buffer, and flags[check] = 1 from any wave that saw a bad element
// (every such store writes the same value, so their order does not matter).
// ---------------------------------------------------------------------------
namespace r2s {

constexpr int kCheckThreads = 256;
constexpr uint32_t kCheckBlocks = 4096;  // enough workgroups per check to stream a large buffer at HBM rate

struct ValueCheckTable {
    r2s_value_check c[R2S_MAX_VALUE_CHECKS];
};

template <typename S>
__device__ __forceinline__ bool bad_element(S x, int kind) {
    if (kind == R2S_CHECK_NONFINITE) return !(fabs(x) <= std::numeric_limits<S>::max());
    if (kind == R2S_CHECK_NONPOSITIVE) return x <= S(0);
    return x < S(0);  // R2S_CHECK_NEGATIVE
}

template <typename S>
__device__ __forceinline__ bool scan_check(const r2s_value_check& c) {
    const S* p = static_cast<const S*>(c.data);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    bool bad = false;
    if (c.kind >= R2S_CHECK_ZERO_ROW) {
        const uint64_t rows = c.n / c.row;
        for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += stride) {
            const S* m = p + r * c.row;
            bool b;
            if (c.kind == R2S_CHECK_K_FOCAL) {
                b = m[0] <= S(0) || m[4] <= S(0);
            } else if (c.kind == R2S_CHECK_K_PINHOLE) {
                b = m[1] != S(0) || m[3] != S(0) || m[6] != S(0) || m[7] != S(0) || m[8] != S(1);
            } else if (c.kind == R2S_CHECK_VIEW_ROW) {
                b = m[12] != S(0) || m[13] != S(0) || m[14] != S(0) || m[15] != S(1);
            } else {
                b = true;
                for (uint32_t k = 0; k < c.row; ++k) b = b && m[k] == S(0);
            }
            bad = bad || b;
        }
    } else {
        for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < c.n; i += stride)
            bad = bad || bad_element(p[i], c.kind);
    }
    return bad;
}

__global__ void __launch_bounds__(kWave) value_checks_clear_kernel(uint32_t n_checks, uint32_t* __restrict__ flags) {
    if (threadIdx.x < n_checks) flags[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(kCheckThreads) value_checks_kernel(ValueCheckTable table,
                                                                     uint32_t* __restrict__ flags) {
    const r2s_value_check& c = table.c[blockIdx.y];
    const bool bad = c.dtype == R2S_F64 ? scan_check<double>(c) : scan_check<float>(c);
    if (__ballot(bad) != 0 && threadIdx.x % kWave == 0) flags[blockIdx.y] = 1u;
}

}  // namespace r2s

extern "C" int r2s_value_checks(const r2s_value_check* checks, uint32_t n_checks, uint32_t* flags,
                                r2s_stream_t stream) {
    using namespace r2s;
    if (n_checks == 0) return R2S_OK;
    R2S_REQUIRE(checks && flags, "checks and flags must be non-null");
    R2S_REQUIRE(n_checks <= R2S_MAX_VALUE_CHECKS, "at most %d value checks per call", R2S_MAX_VALUE_CHECKS);
    ValueCheckTable table = {};
    uint64_t most = 0;
    for (uint32_t i = 0; i < n_checks; ++i) {
        const r2s_value_check& c = checks[i];
        R2S_REQUIRE(c.dtype == R2S_F32 || c.dtype == R2S_F64, "check %u: dtype must be R2S_F32 or R2S_F64", i);
        R2S_REQUIRE(c.kind >= R2S_CHECK_NONFINITE && c.kind <= R2S_CHECK_VIEW_ROW, "check %u: unknown kind %d", i,
                    c.kind);
        R2S_REQUIRE(c.n == 0 || c.data, "check %u: data must be non-null", i);
        R2S_REQUIRE(c.kind < R2S_CHECK_ZERO_ROW || (c.row > 0 && c.n % c.row == 0),
                    "check %u: row must be positive and divide n", i);
        R2S_REQUIRE((c.kind != R2S_CHECK_K_FOCAL && c.kind != R2S_CHECK_K_PINHOLE) || c.row == 9,
                    "check %u: intrinsics rows hold 9 elements", i);
        R2S_REQUIRE(c.kind != R2S_CHECK_VIEW_ROW || c.row == 16, "check %u: viewmat rows hold 16 elements", i);
        table.c[i] = c;
        most = std::max<uint64_t>(most, c.kind >= R2S_CHECK_ZERO_ROW ? c.n / c.row : c.n);
    }
    value_checks_clear_kernel<<<1, kWave, 0, stream>>>(n_checks, flags);
    if (int rc = check_launch("value_checks_clear")) return rc;
    constexpr uint64_t kPerBlock = kCheckThreads * 4;
    const uint64_t want = std::max<uint64_t>(1, (most + kPerBlock - 1) / kPerBlock);
    const uint32_t blocks = uint32_t(std::min<uint64_t>(kCheckBlocks, want));
    value_checks_kernel<<<dim3(blocks, n_checks), kCheckThreads, 0, stream>>>(table, flags);
    return check_launch("value_checks");
}
