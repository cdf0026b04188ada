// Internal helpers shared by the libtt.so translation units (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "tt.h"

namespace tt {

// thread-local message for tt_last_error()
void set_error(const char *fmt, ...);

inline int hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return TT_ERR_HIP;
    }
    return TT_OK;
}

#define TT_REQUIRE(cond, ...)                                                                   \
    do {                                                                                        \
        if (!(cond)) {                                                                          \
            ::tt::set_error(__VA_ARGS__);                                                       \
            return TT_ERR_ARG;                                                                  \
        }                                                                                       \
    } while (0)

#define TT_HIP(call, what)                                                                      \
    do {                                                                                        \
        hipError_t _e = (call);                                                                 \
        if (_e != hipSuccess) return ::tt::hip_check(_e, what);                                 \
    } while (0)

#define TT_LAUNCH_CHECK(what)                                                                   \
    do {                                                                                        \
        hipError_t _e = hipGetLastError();                                                      \
        if (_e != hipSuccess) return ::tt::hip_check(_e, what);                                 \
    } while (0)

constexpr double kC = 299792458.0;  // scipy.constants.c (particle_tracker.py:119)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Stream ids share the numpy twin's
// values (oracle/philox.py).
// ---------------------------------------------------------------------------
constexpr uint32_t kStreamField = 0;
constexpr uint32_t kStreamBeam = 1;

struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c.x;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

// ((x >> 8) + 0.5) * 2^-24: exact in float and double
__device__ __forceinline__ float u01f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * 5.9604644775390625e-8f; }
__device__ __forceinline__ double u01d(uint32_t x) { return ((double)(x >> 8) + 0.5) * 5.9604644775390625e-8; }

// Workgroups are dealt to the 8 XCDs in turn (block b to XCD b % 8), and each
// XCD has its own L2.  Block b takes work block xcd_block(b, nb): every XCD
// works through one contiguous eighth of the range, so with rays ordered by
// entry column its L2 sees neighbouring columns only.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb / 8, r = nb % 8, x = b % 8;
    return x * q + min(x, r) + b / 8;
}

// W - 1 (W = vz / c) at or below which the z-march hands a ray over to the
// time integrator (the rule is stated with the tracer)
constexpr float kWdMin = -0.03f;

}  // namespace tt
