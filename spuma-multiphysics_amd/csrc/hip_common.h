// SPDX-License-Identifier: GPL-3.0-or-later
// Shared host helpers for the HIP library: error checks, wave reductions.
#ifndef REACTIVE_AMD_HIP_COMMON_H
#define REACTIVE_AMD_HIP_COMMON_H
#include <hip/hip_runtime.h>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace ramd {

inline void hipCheck(hipError_t status, const char* what) {
    if (status != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}
#define RAMD_HIP(call) ::ramd::hipCheck((call), #call)

inline void require(bool ok, const std::string& message) {
    if (!ok) throw std::runtime_error(message);
}

inline void writeError(char* error, size_t size, const char* message) {
    if (error && size) std::snprintf(error, size, "%s", message);
}

// Sum over the 64 lanes of a wavefront (every lane gets the result).
__device__ inline uint32_t waveSum(uint32_t v) {
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    return v;
}
__device__ inline double waveMin(double v) {
    for (int offset = 32; offset > 0; offset >>= 1) {
        const double o = __shfl_xor(v, offset, 64);
        v = o < v ? o : v;
    }
    return v;
}

// Device that the calling host thread has current.
inline int currentDevice() {
    int d = 0;
    RAMD_HIP(hipGetDevice(&d));
    return d;
}

}  // namespace ramd
#endif
