// ctlab_common.h -- shared host-side helpers of libctlab_hip.so: status codes,
// the thread-local error message, argument checks and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/ctlab_hip.h"

namespace ctlab {

// Set the thread-local message and return the status (used as `return fail(...)`).
int fail(int status, const char* fmt, ...);
const char* last_error();

inline hipStream_t as_stream(ctlab_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Launch-error check after a kernel launch.
inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(CTLAB_EHIP, "%s: launch failed: %s", what, hipGetErrorString(e));
    return CTLAB_OK;
}

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

// Host-side limits: flat voxel indices are 32-bit, ray indices 64-bit.
constexpr int64_t kMaxVoxels = 0x7fffffffLL;
constexpr int64_t kMaxGridX = 0x7fffffffLL;

// The traversal forms flat voxel indices with 24-bit multiply-adds: every grid
// dimension, and n_y*n_z for 3D grids, must be below 2^24.
inline int check_index_limits(const char* fn, const int64_t* n, int nd) {
    for (int a = 0; a < nd; ++a)
        if (n[a] >= (1LL << 24)) return fail(CTLAB_EINVAL, "%s: grid dimension %d >= 2^24", fn, a);
    if (nd == 3 && n[1] * n[2] >= (1LL << 24))
        return fail(CTLAB_EINVAL, "%s: n_y*n_z = %lld must be below 2^24", fn, (long long)(n[1] * n[2]));
    return CTLAB_OK;
}

// Fills n floats with value on the stream (a kernel, never hipMemsetAsync, so
// that the call is safe inside a captured graph).
int fill_f32(float* p, int64_t n, float value, hipStream_t stream);

// Zeroes the images of a ctlab_scan2d table (when clear_images) and the work
// counter of a persistent multi-scan launch.
int batch2d_clear(const char* fn, const ctlab_scan2d* scans, int64_t n_scan, bool clear_images,
                  unsigned long long* counter, hipStream_t stream);

// 2D back projection of one scan (batch elements along the sinogram's first
// axis) by the LDS tile kernel; the caller has checked the arguments.
int back_project_2d_tiles(const char* fn, const float* sino, int64_t batch, int64_t n_ray, const float* src,
                          const float* dst, const float* minv, const float* b, const int64_t* n, int mode,
                          int zero_rule, int accumulate, float* out, hipStream_t stream);

}  // namespace ctlab

#define CTLAB_REQUIRE(cond, ...)                                   \
    do {                                                           \
        if (!(cond)) return ::ctlab::fail(CTLAB_EINVAL, __VA_ARGS__); \
    } while (0)
