// ctlab_common.hip -- error state, fill kernel and library identification.
#include "ctlab_common.h"

namespace ctlab {

static thread_local std::string g_err;

int fail(int status, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return status;
}

const char* last_error() { return g_err.c_str(); }

__global__ __launch_bounds__(kBlock) void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

int fill_f32(float* p, int64_t n, float value, hipStream_t stream) {
    if (n <= 0) return CTLAB_OK;
    // 256 CUs x 8 blocks covers the chip; the grid-stride loop does the rest.
    int64_t blocks = grid_for(n);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)blocks), dim3(kBlock), 0, stream, p, n, value);
    return check_launch("fill_f32");
}

}  // namespace ctlab

extern "C" int ctlab_version(void) { return 1; }

extern "C" const char* ctlab_last_error(void) { return ctlab::last_error(); }
