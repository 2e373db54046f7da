// ctlab_map.hip -- the elementwise and stencil parts of a MAP iteration with a
// Gaussian likelihood and the quadratic smoothness prior, fused around the
// projector kernels.
//
// The reference minimises -(log N(y; A x, var) + prior_weight * log p(x)) with
// log p(x) = -(w/2) sum_d ||diff_d(x)||^2 (random_variable/gaussian.py:54-63,
// random_variable/image_energies.py:18-29) by preconditioned gradient steps
// x <- x - lr * P(P(g)) (bayesian_estimation/map_estimator.py:113-128).  The
// gradient is g = A^T (A x - y) / var + beta * L x with beta = prior_weight * w
// and L = D^T D the graph Laplacian of the grid with free (non-periodic) ends.
// Here one iteration is: forward projection -> k_residual (r = (A x - y)/var,
// data energy) -> back projection of r -> k_prior_update (prior gradient,
// preconditioned step, prior energy).  The energies go to a double pair on the
// device so that no iteration synchronises with the host; every kernel writes
// all of its output, so the iteration can be captured in a graph and replayed.
#include "ctlab_common.h"

namespace ctlab {

// Block sum of a double into one atomic per block.
__device__ __forceinline__ void block_add(double v, double* __restrict__ dst, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < (int)(blockDim.x / 64); ++k) s += red[k];
        atomicAdd(dst, s);
    }
}

__global__ __launch_bounds__(kBlock) void k_zero_f64(double* __restrict__ p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_residual(const float* __restrict__ ax, const float* __restrict__ y,
                                                     int64_t n, float inv_var, float* __restrict__ r,
                                                     double* __restrict__ energy) {
    __shared__ double red[kBlock / 64];
    double e = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float d = ax[i] - y[i];
        r[i] = d * inv_var;
        e += 0.5 * (double)d * (double)d * (double)inv_var;
    }
    if (energy) block_add(e, energy, red);
}

// One thread per voxel of [n0, n1, n2] (n2 fastest; 2D grids use n0 = 1).
__global__ __launch_bounds__(kBlock) void k_prior_update(const float* __restrict__ x, const float* __restrict__ bp,
                                                         const float* __restrict__ d2, float beta, float lr,
                                                         int n0, int n1, int n2, float* __restrict__ x_out,
                                                         double* __restrict__ energy) {
    __shared__ double red[kBlock / 64];
    const int64_t total = (int64_t)n0 * n1 * n2;
    const int64_t s1 = n2, s0 = (int64_t)n1 * n2;
    double e = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int k = (int)(i % n2), j = (int)((i / n2) % n1), a = (int)(i / s0);
        const float xi = x[i];
        float lap = 0.0f;
        if (a > 0) lap += xi - x[i - s0];
        if (a + 1 < n0) {
            const float dd = x[i + s0] - xi;
            lap -= dd;
            e += (double)dd * dd;
        }
        if (j > 0) lap += xi - x[i - s1];
        if (j + 1 < n1) {
            const float dd = x[i + s1] - xi;
            lap -= dd;
            e += (double)dd * dd;
        }
        if (k > 0) lap += xi - x[i - 1];
        if (k + 1 < n2) {
            const float dd = x[i + 1] - xi;
            lap -= dd;
            e += (double)dd * dd;
        }
        const float g = bp[i] + beta * lap;
        const float p2 = d2 ? d2[i] : 1.0f;
        x_out[i] = xi - lr * (p2 * g);
    }
    if (energy) block_add(0.5 * (double)beta * e, energy, red);
}

// Diagonal majoriser of the prior Hessian beta * L: 2 beta * (number of grid
// neighbours), the row sums of |beta L|.
__global__ __launch_bounds__(kBlock) void k_sqs_weights(const float* __restrict__ ata1, float inv_var, float beta,
                                                        int n0, int n1, int n2, float* __restrict__ d2) {
    const int64_t total = (int64_t)n0 * n1 * n2;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int k = (int)(i % n2), j = (int)((i / n2) % n1), a = (int)(i / ((int64_t)n1 * n2));
        const int deg = (a > 0) + (a + 1 < n0) + (j > 0) + (j + 1 < n1) + (k > 0) + (k + 1 < n2);
        const float h = ata1[i] * inv_var + 2.0f * beta * (float)deg;
        d2[i] = h > 0.0f ? 1.0f / h : 0.0f;
    }
}

static unsigned blocks_for(int64_t n) {
    int64_t b = grid_for(n);
    return (unsigned)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int ctlab_map_residual(const float* ax, const float* y, int64_t n, float inv_var, float* r,
                                  double* energy, ctlab_stream_t stream) {
    const char* fn = "map_residual";
    CTLAB_REQUIRE(n >= 0, "%s: n < 0", fn);
    if (n == 0) return CTLAB_OK;
    CTLAB_REQUIRE(ax && y && r, "%s: null pointer", fn);
    hipLaunchKernelGGL(k_residual, dim3(blocks_for(n)), dim3(kBlock), 0, as_stream(stream), ax, y, n, inv_var, r,
                       energy);
    return check_launch(fn);
}

extern "C" int ctlab_map_prior_update(const float* x, const float* bp, const float* d2, float beta, float lr,
                                      int64_t n0, int64_t n1, int64_t n2, float* x_out, double* energy,
                                      ctlab_stream_t stream) {
    const char* fn = "map_prior_update";
    CTLAB_REQUIRE(n0 >= 1 && n1 >= 1 && n2 >= 1 && n0 * n1 * n2 <= kMaxVoxels, "%s: bad grid", fn);
    CTLAB_REQUIRE(x && bp && x_out && x != x_out, "%s: null pointer or in-place update", fn);
    hipLaunchKernelGGL(k_prior_update, dim3(blocks_for(n0 * n1 * n2)), dim3(kBlock), 0, as_stream(stream), x, bp,
                       d2, beta, lr, (int)n0, (int)n1, (int)n2, x_out, energy);
    return check_launch(fn);
}

extern "C" int ctlab_map_sqs_weights(const float* ata1, float inv_var, float beta, int64_t n0, int64_t n1,
                                     int64_t n2, float* d2, ctlab_stream_t stream) {
    const char* fn = "map_sqs_weights";
    CTLAB_REQUIRE(n0 >= 1 && n1 >= 1 && n2 >= 1 && n0 * n1 * n2 <= kMaxVoxels, "%s: bad grid", fn);
    CTLAB_REQUIRE(ata1 && d2, "%s: null pointer", fn);
    hipLaunchKernelGGL(k_sqs_weights, dim3(blocks_for(n0 * n1 * n2)), dim3(kBlock), 0, as_stream(stream), ata1,
                       inv_var, beta, (int)n0, (int)n1, (int)n2, d2);
    return check_launch(fn);
}

extern "C" int ctlab_zero_f64(double* p, int64_t n, ctlab_stream_t stream) {
    const char* fn = "zero_f64";
    CTLAB_REQUIRE(n >= 0 && n <= kBlock, "%s: n out of range", fn);
    if (n == 0) return CTLAB_OK;
    CTLAB_REQUIRE(p, "%s: null pointer", fn);
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(kBlock), 0, as_stream(stream), p, (int)n);
    return check_launch(fn);
}
