// ctlab_f64.hip -- on-the-fly projections with FP64 values over the FP32
// traversal, for derivative checks.
//
// The segments (voxel index and float32 length) are the traversal's, exactly
// as in the float32 kernels: the operator A is the same matrix.  Only the
// values are double: the forward sums vol[idx] * len in double (one fused
// multiply-add per segment, in segment order), and the back projection adds
// sino * len with double atomics.  The pair is then the exact transpose in
// double precision, so central differences of a loss through it agree with
// the hand-written adjoint to FP64 rounding (torch.autograd.gradcheck can run
// on it), which the float32 pair cannot show beyond float32 rounding.  One ray
// per lane in storage order: these kernels serve checks, not throughput.
#include "ctlab_common.h"
#include "siddon_traverse.h"

namespace ctlab {

template <int ND>
struct DimsF64 {
    int n[ND];
};

template <int ND>
__device__ __forceinline__ void load_ray_f64(const float* __restrict__ src, const float* __restrict__ dst, int64_t r,
                                             float* s, float* d) {
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        s[a] = src[r * ND + a];
        d[a] = dst[r * ND + a];
    }
}

template <int ND, int MODE>
__global__ __launch_bounds__(kBlock) void k_forward_f64(const double* __restrict__ vol, double* __restrict__ out,
                                                        int64_t n_ray, int64_t vol_elems, DimsF64<ND> dims,
                                                        const float* __restrict__ src, const float* __restrict__ dst,
                                                        const float* __restrict__ minv, const float* __restrict__ b,
                                                        int zero_rule) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ray) return;
    const int64_t bi = blockIdx.y;
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, zero_rule);
    float s[ND], d[ND];
    load_ray_f64<ND>(src, dst, r, s, d);
    const double* __restrict__ v = vol + bi * vol_elems;
    double acc = 0.0;
    auto gather = [&](uint32_t idx, float len) { acc = fma(v[idx], (double)len, acc); };
    if (g.diag)
        trace_ray<ND, MODE, true>(g, s, d, gather);
    else
        trace_ray<ND, MODE, false>(g, s, d, gather);
    out[bi * n_ray + r] = acc;
}

template <int ND, int MODE>
__global__ __launch_bounds__(kBlock) void k_back_f64(const double* __restrict__ sino, double* __restrict__ out,
                                                     int64_t n_ray, int64_t vol_elems, DimsF64<ND> dims,
                                                     const float* __restrict__ src, const float* __restrict__ dst,
                                                     const float* __restrict__ minv, const float* __restrict__ b,
                                                     int zero_rule) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ray) return;
    const int64_t bi = blockIdx.y;
    const double val = sino[bi * n_ray + r];
    if (val == 0.0) return;  // every contribution would be exactly zero
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, zero_rule);
    float s[ND], d[ND];
    load_ray_f64<ND>(src, dst, r, s, d);
    double* __restrict__ o = out + bi * vol_elems;
    auto scatter = [&](uint32_t idx, float len) { unsafeAtomicAdd(o + idx, val * (double)len); };
    if (g.diag)
        trace_ray<ND, MODE, true>(g, s, d, scatter);
    else
        trace_ray<ND, MODE, false>(g, s, d, scatter);
}

template <int ND>
static int check_f64(const char* fn, const int64_t* n, int64_t batch, int64_t n_ray, const void* src,
                     const void* dst, const void* minv, const void* b, int mode, const void* in, const void* out,
                     int64_t* vox) {
    int64_t v = 1;
    for (int a = 0; a < ND; ++a) {
        CTLAB_REQUIRE(n[a] >= 1 && n[a] <= 0x7fffffffLL, "%s: grid dimension %d is %lld", fn, a, (long long)n[a]);
        v *= n[a];
        CTLAB_REQUIRE(v <= kMaxVoxels, "%s: grid has more than 2^31-1 voxels", fn);
    }
    int st = check_index_limits(fn, n, ND);
    if (st) return st;
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range [1, 65535]", fn, (long long)batch);
    CTLAB_REQUIRE(n_ray >= 0 && grid_for(n_ray) <= kMaxGridX, "%s: bad n_ray %lld", fn, (long long)n_ray);
    CTLAB_REQUIRE(mode == CTLAB_MODE_TORCH || mode == CTLAB_MODE_OTF, "%s: unknown mode %d", fn, mode);
    CTLAB_REQUIRE(out, "%s: out is null", fn);
    if (n_ray > 0) CTLAB_REQUIRE(src && dst && minv && b && in, "%s: null pointer", fn);
    *vox = v;
    return CTLAB_OK;
}

template <int ND>
static int forward_f64(const char* fn, const double* vol, int64_t batch, const int64_t* n, const float* src,
                       const float* dst, int64_t n_ray, const float* minv, const float* b, int mode, int zero_rule,
                       double* out, hipStream_t stream) {
    int64_t V;
    int st = check_f64<ND>(fn, n, batch, n_ray, src, dst, minv, b, mode, vol, out, &V);
    if (st || n_ray == 0) return st;
    DimsF64<ND> dims;
    for (int a = 0; a < ND; ++a) dims.n[a] = (int)n[a];
    dim3 grid(grid_for(n_ray), (unsigned)batch);
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL((k_forward_f64<ND, MODE_TORCH>), grid, dim3(kBlock), 0, stream, vol, out, n_ray, V, dims,
                           src, dst, minv, b, zero_rule);
    else
        hipLaunchKernelGGL((k_forward_f64<ND, MODE_OTF>), grid, dim3(kBlock), 0, stream, vol, out, n_ray, V, dims,
                           src, dst, minv, b, zero_rule);
    return check_launch(fn);
}

template <int ND>
static int back_f64(const char* fn, const double* sino, int64_t batch, int64_t n_ray, const float* src,
                    const float* dst, const float* minv, const float* b, const int64_t* n, int mode, int zero_rule,
                    int accumulate, double* out, hipStream_t stream) {
    int64_t V;
    int st = check_f64<ND>(fn, n, batch, n_ray, src, dst, minv, b, mode, sino, out, &V);
    if (st) return st;
    if (!accumulate) {  // 0.0 is all-zero bits in double as in float
        st = fill_f32(reinterpret_cast<float*>(out), 2 * V * batch, 0.0f, stream);
        if (st) return st;
    }
    if (n_ray == 0) return CTLAB_OK;
    DimsF64<ND> dims;
    for (int a = 0; a < ND; ++a) dims.n[a] = (int)n[a];
    dim3 grid(grid_for(n_ray), (unsigned)batch);
    if (mode == CTLAB_MODE_TORCH)
        hipLaunchKernelGGL((k_back_f64<ND, MODE_TORCH>), grid, dim3(kBlock), 0, stream, sino, out, n_ray, V, dims,
                           src, dst, minv, b, zero_rule);
    else
        hipLaunchKernelGGL((k_back_f64<ND, MODE_OTF>), grid, dim3(kBlock), 0, stream, sino, out, n_ray, V, dims,
                           src, dst, minv, b, zero_rule);
    return check_launch(fn);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int ctlab_forward_project_3d_otf_f64(const double* volume, int64_t batch, int64_t n_x, int64_t n_y,
                                                int64_t n_z, const float* src, const float* dst, int64_t n_ray,
                                                const float* M_inv, const float* b, int mode, double* out,
                                                ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return forward_f64<3>("forward_project_3d_otf_f64", volume, batch, n, src, dst, n_ray, M_inv, b, mode,
                          CTLAB_ZERO_THRESH, out, as_stream(stream));
}

extern "C" int ctlab_back_project_3d_otf_f64(const double* sino, int64_t batch, int64_t n_ray, const float* src,
                                             const float* dst, const float* M_inv, const float* b, int64_t n_x,
                                             int64_t n_y, int64_t n_z, int mode, int accumulate, double* out,
                                             ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return back_f64<3>("back_project_3d_otf_f64", sino, batch, n_ray, src, dst, M_inv, b, n, mode, CTLAB_ZERO_THRESH,
                       accumulate, out, as_stream(stream));
}

extern "C" int ctlab_forward_project_2d_otf_f64(const double* image, int64_t batch, int64_t n_row, int64_t n_col,
                                                const float* src, const float* dst, int64_t n_ray,
                                                const float* M_inv, const float* b, int mode, int zero_rule,
                                                double* out, ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return forward_f64<2>("forward_project_2d_otf_f64", image, batch, n, src, dst, n_ray, M_inv, b, mode, zero_rule,
                          out, as_stream(stream));
}

extern "C" int ctlab_back_project_2d_otf_f64(const double* sino, int64_t batch, int64_t n_ray, const float* src,
                                             const float* dst, const float* M_inv, const float* b, int64_t n_row,
                                             int64_t n_col, int mode, int zero_rule, int accumulate, double* out,
                                             ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return back_f64<2>("back_project_2d_otf_f64", sino, batch, n_ray, src, dst, M_inv, b, n, mode, zero_rule,
                       accumulate, out, as_stream(stream));
}
