// ctlab_tvals.hip -- projectors over precomputed crossing tables, the table
// builders, and the uint16 compression of the tables.
//
//   compute_intersections_{2d,3d}   sorted crossings [n_ray, K] + INF padding
//                                    (merge of the per-axis sequences; the
//                                    reference bubble-sorts, 3d.cu:90-95)
//   forward/back with float tables  (2d.cu:76-155, 3d.cu:272-376)
//   compress_tvals_3d                uint16 deltas (3d.cu:160-212 or the torch
//                                    variant 3d_torch.py:4-65)
//   forward/back with uint16 tables (3d.cu:99-157, 215-269)
//
// One lane per ray; a lane reads (or writes) its own table row in order, so
// each 128-byte line of the row is reused by the same lane on consecutive
// iterations.  Segment geometry follows the torch consumer's op order
// (segment_from_t), so for the same crossings the voxel sequence equals the
// reference's.
#include "ctlab_common.h"
#include "siddon_traverse.h"
#include "table_walk.h"

namespace ctlab {

template <int ND>
struct DimsT {
    int n[ND];
};

template <int ND>
__device__ __forceinline__ void load_ray_t(const float* __restrict__ src, const float* __restrict__ dst, int64_t r,
                                           float* s, float* d) {
#pragma unroll
    for (int a = 0; a < ND; ++a) {
        s[a] = src[r * ND + a];
        d[a] = dst[r * ND + a];
    }
}

// ---------------------------------------------------------------------------
// Crossing tables
// ---------------------------------------------------------------------------

template <int ND>
__global__ __launch_bounds__(kBlock) void k_intersections(int64_t n_ray, DimsT<ND> dims,
                                                          const float* __restrict__ src,
                                                          const float* __restrict__ dst,
                                                          const float* __restrict__ minv,
                                                          const float* __restrict__ b, int zero_rule, int64_t K,
                                                          float* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ray) return;
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, zero_rule);
    float s[ND], d[ND];
    load_ray_t<ND>(src, dst, r, s, d);
    float* row = out + r * K;
    int64_t k = 0;
    emit_crossings<ND>(g, s, d, [&](float t) {
        if (k < K) row[k] = t;
        ++k;
    });
    const float kInf = __builtin_huge_valf();
    for (; k < K; ++k) row[k] = kInf;
}

// Number of crossings per ray (for sizing trimmed tables).
template <int ND>
__global__ __launch_bounds__(kBlock) void k_count_crossings(int64_t n_ray, DimsT<ND> dims,
                                                            const float* __restrict__ src,
                                                            const float* __restrict__ dst,
                                                            const float* __restrict__ minv,
                                                            const float* __restrict__ b, int zero_rule,
                                                            int32_t* __restrict__ counts) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ray) return;
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, zero_rule);
    float s[ND], d[ND];
    load_ray_t<ND>(src, dst, r, s, d);
    counts[r] = emit_crossings<ND>(g, s, d, [](float) {});
}

// ---------------------------------------------------------------------------
// Projectors over float tables
// ---------------------------------------------------------------------------

template <int ND>
__global__ __launch_bounds__(kBlock) void k_forward_tvals(const float* __restrict__ vol, float* __restrict__ out,
                                                          int64_t n_ray, int64_t vol_elems, DimsT<ND> dims,
                                                          const float* __restrict__ tvals, int64_t K,
                                                          const float* __restrict__ src,
                                                          const float* __restrict__ dst,
                                                          const float* __restrict__ minv,
                                                          const float* __restrict__ b) {
    __shared__ RowStage<float, kStageF32> st;
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t r = r0 + threadIdx.x;
    const bool active = r < n_ray;
    const int64_t bi = blockIdx.y;
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, ZERO_THRESH);
    float s[ND] = {}, d[ND] = {};
    if (active) load_ray_t<ND>(src, dst, r, s, d);
    const float* __restrict__ v = vol + bi * vol_elems;
    // Each gathered value is consumed one segment later, so its load is in
    // flight while the next segment is computed; the chain of fused
    // multiply-adds is the same (the first, fma(0, 0, +0), leaves +0).
    float acc = 0.0f, pv = 0.0f, pl = 0.0f;
    auto gather = [&](int64_t, uint32_t, const float* c, float len) {
        uint32_t idx;
        if (flat_voxel<ND>(g, c, &idx)) {
            acc = __builtin_fmaf(pv, pl, acc);
            pv = v[idx];
            pl = len;
        }
    };
    if (g.diag)
        walk_table<ND, true>(g, s, d, tvals, K, r0, n_ray, active, st, gather);
    else
        walk_table<ND, false>(g, s, d, tvals, K, r0, n_ray, active, st, gather);
    acc = __builtin_fmaf(pv, pl, acc);
    if (active) out[bi * n_ray + r] = acc;
}

template <int ND>
__global__ __launch_bounds__(kBlock) void k_back_tvals(const float* __restrict__ sino, float* __restrict__ out,
                                                       int64_t n_ray, int64_t vol_elems, DimsT<ND> dims,
                                                       const float* __restrict__ tvals, int64_t K,
                                                       const float* __restrict__ src, const float* __restrict__ dst,
                                                       const float* __restrict__ minv,
                                                       const float* __restrict__ b) {
    __shared__ RowStage<float, kStageF32> st;
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t r = r0 + threadIdx.x;
    const int64_t bi = blockIdx.y;
    const float val = (r < n_ray) ? sino[bi * n_ray + r] : 0.0f;
    const bool active = r < n_ray && val != 0.0f;
    // a block without a nonzero value stages no rows (every lane gets the same answer)
    if (!__syncthreads_or(active)) return;
    const Grid<ND> g = load_grid<ND>(dims.n, minv, b, ZERO_THRESH);
    float s[ND] = {}, d[ND] = {};
    if (active) load_ray_t<ND>(src, dst, r, s, d);
    float* __restrict__ o = out + bi * vol_elems;
    auto scatter = [&](int64_t, uint32_t, const float* c, float len) {
        uint32_t idx;
        if (flat_voxel<ND>(g, c, &idx)) unsafeAtomicAdd(o + idx, val * len);
    };
    if (g.diag)
        walk_table<ND, true>(g, s, d, tvals, K, r0, n_ray, active, st, scatter);
    else
        walk_table<ND, false>(g, s, d, tvals, K, r0, n_ray, active, st, scatter);
}

// ---------------------------------------------------------------------------
// uint16 compression
// ---------------------------------------------------------------------------

enum : int { COMPRESS_CUDA = 0, COMPRESS_TORCH = 1 };

template <int SEM>
__global__ __launch_bounds__(kBlock) void k_compress(const float* __restrict__ tvals, int64_t n_ray, int64_t K,
                                                     const float* __restrict__ src, const float* __restrict__ dst,
                                                     float voxel_diag_len, uint16_t* __restrict__ u16,
                                                     float* __restrict__ start, float* __restrict__ scale) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_ray) return;
    const float vx = dst[3 * r] - src[3 * r], vy = dst[3 * r + 1] - src[3 * r + 1],
                vz = dst[3 * r + 2] - src[3 * r + 2];
    // Ray length: sqrtf of the squares summed as an fma chain.  nvcc contracts
    // the reference kernel's sum to it (3d.cu:172-173), and torch.norm(dim=1)
    // of a float32 [n,3] tensor on the CPU computes the same (3d_torch.py:38-42).
    float rl = __builtin_sqrtf(__builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx)));
    if (SEM == COMPRESS_CUDA) {
        rl = fmaxf(rl, 1e-6f);
    } else {
        if (rl < 1e-6f) rl = 1e-6f;  // clamp(min=1e-6) keeps a NaN
    }
    const float sc = (voxel_diag_len / 10000.0f) / rl;
    scale[r] = sc;
    const float inv = 1.0f / sc;
    const float* __restrict__ row = tvals + r * K;
    uint16_t* __restrict__ o = u16 + r * K;
    if (SEM == COMPRESS_CUDA) {
        const float st = row[0];
        start[r] = st;
        o[0] = 0;
        uint32_t kprev = 0;
        for (int64_t i = 1; i < K; ++i) {
            const float t = row[i];
            if (isinf(t)) {
                o[i] = 0;
                continue;
            }
            const float q = roundf((t - st) * inv);  // ties away from zero
            const uint32_t kc = (q >= 0.0f) ? ((q < 4294967296.0f) ? (uint32_t)q : 0xffffffffu) : 0u;
            int64_t diff = (int64_t)(int32_t)kc - (int64_t)(int32_t)kprev;
            diff = diff < 0 ? 0 : (diff > 32000 ? 32000 : diff);
            o[i] = (uint16_t)diff;
            kprev += (uint32_t)diff;
        }
    } else {
        const float st = isinf(row[0]) ? 10.0f : row[0];
        start[r] = st;
        o[0] = 0;
        float kprev = 0.0f;
        for (int64_t i = 1; i < K; ++i) {
            const float t = isinf(row[i]) ? 10.0f : row[i];
            const float kc = __builtin_rintf((t - st) * inv);  // torch.round: ties to even
            float df = kc - kprev;
            df = df < 0.0f ? 0.0f : (df > 32000.0f ? 32000.0f : df);
            o[i] = (uint16_t)(int)df;
            kprev += df;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_forward_u16(const float* __restrict__ vol, float* __restrict__ out,
                                                        int64_t n_ray, int64_t vol_elems, DimsT<3> dims,
                                                        const uint16_t* __restrict__ u16,
                                                        const float* __restrict__ start,
                                                        const float* __restrict__ scale, int64_t K,
                                                        const float* __restrict__ src, const float* __restrict__ dst,
                                                        const float* __restrict__ minv,
                                                        const float* __restrict__ b) {
    __shared__ RowStage<uint16_t, kStageU16> st;
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t r = r0 + threadIdx.x;
    const bool active = r < n_ray;
    const int64_t bi = blockIdx.y;
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    float s[3] = {}, d[3] = {};
    float st0 = 0.0f, sc = 0.0f;
    if (active) {
        load_ray_t<3>(src, dst, r, s, d);
        st0 = start[r];
        sc = scale[r];
    }
    const float* __restrict__ v = vol + bi * vol_elems;
    float acc = 0.0f, pv = 0.0f, pl = 0.0f;  // gathers consumed one segment later (as k_forward_tvals)
    auto gather = [&](int64_t, uint32_t, const float* c, float len) {
        uint32_t idx;
        if (flat_voxel<3>(g, c, &idx)) {
            acc = __builtin_fmaf(pv, pl, acc);
            pv = v[idx];
            pl = len;
        }
    };
    if (g.diag)
        walk_u16<true>(g, s, d, u16, K, r0, n_ray, active, st0, sc, st, gather);
    else
        walk_u16<false>(g, s, d, u16, K, r0, n_ray, active, st0, sc, st, gather);
    acc = __builtin_fmaf(pv, pl, acc);
    if (active) out[bi * n_ray + r] = acc;
}

__global__ __launch_bounds__(kBlock) void k_back_u16(const float* __restrict__ sino, float* __restrict__ out,
                                                     int64_t n_ray, int64_t vol_elems, DimsT<3> dims,
                                                     const uint16_t* __restrict__ u16,
                                                     const float* __restrict__ start,
                                                     const float* __restrict__ scale, int64_t K,
                                                     const float* __restrict__ src, const float* __restrict__ dst,
                                                     const float* __restrict__ minv, const float* __restrict__ b) {
    __shared__ RowStage<uint16_t, kStageU16> st;
    const int64_t r0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t r = r0 + threadIdx.x;
    const int64_t bi = blockIdx.y;
    const float val = (r < n_ray) ? sino[bi * n_ray + r] : 0.0f;
    const bool active = r < n_ray && val != 0.0f;
    // a block without a nonzero value stages no rows (every lane gets the same answer)
    if (!__syncthreads_or(active)) return;
    const Grid<3> g = load_grid<3>(dims.n, minv, b, ZERO_THRESH);
    float s[3] = {}, d[3] = {};
    float st0 = 0.0f, sc = 0.0f;
    if (active) {
        load_ray_t<3>(src, dst, r, s, d);
        st0 = start[r];
        sc = scale[r];
    }
    float* __restrict__ o = out + bi * vol_elems;
    auto scatter = [&](int64_t, uint32_t, const float* c, float len) {
        uint32_t idx;
        if (flat_voxel<3>(g, c, &idx)) unsafeAtomicAdd(o + idx, val * len);
    };
    if (g.diag)
        walk_u16<true>(g, s, d, u16, K, r0, n_ray, active, st0, sc, st, scatter);
    else
        walk_u16<false>(g, s, d, u16, K, r0, n_ray, active, st0, sc, st, scatter);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

template <int ND>
static int check_grid(const char* fn, const int64_t* n, int64_t* V) {
    int64_t vox = 1;
    for (int a = 0; a < ND; ++a) {
        CTLAB_REQUIRE(n[a] >= 1 && n[a] <= 0x7fffffffLL, "%s: grid dimension %d is %lld (must be >= 1)", fn, a,
                      (long long)n[a]);
        vox *= n[a];
        CTLAB_REQUIRE(vox <= kMaxVoxels, "%s: grid has more than 2^31-1 voxels", fn);
    }
    int st = check_index_limits(fn, n, ND);
    if (st) return st;
    *V = vox;
    return CTLAB_OK;
}

static int check_rays(const char* fn, int64_t n_ray, int64_t batch, const void* src, const void* dst) {
    CTLAB_REQUIRE(n_ray >= 0, "%s: n_ray %lld < 0", fn, (long long)n_ray);
    CTLAB_REQUIRE(batch >= 1 && batch <= 65535, "%s: batch %lld out of range [1, 65535]", fn, (long long)batch);
    CTLAB_REQUIRE(grid_for(n_ray) <= kMaxGridX, "%s: too many rays", fn);
    if (n_ray > 0) CTLAB_REQUIRE(src && dst, "%s: src/dst is null", fn);
    return CTLAB_OK;
}

template <int ND>
static DimsT<ND> dims_of(const int64_t* n) {
    DimsT<ND> d;
    for (int a = 0; a < ND; ++a) d.n[a] = (int)n[a];
    return d;
}

template <int ND>
static int intersections(const char* fn, const int64_t* n, const float* src, const float* dst, int64_t n_ray,
                         const float* minv, const float* b, int zero_rule, int64_t K, float* out,
                         hipStream_t stream) {
    int64_t V;
    int st = check_grid<ND>(fn, n, &V);
    if (st) return st;
    st = check_rays(fn, n_ray, 1, src, dst);
    if (st) return st;
    CTLAB_REQUIRE(K >= 0, "%s: K < 0", fn);
    if (n_ray == 0 || K == 0) return CTLAB_OK;
    CTLAB_REQUIRE(out && minv && b, "%s: null pointer", fn);
    hipLaunchKernelGGL((k_intersections<ND>), dim3(grid_for(n_ray)), dim3(kBlock), 0, stream, n_ray, dims_of<ND>(n),
                       src, dst, minv, b, zero_rule, K, out);
    return check_launch(fn);
}

template <int ND>
static int count_crossings(const char* fn, const int64_t* n, const float* src, const float* dst, int64_t n_ray,
                           const float* minv, const float* b, int zero_rule, int32_t* counts, hipStream_t stream) {
    int64_t V;
    int st = check_grid<ND>(fn, n, &V);
    if (st) return st;
    st = check_rays(fn, n_ray, 1, src, dst);
    if (st) return st;
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(counts && minv && b, "%s: null pointer", fn);
    hipLaunchKernelGGL((k_count_crossings<ND>), dim3(grid_for(n_ray)), dim3(kBlock), 0, stream, n_ray,
                       dims_of<ND>(n), src, dst, minv, b, zero_rule, counts);
    return check_launch(fn);
}

template <int ND>
static int forward_tvals(const char* fn, const float* vol, int64_t batch, const int64_t* n, const float* tvals,
                         int64_t K, const float* src, const float* dst, int64_t n_ray, const float* minv,
                         const float* b, float* out, hipStream_t stream) {
    int64_t V;
    int st = check_grid<ND>(fn, n, &V);
    if (st) return st;
    st = check_rays(fn, n_ray, batch, src, dst);
    if (st) return st;
    CTLAB_REQUIRE(K >= 0, "%s: K < 0", fn);
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(vol && out && minv && b && (K == 0 || tvals), "%s: null pointer", fn);
    hipLaunchKernelGGL((k_forward_tvals<ND>), dim3(grid_for(n_ray), (unsigned)batch), dim3(kBlock), 0, stream, vol,
                       out, n_ray, V, dims_of<ND>(n), tvals, K, src, dst, minv, b);
    return check_launch(fn);
}

template <int ND>
static int back_tvals(const char* fn, const float* sino, int64_t batch, const float* tvals, int64_t K,
                      const float* src, const float* dst, int64_t n_ray, const float* minv, const float* b,
                      const int64_t* n, int accumulate, float* out, hipStream_t stream) {
    int64_t V;
    int st = check_grid<ND>(fn, n, &V);
    if (st) return st;
    st = check_rays(fn, n_ray, batch, src, dst);
    if (st) return st;
    CTLAB_REQUIRE(K >= 0, "%s: K < 0", fn);
    CTLAB_REQUIRE(out, "%s: out is null", fn);
    if (!accumulate) {
        st = fill_f32(out, V * batch, 0.0f, stream);
        if (st) return st;
    }
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(sino && minv && b && (K == 0 || tvals), "%s: null pointer", fn);
    hipLaunchKernelGGL((k_back_tvals<ND>), dim3(grid_for(n_ray), (unsigned)batch), dim3(kBlock), 0, stream, sino,
                       out, n_ray, V, dims_of<ND>(n), tvals, K, src, dst, minv, b);
    return check_launch(fn);
}

}  // namespace ctlab

using namespace ctlab;

extern "C" int ctlab_compute_intersections_2d(int64_t n_row, int64_t n_col, const float* src, const float* dst,
                                              int64_t n_ray, const float* M_inv, const float* b, int zero_rule,
                                              int64_t K, float* out, ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return intersections<2>("compute_intersections_2d", n, src, dst, n_ray, M_inv, b, zero_rule, K, out,
                            as_stream(stream));
}

extern "C" int ctlab_compute_intersections_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* src,
                                              const float* dst, int64_t n_ray, const float* M_inv, const float* b,
                                              int zero_rule, int64_t K, float* out, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return intersections<3>("compute_intersections_3d", n, src, dst, n_ray, M_inv, b, zero_rule, K, out,
                            as_stream(stream));
}

extern "C" int ctlab_count_crossings_3d(int64_t n_x, int64_t n_y, int64_t n_z, const float* src, const float* dst,
                                        int64_t n_ray, const float* M_inv, const float* b, int zero_rule,
                                        int32_t* counts, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return count_crossings<3>("count_crossings_3d", n, src, dst, n_ray, M_inv, b, zero_rule, counts,
                              as_stream(stream));
}

extern "C" int ctlab_count_crossings_2d(int64_t n_row, int64_t n_col, const float* src, const float* dst,
                                        int64_t n_ray, const float* M_inv, const float* b, int zero_rule,
                                        int32_t* counts, ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return count_crossings<2>("count_crossings_2d", n, src, dst, n_ray, M_inv, b, zero_rule, counts,
                              as_stream(stream));
}

extern "C" int ctlab_forward_project_2d_tvals(const float* image, int64_t batch, int64_t n_row, int64_t n_col,
                                              const float* tvals, int64_t K, const float* src, const float* dst,
                                              int64_t n_ray, const float* M_inv, const float* b, float* out,
                                              ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return forward_tvals<2>("forward_project_2d_tvals", image, batch, n, tvals, K, src, dst, n_ray, M_inv, b, out,
                            as_stream(stream));
}

extern "C" int ctlab_back_project_2d_tvals(const float* sino, int64_t batch, const float* tvals, int64_t K,
                                           const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                           const float* b, int64_t n_row, int64_t n_col, int accumulate, float* out,
                                           ctlab_stream_t stream) {
    const int64_t n[2] = {n_row, n_col};
    return back_tvals<2>("back_project_2d_tvals", sino, batch, tvals, K, src, dst, n_ray, M_inv, b, n, accumulate,
                         out, as_stream(stream));
}

extern "C" int ctlab_forward_project_3d_tvals(const float* volume, int64_t batch, int64_t n_x, int64_t n_y,
                                              int64_t n_z, const float* tvals, int64_t K, const float* src,
                                              const float* dst, int64_t n_ray, const float* M_inv, const float* b,
                                              float* out, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return forward_tvals<3>("forward_project_3d_tvals", volume, batch, n, tvals, K, src, dst, n_ray, M_inv, b, out,
                            as_stream(stream));
}

extern "C" int ctlab_back_project_3d_tvals(const float* sino, int64_t batch, const float* tvals, int64_t K,
                                           const float* src, const float* dst, int64_t n_ray, const float* M_inv,
                                           const float* b, int64_t n_x, int64_t n_y, int64_t n_z, int accumulate,
                                           float* out, ctlab_stream_t stream) {
    const int64_t n[3] = {n_x, n_y, n_z};
    return back_tvals<3>("back_project_3d_tvals", sino, batch, tvals, K, src, dst, n_ray, M_inv, b, n, accumulate,
                         out, as_stream(stream));
}

extern "C" int ctlab_compress_tvals_3d(const float* tvals, int64_t n_ray, int64_t K, const float* src,
                                       const float* dst, float voxel_diag_len, int semantics, uint16_t* u16,
                                       float* start, float* scale, ctlab_stream_t stream) {
    const char* fn = "compress_tvals_3d";
    int st = check_rays(fn, n_ray, 1, src, dst);
    if (st) return st;
    CTLAB_REQUIRE(K >= 1, "%s: K must be >= 1", fn);
    CTLAB_REQUIRE(semantics == 0 || semantics == 1, "%s: unknown semantics %d", fn, semantics);
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(tvals && u16 && start && scale, "%s: null pointer", fn);
    if (semantics == COMPRESS_CUDA)
        hipLaunchKernelGGL((k_compress<COMPRESS_CUDA>), dim3(grid_for(n_ray)), dim3(kBlock), 0, as_stream(stream),
                           tvals, n_ray, K, src, dst, voxel_diag_len, u16, start, scale);
    else
        hipLaunchKernelGGL((k_compress<COMPRESS_TORCH>), dim3(grid_for(n_ray)), dim3(kBlock), 0, as_stream(stream),
                           tvals, n_ray, K, src, dst, voxel_diag_len, u16, start, scale);
    return check_launch(fn);
}

extern "C" int ctlab_forward_project_3d_u16(const float* volume, int64_t batch, int64_t n_x, int64_t n_y,
                                            int64_t n_z, const uint16_t* u16, const float* start, const float* scale,
                                            int64_t K, const float* src, const float* dst, int64_t n_ray,
                                            const float* M_inv, const float* b, float* out, ctlab_stream_t stream) {
    const char* fn = "forward_project_3d_u16";
    const int64_t n[3] = {n_x, n_y, n_z};
    int64_t V;
    int st = check_grid<3>(fn, n, &V);
    if (st) return st;
    st = check_rays(fn, n_ray, batch, src, dst);
    if (st) return st;
    CTLAB_REQUIRE(K >= 0, "%s: K < 0", fn);
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(volume && out && M_inv && b && start && scale && (K == 0 || u16), "%s: null pointer", fn);
    hipLaunchKernelGGL(k_forward_u16, dim3(grid_for(n_ray), (unsigned)batch), dim3(kBlock), 0, as_stream(stream),
                       volume, out, n_ray, V, dims_of<3>(n), u16, start, scale, K, src, dst, M_inv, b);
    return check_launch(fn);
}

extern "C" int ctlab_back_project_3d_u16(const float* sino, int64_t batch, const uint16_t* u16, const float* start,
                                         const float* scale, int64_t K, const float* src, const float* dst,
                                         int64_t n_ray, const float* M_inv, const float* b, int64_t n_x, int64_t n_y,
                                         int64_t n_z, int accumulate, float* out, ctlab_stream_t stream) {
    const char* fn = "back_project_3d_u16";
    const int64_t n[3] = {n_x, n_y, n_z};
    int64_t V;
    int st = check_grid<3>(fn, n, &V);
    if (st) return st;
    st = check_rays(fn, n_ray, batch, src, dst);
    if (st) return st;
    CTLAB_REQUIRE(K >= 0, "%s: K < 0", fn);
    CTLAB_REQUIRE(out, "%s: out is null", fn);
    if (!accumulate) {
        st = fill_f32(out, V * batch, 0.0f, as_stream(stream));
        if (st) return st;
    }
    if (n_ray == 0) return CTLAB_OK;
    CTLAB_REQUIRE(sino && M_inv && b && start && scale && (K == 0 || u16), "%s: null pointer", fn);
    hipLaunchKernelGGL(k_back_u16, dim3(grid_for(n_ray), (unsigned)batch), dim3(kBlock), 0, as_stream(stream), sino,
                       out, n_ray, V, dims_of<3>(n), u16, start, scale, K, src, dst, M_inv, b);
    return check_launch(fn);
}
